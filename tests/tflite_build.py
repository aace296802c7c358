"""Tiny TFLite flatbuffers built from the published schema (schema.fbs), for
known-answer tests of the lattice interpreters through their loaders.

A model here has the speech lattice's input names for one band (fvnsim0,
fvnsim10_0, fstdnsim0, fvdegenergy0, tau: scalars of shape [1]) and one
output named "predictions"; its graph computes from constant tensors and
ends in a one-hot FULLY_CONNECTED that picks one element of the tensor
under test (a dot product with a single 1 and zeros elsewhere, exact in any
order), so the scalar prediction exposes that element.
"""
from __future__ import annotations

import struct

import numpy as np

F32, I32, BOOL = 0, 2, 6


class _Table:
    def __init__(self, fields):
        self.fields = fields  # [(field id, kind, value)]; kind: "u8" "i8" "i32" "u32" "table" "vec_tables" "vec_i32" "vec_u8" "str"


class _Writer:
    """Forward layout: every referenced object sits after its reference
    (unsigned offsets point forward); each table's vtable sits right before
    it."""

    def __init__(self):
        self.buf = bytearray()
        self.fix = []  # (position of a u32 offset, object)

    def align(self, n):
        while len(self.buf) % n:
            self.buf.append(0)

    def table(self, t: _Table) -> int:
        nf = max((f[0] for f in t.fields), default=-1) + 1
        # inline layout of the fields after the soffset
        layout, size = [], 4
        for fid, kind, val in sorted(t.fields, key=lambda f: f[0]):
            w = {"u8": 1, "i8": 1, "i32": 4, "u32": 4}.get(kind, 4)
            size = (size + w - 1) // w * w
            layout.append((fid, kind, val, size, w))
            size += w
        vt = struct.pack("<HH", 4 + 2 * nf, size) + b"".join(
            struct.pack("<H", next((off for f, _, _, off, _ in layout if f == i), 0)) for i in range(nf))
        self.align(4)
        if (len(self.buf) + len(vt)) % 4:
            self.buf += b"\0\0"
        vpos = len(self.buf)
        self.buf += vt
        tpos = len(self.buf)
        self.buf += bytes(size)
        struct.pack_into("<i", self.buf, tpos, tpos - vpos)
        for fid, kind, val, off, w in layout:
            p = tpos + off
            if kind == "u8":
                struct.pack_into("<B", self.buf, p, val)
            elif kind == "i8":
                struct.pack_into("<b", self.buf, p, val)
            elif kind == "i32":
                struct.pack_into("<i", self.buf, p, val)
            elif kind == "u32":
                struct.pack_into("<I", self.buf, p, val)
            else:
                self.fix.append((p, kind, val))
        return tpos

    def obj(self, kind, val) -> int:
        if kind == "table":
            return self.table(val)
        if kind == "str":
            self.align(4)
            p = len(self.buf)
            raw = val.encode()
            self.buf += struct.pack("<I", len(raw)) + raw + b"\0"
            return p
        if kind == "vec_i32":
            self.align(4)
            p = len(self.buf)
            self.buf += struct.pack(f"<I{len(val)}i", len(val), *val)
            return p
        if kind == "vec_u8":
            self.align(16)
            self.buf += b"\0" * 12  # so the bytes after the length start 16-aligned
            p = len(self.buf)
            self.buf += struct.pack("<I", len(val)) + bytes(val)
            return p
        if kind == "vec_tables":
            self.align(4)
            p = len(self.buf)
            self.buf += struct.pack("<I", len(val)) + bytes(4 * len(val))
            for i, t in enumerate(val):
                self.fix.append((p + 4 + 4 * i, "table", t))
            return p
        raise ValueError(kind)

    def finish(self, root: _Table) -> bytes:
        self.buf += bytes(4) + b"TFL3"  # root offset, file identifier
        self.fix.append((0, "table", root))
        while self.fix:
            at, kind, val = self.fix.pop(0)
            p = self.obj(kind, val)
            struct.pack_into("<I", self.buf, at, p - at)
        return bytes(self.buf)


Table = _Table


def _tensor(name, ttype, shape, buffer):
    return _Table([(0, "vec_i32", list(shape)), (1, "u8", ttype), (2, "u32", buffer), (3, "str", name)])


class Graph:
    """Builds one model: inputs of the lattice's names, constants, ops."""

    def __init__(self):
        self.tensors = []   # (name, type, shape, buffer index)
        self.buffers = [b""]  # buffer 0: empty, as the schema asks
        self.ops = []
        self.codes = []
        self.index = {}
        for name in ("fvnsim0", "fvnsim10_0", "fstdnsim0", "fvdegenergy0", "tau"):
            self.tensor(name, F32, [1])
        self.inputs = [0, 1, 2, 3, 4]

    def tensor(self, name, ttype, shape, data: bytes | None = None) -> int:
        b = 0
        if data is not None:
            self.buffers.append(data)
            b = len(self.buffers) - 1
        self.tensors.append((name, ttype, list(shape), b))
        self.index[name] = len(self.tensors) - 1
        return self.index[name]

    def const(self, name, arr: np.ndarray) -> int:
        if arr.dtype == np.float32:
            t = F32
        elif arr.dtype == np.int32:
            t = I32
        else:
            t = BOOL
            arr = arr.astype(np.uint8)
        return self.tensor(name, t, arr.shape, arr.astype(arr.dtype.newbyteorder("<")).tobytes())

    def op(self, code, ins, outs, options=None, options_type=0):
        """One operator.  options_type is written as the BuiltinOptions union
        tag; the interpreters under test (and tflite_np) read the options
        table by the operator's code, not by the tag."""
        if code not in self.codes:
            self.codes.append(code)
        self.ops.append((self.codes.index(code), [self.index[i] if i else -1 for i in ins],
                         [self.index[o] for o in outs], options, options_type))

    def pick(self, name, length, k):
        """predictions = element k of `name` (flattened, `length` long)."""
        self.tensor(name + "_row", self.tensors[self.index[name]][1], [1, length])
        self.op(22, [name], [name + "_row"], _Table([(0, "vec_i32", [1, length])]), 17)
        w = np.zeros((1, length), np.float32)
        w[0, k] = 1.0
        self.const("onehot", w)
        self.tensor("predictions", F32, [1, 1])
        self.op(9, [name + "_row", "onehot", None], ["predictions"], _Table([(0, "i8", 0), (2, "u8", 0)]), 8)

    def bytes(self) -> bytes:
        op_tables = []
        for ci, ins, outs, opt, opt_type in self.ops:
            fields = [(0, "u32", ci), (1, "vec_i32", ins), (2, "vec_i32", outs)]
            if opt is not None:
                fields += [(3, "u8", opt_type), (4, "table", opt)]
            op_tables.append(_Table(fields))
        sub = _Table([(0, "vec_tables", [_tensor(n, t, s, b) for n, t, s, b in self.tensors]),
                      (1, "vec_i32", self.inputs), (2, "vec_i32", [self.index["predictions"]]),
                      (3, "vec_tables", op_tables)])
        codes = [_Table([(0, "i8", min(c, 127)), (2, "i32", 1), (3, "i32", c)]) for c in self.codes]
        bufs = [_Table([(0, "vec_u8", list(b))] if b else []) for b in self.buffers]
        model = _Table([(0, "u32", 3), (1, "vec_tables", codes), (2, "vec_tables", [sub]), (4, "vec_tables", bufs)])
        return _Writer().finish(model)
