"""An f64 numpy evaluation of TFLite flatbuffer models, for checks.

Written from TensorFlow Lite's published schema (schema.fbs: Model,
SubGraph, Tensor, Operator, OperatorCode, Buffer, SignatureDef and the
builtin option tables, by field id) and the documented semantics of the
builtin operators the speech lattice model uses.  It shares nothing with
the oracle's or the library's C/C++ interpreters: its own flatbuffer
reader, numpy broadcasting, and every product and sum in float64 (numpy's
own summation order), so it pins what the model computes, not how either
interpreter rounds it.
"""
from __future__ import annotations

import struct

import numpy as np

# BuiltinOperator codes (schema.fbs)
ADD, CONCATENATION, FULLY_CONNECTED, MUL, RELU, RELU6, RESHAPE = 0, 2, 9, 18, 19, 21, 22
SUB, SPLIT, CAST, MAXIMUM, MINIMUM, EXPAND_DIMS, EQUAL, SHAPE = 41, 49, 53, 55, 57, 70, 71, 77
FILL, ABS, BATCH_MATMUL = 94, 101, 126
# TensorType
_DTYPES = {0: np.float64, 2: np.int32, 3: np.uint8, 4: np.int64, 6: np.bool_}
_RAW = {0: "<f4", 2: "<i4", 3: "u1", 4: "<i8", 6: "u1"}


def _fc(flat: np.ndarray, w: np.ndarray) -> np.ndarray:
    """FULLY_CONNECTED's product: rows of `flat` against rows of the
    [out, in] weights (numpy's own summation order)."""
    return flat @ w.T


class _Buf:
    """Little-endian flatbuffer access by table position and field id."""

    def __init__(self, data: bytes):
        self.d = data

    def u8(self, p): return self.d[p]
    def i8(self, p): return struct.unpack_from("<b", self.d, p)[0]
    def u16(self, p): return struct.unpack_from("<H", self.d, p)[0]
    def i32(self, p): return struct.unpack_from("<i", self.d, p)[0]
    def u32(self, p): return struct.unpack_from("<I", self.d, p)[0]
    def u64(self, p): return struct.unpack_from("<Q", self.d, p)[0]

    def root(self):
        return self.u32(0)

    def field(self, t, fid):
        vt = t - self.i32(t)
        vlen = self.u16(vt)
        at = 4 + 2 * fid
        if at + 2 > vlen:
            return None
        off = self.u16(vt + at)
        return t + off if off else None

    def scalar(self, t, fid, kind, default=0):
        p = self.field(t, fid)
        if p is None:
            return default
        return getattr(self, kind)(p)

    def deref(self, p):
        return p + self.u32(p)

    def table(self, t, fid):
        p = self.field(t, fid)
        return None if p is None else self.deref(p)

    def vector(self, t, fid):
        """(first element position, length) or None."""
        p = self.field(t, fid)
        if p is None:
            return None
        v = self.deref(p)
        return v + 4, self.u32(v)

    def tables(self, t, fid):
        v = self.vector(t, fid)
        if v is None:
            return []
        start, n = v
        return [self.deref(start + 4 * i) for i in range(n)]

    def ints(self, t, fid):
        v = self.vector(t, fid)
        if v is None:
            return None
        start, n = v
        return list(struct.unpack_from(f"<{n}i", self.d, start))

    def string(self, t, fid):
        v = self.vector(t, fid)
        if v is None:
            return None
        start, n = v
        return self.d[start:start + n].decode()


class Model:
    """One subgraph's tensors, constants and operators."""

    def __init__(self, data: bytes):
        b = _Buf(data)
        self.b = b
        m = b.root()
        codes = []
        for oc in b.tables(m, 1):
            dep = b.scalar(oc, 0, "i8")
            code = b.scalar(oc, 3, "i32")
            codes.append(max(dep, code))
        g = b.tables(m, 2)[0]
        buffers = b.tables(m, 4)
        self.tensors = []
        self.consts = {}
        for i, t in enumerate(b.tables(g, 0)):
            shape = b.ints(t, 0) or []
            ttype = b.scalar(t, 1, "u8")
            name = b.string(t, 3)
            self.tensors.append((name, ttype, shape))
            bi = b.scalar(t, 2, "u32")
            if bi == 0 or bi >= len(buffers):
                continue
            dv = b.vector(buffers[bi], 0)
            raw = None
            if dv is not None and dv[1] > 0:
                raw = data[dv[0]:dv[0] + dv[1]]
            else:
                off, size = b.scalar(buffers[bi], 1, "u64"), b.scalar(buffers[bi], 2, "u64")
                if off > 1 and size > 0:
                    raw = data[off:off + size]
            if raw is None:
                continue
            arr = np.frombuffer(raw, dtype=_RAW[ttype]).astype(_DTYPES[ttype]).reshape(shape)
            self.consts[i] = arr
        self.ops = []
        for op in b.tables(g, 3):
            code = codes[b.scalar(op, 0, "u32")]
            ins, outs = b.ints(op, 1) or [], b.ints(op, 2) or []
            opt = b.table(op, 4)
            self.ops.append((code, ins, outs, opt))
        self.inputs, self.outputs = {}, {}
        for sig in b.tables(m, 7):
            for fid, dst in ((0, self.inputs), (1, self.outputs)):
                for tm in b.tables(sig, fid):
                    dst[b.string(tm, 0)] = b.scalar(tm, 1, "u32")
        if not self.inputs:
            self.inputs = {self.tensors[i][0]: i for i in b.ints(g, 1)}
            self.outputs = {self.tensors[i][0]: i for i in b.ints(g, 2)}

    @staticmethod
    def _act(x, a):
        if a == 1:
            return np.maximum(x, 0.0)
        if a == 2:
            return np.clip(x, -1.0, 1.0)
        if a == 3:
            return np.clip(x, 0.0, 6.0)
        assert a == 0, a
        return x

    def run(self, feeds: dict, output: str, dtype=np.float64) -> np.ndarray:
        """Float tensors in `dtype` (float64 by default; float32 rounds every
        op as TFLite's reference kernels do, in numpy's summation order)."""
        b = self.b
        v = {i: (c.astype(dtype) if c.dtype == np.float64 else c) for i, c in self.consts.items()}
        for name, x in feeds.items():
            v[self.inputs[name]] = np.asarray(x, dtype=dtype)
        for code, ins, outs, opt in self.ops:
            x = [v[i] if i >= 0 else None for i in ins]
            if code in (ADD, SUB, MUL, MINIMUM, MAXIMUM):
                f = {ADD: np.add, SUB: np.subtract, MUL: np.multiply, MINIMUM: np.minimum, MAXIMUM: np.maximum}[code]
                a = b.scalar(opt, 0, "i8") if opt is not None and code in (ADD, SUB, MUL) else 0
                r = self._act(f(x[0], x[1]), a)
            elif code == EQUAL:
                r = np.equal(x[0], x[1])
            elif code in (RELU, RELU6, ABS):
                r = {RELU: lambda t: np.maximum(t, 0.0), RELU6: lambda t: np.clip(t, 0.0, 6.0), ABS: np.abs}[code](x[0])
            elif code == CAST:
                to = _DTYPES[self.tensors[outs[0]][1]]
                r = x[0].astype(dtype if to == np.float64 else to)
            elif code == CONCATENATION:
                axis, a = b.scalar(opt, 0, "i32"), b.scalar(opt, 1, "i8")
                r = self._act(np.concatenate(x, axis=axis), a)
            elif code == FULLY_CONNECTED:
                w = x[1]
                a = b.scalar(opt, 0, "i8") if opt is not None else 0
                keep = bool(b.scalar(opt, 2, "u8")) if opt is not None else False
                flat = x[0].reshape(-1, w.shape[1])
                r = _fc(flat, w)
                if len(x) > 2 and x[2] is not None:
                    r = r + x[2]
                r = self._act(r, a)
                if keep:
                    r = r.reshape(list(x[0].shape[:-1]) + [w.shape[0]])
            elif code == BATCH_MATMUL:
                adj_x = bool(b.scalar(opt, 0, "u8")) if opt is not None else False
                adj_y = bool(b.scalar(opt, 1, "u8")) if opt is not None else False
                p = np.swapaxes(x[0], -1, -2) if adj_x else x[0]
                q = np.swapaxes(x[1], -1, -2) if adj_y else x[1]
                r = np.matmul(p, q)
            elif code == RESHAPE:
                shape = list(x[1].astype(np.int64).ravel()) if len(x) > 1 and x[1] is not None else b.ints(opt, 0)
                r = x[0].reshape(shape)
            elif code == EXPAND_DIMS:
                r = np.expand_dims(x[0], int(x[1].ravel()[0]))
            elif code == SPLIT:
                r = np.split(x[1], len(outs), axis=int(x[0].ravel()[0]))
            elif code == SHAPE:
                r = np.array(x[0].shape, dtype=np.int32)
            elif code == FILL:
                r = np.full(tuple(int(d) for d in x[0].ravel()), x[1].ravel()[0])
            else:
                raise NotImplementedError(f"builtin op {code}")
            if code == SPLIT:
                for o, part in zip(outs, r):
                    v[o] = part
            else:
                v[outs[0]] = r
        return v[self.outputs[output]]

    def op_codes(self):
        return [c for c, _, _, _ in self.ops]


def lattice_mos(model: Model, fvnsim, fvnsim10, fstdnsim, fvdegenergy, tau: float = 0.5, dtype=np.float64) -> float:
    """The speech lattice mapping's inputs (one scalar per feature and band,
    named fvnsim<b>, fvnsim10_<b>, fstdnsim<b>, fvdegenergy<b>, and tau),
    evaluated in `dtype`; the 'predictions' output."""
    feeds = {"tau": np.array([tau])}
    for name, vals in (("fvnsim", fvnsim), ("fvnsim10_", fvnsim10), ("fstdnsim", fstdnsim),
                       ("fvdegenergy", fvdegenergy)):
        for k, x in enumerate(vals):
            feeds[f"{name}{k}"] = np.array([float(x)])
    return float(np.asarray(model.run(feeds, "predictions", dtype)).ravel()[0])
