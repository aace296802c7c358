"""Pins of the oracle's aggregation and MOS mappings (CPU).

* buildResult (visqol.zig:595-709): fvnsim and fvdegenergy against numpy
  means, fvnsim10 against the mean of the lowest max(1, floor(0.1 P))
  values, fstdnsim against np.std(ddof=1) of per-patch samples built with
  the given means, sample deviations and frame counts, and MOS = 1 below
  vnsim 0.15.
* The speech exponential mapping (visqol.zig:85-96) by its closed form's
  properties: log(mos - a) affine in the mean nsim with slope b, the scaled
  score exactly the unscaled one times the f32-rounded 1.245063, both
  clamps.
* The TF-Lattice interpreter (tflite.zig:799-828 known answers; lattice.zig
  via the model): the reference's broadcast sub / equal and batch-matmul +
  reshape results and its truncated-model rejections, through tiny models
  built from the TFLite schema and run by the oracle, by this library's
  interpreter and by an independent numpy evaluation; the speech model
  itself against that numpy evaluation in f64 (|dMOS| <= 1e-4: the f32
  interpreters differ from exact arithmetic by their f32 rounding) and,
  with the 4096-term dot products summed in the interpreters' sequential
  order, in f32 (|dMOS| <= 2e-5); and piecewise linearity along single
  features.
"""
import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import tflite_build as tb  # noqa: E402
import tflite_np  # noqa: E402

MODEL = ROOT.parent / "akvisqol_amd" / "models" / "lattice_speech.tflite"


@pytest.fixture(scope="module")
def vo():
    from oracle import vo as v

    return v


# ---------------------------------------------------------------- aggregation
def _matches(rng, n, nb, fd, patch_frames=30):
    means = rng.uniform(0.2, 1.0, (n, nb))
    sds = rng.uniform(0.0, 0.3, (n, nb))
    energy = rng.uniform(0.0, 60.0, (n, nb))
    start = np.arange(n) * patch_frames * fd * 1.5
    end = start + patch_frames * fd
    return means, sds, energy, start, end


@pytest.mark.parametrize("n", [1, 5, 9, 10, 16, 23])
def test_fvnsim_fvnsim10_fvdegenergy(vo, n):
    rng = np.random.default_rng(n)
    fd = 960 / 48000
    means, sds, energy, start, end = _matches(rng, n, 32, fd)
    res = vo.Oracle(mode=0).build_result(means, sds, energy, start, end, fd)
    np.testing.assert_allclose(res["fvnsim"], means.mean(axis=0), rtol=1e-14)
    np.testing.assert_allclose(res["fvdegenergy"], energy.mean(axis=0), rtol=1e-14)
    k = max(1, int(np.floor(0.1 * n)))
    np.testing.assert_allclose(res["fvnsim10"], np.sort(means, axis=0)[:k].mean(axis=0), rtol=1e-14)
    np.testing.assert_allclose(res["vnsim"], res["fvnsim"].mean(), rtol=1e-14)


def _samples(mean, sd, count, rng):
    """count values with exactly this mean and sample deviation (ddof=1)."""
    if count == 1:
        return np.array([mean])
    z = rng.standard_normal(count)
    z -= z.mean()
    z /= np.std(z, ddof=1)
    return mean + sd * z


@pytest.mark.parametrize("n,patch_frames", [(4, 30), (16, 30), (7, 20)])
def test_fstdnsim_is_the_pooled_sample_deviation(vo, n, patch_frames):
    rng = np.random.default_rng(100 + n)
    fd = 960 / 48000 if patch_frames == 30 else 320 / 16000
    nb = 32 if patch_frames == 30 else 21
    means, sds, energy, start, end = _matches(rng, n, nb, fd, patch_frames)
    # patches starting at one time span exactly the same seconds, so they
    # have one frame count and the pooled mean is the mean of means (with
    # distinct starts, end - start rounds differently and the counts can
    # differ by one, as in the reference)
    start = np.zeros(n)
    end = start + patch_frames * fd
    mode = 0 if nb == 32 else 1
    res = vo.Oracle(mode=mode, use_lattice_model=False).build_result(means, sds, energy, start, end, fd)
    counts = np.ceil((end - start) / fd).astype(int)
    assert len(set(counts)) == 1
    for b in range(nb):
        pooled = np.concatenate([_samples(means[i, b], sds[i, b], counts[i], rng) for i in range(n)])
        assert res["fstdnsim"][b] == pytest.approx(np.std(pooled, ddof=1), rel=1e-10, abs=1e-13)


def test_mos_is_one_below_vnsim_015(vo):
    rng = np.random.default_rng(7)
    fd = 960 / 48000
    n, nb = 8, 32
    _, sds, energy, start, end = _matches(rng, n, nb, fd)
    orc = vo.Oracle(mode=0)
    low = orc.build_result(np.full((n, nb), 0.1499), sds, energy, start, end, fd)
    assert low["vnsim"] < 0.15 and low["moslqo"] == 1.0
    high = orc.build_result(np.full((n, nb), 0.9), sds, energy, start, end, fd)
    assert high["moslqo"] > 1.0


# ---------------------------------------------------------------- exponential mapping
A, B, X0 = -262.847869, 0.0154302525, -361.063949  # visqol.zig:86-88


def _speech_mos(vo, mean, unscaled):
    fd = 320 / 16000
    n, nb = 4, 21
    means = np.full((n, nb), mean)
    z = np.zeros((n, nb))
    start = np.arange(n) * 0.6
    orc = vo.Oracle(mode=1, use_lattice_model=False, use_unscaled_speech_mos_mapping=unscaled)
    return orc.build_result(means, z, z, start, start + 0.4, fd)["moslqo"]


def test_exponential_mapping_closed_form(vo):
    xs = np.array([0.55, 0.6, 0.7, 0.8, 0.9])
    u = np.array([_speech_mos(vo, x, True) for x in xs])
    assert np.all((u > 1.0) & (u < 5.0))
    # log(mos - a) is affine in the mean with slope b and intercept -b*x0
    slope, icept = np.polyfit(xs, np.log(u - A), 1)
    assert slope == pytest.approx(B, rel=1e-9)
    assert icept == pytest.approx(-B * X0, rel=1e-9)
    # the scale upstream keeps in a float: 1.245063 rounded to f32, applied once
    scale = float(np.float32(1.245063))
    assert scale != 1.245063
    for x, mu in zip(xs, u):
        s = _speech_mos(vo, x, False)
        if s < 5.0:
            assert s == mu * scale


def test_exponential_mapping_clamps(vo):
    assert _speech_mos(vo, 0.16, True) == 1.0   # a + exp(...) below 1
    assert _speech_mos(vo, 1.2, False) == 5.0    # scaled past 5
    assert _speech_mos(vo, 1.2, True) < 5.0      # unscaled, the same mean stays below
    assert _speech_mos(vo, 0.1, False) == 1.0    # vnsim < 0.15


# ---------------------------------------------------------------- TF-Lattice interpreter
def _eval_all(model: bytes):
    """(oracle, this library, numpy f64) predictions of a one-band model."""
    from akvisqol_amd import native
    from oracle import vo

    one = np.array([0.5])
    y = ctypes.c_double()
    rc = vo.lib().vo_lattice_predict(model, len(model), *(vo.dptr(vo.f64(one)) for _ in range(4)), 1, ctypes.byref(y))
    assert rc == 0
    got = np.zeros(1)
    p = native._dptr(one)
    assert native.lib.visqol_hip_lattice_predict(model, len(model), 1, 1, p, p, p, p, native._dptr(got)) == 0
    ref = tflite_np.lattice_mos(tflite_np.Model(model), [0.5], [0.5], [0.5], [0.5])
    return y.value, float(got[0]), ref


def _sub_model(k):
    g = tb.Graph()
    g.const("x", np.array([[1.0], [2.0]], np.float32))
    g.const("y", np.array([10.0, 20.0, 30.0], np.float32))
    g.tensor("d", tb.F32, [2, 3])
    g.op(41, ["x", "y"], ["d"], tb.Table([(0, "i8", 0)]), 28)  # SUB, SubOptions
    g.pick("d", 6, k)
    return g.bytes()


@pytest.mark.parametrize("k,want", list(enumerate([-9, -19, -29, -8, -18, -28])))
def test_broadcast_sub_kat(k, want):
    # tflite.zig:799-808: [2,1] {1,2} - [3] {10,20,30} -> [2,3]
    assert _eval_all(_sub_model(k)) == (want, want, want)


@pytest.mark.parametrize("k,want", [(0, 0.0), (1, 1.0)])
def test_broadcast_equal_kat(k, want):
    # tflite.zig:809-810: [2,1] {1,2} == scalar 2 -> {false, true}
    g = tb.Graph()
    g.const("x", np.array([[1.0], [2.0]], np.float32))
    g.const("two", np.array(2.0, np.float32).reshape(()))
    g.tensor("e", tb.BOOL, [2, 1])
    g.op(71, ["x", "two"], ["e"])  # EQUAL
    g.tensor("ef", tb.F32, [2, 1])
    g.op(53, ["e"], ["ef"])  # CAST bool -> f32
    g.pick("ef", 2, k)
    assert _eval_all(g.bytes()) == (want, want, want)


def _bmm_graph():
    g = tb.Graph()
    g.const("a", np.array([1.0, 2.0], np.float32).reshape(1, 2, 1))
    g.const("b", np.array([3.0, 4.0], np.float32).reshape(1, 1, 2))
    g.tensor("p", tb.F32, [1, 2, 2])
    g.op(126, ["a", "b"], ["p"], tb.Table([(0, "u8", 0), (1, "u8", 0)]))  # BATCH_MATMUL, adj_x = adj_y = false
    g.tensor("r", tb.F32, [1, 4, 1])
    g.op(22, ["p"], ["r"], tb.Table([(0, "vec_i32", [-1, 4, 1])]), 17)  # RESHAPE new_shape {-1,4,1}
    return g


@pytest.mark.parametrize("k,want", list(enumerate([3, 4, 6, 8])))
def test_batch_matmul_outer_product_kat(k, want):
    # tflite.zig:813-822: [1,2,1] {1,2} x [1,1,2] {3,4} -> [1,2,2] {3,4,6,8}
    g = _bmm_graph()
    g.pick("r", 4, k)
    assert _eval_all(g.bytes()) == (want, want, want)


@pytest.mark.parametrize("k,want", list(enumerate([1, 4, 1])))
def test_reshape_infers_minus_one_kat(k, want):
    # tflite.zig:821-822: reshape [-1,4,1] of the [1,2,2] product -> shape {1,4,1}
    g = _bmm_graph()
    g.tensor("s", tb.I32, [3])
    g.op(77, ["r"], ["s"])  # SHAPE
    g.tensor("sf", tb.F32, [3])
    g.op(53, ["s"], ["sf"])  # CAST i32 -> f32
    g.pick("sf", 3, k)
    assert _eval_all(g.bytes()) == (want, want, want)


@pytest.mark.parametrize("model", [bytes([1, 2]), bytes([8, 0, 0, 0, 0, 0, 0, 0])])
def test_truncated_models_rejected(vo, model):
    # tflite.zig:825-828: InvalidModel for truncated buffers
    from akvisqol_amd import native

    one = np.array([0.5])
    y = ctypes.c_double()
    assert vo.lib().vo_lattice_predict(model, len(model), *(vo.dptr(vo.f64(one)) for _ in range(4)), 1,
                                       ctypes.byref(y)) != 0
    got = np.zeros(1)
    p = native._dptr(one)
    assert native.lib.visqol_hip_lattice_predict(model, len(model), 1, 1, p, p, p, p, native._dptr(got)) == \
        native.E_MODEL


@pytest.fixture(scope="module")
def lattice():
    data = MODEL.read_bytes()
    return data, tflite_np.Model(data)


@pytest.fixture(scope="module")
def feature_vectors(vo):
    """200 vectors: features of scored speech pairs, perturbations of them
    (scaled toward both ends of the similarity range), and edge vectors."""
    from akvisqol_amd import synth

    orc = vo.Oracle(mode=1, use_lattice_model=True)
    base = []
    for s in range(10):
        r, d = synth.speech_pair(s, duration=3.0, snr_db=5.0 + 4.0 * s)
        res, _ = orc.measure(r, d, 16000)
        base.append([res[k] for k in ("fvnsim", "fvnsim10", "fstdnsim", "fvdegenergy")])
    rng = np.random.default_rng(2024)
    out = [b for b in base]
    while len(out) < 190:
        f0, f1, f2, f3 = base[len(out) % len(base)]
        sc = rng.uniform(0.5, 1.8)
        out.append([np.clip(f0 * sc * rng.uniform(0.9, 1.1, 21), 0.0, 1.0), f1 * sc * rng.uniform(0.9, 1.1, 21),
                    f2 * rng.uniform(0.3, 1.7, 21), f3 * rng.uniform(0.6, 1.4, 21)])
    ones, zeros = np.ones(21), np.zeros(21)
    out += [[ones, ones, zeros, np.full(21, 50.0)], [zeros, zeros, zeros, zeros], [ones, ones, ones, ones],
            [ones, zeros, zeros, np.full(21, 100.0)], [np.full(21, 0.5)] * 3 + [np.full(21, 20.0)],
            [ones * 0.95, ones * 0.9, ones * 0.05, ones * 30.0], [ones * 0.3, -ones * 0.2, ones * 0.5, ones * 10.0],
            [ones * 0.7, ones * 0.6, ones * 0.2, ones * 200.0], [ones * 1.2, ones * 1.1, zeros, ones * 40.0],
            [ones * 0.8, ones * 0.7, ones * 0.1, -ones * 5.0]]
    return out


def _oracle_lattice(vo, data, f):
    y = ctypes.c_double()
    assert vo.lib().vo_lattice_predict(data, len(data), *(vo.dptr(vo.f64(np.asarray(x, float))) for x in f), 21,
                                       ctypes.byref(y)) == 0
    return y.value


def test_lattice_model_against_numpy_f64(vo, lattice, feature_vectors):
    data, model = lattice
    diffs, spread = [], []
    for f in feature_vectors:
        want = tflite_np.lattice_mos(model, *f)
        got = _oracle_lattice(vo, data, f)
        diffs.append(abs(got - want))
        spread.append(got)
    print(f"lattice vs numpy f64: max |dMOS| {max(diffs):.3g} over {len(diffs)} vectors, "
          f"MOS {min(spread):.3f}..{max(spread):.3f}")
    assert len(diffs) == 200
    assert max(spread) - min(spread) > 2.0  # the vectors exercise the model's range
    assert max(diffs) <= 1e-4


def test_lattice_model_against_numpy_f32_sequential(vo, lattice, feature_vectors, monkeypatch):
    """The same model in f32 with the 4096-term fully connected dots summed
    in sequential order with one rounding per step (TFLite's builtin kernel
    order, tflite.zig:665-671): the remaining difference is the order of the
    short dots and numpy's own roundings."""
    data, model = lattice

    def fc_sequential(flat, w):
        if flat.dtype != np.float32 or w.shape[1] < 1000:
            return flat @ w.T
        acc = np.zeros((flat.shape[0], w.shape[0]), np.float32)
        xs, ws = flat.astype(np.float64), w.astype(np.float64)
        for k in range(w.shape[1]):
            acc = (acc.astype(np.float64) + xs[:, k:k + 1] * ws[None, :, k]).astype(np.float32)
        return acc

    monkeypatch.setattr(tflite_np, "_fc", fc_sequential)
    worst = 0.0
    for f in feature_vectors[::40]:
        want = tflite_np.lattice_mos(model, *f, dtype=np.float32)
        worst = max(worst, abs(_oracle_lattice(vo, data, f) - want))
    print(f"lattice vs numpy f32 (sequential dots): max |dMOS| {worst:.3g}")
    assert worst <= 2e-5


def test_lattice_piecewise_linear_along_features(vo, lattice, feature_vectors):
    """Calibrators are piecewise linear and the lattice multilinear, so along
    one feature the score is piecewise linear: second differences on a fine
    grid vanish (to the f32 rounding) except next to a few kinks."""
    data, _ = lattice
    rng = np.random.default_rng(9)
    for trial in range(4):
        f = [np.array(x, float) for x in feature_vectors[trial * 7]]
        fi, b = int(rng.integers(0, 4)), int(rng.integers(0, 21))
        lo = f[fi][b] - (0.2 if fi < 3 else 10.0)
        ts = np.linspace(lo, lo + (0.4 if fi < 3 else 20.0), 201)
        ys = []
        for t in ts:
            g = [x.copy() for x in f]
            g[fi][b] = t
            ys.append(_oracle_lattice(vo, data, g))
        d2 = np.abs(np.diff(np.array(ys), 2))
        kinks = int(np.sum(d2 > 1e-5))
        print(f"feature {fi} band {b}: {kinks} of {d2.size} second differences above 1e-5")
        assert kinks <= 40
