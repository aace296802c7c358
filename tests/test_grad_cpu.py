"""Host pieces of the gradient against central differences of the oracle.

measureBatchGrad (backend.zig:101-110) holds every discrete choice fixed and
differentiates the rest.  The two host stages checked here are the NSIM
patch statistics (nsim.zig:427-510, the oracle's vo_patch_similarity) and
the dB preparation (spectrogram.zig:14-80, vo_prepare_for_comparison): the
library's backward of each is compared cell by cell with central differences
of the oracle's forward, to the bar the reference's own backward tests use
(pipeline.zig:1704-1791: worst difference below 1e-6 of the largest slope).
The 7x7 patch beside a silent reference is the reference's own fixture
(pipeline.zig:1730-1746).
"""
import ctypes

import numpy as np
import pytest

from oracle import vo
from oracle.vo import dptr, f64

L = vo.lib()


@pytest.fixture(scope="module")
def native():
    from akvisqol_amd import native as n

    return n


def sim_loss(ref, deg, gm, gs, ge):
    nr, nc = ref.shape
    means, sds, en, mm = np.zeros(nr), np.zeros(nr), np.zeros(nr), ctypes.c_double()
    L.vo_patch_similarity(dptr(f64(ref.ravel())), dptr(f64(deg.ravel())), nr, nc, dptr(means), dptr(sds), dptr(en),
                          ctypes.byref(mm))
    return float(gm @ means + gs @ sds + ge @ en)


def central(f, x, k, h):
    y = x.copy()
    y.flat[k] = x.flat[k] + h
    up = f(y)
    y.flat[k] = x.flat[k] - h
    return (up - f(y)) / (2 * h)


def check_sim(native, ref, deg, gm, gs, ge, h=1e-6, cells=None):
    gr, gd = native.patch_similarity_grad(ref, deg, gm, gs, ge)
    cells = range(ref.size) if cells is None else cells
    worst, largest = 0.0, 0.0
    for k in cells:
        for side, got in ((0, gr), (1, gd)):
            if side == 0:
                cd = central(lambda y: sim_loss(y, deg, gm, gs, ge), ref, k, h)
            else:
                cd = central(lambda y: sim_loss(ref, y, gm, gs, ge), deg, k, h)
            worst = max(worst, abs(cd - got.flat[k]))
            largest = max(largest, abs(cd))
    assert largest > 0
    assert worst < 1e-6 * largest, (worst, largest)


@pytest.mark.parametrize("shape,seed", [((4, 5), 0), ((32, 20), 1), ((7, 1), 2), ((1, 9), 3)])
def test_patch_similarity_grad_matches_central_differences(native, shape, seed):
    rng = np.random.default_rng(seed)
    ref = rng.uniform(0, 60, shape)
    deg = ref + rng.normal(0, 6, shape)
    gm, gs, ge = rng.normal(size=(3, shape[0]))
    cells = None if ref.size <= 60 else rng.choice(ref.size, 60, replace=False)
    check_sim(native, ref, deg, gm, gs, ge, cells=cells)


def test_patch_similarity_grad_optional_slopes(native):
    rng = np.random.default_rng(4)
    ref, deg = rng.uniform(0, 50, (2, 6, 8))
    z = np.zeros(6)
    gm = rng.normal(size=6)
    # NULL slopes count as zero
    a = native.patch_similarity_grad(ref, deg, gm, None, None)
    b = native.patch_similarity_grad(ref, deg, gm, z, z)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # the energy term is the band mean of the degraded cells
    gr, gd = native.patch_similarity_grad(ref, deg, None, None, np.arange(6.0))
    assert not gr.any()
    assert np.array_equal(gd, np.repeat(np.arange(6.0)[:, None] / 8, 8, axis=1))


# The reference's fixture: ref (7 bands x 7 frames) then deg, contrabassoon
# around band 22, frame 364 (pipeline.zig:1730-1746).
SILENT_REF_FIXTURE = np.array([
    2.96856108177539820, 3.67188762495687370, 4.75163221275552400, 5.38719267570785800, 5.44681217810485700,
    4.48020995284284100, 3.84323495673001500,
    2.16903693348019300, 1.70620203480871170, 2.54008730917163900, 3.17797227061273450, 3.44683253742547900,
    2.99890096590598180, 1.73913841381336450,
    0.55345063100589440, 0.18153692307269154, 0.43106553359726973, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0,
    0.12846954386578346, 0.28734778774285274, 0.22767008222064078, 0, 0, 0, 0,
    0.30574495140761115, 0.25902979977644947, 0, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0,
    3.42049712657014740, 4.17133489097896200, 5.54906142634718700, 6.21551341983749950, 6.44507205162854300,
    5.75721069191466500, 4.17077133206959200,
    2.21469770025635900, 1.43181593933577030, 2.42618496219024140, 2.99547772363823840, 3.18128146171978440,
    3.34061046405459900, 2.11063431705386500,
    1.76639321404701380, 0.91498238961293050, 1.18047283437498150, 1.25381597614732240, 0.37105234702722356,
    0.02268735575298564, 0,
    0, 0, 0, 0.00432372571879114, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0,
]).reshape(2, 7, 7)


def test_patch_similarity_grad_beside_silent_reference(native):
    # loss = Σ band means, slopes w.r.t. the degraded cells (pipeline.zig:1758-1790)
    ref, deg = SILENT_REF_FIXTURE
    ones, z = np.ones(7), np.zeros(7)
    _, gd = native.patch_similarity_grad(ref, deg, ones, z, z)
    worst, largest = 0.0, 0.0
    for k in range(deg.size):
        cd = central(lambda y: sim_loss(ref, y, ones, z, z), deg, k, 1e-8)
        worst = max(worst, abs(cd - gd.flat[k]))
        largest = max(largest, abs(cd))
    assert worst < 1e-6 * largest, (worst, largest)


def prep_loss(raw_ref, raw_deg, g1, g2):
    r, d = f64(raw_ref.ravel()).copy(), f64(raw_deg.ravel()).copy()
    nb, nfr = raw_ref.shape
    L.vo_prepare_for_comparison(dptr(r), dptr(d), nb, nfr, raw_deg.shape[1])
    return float(g1.ravel() @ r + g2.ravel() @ d)


def check_prep(native, raw_ref, raw_deg, seed):
    rng = np.random.default_rng(seed)
    g1, g2 = rng.normal(size=raw_ref.shape), rng.normal(size=raw_deg.shape)
    gr, gd = native.prepare_grad(raw_ref, raw_deg, g1, g2)
    worst, largest = 0.0, 0.0
    for side, x, got in ((0, raw_ref, gr), (1, raw_deg, gd)):
        for k in range(x.size):
            h = 1e-7 * abs(x.flat[k]) if x.flat[k] != 0 else 0.0
            if h == 0.0:
                assert got.flat[k] == 0.0  # an exact zero is the epsilon cell: it passes nothing
                continue
            f = (lambda y: prep_loss(y, raw_deg, g1, g2)) if side == 0 else (lambda y: prep_loss(raw_ref, y, g1, g2))
            cd = central(f, x, k, h)
            # slopes here scale as 1/x: compare relative to the cell's own size
            worst = max(worst, abs(cd - got.flat[k]) * abs(x.flat[k]))
            largest = max(largest, abs(cd) * abs(x.flat[k]))
    assert largest > 0
    assert worst < 1e-6 * largest, (worst, largest)
    return gr, gd


def test_prepare_grad_no_floors(native):
    # every cell within 45 dB of its frame's maximum: the joint minimum is a live cell
    rng = np.random.default_rng(10)
    raw_ref = np.exp(rng.uniform(-4, 0, (6, 9)))
    raw_deg = np.exp(rng.uniform(-4, 0, (6, 11)))
    gr, gd = check_prep(native, raw_ref, raw_deg, 11)
    # scaling every raw cell by one factor moves every dB value, floor and
    # the joint minimum alike, so the prepared values do not move: the
    # slope along x itself is zero (the joint minimum's cell carries minus
    # the sum of the rest)
    a, b = native.prepare_grad(raw_ref, raw_deg, np.ones_like(raw_ref), np.ones_like(raw_deg))
    assert abs((a * raw_ref).sum() + (b * raw_deg).sum()) < 1e-9


def test_prepare_grad_with_floors_and_zeros(native):
    # quiet cells below the per-frame floor, cells under the absolute floor
    # (< 10^-4.5) and exact zeros, in frames shared by both sides and in the
    # degraded side's extra frames
    rng = np.random.default_rng(12)
    raw_ref = np.exp(rng.uniform(-3, 0, (8, 10)))
    raw_deg = np.exp(rng.uniform(-3, 0, (8, 13)))
    raw_ref[2, 3] = 1e-4     # -40 dB: under a frame floor near 10 dB - 45
    raw_ref[0, 3] = 30.0     # the frame's maximum, ~14.8 dB
    raw_deg[5, 7] = 1e-6     # -60 dB: absolute floor
    raw_deg[1, 11] = 2e-6    # extra frame, absolute floor
    raw_ref[4, 0] = 0.0
    raw_deg[3, 12] = 0.0
    check_prep(native, raw_ref, raw_deg, 13)
