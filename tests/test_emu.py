"""The gfx950 kernels' logic, run through their CPU emulation (libvgc_emu.so:
the same clip_cell.h / common.h / pipeline.h sources with the wave layer
emulated lane by lane), against the oracle.  The GPU tests then check that the
device reproduces the emulation bit for bit."""
import ctypes
import os

import numpy as np
import pytest
from scipy.spatial import cKDTree

import geogram_amd as ga
import oracle

REL = 1e-10


def compare(r, o, n, rel=REL, x=None, w=None, box=None):
    """Volumes, centroids, facet areas and neighbour sets against the oracle.

    o is the oracle in geogram's absolute frame (planes formed from absolute
    coordinates, as copy_Laguerre_cell_facet_from_Delaunay does); its plane
    offsets round to ~1.3e-15 absolute, i.e. up to ~4e-15 / vol^(2/3)
    relative in volume on tiny cells, so it is held to that bound.  When x is
    given, the oracle is rerun in the seed-centred frame and held to rel."""
    vol_o = np.maximum(o["vol"], 1e-300)
    cond = 4e-15 / vol_o ** (2 / 3)
    assert np.all(np.abs(r.vol - o["vol"]) <= vol_o * (rel + cond)), "volume vs absolute frame"
    if x is not None:
        kw = {} if box is None else {"box": box}
        oc = oracle.cells(x, w=w, k0=20, centred=True, **kw)
        vrel = np.abs(r.vol - oc["vol"]) / np.maximum(oc["vol"], 1e-300)
        assert vrel.max() < rel, f"volume rel err {vrel.max()} at {vrel.argmax()}"
        o = oc
        cond = 0.0 * cond
    nz = o["vol"] > 0
    cen = r.mg[nz] / r.vol[nz, None]
    ocen = o["mg"][nz] / o["vol"][nz, None]
    assert (np.abs(cen - ocen) / np.cbrt(o["vol"][nz])[:, None]).max() < rel + cond[nz].max(initial=0)
    bad = []
    for i in range(n):
        a = r.neighbors(i)
        b = o["nbr"][i, : o["cnt"][i]]
        if set(a.tolist()) != set(b.tolist()):
            bad.append(i)
            continue
        da = dict(zip(a.tolist(), r.facet_areas(i).tolist()))
        for j, ar in zip(b.tolist(), o["farea"][i, : o["cnt"][i]].tolist()):
            # facet areas: 1e-10 relative, or 2e-11 of the cell's face scale
            # (sliver facets between ill-conditioned corners)
            assert abs(da[j] - ar) <= (rel + cond[i]) * ar + 2e-11 * o["vol"][i] ** (2 / 3)
    assert not bad, f"{len(bad)} neighbour sets differ, first {bad[:5]}"


@pytest.mark.parametrize("kind,n,k0", [(0, 2000, 32), (0, 4000, 64), (1, 2500, 48), (2, 2000, 16),
                                       (0, 4000, 0), (1, 5000, 0), (2, 5000, 0)])
def test_cells_vs_oracle(emu_ctx, kind, n, k0):
    x = ga.sample_points(kind, 100 + kind, n)
    r = emu_ctx.cells(x, k0=k0)
    assert r.stats["nb_failed"] == 0 and r.stats["nb_uncertain"] == 0
    o = oracle.cells(x, k0=20)
    compare(r, o, n, x=x)
    assert abs(r.vol.sum() - 1.0) < 1e-12


def test_laguerre_vs_oracle(emu_ctx):
    rng = np.random.default_rng(3)
    n = 2500
    x = rng.random((n, 3))
    w = rng.random(n) * 4e-4
    for k0 in (32, 0):
        r = emu_ctx.cells(x, weights=w, k0=k0)
        o = oracle.cells(x, w=w, k0=20)
        compare(r, o, n, x=x, w=w)
    assert abs(r.vol.sum() - 1.0) < 1e-12


def test_laguerre_empty_cells(emu_ctx):
    rng = np.random.default_rng(4)
    n = 800
    x = rng.random((n, 3))
    w = np.zeros(n)
    w[:10] = 0.05   # ten heavy seeds swallow their neighbours
    r = emu_ctx.cells(x, weights=w, k0=32)
    o = oracle.cells(x, w=w, k0=20, stride=512)
    assert r.stats["nb_empty"] == o["stats"]["nb_empty"] > 0
    nz = o["vol"] > 0
    assert np.array_equal(r.vol == 0, ~nz)
    assert (np.abs(r.vol[nz] - o["vol"][nz]) / o["vol"][nz]).max() < REL
    assert abs(r.vol.sum() - 1.0) < 1e-12


def test_batch_of_instances(emu_ctx):
    sizes = [300, 1700, 50, 1000]
    xs = [ga.sample_points(i % 3, 40 + i, s) for i, s in enumerate(sizes)]
    boxes = np.array([[0, 0, 0, 1, 1, 1]] * len(sizes), dtype=float)
    boxes[2] = [-1, -1, -1, 2, 2, 2]
    x = np.vstack(xs)
    r = emu_ctx.cells(x, k0=32, inst_n=sizes, boxes=boxes)
    off = np.r_[0, np.cumsum(sizes)]
    for i, s in enumerate(sizes):
        o = oracle.cells(xs[i], box=boxes[i], k0=20)
        sl = slice(off[i], off[i + 1])
        assert (np.abs(r.vol[sl] - o["vol"]) / o["vol"]).max() < REL
        assert abs(r.vol[sl].sum() - np.prod(boxes[i, 3:] - boxes[i, :3])) < 1e-11
        for c in range(s):
            got = set((r.neighbors(off[i] + c) - off[i]).tolist())
            assert got == set(o["nbr"][c, : o["cnt"][c]].tolist())


def test_small_and_coincident(emu_ctx):
    # n smaller than k0, and a coincident pair (delaunay_nn.cpp:125-133)
    rng = np.random.default_rng(5)
    for n in (1, 2, 5, 17):
        x = rng.random((n, 3))
        r = emu_ctx.cells(x, k0=32)
        o = oracle.cells(x, k0=20)
        assert np.allclose(r.vol, o["vol"], rtol=1e-12)
        assert abs(r.vol.sum() - 1) < 1e-13
    x = rng.random((300, 3))
    x[200] = x[17]
    r = emu_ctx.cells(x, k0=32)
    o = oracle.cells(x, k0=20)
    assert r.stats["nb_disconnected"] == 1
    assert np.allclose(r.vol, o["vol"], rtol=1e-12)
    assert oracle.neighbour_sets(o) == [frozenset(r.neighbors(i).tolist()) for i in range(300)]


def test_lattice_ties_exact_pass(emu_ctx):
    # exact lattices: every corner is shared by 8 cells, so many signs are ties
    # (det4 = 0); the cells are re-clipped with exact signs and the tie rule
    # (ties conflict, convex_cell.cpp:905-911)
    for m in (5, 8):
        g = (np.arange(m) + 0.5) / m
        x = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T.copy()
        r = emu_ctx.cells(x, k0=32)
        o = oracle.cells(x, k0=20)
        assert r.stats["nb_pass_e"] > 0
        assert r.stats["nb_uncertain"] == 0 and r.stats["nb_failed"] == 0
        assert np.allclose(r.vol, 1.0 / m**3, rtol=1e-14)
        if m == 8:
            # dyadic coordinates: seed-relative and absolute planes are both
            # exact, so the tie resolution, hence the neighbour sets (zero-area
            # facets included), match the oracle's exactly
            assert oracle.neighbour_sets(o) == [frozenset(r.neighbors(i).tolist())
                                                for i in range(x.shape[0])]


def test_det4_sign_exact_integer():
    # the device's fixed-point evaluator against rational arithmetic
    from tests.test_oracle import det4_exact
    lib = ctypes.CDLL(ga._lib.EMU_PATH)
    lib.vgc_emu_det4_sign.argtypes = [ctypes.c_void_p]
    rng = np.random.default_rng(21)
    for t in range(400):
        R = rng.standard_normal((4, 4)) * (10.0 ** rng.integers(-8, 3, (4, 4)))
        if t % 2:
            # exactly singular with dyadic combination
            R[3] = 2 * R[0] - 0.5 * R[1] + R[2]
            R[3] = R[3]  # rounded combination: the sign is decided exactly
        R = np.ascontiguousarray(R)
        assert lib.vgc_emu_det4_sign(R.ctypes.data) == det4_exact(R)
    # exact zeros
    for _ in range(100):
        R = np.round(rng.standard_normal((3, 4)) * 2**10) / 2**10
        M = np.ascontiguousarray(np.vstack([R, R[0] + R[1] - R[2]]))
        assert det4_exact(M) == 0 and lib.vgc_emu_det4_sign(M.ctypes.data) == 0


def test_lloyd_vs_oracle(emu_ctx):
    x = ga.sample_points(0, 77, 3000)
    xg, st = emu_ctx.lloyd(x, 4, k0=32)
    xo, _ = oracle.lloyd(x, 4)
    assert np.abs(xg - xo).max() < 1e-12


def test_knn_vs_kdtree(emu_ctx):
    x = ga.sample_points(1, 9, 5000)
    for k in (16, 64):
        idx, d2 = emu_ctx.knn(x, k)
        dd, ii = cKDTree(x).query(x, k + 1)
        # FP32 ranking: sets agree except at FP64 near-ties of the k-th distance
        for i in range(x.shape[0]):
            if set(idx[i]) != set(ii[i, 1:]):
                kth, nxt = cKDTree(x).query(x[i], k + 2)[0][k:k + 2]
                assert (nxt - kth) <= 1e-5 * kth
        assert np.allclose(d2, dd[:, 1:] ** 2, rtol=2e-5)


def test_random_float64_is_mt19937_64():
    # C++ [rand.predef]: the 10000th output of a default-constructed
    # mt19937_64 (seed 5489) is 9981545732273789042
    out = ga.random_float64(5489, 10000)
    assert out[9999] == np.float64(9981545732273789042) / 2.0**64
    assert np.all((out >= 0) & (out < 1))
    assert np.array_equal(ga.random_float64(1, 5), ga.random_float64(1, 5))


def _transport_weights(emu_ctx, seed, n, newton):
    """Laguerre weights of a damped Newton iterate of the semi-discrete
    transport onto a 16-Gaussian mixture with log-normal masses: a weight
    field of wide spread whose cells sit away from their seeds."""
    from geogram_amd import warmstart
    x, nu = warmstart.make_instance(seed, n)
    w, _ = emu_ctx.ot_solve(x, nu, tol=1e-9, max_newton=newton)
    return x, w


def test_transport_weights_power_filter(emu_ctx):
    x, w = _transport_weights(emu_ctx, 77, 2500, 3)
    assert np.ptp(w) > 1e-2
    o = oracle.cells(x, w=w, centred=True, stride=512)
    nz = o["vol"] > 0

    def check(r):
        assert r.stats["nb_failed"] == 0
        assert np.abs(r.vol - o["vol"]).max() <= 1e-10 * o["vol"].max()
        assert np.all(np.abs(r.vol[nz] - o["vol"][nz]) <= 1e-10 * o["vol"][nz])
        bad = [i for i in range(len(x))
               if set(r.neighbors(i).tolist()) != set(o["nbr"][i, :o["cnt"][i]].tolist())]
        assert not bad, f"{len(bad)} neighbour sets differ"

    # in-kernel gathering: power-filtered depth-first rings
    check(emu_ctx.cells(x, weights=w))
    # first pass from 16-NN lists: most cells go to the list tiers, whose
    # radius queries carry the power filter of the partial cell
    r = emu_ctx.cells(x, weights=w, k0=16)
    s = r.stats
    assert s["nb_pass_b"] > len(x) // 4
    # far fewer entries than the seeds within the max-weight radius
    assert s["radius_entries"] < 0.2 * len(x) * s["nb_pass_b"]
    check(r)
    # list tiers cut into small chunks: same cells bit for bit
    emu_ctx.lib.cdll.vgc_emu_set_list_chunk(emu_ctx.ctx, ctypes.c_longlong(300))
    try:
        r2 = emu_ctx.cells(x, weights=w, k0=16)
    finally:
        emu_ctx.lib.cdll.vgc_emu_set_list_chunk(emu_ctx.ctx, ctypes.c_longlong(1 << 27))
    assert np.array_equal(r2.vol, r.vol) and np.array_equal(r2.mg, r.mg)
    assert np.array_equal(r2.nbr, r.nbr) and np.array_equal(r2.farea, r.farea)


def test_cells_beyond_256_corners(emu_ctx):
    # a heavy seed in a uniform cloud: a ball-like cell with 185 facets and
    # ~370 corners runs in the 512-corner tier
    x = ga.sample_points(0, 41, 20000)
    x[0] = [0.5, 0.5, 0.5]
    w = np.zeros(len(x))
    w[0] = 0.012
    r = emu_ctx.cells(x, weights=w)
    assert r.stats["nb_failed"] == 0 and len(r.neighbors(0)) == 185
    o = oracle.cells(x, w=w, centred=True, stride=512)
    assert np.all(np.abs(r.vol - o["vol"]) <= 1e-10 * o["vol"])
    assert set(r.neighbors(0).tolist()) == set(o["nbr"][0, :o["cnt"][0]].tolist())
    da = dict(zip(r.neighbors(0).tolist(), r.facet_areas(0).tolist()))
    for j, a in zip(o["nbr"][0, :o["cnt"][0]].tolist(), o["farea"][0, :o["cnt"][0]].tolist()):
        assert abs(da[j] - a) <= 1e-10 * a + 1e-14
    # 295 and 664 facets: over the 256 plane slots, computed in octant
    # pieces (split again where a piece is still over the limit)
    for w0, nf in ((0.02, 295), (0.05, 664)):
        w[0] = w0
        r = emu_ctx.cells(x, weights=w)
        assert r.status == ga.VGC_OK and r.stats["nb_failed"] == 0 and r.stats["nb_pieces"] == 1
        o = oracle.cells(x, w=w, centred=True, stride=1024)
        assert o["cnt"][0] == nf and len(r.neighbors(0)) == nf
        vo = np.maximum(o["vol"], 1e-300)
        assert np.all(np.abs(r.vol - o["vol"]) <= 1e-10 * vo)
        assert np.abs(r.mg[0] / r.vol[0] - o["mg"][0] / o["vol"][0]).max() < 1e-12
        bad = [i for i in range(len(x))
               if set(r.neighbors(i).tolist()) != set(o["nbr"][i, :o["cnt"][i]].tolist())]
        assert not bad
        da = dict(zip(r.neighbors(0).tolist(), r.facet_areas(0).tolist()))
        for j, a in zip(o["nbr"][0, :o["cnt"][0]].tolist(), o["farea"][0, :o["cnt"][0]].tolist()):
            assert abs(da[j] - a) <= 1e-9 * a + 1e-13


def test_lloyd_batch_equals_separate_runs(emu_ctx):
    sizes = [500, 3000, 64, 1500]
    x = np.vstack([ga.sample_points(i % 3, 60 + i, s) for i, s in enumerate(sizes)])
    boxes = np.array([[0, 0, 0, 1, 1, 1]] * len(sizes), dtype=float)
    boxes[1] = [-1.0, 0.0, 0.0, 2.0, 1.0, 1.5]
    x[500:3500] = boxes[1, :3] + x[500:3500] * (boxes[1, 3:] - boxes[1, :3])
    xb, st = emu_ctx.lloyd(x, 3, inst_n=sizes, boxes=boxes)
    o = 0
    for s, b in zip(sizes, boxes):
        xi, _ = emu_ctx.lloyd(x[o:o + s], 3, box=b)
        assert np.array_equal(xi, xb[o:o + s])
        o += s


def _clusters_and_outliers(seed=3, sig=0.005):
    """Two tight clusters of 1495 seeds and 10 uniform outliers: the outliers'
    cells reach far past the clusters, whose seeds fill their filtered rings."""
    rng = np.random.default_rng(seed)
    c = np.array([[0.3, 0.3, 0.3], [0.7, 0.6, 0.5]])
    x = np.vstack([c[0] + sig * rng.standard_normal((1495, 3)),
                   c[1] + sig * rng.standard_normal((1495, 3)), rng.random((10, 3))])
    return np.clip(x, 1e-3, 1 - 1e-3)


def test_second_gather_with_corner_test(emu_ctx):
    """Voronoi cells that use up the first pass's filtered rings are gathered
    again with the corner test (seeds that cannot cut any corner dropped)."""
    x = _clusters_and_outliers()
    n = x.shape[0]
    r = emu_ctx.cells(x)
    assert r.stats["nb_regather"] > 0
    assert r.stats["nb_failed"] == 0
    o = oracle.cells(x, k0=20)
    compare(r, o, n, x=x)
    assert abs(r.vol.sum() - 1.0) < 1e-12


@pytest.mark.parametrize("kind", [0, 2])
def test_cvt_energy_and_gradient(emu_ctx, kind):
    """vgc_cvt_func_grad: energy = sum of the oracle's per-cell second moments
    (a different tetrahedral decomposition: apex at the seed), gradient =
    2 (m p - mg); and the gradient is the derivative of the energy (central
    differences)."""
    n = 1200
    x = ga.sample_points(kind, 60 + kind, n)
    f, g = emu_ctx.cvt_func_grad(x)
    o = oracle.cells(x, e2=True, centred=True)
    assert abs(f - o["e2"].sum()) <= 1e-12 * o["e2"].sum()
    go = 2.0 * (o["vol"][:, None] * x - o["mg"])
    assert np.abs(g - go).max() <= 1e-10 * np.abs(go).max()
    rng = np.random.default_rng(kind)
    for i in rng.choice(n, 4, replace=False):
        for c in range(3):
            h = 1e-6
            xp, xm = x.copy(), x.copy()
            xp[i, c] += h
            xm[i, c] -= h
            fd = (emu_ctx.cvt_func_grad(xp)[0] - emu_ctx.cvt_func_grad(xm)[0]) / (2 * h)
            assert abs(fd - g[i, c]) <= 1e-6 * np.abs(g).max()


# ------------------------------------ first pass with FP32 corner state

def _same_cells(a, b):
    """The same bits in every output (the CSR up to its end)."""
    for f in ("vol", "mg", "nbr_ptr"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    m = int(a.nbr_ptr[-1])
    assert np.array_equal(a.nbr[:m], b.nbr[:m]), "nbr"
    assert np.array_equal(a.farea[:m], b.farea[:m]), "farea"


@pytest.mark.parametrize("case", ["uniform", "mixture", "laguerre", "periodic", "coincident",
                                  "batch"])
def test_fp32_first_pass_equals_fp64_bitwise(emu_ctx, case):
    # FP32 corners take the FP64 pass's decisions (signs certified against
    # the FP64 planes) and rebuild the FP64 corners for the integration, so
    # the outputs are the same bits; uncertain cells are redone in FP64
    rng = np.random.default_rng(11)
    kw = {}
    w = None
    if case == "mixture":
        x = ga.sample_points(1, 12, 6000)
    elif case == "laguerre":
        x = rng.random((6000, 3))
        w = rng.random(6000) * 2e-4
    elif case == "periodic":
        x = rng.random((3000, 3))
        kw["periodic"] = True
    elif case == "coincident":
        x = rng.random((3000, 3))
        x[2000:2010] = x[100:110]
    elif case == "batch":
        x = rng.random((5000, 3))
        kw["inst_n"] = [1000, 2500, 1500]
        kw["boxes"] = np.tile([0.0, 0, 0, 1, 1, 1], (3, 1))
    else:
        x = rng.random((6000, 3))
    a = emu_ctx.cells(x, w, **kw)
    b = emu_ctx.cells(x, w, fp64_first=True, **kw)
    _same_cells(a, b)
    assert a.stats["nb_failed"] == b.stats["nb_failed"] == 0
    assert b.stats["nb_fp64_redo"] == 0
    assert a.stats["nb_fp64_redo"] < 0.02 * len(x)


def test_fp32_first_pass_lattice_all_redone(emu_ctx):
    # on an exact lattice every cell has ties: FP32 certifies none of them,
    # the FP64 pass redoes every cell and sends them to the exact pass
    g = (np.arange(6) + 0.5) / 6
    x = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T.copy()
    a = emu_ctx.cells(x)
    b = emu_ctx.cells(x, fp64_first=True)
    _same_cells(a, b)
    assert a.stats["nb_fp64_redo"] == len(x)
    assert a.stats["nb_pass_e"] == b.stats["nb_pass_e"] == len(x)
