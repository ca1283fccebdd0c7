"""GPU parity of the Voronoi / Laguerre cell path against the oracle.

The oracle (oracle/) restates geogram's VBW::ConvexCell + security-radius loop
with exact predicates; volumes / centroids / facet areas must agree within
1e-10 relative and neighbour sets exactly."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

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


@pytest.mark.parametrize("kind,n,k0", [(0, 3000, 32), (1, 5000, 64), (0, 20000, 48), (1, 20000, 0),
                                       (2, 20000, 0)])
def test_voronoi_vs_oracle(gpu_ctx, kind, n, k0):
    import geogram_amd as ga

    x = ga.sample_points(kind, 11 + kind, n)
    r = gpu_ctx.cells(x, k0=k0)
    print(r.stats)
    o = oracle.cells(x, k0=20)
    compare(r, o, n, x=x)
    assert abs(r.vol.sum() - 1.0) < 1e-12


@pytest.mark.parametrize("kind,n", [(1, 3000), (2, 2000)])
def test_gpu_matches_emulation_bitwise(gpu_ctx, emu_ctx, kind, n):
    # kind 2 (16 tight clusters, few seeds): large cells beside the clusters,
    # many filtered rings per cell
    import geogram_amd as ga

    x = ga.sample_points(kind, 5, n)
    g = gpu_ctx.cells(x, k0=0)
    e = emu_ctx.cells(x, k0=0)
    assert np.array_equal(g.vol, e.vol)
    assert np.array_equal(g.mg, e.mg)
    assert np.array_equal(g.nbr_ptr, e.nbr_ptr)
    assert np.array_equal(g.nbr, e.nbr)
    assert np.array_equal(g.farea, e.farea)


def _same(g, e):
    assert np.array_equal(g.vol, e.vol)
    assert np.array_equal(g.mg, e.mg)
    assert np.array_equal(g.nbr_ptr, e.nbr_ptr)
    assert np.array_equal(g.nbr, e.nbr)
    assert np.array_equal(g.farea, e.farea)


def test_gpu_matches_emulation_laguerre_and_lattice(gpu_ctx, emu_ctx):
    rng = np.random.default_rng(6)
    x = rng.random((3000, 3))
    w = rng.random(3000) * 4e-4
    w[:5] = 0.05  # huge cells: 128 / 256 corner tiers and empty cells
    _same(gpu_ctx.cells(x, weights=w, k0=32), emu_ctx.cells(x, weights=w, k0=32))
    _same(gpu_ctx.cells(x, weights=w, k0=0), emu_ctx.cells(x, weights=w, k0=0))
    m = 8
    g = (np.arange(m) + 0.5) / m
    lat = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T.copy()
    rg = gpu_ctx.cells(lat, k0=0)
    assert rg.stats["nb_pass_e"] > 0 and rg.stats["nb_uncertain"] == 0
    _same(rg, emu_ctx.cells(lat, k0=0))


def test_gpu_matches_emulation_transport_weights(gpu_ctx, emu_ctx):
    # weights of a Newton iterate of transport onto a Gaussian mixture: the
    # power-filtered rings and list tiers, bit for bit
    from geogram_amd import warmstart
    x, nu = warmstart.make_instance(77, 2500)
    w, _ = emu_ctx.ot_solve(x, nu, tol=1e-9, max_newton=3)
    rg = gpu_ctx.cells(x, weights=w)
    assert rg.stats["nb_failed"] == 0
    _same(rg, emu_ctx.cells(x, weights=w))
    rg = gpu_ctx.cells(x, weights=w, k0=16)
    assert rg.stats["nb_pass_b"] > 0 and rg.stats["nb_failed"] == 0
    _same(rg, emu_ctx.cells(x, weights=w, k0=16))


def test_gpu_matches_emulation_512_corner_tier(gpu_ctx, emu_ctx):
    import geogram_amd as ga

    x = ga.sample_points(0, 41, 20000)
    x[0] = [0.5, 0.5, 0.5]
    w = np.zeros(len(x))
    w[0] = 0.012  # 185 facets: beyond 256 corners
    rg = gpu_ctx.cells(x, weights=w)
    assert rg.stats["nb_failed"] == 0 and len(rg.neighbors(0)) == 185
    _same(rg, emu_ctx.cells(x, weights=w))


def test_gpu_batch_and_lloyd(gpu_ctx, emu_ctx):
    import geogram_amd as ga

    sizes = [500, 3000, 64, 1500]
    x = np.vstack([ga.sample_points(i % 3, 60 + i, s) for i, s in enumerate(sizes)])
    boxes = np.array([[0, 0, 0, 1, 1, 1]] * len(sizes), dtype=float)
    _same(gpu_ctx.cells(x, k0=32, inst_n=sizes, boxes=boxes),
          emu_ctx.cells(x, k0=32, inst_n=sizes, boxes=boxes))
    _same(gpu_ctx.cells(x, k0=0, inst_n=sizes, boxes=boxes),
          emu_ctx.cells(x, k0=0, inst_n=sizes, boxes=boxes))
    gb, _ = gpu_ctx.lloyd(x, 3, inst_n=sizes, boxes=boxes)
    eb, _ = emu_ctx.lloyd(x, 3, inst_n=sizes, boxes=boxes)
    assert np.array_equal(gb, eb)
    # the autotuner's work-group size does not change the results
    ga_, st = gpu_ctx.lloyd(x, 3, inst_n=sizes, boxes=boxes, autotune=True)
    assert st["tuned_waves"] in (1, 4) and np.array_equal(ga_, gb)
    y = ga.sample_points(0, 3, 4000)
    gy, _ = gpu_ctx.lloyd(y, 3, k0=0)
    ey, _ = emu_ctx.lloyd(y, 3, k0=0)
    assert np.array_equal(gy, ey)
    oy, _ = oracle.lloyd(y, 3)
    assert np.abs(gy - oy).max() < 1e-12


def test_gpu_knn(gpu_ctx, emu_ctx):
    import geogram_amd as ga

    x = ga.sample_points(1, 8, 20000)
    gi, gd = gpu_ctx.knn(x, 32)
    ei, ed = emu_ctx.knn(x, 32)
    assert np.array_equal(gi, ei) and np.array_equal(gd, ed)


def test_c2_two_million_invariants(gpu_ctx):
    """Full-size C2 workload: 2M seeds, half uniform, half 8-Gaussian mixture.
    Size-independent checks: volumes sum to the cube within 1e-12, every cell
    computed, neighbour relation symmetric; plus 1500 random cells against the
    oracle (which recomputes them with its own kd-tree over all 2M seeds)."""
    import geogram_amd as ga

    n = 2_000_000
    x = ga.sample_points(1, 2, n)
    r = gpu_ctx.cells(x, k0=0)
    st = r.stats
    assert st["nb_failed"] == 0 and st["nb_uncertain"] == 0
    assert abs(r.vol.sum() - 1.0) < 1e-12
    ptr = r.nbr_ptr.astype(np.int64)
    rows = np.repeat(np.arange(n), np.diff(ptr))
    a = rows.astype(np.int64) * n + r.nbr
    b = r.nbr.astype(np.int64) * n + rows
    assert np.array_equal(np.sort(a), np.sort(b))
    sample = np.random.default_rng(0).choice(n, 1500, replace=False)
    # seed-centred oracle frame: 1e-10 relative
    oc = oracle.cells(x, k0=20, subset=sample, centred=True)
    vrel = np.abs(r.vol[sample] - oc["vol"]) / oc["vol"]
    assert vrel.max() < 1e-10
    # reference-faithful absolute frame: its plane offset |pj|^2 - |pi|^2
    # rounds to ~1.3e-15 absolute, i.e. ~4e-15 / vol^(2/3) relative in volume
    # on the tiny cells of the sigma = 0.01 clusters
    o = oracle.cells(x, k0=20, subset=sample)
    vg = r.vol[sample]
    assert np.all(np.abs(vg - o["vol"]) <= o["vol"] * (1e-10 + 4e-15 / o["vol"] ** (2 / 3)))
    for t, i in enumerate(sample):
        assert set(r.neighbors(i).tolist()) == set(o["nbr"][t, : o["cnt"][t]].tolist())
        assert set(r.neighbors(i).tolist()) == set(oc["nbr"][t, : oc["cnt"][t]].tolist())


def test_gpu_second_gather_matches_emulation(gpu_ctx, emu_ctx):
    """Voronoi cells handed to the second gather (corner test) and weighted
    cells (corner test in the first pass): GPU equals emulation bit for bit."""
    from tests.test_emu import _clusters_and_outliers
    x = _clusters_and_outliers()
    g = gpu_ctx.cells(x)
    e = emu_ctx.cells(x)
    assert g.stats["nb_regather"] == e.stats["nb_regather"] > 0
    _same(g, e)
    rng = np.random.default_rng(8)
    w = rng.random(x.shape[0]) * 1e-4
    _same(gpu_ctx.cells(x, weights=w), emu_ctx.cells(x, weights=w))


def test_gpu_cvt_energy_matches_emulation(gpu_ctx, emu_ctx):
    import geogram_amd as ga
    import torch
    x = ga.sample_points(1, 14, 20000)
    fe, ge = emu_ctx.cvt_func_grad(x)
    fg, gg = gpu_ctx.cvt_func_grad(x)
    assert fg == fe and np.array_equal(gg, ge)
    fd, gd = gpu_ctx.cvt_func_grad(torch.from_numpy(x).cuda())
    assert fd == fe and np.array_equal(gd.cpu().numpy(), ge)
