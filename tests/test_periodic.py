"""Periodic domain (VGC_PERIODIC: the box is one period of a 3-torus), the
counterpart of PeriodicDelaunay3d(periodic=true) (delaunay/periodic_delaunay_3d.h,
delaunay/periodic.h) for cells, Lloyd, the CVT energy / gradient and
semi-discrete transport.

Pinned three ways: the oracle's periodic cells (the seeds and their 26
translated copies, neighbour ids modulo n) against geogram's own periodic
triangulation compiled from its sources (oracle/_ref, skipped where it is not
built); the invariants of a torus (the cells tile it: volumes add up to the
box volume, neighbour relations are symmetric, a seed set translated by a
period has the same cells); the emulated kernels and then the GPU against
the oracle."""
import numpy as np
import pytest

import geogram_amd as ga
import oracle

BOX2 = np.array([-1.0, 0.5, 2.0, 1.0, 2.5, 3.5])


def _wrap(c, box):
    """Seed positions folded into [lo, hi) (the Lloyd update's wrap)."""
    lo, hi = box[:3], box[3:]
    L = hi - lo
    c = c - L * np.floor((c - lo) / L)
    c = np.where(c >= hi, c - L, c)
    return np.where(c < lo, lo, c)


def _sets(cnt, nbr, i):
    return set(nbr[i, : cnt[i]].tolist())


def _compare(r, o, n, rel=1e-10):
    vol = np.maximum(o["vol"], 1e-300)
    vr = np.abs(r.vol - o["vol"]) / vol
    assert vr.max() < rel, f"volume rel err {vr.max()} at {vr.argmax()}"
    nz = o["vol"] > 0
    cen = r.mg[nz] / r.vol[nz, None]
    oc = o["mg"][nz] / o["vol"][nz, None]
    assert (np.abs(cen - oc) / np.cbrt(o["vol"][nz])[:, None]).max() < rel
    bad = [i for i in range(n) if set(r.neighbors(i).tolist()) != _sets(o["cnt"], o["nbr"], i)]
    assert not bad, f"{len(bad)} neighbour sets differ, first {bad[:5]}"
    # facet areas per neighbouring seed: with few seeds two images of one
    # seed can both bound a cell, so areas are compared as sorted lists
    for i in range(0, n, max(1, n // 500)):
        ga_, go_ = {}, {}
        for j, a in zip(r.neighbors(i).tolist(), r.facet_areas(i).tolist()):
            ga_.setdefault(j, []).append(a)
        for j, a in zip(o["nbr"][i, : o["cnt"][i]].tolist(), o["farea"][i, : o["cnt"][i]].tolist()):
            go_.setdefault(j, []).append(a)
        for j in go_:
            assert len(ga_[j]) == len(go_[j])
            for x, y in zip(sorted(ga_[j]), sorted(go_[j])):
                assert abs(x - y) <= rel * y + 2e-11 * o["vol"][i] ** (2 / 3)


# ---------------------------------------------------------------- oracle pins
def test_oracle_periodic_invariants():
    rng = np.random.default_rng(3)
    x = BOX2[:3] + rng.random((1500, 3)) * (BOX2[3:] - BOX2[:3])
    o = oracle.periodic_cells(x, box=BOX2)
    V = np.prod(BOX2[3:] - BOX2[:3])
    assert abs(o["vol"].sum() - V) < 1e-12 * V
    # symmetric neighbour relation, equal facet areas on both sides
    for i in range(0, 1500, 7):
        for j, a in zip(o["nbr"][i, : o["cnt"][i]].tolist(), o["farea"][i, : o["cnt"][i]].tolist()):
            row = dict(zip(o["nbr"][j, : o["cnt"][j]].tolist(), o["farea"][j, : o["cnt"][j]].tolist()))
            assert i in row and abs(row[i] - a) <= 1e-11 * max(a, 1e-6)
    # translation by part of a period (folded back into the box) only
    # translates the cells: volumes and neighbour sets unchanged
    t = np.array([0.37, 0.61, 0.13]) * (BOX2[3:] - BOX2[:3])
    o2 = oracle.periodic_cells(_wrap(x + t, BOX2), box=BOX2)
    assert np.allclose(o2["vol"], o["vol"], rtol=1e-9, atol=0)
    assert all(_sets(o["cnt"], o["nbr"], i) == _sets(o2["cnt"], o2["nbr"], i) for i in range(1500))


def test_oracle_periodic_lattice_closed_form():
    # a 4x4x4 lattice (jittered by 1e-9 to avoid exact degeneracy) on the
    # torus: 64 cubes of volume 1/64, each with its 6 face neighbours
    g = (np.arange(4) + 0.5) / 4
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    x = x + 1e-9 * np.random.default_rng(1).standard_normal(x.shape)
    o = oracle.periodic_cells(x)
    assert np.allclose(o["vol"], 1 / 64, rtol=1e-6)
    idx = np.arange(64).reshape(4, 4, 4)
    for a in range(4):
        for b in range(4):
            for c in range(4):
                i = idx[a, b, c]
                faces = {idx[(a + 1) % 4, b, c], idx[(a - 1) % 4, b, c], idx[a, (b + 1) % 4, c],
                         idx[a, (b - 1) % 4, c], idx[a, b, (c + 1) % 4], idx[a, b, (c - 1) % 4]}
                big = {j for j, f in zip(o["nbr"][i, : o["cnt"][i]], o["farea"][i, : o["cnt"][i]])
                       if f > 1e-4}
                assert big == faces


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("weighted", [False, True])
def test_oracle_periodic_vs_reference(weighted):
    rng = np.random.default_rng(5 + weighted)
    x = BOX2[:3] + rng.random((6000, 3)) * (BOX2[3:] - BOX2[:3])
    w = rng.random(6000) * 1e-3 if weighted else None
    r = oracle.ref_cells(x, w, box=BOX2, periodic=True)
    o = oracle.periodic_cells(x, w, box=BOX2, centred=False)
    vol = np.maximum(o["vol"], 1e-300)
    assert np.all(np.abs(r["vol"] - o["vol"]) <= vol * (1e-11 + 2e-14 / vol ** (2 / 3)))
    assert abs(r["vol"].sum() - np.prod(BOX2[3:] - BOX2[:3])) < 1e-11
    nz = o["vol"] > 0
    cr = r["mg"][nz] / r["vol"][nz, None]
    co = o["mg"][nz] / o["vol"][nz, None]
    assert (np.abs(cr - co) / np.cbrt(o["vol"][nz])[:, None]).max() < 1e-9
    bad = [i for i in range(6000) if _sets(r["cnt"], r["nbr"], i) != _sets(o["cnt"], o["nbr"], i)]
    assert not bad, f"{len(bad)} neighbour sets differ, first {bad[:5]}"


# ------------------------------------------------------- emulated kernels
@pytest.mark.parametrize("kind,n", [(0, 3000), (1, 3000), (2, 2000)])
def test_emu_periodic_cells(emu_ctx, kind, n):
    x = ga.sample_points(kind, 40 + kind, n)
    r = emu_ctx.cells(x, periodic=True)
    assert r.stats["nb_failed"] == 0
    assert abs(r.vol.sum() - 1.0) < 1e-12
    _compare(r, oracle.periodic_cells(x), n)


def test_emu_periodic_laguerre_box(emu_ctx):
    rng = np.random.default_rng(8)
    x = BOX2[:3] + rng.random((3000, 3)) * (BOX2[3:] - BOX2[:3])
    w = rng.random(3000) * 2e-3
    r = emu_ctx.cells(x, w, box=BOX2, periodic=True)
    _compare(r, oracle.periodic_cells(x, w, box=BOX2), 3000)


def test_emu_periodic_few_seeds_margin_retry(emu_ctx):
    # 40 seeds: cells are a large part of the period, the first image margin
    # is too thin and the reach check makes it grow
    x = ga.sample_points(0, 9, 40)
    r = emu_ctx.cells(x, periodic=True)
    assert abs(r.vol.sum() - 1.0) < 1e-12
    _compare(r, oracle.periodic_cells(x), 40)


def test_emu_periodic_lloyd_and_energy(emu_ctx):
    x = ga.sample_points(1, 12, 2000)
    o = oracle.periodic_cells(x, e2=True)
    y, st = emu_ctx.lloyd(x, 1, periodic=True)
    exp = _wrap(o["mg"] / o["vol"][:, None], np.array(ga.UNIT_BOX))
    d = np.abs(y - exp)
    d = np.minimum(d, 1.0 - d)  # a centroid on the seam may fold either way
    assert d.max() < 1e-12
    assert np.all((y >= 0.0) & (y < 1.0))
    f, g = emu_ctx.cvt_func_grad(x, periodic=True)
    assert abs(f - o["e2"].sum()) < 1e-10 * o["e2"].sum()
    go = 2.0 * (o["vol"][:, None] * x - o["mg"])
    assert np.abs(g - go).max() < 1e-10 * np.abs(go).max()


def test_emu_periodic_transport(emu_ctx):
    rng = np.random.default_rng(21)
    x = rng.random((3000, 3))
    nu = np.full(3000, 1.0 / 3000)
    w, st = emu_ctx.ot_solve(x, nu, tol=1e-6, periodic=True)
    assert st["converged"]
    o = oracle.periodic_cells(x, w)
    # converged: max_i |m_i - nu_i| / nu_i <= 1e-6, up to the frames' rounding
    assert (np.abs(o["vol"] - nu) / nu).max() < 1.01e-6
    assert abs(o["vol"].sum() - 1.0) < 1e-12


# --------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_gpu_periodic_cells_vs_oracle(gpu_ctx, emu_ctx, kind):
    x = ga.sample_points(kind, 60 + kind, 20000)
    r = gpu_ctx.cells(x, periodic=True)
    assert r.stats["nb_failed"] == 0
    assert abs(r.vol.sum() - 1.0) < 1e-11
    _compare(r, oracle.periodic_cells(x), 20000)
    e = emu_ctx.cells(x, periodic=True)
    assert np.array_equal(e.vol, r.vol) and np.array_equal(e.mg, r.mg)


@pytest.mark.gpu
def test_gpu_periodic_laguerre_lloyd_transport(gpu_ctx):
    rng = np.random.default_rng(31)
    x = BOX2[:3] + rng.random((20000, 3)) * (BOX2[3:] - BOX2[:3])
    w = rng.random(20000) * 2e-4
    r = gpu_ctx.cells(x, w, box=BOX2, periodic=True)
    _compare(r, oracle.periodic_cells(x, w, box=BOX2), 20000)
    y, _ = gpu_ctx.lloyd(x, 1, box=BOX2, periodic=True)
    o = oracle.periodic_cells(x, box=BOX2)
    exp = _wrap(o["mg"] / o["vol"][:, None], BOX2)
    L = BOX2[3:] - BOX2[:3]
    d = np.abs(y - exp)
    assert np.minimum(d, L - d).max() < 1e-11
    V = np.prod(L)
    nu = np.full(20000, V / 20000)
    wt, st = gpu_ctx.ot_solve(x, nu, box=BOX2, tol=1e-6, periodic=True)
    assert st["converged"]
    oo = oracle.periodic_cells(x, wt, box=BOX2)
    assert (np.abs(oo["vol"] - nu) / nu).max() < 1.01e-6
