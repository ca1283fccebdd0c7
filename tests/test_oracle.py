"""Pins the oracle (oracle/) before it is used to judge the GPU path.

Each check is independent of the oracle's own code: exact rational arithmetic
for the predicate, scipy's kd-tree for kNN, Qhull half-space intersection for
the cells, closed forms (lattice cubes, volume of the domain), and the
reference's documented behaviours (security radius off => overlapping cells,
coincident seeds disconnected)."""
from fractions import Fraction

import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle
from tests.qhull_ref import qhull_cell


# ------------------------------------------------------------- predicates

def det4_exact(rows):
    M = [[Fraction(v) for v in r] for r in rows]

    def det(m):
        if len(m) == 1:
            return m[0][0]
        s = 0
        for j in range(len(m)):
            minor = [r[:j] + r[j + 1:] for r in m[1:]]
            s += (-1) ** j * m[0][j] * det(minor)
        return s

    d = det(M)
    return (d > 0) - (d < 0)


def test_det4_sign_random_and_degenerate():
    rng = np.random.default_rng(0)
    for _ in range(300):
        R = rng.standard_normal((4, 4))
        assert oracle.det4_sign(*R) == det4_exact(R)
    # exactly singular: last row a combination of the others with exact
    # dyadic coefficients (exact in floating point)
    for _ in range(200):
        R = np.round(rng.standard_normal((3, 4)) * 2**20) / 2**20
        a, b, c = (rng.integers(-4, 5, 3)).astype(float)
        r3 = a * R[0] + b * R[1] + c * R[2]
        M = np.vstack([R, r3])
        assert det4_exact(M) == 0
        assert oracle.det4_sign(*M) == 0
        # one ulp away: the sign must follow the perturbation
        M2 = M.copy()
        M2[3, 3] = np.nextafter(M2[3, 3], np.inf)
        assert oracle.det4_sign(*M2) == det4_exact(M2)


def test_det4_near_degenerate_planes():
    # four planes through (almost) one point: the conflict test of a corner
    # lying on the new plane
    rng = np.random.default_rng(1)
    p = rng.random(3)
    for _ in range(200):
        rows = []
        for _k in range(4):
            nrm = rng.standard_normal(3)
            rows.append(np.r_[nrm, -nrm @ p])
        assert oracle.det4_sign(*rows) == det4_exact(rows)
        assert oracle.det4_sign(*rows, exact_only=True) == det4_exact(rows)


# -------------------------------------------------------------------- kNN

@pytest.mark.parametrize("n,k", [(2000, 20), (5000, 64)])
def test_knn_vs_kdtree(n, k):
    rng = np.random.default_rng(n)
    x = rng.random((n, 3))
    idx, d2 = oracle.knn(x, k)
    dd, ii = cKDTree(x).query(x, k + 1)
    assert np.allclose(d2, dd[:, 1:] ** 2, rtol=1e-12, atol=0)
    for i in range(n):
        assert set(idx[i]) == set(ii[i, 1:])


# ------------------------------------------------------------------ cells

def _check_vs_qhull(x, w, res, cells, tol=1e-9):
    for i in cells:
        vol, cen, areas = qhull_cell(x, i, w)
        assert abs(res["vol"][i] - vol) <= tol * vol
        c = res["mg"][i] / res["vol"][i]
        assert np.abs(c - cen).max() <= tol * np.cbrt(vol) * 10
        mine = dict(zip(res["nbr"][i, : res["cnt"][i]].tolist(),
                        res["farea"][i, : res["cnt"][i]].tolist()))
        big = {j for j, a in areas.items() if a > 1e-9 * vol ** (2 / 3)}
        assert big <= set(mine)
        for j in mine:
            assert abs(mine[j] - areas.get(j, 0.0)) <= 1e-8 * vol ** (2 / 3)


def test_voronoi_cells_vs_qhull():
    rng = np.random.default_rng(7)
    x = rng.random((400, 3))
    res = oracle.cells(x, k0=20)
    _check_vs_qhull(x, None, res, range(0, 400, 7))
    assert abs(res["vol"].sum() - 1.0) < 1e-13


def test_laguerre_cells_vs_qhull():
    rng = np.random.default_rng(8)
    n = 300
    x = rng.random((n, 3))
    w = rng.random(n) * 2e-3   # comparable to squared spacing (~4.5e-3)
    res = oracle.cells(x, w=w, k0=20)
    _check_vs_qhull(x, w, res, range(0, n, 5))
    assert abs(res["vol"].sum() - 1.0) < 1e-13


def test_security_radius_equals_all_planes():
    rng = np.random.default_rng(9)
    x = rng.random((600, 3))
    a = oracle.cells(x, k0=20, mode=oracle.MODE_SR)
    b = oracle.cells(x, mode=oracle.MODE_ALL)
    assert np.array_equal(a["vol"], b["vol"])
    assert oracle.neighbour_sets(a) == oracle.neighbour_sets(b)
    assert a["stats"]["nb_knn_grow"] > 0


def test_security_radius_off_overlaps():
    # the stock Lloyd_iterations disables the security radius (voronoi/CVT.cpp:139):
    # with only 20 neighbours the cells overlap and the volumes sum above 1
    rng = np.random.default_rng(10)
    x = rng.random((20000, 3))
    res = oracle.cells(x, k0=20, mode=oracle.MODE_NO_GROWTH)
    assert res["vol"].sum() > 1.01
    res = oracle.cells(x, k0=20, mode=oracle.MODE_SR)
    assert abs(res["vol"].sum() - 1.0) < 1e-12


def test_lattice_cubes_closed_form():
    m = 6
    h = 1.0 / m
    g = (np.arange(m) + 0.5) * h
    x = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T.copy()
    res = oracle.cells(x, k0=20)
    assert np.allclose(res["vol"], h**3, rtol=1e-13, atol=0)
    assert np.allclose(res["mg"] / res["vol"][:, None], x, atol=1e-14)
    for i in range(x.shape[0]):
        fa = res["farea"][i, : res["cnt"][i]]
        big = fa[fa > 1e-12]
        assert np.allclose(big, h * h, rtol=1e-12)
        # face neighbours of the cube: seeds at distance h
        nb = res["nbr"][i, : res["cnt"][i]][fa > 1e-12]
        assert np.allclose(np.linalg.norm(x[nb] - x[i], axis=1), h)


def test_coincident_seeds_rule():
    # delaunay/delaunay_nn.cpp:125-133: the later copy has no neighbours
    rng = np.random.default_rng(11)
    x = rng.random((200, 3))
    x[150] = x[40]
    res = oracle.cells(x, k0=20)
    assert res["cnt"][150] == 0 and abs(res["vol"][150] - 1.0) < 1e-15
    assert res["stats"]["nb_disconnected"] == 1
    assert 150 not in set(res["nbr"][40, : res["cnt"][40]].tolist())


def test_lloyd_moves_to_centroids_and_lowers_energy():
    rng = np.random.default_rng(12)
    x = rng.random((1500, 3))

    def energy(p):
        # CVT energy sum_i int |x - p_i|^2 by Monte Carlo on a fixed sample
        s = np.random.default_rng(0).random((40000, 3))
        d, _ = cKDTree(p).query(s)
        return np.mean(d**2)

    x1, _ = oracle.lloyd(x, 1)
    r0 = oracle.cells(x)
    assert np.allclose(x1, r0["mg"] / r0["vol"][:, None], atol=1e-15)
    x5, _ = oracle.lloyd(x, 5)
    assert energy(x5) < energy(x1) < energy(x)


def test_second_moments_closed_forms_and_sampling():
    """Oracle cell integrals of |x - p_i|^2 (the CVT energy terms): one seed
    owns the box (sum over axes of ((1-p)^3 + p^3)/3); the centred 8^3
    lattice gives h^5/4 per cell; random seeds agree with a Monte Carlo
    estimate of the energy."""
    p = np.array([[0.3, 0.6, 0.8]])
    o = oracle.cells(p, k0=1, e2=True)
    assert abs(o["e2"][0] - sum(((1 - a) ** 3 + a ** 3) / 3 for a in p[0])) < 1e-15
    g = (np.indices((8, 8, 8)).reshape(3, -1).T + 0.5) / 8
    o = oracle.cells(g, e2=True, centred=True)
    assert np.allclose(o["e2"], (1 / 8) ** 5 / 4, rtol=1e-12, atol=0)
    x = np.random.default_rng(21).random((400, 3))
    o = oracle.cells(x, e2=True, centred=True)
    s = np.random.default_rng(5).random((400000, 3))
    d, _ = cKDTree(x).query(s)
    mc = np.mean(d ** 2)
    assert abs(o["e2"].sum() - mc) < 4 * np.std(d ** 2) / np.sqrt(len(s))
