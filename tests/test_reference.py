"""The oracle against the reference itself: geogram's own Delaunay +
ConvexCell cells and kd-tree kNN, compiled from its sources into
oracle/_ref/libgeoref.so by oracle/build_ref.sh (skipped where it is not
built).  The reference computes each cell from its Delaunay neighbourhood
and the box clips, the oracle by the security-radius loop; both form planes
in absolute coordinates, so they agree to the rounding of those planes."""
import numpy as np
import pytest

import geogram_amd as ga
import oracle

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")


def _agree(r, o, n, rel=1e-12):
    vol = np.maximum(o["vol"], 1e-300)
    tol = vol * (rel + 4e-15 / vol ** (2 / 3))
    assert np.all(np.abs(r["vol"] - o["vol"]) <= tol)
    nz = o["vol"] > 0
    cr = r["mg"][nz] / r["vol"][nz, None]
    co = o["mg"][nz] / o["vol"][nz, None]
    assert (np.abs(cr - co) / np.cbrt(o["vol"][nz])[:, None]).max() < 1e-10
    for i in range(n):
        a = dict(zip(r["nbr"][i, : r["cnt"][i]].tolist(), r["farea"][i, : r["cnt"][i]].tolist()))
        b = dict(zip(o["nbr"][i, : o["cnt"][i]].tolist(), o["farea"][i, : o["cnt"][i]].tolist()))
        assert set(a) == set(b), f"cell {i}"
        for j in b:
            assert abs(a[j] - b[j]) <= 1e-10 * b[j] + 1e-11 * vol[i] ** (2 / 3)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_voronoi_cells_oracle_vs_reference(kind):
    x = ga.sample_points(kind, 70 + kind, 20000)
    r = oracle.ref_cells(x)
    o = oracle.cells(x, k0=20)
    _agree(r, o, len(x))
    assert abs(r["vol"].sum() - 1.0) < 1e-12


def test_laguerre_cells_oracle_vs_reference():
    rng = np.random.default_rng(12)
    x = rng.random((8000, 3))
    # weights small enough that no cell is empty: with hidden seeds geogram's
    # parallel triangulation may stop on them (has_empty_cells) or not,
    # depending on the thread count
    w = rng.random(8000) * 1e-4
    r = oracle.ref_cells(x, w)
    o = oracle.cells(x, w=w, k0=20)
    _agree(r, o, len(x))


def test_knn_oracle_vs_reference():
    # geogram's kd-tree returns the point itself first; equal distances may
    # come in another order (kd_tree.h:176-190), so rows are compared as
    # (distance, set) beyond the first entry
    x = ga.sample_points(1, 9, 20000)
    ri, rd = oracle.ref_knn(x, 21)
    assert np.array_equal(ri[:, 0], np.arange(len(x))) and np.all(rd[:, 0] == 0)
    oi, od = oracle.knn(x, 20)
    assert np.array_equal(rd[:, 1:], od)
    assert np.array_equal(np.sort(ri[:, 1:], axis=1), np.sort(oi, axis=1))
