"""Cell meshes and primal tetrahedra (vgc_cell_mesh, vgc_rdt: ConvexCell::
append_to_mesh / save, RestrictedVoronoiDiagram::compute_RDT).

The oracle's polyhedra are pinned by closed checks — Euler's formula,
outward orientation, the divergence-theorem volume against the cell volume —
and its tetrahedra by Qhull: for points in general position the primal
tetrahedra of the box-clipped cells are exactly the Delaunay tetrahedra whose
circumcentre lies inside the box.  The emulated kernels are then held to the
oracle cell by cell."""
import numpy as np
import pytest
from scipy.spatial import Delaunay

import geogram_amd as ga
import oracle


def polygon_normal(P):
    return 0.5 * sum(np.cross(P[a], P[(a + 1) % len(P)]) for a in range(len(P)))


def polyhedron_volume(V, faces):
    # divergence theorem over a fan of each (planar) polygon
    vol = 0.0
    for f in faces:
        P = V[f]
        for a in range(1, len(f) - 1):
            vol += np.dot(P[0], np.cross(P[a], P[a + 1])) / 6.0
    return vol


def circumcentres(x, T):
    A = x[T[:, 1:]] - x[T[:, :1]]
    b = 0.5 * np.einsum("tij,tij->ti", A, A)
    return x[T[:, 0]] + np.linalg.solve(A, b[..., None])[..., 0]


def orient3d(p0, p1, p2, p3):
    return np.linalg.det(np.stack([p1 - p0, p2 - p0, p3 - p0]))


def test_oracle_polyhedra_closed_checks():
    x = ga.sample_points(1, 3, 600)
    o = oracle.cells(x, centred=True)
    m = oracle.cell_meshes(x)
    for i, c in enumerate(m):
        V, faces = c["v"], [f for _, f in c["faces"]]
        E = sum(len(f) for f in faces) // 2
        assert len(V) - E + len(faces) == 2
        cen = V.mean(axis=0)
        for f in faces:
            assert np.dot(polygon_normal(V[f]), V[f].mean(axis=0) - cen) > 0
        assert abs(polyhedron_volume(V, faces) - o["vol"][i]) <= 1e-12
        nbrs = sorted(nb for nb, _ in c["faces"] if nb >= 0)
        assert nbrs == sorted(o["nbr"][i, : o["cnt"][i]].tolist())


def test_oracle_tetrahedra_are_delaunay_with_inner_circumcentres():
    x = np.random.default_rng(7).random((3000, 3))
    m = oracle.cell_meshes(x)
    tets = np.array([t for c in m for t in c["tets"]])
    D = Delaunay(x)
    cc = circumcentres(x, D.simplices)
    inside = np.all((cc > 0.0) & (cc < 1.0), axis=1)
    ref = {tuple(sorted(t)) for t in D.simplices[inside].tolist()}
    got = {tuple(sorted(t)) for t in tets.tolist()}
    assert got == ref
    for t in tets[:2000]:
        assert orient3d(*x[t]) > 0


def _face_key(V, f, nb):
    return nb, V[f]


def _same_cycle(P, Q, tol):
    """Same cyclic sequence of points (same orientation), up to tol."""
    if len(P) != len(Q):
        return False
    return any(np.abs(np.roll(Q, -k, axis=0) - P).max() <= tol for k in range(len(Q)))


def compare_meshes(r, m, x, tol=1e-12):
    n = len(m)
    vp = r["cell_vptr"].astype(np.int64)
    fp = r["cell_fptr"].astype(np.int64)
    fptr = r["facet_ptr"].astype(np.int64)
    for i in range(n):
        c = m[i]
        V = r["vertices"]
        assert vp[i + 1] - vp[i] == len(c["v"])
        mine = []
        for f in range(fp[i], fp[i + 1]):
            verts = r["facet_vertices"][fptr[f]:fptr[f + 1]]
            assert np.all((verts >= vp[i]) & (verts < vp[i + 1]))
            mine.append(_face_key(V, verts, int(r["facet_nbr"][f])))
        theirs = [_face_key(c["v"], np.array(f), nb) for nb, f in c["faces"]]
        assert len(mine) == len(theirs)
        for nb, P in theirs:
            ok = [Q for nb2, Q in mine if nb2 == nb and _same_cycle(P, Q, tol)]
            assert ok, f"cell {i}: facet of {nb} not matched"


def compare_tets(tets, m, x):
    """Same tetrahedra as the oracle's (the order of the last three seeds
    follows each implementation's corner), all positively oriented (the
    orient_3d swap of compute_RDT; exact zeros do not occur here)."""
    mine = {tuple(sorted(t)) for t in tets.tolist()}
    theirs = {tuple(sorted(t)) for c in m for t in c["tets"]}
    assert len(mine) == len(tets) and mine == theirs
    P = x[tets]
    det = np.linalg.det(np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]], 1))
    assert np.all(det > 0)


@pytest.mark.parametrize("kind", [0, 1])
def test_emulated_cell_mesh_vs_oracle(emu_ctx, kind):
    x = ga.sample_points(kind, 50 + kind, 2000)
    r = emu_ctx.cell_mesh(x, tets=True)
    m = oracle.cell_meshes(x)
    compare_meshes(r, m, x)
    compare_tets(r["tets"], m, x)
    t2 = emu_ctx.rdt(x)
    assert np.array_equal(t2, r["tets"])


def test_emulated_laguerre_mesh_and_shrink(emu_ctx):
    rng = np.random.default_rng(9)
    x = rng.random((1500, 3))
    w = rng.random(1500) * 3e-4
    r = emu_ctx.cell_mesh(x, weights=w, tets=True)
    m = oracle.cell_meshes(x, w=w)
    compare_meshes(r, m, x)
    compare_tets(r["tets"], m, x)
    # shrink: vertices pulled towards each cell's centroid
    cr = emu_ctx.cells(x, weights=w, neighbors=False)
    s = emu_ctx.cell_mesh(x, weights=w, shrink=0.25)
    vp = s["cell_vptr"].astype(np.int64)
    for i in (0, 17, 900):
        g = cr.mg[i] / cr.vol[i]
        V0 = r["vertices"][vp[i]:vp[i + 1]]
        assert np.abs(s["vertices"][vp[i]:vp[i + 1]] - (0.25 * g + 0.75 * V0)).max() < 1e-12
