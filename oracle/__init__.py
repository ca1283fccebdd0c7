"""ORACLE — test infrastructure only.

ctypes bindings to ``liboracle.so``, the plain-C restatement of geogram's
meshless Voronoi / Laguerre cell path (see ``oracle.h`` for the reference
files and lines each part follows).  Only ``tests/``, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py`` may import this module; the product
package never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class Stats(ctypes.Structure):
    _fields_ = [
        ("nb_clip", ctypes.c_longlong),
        ("nb_knn_grow", ctypes.c_longlong),
        ("nb_disconnected", ctypes.c_longlong),
        ("nb_empty", ctypes.c_longlong),
        ("nb_bad_border", ctypes.c_longlong),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_knn.argtypes = [dp, ctypes.c_int, ctypes.c_int, ip, dp]
        L.orc_knn_query.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ip, dp]
        L.orc_cells.argtypes = [dp, dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int,
                                dp, dp, ip, ip, dp, ctypes.c_int, ctypes.POINTER(Stats)]
        L.orc_cells_subset.argtypes = [dp, dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int,
                                       ip, ctypes.c_int, dp, dp, ip, ip, dp, ctypes.c_int,
                                       ctypes.POINTER(Stats)]
        L.orc_lloyd.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(Stats)]
        L.orc_det4_sign.argtypes = [dp, dp, dp, dp]
        L.orc_surface_centroids.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int, ip, ctypes.c_int,
                                            dp, ctypes.c_int, dp, dp,
                                            ctypes.POINTER(ctypes.c_longlong)]
        L.orc_det4_sign_exact.argtypes = [dp, dp, dp, dp]
        L.orc_det3_sign.argtypes = [dp, dp, dp]
        L.orc_exact_fallbacks.restype = ctypes.c_longlong
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_local_planes.argtypes = [ctypes.c_int]
        L.orc_set_e2_output.argtypes = [ctypes.c_void_p]
        L.orc_set_mesh_output.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


UNIT_BOX = np.array([0.0, 0.0, 0.0, 1.0, 1.0, 1.0])

MODE_NO_GROWTH = 0
MODE_SR = 1
MODE_ALL = 2


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def knn(xyz, k):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    idx = np.zeros((n, k), dtype=np.int32)
    d2 = np.zeros((n, k), dtype=np.float64)
    err = lib().orc_knn(_dp(xyz), n, k, _ip(idx), _dp(d2))
    if err:
        raise RuntimeError(f"orc_knn failed ({err})")
    return idx, d2


def knn_query(xyz, q, k):
    """Exact k nearest points of xyz to each query point (none excluded),
    sorted by (squared distance, index)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    idx = np.zeros((q.shape[0], k), dtype=np.int32)
    d2 = np.zeros((q.shape[0], k), dtype=np.float64)
    err = lib().orc_knn_query(_dp(xyz), xyz.shape[0], _dp(q), q.shape[0], k, _ip(idx), _dp(d2))
    if err:
        raise RuntimeError(f"orc_knn_query failed ({err})")
    return idx, d2


def cells(xyz, w=None, box=UNIT_BOX, k0=20, mode=MODE_SR, stride=96, subset=None,
          centred=False, e2=False):
    """Per-cell volume, mg (= integral of x), neighbour lists and facet areas.

    centred=True forms the planes in the seed-centred frame (same algorithm).
    e2=True adds "e2": the integral of |x - p_i|^2 over each cell (the CVT
    energy terms, RVD.cpp:792-866).
    Returns dict(vol, mg, cnt, nbr, farea, stats); nbr/farea are (n, stride)."""
    lib().orc_set_local_planes(1 if centred else 0)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    box = np.ascontiguousarray(box, dtype=np.float64)
    wp = None
    if w is not None:
        w = np.ascontiguousarray(w, dtype=np.float64)
        wp = _dp(w)
    if subset is None:
        nc = n
    else:
        subset = np.ascontiguousarray(subset, dtype=np.int32)
        nc = subset.shape[0]
    vol = np.zeros(nc)
    mg = np.zeros((nc, 3))
    cnt = np.zeros(nc, dtype=np.int32)
    nbr = np.full((nc, stride), -1, dtype=np.int32)
    fa = np.zeros((nc, stride))
    st = Stats()
    e2a = np.zeros(nc) if e2 else None
    lib().orc_set_e2_output(_dp(e2a) if e2 else None)
    if subset is None:
        err = lib().orc_cells(_dp(xyz), wp, n, _dp(box), int(k0), int(mode), _dp(vol),
                              _dp(mg), _ip(cnt), _ip(nbr), _dp(fa), stride, ctypes.byref(st))
    else:
        err = lib().orc_cells_subset(_dp(xyz), wp, n, _dp(box), int(k0), int(mode),
                                     _ip(subset), nc, _dp(vol), _dp(mg), _ip(cnt), _ip(nbr),
                                     _dp(fa), stride, ctypes.byref(st))
    lib().orc_set_e2_output(None)
    if err:
        raise RuntimeError(f"orc_cells failed ({err})")
    out = dict(vol=vol, mg=mg, cnt=cnt, nbr=nbr, farea=fa, stats=st.as_dict())
    if e2:
        out["e2"] = e2a
    return out


def cell_meshes(xyz, w=None, box=UNIT_BOX, subset=None, centred=True, istride=2048,
                vstride=768):
    """Polyhedron and primal tetrahedra of each cell (ConvexCell::save /
    append_to_mesh, compute_RDT).  Returns a list, per cell, of dicts with
    "v" (nv, 3) vertices, "faces" [(neighbour or -1, [vertex numbers])] and
    "tets" [(i, j, k, l)]."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    nc = xyz.shape[0] if subset is None else len(subset)
    ib = np.zeros((nc, istride), dtype=np.int32)
    vb = np.zeros((nc, vstride))
    lib().orc_set_mesh_output(ib.ctypes.data_as(ctypes.c_void_p), istride,
                              vb.ctypes.data_as(ctypes.c_void_p), vstride)
    try:
        cells(xyz, w=w, box=box, subset=subset, centred=centred, stride=256)
    finally:
        lib().orc_set_mesh_output(None, 0, None, 0)
    out = []
    for c in range(nc):
        nv, nf, nt = ib[c, :3]
        if nv < 0:
            raise RuntimeError("cell_meshes: stride too small")
        pos, faces = 3, []
        for _ in range(nf):
            nb, sz = int(ib[c, pos]), int(ib[c, pos + 1])
            faces.append((nb, ib[c, pos + 2:pos + 2 + sz].tolist()))
            pos += 2 + sz
        tets = [tuple(ib[c, pos + 4 * t:pos + 4 * t + 4].tolist()) for t in range(nt)]
        out.append({"v": vb[c, :3 * nv].reshape(-1, 3).copy(), "faces": faces, "tets": tets})
    return out


def periodic_cells(xyz, w=None, box=UNIT_BOX, centred=True, stride=128, e2=False):
    """Cells of the 3-torus of box (one period): the seeds and their 26
    translated copies x + s L (s in {-1, 0, 1}^3), the security-radius loop of
    every seed of the period over all 27 n, neighbour ids taken modulo n
    (PeriodicDelaunay3d's periodic_vertex_real, periodic.h:85-88)."""
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    box = np.asarray(box, dtype=np.float64)
    L = box[3:] - box[:3]
    shifts = [(0, 0, 0)] + [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)
                            if (a, b, c) != (0, 0, 0)]
    X = np.vstack([x + np.array(s, dtype=np.float64) * L for s in shifts])
    W = None if w is None else np.tile(np.asarray(w, dtype=np.float64), len(shifts))
    big = np.concatenate([box[:3] - L, box[3:] + L])
    o = cells(X, w=W, box=big, subset=np.arange(n, dtype=np.int32), centred=centred,
              stride=stride, e2=e2)
    o["nbr"] = np.where(o["nbr"] >= 0, o["nbr"] % n, -1)
    return o


def surface_centroids(xyz, V, F, vw=None, brute=False):
    """Voronoi cells of the seeds restricted to the triangle mesh (V, F):
    (mg (n, 3), m (n,), polygons).  vw: per-vertex density (linear on each
    triangle).  brute: every facet clipped by all bisectors of every seed."""
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
    F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
    n = x.shape[0]
    mg = np.zeros((n, 3))
    m = np.zeros(n)
    npoly = ctypes.c_longlong(0)
    wp = None
    if vw is not None:
        vw = np.ascontiguousarray(vw, dtype=np.float64)
        wp = _dp(vw)
    err = lib().orc_surface_centroids(_dp(x), n, _dp(V), V.shape[0], _ip(F), F.shape[0], wp,
                                      1 if brute else 0, _dp(mg), _dp(m), ctypes.byref(npoly))
    if err:
        raise RuntimeError(f"orc_surface_centroids failed ({err})")
    return mg, m, int(npoly.value)


def lloyd(xyz, iters, box=UNIT_BOX, k0=20):
    x = np.array(xyz, dtype=np.float64, copy=True).reshape(-1, 3)
    box = np.ascontiguousarray(box, dtype=np.float64)
    st = Stats()
    err = lib().orc_lloyd(_dp(x), x.shape[0], _dp(box), int(iters), int(k0), ctypes.byref(st))
    if err:
        raise RuntimeError(f"orc_lloyd failed ({err})")
    return x, st.as_dict()


def det4_sign(r0, r1, r2, r3, exact_only=False):
    rows = [np.ascontiguousarray(r, dtype=np.float64) for r in (r0, r1, r2, r3)]
    f = lib().orc_det4_sign_exact if exact_only else lib().orc_det4_sign
    return int(f(*[_dp(r) for r in rows]))


def det3_sign(r0, r1, r2):
    rows = [np.ascontiguousarray(r, dtype=np.float64) for r in (r0, r1, r2)]
    return int(lib().orc_det3_sign(*[_dp(r) for r in rows]))


def exact_fallbacks() -> int:
    return int(lib().orc_exact_fallbacks())


def neighbour_sets(res):
    """List of frozensets of neighbour ids per cell."""
    return [frozenset(res["nbr"][i, : res["cnt"][i]].tolist()) for i in range(len(res["cnt"]))]


# ------------------------------------------------- the reference's own code
_REF_PATH = os.path.join(_HERE, "_ref", "libgeoref.so")
_ref = None


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def ref_lib():
    """oracle/_ref/libgeoref.so (built by oracle/build_ref.sh from geogram's
    sources): geogram's Delaunay + ConvexCell cells and its kd-tree kNN."""
    global _ref
    if _ref is None:
        L = ctypes.CDLL(_REF_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.georef_delaunay_cells.argtypes = [dp, dp, ctypes.c_int, dp, ctypes.c_int, dp, dp, ip,
                                            ip, dp, ctypes.c_int, dp]
        L.georef_knn.argtypes = [dp, ctypes.c_int, ctypes.c_int, ip, dp]
        L.georef_periodic_cells.argtypes = L.georef_delaunay_cells.argtypes
        L.georef_max_threads.restype = ctypes.c_int
        _ref = L
    return _ref


def ref_cells(xyz, w=None, box=UNIT_BOX, threads=0, stride=128, periodic=False):
    """geogram's fastest CPU path (PeriodicDelaunay3d + ConvexCell + box
    clips; periodic: PeriodicDelaunay3d on the torus of the box):
    dict(vol, mg, cnt, nbr, farea, t_delaunay_s, t_cells_s)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    box = np.ascontiguousarray(box, dtype=np.float64)
    wp = None
    if w is not None:
        w = np.ascontiguousarray(w, dtype=np.float64)
        wp = _dp(w)
    vol = np.zeros(n)
    mg = np.zeros((n, 3))
    cnt = np.zeros(n, dtype=np.int32)
    nbr = np.full((n, stride), -1, dtype=np.int32)
    fa = np.zeros((n, stride))
    t = np.zeros(2)
    f = ref_lib().georef_periodic_cells if periodic else ref_lib().georef_delaunay_cells
    err = f(_dp(xyz), wp, n, _dp(box), int(threads), _dp(vol), _dp(mg), _ip(cnt), _ip(nbr),
            _dp(fa), stride, _dp(t))
    if err == -3:
        raise ValueError("georef_delaunay_cells: a weighted seed has an empty cell "
                         "(geogram's triangulation stops)")
    if err:
        raise RuntimeError(f"georef_delaunay_cells failed ({err})")
    return dict(vol=vol, mg=mg, cnt=cnt, nbr=nbr, farea=fa, t_delaunay_s=float(t[0]),
                t_cells_s=float(t[1]))


def ref_knn(xyz, k):
    """geogram's kd-tree: the k nearest points of each point, itself first."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    idx = np.zeros((n, k), dtype=np.int32)
    d2 = np.zeros((n, k))
    ref_lib().georef_knn(_dp(xyz), n, k, _ip(idx), _dp(d2))
    return idx, d2
