"""geogram_amd — MI355X-native meshless Voronoi / Laguerre cells.

Python mirror of geogram's entry points for the meshless cell path, over the
C ABI of ``libvgc.so`` (include/vgc.h):

* ``NearestNeighborSearch`` — points/nn_search.h (set_points,
  get_nearest_neighbors), batched over all points;
* ``RestrictedVoronoiDiagram.compute_centroids`` — voronoi/RVD.h:284-290
  (caller-zeroed mg / m accumulated, RVD.cpp:500-528), volumetric, box domain;
* ``CentroidalVoronoiTesselation`` — voronoi/CVT.h (set_points,
  Lloyd_iterations, embedding), security radius on;
* ``cells`` — per-cell volume, centroid, neighbour CSR and facet areas, the
  outputs of VBW::ConvexCell (voronoi/convex_cell.h);
* ``OptimalTransportMap3d`` — the semi-discrete optimal-transport solver
  called by examples/exploragram/compute_OTM/main.cpp:283-297.

Arrays may be numpy (host) or torch CUDA tensors (device-resident, no PCIe).
Everything runs on the GPU; without ``libvgc.so`` or a GPU the calls raise.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import (OtStats, Stats, VgcError, VGC_COUNT_WORK, VGC_DEVICE_PTRS, VGC_NO_NEIGHBORS, VGC_OK,
                   VGC_PARTIAL)

UNIT_BOX = (0.0, 0.0, 0.0, 1.0, 1.0, 1.0)

__all__ = [
    "Context", "CellResult", "NearestNeighborSearch", "RestrictedVoronoiDiagram",
    "CentroidalVoronoiTesselation", "OptimalTransportMap3d", "default_context", "cells", "knn",
    "knn_search",
    "lloyd", "ot_solve",
    "random_float64", "sample_points", "Stats", "VgcError", "UNIT_BOX",
]


def _is_torch_cuda(a):
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(a, torch.Tensor) and a.is_cuda


def _ptr(a):
    if a is None:
        return None
    if _is_torch_cuda(a):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


class CellResult:
    """Per-seed outputs in the caller's numbering.

    vol[n]; mg[n,3] (integral of x, so centroid = mg / vol); nbr_ptr[n+1];
    nbr[nnz] (neighbour seed ids); farea[nnz] (facet areas); stats (dict)."""

    def __init__(self, vol, mg, nbr_ptr, nbr, farea, stats, status):
        self.vol, self.mg, self.nbr_ptr, self.nbr, self.farea = vol, mg, nbr_ptr, nbr, farea
        self.stats = stats
        self.status = status

    @property
    def centroid(self):
        return self.mg / self.vol[:, None]

    def neighbors(self, i):
        return self.nbr[self.nbr_ptr[i]:self.nbr_ptr[i + 1]]

    def facet_areas(self, i):
        return self.farea[self.nbr_ptr[i]:self.nbr_ptr[i + 1]]


class Context:
    """A device context (workspace reused across calls)."""

    def __init__(self, device: int = 0, lib=None):
        self.lib = lib if lib is not None else _lib.gpu()
        self.device = device
        self.ctx = self.lib.create(device)

    @classmethod
    def emulated(cls):
        """CPU emulation of the kernels (test-suite only)."""
        return cls(0, _lib.emu())

    def close(self):
        if self.ctx:
            self.lib.destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self, dev):
        """Device arrays: the library waits for the work queued before the
        call on torch's current stream (vgc_set_stream)."""
        if dev and self.lib.prefix == "vgc_":
            import torch
            self.lib.cdll.vgc_set_stream(self.ctx,
                                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))

    def _check(self, rc, allow_partial=True):
        if rc == VGC_OK or (rc == VGC_PARTIAL and allow_partial):
            return rc
        raise VgcError(rc, self.lib.last_error(self.ctx))

    # ---------------------------------------------------------------- cells
    def cells(self, points, weights=None, box=UNIT_BOX, k0=0, neighbors=True,
              inst_n=None, boxes=None, count_work=False, autotune=False, waves4=False,
              periodic=False, fp64_first=False):
        """Voronoi (weights None) or Laguerre cells clipped to box.

        For a batch of instances give inst_n (seed counts, consecutive in
        points) and boxes (6 per instance).  autotune: pick the first pass's
        work-group size by timing (VGC_AUTOTUNE; results do not depend on it).
        waves4: run the first pass with 4-wavefront work-groups (VGC_WAVES4).
        periodic: the box is one period of a torus (VGC_PERIODIC).
        fp64_first: first pass with FP64 corner state (VGC_FP64_FIRST; the
        default FP32 pass gives identical results)."""
        dev = _is_torch_cuda(points)
        self._stream(dev)
        if dev:
            import torch
            pts = points.contiguous().to(torch.float64).reshape(-1, 3)
            w = None if weights is None else weights.contiguous().to(torch.float64)
            n = pts.shape[0]
            vol = torch.empty(n, dtype=torch.float64, device=pts.device)
            mg = torch.empty((n, 3), dtype=torch.float64, device=pts.device)
            ptr = torch.empty(n + 1, dtype=torch.int64, device=pts.device)
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
            n = pts.shape[0]
            vol = np.empty(n)
            mg = np.empty((n, 3))
            ptr = np.empty(n + 1, dtype=np.uint64)
        flags = ((VGC_DEVICE_PTRS if dev else 0) | (VGC_COUNT_WORK if count_work else 0)
                 | (_lib.VGC_AUTOTUNE if autotune else 0) | (_lib.VGC_WAVES4 if waves4 else 0)
                 | (_lib.VGC_PERIODIC if periodic else 0) | (_lib.VGC_FP64_FIRST if fp64_first else 0))
        st = Stats()
        if inst_n is None:
            boxa = np.ascontiguousarray(box, dtype=np.float64)
            call = lambda nb, fa, cap: self.lib.fn("cells")(
                self.ctx, _ptr(pts), _ptr(w), n, boxa.ctypes.data_as(ctypes.c_void_p), int(k0),
                flags | (0 if neighbors else VGC_NO_NEIGHBORS), _ptr(vol), _ptr(mg), _ptr(ptr),
                nb, fa, cap, ctypes.byref(st))
        else:
            inst = np.ascontiguousarray(inst_n, dtype=np.uint32)
            boxa = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1)
            call = lambda nb, fa, cap: self.lib.fn("cells_batch")(
                self.ctx, _ptr(pts), _ptr(w), inst.ctypes.data_as(ctypes.c_void_p), len(inst),
                boxa.ctypes.data_as(ctypes.c_void_p), int(k0),
                flags | (0 if neighbors else VGC_NO_NEIGHBORS), _ptr(vol), _ptr(mg), _ptr(ptr),
                nb, fa, cap, ctypes.byref(st))
        if not neighbors:
            rc = self._check(call(None, None, 0))
            return CellResult(vol, mg, ptr, None, None, st.as_dict(), rc)
        # capacity guess (about 16 neighbours per cell), retry on VGC_ERR_CAPACITY
        cap = 24 * n + 64
        for _ in range(2):
            if dev:
                import torch
                nbr = torch.empty(cap, dtype=torch.int32, device=pts.device)
                fa = torch.empty(cap, dtype=torch.float64, device=pts.device)
            else:
                nbr = np.empty(cap, dtype=np.int32)
                fa = np.empty(cap)
            rc = call(_ptr(nbr), _ptr(fa), cap)
            if rc == _lib.VGC_ERR_CAPACITY:
                cap = int(st.nb_neighbors)
                continue
            self._check(rc)
            nnz = int(st.nb_neighbors)
            return CellResult(vol, mg, ptr, nbr[:nnz], fa[:nnz], st.as_dict(), rc)
        raise VgcError(rc, self.lib.last_error(self.ctx))

    # ------------------------------------------------------- meshes / RDT
    def cell_mesh(self, points, weights=None, box=UNIT_BOX, shrink=0.0, tets=False):
        """The cells as polyhedra (ConvexCell::append_to_mesh; vgc_cell_mesh).
        Returns a dict: cell_vptr, cell_fptr (n + 1 each), vertices (V, 3),
        facet_ptr (F + 1), facet_vertices, facet_nbr (-1: box), and with
        tets=True the primal tetrahedra (T, 4).  Host arrays."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        n = pts.shape[0]
        boxa = np.ascontiguousarray(box, dtype=np.float64)
        need = _lib.MeshSizes()
        st = Stats()
        f = self.lib.fn("cell_mesh")
        rc = f(self.ctx, _ptr(pts), _ptr(w), n, boxa.ctypes.data_as(ctypes.c_void_p), float(shrink),
               0, None, None, None, None, None, None, None, None, ctypes.byref(need),
               ctypes.byref(st))
        self._check(rc, allow_partial=False)
        cap = _lib.MeshSizes(need.nb_vertices, need.nb_facets, need.nb_facet_vertices,
                             need.nb_tets)
        out = {"cell_vptr": np.empty(n + 1, np.uint64), "cell_fptr": np.empty(n + 1, np.uint64),
               "vertices": np.empty((need.nb_vertices, 3)),
               "facet_ptr": np.empty(need.nb_facets + 1, np.uint64),
               "facet_vertices": np.empty(need.nb_facet_vertices, np.int32),
               "facet_nbr": np.empty(need.nb_facets, np.int32)}
        if tets:
            out["tets"] = np.empty((need.nb_tets, 4), np.int32)
        rc = f(self.ctx, _ptr(pts), _ptr(w), n, boxa.ctypes.data_as(ctypes.c_void_p), float(shrink),
               0, _ptr(out["cell_vptr"]), _ptr(out["cell_fptr"]), _ptr(out["vertices"]),
               _ptr(out["facet_ptr"]), _ptr(out["facet_vertices"]), _ptr(out["facet_nbr"]),
               _ptr(out["tets"]) if tets else None, ctypes.byref(cap), ctypes.byref(need),
               ctypes.byref(st))
        self._check(rc)
        out["stats"] = st.as_dict()
        return out

    def rdt(self, points, weights=None, box=UNIT_BOX):
        """Primal tetrahedra of the cells (compute_RDT, volumetric; vgc_rdt):
        a (T, 4) int32 array of seed indices."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        boxa = np.ascontiguousarray(box, dtype=np.float64)
        nt = _lib._u64(0)
        st = Stats()
        f = self.lib.fn("rdt")
        rc = f(self.ctx, _ptr(pts), _ptr(w), pts.shape[0], boxa.ctypes.data_as(ctypes.c_void_p), 0,
               None, 0, ctypes.byref(nt), ctypes.byref(st))
        if rc != _lib.VGC_ERR_CAPACITY:
            self._check(rc, allow_partial=False)
        t = np.empty((nt.value, 4), np.int32)
        rc = f(self.ctx, _ptr(pts), _ptr(w), pts.shape[0], boxa.ctypes.data_as(ctypes.c_void_p), 0,
               _ptr(t), nt.value, ctypes.byref(nt), ctypes.byref(st))
        self._check(rc)
        return t

    # --------------------------------------------------------- surface RVD
    def surface_centroids(self, points, vertices, triangles, vertex_weights=None):
        """Voronoi cells of the seeds restricted to a triangle mesh
        (compute_centroids_on_surface, RVD.cpp:371-415; vgc_surface_centroids):
        returns (mg (n, 3), m (n,), stats); mg / m is each restricted cell's
        centroid.  vertex_weights: a density per vertex, linear on each
        triangle.  Arrays on the GPU (torch) stay there."""
        dev = _is_torch_cuda(points)
        self._stream(dev)
        if dev:
            import torch
            pts = points.contiguous().to(torch.float64).reshape(-1, 3)
            V = vertices.contiguous().to(torch.float64).reshape(-1, 3)
            F = triangles.contiguous().to(torch.int32).reshape(-1, 3)
            vw = None if vertex_weights is None else vertex_weights.contiguous().to(torch.float64)
            n = pts.shape[0]
            mg = torch.empty((n, 3), dtype=torch.float64, device=pts.device)
            m = torch.empty(n, dtype=torch.float64, device=pts.device)
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            V = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
            F = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
            vw = (None if vertex_weights is None
                  else np.ascontiguousarray(vertex_weights, dtype=np.float64))
            n = pts.shape[0]
            mg = np.empty((n, 3))
            m = np.empty(n)
        st = _lib.SurfaceStats()
        rc = self.lib.fn("surface_centroids")(self.ctx, _ptr(pts), n, _ptr(V), V.shape[0], _ptr(F),
                                              F.shape[0], _ptr(vw),
                                              VGC_DEVICE_PTRS if dev else 0, _ptr(mg), _ptr(m),
                                              ctypes.byref(st))
        self._check(rc, allow_partial=False)
        return mg, m, st.as_dict()

    # ------------------------------------------------------------------ kNN
    def knn(self, points, k):
        """k nearest other points of every point, exact (FP64 squared
        distances, ties by index), with the coincident-point rule of
        Delaunay_NearestNeighbors (vgc_knn).  Returns (idx int32 (n, k),
        d2 float64 (n, k)); rows that end early are padded with -1."""
        dev = _is_torch_cuda(points)
        self._stream(dev)
        if dev:
            import torch
            pts = points.contiguous().to(torch.float64).reshape(-1, 3)
            n = pts.shape[0]
            idx = torch.empty((n, k), dtype=torch.int32, device=pts.device)
            d2 = torch.empty((n, k), dtype=torch.float64, device=pts.device)
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            n = pts.shape[0]
            idx = np.empty((n, k), dtype=np.int32)
            d2 = np.empty((n, k), dtype=np.float64)
        rc = self.lib.fn("knn")(self.ctx, _ptr(pts), n, int(k), VGC_DEVICE_PTRS if dev else 0,
                                _ptr(idx), _ptr(d2))
        self._check(rc, allow_partial=False)
        return idx, d2

    def knn_search(self, points, queries, k):
        """k nearest points of each query point (NearestNeighborSearch::
        get_nearest_neighbors by coordinates, vgc_knn_search): exact FP64
        squared distances, ascending, ties by index.  Returns (idx, d2),
        each (nq, k)."""
        dev = _is_torch_cuda(points)
        self._stream(dev)
        if dev:
            import torch
            pts = points.contiguous().to(torch.float64).reshape(-1, 3)
            q = queries.contiguous().to(torch.float64).reshape(-1, 3)
            nq = q.shape[0]
            idx = torch.empty((nq, k), dtype=torch.int32, device=pts.device)
            d2 = torch.empty((nq, k), dtype=torch.float64, device=pts.device)
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            q = np.ascontiguousarray(queries, dtype=np.float64).reshape(-1, 3)
            nq = q.shape[0]
            idx = np.empty((nq, k), dtype=np.int32)
            d2 = np.empty((nq, k), dtype=np.float64)
        rc = self.lib.fn("knn_search")(self.ctx, _ptr(pts), pts.shape[0], _ptr(q), nq, int(k),
                                       VGC_DEVICE_PTRS if dev else 0, _ptr(idx), _ptr(d2))
        self._check(rc, allow_partial=False)
        return idx, d2

    # ------------------------------------------------------------------- OT
    def ot_solve(self, points, nu, w0=None, box=UNIT_BOX, tol=1e-4, max_newton=100,
                 judge_start=False, verbose=False, inexact=False, periodic=False, repair=False):
        """Laguerre weights w with cell masses nu (uniform density on box).
        Returns (w, stats); w0 is the start (None: zeros).  judge_start: keep
        w0 only if its mass residual beats the zero start's.  verbose: one
        line per Newton step on stderr.  inexact: CG tolerance of each Newton
        step from its mass error (VGC_OT_INEXACT)."""
        dev = _is_torch_cuda(points)
        self._stream(dev)
        if dev:
            import torch
            pts = points.contiguous().to(torch.float64).reshape(-1, 3)
            nu_ = nu.contiguous().to(torch.float64)
            w = (torch.zeros(pts.shape[0], dtype=torch.float64, device=pts.device) if w0 is None
                 else w0.detach().clone().to(torch.float64).contiguous())
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            nu_ = np.ascontiguousarray(nu, dtype=np.float64)
            w = np.zeros(pts.shape[0]) if w0 is None else np.array(w0, dtype=np.float64, copy=True)
        boxa = np.ascontiguousarray(box, dtype=np.float64)
        st = OtStats()
        rc = self.lib.fn("ot_solve")(self.ctx, _ptr(pts), _ptr(nu_), _ptr(w), pts.shape[0],
                                     boxa.ctypes.data_as(ctypes.c_void_p), float(tol),
                                     int(max_newton),
                                     (VGC_DEVICE_PTRS if dev else 0)
                                     | (_lib.VGC_OT_JUDGE_START if judge_start else 0)
                                     | (_lib.VGC_OT_REPAIR if repair else 0)
                                     | (_lib.VGC_OT_VERBOSE if verbose else 0)
                                     | (_lib.VGC_OT_INEXACT if inexact else 0)
                                     | (_lib.VGC_PERIODIC if periodic else 0),
                                     ctypes.byref(st))
        self._check(rc)
        return w, st.as_dict()

    def ot_solve_levels(self, points, nu, levels, box=UNIT_BOX, tol=1e-4, max_newton=100,
                        verbose=False, inexact=False, repair=False, convex=False):
        """Multilevel transport (optimize_levels): points in level order,
        levels = [0, ..., n] (see brio_levels).  convex: start each level from
        the convex extension of the previous plan (VGC_OT_CONVEX_LEVELS)
        instead of copied weights.  Returns (w, per-level stats)."""
        dev = _is_torch_cuda(points)
        self._stream(dev)
        if dev:
            import torch
            pts = points.contiguous().to(torch.float64).reshape(-1, 3)
            nu_ = nu.contiguous().to(torch.float64)
            w = torch.zeros(pts.shape[0], dtype=torch.float64, device=pts.device)
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            nu_ = np.ascontiguousarray(nu, dtype=np.float64)
            w = np.zeros(pts.shape[0])
        lv = np.ascontiguousarray(levels, dtype=np.uint32)
        nl = len(lv) - 1
        boxa = np.ascontiguousarray(box, dtype=np.float64)
        sts = (OtStats * max(nl, 1))()
        rc = self.lib.fn("ot_solve_levels")(self.ctx, _ptr(pts), _ptr(nu_), _ptr(w), pts.shape[0],
                                            boxa.ctypes.data_as(ctypes.c_void_p),
                                            lv.ctypes.data_as(ctypes.c_void_p), nl, float(tol),
                                            int(max_newton),
                                            (VGC_DEVICE_PTRS if dev else 0)
                                            | (_lib.VGC_OT_VERBOSE if verbose else 0)
                                            | (_lib.VGC_OT_INEXACT if inexact else 0)
                                            | (_lib.VGC_OT_REPAIR if repair else 0)
                                            | (_lib.VGC_OT_CONVEX_LEVELS if convex else 0), sts)
        self._check(rc)
        return w, [sts[k].as_dict() for k in range(nl)]

    def ot_solve_batch(self, points, nu, inst_n, boxes, w0=None, tol=1e-4, max_newton=100,
                       verbose=False, autotune=False, inexact=False):
        """Independent transport problems at once (consecutive seeds, one box
        each; each instance's nu sums to its box volume).  Returns (w, list of
        per-instance stats)."""
        dev = _is_torch_cuda(points)
        self._stream(dev)
        if dev:
            import torch
            pts = points.contiguous().to(torch.float64).reshape(-1, 3)
            nu_ = nu.contiguous().to(torch.float64)
            w = (torch.zeros(pts.shape[0], dtype=torch.float64, device=pts.device) if w0 is None
                 else w0.detach().clone().to(torch.float64).contiguous())
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            nu_ = np.ascontiguousarray(nu, dtype=np.float64)
            w = np.zeros(pts.shape[0]) if w0 is None else np.array(w0, dtype=np.float64, copy=True)
        ia = np.ascontiguousarray(inst_n, dtype=np.uint32)
        ba = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
        if ia.sum() != pts.shape[0] or len(ba) != len(ia):
            raise ValueError("inst_n must add up to the seed count, one box per instance")
        st = (OtStats * len(ia))()
        flags = ((VGC_DEVICE_PTRS if dev else 0) | (_lib.VGC_OT_VERBOSE if verbose else 0)
                 | (_lib.VGC_AUTOTUNE if autotune else 0) | (_lib.VGC_OT_INEXACT if inexact else 0))
        rc = self.lib.fn("ot_solve_batch")(self.ctx, _ptr(pts), _ptr(nu_), _ptr(w),
                                           ia.ctypes.data_as(ctypes.c_void_p), len(ia),
                                           ba.ctypes.data_as(ctypes.c_void_p), float(tol),
                                           int(max_newton), flags, st)
        self._check(rc)
        return w, [s.as_dict() for s in st]

    # ------------------------------------------------- learned warm start
    def ot_features(self, points, nu, box=UNIT_BOX):
        """Per-seed MLP inputs (n x 6, float32) of the learned OT warm start."""
        dev = _is_torch_cuda(points)
        self._stream(dev)
        if dev:
            import torch
            pts = points.contiguous().to(torch.float64).reshape(-1, 3)
            nu_ = nu.contiguous().to(torch.float64)
            f = torch.empty((pts.shape[0], _lib.VGC_MLP_FEATURES), dtype=torch.float32,
                            device=pts.device)
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            nu_ = np.ascontiguousarray(nu, dtype=np.float64)
            f = np.empty((pts.shape[0], _lib.VGC_MLP_FEATURES), dtype=np.float32)
        boxa = np.ascontiguousarray(box, dtype=np.float64)
        rc = self.lib.fn("ot_features")(self.ctx, _ptr(pts), _ptr(nu_), pts.shape[0],
                                        boxa.ctypes.data_as(ctypes.c_void_p),
                                        VGC_DEVICE_PTRS if dev else 0, _ptr(f))
        self._check(rc, allow_partial=False)
        return f

    def ot_convex_start(self, points, pieces, s):
        """Start weights w_i = |p_i|^2 - s log sum_j exp(piece_j(p_i)/s)
        (vgc_ot_convex_start); pieces is (m, 9): a, g[3], mu, b[3], R."""
        dev = _is_torch_cuda(points)
        self._stream(dev)
        if dev:
            import torch
            pts = points.contiguous().to(torch.float64).reshape(-1, 3)
            pc = pieces.contiguous().to(device=pts.device, dtype=torch.float64).reshape(-1, 9)
            w = torch.empty(pts.shape[0], dtype=torch.float64, device=pts.device)
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            pc = np.ascontiguousarray(pieces, dtype=np.float64).reshape(-1, 9)
            w = np.empty(pts.shape[0])
        rc = self.lib.fn("ot_convex_start")(self.ctx, _ptr(pts), pts.shape[0], _ptr(pc),
                                            pc.shape[0], float(s), VGC_DEVICE_PTRS if dev else 0,
                                            _ptr(w))
        self._check(rc, allow_partial=False)
        return w

    def mlp_forward(self, params, feats):
        dev = _is_torch_cuda(feats)
        self._stream(dev)
        n = feats.shape[0]
        if dev:
            import torch
            out = torch.empty(n, dtype=torch.float32, device=feats.device)
        else:
            params = np.ascontiguousarray(params, dtype=np.float32)
            feats = np.ascontiguousarray(feats, dtype=np.float32)
            out = np.empty(n, dtype=np.float32)
        rc = self.lib.fn("mlp_forward")(self.ctx, _ptr(params), _ptr(feats), n,
                                        VGC_DEVICE_PTRS if dev else 0, _ptr(out))
        self._check(rc, allow_partial=False)
        return out

    def mlp_grad(self, params, feats, target, grad=None):
        """Gradient of the mean squared error of the MLP over the rows;
        returns (grad, loss)."""
        dev = _is_torch_cuda(feats)
        self._stream(dev)
        n = feats.shape[0]
        if dev:
            import torch
            if grad is None:
                grad = torch.empty(_lib.VGC_MLP_PARAMS, dtype=torch.float32, device=feats.device)
        else:
            params = np.ascontiguousarray(params, dtype=np.float32)
            feats = np.ascontiguousarray(feats, dtype=np.float32)
            target = np.ascontiguousarray(target, dtype=np.float32)
            grad = np.empty(_lib.VGC_MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        rc = self.lib.fn("mlp_grad")(self.ctx, _ptr(params), _ptr(feats), _ptr(target), n,
                                     VGC_DEVICE_PTRS if dev else 0, _ptr(grad), ctypes.byref(loss))
        self._check(rc, allow_partial=False)
        return grad, loss.value

    def mlp_adam(self, params, grad, m, v, step, lr):
        """One in-place Adam step on device tensors (params, m, v)."""
        self._stream(True)
        rc = self.lib.fn("mlp_adam")(self.ctx, _ptr(params), _ptr(grad), _ptr(m), _ptr(v),
                                     int(step), float(lr), VGC_DEVICE_PTRS)
        self._check(rc, allow_partial=False)

    # ---------------------------------------------------------------- Lloyd
    def cvt_func_grad(self, points, box=UNIT_BOX, periodic=False):
        """CVT energy f = sum_i integral over cell i of |x - p_i|^2 and its
        gradient g_i = 2 (m_i p_i - mg_i) (compute_CVT_func_grad_in_volume,
        RVD.cpp:792-899).  Returns (f, g); g is an (n, 3) array of the same
        kind as points."""
        dev = _is_torch_cuda(points)
        self._stream(dev)
        if dev:
            import torch
            x = points.detach().to(torch.float64).contiguous().reshape(-1, 3)
            g = torch.zeros_like(x)
        else:
            x = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            g = np.zeros_like(x)
        f = ctypes.c_double(0.0)
        st = Stats()
        boxa = np.ascontiguousarray(box, dtype=np.float64)
        rc = self.lib.fn("cvt_func_grad")(self.ctx, _ptr(x), x.shape[0],
                                          boxa.ctypes.data_as(ctypes.c_void_p),
                                          (VGC_DEVICE_PTRS if dev else 0)
                                          | (_lib.VGC_PERIODIC if periodic else 0),
                                          ctypes.byref(f), _ptr(g),
                                          ctypes.byref(st))
        self._check(rc)
        return f.value, g

    def cvt_newton(self, points, iters, m=7, box=UNIT_BOX, periodic=False):
        """Quasi-Newton CVT (Newton_iterations, CVT.cpp:272-302): L-BFGS with m
        pairs and a More-Thuente line search on the CVT energy; returns the
        new points (same kind of array) and the optimiser's statistics."""
        dev = _is_torch_cuda(points)
        self._stream(dev)
        if dev:
            import torch
            x = points.detach().clone().to(torch.float64).contiguous().reshape(-1, 3)
        else:
            x = np.array(points, dtype=np.float64, copy=True).reshape(-1, 3)
        st = _lib.OptStats()
        boxa = np.ascontiguousarray(box, dtype=np.float64)
        rc = self.lib.fn("cvt_newton")(self.ctx, _ptr(x), x.shape[0],
                                       boxa.ctypes.data_as(ctypes.c_void_p), int(iters), int(m),
                                       (VGC_DEVICE_PTRS if dev else 0)
                                       | (_lib.VGC_PERIODIC if periodic else 0), ctypes.byref(st))
        self._check(rc)
        return x, st.as_dict()

    def lloyd(self, points, iters, box=UNIT_BOX, k0=0, inst_n=None, boxes=None, autotune=False,
              periodic=False):
        """iters Lloyd iterations; returns the new points (same kind of array).
        inst_n / boxes: several independent instances at once (consecutive
        seeds, one box each).  autotune: pick the first pass's work-group size
        by timing (results do not depend on it)."""
        dev = _is_torch_cuda(points)
        self._stream(dev)
        if dev:
            import torch
            x = points.detach().clone().to(torch.float64).contiguous().reshape(-1, 3)
        else:
            x = np.array(points, dtype=np.float64, copy=True).reshape(-1, 3)
        st = Stats()
        flags = ((VGC_DEVICE_PTRS if dev else 0) | (_lib.VGC_AUTOTUNE if autotune else 0)
                 | (_lib.VGC_PERIODIC if periodic else 0))
        if inst_n is None:
            boxa = np.ascontiguousarray(box, dtype=np.float64)
            rc = self.lib.fn("lloyd")(self.ctx, _ptr(x), x.shape[0],
                                      boxa.ctypes.data_as(ctypes.c_void_p), int(iters), int(k0),
                                      flags, ctypes.byref(st))
        else:
            ia = np.ascontiguousarray(inst_n, dtype=np.uint32)
            ba = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
            if ia.sum() != x.shape[0] or len(ba) != len(ia):
                raise ValueError("inst_n must add up to the seed count, one box per instance")
            rc = self.lib.fn("lloyd_batch")(self.ctx, _ptr(x), ia.ctypes.data_as(ctypes.c_void_p),
                                            len(ia), ba.ctypes.data_as(ctypes.c_void_p),
                                            int(iters), int(k0), flags, ctypes.byref(st))
        self._check(rc)
        return x, st.as_dict()


_default = {}


def default_context(device: int = 0) -> Context:
    if device not in _default:
        _default[device] = Context(device)
    return _default[device]


def cells(points, weights=None, box=UNIT_BOX, k0=0, neighbors=True, device=0):
    return default_context(device).cells(points, weights, box, k0, neighbors)


def knn(points, k, device=0):
    return default_context(device).knn(points, k)


def knn_search(points, queries, k, device=0):
    return default_context(device).knn_search(points, queries, k)


def lloyd(points, iters, box=UNIT_BOX, k0=0, device=0):
    return default_context(device).lloyd(points, iters, box, k0)


def ot_solve(points, nu, w0=None, box=UNIT_BOX, tol=1e-4, max_newton=100, device=0):
    return default_context(device).ot_solve(points, nu, w0, box, tol, max_newton)


def brio_levels(n: int, ratio: float = 0.125, min_level: int = 2000, seed: int = 0):
    """A random order of n seeds and its levels for ot_solve_levels, as a
    biased randomized insertion order cuts it: the last level holds all n, each
    level before it a fraction `ratio` of the next, down to min_level seeds.
    Returns (permutation, levels) with levels[0] = 0 and levels[-1] = n."""
    perm = np.random.default_rng(seed).permutation(n)
    sizes = [n]
    while int(sizes[-1] * ratio) >= min_level:
        sizes.append(int(sizes[-1] * ratio))
    return perm, np.array([0] + sizes[::-1], dtype=np.uint32)


def random_float64(seed: int, count: int) -> np.ndarray:
    """std::mt19937_64(seed) + uniform_real_distribution<double>(0,1) — the
    stream of Numeric::random_reset / random_float64 (basic/numeric.cpp:76-103)."""
    out = np.empty(count)
    rc = _lib.gpu().cdll.vgc_random_float64(seed, out.ctypes.data_as(ctypes.c_void_p), count)
    if rc != VGC_OK:
        raise VgcError(rc, "vgc_random_float64")
    return out


def sample_points(kind: int, seed: int, n: int) -> np.ndarray:
    """Synthetic seeds (0: uniform cube, 1: half uniform + 8-Gaussian mixture,
    2: 16-Gaussian mixture), see include/vgc.h."""
    out = np.empty((n, 3))
    rc = _lib.gpu().cdll.vgc_sample_points(kind, seed, n, out.ctypes.data_as(ctypes.c_void_p))
    if rc != VGC_OK:
        raise VgcError(rc, "vgc_sample_points")
    return out


# ------------------------------------------------------- reference mirrors

class NearestNeighborSearch:
    """points/nn_search.h: set_points, then exact nearest neighbours (FP64
    squared distances, ascending, ties by index) of query points
    (get_nearest_neighbors by coordinates, nn_search.h:129-134, batched over
    the rows of the query array) or of stored points by index (the point
    itself included, nn_search.h:180-185)."""

    def __init__(self, dimension: int = 3, ctx: Context | None = None):
        if dimension != 3:
            raise ValueError("only dimension 3 is supported")
        self.ctx = ctx
        self.points = None

    def set_points(self, points):
        self.points = points

    def nb_points(self):
        return self.points.shape[0]

    def get_nearest_neighbors(self, nb_neighbors: int, query_points):
        """query_points: (nq, 3) coordinates, or an int / int array of stored
        point indices.  Returns (neighbors, sq_dist), each (nq, nb_neighbors)."""
        ctx = self.ctx or default_context()
        q = query_points
        if isinstance(q, (int, np.integer)) or (isinstance(q, np.ndarray) and q.dtype.kind in "iu"):
            q = self.points[np.atleast_1d(q)]
        return ctx.knn_search(self.points, q, nb_neighbors)

    def get_nearest_neighbor(self, query_point):
        idx, _ = self.get_nearest_neighbors(1, np.asarray(query_point, dtype=np.float64).reshape(1, 3))
        return int(idx[0, 0])

    def get_neighbors_of_all(self, nb_neighbors: int):
        """Every stored point's nearest others with the coincident-point rule
        of Delaunay_NearestNeighbors (delaunay/delaunay_nn.cpp:105-143)."""
        ctx = self.ctx or default_context()
        return ctx.knn(self.points, nb_neighbors)


class RestrictedVoronoiDiagram:
    """Restricted Voronoi diagram of seeds (voronoi/RVD.h): in a box domain
    (volumetric), or on a triangulated surface when surface=(vertices,
    triangles[, vertex_weights]) is given (volumetric() false).
    compute_centroids accumulates into caller-zeroed arrays like
    compute_centroids_in_volume / _on_surface (RVD.cpp:500-528, 371-415)."""

    def __init__(self, points, box=UNIT_BOX, ctx: Context | None = None, k0: int = 64,
                 surface=None):
        self.points = points
        self.box = box
        self.ctx = ctx
        self.k0 = k0
        self.check_SR = True
        self.surface = surface

    def volumetric(self) -> bool:
        return self.surface is None

    def compute_centroids_on_surface(self, mg, m):
        if self.surface is None:
            raise ValueError("no surface: construct with surface=(vertices, triangles)")
        ctx = self.ctx or default_context()
        V, F = self.surface[0], self.surface[1]
        vw = self.surface[2] if len(self.surface) > 2 else None
        g, mm, st = ctx.surface_centroids(self.points, V, F, vw)
        mg += g.reshape(mg.shape)
        m += mm
        return st

    def set_check_SR(self, x: bool):
        if not x:
            raise ValueError("the GPU path always applies the security radius")

    def compute_centroids(self, mg, m):
        if not self.volumetric():
            return self.compute_centroids_on_surface(mg, m)
        ctx = self.ctx or default_context()
        r = ctx.cells(self.points, None, self.box, self.k0, neighbors=False)
        mg += r.mg.reshape(mg.shape)
        m += r.vol
        return r

    def compute_RDT(self):
        """RVD.h:504-510 (volumetric): the primal tetrahedra, (T, 4) seed
        indices, oriented by orient_3d."""
        ctx = self.ctx or default_context()
        return ctx.rdt(self.points, None, self.box)

    def compute_RVD_cells(self, shrink: float = 0.0):
        """The cells as polyhedra (the cube mode of compute_RVD, ConvexCell::
        append_to_mesh); see Context.cell_mesh."""
        ctx = self.ctx or default_context()
        return ctx.cell_mesh(self.points, None, self.box, shrink)

    def compute_CVT_func_grad(self, g):
        """RestrictedVoronoiDiagram::compute_CVT_func_grad_in_volume
        (RVD.h:318): returns the energy and adds its gradient to g (n x 3)."""
        ctx = self.ctx or default_context()
        f, gg = ctx.cvt_func_grad(self.points, self.box)
        g += gg.reshape(g.shape)
        return f


class CentroidalVoronoiTesselation:
    """voronoi/CVT.h: set_points, Lloyd_iterations, embedding (box domain)."""

    def __init__(self, box=UNIT_BOX, ctx: Context | None = None, k0: int = 64):
        self.box = box
        self.ctx = ctx
        self.k0 = k0
        self.points_ = None
        self.last_stats = None

    def set_points(self, points):
        self.points_ = np.array(points, dtype=np.float64).reshape(-1, 3) \
            if not _is_torch_cuda(points) else points.clone()

    def nb_points(self):
        return self.points_.shape[0]

    def embedding(self, i):
        return self.points_[i]

    def Lloyd_iterations(self, nb_iter: int):
        ctx = self.ctx or default_context()
        self.points_, self.last_stats = ctx.lloyd(self.points_, nb_iter, self.box, self.k0)

    def Newton_iterations(self, nb_iter: int, m: int = 7):
        """voronoi/CVT.h:164: quasi-Newton (L-BFGS, m pairs) on the CVT energy."""
        ctx = self.ctx or default_context()
        self.points_, self.last_stats = ctx.cvt_newton(self.points_, nb_iter, m, self.box)


class OptimalTransportMap3d:
    """Semi-discrete optimal transport from the uniform density on a box to
    Dirac masses at the seeds — the calls compute_OTM makes on Exploragram's
    OptimalTransportMap3d (examples/exploragram/compute_OTM/main.cpp:283-297):
    set_points, set_epsilon, optimize.  The source domain is a box here
    (the reference takes a tetrahedral mesh)."""

    def __init__(self, box=UNIT_BOX, ctx: Context | None = None):
        self.box = box
        self.ctx = ctx
        self.epsilon = 1e-4
        self.points_ = None
        self.nu = None
        self.weights = None
        self.last_stats = None
        self.learned_params = None
        self.coarse_ratio = None

    def set_points(self, points, masses=None):
        """Seeds and their target masses (default: equal, summing to the box volume)."""
        self.points_ = points
        n = points.shape[0]
        vol = float(np.prod(np.asarray(self.box[3:]) - np.asarray(self.box[:3])))
        if masses is None:
            if _is_torch_cuda(points):
                import torch
                masses = torch.full((n,), vol / n, dtype=torch.float64, device=points.device)
            else:
                masses = np.full(n, vol / n)
        self.nu = masses
        self.weights = None

    def set_epsilon(self, eps: float):
        """Target max relative mass error of a cell."""
        self.epsilon = eps

    def set_initial_weights(self, w0):
        self.weights = w0

    def use_learned_warm_start(self, params=None):
        """Start Newton from the weights predicted by the trained MLP
        (geogram_amd.warmstart); the zero start stays the fallback when the
        prediction leaves a cell empty or has a larger mass residual.
        params: None = the packaged network, False = off."""
        if params is False:
            self.learned_params = None
            return
        from . import warmstart
        self.learned_params = warmstart.load_params() if params is None else params

    def use_coarse_warm_start(self, ratio=256):
        """Start Newton from a coarse transport plan (warmstart.coarse_start:
        a convex start, no empty cell), judged against the zero start.
        ratio: fine seeds per coarse seed; None = off."""
        self.coarse_ratio = ratio

    def optimize(self, max_iter: int = 100):
        ctx = self.ctx or default_context()
        w0, judge = self.weights, False
        if w0 is None and getattr(self, "coarse_ratio", None):
            from . import warmstart
            w0, _ = warmstart.coarse_start(ctx, self.points_, self.nu, self.box,
                                           ratio=self.coarse_ratio)
            judge = True
        elif w0 is None and self.learned_params is not None:
            from . import warmstart
            w0 = warmstart.predict_weights(ctx, self.points_, self.nu, self.box,
                                           self.learned_params)
            judge = True
        self.weights, self.last_stats = ctx.ot_solve(self.points_, self.nu, w0, self.box,
                                                     self.epsilon, max_iter, judge_start=judge)
        return self.weights

    def optimize_levels(self, levels, max_iter: int = 100, convex: bool = False):
        """optimize_levels (compute_OTM/main.cpp:294): the points set with
        set_points are in level order, levels = [0, ..., n].  convex: each
        level starts from the convex extension of the previous plan."""
        ctx = self.ctx or default_context()
        self.weights, self.last_stats = ctx.ot_solve_levels(self.points_, self.nu, levels,
                                                            self.box, self.epsilon, max_iter,
                                                            convex=convex)
        return self.weights
