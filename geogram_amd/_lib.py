"""Loader of the C ABI (include/vgc.h).

``libvgc.so`` is the gfx950 build and the only library the package uses.  If it
is missing, or no GPU is visible, calls raise: there is no CPU fallback.
``emu()`` loads ``libvgc_emu.so``, the CPU emulation of the same kernels, which
only the test-suite uses.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvgc.so")
EMU_PATH = os.path.join(_HERE, "libvgc_emu.so")

VGC_OK = 0
VGC_PARTIAL = 1
VGC_ERR_ARG = -1
VGC_ERR_DEVICE = -2
VGC_ERR_CAPACITY = -3
VGC_ERR_INTERNAL = -4

VGC_DEVICE_PTRS = 1
VGC_NO_NEIGHBORS = 2
VGC_COUNT_WORK = 4
VGC_OT_JUDGE_START = 8
VGC_OT_VERBOSE = 16
VGC_AUTOTUNE = 32
VGC_OT_INEXACT = 64
VGC_WAVES4 = 128
VGC_PERIODIC = 256
VGC_OT_REPAIR = 512
VGC_OT_CONVEX_LEVELS = 2048
VGC_FP64_FIRST = 1024

VGC_MLP_PARAMS = 4673
VGC_START_PIECE = 9
VGC_MLP_FEATURES = 6

_STAT_U64 = ("nb_cells nb_neighbors nb_pass_b nb_pass_c nb_pass_e nb_uncertain nb_failed nb_empty "
             "nb_disconnected nb_clips nb_tests radius_entries").split()
_STAT_F64 = "t_build_ms t_knn_ms t_clip_ms t_pass_bc_ms t_compact_ms t_total_ms".split()
_STAT_F64B = "t_clip_a_gpu_ms t_knn_gpu_ms".split()
_STAT_U64B = ("a_cells a_tests a_live a_new a_tets a_facets nb_fail_capacity why_rings why_walk "
              "why_narrow nb_regather tuned_waves nb_periodic_retries nb_pieces").split()


class Stats(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_uint64) for k in _STAT_U64] + [(k, ctypes.c_double) for k in _STAT_F64]
                + [(k, ctypes.c_double) for k in _STAT_F64B] + [(k, ctypes.c_uint64) for k in _STAT_U64B]
                + [("nb_fp64_redo", ctypes.c_uint64), ("t_redo_gpu_ms", ctypes.c_double)])

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OtStats(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_uint32) for k in ("newton_iters", "cg_iters", "cell_passes", "halvings",
                                                 "used_fallback_start", "converged",
                                                 "repair_rounds", "pad")]
                + [(k, ctypes.c_double) for k in ("err0", "final_err", "t_total_ms", "t_cells_ms",
                                                   "t_cg_ms")])

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MeshSizes(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("nb_vertices", "nb_facets", "nb_facet_vertices",
                                                 "nb_tets")]


class SurfaceStats(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_uint64) for k in ("nb_items", "nb_split", "nb_redo", "nb_records")]
                + [("k_final", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("t_ms", ctypes.c_double)])

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class OptStats(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_uint32) for k in ("iterations", "evaluations")]
                + [("line_search_info", ctypes.c_int32), ("failed_cells", ctypes.c_uint32)]
                + [(k, ctypes.c_double) for k in ("f0", "f", "gnorm0", "gnorm", "t_total_ms")])

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class VgcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vgc error {code}: {msg}")
        self.code = code


_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64


def _declare(L, prefix):
    ctx_t = _vp
    L[prefix + "last_error"].restype = ctypes.c_char_p
    L[prefix + "last_error"].argtypes = [ctx_t]
    f = L[prefix + "knn"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _u32, _u32, _u32, _vp, _vp]
    f = L[prefix + "knn_search"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _u32, _vp, _u32, _u32, _u32, _vp, _vp]
    f = L[prefix + "cells"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _vp, _u32, _vp, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _u64,
                  ctypes.POINTER(Stats)]
    f = L[prefix + "cells_batch"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _vp, _vp, _u32, _vp, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _u64,
                  ctypes.POINTER(Stats)]
    f = L[prefix + "lloyd"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _u32, _vp, _u32, _u32, _u32, ctypes.POINTER(Stats)]
    f = L[prefix + "cvt_func_grad"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _u32, _vp, _u32, _vp, _vp, ctypes.POINTER(Stats)]
    f = L[prefix + "cell_mesh"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _vp, _u32, _vp, ctypes.c_double, _u32, _vp, _vp, _vp, _vp, _vp, _vp,
                  _vp, ctypes.POINTER(MeshSizes), ctypes.POINTER(MeshSizes), ctypes.POINTER(Stats)]
    f = L[prefix + "rdt"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _vp, _u32, _vp, _u32, _vp, _u64, ctypes.POINTER(_u64),
                  ctypes.POINTER(Stats)]
    f = L[prefix + "surface_centroids"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp,
                  ctypes.POINTER(SurfaceStats)]
    f = L[prefix + "cvt_newton"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _u32, _vp, _u32, _u32, _u32, ctypes.POINTER(OptStats)]
    f = L[prefix + "lloyd_batch"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _vp, _u32, _vp, _u32, _u32, _u32, ctypes.POINTER(Stats)]
    f = L[prefix + "ot_solve"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _vp, _vp, _u32, _vp, ctypes.c_double, _u32, _u32,
                  ctypes.POINTER(OtStats)]
    f = L[prefix + "ot_solve_levels"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _vp, _vp, _u32, _vp, _vp, _u32, ctypes.c_double, _u32, _u32, _vp]
    f = L[prefix + "ot_solve_batch"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _vp, _vp, _vp, _u32, _vp, ctypes.c_double, _u32, _u32,
                  ctypes.POINTER(OtStats)]
    f = L[prefix + "ot_features"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _vp, _u32, _vp, _u32, _vp]
    f = L[prefix + "ot_convex_start"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _u32, _vp, _u32, ctypes.c_double, _u32, _vp]
    f = L[prefix + "mlp_forward"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _vp, _u32, _u32, _vp]
    f = L[prefix + "mlp_grad"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _vp, _vp, _u32, _u32, _vp, ctypes.POINTER(ctypes.c_double)]
    f = L[prefix + "mlp_adam"]
    f.restype = ctypes.c_int
    f.argtypes = [ctx_t, _vp, _vp, _vp, _vp, _u32, ctypes.c_float, _u32]


class _Lib:
    """Uniform view of libvgc (prefix vgc_) or libvgc_emu (prefix vgc_emu_)."""

    def __init__(self, path, prefix):
        if not os.path.exists(path):
            raise ImportError(
                f"{os.path.basename(path)} is not built; run "
                "`python -c 'import __graft_entry__; __graft_entry__.build()'`")
        self.path = path
        self.cdll = ctypes.CDLL(path)
        self.prefix = prefix
        self._f = {}
        names = ["last_error", "knn", "knn_search", "cells", "cells_batch", "lloyd", "lloyd_batch",
                 "cvt_func_grad", "cvt_newton", "cell_mesh", "rdt", "surface_centroids", "ot_solve", "ot_solve_levels",
                 "ot_solve_batch", "ot_features", "ot_convex_start",
                 "mlp_forward", "mlp_grad", "mlp_adam", "create", "destroy"]
        for nm in names:
            self._f[prefix + nm] = getattr(self.cdll, prefix + nm)
        _declare(self._f, prefix)
        if prefix == "vgc_":
            self._f["vgc_create"].argtypes = [ctypes.POINTER(_vp), ctypes.c_int]
            self.cdll.vgc_random_float64.argtypes = [_u64, _vp, _u64]
            self.cdll.vgc_sample_points.argtypes = [_u32, _u64, _u32, _vp]
            self.cdll.vgc_set_stream.argtypes = [_vp, _vp]
        else:
            self._f["vgc_emu_create"].argtypes = [ctypes.POINTER(_vp)]
        self._f[prefix + "destroy"].argtypes = [_vp]
        self._f[prefix + "destroy"].restype = None

    def fn(self, name):
        return self._f[self.prefix + name]

    def create(self, device=0):
        ctx = _vp()
        if self.prefix == "vgc_":
            rc = self.fn("create")(ctypes.byref(ctx), int(device))
        else:
            rc = self.fn("create")(ctypes.byref(ctx))
        if rc != VGC_OK:
            raise VgcError(rc, f"cannot create a context on device {device} "
                               "(no GPU visible, or HIP runtime failure)")
        return ctx

    def destroy(self, ctx):
        self.fn("destroy")(ctx)

    def last_error(self, ctx):
        s = self.fn("last_error")(ctx)
        return s.decode() if s else ""


_gpu = None
_emu = None


def _bind_torch_runtime():
    """torch ships its own HIP runtime (torch/lib/libamdhip64.so).  A process
    can drive the GPU through one HIP runtime only: if libvgc.so were loaded
    first it would map /opt/rocm's, and torch's would then find no GPU.
    Importing torch first makes the dynamic loader bind libvgc.so to the
    runtime already mapped (same SONAME), so both share one runtime and torch
    device pointers are valid in the library."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def gpu() -> _Lib:
    global _gpu
    if _gpu is None:
        _bind_torch_runtime()
        _gpu = _Lib(LIB_PATH, "vgc_")
    return _gpu


def emu() -> _Lib:
    global _emu
    if _emu is None:
        _emu = _Lib(EMU_PATH, "vgc_emu_")
    return _emu
