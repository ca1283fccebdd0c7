"""The C ABI: libvgc.so loads and exports every entry point include/vgc.h
declares; the C++ mirror (include/vgc_geogram.hpp) compiles and links against
it.  No compute on the CPU: without a GPU, context creation must fail cleanly."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "geogram_amd", "libvgc.so")


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "vgc.h")).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(vgc_\w+)\s*\(", src, re.M)))


def test_exports_every_declared_entry_point():
    names = declared_symbols()
    assert {"vgc_create", "vgc_destroy", "vgc_last_error", "vgc_knn", "vgc_cells",
            "vgc_cells_batch", "vgc_lloyd", "vgc_ot_solve", "vgc_random_float64",
            "vgc_sample_points", "vgc_cvt_func_grad", "vgc_lloyd_batch",
            "vgc_ot_solve_batch", "vgc_ot_solve_levels"} <= set(names)
    lib = ctypes.CDLL(LIB)
    for nm in names:
        assert hasattr(lib, nm), nm


def test_no_gpu_means_clean_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    lib = ctypes.CDLL(LIB)
    ctx = ctypes.c_void_p()
    rc = lib.vgc_create(ctypes.byref(ctx), 0)
    assert rc == -2 and not ctx.value  # VGC_ERR_DEVICE
    import geogram_amd as ga
    with pytest.raises(ga.VgcError):
        ga.Context(0)


def build_cpp_test(tmp_path):
    exe = str(tmp_path / "test_api")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "test_api.cpp"),
                    "-L", os.path.join(ROOT, "geogram_amd"), "-lvgc",
                    "-Wl,-rpath," + os.path.join(ROOT, "geogram_amd"), "-o", exe], check=True)
    return exe


def test_cpp_mirror_builds_and_fails_cleanly_without_gpu(tmp_path):
    import torch

    exe = build_cpp_test(tmp_path)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible (tests/test_gpu_cells.py runs the program)")
    assert subprocess.run([exe]).returncode == 3


@pytest.mark.gpu
def test_cpp_mirror_runs_on_gpu(tmp_path):
    exe = build_cpp_test(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout, r.stderr)
    assert r.returncode == 0


REFERENCE_LIB = "/root/reference/src/lib"


@pytest.mark.skipif(not os.path.isdir(REFERENCE_LIB), reason="geogram sources not present")
def test_geogram_shim_compiles_against_reference_headers():
    """geogram_amd/shim/geogram_vgc.cpp — the translation unit a geogram
    maintainer adds (NearestNeighborSearch "VGC", RestrictedVoronoiDiagram
    and CentroidalVoronoiTesselation subclasses, the OptimalTransportMap3d
    calls) — compiles against geogram's own headers with every pure virtual
    implemented."""
    shim = os.path.join(ROOT, "geogram_amd", "shim", "geogram_vgc.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I", REFERENCE_LIB, "-I", os.path.join(ROOT, "include"), shim],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
