"""C-ABI boundary (CPU): the libraries load, every entry point declared in include/*.h is exported,
and argument errors come back as status codes with a message (no crash, outputs untouched)."""

import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import repulsor_amd as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "repulsor_amd", "lib", "librepulsor_hip.so")


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rp_[a-z_0-9]+)\s*\(", text)))


def test_all_declared_symbols_exported():
    names = _declared("repulsor_hip.h")
    assert "rp_tp0_compute" in names and "rp_multiply_metric" in names
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (rp_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_library_loads_and_reports_version():
    lib = C.CDLL(LIB)
    lib.rp_version.restype = C.c_char_p
    assert b"gfx950" in lib.rp_version()


def _lib():
    lib = C.CDLL(LIB)
    lib.rp_last_error.restype = C.c_char_p
    return lib


def test_null_context_errors():
    lib = _lib()
    d = C.POINTER(C.c_double)
    Y = np.full((4, 3), 7.0)
    lib.rp_multiply_metric.argtypes = [C.c_void_p, C.c_double, d, C.c_int64, C.c_double, d, C.c_int64, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_int]
    st = lib.rp_multiply_metric(None, 1.0, Y.ctypes.data_as(d), 3, 0.0, Y.ctypes.data_as(d), 3, 3, 1, 1, 1, 0)
    assert st == R.RP_EINVAL and b"ctx" in lib.rp_last_error()
    # nrhs != 3 is rejected before anything is touched
    st = lib.rp_multiply_metric(None, 1.0, Y.ctypes.data_as(d), 3, 0.0, Y.ctypes.data_as(d), 3, 2, 1, 1, 1, 0)
    assert st == R.RP_EINVAL
    assert np.all(Y == 7.0)
    lib.rp_tp0_compute.argtypes = [C.c_void_p, C.c_uint, C.c_double, C.c_double, C.c_int, d]
    assert lib.rp_tp0_compute(None, 1, 6.0, 12.0, 0, None) == R.RP_ENOTREADY
    lib.rp_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    assert lib.rp_ctx_create(0, None) == R.RP_EINVAL


def test_context_creation_status():
    """Without a HIP device rp_ctx_create returns RP_EHIP with a message; with one it succeeds."""
    from conftest import gpu_available
    lib = _lib()
    ctx = C.c_void_p()
    lib.rp_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    st = lib.rp_ctx_create(0, C.byref(ctx))
    if gpu_available():
        assert st == R.RP_OK and ctx.value
        lib.rp_ctx_destroy.argtypes = [C.c_void_p]
        assert lib.rp_ctx_destroy(ctx) == R.RP_OK
    else:
        assert st == R.RP_EHIP and len(lib.rp_last_error()) > 0 and not ctx.value


def test_host_api_fails_loudly_without_device():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("device present: covered by the GPU tests")
    X, S = R.make_torus(12, 8, 1.0, 0.4, 0.0, 1)
    M = R.SimplicialMesh(X, S)
    M.cluster_tree_settings.split_threshold = 4
    M.build()
    E = R.TangentPointEnergy0(6.0, 12.0)
    with pytest.raises(R.RepulsorError):
        E.Value(M)
