"""The allocation-free energy entry point (rp_tp0_energy_noalloc) against rp_tp0_compute(RP_ENERGY),
both called through the C-ABI on the same device context: the energies must be equal bit for bit (same
kernels, same summation order). Two calls in a row on one device share its scratch array, the second on
a smaller mesh; a mesh whose very-near records exceed the scratch is refused with RP_EINVAL."""

import ctypes as C
import os

import numpy as np
import pytest

import repulsor_amd as R
from repulsor_amd import H
from helpers import vf_shells

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "repulsor_amd", "lib", "librepulsor_hip.so")
_d = C.POINTER(C.c_double)
RP_DEVICE_PTR = 1          # pointer-location flag of include/repulsor_hip.h


def _lib():
    lib = C.CDLL(LIB)
    lib.rp_last_error.restype = C.c_char_p
    lib.rp_tp0_compute.argtypes = [C.c_void_p, C.c_uint, C.c_double, C.c_double, C.c_int, _d]
    lib.rp_tp0_energy_noalloc.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_int]
    return lib


def _context(M):
    ctx = H.rpm_mesh_context(M._h)
    assert ctx, H.rpm_last_error()
    return ctx


def _energies(lib, ctx, q, p, precision):
    ref = C.c_double(0.0)
    assert lib.rp_tp0_compute(ctx, 1, q, p, precision, C.byref(ref)) == R.RP_OK, lib.rp_last_error()
    host = C.c_double(-1.0)
    assert lib.rp_tp0_energy_noalloc(ctx, q, p, precision, C.addressof(host), 0) == R.RP_OK, lib.rp_last_error()
    # device output
    hip = C.CDLL("libamdhip64.so")
    dptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(dptr), C.c_size_t(8)) == 0
    try:
        assert lib.rp_tp0_energy_noalloc(ctx, q, p, precision, dptr, RP_DEVICE_PTR) == R.RP_OK, lib.rp_last_error()
        dev = C.c_double(-2.0)
        assert hip.hipMemcpy(C.byref(dev), dptr, C.c_size_t(8), 2) == 0    # hipMemcpyDeviceToHost
    finally:
        hip.hipFree(dptr)
    return ref.value, host.value, dev.value


def test_noalloc_energy_equals_tp0_compute(require_gpu):
    lib = _lib()
    # first call: two nested shells with a very-near field (records in the scratch array)
    M1 = vf_shells(3, 0.004)
    b1 = M1.build()
    assert b1["verynear_nnz"] > 0
    ctx1 = _context(M1)
    for prec in (0, 1):
        ref, host, dev = _energies(lib, ctx1, 6.0, 12.0, prec)
        print(f"shells precision {prec}: rp_tp0_compute {ref!r} noalloc {host!r} / {dev!r}", flush=True)
        assert host == ref and dev == ref
    # second call, smaller mesh, same device (the scratch keeps the first call's contents)
    X, S = R.make_torus(40, 16, 1.0, 0.5, 1e-3, 1)
    M2 = R.SimplicialMesh(X, S)
    M2.cluster_tree_settings.split_threshold = 16
    M2.build()
    ctx2 = _context(M2)
    refs = {}
    for prec in (0, 1):
        ref, host, dev = _energies(lib, ctx2, 6.0, 12.0, prec)
        print(f"torus_1280 precision {prec}: rp_tp0_compute {ref!r} noalloc {host!r} / {dev!r}", flush=True)
        assert host == ref and dev == ref
        refs[prec] = ref
    # the product path (FP64) agrees with the direct call
    assert R.TangentPointEnergy0(6.0, 12.0).Value(M2) == refs[0]


def test_noalloc_energy_refuses_oversized_scratch(require_gpu):
    lib = _lib()
    M = vf_shells(4, 0.001)
    b = M.build()
    need = 24 * b["verynear_nnz"]
    assert need > 2 * 1024 * 1024, need
    ctx = _context(M)
    out = C.c_double(123.0)
    st = lib.rp_tp0_energy_noalloc(ctx, 6.0, 12.0, 0, C.addressof(out), 0)
    msg = lib.rp_last_error().decode()
    print(f"oversized: status {st}: {msg}", flush=True)
    assert st == R.RP_EINVAL and "scratch" in msg
    assert out.value == 123.0
