"""GPU triangular solves with hand-built supernodal factors (rp_upload_h1_factor + rp_h1_solve through
the C-ABI), against a dense solve of H = L L^T.

* The first factor is shaped so that its lowest level is a wide level (supernodes wider than 128
  columns) whose longest supernode has more than 2560 rows while another has between 512 and 2560: the
  backward kernel of such a level must size its staging buffer for the tasks that stage v at once and
  chunk the others. One supernode's tasks fall on both sides of that rule (r - k above and below 2560).
* The second has a root supernode of 1536 columns (supernodes may be up to 2048 wide, as chains of the
  elimination tree are amalgamated) above two children of 300 and 364 columns.
"""

import ctypes as C
import os

import numpy as np
import pytest

import repulsor_amd as R

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "repulsor_amd", "lib", "librepulsor_hip.so")

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)


def _ptr(a, t):
    return a.ctypes.data_as(t)


FACTORS = {
    # supernodes (first column, width, update-row ranges, parent) in postorder, and the column count
    "wide_level_long_and_medium": ([
        (0,    256, [(700, 3016)],              2),   # r = 256 + 2316 = 2572: tasks with k < 12 chunk, the rest stage at once
        (256,  200, [(456, 968), (968, 1268)],  2),   # r = 1012, on the same level
        (456,  512, [(968, 3016)],              3),   # r = 2560
        (968,  512, [(1480, 3016)],             4),
        (1480, 512, [(1992, 3016)],             5),
        (1992, 512, [(2504, 3016)],             6),
        (2504, 512, [],                         -1),
    ], 3016),
    "root_1536_columns": ([
        (0,    300, [(664, 2200)],              2),   # r = 1836
        (300,  364, [(664, 1800)],              2),   # r = 1500
        (664,  1536, [],                        -1),
    ], 2200),
}


def _factor(name):
    sn, n = FACTORS[name]
    rng = np.random.default_rng(21)
    L = np.zeros((n, n))
    sn_begin, row_ptr, rows, val_ptr, vals, parent = [0], [0], [], [0], [], []
    for f, w, upd, par in sn:
        R_s = list(range(f, f + w)) + [j for a, b in upd for j in range(a, b)]
        r = len(R_s)
        Lss = np.tril(rng.standard_normal((w, w)) * (0.3 / np.sqrt(w)), -1) + np.diag(1.0 + rng.random(w))
        Loff = rng.standard_normal((r - w, w)) * (0.3 / np.sqrt(n))
        L[f:f + w, f:f + w] = Lss
        L[np.array(R_s[w:], dtype=np.int64)[:, None], np.arange(f, f + w)[None, :]] = Loff
        blk = np.zeros((r, w))
        blk[:w] = np.linalg.inv(Lss)
        blk[w:] = Loff
        vals.append(blk.T.reshape(-1))              # column-major r x w
        rows += R_s
        sn_begin.append(f + w)
        row_ptr.append(row_ptr[-1] + r)
        val_ptr.append(val_ptr[-1] + r * w)
        parent.append(par)
    return (n, L, np.array(sn_begin, np.int32), np.array(row_ptr, np.int64), np.array(rows, np.int32),
            np.array(val_ptr, np.int64), np.concatenate(vals), np.array(parent, np.int32))


@pytest.mark.parametrize("name", list(FACTORS))
def test_h1_solve_hand_built_factor(name, require_gpu):
    n, L, sn_begin, row_ptr, rows, val_ptr, vals, parent = _factor(name)
    ns = len(parent)
    rng = np.random.default_rng(22)
    perm = rng.permutation(n).astype(np.int32)        # new -> old
    lib = C.CDLL(LIB)
    lib.rp_last_error.restype = C.c_char_p
    ctx = C.c_void_p()
    lib.rp_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    assert lib.rp_ctx_create(0, C.byref(ctx)) == R.RP_OK
    try:
        tri = np.array([[0, 1, 2]], np.int32)
        lib.rp_upload_mesh.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_int64, _i32, _d]
        assert lib.rp_upload_mesh(ctx, 2, n, 1, _ptr(tri, _i32), None) == R.RP_OK, lib.rp_last_error()
        lib.rp_upload_h1_factor.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _i32, _i32, _i64, _i32, _i64, _d, _i32]
        st = lib.rp_upload_h1_factor(ctx, n, ns, _ptr(perm, _i32), _ptr(sn_begin, _i32), _ptr(row_ptr, _i64),
                                     _ptr(rows, _i32), _ptr(val_ptr, _i64), _ptr(vals, _d), _ptr(parent, _i32))
        assert st == R.RP_OK, lib.rp_last_error()
        lib.rp_h1_solve.argtypes = [C.c_void_p, C.c_double, _d, C.c_int64, C.c_double, _d, C.c_int64, C.c_int, C.c_int]
        Hn = L @ L.T
        for call in range(2):
            B = rng.standard_normal((n, 3))
            X0 = rng.standard_normal((n, 3))
            X = X0.copy()
            st = lib.rp_h1_solve(ctx, 2.0, _ptr(B, _d), 3, 0.5, _ptr(X, _d), 3, 3, 0)
            assert st == R.RP_OK, lib.rp_last_error()
            xn = np.linalg.solve(Hn, B[perm])
            ref = np.empty_like(xn)
            ref[perm] = xn
            err = np.linalg.norm(X - (2.0 * ref + 0.5 * X0)) / np.linalg.norm(2.0 * ref + 0.5 * X0)
            print(f"call {call}: relative error {err:.2e}", flush=True)
            assert err <= 1e-11
    finally:
        lib.rp_ctx_destroy.argtypes = [C.c_void_p]
        lib.rp_ctx_destroy(ctx)
