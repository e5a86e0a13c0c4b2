"""ORACLE -- test infrastructure only.

Python binding of the CPU restatement of the reference's tangent-point path (oracle/tp0_oracle.cpp,
built into oracle/_build/liboracle.so by oracle/Makefile). Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module, as the checker or the timed CPU baseline.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name):
    path = os.path.join(_HERE, "_build", name)
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: run make -C oracle")
    return C.CDLL(path)


_lib = _load("liboracle.so")
# Release-flag build of the same restatement (-O3 -ffast-math), timed as the CPU baseline.


def fast_tp0():
    """orc_tp0_compute from the -O3 -ffast-math build (CPU baseline timing only)."""
    return fast("orc_tp0_compute")


_fast_lib = None


def fast(name):
    """A function of the -O3 -ffast-math build with the signature of the checker build's (CPU
    baseline timing only)."""
    global _fast_lib
    if _fast_lib is None:
        _fast_lib = _load("liboracle_fast.so")
    f = getattr(_fast_lib, name)
    ref = getattr(_lib, name)
    f.restype = ref.restype
    f.argtypes = ref.argtypes
    return f


def fast_mt(name, threads):
    """The threaded variant (<name>_mt: rows split by cost over `threads` std::threads, per-thread
    accumulators) of a function of the -O3 -ffast-math build, called with the checker build's signature
    (CPU baseline timing only)."""
    global _fast_lib
    if _fast_lib is None:
        _fast_lib = _load("liboracle_fast.so")
    f = getattr(_fast_lib, name + "_mt")
    ref = getattr(_lib, name)
    f.restype = ref.restype
    f.argtypes = list(ref.argtypes) + [C.c_int]
    return lambda *args: f(*args, int(threads))
_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)


def _p(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


def _f(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_simplex = _f("orc_simplex_data", None, [C.c_int, C.c_int64, _d, _i32, _d, _d, _d, _d, _i32])
_tp0 = _f("orc_tp0_compute", C.c_double, [C.c_int, C.c_int64, C.c_int64, _d, _d, _i64, _i32, _i64, _i64, _i32, _i64,
                                          C.c_double, C.c_double, C.c_int, _d, _d, _d, _d, _d, _d])
_collect = _f("orc_collect_derivatives", None, [C.c_int, C.c_int64, C.c_int64, _i32, _i32, _i32, _i32, _i32, _d, _d, _d, _d])
_hulls = _f("orc_dnearfar_to_hulls", None, [C.c_int, C.c_int64, _d, _i32, _d, _d, _d, _d])
_assemble = _f("orc_assemble", None, [C.c_int, C.c_int64, C.c_int64, _i32, _d, C.c_double, C.c_double, _d, C.c_int64])
_mm = _f("orc_multiply_metric", None, [C.c_int, C.c_int64, C.c_int64, C.c_int64, _i32, _i32, _i32, _i32, _i32, _d,
                                       _i64, _i32, _d, _d, _i64, _i32, _d, _d,
                                       C.c_double, _d, C.c_int64, C.c_double, _d, C.c_int64, C.c_int, C.c_int])
_allpairs = _f("orc_allpairs_energy", C.c_double, [C.c_int, C.c_int64, _d, C.c_double, C.c_double])
_allpairs_metric = _f("orc_allpairs_metric", None, [C.c_int, C.c_int64, _d, C.c_double, C.c_double, _d])


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def c64i(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def simplex_data(V, S, charges=None):
    """SimplexDataKernel restatement: near (T x ND), far (T x 10), diffop (T x 3 x (n+1)), sorted simplices."""
    V, S = c64(V), c32(S)
    n = S.shape[1] - 1
    T = S.shape[0]
    nd = 1 + 3 * (n + 1) + 6
    near = np.zeros((T, nd)); far = np.zeros((T, 10)); dop = np.zeros((T, 3, n + 1)); srt = np.zeros_like(S)
    _simplex(n, T, _p(V), _p(S, _i32), _p(None if charges is None else c64(charges)), _p(near), _p(far), _p(dop), _p(srt, _i32))
    return near, far, dop, srt


def tp0_compute(n, P_near, C_far, near_pat, far_pat, q, p, flags, impl=None):
    """TP0_Traversor::Compute restatement. *_pat = (rp, ci, diag) full symmetric CSR. flags 1|2|4."""
    P_near, C_far = c64(P_near), c64(C_far)
    T, Cn = P_near.shape[0], C_far.shape[0]
    nrp, nci, ndg = c64i(near_pat[0]), c32(near_pat[1]), c64i(near_pat[2])
    frp, fci, fdg = c64i(far_pat[0]), c32(far_pat[1]), c64i(far_pat[2])
    nd = P_near.shape[1]
    out = {}
    PDn = np.zeros((T, nd)) if flags & 2 else None
    CDf = np.zeros((Cn, 10)) if flags & 2 else None
    nfv = np.zeros((max(1, len(nci)), 2)) if flags & 4 else None
    ffv = np.zeros((max(1, len(fci)), 2)) if flags & 4 else None
    nfd = np.zeros((T, 2)) if flags & 4 else None
    ffd = np.zeros((Cn, 2)) if flags & 4 else None
    en = (impl or _tp0)(n, T, Cn, _p(P_near), _p(C_far), _p(nrp, _i64), _p(nci if len(nci) else np.zeros(1, np.int32), _i32), _p(ndg, _i64),
              _p(frp, _i64), _p(fci if len(fci) else np.zeros(1, np.int32), _i32), _p(fdg, _i64), q, p, flags,
              _p(PDn), _p(CDf), _p(nfv), _p(ffv), _p(nfd), _p(ffd))
    out["energy"] = en
    if flags & 2:
        out["P_D_near_tree"], out["C_D_far"] = PDn, CDf
    if flags & 4:
        out["NF"], out["FF"], out["NF_diag"], out["FF_diag"] = nfv[: len(nci)], ffv[: len(fci)], nfd, ffd
    return out


def collect_derivatives(n, tree, P_D_near_tree, C_D_far):
    T = len(tree["P_ordering"])
    Cn = len(tree["C_left"])
    nd = 1 + 3 * (n + 1) + 6
    Cd = c64(C_D_far).copy()
    PDn = np.zeros((T, nd)); PDf = np.zeros((T, 10))
    _collect(n, T, Cn, _p(c32(tree["C_left"]), _i32), _p(c32(tree["C_right"]), _i32), _p(c32(tree["C_begin"]), _i32),
             _p(c32(tree["C_end"]), _i32), _p(c32(tree["P_inverse_ordering"]), _i32), _p(c64(P_D_near_tree)), _p(Cd), _p(PDn), _p(PDf))
    return PDn, PDf


def dnearfar_to_hulls(V, S, N, F, charges=None):
    V, S = c64(V), c32(S)
    n = S.shape[1] - 1
    B = np.zeros((S.shape[0], n + 1, 3))
    _hulls(n, S.shape[0], _p(V), _p(S, _i32), _p(None if charges is None else c64(charges)), _p(c64(N)), _p(c64(F)), _p(B))
    return B


def assemble(S, nV, B, alpha=1.0, beta=0.0, Y=None):
    S = c32(S)
    n = S.shape[1] - 1
    Y = np.zeros((nV, 3)) if Y is None else Y
    _assemble(n, nV, S.shape[0], _p(S, _i32), _p(c64(B)), alpha, beta, _p(Y), Y.shape[1])
    return Y


def differential(n, V, S, tree, P_D_near_tree, C_D_far, charges=None):
    """Full Assemble_ClusterTree_Derivatives restatement: collect, chain rule, assemble."""
    N, F = collect_derivatives(n, tree, P_D_near_tree, C_D_far)
    B = dnearfar_to_hulls(V, S, N, F, charges)
    return assemble(S, len(V), B), N, F


def multiply_metric(n, nV, tree, pre_cols, pre_vals, near_pat, nf_vals, nf_diag, far_pat, ff_vals, ff_diag,
                    alpha, X, beta, Y, nf=True, ff=True, impl=None):
    T = len(tree["P_ordering"])
    Cn = len(tree["C_left"])
    X = c64(X)
    assert Y.dtype == np.float64 and Y.flags.c_contiguous
    z = np.zeros(1, np.int32)
    nci = c32(near_pat[1]); fci = c32(far_pat[1])
    (impl or _mm)(n, nV, T, Cn, _p(c32(tree["C_left"]), _i32), _p(c32(tree["C_right"]), _i32), _p(c32(tree["C_begin"]), _i32),
        _p(c32(tree["C_end"]), _i32), _p(c32(pre_cols), _i32), _p(c64(pre_vals)),
        _p(c64i(near_pat[0]), _i64), _p(nci if len(nci) else z, _i32), _p(c64(nf_vals) if len(nci) else np.zeros(2)), _p(c64(nf_diag)),
        _p(c64i(far_pat[0]), _i64), _p(fci if len(fci) else z, _i32), _p(c64(ff_vals) if len(fci) else np.zeros(2)), _p(c64(ff_diag)),
        alpha, _p(X), X.shape[1], beta, _p(Y), Y.shape[1], int(nf), int(ff))
    return Y


def allpairs_energy(n, P_far, q, p):
    P_far = c64(P_far)
    return _allpairs(n, P_far.shape[0], _p(P_far), q, p)


def allpairs_metric(n, P_far, q, p):
    P_far = c64(P_far)
    T = P_far.shape[0]
    out = np.zeros((T, T, 2))
    _allpairs_metric(n, T, _p(P_far), q, p, _p(out))
    return out


_pl = _f("orc_pl_compute", None, [C.c_int, C.c_int64, C.c_int64, _d, _d, _i64, _i32, _i64, _i64, _i32, _i64,
                                  C.c_double, _d, _d, _d, _d])
_plm = _f("orc_pl_multiply", None, [C.c_int, C.c_int64, C.c_int64, C.c_int64, _i32, _i32, _i32, _i32, _i32, _d,
                                    _i64, _i32, _d, _d, _i64, _i32, _d, _d,
                                    C.c_double, _d, C.c_int64, C.c_double, _d, C.c_int64, C.c_int, C.c_int])


def pl_compute(n, P_near, C_far, near_pat, far_pat, s, impl=None):
    """PseudoLaplacian (LowOrder) values restatement: NF (N_nf), FF (N_ff), NF_diag (T), FF_diag (C)."""
    P_near, C_far = c64(P_near), c64(C_far)
    T, Cn = P_near.shape[0], C_far.shape[0]
    nrp, nci, ndg = c64i(near_pat[0]), c32(near_pat[1]), c64i(near_pat[2])
    frp, fci, fdg = c64i(far_pat[0]), c32(far_pat[1]), c64i(far_pat[2])
    z = np.zeros(1, np.int32)
    nfv = np.zeros(max(1, len(nci))); ffv = np.zeros(max(1, len(fci)))
    nfd = np.zeros(T); ffd = np.zeros(Cn)
    (impl or _pl)(n, T, Cn, _p(P_near), _p(C_far), _p(nrp, _i64), _p(nci if len(nci) else z, _i32), _p(ndg, _i64),
        _p(frp, _i64), _p(fci if len(fci) else z, _i32), _p(fdg, _i64), s, _p(nfv), _p(ffv), _p(nfd), _p(ffd))
    return {"NF": nfv[: len(nci)], "FF": ffv[: len(fci)], "NF_diag": nfd, "FF_diag": ffd}


def pl_multiply(n, nV, tree, pre_cols, pre_vals, near_pat, pl, far_pat, alpha, X, beta, Y, nf=True, ff=True, impl=None):
    T = len(tree["P_ordering"])
    Cn = len(tree["C_left"])
    X = c64(X)
    assert Y.dtype == np.float64 and Y.flags.c_contiguous
    z = np.zeros(1, np.int32)
    nci = c32(near_pat[1]); fci = c32(far_pat[1])
    (impl or _plm)(n, nV, T, Cn, _p(c32(tree["C_left"]), _i32), _p(c32(tree["C_right"]), _i32), _p(c32(tree["C_begin"]), _i32),
         _p(c32(tree["C_end"]), _i32), _p(c32(pre_cols), _i32), _p(c64(pre_vals)),
         _p(c64i(near_pat[0]), _i64), _p(nci if len(nci) else z, _i32), _p(c64(pl["NF"]) if len(nci) else np.zeros(1)),
         _p(c64(pl["NF_diag"])),
         _p(c64i(far_pat[0]), _i64), _p(fci if len(fci) else z, _i32), _p(c64(pl["FF"]) if len(fci) else np.zeros(1)),
         _p(c64(pl["FF_diag"])),
         alpha, _p(X), X.shape[1], beta, _p(Y), Y.shape[1], int(nf), int(ff))
    return Y


_vf = _f("orc_vf_compute", C.c_double, [C.c_int, C.c_int64, _d, _i64, _i32, _i64, C.c_double, C.c_int, C.c_double, C.c_double,
                                        C.c_int, _d, _d, _d, _d])
_plvf = _f("orc_pl_vf", None, [C.c_int, C.c_int64, _d, _i64, _i32, _i64, C.c_double, C.c_int, C.c_double, _d, _d, _d])


def vf_compute(n, P_near, vf_pat, theta, max_refinement, q, p, flags):
    """TP0 very-near field restatement. vf_pat = (rp, ci, diag) full symmetric CSR. flags 1|2|4, +8 for
    the w-weighted energy (a pinning aid, not the reference's sum). Returns energy, P_D_near_tree
    (T x ND, VF part only), VF (nnz x 2), VF_diag (T x 2), evaluations, max_level."""
    P_near = c64(P_near)
    T, nd = P_near.shape
    rp, ci, dg = c64i(vf_pat[0]), c32(vf_pat[1]), c64i(vf_pat[2])
    PDn = np.zeros((T, nd)); vals = np.zeros((max(1, len(ci)), 2)); diag = np.zeros((T, 2)); st = np.zeros(2)
    en = _vf(n, T, _p(P_near), _p(rp, _i64), _p(ci if len(ci) else np.zeros(1, np.int32), _i32), _p(dg, _i64),
             float(theta), int(max_refinement), float(q), float(p), int(flags), _p(PDn), _p(vals), _p(diag), _p(st))
    return {"energy": en, "P_D_near_tree": PDn, "VF": vals[: len(ci)], "VF_diag": diag,
            "evaluations": int(st[0]), "max_level": int(st[1])}


def pl_vf(n, P_near, vf_pat, theta, max_refinement, s):
    """Pseudo-Laplacian very-near field restatement: VF (nnz), VF_diag (T)."""
    P_near = c64(P_near)
    T = P_near.shape[0]
    rp, ci, dg = c64i(vf_pat[0]), c32(vf_pat[1]), c64i(vf_pat[2])
    vals = np.zeros(max(1, len(ci))); diag = np.zeros(T); st = np.zeros(2)
    _plvf(n, T, _p(P_near), _p(rp, _i64), _p(ci if len(ci) else np.zeros(1, np.int32), _i32), _p(dg, _i64),
          float(theta), int(max_refinement), float(s), _p(vals), _p(diag), _p(st))
    return {"VF": vals[: len(ci)], "VF_diag": diag, "evaluations": int(st[0])}


def merge_patterns(pat_a, vals_a, pat_b, vals_b):
    """Row-wise union of two disjoint CSR patterns with their values (for applying NF and VF blocks in
    one near-field product). Returns ((rp, ci, diag), vals); columns sorted per row."""
    rpa, cia = np.asarray(pat_a[0]), np.asarray(pat_a[1])
    rpb, cib = np.asarray(pat_b[0]), np.asarray(pat_b[1])
    vals_a, vals_b = np.asarray(vals_a), np.asarray(vals_b)
    T = len(rpa) - 1
    rp = [0]; ci = []; vv = []; dg = []
    for i in range(T):
        cols = np.concatenate([cia[rpa[i]:rpa[i + 1]], cib[rpb[i]:rpb[i + 1]]])
        vals = np.concatenate([vals_a[rpa[i]:rpa[i + 1]], vals_b[rpb[i]:rpb[i + 1]]])
        o = np.argsort(cols, kind="stable")
        cols, vals = cols[o], vals[o]
        dg.append(rp[-1] + int(np.searchsorted(cols, i, side="right")))
        ci.append(cols); vv.append(vals); rp.append(rp[-1] + len(cols))
    ci = np.concatenate(ci).astype(np.int32) if ci else np.zeros(0, np.int32)
    vv = np.concatenate(vv) if vv else np.zeros((0,) + vals_a.shape[1:])
    return (np.array(rp, np.int64), ci, np.array(dg, np.int64)), vv


_sd2 = _f("orc_simplex_dist2", C.c_double, [C.c_int, _d, _d])


def simplex_dist2(P, Q):
    """Squared distance of two disjoint simplices (2 or 3 points each, rows = points)."""
    P, Q = c64(P), c64(Q)
    return _sd2(P.shape[0], _p(P), _p(Q))


_ssp = _f("orc_safe_step_pair", C.c_double, [C.c_int, _d, _d, _d, _d, C.c_double, C.c_double])
_mhm = _f("orc_moving_hulls_meet", C.c_int, [C.c_int, _d, _d, _d, _d, C.c_double, C.c_double])
_mss = _f("orc_max_safe_step", C.c_double, [C.c_int, C.c_int64, C.c_int64, _d, _d, _i32, C.c_double, C.c_double, _d])


def safe_step_pair(p, u, q, v, tinit, eps=0.0625):
    """The reference's bisection for one pair of moving simplices (rows = points)."""
    p, u, q, v = c64(p), c64(u), c64(q), c64(v)
    return _ssp(p.shape[0], _p(p), _p(u), _p(q), _p(v), float(tinit), float(eps))


def moving_hulls_meet(p, u, q, v, a, b):
    p, u, q, v = c64(p), c64(u), c64(q), c64(v)
    return bool(_mhm(p.shape[0], _p(p), _p(u), _p(q), _p(v), float(a), float(b)))


def max_safe_step(X, U, S, t_max, eps=0.0625):
    """MaximumSafeStepSize over all non-adjacent pairs (brute force over pairs with meeting swept boxes).
    Returns (t, candidate pairs, hull tests)."""
    X, U, S = c64(X), c64(U), c32(S)
    st = np.zeros(2)
    t = _mss(S.shape[1] - 1, X.shape[0], S.shape[0], _p(X), _p(U), _p(S, _i32), float(t_max), float(eps), _p(st))
    return t, int(st[0]), int(st[1])
