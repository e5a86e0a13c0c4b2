"""MI355X tangent-point library: Python bindings of the header-only C++ host API.

The classes mirror the reference's C++ surface for the tangent-point hot path
(HenrikSchumacher/Repulsor: SimplicialMesh, TangentPointEnergy0, TangentPointMetric0).
Compute runs in librepulsor_hip.so (gfx950 HIP kernels behind a C-ABI); the host side
(cluster tree, block cluster tree) runs in librepulsor_host.so. Both are built in-tree by
``__graft_entry__.build()`` (``make -C repulsor_amd/csrc``). Importing this package fails
loudly when the libraries are missing; there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")

RP_OK, RP_EINVAL, RP_EHIP, RP_ENOMEM, RP_ESTAGNATION, RP_EUNSUPPORTED, RP_ENOTREADY = range(7)
RP_SOLVE_NO_PRECONDITIONER, RP_SOLVE_PC_NORM, RP_SOLVE_GMRES, RP_SOLVE_RECORD = 1, 2, 4, 8
PRECISIONS = {"fp64": 0, "mixed": 1, "mixed-refined": 2, "bf16": 3}


class RepulsorError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


def _load():
    hip_path = os.path.join(LIB_DIR, "librepulsor_hip.so")
    host_path = os.path.join(LIB_DIR, "librepulsor_host.so")
    for p in (hip_path, host_path):
        if not os.path.exists(p):
            raise ImportError(
                f"{p} is missing: build the HIP extension first "
                "(python -c 'import __graft_entry__ as g; g.build()' or make -C repulsor_amd/csrc)")
    hip = C.CDLL(hip_path, mode=C.RTLD_GLOBAL)
    host = C.CDLL(host_path, mode=C.RTLD_GLOBAL)
    return hip, host


_hip, _host = _load()

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_u32 = C.POINTER(C.c_uint32)
_vp = C.c_void_p


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_host_sigs = {
    "rpm_last_error": (C.c_char_p, []),
    "rpm_gen_torus": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64, _d, _i32, _i64]),
    "rpm_gen_icosphere": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_int, _d, _i32, _i64]),
    "rpm_gen_trefoil": (C.c_int, [C.c_int64, C.c_int, C.c_double, C.c_uint64, _d, _i32, _i64]),
    "rpm_mesh_new": (_vp, [C.c_int, _d, C.c_int64, _i32, C.c_int64, _d, C.c_int]),
    "rpm_mesh_free": (None, [_vp]),
    "rpm_mesh_settings": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]),
    "rpm_mesh_build": (C.c_int, [_vp, _i64]),
    "rpm_mesh_tree": (C.c_int, [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _i32]),
    "rpm_mesh_tree_data": (C.c_int, [_vp, _d, _d, _d, _d, _d, _i32, _d]),
    "rpm_mesh_pattern": (C.c_int, [_vp, C.c_int, _i64, _i32, _i64, _i64, C.c_int32]),
    "rpm_mesh_leaf_blocks": (C.c_int, [_vp, _i32, _i32, _i32]),
    "rpm_mesh_tiles": (C.c_int, [_vp, _i64, _i32, C.POINTER(C.c_uint32), _i64]),
    "rpm_mesh_build_device_ms": (C.c_int, [_vp, _d]),
    "rpm_mesh_semistatic_update": (C.c_int, [_vp, _d, C.c_int]),
    "rpm_mesh_coords": (C.c_int, [_vp, _d]),
    "rpm_mesh_device": (C.c_int, [_vp]),
    "rpm_mesh_timing": (C.c_int, [_vp, C.c_int, _d]),
    "rpm_mesh_context": (_vp, [_vp]),
    "rpm_op_new": (_vp, [C.c_double, C.c_double, C.c_int]),
    "rpm_op_free": (None, [_vp]),
    "rpm_energy_value": (C.c_int, [_vp, _vp, _d]),
    "rpm_energy_differential": (C.c_int, [_vp, _vp, C.c_double, C.c_double, _d, C.c_int64, _d, C.c_int]),
    "rpm_energy_moment_derivatives": (C.c_int, [_vp, _vp, _d, _d]),
    "rpm_metric_values": (C.c_int, [_vp, _vp, _d, _d, _d, _d]),
    "rpm_metric_multiply": (C.c_int, [_vp, _vp, C.c_double, _d, C.c_int64, C.c_double, _d, C.c_int64,
                                      C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "rpm_metric_assemble": (C.c_int, [_vp, _vp]),
    "rpm_timer_report": (C.c_int, [C.c_char_p, C.c_int64]),
    "rpm_h1_solve_host": (C.c_int, [_vp, _d, _d, _d]),
    "rpm_h1_multiply_host": (C.c_int, [_vp, _d, _d]),
    "rpm_nd_ordering": (C.c_int, [_vp, _i32]),
    "rpm_h1_csr": (C.c_int, [_vp, _i64, _i32, _d, _i64]),
    "rpm_metric_precondition": (C.c_int, [_vp, _vp, C.c_double, _d, C.c_int64, C.c_double, _d, C.c_int64, C.c_int, C.c_int]),
    "rpm_metric_solve": (C.c_int, [_vp, _vp, C.c_double, _d, C.c_int64, C.c_double, _d, C.c_int64, C.c_int, C.c_int,
                                   C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), _d]),
    "rpm_pl_values": (C.c_int, [_vp, _vp, _d, _d, _d, _d]),
    "rpm_mesh_device_state_rows": (C.c_int, [_vp, _d, _d]),
    "rpm_mesh_max_safe_step": (C.c_int, [_vp, _d, C.c_double, C.c_double, C.c_int, _d, _i64]),
    "rpm_vf_metric_values": (C.c_int, [_vp, _vp, _d, _d]),
    "rpm_pl_vf_values": (C.c_int, [_vp, _vp, _d, _d]),
    "rpm_vf_stats": (C.c_int, [_vp, _i64, C.POINTER(C.c_int)]),
    "rpm_pl_multiply": (C.c_int, [_vp, _vp, C.c_double, _d, C.c_int64, C.c_double, _d, C.c_int64, C.c_int]),
    "rpm_h1_solve_device": (C.c_int, [_vp, C.c_double, _d, C.c_int64, C.c_double, _d, C.c_int64, C.c_int]),
    "rpm_op_set_pc_weight": (None, [_vp, C.c_double]),
    "rpm_op_set_lp_guard": (None, [_vp, C.c_int, C.c_double]),
    "rpm_op_set_pc_precision": (None, [_vp, C.c_int]),
    "rpm_op_set_learned": (None, [_vp, C.c_int, C.POINTER(C.c_float), C.c_int]),
    "rpm_learned_feature_count": (C.c_int, []),
    "rpm_learned_features": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "rpm_pc_learn_step": (C.c_int, [_vp, _vp, C.c_float, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.POINTER(C.c_float), _d, _d]),
    "rpm_pc_scaling": (C.c_int, [_vp, _vp, _d]),
    "rpm_pc_weight": (C.c_int, [_vp, _vp, _d]),
    "rpm_metric_diagonal": (C.c_int, [_vp, _vp, _d]),
}
H = type("H", (), {})()
for _n, (_r, _a) in _host_sigs.items():
    setattr(H, _n, _sig(_host, _n, _r, _a))
_sig(_hip, "rp_last_error", C.c_char_p, [])
_sig(_hip, "rp_version", C.c_char_p, [])


def _p(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


def _check(st):
    if st != RP_OK:
        raise RepulsorError(st, H.rpm_last_error().decode(errors="replace"))


def version() -> str:
    return _hip.rp_version().decode()


def timer_report() -> str:
    n = H.rpm_timer_report(None, 0)
    buf = C.create_string_buffer(n + 1)
    H.rpm_timer_report(buf, n + 1)
    return buf.value.decode()


# ---------------------------------------------------------------------------------------- generators
def _gen(fn, *args):
    counts = np.zeros(2, dtype=np.int64)
    fn(*args, None, None, _p(counts, _i64))
    V, T = int(counts[0]), int(counts[1])
    return V, T, counts


def make_torus(Nu, Nv, R, r, sigma, seed):
    """Torus grid Nu x Nv split into 2 Nu Nv triangles, jittered by sigma N(0,1) along the normal."""
    counts = np.zeros(2, dtype=np.int64)
    H.rpm_gen_torus(Nu, Nv, R, r, sigma, seed, None, None, _p(counts, _i64))
    V, T = int(counts[0]), int(counts[1])
    X = np.zeros((V, 3)); S = np.zeros((T, 3), dtype=np.int32)
    H.rpm_gen_torus(Nu, Nv, R, r, sigma, seed, _p(X), _p(S, _i32), _p(counts, _i64))
    return X, S


def make_icosphere(level, lmax, amplitude, seed, threads=8):
    """Icosahedron subdivided `level` times, radially perturbed by seeded real spherical harmonics l <= lmax."""
    counts = np.zeros(2, dtype=np.int64)
    # The generator is deterministic; the first call only sizes the arrays.
    V = 10 * 4 ** level + 2
    T = 20 * 4 ** level
    X = np.zeros((V, 3)); S = np.zeros((T, 3), dtype=np.int32)
    H.rpm_gen_icosphere(level, lmax, amplitude, seed, threads, _p(X), _p(S, _i32), _p(counts, _i64))
    assert counts[0] == V and counts[1] == T
    return X, S


def make_trefoil(N, modes, amplitude, seed):
    """Closed trefoil polygon with N edges plus a seeded Fourier perturbation of max norm `amplitude`."""
    counts = np.zeros(2, dtype=np.int64)
    X = np.zeros((N, 3)); S = np.zeros((N, 2), dtype=np.int32)
    H.rpm_gen_trefoil(N, modes, amplitude, seed, _p(X), _p(S, _i32), _p(counts, _i64))
    return X, S


# ---------------------------------------------------------------------------------------- mesh
class _Settings:
    def __init__(self, **kw):
        self.__dict__.update(kw)


class SimplicialMesh:
    """SimplicialMesh<n,3> (reference src/SimplicialMesh.hpp). coords: V x 3, simplices: T x (n+1)."""

    def __init__(self, coords, simplices, charges=None, thread_count: Optional[int] = None, device: int = 0):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        simplices = np.ascontiguousarray(simplices, dtype=np.int32)
        self.dom_dim = simplices.shape[1] - 1
        self.V, self.T = coords.shape[0], simplices.shape[0]
        self._charges = None if charges is None else np.ascontiguousarray(charges, dtype=np.float64)
        tc = thread_count or min(16, os.cpu_count() or 1)
        self._h = H.rpm_mesh_new(self.dom_dim, _p(coords), self.V, _p(simplices, _i32), self.T, _p(self._charges), tc)
        if not self._h:
            raise RepulsorError(RP_EINVAL, H.rpm_last_error().decode())
        self.simplices = simplices
        self.device = device
        # Library defaults (reference src/Settings.hpp:5-54).
        self.cluster_tree_settings = _Settings(split_threshold=1, device_build=False)
        self.block_cluster_tree_settings = _Settings(far_field_separation_parameter=0.5)
        self.adaptivity_settings = _Settings(theta=10.0, max_refinement=30)
        self._built = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and H is not None:          # the module's globals may already be cleared at interpreter exit
            H.rpm_mesh_free(h)
            self._h = None

    def VertexCount(self):
        return self.V

    def SimplexCount(self):
        return self.T

    def _apply_settings(self):
        _check(H.rpm_mesh_settings(self._h, int(self.cluster_tree_settings.split_threshold),
                                   float(self.block_cluster_tree_settings.far_field_separation_parameter),
                                   float(self.adaptivity_settings.theta), int(self.adaptivity_settings.max_refinement),
                                   int(self.device), int(bool(self.cluster_tree_settings.device_build))))

    def build(self):
        """Builds (once) the cluster tree and the block cluster tree; returns the counts."""
        if self._built is None:
            self._apply_settings()
            c = np.zeros(12, dtype=np.int64)
            _check(H.rpm_mesh_build(self._h, _p(c, _i64)))
            keys = ["V", "T", "C", "L", "depth", "near_nnz", "far_nnz", "verynear_nnz", "inter_nnz",
                    "leaf_blocks", "tile_cols", "tile_slots"]
            self._built = {k: int(v) for k, v in zip(keys, c)}
        return self._built

    def tree(self):
        b = self.build()
        Cn, T, L = b["C"], b["T"], b["L"]
        out = {k: np.zeros(Cn, dtype=np.int32) for k in ("C_left", "C_right", "C_begin", "C_end")}
        out["P_ordering"] = np.zeros(T, dtype=np.int32)
        out["P_inverse_ordering"] = np.zeros(T, dtype=np.int32)
        out["leaf_clusters"] = np.zeros(L, dtype=np.int32)
        _check(H.rpm_mesh_tree(self._h, *[_p(out[k], _i32) for k in
                                          ("C_left", "C_right", "C_begin", "C_end", "P_ordering",
                                           "P_inverse_ordering", "leaf_clusters")]))
        return out

    def tree_data(self):
        b = self.build()
        n1 = self.dom_dim + 1
        nd = 1 + 3 * n1 + 6
        out = {
            "P_near": np.zeros((b["T"], nd)), "P_far": np.zeros((b["T"], 10)), "C_far": np.zeros((b["C"], 10)),
            "C_serialized": np.zeros((b["C"], 7)), "P_serialized": np.zeros((b["T"], 4 + 3 * n1)),
            "pre_cols": np.zeros((b["T"], n1), dtype=np.int32), "pre_vals": np.zeros((b["T"], 4, n1)),
        }
        _check(H.rpm_mesh_tree_data(self._h, _p(out["P_near"]), _p(out["P_far"]), _p(out["C_far"]),
                                    _p(out["C_serialized"]), _p(out["P_serialized"]), _p(out["pre_cols"], _i32),
                                    _p(out["pre_vals"])))
        return out

    def pattern(self, kind: str, row_end: int = -1):
        """Full symmetric CSR (rp, ci, diag) of 'near', 'far' or 'verynear' (reference BinaryMatrixCSR layout).
        For 'near', row_end >= 0 keeps only the rows [0, row_end)."""
        k = {"near": 0, "far": 1, "verynear": 2}[kind]
        self.build()
        nnz = np.zeros(1, dtype=np.int64)
        _check(H.rpm_mesh_pattern(self._h, k, None, None, None, _p(nnz, _i64), row_end))
        rows = self._built["C"] if k == 1 else self._built["T"]
        rp = np.zeros(rows + 1, dtype=np.int64)
        ci = np.zeros(max(1, int(nnz[0])), dtype=np.int32)
        dg = np.zeros(rows, dtype=np.int64)
        _check(H.rpm_mesh_pattern(self._h, k, _p(rp, _i64), _p(ci, _i32), _p(dg, _i64), _p(nnz, _i64), row_end))
        return rp, ci[: int(nnz[0])], dg

    def leaf_blocks(self):
        b = self.build()
        blocks = np.zeros((b["leaf_blocks"], 2), dtype=np.int32)
        vf = np.zeros((b["verynear_nnz"] // 2, 2), dtype=np.int32)
        it = np.zeros((b["inter_nnz"] // 2, 2), dtype=np.int32)
        _check(H.rpm_mesh_leaf_blocks(self._h, _p(blocks, _i32), _p(vf, _i32), _p(it, _i32)))
        return blocks, vf, it

    def tiles(self):
        """Near tiles: col_ptr (leaves + 1), col_prim, col_mask (tile columns), val_ptr (leaves + 1)."""
        b = self.build()
        cp = np.zeros(b["L"] + 1, dtype=np.int64)
        prim = np.zeros(b["tile_cols"], dtype=np.int32)
        mask = np.zeros(b["tile_cols"], dtype=np.uint32)
        vp = np.zeros(b["L"] + 1, dtype=np.int64)
        _check(H.rpm_mesh_tiles(self._h, _p(cp, _i64), _p(prim, _i32), mask.ctypes.data_as(C.POINTER(C.c_uint32)), _p(vp, _i64)))
        return cp, prim, mask, vp

    def build_device_ms(self):
        """Device time (ms) of the GPU build of the cluster tree and block cluster tree, or -1."""
        ms = np.zeros(1)
        _check(H.rpm_mesh_build_device_ms(self._h, _p(ms)))
        return float(ms[0])

    def VeryNearStats(self):
        """(evaluations, deepest level) of the very-near field of the last tangent-point compute."""
        ev = np.zeros(1, dtype=np.int64)
        lv = C.c_int(0)
        _check(H.rpm_vf_stats(self._h, _p(ev, _i64), C.byref(lv)))
        return int(ev[0]), int(lv.value)

    def SemiStaticUpdate(self, coords=None, device_ptr=None):
        """New vertex coordinates (V x 3) from host memory, or from a device pointer (V x 3 row-major
        doubles in HBM, an int or ctypes pointer): the per-state data is then computed on the GPU."""
        if device_ptr is not None:
            _check(H.rpm_mesh_semistatic_update(self._h, C.cast(C.c_void_p(int(device_ptr) if not isinstance(device_ptr, C.c_void_p) else device_ptr.value), _d), 1))
            return
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        assert coords.shape == (self.V, 3)
        _check(H.rpm_mesh_semistatic_update(self._h, _p(coords), 0))

    def MaximumSafeStepSize(self, vecs=None, max_time=1.0, TOL=0.0625, return_stats=False, device_ptr=None):
        """Largest t in [0, max_time] such that moving the vertices by t vecs (V x 3) causes no collision
        between simplices without a common vertex (reference SimplicialMesh::MaximumSafeStepSize).
        device_ptr: address of V x 3 row-major doubles on the mesh's GPU instead of vecs."""
        self.build()
        t = np.zeros(1); st = np.zeros(3, dtype=np.int64)
        if device_ptr is not None:
            _check(H.rpm_mesh_max_safe_step(self._h, C.cast(C.c_void_p(int(device_ptr) if not isinstance(device_ptr, C.c_void_p) else device_ptr.value), _d), float(max_time), float(TOL), 1, _p(t), _p(st, _i64)))
            return (float(t[0]), {"leaf_pairs": int(st[0]), "pairs": int(st[1]), "hull_tests": int(st[2])}) if return_stats else float(t[0])
        U = np.ascontiguousarray(vecs, dtype=np.float64)
        assert U.shape == (self.V, 3)
        _check(H.rpm_mesh_max_safe_step(self._h, _p(U), float(max_time), float(TOL), 0, _p(t), _p(st, _i64)))
        return (float(t[0]), {"leaf_pairs": int(st[0]), "pairs": int(st[1]), "hull_tests": int(st[2])}) if return_stats else float(t[0])

    def device_state_rows(self):
        """Near rows (T x ND) and cluster rows (C x 10) computed on the device by the last update."""
        b = self.build()
        nd = 1 + 3 * (self.dom_dim + 1) + 6
        N = np.zeros((b["T"], nd)); Cf = np.zeros((b["C"], 10))
        _check(H.rpm_mesh_device_state_rows(self._h, _p(N), _p(Cf)))
        return N, Cf

    def VertexCoordinates(self):
        X = np.zeros((self.V, 3))
        _check(H.rpm_mesh_coords(self._h, _p(X)))
        return X

    # ---------------------------------------------------------------- H^1 metric (host side)
    def h1_csr(self):
        self.build()
        nnz = np.zeros(1, dtype=np.int64)
        _check(H.rpm_h1_csr(self._h, None, None, None, _p(nnz, _i64)))
        rp = np.zeros(self.V + 1, dtype=np.int64); ci = np.zeros(int(nnz[0]), dtype=np.int32); v = np.zeros(int(nnz[0]))
        _check(H.rpm_h1_csr(self._h, _p(rp, _i64), _p(ci, _i32), _p(v), _p(nnz, _i64)))
        return rp, ci, v

    def nd_ordering(self):
        self.build()
        p = np.zeros(self.V, dtype=np.int32)
        _check(H.rpm_nd_ordering(self._h, _p(p, _i32)))
        return p

    def h1_solve_host(self, B):
        """X = H^{-1} B with the host supernodal factor (B: V x 3); returns (X, stats)."""
        self.build()
        B = np.ascontiguousarray(B, dtype=np.float64)
        X = np.zeros_like(B)
        st = np.zeros(5)
        _check(H.rpm_h1_solve_host(self._h, _p(B), _p(X), _p(st)))
        keys = ["nnz_H", "nnz_L", "supernodes", "levels", "flops"]
        return X, dict(zip(keys, st.tolist()))

    def h1_solve(self, B, alpha=1.0, beta=0.0, X=None):
        """X = alpha H^{-1} B + beta X with the factor's triangular solves on the GPU (B: V x 3)."""
        self.build()
        B = np.ascontiguousarray(B, dtype=np.float64)
        if X is None:
            X = np.zeros_like(B)
        _check(H.rpm_h1_solve_device(self._h, alpha, _p(B), B.shape[1], beta, _p(X), X.shape[1], 0))
        return X

    def h1_multiply_host(self, X):
        self.build()
        X = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.zeros_like(X)
        _check(H.rpm_h1_multiply_host(self._h, _p(X), _p(Y)))
        return Y

    def upload(self):
        """Builds the hierarchy and makes the device copies current (no compute)."""
        self.build()
        _check(H.rpm_mesh_device(self._h))

    def last_timing(self, which: int) -> float:
        ms = np.zeros(1)
        _check(H.rpm_mesh_timing(self._h, which, _p(ms)))
        return float(ms[0])


class _Op:
    def __init__(self, q, p=None, precision="fp64"):
        if p is None:  # (p) constructor: q = p, p = 2p (reference TangentPointEnergy0.hpp:35-39)
            q, p = q, 2 * q
        self.q, self.p = float(q), float(p)
        self.precision = precision
        self._h = H.rpm_op_new(self.q, self.p, PRECISIONS[precision])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and H is not None:
            H.rpm_op_free(h)
            self._h = None


class TangentPointEnergy0(_Op):
    """TangentPointEnergy0 (reference src/Energies/TangentPointEnergy0.hpp, EnergyBase.hpp)."""

    def Value(self, M: SimplicialMesh) -> float:
        M.build()
        out = np.zeros(1)
        _check(H.rpm_energy_value(self._h, M._h, _p(out)))
        return float(out[0])

    def Differential(self, M: SimplicialMesh, alpha=1.0, beta=0.0, Y=None, return_energy=False):
        """Y = alpha dE + beta Y (V x 3). Returns Y (and the energy if return_energy)."""
        M.build()
        if Y is None:
            Y = np.zeros((M.V, 3))
        assert Y.dtype == np.float64 and Y.flags.c_contiguous and Y.shape[0] == M.V
        en = np.zeros(1)
        _check(H.rpm_energy_differential(self._h, M._h, alpha, beta, _p(Y), Y.shape[1], _p(en), 0))
        return (Y, float(en[0])) if return_energy else Y

    def MomentDerivatives(self, M: SimplicialMesh):
        """Derivatives w.r.t. the near moments (T x ND) and far moments (T x 10), original order."""
        nd = 1 + 3 * (M.dom_dim + 1) + 6
        N = np.zeros((M.T, nd)); F = np.zeros((M.T, 10))
        _check(H.rpm_energy_moment_derivatives(self._h, M._h, _p(N), _p(F)))
        return N, F


class TangentPointMetric0(_Op):
    """TangentPointMetric0 (reference src/Metrics/TangentPointMetric0.hpp, MetricDimRestricted.hpp)."""

    def MetricValues(self, M: SimplicialMesh):
        b = M.build()
        out = {"NF": np.zeros((b["near_nnz"], 2)), "NF_diag": np.zeros((b["T"], 2)),
               "FF": np.zeros((b["far_nnz"], 2)), "FF_diag": np.zeros((b["C"], 2))}
        _check(H.rpm_metric_values(self._h, M._h, _p(out["NF"]), _p(out["NF_diag"]), _p(out["FF"]), _p(out["FF_diag"])))
        return out

    def VeryNearMetricValues(self, M: SimplicialMesh):
        """Very-near-field metric values: VF (nnz x 2, CSR order of M.pattern('verynear')) and VF_diag
        (T x 2, tree order)."""
        b = M.build()
        out = {"VF": np.zeros((max(1, b["verynear_nnz"]), 2)), "VF_diag": np.zeros((b["T"], 2))}
        _check(H.rpm_vf_metric_values(self._h, M._h, _p(out["VF"]), _p(out["VF_diag"])))
        out["VF"] = out["VF"][: b["verynear_nnz"]]
        return out

    def VeryNearPseudoLaplacianValues(self, M: SimplicialMesh):
        """Very-near-field values of L_{2-s}: VF (nnz) and VF_diag (T, tree order)."""
        b = M.build()
        out = {"VF": np.zeros(max(1, b["verynear_nnz"])), "VF_diag": np.zeros(b["T"])}
        _check(H.rpm_pl_vf_values(self._h, M._h, _p(out["VF"]), _p(out["VF_diag"])))
        out["VF"] = out["VF"][: b["verynear_nnz"]]
        return out

    def assemble(self, M: SimplicialMesh):
        M.build()
        _check(H.rpm_metric_assemble(self._h, M._h))

    def MultiplyMetric(self, M: SimplicialMesh, alpha, X, beta, Y, nrhs=3, VF=True, NF=True, FF=True):
        """Y = alpha A X + beta Y; X, Y are V x nrhs (nrhs must be 3, else Y is left untouched)."""
        M.build()
        X = np.ascontiguousarray(X, dtype=np.float64)
        assert Y.dtype == np.float64 and Y.flags.c_contiguous
        st = H.rpm_metric_multiply(self._h, M._h, alpha, _p(X), X.shape[1], beta, _p(Y), Y.shape[1],
                                   nrhs, int(VF), int(NF), int(FF), 0)
        _check(st)
        return Y

    def MultiplyMetricInner(self, M: SimplicialMesh, alpha, X, beta, Y):
        """Y = alpha A32 X + beta Y with the FP32 copies of a "mixed-refined" metric (the product the CG
        iterations of Solve use; no reference counterpart)."""
        M.build()
        X = np.ascontiguousarray(X, dtype=np.float64)
        assert Y.dtype == np.float64 and Y.flags.c_contiguous
        _check(H.rpm_metric_multiply(self._h, M._h, alpha, _p(X), X.shape[1], beta, _p(Y), Y.shape[1],
                                     3, 1, 2, 1, 0))
        return Y

    def MetricNorm(self, M: SimplicialMesh, X):
        Y = np.zeros_like(X)
        self.MultiplyMetric(M, 1.0, X, 0.0, Y)
        return float(np.sqrt(np.sum(X * Y)))

    # ------------------------------------------------------------------ preconditioner and solve
    def PseudoLaplacianValues(self, M: SimplicialMesh):
        """Values of L_{2-s} (near CSR order, scalar per nnz), its diagonals and the far values."""
        b = M.build()
        out = {"NF": np.zeros(b["near_nnz"]), "NF_diag": np.zeros(b["T"]),
               "FF": np.zeros(b["far_nnz"]), "FF_diag": np.zeros(b["C"])}
        _check(H.rpm_pl_values(self._h, M._h, _p(out["NF"]), _p(out["NF_diag"]), _p(out["FF"]), _p(out["FF_diag"])))
        return out

    def MultiplyPseudoLaplacian(self, M: SimplicialMesh, alpha, X, beta, Y):
        """Y = alpha L_{2-s} X + beta Y (V x 3)."""
        M.build()
        X = np.ascontiguousarray(X, dtype=np.float64)
        assert Y.dtype == np.float64 and Y.flags.c_contiguous
        _check(H.rpm_pl_multiply(self._h, M._h, alpha, _p(X), X.shape[1], beta, _p(Y), Y.shape[1], 0))
        return Y

    def MultiplyPreconditioner(self, M: SimplicialMesh, alpha, X, beta, Y, nrhs=3):
        """Y = alpha H^-1 L_{2-s} H^-1 X + beta Y (reference TangentPointMetric0.hpp:88-126)."""
        M.build()
        X = np.ascontiguousarray(X, dtype=np.float64)
        assert Y.dtype == np.float64 and Y.flags.c_contiguous
        _check(H.rpm_metric_precondition(self._h, M._h, alpha, _p(X), X.shape[1], beta, _p(Y), Y.shape[1], nrhs, 0))
        return Y

    def SetPreconditionerPrecision(self, precision):
        """Precision ("fp64", "mixed" or "bf16": FP32 values stored in BF16) of the pseudo-Laplacian
        inside the preconditioner; the metric operator keeps its own precision (no reference counterpart)."""
        H.rpm_op_set_pc_precision(self._h, PRECISIONS[precision])

    def SetPreconditionerDiagonalWeight(self, w):
        """P = P0 + w diag(A)^-1: w = 0 the reference P0 (default), w > 0 fixed, w < 0 estimated per state
        as lambda_max(P0 A) / 4. The diagonal term has no counterpart in the reference."""
        self.pc_weight = float(w)
        H.rpm_op_set_pc_weight(self._h, self.pc_weight)

    def PreconditionerWeightInUse(self, M: SimplicialMesh):
        w = np.zeros(1)
        _check(H.rpm_pc_weight(self._h, M._h, _p(w)))
        return float(w[0])

    def MetricDiagonal(self, M: SimplicialMesh):
        """Exact diagonal of the assembled metric (V values, the same for the three columns)."""
        M.build()
        d = np.zeros(M.V)
        _check(H.rpm_metric_diagonal(self._h, M._h, _p(d)))
        return d

    def PreconditionerNorm(self, M: SimplicialMesh, X):
        Y = np.zeros_like(X)
        self.MultiplyPreconditioner(M, 1.0, X, 0.0, Y)
        return float(np.sqrt(np.sum(X * Y)))

    def Solve(self, M: SimplicialMesh, alpha, B, beta, X, max_iter=1000, tolerance=1e-6, nrhs=3,
              preconditioner=True, measure="l2", method="pcg", raise_on_stagnation=True, record=False):
        """X = alpha A^+ B + beta X on the GPU; the record is kept for Solver_IterationCount and
        Solver_RelativeResiduals (reference MetricDimRestricted.hpp:232-289, MetricBase.hpp:126-134).
        method "pcg" with measure "l2" (|B - A X|_2 / |B|_2 <= tolerance per column) or "pc" (the same
        ratio in the preconditioner norm sqrt(<r, P r>)); method "gmres": the reference's
        left-preconditioned GMRES, measure |P (B - A X)|_2 / |P B|_2. record keeps the PCG iterates on
        the device for PreconditionerLearnStep."""
        M.build()
        B = np.ascontiguousarray(B, dtype=np.float64)
        assert X.dtype == np.float64 and X.flags.c_contiguous
        assert measure in ("l2", "pc") and method in ("pcg", "gmres")
        it = (C.c_int * 2)()
        rel = np.zeros(3)
        flags = (0 if preconditioner else RP_SOLVE_NO_PRECONDITIONER) | (RP_SOLVE_PC_NORM if measure == "pc" else 0)
        flags |= (RP_SOLVE_GMRES if method == "gmres" else 0) | (RP_SOLVE_RECORD if record else 0)
        st = H.rpm_metric_solve(self._h, M._h, alpha, _p(B), B.shape[1], beta, _p(X), X.shape[1], nrhs,
                                int(max_iter), float(tolerance), flags, 0, it, _p(rel))
        self._iter, self._relres, self._fell_back = int(it[0]), rel, bool(it[1])
        if st == RP_ESTAGNATION and not raise_on_stagnation:
            return X
        _check(st)
        return X

    def Solver_IterationCount(self):
        return getattr(self, "_iter", 0)

    def Solver_RelativeResiduals(self):
        return getattr(self, "_relres", np.zeros(3))

    def SetLearnedPreconditionerGuard(self, check_iteration=12, reference_rate=0.5):
        """A learned PCG solve whose largest relative residual at `check_iteration` is above
        reference_rate ** (check_iteration / 1.5) (the base preconditioner's rate, 1.5 times the iterations)
        stops and falls back to P0 + w diag(A)^-1; 0 keeps only the fallback at max_iter."""
        H.rpm_op_set_lp_guard(self._h, int(check_iteration), float(reference_rate))

    def Solver_FellBack(self):
        """Whether the last Solve fell back from the learned preconditioner to P0 + w diag(A)^-1."""
        return getattr(self, "_fell_back", False)

    # ------------------------------------------------------------------ learned preconditioner
    def SetLearnedPreconditioner(self, params=None, path=None, bf16=True):
        """P = D^{1/2} (P0 + w diag(A)^-1) D^{1/2} with the node MLP; params (float32, 64 F + 130) or a
        model file written by repulsor_amd.learned_pc.save_model. None removes the model."""
        from . import learned_pc
        if path is not None:
            params = learned_pc.load_model(path)
        if params is None:
            H.rpm_op_set_learned(self._h, 0, None, 0)
            self._lp = None
            return
        self._lp = np.ascontiguousarray(params, dtype=np.float32)
        F = learned_pc.feature_count()
        assert self._lp.size == learned_pc.param_count(F)
        H.rpm_op_set_learned(self._h, F, self._lp.ctypes.data_as(C.POINTER(C.c_float)), int(bool(bf16)))

    def PreconditionerLearnStep(self, M: SimplicialMesh, params, m, v, step, lr):
        """One Adam step of the model on the pairs recorded by the last Solve(record=True); params, m, v
        (float32) are updated in place. Returns (loss, gradient)."""
        fp = C.POINTER(C.c_float)
        loss = np.zeros(1)
        grad = np.zeros(params.size)
        _check(H.rpm_pc_learn_step(self._h, M._h, float(lr), int(step), params.ctypes.data_as(fp), m.ctypes.data_as(fp),
                                   v.ctypes.data_as(fp), _p(loss), _p(grad)))
        return float(loss[0]), grad

    def LearnedScaling(self, M: SimplicialMesh):
        d = np.zeros(M.V)
        _check(H.rpm_pc_scaling(self._h, M._h, _p(d)))
        return d
