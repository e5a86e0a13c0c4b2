"""The benchmark meshes at full size on the GPU.

* sphere_81920 (icosphere level 6, 81,920 triangles, q = 6, p = 12) and trefoil_200000 (200,000 edges,
  q = 2, p = 4): FP64 energy, differential and MultiplyMetric (3 columns) against the oracle at relative
  1e-10; the FP32-mixed path's errors against the same oracle are printed and bounded: energy <= 1e-7
  [2.0e-9, 2.0e-8 measured], metric product <= 1e-6 [4.6e-8, 8.2e-8], differential <= 2e-5 for the
  surface [3.9e-6] and <= 1e-2 for the curve [3.6e-3]. The differential is a sum of pair terms that
  cancel: the FP64 path's own differential error (1.8e-13 surface, 1.8e-11 curve, against 1.4e-16 and
  1.8e-16 for the energy) shows the amplification, about 1e3 for the surface and 1e5 for the 200,000-edge
  curve, so FP32 pair terms (eps 6e-8) cannot give the curve's differential to better than about 1e-3.
* sphere_81920: the mixed-refined solve (FP32 near values inside PCG, FP64 residual as the judge) ends
  at FP64 residual <= 1e-6 with at most 4 more iterations than the FP64 solve.
* torus_153600 (320 x 240, R = 1, r = 0.4): MultiplyMetric against the oracle at 1e-10, then the
  metric solve A X = dE to 1e-6 with the reference preconditioner P0, the two-level P0 + w diag(A)^-1
  and the learned model; every solve's residual is recomputed with the FP64 product (<= 1e-6) and
  iterations and times are printed.
* sphere_1310720 (icosphere level 8): properties of the GPU output that hold at any size: the
  differential is translation and rotation invariant and satisfies Euler's relation
  sum_v x_v . dE/dx_v = (2n + q - p) E; the metric is symmetric and annihilates constants; the solve
  reaches FP64 residual <= 1e-6. The FP32-mixed energy and differential are compared with the FP64 GPU
  path (itself oracle-checked at the smaller sizes): energy <= 1e-7, differential <= 5e-5.
"""

import os
import time

import numpy as np
import pytest

import repulsor_amd as R
from repulsor_amd.workloads import WORKLOADS, random_field
from helpers import oracle_energy_diff, oracle_multiply, rel

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
from repulsor_amd.learned_pc import DEFAULT_MODEL as MODEL  # noqa: E402

_meshes = {}


def _mesh(name):
    if name not in _meshes:
        _meshes.clear()          # one large mesh at a time
        wl = WORKLOADS[name]
        M = wl.mesh()
        b = M.build()
        print(f"{name}: T = {b['T']}, near nnz {b['near_nnz']}, far nnz {b['far_nnz']}", flush=True)
        _meshes[name] = (wl, M)
    return _meshes[name]


def _fp64_residual(M, A, B, X):
    AX = np.zeros_like(X)
    A.MultiplyMetric(M, 1.0, X, 0.0, AX)
    return np.linalg.norm(B - AX, axis=0) / np.linalg.norm(B, axis=0)


# ------------------------------------------------------------------------------------------------ cfg2, cfg3
@pytest.mark.parametrize("name", ["sphere_81920", "trefoil_200000"])
def test_fullsize_parity_fp64_and_mixed(name):
    wl, M = _mesh(name)
    en_o, G_o, _, _ = oracle_energy_diff(M, wl.q, wl.p)
    E = R.TangentPointEnergy0(wl.q, wl.p)
    G, en = E.Differential(M, return_energy=True)
    e_en, e_g = abs(en - en_o) / abs(en_o), rel(G, G_o)
    X = random_field(M.V, seed=11)
    AX_o = oracle_multiply(M, wl.q, wl.p, X)
    A = R.TangentPointMetric0(wl.q, wl.p)
    AX = np.zeros_like(X)
    A.MultiplyMetric(M, 1.0, X, 0.0, AX)
    e_a = rel(AX, AX_o)
    print(f"{name} FP64 vs oracle: energy {e_en:.2e}, differential {e_g:.2e}, MultiplyMetric {e_a:.2e}", flush=True)
    assert e_en <= 1e-10 and e_g <= 1e-10 and e_a <= 1e-10
    Em = R.TangentPointEnergy0(wl.q, wl.p, precision="mixed")
    Gm, enm = Em.Differential(M, return_energy=True)
    Am = R.TangentPointMetric0(wl.q, wl.p, precision="mixed")
    AXm = np.zeros_like(X)
    Am.MultiplyMetric(M, 1.0, X, 0.0, AXm)
    m_en, m_g, m_a = abs(enm - en_o) / abs(en_o), rel(Gm, G_o), rel(AXm, AX_o)
    print(f"{name} FP32-mixed vs oracle: energy {m_en:.2e}, differential {m_g:.2e}, MultiplyMetric {m_a:.2e}", flush=True)
    assert m_en <= 1e-7
    assert m_g <= (1e-2 if M.dom_dim == 1 else 2e-5)
    assert m_a <= 1e-6


def test_fullsize_refined_solve_sphere_81920():
    wl, M = _mesh("sphere_81920")
    G = R.TangentPointEnergy0(wl.q, wl.p).Differential(M)
    its = {}
    for prec in ("fp64", "mixed-refined"):
        A = R.TangentPointMetric0(wl.q, wl.p, precision=prec)
        A.SetPreconditionerDiagonalWeight(-1.0)
        X = np.zeros_like(G)
        t = time.perf_counter()
        A.Solve(M, 1.0, G, 0.0, X, max_iter=500, tolerance=1e-6)
        dt = time.perf_counter() - t
        res = _fp64_residual(M, R.TangentPointMetric0(wl.q, wl.p), G, X)
        its[prec] = A.Solver_IterationCount()
        print(f"sphere_81920 {prec} solve: {its[prec]} iterations, {1e3 * dt:.1f} ms, FP64 residual {res.max():.2e}", flush=True)
        assert np.all(res <= 1e-6)
    assert its["mixed-refined"] <= its["fp64"] + 4


# ------------------------------------------------------------------------------------------------ cfg4
def test_fullsize_torus_153600_preconditioners():
    wl, M = _mesh("torus_153600")
    X = random_field(M.V, seed=11)
    A = R.TangentPointMetric0(wl.q, wl.p)
    AX = np.zeros_like(X)
    A.MultiplyMetric(M, 1.0, X, 0.0, AX)
    e_a = rel(AX, oracle_multiply(M, wl.q, wl.p, X))
    print(f"torus_153600 MultiplyMetric vs oracle {e_a:.2e}", flush=True)
    assert e_a <= 1e-10
    G = R.TangentPointEnergy0(wl.q, wl.p).Differential(M)
    rows = {}
    for label in ("P0", "two-level", "learned"):
        S = R.TangentPointMetric0(wl.q, wl.p)
        if label == "two-level":
            S.SetPreconditionerDiagonalWeight(-1.0)
        if label == "learned":
            S.SetLearnedPreconditioner(path=MODEL)
        Xs = np.zeros_like(G)
        S.Solve(M, 1.0, G, 0.0, Xs, max_iter=3000, tolerance=1e-6)      # warm-up: factor, assembly, model
        Xs[:] = 0.0
        t = time.perf_counter()
        S.Solve(M, 1.0, G, 0.0, Xs, max_iter=3000, tolerance=1e-6)
        dt = time.perf_counter() - t
        res = _fp64_residual(M, A, G, Xs)
        rows[label] = (S.Solver_IterationCount(), dt)
        print(f"torus_153600 {label}: {rows[label][0]} PCG iterations, {1e3 * dt:.1f} ms, FP64 residual {res.max():.2e}"
              + (f", fell back {S.Solver_FellBack()}" if label == "learned" else ""), flush=True)
        assert np.all(res <= 1e-6)
    assert rows["two-level"][0] < rows["P0"][0]


# ------------------------------------------------------------------------------------------------ cfg5
def test_fullsize_sphere_1310720_properties():
    wl, M = _mesh("sphere_1310720")
    n = M.dom_dim
    Xc = M.VertexCoordinates()
    E = R.TangentPointEnergy0(wl.q, wl.p)
    G, en = E.Differential(M, return_energy=True)
    nG, nX = np.linalg.norm(G), np.linalg.norm(Xc)
    trans = np.linalg.norm(G.sum(axis=0)) / (nG * np.sqrt(M.V))
    rot = np.linalg.norm(np.cross(Xc, G).sum(axis=0)) / (nG * nX)
    euler = abs(np.sum(Xc * G) - (2 * n + wl.q - wl.p) * en) / abs(en)
    print(f"sphere_1310720 energy {en:.12e}: translation {trans:.2e}, rotation {rot:.2e}, Euler {euler:.2e}", flush=True)
    assert trans <= 1e-12 and rot <= 1e-12 and euler <= 1e-10
    Gm, enm = R.TangentPointEnergy0(wl.q, wl.p, precision="mixed").Differential(M, return_energy=True)
    m_en, m_g = abs(enm - en) / abs(en), rel(Gm, G)
    print(f"sphere_1310720 FP32-mixed vs FP64 (GPU, oracle-checked at smaller sizes): energy {m_en:.2e}, "
          f"differential {m_g:.2e}", flush=True)
    assert m_en <= 1e-7 and m_g <= 5e-5          # measured 2.5e-9 and 9.2e-6
    A = R.TangentPointMetric0(wl.q, wl.p)
    U, V = random_field(M.V, seed=11), random_field(M.V, seed=12)
    AU, AV = np.zeros_like(U), np.zeros_like(V)
    A.MultiplyMetric(M, 1.0, U, 0.0, AU)
    A.MultiplyMetric(M, 1.0, V, 0.0, AV)
    sym = abs(np.sum(U * AV) - np.sum(AU * V)) / (np.linalg.norm(AU) * np.linalg.norm(V))
    A1 = np.zeros_like(U)
    A.MultiplyMetric(M, 1.0, np.ones_like(U) * np.array([1.0, -2.0, 0.5]), 0.0, A1)
    null = np.linalg.norm(A1) / (np.linalg.norm(AU) * np.sqrt(M.V) / np.linalg.norm(U))
    print(f"sphere_1310720 metric: symmetry {sym:.2e}, A 1 {null:.2e}", flush=True)
    assert sym <= 1e-12 and null <= 1e-12
    for prec in ("fp64", "mixed-refined"):
        S = R.TangentPointMetric0(wl.q, wl.p, precision=prec)
        S.SetPreconditionerDiagonalWeight(-1.0)
        Xs = np.zeros_like(G)
        t = time.perf_counter()
        S.Solve(M, 1.0, G, 0.0, Xs, max_iter=500, tolerance=1e-6)
        dt = time.perf_counter() - t
        res = _fp64_residual(M, A, G, Xs)
        print(f"sphere_1310720 {prec} solve: {S.Solver_IterationCount()} iterations, {1e3 * dt:.1f} ms "
              f"(first call: includes the H1 factor), FP64 residual {res.max():.2e}", flush=True)
        assert np.all(res <= 1e-6)
