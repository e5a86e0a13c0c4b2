"""Mixed precision (FP32 pair arithmetic after an FP64 shift to a per-leaf origin, FP64 accumulation)
against the FP64 oracle: energy, differential, metric product, pseudo-Laplacian product, and a
solve whose residual is measured with the FP64 oracle metric.

Stated errors (relative, 2-norm; measured on MI355X in brackets): energy <= 1e-7 [0.2-4.2e-9],
differential <= 1e-5 for surfaces [0.7-1.4e-6] and <= 2e-4 for curves [7.2e-5], metric product
<= 1e-6 [4.5-8.3e-8], pseudo-Laplacian product <= 1e-6 [1.4-2.5e-8]. The near-field FP32 kernel forms
<v, P v> as |P v|^2 (P a projector), which keeps the rounding of close, almost tangential pairs at
about eps / theta instead of eps / theta^2; larger meshes are covered in test_gpu_fullsize.py. The mixed solve stops at
1e-6 on its FP32-valued operator; its residual against the FP64 operator is <= 3e-6 [1.2-1.7e-6].

"mixed-refined" (FP64 metric plus FP32 copies of the near values for the CG iterations): the FP64
product is bitwise the FP64 metric's, the FP32-copy product is within 1e-6 of it [0.6-1.9e-8], and
the solve, judged by the FP64 residual, ends below 1e-6 in the FP64 oracle's measure [9.0-9.1e-7]
with at most 4 more iterations than the FP64 solve [the same counts].
"""

import numpy as np
import pytest

import oracle as O
import repulsor_amd as R
from repulsor_amd.workloads import WORKLOADS, random_field
from helpers import OracleSolve, oracle_energy_diff, oracle_multiply, oracle_state, rel

pytestmark = pytest.mark.gpu

CASES = ["torus_2400", "sphere_1280", "trefoil_3000"]


def _mesh(name):
    wl = WORKLOADS[name]
    M = wl.mesh()
    M.build()
    return wl, M


@pytest.mark.parametrize("name", CASES)
def test_mixed_energy_differential(name):
    wl, M = _mesh(name)
    E = R.TangentPointEnergy0(wl.q, wl.p, precision="mixed")
    Y, en = E.Differential(M, return_energy=True)
    en_ref, Y_ref, _, _ = oracle_energy_diff(M, wl.q, wl.p)
    e_en, e_d = abs(en - en_ref) / abs(en_ref), rel(Y, Y_ref)
    print(f"case {name}: mixed energy rel err {e_en:.2e}, differential {e_d:.2e}", flush=True)
    assert e_en <= 1e-7
    assert e_d <= (2e-4 if M.dom_dim == 1 else 1e-5)


@pytest.mark.parametrize("name", CASES)
def test_mixed_metric_products(name):
    wl, M = _mesh(name)
    A = R.TangentPointMetric0(wl.q, wl.p, precision="mixed")
    X = random_field(M.V, seed=11)
    Y = np.zeros_like(X)
    A.MultiplyMetric(M, 1.0, X, 0.0, Y)
    e_a = rel(Y, oracle_multiply(M, wl.q, wl.p, X))
    Z = np.zeros_like(X)
    A.MultiplyPseudoLaplacian(M, 1.0, X, 0.0, Z)
    tr, td, npat, fpat = oracle_state(M)
    pl = O.pl_compute(M.dom_dim, td["P_near"], td["C_far"], npat, fpat, 2.0 - (wl.p - M.dom_dim) / wl.q)
    Zr = np.zeros_like(X)
    O.pl_multiply(M.dom_dim, M.V, tr, td["pre_cols"], td["pre_vals"], npat, pl, fpat, 1.0, X, 0.0, Zr)
    e_l = rel(Z, Zr)
    print(f"case {name}: mixed metric product rel err {e_a:.2e}, pseudo-Laplacian {e_l:.2e}", flush=True)
    assert e_a <= 1e-6
    assert e_l <= 1e-6


@pytest.mark.parametrize("name", ["torus_2400", "sphere_1280"])
def test_mixed_solve(name):
    wl, M = _mesh(name)
    E = R.TangentPointEnergy0(wl.q, wl.p)
    B = E.Differential(M)
    A = R.TangentPointMetric0(wl.q, wl.p, precision="mixed")
    A.SetPreconditionerDiagonalWeight(-1.0)
    X = np.zeros_like(B)
    A.Solve(M, 1.0, B, 0.0, X, max_iter=500, tolerance=1e-6)
    S = OracleSolve(M, wl.q, wl.p)
    true = np.linalg.norm(B - S.A(X), axis=0) / np.linalg.norm(B, axis=0)
    print(f"case {name}: mixed solve {A.Solver_IterationCount()} iterations, FP64 residual {true.max():.2e}", flush=True)
    assert np.all(true <= 3e-6), true


@pytest.mark.parametrize("name", CASES)
def test_refined_metric_copies(name):
    """"mixed-refined": the FP64 values are those of the FP64 metric (same kernel, bit for bit); the
    FP32 copies and their diagonal give a product within FP32 rounding of the FP64 one, also on a
    smooth field (the vertex coordinates), where a diagonal inconsistent with the copies would show."""
    wl, M = _mesh(name)
    A64 = R.TangentPointMetric0(wl.q, wl.p)
    Ar = R.TangentPointMetric0(wl.q, wl.p, precision="mixed-refined")
    rng = np.random.default_rng(5)
    for label, X in (("random", random_field(M.V, seed=11)),
                     ("coordinates", M.VertexCoordinates() + 1e-3 * rng.standard_normal((M.V, 3)))):
        Y64 = np.zeros_like(X); A64.MultiplyMetric(M, 1.0, X, 0.0, Y64)
        Yr = np.zeros_like(X); Ar.MultiplyMetric(M, 1.0, X, 0.0, Yr)
        assert np.array_equal(Yr, Y64), label
        Yi = np.zeros_like(X); Ar.MultiplyMetricInner(M, 1.0, X, 0.0, Yi)
        e_i = rel(Yi, Y64)
        e_o = rel(Y64, oracle_multiply(M, wl.q, wl.p, X))
        print(f"case {name} {label}: FP32-copy product rel err {e_i:.2e} (FP64 vs oracle {e_o:.2e})", flush=True)
        assert e_o <= 1e-10
        assert 0 < e_i <= 1e-6


def test_refined_copies_need_refined_metric():
    wl, M = _mesh("torus_2400")
    A = R.TangentPointMetric0(wl.q, wl.p)
    X = random_field(M.V, seed=3)
    with pytest.raises(RuntimeError, match="RP_MIXED_REFINED"):
        A.MultiplyMetricInner(M, 1.0, X, 0.0, np.zeros_like(X))


@pytest.mark.parametrize("name", ["torus_2400", "sphere_1280"])
@pytest.mark.parametrize("pc_norm", [False, True])
def test_refined_solve(name, pc_norm):
    """CG on the FP32 copies, residual replaced by the FP64 one: the FP64 oracle residual meets the
    tolerance, as in the FP64 solve, and the reported residual is the FP64 one."""
    wl, M = _mesh(name)
    E = R.TangentPointEnergy0(wl.q, wl.p)
    B = E.Differential(M)
    S = OracleSolve(M, wl.q, wl.p)
    its = {}
    for prec in ("fp64", "mixed-refined"):
        A = R.TangentPointMetric0(wl.q, wl.p, precision=prec)
        A.SetPreconditionerDiagonalWeight(-1.0)
        X = np.zeros_like(B)
        A.Solve(M, 1.0, B, 0.0, X, max_iter=500, tolerance=1e-6, measure="pc" if pc_norm else "l2")
        its[prec] = A.Solver_IterationCount()
        true = np.linalg.norm(B - S.A(X), axis=0) / np.linalg.norm(B, axis=0)
        rep = np.asarray(A.Solver_RelativeResiduals())
        print(f"case {name} pc_norm={pc_norm} {prec}: {its[prec]} iterations, FP64 residual {true.max():.2e}, "
              f"reported {rep.max():.2e}", flush=True)
        if not pc_norm:
            assert np.all(true <= 1e-6), true
            assert np.allclose(rep, true, rtol=1e-3, atol=0), (rep, true)
        else:
            assert np.all(rep <= 1e-6), rep
    assert its["mixed-refined"] <= its["fp64"] + 4


def test_refined_solve_iteration_cap():
    """Iteration cap reached while CG still multiplies with the FP32 copies: the solve reports
    stagnation, and the residual it reports is the FP64 one of the iterate it returns."""
    wl, M = _mesh("torus_2400")
    E = R.TangentPointEnergy0(wl.q, wl.p)
    B = E.Differential(M)
    A = R.TangentPointMetric0(wl.q, wl.p, precision="mixed-refined")
    A.SetPreconditionerDiagonalWeight(-1.0)
    X = np.zeros_like(B)
    A.Solve(M, 1.0, B, 0.0, X, max_iter=3, tolerance=1e-6, raise_on_stagnation=False)
    S = OracleSolve(M, wl.q, wl.p)
    true = np.linalg.norm(B - S.A(X), axis=0) / np.linalg.norm(B, axis=0)
    rep = np.asarray(A.Solver_RelativeResiduals())
    print(f"capped refined solve: {A.Solver_IterationCount()} iterations, FP64 residual {true.max():.2e}, "
          f"reported {rep.max():.2e}", flush=True)
    assert A.Solver_IterationCount() == 3
    assert np.all(true > 1e-6)
    assert np.allclose(rep, true, rtol=1e-6, atol=0), (rep, true)
    with pytest.raises(R.RepulsorError):
        A.Solve(M, 1.0, B, 0.0, X, max_iter=3, tolerance=1e-6)
