"""Very-near field (VF) on the GPU against the oracle, through the C-ABI.

Meshes: two nested icospheres 0.004 apart (16,580 VF pairs, adaptive subdivision to level 3) and two
stacked circles 3e-4 apart (1,200 VF pairs, level 3).

* evaluation count and deepest level equal the oracle's (the same MAC decisions);
* energy (the reference's unweighted VF sum) and vertex differential: relative 1e-10;
* moment derivatives: relative 1e-10;
* VF metric values and diagonal: relative 1e-12; MultiplyMetric with the VF blocks and with the VF
  flag off: relative 1e-10; the exact metric diagonal includes the VF diagonal;
* pseudo-Laplacian VF values: relative 1e-12; preconditioner product: relative 1e-9;
* a metric solve on a milder near-contact mesh converges, measured with the oracle's metric;
* mixed precision keeps the VF part in FP64: the energy stays within the mixed-precision tolerance.
"""

import numpy as np
import pytest

import oracle as O
import repulsor_amd as R
from repulsor_amd.workloads import random_field
from helpers import OracleSolve, oracle_energy_diff, oracle_multiply, oracle_vf, rel, vf_circles, vf_shells

pytestmark = pytest.mark.gpu

CASES = {
    "shells": (lambda: vf_shells(3, 0.004), 6.0, 12.0),
    "circles": (lambda: vf_circles(400, 3e-4), 2.0, 4.0),
}


def _mesh(name):
    make, q, p = CASES[name]
    M = make()
    b = M.build()
    assert b["verynear_nnz"] > 0
    return M, q, p


@pytest.mark.parametrize("name", list(CASES))
def test_vf_energy_and_differential(name):
    M, q, p = _mesh(name)
    E = R.TangentPointEnergy0(q, p)
    Y, en = E.Differential(M, return_energy=True)
    ev, lv = M.VeryNearStats()
    en_o, Y_o, N_o, _ = oracle_energy_diff(M, q, p)
    vf, _ = oracle_vf(M, q, p, 1)
    print(f"{name}: energy {en:.12e} oracle {en_o:.12e} (VF part {vf['energy']:.4e}); evaluations {ev} "
          f"(oracle {vf['evaluations']}), deepest level {lv} (oracle {vf['max_level']})", flush=True)
    assert ev == vf["evaluations"] and lv == vf["max_level"]
    assert abs(en - en_o) <= 1e-10 * abs(en_o)
    assert abs(E.Value(M) - en_o) <= 1e-10 * abs(en_o)
    assert rel(Y, Y_o) <= 1e-10
    N, _ = E.MomentDerivatives(M)
    assert rel(N, N_o) <= 1e-10


@pytest.mark.parametrize("name", list(CASES))
def test_vf_metric_values_and_product(name):
    M, q, p = _mesh(name)
    A = R.TangentPointMetric0(q, p)
    got = A.VeryNearMetricValues(M)
    vf, _ = oracle_vf(M, q, p, 4)
    assert rel(got["VF"], vf["VF"]) <= 1e-12
    # device diagonal in tree order, oracle's in tree order
    assert rel(got["VF_diag"], vf["VF_diag"]) <= 1e-12
    X = random_field(M.V, seed=31)
    Y = np.zeros_like(X)
    A.MultiplyMetric(M, 1.0, X, 0.0, Y)
    Y_o = oracle_multiply(M, q, p, X)
    print(f"{name}: |A X| {np.linalg.norm(Y):.6e}, rel {rel(Y, Y_o):.2e}", flush=True)
    assert rel(Y, Y_o) <= 1e-10
    Y2 = np.zeros_like(X)
    A.MultiplyMetric(M, 1.0, X, 0.0, Y2, VF=False)
    assert rel(Y2, oracle_multiply(M, q, p, X, vf=False)) <= 1e-10
    Y3 = np.zeros_like(X)
    A.MultiplyMetric(M, 1.0, X, 0.0, Y3, NF=False, FF=False)
    assert rel(Y3, oracle_multiply(M, q, p, X, nf=False, ff=False)) <= 1e-10
    # exact diagonal (includes the VF diagonal blocks) against probing the oracle's product
    d = A.MetricDiagonal(M)
    S = OracleSolve(M, q, p)
    idx = np.linspace(0, M.V - 1, 12).astype(int)
    for v in idx:
        e = np.zeros((M.V, 3)); e[v, 0] = 1.0
        assert abs(d[v] - S.A(e)[v, 0]) <= 1e-10 * abs(d[v])


@pytest.mark.parametrize("name", list(CASES))
def test_vf_pseudo_laplacian_and_preconditioner(name):
    M, q, p = _mesh(name)
    n = M.dom_dim
    A = R.TangentPointMetric0(q, p)
    got = A.VeryNearPseudoLaplacianValues(M)
    s = (p - n) / q
    ref = O.pl_vf(n, M.tree_data()["P_near"], M.pattern("verynear"), M.adaptivity_settings.theta, 20, 2.0 - s)
    assert rel(got["VF"], ref["VF"]) <= 1e-12
    assert rel(got["VF_diag"], ref["VF_diag"]) <= 1e-12
    X = random_field(M.V, seed=32)
    PX = np.zeros_like(X)
    A.MultiplyPreconditioner(M, 1.0, X, 0.0, PX)
    S = OracleSolve(M, q, p)
    assert rel(PX, S.P(X)) <= 1e-9


def test_vf_solve_converges():
    """Near-contact meshes make the metric very ill-conditioned (the reference preconditioner P0
    stalls on all of these meshes); the two-level preconditioner converges on shells 0.006 apart."""
    M, q, p = vf_shells(2, 0.006), 6.0, 12.0
    assert M.build()["verynear_nnz"] > 0
    E = R.TangentPointEnergy0(q, p)
    B = E.Differential(M)
    A = R.TangentPointMetric0(q, p)
    A.SetPreconditionerDiagonalWeight(-1.0)
    X = np.zeros_like(B)
    A.Solve(M, 1.0, B, 0.0, X, max_iter=600, tolerance=1e-6, raise_on_stagnation=False)
    S = OracleSolve(M, q, p)
    _, Bo, _, _ = oracle_energy_diff(M, q, p)
    true = np.linalg.norm(Bo - S.A(X), axis=0) / np.linalg.norm(Bo, axis=0)
    print(f"shells solve: {A.Solver_IterationCount()} iterations, true relative residuals {true}", flush=True)
    assert np.all(true <= 2e-6)


def test_vf_mixed_precision_energy():
    M, q, p = _mesh("shells")
    E = R.TangentPointEnergy0(q, p)
    en64 = E.Value(M)
    en32 = R.TangentPointEnergy0(q, p, precision="mixed").Value(M)
    print(f"shells mixed energy rel err {abs(en32 - en64) / abs(en64):.2e}", flush=True)
    assert abs(en32 - en64) <= 1e-7 * abs(en64)


def test_vf_semistatic_update():
    M, q, p = _mesh("circles")
    E = R.TangentPointEnergy0(q, p)
    E.Value(M)
    X = M.VertexCoordinates()
    rng = np.random.default_rng(5)
    X2 = X + 1e-5 * rng.standard_normal(X.shape)
    M.SemiStaticUpdate(X2)
    en = E.Value(M)
    en_o, Y_o, _, _ = oracle_energy_diff(M, q, p)
    assert abs(en - en_o) <= 1e-10 * abs(en_o)
    assert rel(E.Differential(M), Y_o) <= 1e-10
