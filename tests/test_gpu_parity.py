"""GPU parity of the HIP path against the oracle (CPU restatement of the reference), same mesh,
same hierarchy, same q, p, theta.

Tolerances (FP64 path): energy and differential relative 1e-10, MultiplyMetric relative 1e-10
in the 2-norm. Metric blocks are compared entrywise at relative 1e-12 of the largest entry.
"""

import numpy as np
import pytest

import repulsor_amd as R
from repulsor_amd.workloads import WORKLOADS, random_field
from helpers import oracle_energy_diff, oracle_metric, oracle_multiply, rel

pytestmark = pytest.mark.gpu

CASES = ["torus_2400", "sphere_1280", "trefoil_3000"]


def _mesh(name):
    wl = WORKLOADS[name]
    M = wl.mesh()
    M.build()
    return wl, M


@pytest.mark.parametrize("name", CASES)
def test_energy_value(name):
    wl, M = _mesh(name)
    print(f"case {name}", flush=True)
    E = R.TangentPointEnergy0(wl.q, wl.p)
    en_gpu = E.Value(M)
    en_ref, _, _, _ = oracle_energy_diff(M, wl.q, wl.p)
    assert abs(en_gpu - en_ref) <= 1e-10 * abs(en_ref)


@pytest.mark.parametrize("name", CASES)
def test_differential(name):
    wl, M = _mesh(name)
    print(f"case {name}", flush=True)
    E = R.TangentPointEnergy0(wl.q, wl.p)
    Y, en = E.Differential(M, return_energy=True)
    en_ref, Y_ref, N_ref, F_ref = oracle_energy_diff(M, wl.q, wl.p)
    assert abs(en - en_ref) <= 1e-10 * abs(en_ref)
    assert rel(Y, Y_ref) <= 1e-10
    N, F = E.MomentDerivatives(M)
    assert rel(N, N_ref) <= 1e-10
    assert rel(F, F_ref) <= 1e-10


def test_differential_alpha_beta():
    wl, M = _mesh("torus_2400")
    E = R.TangentPointEnergy0(wl.q, wl.p)
    G = E.Differential(M)
    Y0 = random_field(M.V, seed=5)
    Y = Y0.copy()
    E.Differential(M, alpha=-0.5, beta=2.0, Y=Y)
    assert rel(Y, -0.5 * G + 2.0 * Y0) <= 1e-13


@pytest.mark.parametrize("name", CASES)
def test_metric_values(name):
    wl, M = _mesh(name)
    print(f"case {name}", flush=True)
    A = R.TangentPointMetric0(wl.q, wl.p)
    vals = A.MetricValues(M)
    o = oracle_metric(M, wl.q, wl.p)[0]
    for key in ("NF", "FF", "NF_diag", "FF_diag"):
        ref = o[key]
        scale = max(1.0, float(np.max(np.abs(ref)))) if ref.size else 1.0
        assert np.max(np.abs(vals[key] - ref)) <= 1e-12 * scale, key


@pytest.mark.parametrize("name", CASES)
def test_multiply_metric(name):
    wl, M = _mesh(name)
    print(f"case {name}", flush=True)
    A = R.TangentPointMetric0(wl.q, wl.p)
    X = random_field(M.V, seed=11)
    Y = np.zeros_like(X)
    A.MultiplyMetric(M, 1.0, X, 0.0, Y)
    Y_ref = oracle_multiply(M, wl.q, wl.p, X)
    assert rel(Y, Y_ref) <= 1e-10
    # alpha/beta and the flags
    Y0 = random_field(M.V, seed=12)
    Y2 = Y0.copy()
    A.MultiplyMetric(M, 2.0, X, -1.0, Y2)
    assert rel(Y2, 2.0 * Y_ref - Y0) <= 1e-10
    Ynf = np.zeros_like(X)
    A.MultiplyMetric(M, 1.0, X, 0.0, Ynf, FF=False)
    assert rel(Ynf, oracle_multiply(M, wl.q, wl.p, X, ff=False)) <= 1e-10


def test_multiply_metric_nrhs_error():
    wl, M = _mesh("torus_2400")
    A = R.TangentPointMetric0(wl.q, wl.p)
    X = random_field(M.V, nrhs=2)
    Y = np.full((M.V, 2), 7.0)
    with pytest.raises(R.RepulsorError):
        A.MultiplyMetric(M, 1.0, X, 0.0, Y, nrhs=2)
    assert np.all(Y == 7.0)


def test_semistatic_update_parity():
    wl, M = _mesh("sphere_1280")
    E = R.TangentPointEnergy0(wl.q, wl.p)
    e0 = E.Value(M)
    X = M.VertexCoordinates()
    X2 = X + 1e-3 * random_field(M.V, seed=3)
    M.SemiStaticUpdate(X2)
    e1 = E.Value(M)
    assert e1 != e0
    Y, en = E.Differential(M, return_energy=True)
    en_ref, Y_ref, _, _ = oracle_energy_diff(M, wl.q, wl.p, X=X2)
    assert abs(en - en_ref) <= 1e-10 * abs(en_ref)
    assert abs(e1 - en_ref) <= 1e-10 * abs(en_ref)
    assert rel(Y, Y_ref) <= 1e-10


def test_repeatability_bitwise():
    wl, M = _mesh("sphere_1280")
    E = R.TangentPointEnergy0(wl.q, wl.p)
    Y1, e1 = E.Differential(M, return_energy=True)
    Y2, e2 = E.Differential(M, return_energy=True)
    assert e1 == e2
    assert np.array_equal(Y1, Y2)


# ------------------------------------------------------------------------------------------------ real exponents
# The run-time exponent kernel (KIND_GENERIC: real q/2 and p/2 through pow, q = 0) against the oracle,
# which restates the reference's Power as binary powering for integer and std::pow for real exponents
# (TP0_Traversor.hpp:169-245). Tolerances as above: the device pow and the host std::pow differ by ulps.
GENERIC = [
    ("torus_2400", 5.5, 11.3, True),       # real q/2 and p/2, metric s_exp through pow
    ("trefoil_3000", 3.0, 5.5, True),       # curve, odd q
    ("torus_2400", 0.0, 4.5, False),        # q = 0: E = 2 r^-p, no projector terms (metric undefined: s = inf)
]


@pytest.mark.parametrize("name,q,p,metric", GENERIC)
def test_generic_exponents(name, q, p, metric):
    wl, M = _mesh(name)
    print(f"case {name} q={q} p={p}", flush=True)
    E = R.TangentPointEnergy0(q, p)
    Y, en = E.Differential(M, return_energy=True)
    en_ref, Y_ref, _, _ = oracle_energy_diff(M, q, p)
    assert abs(en - en_ref) <= 1e-10 * abs(en_ref)
    assert rel(Y, Y_ref) <= 1e-10
    if metric:
        A = R.TangentPointMetric0(q, p)
        X = random_field(M.V, seed=21)
        Z = np.zeros_like(X)
        A.MultiplyMetric(M, 1.0, X, 0.0, Z)
        assert rel(Z, oracle_multiply(M, q, p, X)) <= 1e-10


def test_single_parameter_constructor_odd_p():
    """TangentPointEnergy0(p) means q = p, p = 2p (TangentPointEnergy0.hpp:35-39): p = 3 gives q/2 = 1.5."""
    wl, M = _mesh("sphere_1280")
    E = R.TangentPointEnergy0(3.0)
    assert (E.q, E.p) == (3.0, 6.0)
    Y, en = E.Differential(M, return_energy=True)
    en_ref, Y_ref, _, _ = oracle_energy_diff(M, 3.0, 6.0)
    assert abs(en - en_ref) <= 1e-10 * abs(en_ref)
    assert rel(Y, Y_ref) <= 1e-10
