"""Pins of the metric-solve pieces on the CPU: the H^1 metric, its nested-dissection ordering and
supernodal Cholesky factor (host C++), the oracle's pseudo-Laplacian, the preconditioner
P0 = H^-1 L_{2-s} H^-1 composed from them, and preconditioned CG on the oracle operators.

Pins (none retypes the implementation's formula):
* H = K + M: H 1 = M 1 = (1/3) sum of incident triangle areas (1/2 for edges), and for a linear
  function f(x) = c.x the stiffness form equals sum_t a_t |P_t c|^2 (P_t the tangent projector,
  the P1 interpolant of a linear function is exact on flat simplices), the mass form equals
  sum_t a_t/12 ((sum f)^2 + sum f^2);
* the ordering is a permutation; the factor solves H x = b to 1e-12 and agrees with scipy's
  sparse LU solve;
* pseudo-Laplacian: with the far field disabled, <X, L X> equals the discrete Gagliardo form
  1/2 sum_{i != j near} a_i a_j |u_i - u_j|^2 |x_i - x_j|^{-(n + 2 s')} with u_i the vertex mean over
  simplex i (the quadratic form of the fractional Laplacian of order s'); with the far field on,
  L 1 = 0 and <X, L Y> = <L X, Y>;
* P0 is symmetric, positive semidefinite and annihilates H 1;
* PCG with P0 on the oracle metric reaches the true residual |B - A X| / |B| <= 1e-6.
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
import repulsor_amd as R
from helpers import OracleSolve, h1_matrix, oracle_energy_diff, oracle_state, rel


def _mesh(X, S, leaf=8, far_theta=0.5, near_theta=10.0):
    M = R.SimplicialMesh(X, S)
    M.cluster_tree_settings.split_threshold = leaf
    M.block_cluster_tree_settings.far_field_separation_parameter = far_theta
    M.adaptivity_settings.theta = near_theta
    M.build()
    return M


@pytest.fixture(scope="module")
def torus():
    X, S = R.make_torus(40, 20, 1.0, 0.5, 1e-3, 1)
    return X, S, _mesh(X, S)


@pytest.fixture(scope="module")
def curve():
    X, S = R.make_trefoil(600, 16, 0.02, 3)
    return X, S, _mesh(X, S)


def _areas(X, S):
    if S.shape[1] == 3:
        e1 = X[S[:, 1]] - X[S[:, 0]]
        e2 = X[S[:, 2]] - X[S[:, 0]]
        nu = np.cross(e1, e2)
        a = 0.5 * np.linalg.norm(nu, axis=1)
        nn = nu / np.linalg.norm(nu, axis=1)[:, None]
        P = np.eye(3)[None] - nn[:, :, None] * nn[:, None, :]
    else:
        t = X[S[:, 1]] - X[S[:, 0]]
        a = np.linalg.norm(t, axis=1)
        tt = t / a[:, None]
        P = tt[:, :, None] * tt[:, None, :]
    return a, P


@pytest.mark.parametrize("which", ["torus", "curve"])
def test_h1_metric_closed_forms(which, torus, curve):
    X, S, M = torus if which == "torus" else curve
    H = h1_matrix(M)
    a, P = _areas(X, S)
    n1 = S.shape[1]
    # H 1 = M 1: every vertex receives a_t / (n+1) from each incident simplex
    ref = np.zeros(M.V)
    np.add.at(ref, S.ravel(), np.repeat(a / n1, n1))
    assert rel(H @ np.ones(M.V), ref) <= 1e-13
    assert np.max(np.abs(H - H.T)) <= 1e-14 * np.max(np.abs(H))
    # linear function f = c.x: stiffness form = sum a_t |P_t c|^2, mass form in closed form
    c = np.array([0.3, -1.1, 0.7])
    f = X @ c
    fs = f[S]
    if n1 == 3:
        mass = np.sum(a / 12.0 * (fs.sum(axis=1) ** 2 + (fs ** 2).sum(axis=1)))
    else:
        mass = np.sum(a / 6.0 * (fs.sum(axis=1) ** 2 + (fs ** 2).sum(axis=1)))
    stiff = np.sum(a * (np.einsum("tij,j->ti", P, c) @ c))
    assert abs(f @ (H @ f) - (stiff + mass)) <= 1e-12 * (stiff + mass)


@pytest.mark.parametrize("which", ["torus", "curve"])
def test_numeric_refactorisation_after_update(which):
    """A moved mesh keeps its H^1 pattern: SemiStaticUpdate keeps the factor's symbolic structure and its
    next use refactorises the values only (SparseCholeskyNumeric). The solve after the move must solve the
    moved mesh's H^1 system, not the old one (which differs by far more than the tolerance)."""
    X, S = R.make_torus(40, 20, 1.0, 0.5, 1e-3, 1) if which == "torus" else R.make_trefoil(600, 16, 0.02, 3)
    M = _mesh(X, S)
    rng = np.random.default_rng(11)
    B = rng.standard_normal((M.V, 3))
    X_old, _ = M.h1_solve_host(B)
    H_old = h1_matrix(M)
    X2 = X + 0.01 * rng.standard_normal(X.shape)
    M.SemiStaticUpdate(X2)
    X_new, _ = M.h1_solve_host(B)
    H_new = h1_matrix(M)
    assert rel(H_new @ X_new, B) <= 1e-12
    assert rel(X_new, spla.spsolve(H_new.tocsc(), B)) <= 1e-10
    assert rel(H_new @ X_old, B) > 1e-6      # the move changed the system


@pytest.mark.parametrize("which", ["torus", "curve"])
def test_cholesky_factor_solves(which, torus, curve):
    X, S, M = torus if which == "torus" else curve
    perm = M.nd_ordering()
    assert np.array_equal(np.sort(perm), np.arange(M.V))
    rng = np.random.default_rng(7)
    B = rng.standard_normal((M.V, 3))
    Xs, stats = M.h1_solve_host(B)
    H = h1_matrix(M)
    assert rel(H @ Xs, B) <= 1e-12
    ref = spla.spsolve(H.tocsc(), B)
    assert rel(Xs, ref) <= 1e-10
    assert stats["nnz_L"] < 0.2 * M.V * M.V   # the ordering keeps the factor sparse


def _pl(M, s_pl, far=True):
    tr, td, npat, fpat = oracle_state(M)
    pl = O.pl_compute(M.dom_dim, td["P_near"], td["C_far"], npat, fpat, s_pl)
    return pl, tr, td, npat, fpat


def _pl_apply(M, s_pl, X):
    pl, tr, td, npat, fpat = _pl(M, s_pl)
    Y = np.zeros_like(X)
    return O.pl_multiply(M.dom_dim, M.V, tr, td["pre_cols"], td["pre_vals"], npat, pl, fpat, 1.0, X, 0.0, Y)


@pytest.mark.parametrize("which,s_pl", [("torus", 1.0 / 3.0), ("curve", 0.5)])
def test_pseudo_laplacian_is_gagliardo_form(which, s_pl, torus, curve):
    X, S, _ = torus if which == "torus" else curve
    M = _mesh(X, S, far_theta=0.0)
    n = M.dom_dim
    tr, td, npat, fpat = oracle_state(M)
    assert len(fpat[1]) == 0
    rng = np.random.default_rng(2)
    U = rng.standard_normal((M.V, 3))
    LU = _pl_apply(M, s_pl, U)
    a = td["P_far"][:, 0]
    x = td["P_far"][:, 1:4]
    ubar = U[td["pre_cols"]].mean(axis=1)
    rp, ci, _ = npat
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    d2 = np.sum((x[rows] - x[ci]) ** 2, axis=1)
    w = a[rows] * a[ci] * d2 ** (-(n / 2.0 + s_pl))
    q_ref = 0.5 * np.sum(w * np.sum((ubar[rows] - ubar[ci]) ** 2, axis=1))
    assert abs(np.sum(U * LU) - q_ref) <= 1e-11 * q_ref


@pytest.mark.parametrize("which,s_pl", [("torus", 1.0 / 3.0), ("curve", 0.5)])
def test_pseudo_laplacian_nullspace_symmetry(which, s_pl, torus, curve):
    _, _, M = torus if which == "torus" else curve
    rng = np.random.default_rng(4)
    X = rng.standard_normal((M.V, 3)); Y = rng.standard_normal((M.V, 3))
    LX, LY = _pl_apply(M, s_pl, X), _pl_apply(M, s_pl, Y)
    assert abs(np.sum(X * LY) - np.sum(LX * Y)) <= 1e-12 * np.linalg.norm(LX) * np.linalg.norm(Y)
    L1 = _pl_apply(M, s_pl, np.ones((M.V, 3)))
    assert np.linalg.norm(L1) <= 1e-12 * np.linalg.norm(LX) * np.sqrt(M.V) / np.linalg.norm(X)


def test_preconditioner_properties(torus):
    _, _, M = torus
    S = OracleSolve(M, 6.0, 12.0)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((M.V, 3)); Y = rng.standard_normal((M.V, 3))
    PX, PY = S.P(X), S.P(Y)
    assert abs(np.sum(X * PY) - np.sum(PX * Y)) <= 1e-10 * np.linalg.norm(PX) * np.linalg.norm(Y)
    assert np.sum(X * PX) > 0
    H1 = h1_matrix(M) @ np.ones((M.V, 3))
    assert np.linalg.norm(S.P(H1)) <= 1e-10 * np.linalg.norm(PX) * np.linalg.norm(H1) / np.linalg.norm(X)


def test_oracle_pcg_converges(torus):
    _, _, M = torus
    S = OracleSolve(M, 6.0, 12.0)
    _, B, _, _ = oracle_energy_diff(M, 6.0, 12.0)
    X, it = S.pcg(B, tol=1e-6)
    res = np.linalg.norm(B - S.A(X), axis=0) / np.linalg.norm(B, axis=0)
    assert np.all(res <= 1.5e-6), res
    assert it < 200


def test_curve_alternating_mode_is_invisible_to_p0(curve):
    """On a closed curve with an even vertex count the edge-average pre-operator annihilates the
    alternating vertex mode, hence L_{2-s} and P0 do as well (why curve solves use the
    preconditioner-norm measure)."""
    X, S, M = curve
    assert M.V % 2 == 0 and np.array_equal(S[:, 1], (S[:, 0] + 1) % M.V)
    alt = np.repeat(((-1.0) ** np.arange(M.V))[:, None], 3, axis=1)
    LX = _pl_apply(M, 0.5, np.ascontiguousarray(alt))
    assert np.linalg.norm(LX) <= 1e-13 * np.linalg.norm(_pl_apply(M, 0.5, np.ascontiguousarray(np.roll(alt, 1, axis=0) + X)))
    So = OracleSolve(M, 2.0, 4.0)
    alt_h = np.ascontiguousarray(h1_matrix(M) @ alt)
    rng = np.random.default_rng(1)
    Z = rng.standard_normal((M.V, 3))
    assert np.linalg.norm(So.P(alt_h)) <= 1e-10 * np.linalg.norm(So.P(Z)) * np.linalg.norm(alt_h) / np.linalg.norm(Z)
