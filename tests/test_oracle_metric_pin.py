"""Pins of the oracle's tangent-point metric (TangentPointMetric0 pair values and product), CPU only.

The metric blocks of the reference are (TP0_Kernel_NF.hpp:79-80, 315-331)

    a1 = -2 r^-(n + 2(s - 1)),   a0 = 1/2 (<N_i v>^2 + <N_j v>^2) r^-4 a1,   s = (p - n) / q,

with v = x_j - x_i between barycenters and N the normal projector (triangles: n n^T with the unit
normal from a cross product; edges: I - t t^T with the unit tangent). Applied through the pre/post
operators (area x mean value, area x surface gradient) with the diagonal rule
diag_i = -sum_j (a_j / a_i) block_ij, the quadratic form of the metric with the far field off is the
discrete Sobolev-Slobodeckij form

    <U, A U> = 1/2 sum_{i != j} a_i a_j [ -a0_ij |U_i - U_j|^2 - a1_ij |grad U_i - grad U_j|^2 ],

U_i the mean of U over simplex i and grad U_i its surface gradient. Everything on the right-hand side
is computed here from the vertex coordinates alone: cross-product normals, edge-rotation gradients
(grad lambda_k = n x e_k / (2 area), e_k the edge opposite vertex k), no call into the oracle's
simplex data or pair routine. A plausible mistake in the oracle's metric (r^2 for r^4 in a0, a
dropped 1/2, the exponent shifted by 1/3) changes the form by far more than the tolerance; the
tests assert that margin too, so they discriminate.
"""

import numpy as np
import pytest

import oracle as O
import repulsor_amd as R
from helpers import oracle_multiply, oracle_state


def _geometry(X, S):
    """Areas / lengths, barycenters, normal projectors (T x 3 x 3) and P1 gradient operators
    (T x (n+1) x 3: gradient of the k-th barycentric coordinate), from coordinates only."""
    n = S.shape[1] - 1
    P = X[S]                                         # T x (n+1) x 3
    x = P.mean(axis=1)
    if n == 2:
        e1 = P[:, 1] - P[:, 0]
        e2 = P[:, 2] - P[:, 0]
        N = np.cross(e1, e2)
        a2 = np.linalg.norm(N, axis=1)
        nh = N / a2[:, None]
        a = 0.5 * a2
        proj = nh[:, :, None] * nh[:, None, :]
        G = np.empty((len(S), 3, 3))
        for k in range(3):
            opp = P[:, (k + 2) % 3] - P[:, (k + 1) % 3]   # edge opposite vertex k, counterclockwise
            G[:, k] = np.cross(nh, opp) / a2[:, None]
    else:
        e = P[:, 1] - P[:, 0]
        a = np.linalg.norm(e, axis=1)
        t = e / a[:, None]
        proj = np.eye(3)[None] - t[:, :, None] * t[:, None, :]
        G = np.empty((len(S), 2, 3))
        G[:, 1] = e / (a * a)[:, None]
        G[:, 0] = -G[:, 1]
    return a, x, proj, G


def _kernels(x, proj, n, q, p, r4_power=2.0, half=0.5, exp_shift=0.0):
    """Dense (a0, a1) of every pair from the Sobolev-Slobodeckij form; the keyword arguments plant
    the mistakes the tests must detect (defaults: the correct form)."""
    s = (p - n) / q
    v = x[None, :, :] - x[:, None, :]                # v_ij = x_j - x_i
    r2 = np.sum(v * v, axis=2)
    np.fill_diagonal(r2, 1.0)
    r = np.sqrt(r2)
    a1 = -2.0 * r ** (-(n + 2.0 * (s - 1.0)) + exp_shift)
    phi = np.einsum("ija,iab,ijb->ij", v, proj, v)   # <v, N_i v>
    psi = np.einsum("ija,jab,ijb->ij", v, proj, v)   # <v, N_j v>
    a0 = half * (phi + psi) * r ** (-2.0 * r4_power) * a1
    np.fill_diagonal(a0, 0.0)
    np.fill_diagonal(a1, 0.0)
    return a0, a1


def _slobodeckij_form(a, Ubar, gU, a0, a1):
    w = a[:, None] * a[None, :]
    dm = np.sum((Ubar[:, None, :] - Ubar[None, :, :]) ** 2, axis=2)
    dg = np.sum((gU[:, None] - gU[None, :]) ** 2, axis=(2, 3))
    return 0.5 * np.sum(w * (-a0 * dm - a1 * dg))


def _mesh(X, S, leaf=4):
    M = R.SimplicialMesh(X, S)
    M.cluster_tree_settings.split_threshold = leaf
    M.block_cluster_tree_settings.far_field_separation_parameter = 0.0   # far field off
    M.adaptivity_settings.theta = 1.0e4                                  # every separate pair is Near
    M.build()
    return M


@pytest.fixture(scope="module")
def torus():
    return R.make_torus(24, 12, 1.0, 0.5, 1e-3, 1), 6.0, 12.0


@pytest.fixture(scope="module")
def curve():
    return R.make_trefoil(400, 16, 0.02, 3), 2.0, 4.0


# ------------------------------------------------------------------------------------------------ single pair
def test_single_pair_metric_closed_form():
    """Two triangles in perpendicular planes: (a0, a1) by hand, q = 6, p = 12 (s = 5/3): a1 = -2 r^(-10/3),
    a0 = (<n1 v>^2 + <n2 v>^2) / (2 r^4) a1 with n1 = e_z, n2 = e_x."""
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [3, 0, 1], [3, 1, 1], [3, 0, 2.0]])
    S = np.array([[0, 1, 2], [3, 4, 5]], dtype=np.int32)
    _, far, _, _ = O.simplex_data(X, S)
    v = X[3:].mean(0) - X[:3].mean(0)              # (3, 1/3 - 1/3, 4/3 - 0) = (3, 0, 4/3)
    r = np.sqrt(v @ v)
    a1 = -2.0 * r ** (-10.0 / 3.0)
    a0 = 0.5 * (v[2] ** 2 + v[0] ** 2) / r ** 4 * a1
    m = O.allpairs_metric(2, far, 6.0, 12.0)
    assert abs(m[0, 1, 0] - a0) <= 1e-14 * abs(a0)
    assert abs(m[0, 1, 1] - a1) <= 1e-14 * abs(a1)
    assert np.array_equal(m[0, 1], m[1, 0])


def test_single_pair_metric_product_closed_form():
    """Two separated triangles through the whole oracle path (tree, pre-operator, block product,
    diagonal rule, post-operator): <U, A U> = -a b [a0 |U_1 - U_2|^2 + a1 |grad U_1 - grad U_2|^2]."""
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [3, 0, 1], [3, 1, 1], [3, 0, 2.0]])
    S = np.array([[0, 1, 2], [3, 4, 5]], dtype=np.int32)
    M = _mesh(X, S, leaf=1)
    rng = np.random.default_rng(5)
    U = rng.standard_normal((6, 3))
    q_or = np.sum(U * oracle_multiply(M, 6.0, 12.0, U))
    a, x, proj, G = _geometry(X, S)
    a0, a1 = _kernels(x, proj, 2, 6.0, 12.0)
    gU = np.einsum("tkc,tkd->tcd", G, U[S])       # T x 3 (ambient) x 3 (columns)
    ref = -a[0] * a[1] * (a0[0, 1] * np.sum((U[S[0]].mean(0) - U[S[1]].mean(0)) ** 2)
                          + a1[0, 1] * np.sum((gU[0] - gU[1]) ** 2))
    assert abs(q_or - ref) <= 1e-12 * abs(ref)


# ------------------------------------------------------------------------------------------------ quadratic form
@pytest.mark.parametrize("which", ["torus", "curve"])
def test_metric_quadratic_form_is_sobolev_slobodeckij(which, torus, curve):
    (X, S), q, p = torus if which == "torus" else curve
    n = S.shape[1] - 1
    M = _mesh(X, S)
    b = M.build()
    assert b["far_nnz"] == 0 and b["verynear_nnz"] == 0
    rng = np.random.default_rng(7)
    a, x, proj, G = _geometry(X, S)
    for trial in range(2):
        U = rng.standard_normal((M.V, 3)) if trial == 0 else np.sin(2.0 * X) + X[:, [1, 2, 0]] ** 2
        q_or = np.sum(U * oracle_multiply(M, q, p, U))
        Ubar = U[S].mean(axis=1)
        gU = np.einsum("tkc,tkd->tcd", G, U[S])
        a0, a1 = _kernels(x, proj, n, q, p)
        ref = _slobodeckij_form(a, Ubar, gU, a0, a1)
        assert abs(q_or - ref) <= 1e-11 * abs(ref), (which, trial, q_or, ref)
        # the planted mistakes move the form far outside the tolerance
        for kw in ({"r4_power": 1.0}, {"half": 1.0}, {"exp_shift": 2.0 / 3.0}, {"exp_shift": -2.0 / 3.0}):
            m0, m1 = _kernels(x, proj, n, q, p, **kw)
            bad = _slobodeckij_form(a, Ubar, gU, m0, m1)
            assert abs(bad - ref) > 1e-6 * abs(ref), kw


def test_metric_pair_values_equal_independent_kernels(torus):
    """Every near block the oracle assembles equals the independently written (a0, a1)."""
    (X, S), q, p = torus
    M = _mesh(X, S)
    tr, td, npat, fpat = oracle_state(M)
    o = O.tp0_compute(2, td["P_near"], td["C_far"], npat, fpat, q, p, 4)
    perm = np.asarray(tr["P_ordering"])          # tree position -> simplex
    _, x, proj, _ = _geometry(X, S)
    a0, a1 = _kernels(x, proj, 2, q, p)
    rp, ci, _ = npat
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    i, j = perm[rows], perm[np.asarray(ci)]
    scale = np.max(np.abs(a1))
    assert np.max(np.abs(o["NF"][:, 0] - a0[i, j])) <= 1e-12 * np.max(np.abs(a0))
    assert np.max(np.abs(o["NF"][:, 1] - a1[i, j])) <= 1e-12 * scale
