"""Pins of the oracle (CPU restatement of TangentPointEnergy0 / TangentPointMetric0), CPU only.

* energy: with the far field disabled (far theta = 0) the hierarchical energy equals the
  brute-force sum over all unordered pairs (TangentPointEnergy_AllPairs idea); with singleton leaves
  the far-field cluster formula equals the primitive formula; the far-field error vanishes as
  theta -> 0; closed form of a single pair;
* differential: central finite differences of the energy at a frozen hierarchy
  (SemiStaticUpdate), translation / rotation invariance and Euler homogeneity
  sum_v x_v . dE/dx_v = (2n + q - p) E;
* chain rule (DNearFarToHulls): finite differences of the moment map computed by the oracle's own
  simplex-data restatement;
* metric: symmetry <X, AY> = <AX, Y>, A 1 = 0, positive semidefinite, linearity, and the near-field
  blocks equal a dense all-pairs assembly.
"""

import numpy as np
import pytest

import oracle as O
import repulsor_amd as R
from helpers import oracle_energy_diff, oracle_metric, oracle_multiply, oracle_state, rel


def _mesh(X, S, leaf=4, far_theta=0.5, near_theta=10.0):
    M = R.SimplicialMesh(X, S)
    M.cluster_tree_settings.split_threshold = leaf
    M.block_cluster_tree_settings.far_field_separation_parameter = far_theta
    M.adaptivity_settings.theta = near_theta
    M.build()
    return M


@pytest.fixture(scope="module")
def torus():
    return R.make_torus(24, 12, 1.0, 0.5, 1e-3, 1)


@pytest.fixture(scope="module")
def curve():
    return R.make_trefoil(400, 16, 0.02, 3)


# ------------------------------------------------------------------------------------------------ energy
def test_single_pair_closed_form():
    """Two triangles: E = a b (|phi|^{q/2} + |psi|^{q/2}) / r^p with q = 6, p = 12, computed by hand."""
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [3, 0, 1], [3, 1, 1], [3, 0, 2.0]])
    S = np.array([[0, 1, 2], [3, 4, 5]], dtype=np.int32)
    near, far, _, _ = O.simplex_data(X, S)
    a, b = 0.5, 0.5
    x = X[:3].mean(0); y = X[3:].mean(0)
    v = y - x
    n1 = np.array([0, 0, 1.0]); n2 = np.array([1.0, 0, 0])
    phi = (v @ n1) ** 2; psi = (v @ n2) ** 2
    E = a * b * (phi ** 3 + psi ** 3) / (v @ v) ** 6
    assert abs(O.allpairs_energy(2, far, 6, 12) - E) <= 1e-14 * E


@pytest.mark.parametrize("which", ["torus", "curve"])
def test_energy_equals_brute_force_without_far_field(which, torus, curve):
    X, S = torus if which == "torus" else curve
    q, p = (6.0, 12.0) if which == "torus" else (2.0, 4.0)
    M = _mesh(X, S, leaf=4, far_theta=0.0)
    b = M.build()
    assert b["far_nnz"] == 0
    en, _, _, _ = oracle_energy_diff(M, q, p)
    near, far, _, _ = O.simplex_data(X, S)
    ap = O.allpairs_energy(S.shape[1] - 1, far, q, p)
    assert abs(en - ap) <= 1e-12 * abs(ap)


def test_far_field_singleton_leaves_is_primitive_formula(torus):
    """split_threshold = 1: a far block (C, D) of two leaves is the primitive pair (i, j), so the
    hierarchical energy equals the brute force up to the near-field barycenter rounding."""
    X, S = torus
    M = _mesh(X, S, leaf=1, far_theta=0.5)
    tr, td, npat, fpat = oracle_state(M)
    o = O.tp0_compute(2, td["P_near"], td["C_far"], npat, fpat, 6, 12, 1)
    # restrict the far pattern to leaf-leaf blocks and compare block by block with the pair formula
    L = tr["C_left"]
    frp, fci, fdg = fpat
    C_far = td["C_far"]; P_far = td["P_far"]
    checked = 0
    for c in range(len(frp) - 1):
        if L[c] >= 0:
            continue
        for k in range(fdg[c], frp[c + 1]):
            d = fci[k]
            if L[d] >= 0:
                continue
            i, j = tr["C_begin"][c], tr["C_begin"][d]
            assert np.allclose(C_far[c], P_far[i], rtol=0, atol=0)
            assert np.allclose(C_far[d], P_far[j], rtol=0, atol=0)
            checked += 1
    assert checked > 0
    assert np.isfinite(o["energy"])


def test_far_field_error_vanishes_with_theta(torus):
    X, S = torus
    near, far, _, _ = O.simplex_data(X, S)
    ap = O.allpairs_energy(2, far, 6, 12)
    errs = []
    for th in (0.5, 0.25, 0.125):
        M = _mesh(X, S, leaf=4, far_theta=th)
        en, _, _, _ = oracle_energy_diff(M, 6, 12)
        errs.append(abs(en - ap) / ap)
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-3


# ------------------------------------------------------------------------------------------------ differential
@pytest.mark.parametrize("which", ["torus", "curve"])
def test_differential_finite_differences(which, torus, curve):
    X, S = torus if which == "torus" else curve
    q, p = (6.0, 12.0) if which == "torus" else (2.0, 4.0)
    M = _mesh(X, S, leaf=4, far_theta=0.5)
    en, G, _, _ = oracle_energy_diff(M, q, p)
    rng = np.random.default_rng(0)
    def cfd(D, h):
        M.SemiStaticUpdate(X + h * D)
        ep, _, _, _ = oracle_energy_diff(M, q, p, X=X + h * D)
        M.SemiStaticUpdate(X - h * D)
        em, _, _, _ = oracle_energy_diff(M, q, p, X=X - h * D)
        return (ep - em) / (2 * h)

    for trial in range(2):
        D = rng.standard_normal(X.shape)
        h = 1e-5
        fd = (4 * cfd(D, h / 2) - cfd(D, h)) / 3          # Richardson: O(h^4)
        assert abs(fd - np.sum(G * D)) <= 1e-7 * np.linalg.norm(G) * np.linalg.norm(D)
    M.SemiStaticUpdate(X)


@pytest.mark.parametrize("which", ["torus", "curve"])
def test_differential_invariants(which, torus, curve):
    X, S = torus if which == "torus" else curve
    q, p = (6.0, 12.0) if which == "torus" else (2.0, 4.0)
    n = S.shape[1] - 1
    M = _mesh(X, S, leaf=4, far_theta=0.5)
    en, G, _, _ = oracle_energy_diff(M, q, p)
    scale = np.linalg.norm(G) * np.linalg.norm(X)
    assert np.linalg.norm(G.sum(axis=0)) <= 1e-12 * np.linalg.norm(G) * np.sqrt(len(X))      # translation
    assert np.linalg.norm(np.cross(X, G).sum(axis=0)) <= 1e-12 * scale                        # rotation
    assert abs(np.sum(X * G) - (2 * n + q - p) * en) <= 1e-11 * max(abs(en), scale)           # homogeneity


def test_chain_rule_finite_differences(torus):
    """orc_dnearfar_to_hulls == transpose Jacobian of the moment map (a, a hull, a P), (a, a x, a P)."""
    X, S = torus
    S = S[:50]
    rng = np.random.default_rng(1)
    N = rng.standard_normal((50, 16))
    F = rng.standard_normal((50, 10))
    B = O.dnearfar_to_hulls(X, S, N, F)

    def moments(Xp):
        near, far, _, _ = O.simplex_data(Xp, S)
        a = near[:, :1]
        mn = np.concatenate([a, a * near[:, 1:]], axis=1)
        mf = np.concatenate([a, a * far[:, 1:]], axis=1)
        return np.sum(N * mn) + np.sum(F * mf)

    h = 1e-6
    for t in (0, 7, 31):
        for j in range(3):
            for c in range(3):
                v = S[t, j]
                Xp = X.copy(); Xp[v, c] += h
                Xm = X.copy(); Xm[v, c] -= h
                # only simplex t's contribution: restrict moments to t by zeroing others
                Nt = np.zeros_like(N); Nt[t] = N[t]
                Ft = np.zeros_like(F); Ft[t] = F[t]
                nearp, farp, _, _ = O.simplex_data(Xp, S[t:t + 1])
                nearm, farm, _, _ = O.simplex_data(Xm, S[t:t + 1])

                def mom(near, far):
                    a = near[:, :1]
                    return np.sum(N[t] * np.concatenate([a, a * near[:, 1:]], 1)) + np.sum(F[t] * np.concatenate([a, a * far[:, 1:]], 1))

                fd = (mom(nearp, farp) - mom(nearm, farm)) / (2 * h)
                assert abs(fd - B[t, j, c]) <= 1e-6 * max(1.0, abs(fd))


# ------------------------------------------------------------------------------------------------ metric
@pytest.fixture(scope="module")
def metric_case(torus):
    X, S = torus
    M = _mesh(X, S, leaf=4, far_theta=0.5)
    return M


def _apply(M, X):
    return oracle_multiply(M, 6.0, 12.0, X)


def test_metric_symmetry_nullspace_psd(metric_case):
    M = metric_case
    rng = np.random.default_rng(3)
    X = rng.standard_normal((M.V, 3)); Y = rng.standard_normal((M.V, 3))
    AX, AY = _apply(M, X), _apply(M, Y)
    assert abs(np.sum(X * AY) - np.sum(AX * Y)) <= 1e-12 * np.linalg.norm(AX) * np.linalg.norm(Y)
    A1 = _apply(M, np.ones((M.V, 3)) * np.array([1.0, -2.0, 0.5]))
    assert np.linalg.norm(A1) <= 1e-12 * np.linalg.norm(AX) * np.sqrt(M.V) / np.linalg.norm(X)
    for _ in range(3):
        Z = rng.standard_normal((M.V, 3))
        assert np.sum(Z * _apply(M, Z)) >= 0
    assert rel(_apply(M, 2 * X - 3 * Y), 2 * AX - 3 * AY) <= 1e-13


def test_near_metric_blocks_equal_dense_assembly(torus):
    X, S = torus
    M = _mesh(X, S, leaf=4, far_theta=0.0)
    o, tr, td, npat, fpat = oracle_metric(M, 6.0, 12.0)
    dense = O.allpairs_metric(2, td["P_far"], 6.0, 12.0)
    rp, ci, dg = npat
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    assert np.max(np.abs(o["NF"] - dense[rows, ci])) <= 1e-12 * np.max(np.abs(dense))
    # diagonal rule: diag_i = -sum_j (a_j / a_i) block_ij
    a = td["P_far"][:, 0]
    dref = -(dense * (a[None, :, None] / a[:, None, None])).sum(axis=1)
    assert np.max(np.abs(o["NF_diag"] - dref)) <= 1e-11 * np.max(np.abs(dref))
