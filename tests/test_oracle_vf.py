"""Pins of the oracle's very-near-field (VF) restatement, CPU only.

* the exact simplex distance used by the MAC against dense sampling of both simplices;
* the adaptive quadrature: as theta shrinks, the weighted energy, the metric value a1 and the
  pseudo-Laplacian value of a very-near triangle pair converge to a Richardson-extrapolated uniform
  midpoint rule built from independent geometry (midpoint subdivision of the vertex coordinates);
* the differential against central differences of the weighted energy, per moment (area, hull
  vertices, projector) of a few primitives;
* the reference's unweighted energy sum (TP0_Kernel_VF.hpp:273): with every evaluation at level 1
  the unweighted energy is (children per simplex)^2 times the weighted one;
* the metric blocks with their diagonal annihilate constants, lower entries equal upper entries,
  and the same for the pseudo-Laplacian.
"""

import numpy as np
import pytest

import oracle as O
from helpers import vf_circles, vf_shells

Q6, P12 = 6.0, 12.0


def _sample(P, n):
    if len(P) == 2:
        t = np.linspace(0.0, 1.0, n + 1)
        return (1 - t)[:, None] * P[0] + t[:, None] * P[1]
    pts = [(1 - i / n - j / n) * P[0] + (i / n) * P[1] + (j / n) * P[2] for i in range(n + 1) for j in range(n + 1 - i)]
    return np.array(pts)


def test_simplex_distance_against_sampling():
    rng = np.random.default_rng(0)
    for m in (2, 3):
        for trial in range(30):
            P = rng.standard_normal((m, 3))
            d = rng.standard_normal(3)
            d *= (3.0 + rng.uniform(0, 2)) / np.linalg.norm(d)
            Q = rng.standard_normal((m, 3)) * rng.uniform(0.2, 1.0) + d
            if trial % 3 == 0:      # nearly parallel, close: a shifted copy
                Q = P + d / np.linalg.norm(d) * 0.05 + 0.01 * rng.standard_normal((m, 3))
            n = 80 if m == 2 else 50
            SP, SQ = _sample(P, n), _sample(Q, n)
            brute = np.min(np.sum((SP[:, None, :] - SQ[None, :, :]) ** 2, axis=2))
            d2 = O.simplex_dist2(P, Q)
            diam = max(np.max(np.linalg.norm(P[:, None] - P[None], axis=2)), np.max(np.linalg.norm(Q[:, None] - Q[None], axis=2)))
            assert d2 <= brute * (1 + 1e-12) + 1e-300
            assert np.sqrt(brute) - np.sqrt(d2) <= 2.0 * diam / n


def _subtris(P, L):
    tris = [np.asarray(P, float)]
    for _ in range(L):
        nt = []
        for t in tris:
            a, b, c = t
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            nt += [np.array([a, ab, ca]), np.array([ab, b, bc]), np.array([ca, bc, c]), np.array([ab, bc, ca])]
        tris = nt
    return np.array(tris)


def _uniform(PA, PB, L, q, p, sexp_m, sexp_pl):
    """Uniform midpoint product rule at level L (4^L sub-triangles each), normalized by the areas."""
    TA, TB = _subtris(PA, L), _subtris(PB, L)
    ca, cb = TA.mean(1), TB.mean(1)

    def proj(P):
        nrm = np.cross(P[1] - P[0], P[2] - P[0])
        nrm /= np.linalg.norm(nrm)
        return np.outer(nrm, nrm)

    Pm, Qm = proj(PA), proj(PB)
    w = 1.0 / (len(TA) * len(TB))
    E = a1 = pl = 0.0
    for i in range(len(ca)):
        v = cb - ca[i]
        r2 = np.sum(v * v, 1)
        phi = np.einsum("ij,jk,ik->i", v, Pm, v)
        psi = np.einsum("ij,jk,ik->i", v, Qm, v)
        E += w * np.sum((np.abs(phi) ** (q / 2) + np.abs(psi) ** (q / 2)) * r2 ** (-p / 2))
        a1 += w * np.sum(-2 * r2 ** sexp_m)
        pl += w * np.sum(-r2 ** sexp_pl)
    return np.array([E, a1, pl])


def test_adaptive_quadrature_converges_to_the_double_integral():
    PA = np.array([[0, 0, 0], [1, 0, 0], [0.2, 0.9, 0.0]])
    PB = np.array([[0.1, 0.05, 0.05], [1.1, 0.1, 0.054], [0.3, 1.0, 0.047]])
    near, _, _, _ = O.simplex_data(np.vstack([PA, PB]), np.array([[0, 1, 2], [3, 4, 5]], np.int32))
    a = near[:, 0]
    pat = (np.array([0, 1, 2], np.int64), np.array([1, 0], np.int32), np.array([0, 2], np.int64))
    n = 2
    s = (P12 - n) / Q6
    sexp_m = -(s - 1) - n / 2
    s_pl = 2 - s
    sexp_pl = -n / 2 - s_pl
    # the pair is very near at the reference's theta = 10: max r^2 >= 100 dist^2
    assert max(((P - P.mean(0)) ** 2).sum(1).max() for P in (PA, PB)) >= 100 * O.simplex_dist2(PA, PB)
    u5 = _uniform(PA, PB, 5, Q6, P12, sexp_m, sexp_pl)
    u6 = _uniform(PA, PB, 6, Q6, P12, sexp_m, sexp_pl)
    ref = (4 * u6 - u5) / 3
    errs = []
    for th in (2.0, 1.0, 0.5):
        o = O.vf_compute(n, near, pat, th, 30, Q6, P12, 1 | 4 | 8)
        pl = O.pl_vf(n, near, pat, th, 30, s_pl)
        got = np.array([o["energy"] / (a[0] * a[1]), o["VF"][0, 1], pl["VF"][0]])
        errs.append(np.abs(got / ref - 1))
        assert o["VF"][0, 1] == o["VF"][1, 1]
    errs = np.array(errs)
    print("relative errors (energy, a1, pl) at theta 2, 1, 0.5:\n", errs)
    assert np.all(errs[-1] <= 0.015)
    assert np.all(errs[2, 1:] < errs[1, 1:]) and np.all(errs[1, 1:] < errs[0, 1:])


def _vf_state(M):
    M.build()
    return M.tree_data()["P_near"], M.pattern("verynear")


@pytest.mark.parametrize("make,n,q,p,prims", [
    (lambda: vf_circles(400, 3e-4), 1, 2.0, 4.0, 3),
    (lambda: vf_shells(2, 0.006), 2, Q6, P12, 2),
])
def test_vf_differential_matches_weighted_energy(make, n, q, p, prims):
    M = make()
    near, pat = _vf_state(M)
    assert len(pat[1]) > 0
    o = O.vf_compute(n, near, pat, 10.0, 30, q, p, 2)
    D = o["P_D_near_tree"]
    S = n + 1
    rows = [i for i in range(near.shape[0]) if pat[0][i + 1] > pat[0][i]][:prims]

    def energy(X):
        return O.vf_compute(n, X, pat, 10.0, 30, q, p, 1 | 8)["energy"]

    def fd(i, k, h):
        Xp = near.copy(); Xp[i, k] += h
        Xm = near.copy(); Xm[i, k] -= h
        return (energy(Xp) - energy(Xm)) / (2 * h)

    a = near[:, 0]
    for i in rows:
        hull = near[i, 1:1 + 3 * S]
        P = near[i, 1 + 3 * S:]
        dhull = D[i, 1:1 + 3 * S]
        dP = D[i, 1 + 3 * S:]
        expect = {0: D[i, 0] + np.dot(dhull, hull) + np.dot(dP, P)}
        for k in range(3 * S):
            expect[1 + k] = a[i] * dhull[k]
        for k in range(6):
            expect[1 + 3 * S + k] = a[i] * dP[k]
        scale = max(abs(v) for v in expect.values())
        for k, ev in expect.items():
            h = 1e-7 * max(1.0, abs(near[i, k])) if k > 0 else 1e-7 * a[i]
            g = fd(i, k, h)
            assert abs(g - ev) <= 1e-5 * scale, (i, k, g, ev)


def test_reference_energy_sum_omits_the_weight():
    """Curves at a gap where every evaluation is at level 1 (weight 1/4): the reference's unweighted
    energy is 4 times the weighted quadrature; triangles at level 1 (weight 1/16): 16 times."""
    for M, n, q, p, G in [(vf_circles(400, 1e-3), 1, 2.0, 4.0, 4), (vf_shells(3, 0.01), 2, Q6, P12, 16)]:
        near, pat = _vf_state(M)
        o_ref = O.vf_compute(n, near, pat, 10.0, 30, q, p, 1)
        o_w = O.vf_compute(n, near, pat, 10.0, 30, q, p, 1 | 8)
        assert o_ref["max_level"] == 1 and o_ref["evaluations"] == G * len(pat[1]) // 2
        assert abs(o_ref["energy"] - G * o_w["energy"]) <= 1e-12 * abs(o_ref["energy"])


@pytest.mark.parametrize("make,n,q,p", [
    (lambda: vf_circles(400, 3e-4), 1, 2.0, 4.0),
    (lambda: vf_shells(2, 0.006), 2, Q6, P12),
])
def test_vf_blocks_annihilate_constants_and_are_symmetric(make, n, q, p):
    M = make()
    near, pat = _vf_state(M)
    rp, ci, _ = pat
    o = O.vf_compute(n, near, pat, 10.0, 30, q, p, 4)
    s = (p - n) / q
    pl = O.pl_vf(n, near, pat, 10.0, 20, 2.0 - s)
    a = near[:, 0]
    T = len(a)
    for vals, diag in [(o["VF"][:, 0], o["VF_diag"][:, 0]), (o["VF"][:, 1], o["VF_diag"][:, 1]), (pl["VF"], pl["VF_diag"])]:
        # primitive-level product on the constant: y_i = diag_i a_i + sum_j v_ij a_j = 0
        y = diag * a
        for i in range(T):
            y[i] += np.dot(vals[rp[i]:rp[i + 1]], a[ci[rp[i]:rp[i + 1]]])
        scale = np.max(np.abs(diag * a))
        assert np.max(np.abs(y)) <= 1e-12 * scale
        for i in range(T):
            for k in range(rp[i], rp[i + 1]):
                j = ci[k]
                kk = rp[j] + np.searchsorted(ci[rp[j]:rp[j + 1]], i)
                assert vals[k] == vals[kk]
