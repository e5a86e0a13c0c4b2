"""Pins of the oracle's MaximumSafeStepSize restatement, CPU only.

* the exact space-time hull test (LP phase one) agrees with scipy's linprog on random hull pairs;
* the step of one pair never exceeds the exact first-contact time of the two moving simplices
  (roots of the coplanarity cubics of vertex-face and edge-edge contacts, checked for contact) and
  is within the bisection tolerance of it;
* on two approaching sheets (triangles) and two approaching polygons (edges) the mesh step is safe
  and within the tolerance of the exact first contact; moving apart gives t_max exactly.
"""

import numpy as np
import pytest
from scipy.optimize import linprog

import oracle as O

EPS = 0.0625


def _hulls_meet_linprog(p, u, q, v, a, b):
    m = p.shape[0]
    P = [np.append(p[i] + t * u[i], t) for i in range(m) for t in (a, b)]
    Q = [np.append(q[i] + t * v[i], t) for i in range(m) for t in (a, b)]
    A = np.zeros((6, 4 * m)); rhs = np.array([1, 1, 0, 0, 0, 0.0])
    for k, x in enumerate(P):
        A[0, k] = 1; A[2:, k] = x
    for k, x in enumerate(Q):
        A[1, 2 * m + k] = 1; A[2:, 2 * m + k] = -x
    r = linprog(np.zeros(4 * m), A_eq=A, b_eq=rhs, bounds=[(0, None)] * (4 * m), method="highs")
    return r.status == 0


def test_hull_test_matches_linprog():
    rng = np.random.default_rng(1)
    agree = 0
    for trial in range(300):
        m = 3 if trial % 2 == 0 else 2
        p = rng.standard_normal((m, 3)); q = rng.standard_normal((m, 3)) + rng.uniform(0, 3) * rng.standard_normal(3)
        u = rng.standard_normal((m, 3)); v = rng.standard_normal((m, 3))
        a = rng.uniform(0, 0.5); b = a + rng.uniform(0.01, 1.0)
        assert O.moving_hulls_meet(p, u, q, v, a, b) == _hulls_meet_linprog(p, u, q, v, a, b)
        agree += 1
    assert agree == 300


def _cubic_roots(f, t_max):
    """Real roots in [0, t_max] of the cubic f (given as a callable), from 4 samples."""
    ts = np.linspace(0.0, t_max, 4)
    c = np.polyfit(ts, [f(t) for t in ts], 3)
    if np.all(np.abs(c) < 1e-300):
        return []
    r = np.roots(c)
    r = r[np.abs(r.imag) <= 1e-9 * max(1.0, t_max)].real
    return sorted(x for x in r if -1e-12 <= x <= t_max * (1 + 1e-12))


def _pair_first_contact(P, U, Q, V, t_max):
    """Exact first contact time of two linearly moving simplices (points rows), or inf."""
    best = np.inf

    def at(X, W, t):
        return X + t * W

    def vf(p, up, tri, utri):
        def f(t):
            A, B, Cc = at(tri, utri, t)
            return np.dot(np.cross(B - A, Cc - A), at(p, up, t) - A)
        for t in _cubic_roots(f, t_max):
            A, B, Cc = at(tri, utri, t)
            x = at(p, up, t)
            M = np.stack([B - A, Cc - A], 1)
            lam, *_ = np.linalg.lstsq(M, x - A, rcond=None)
            if lam.min() >= -1e-9 and lam.sum() <= 1 + 1e-9 and np.linalg.norm(M @ lam - (x - A)) <= 1e-9:
                return t
        return np.inf

    def ee(e, ue, g, ug):
        def f(t):
            a0, a1 = at(e, ue, t); b0, b1 = at(g, ug, t)
            return np.dot(np.cross(a1 - a0, b1 - b0), b0 - a0)
        for t in _cubic_roots(f, t_max):
            a0, a1 = at(e, ue, t); b0, b1 = at(g, ug, t)
            d1, d2, r = a1 - a0, b1 - b0, a0 - b0
            A = np.array([[d1 @ d1, -d1 @ d2], [-d1 @ d2, d2 @ d2]])
            if abs(np.linalg.det(A)) < 1e-18:
                continue
            s, w = np.linalg.solve(A, [-(d1 @ r), d2 @ r])
            if -1e-9 <= s <= 1 + 1e-9 and -1e-9 <= w <= 1 + 1e-9 and np.linalg.norm(a0 + s * d1 - b0 - w * d2) <= 1e-9:
                return t
        return np.inf

    m = P.shape[0]
    if m == 3:
        for i in range(3):
            best = min(best, vf(P[i], U[i], Q, V), vf(Q[i], V[i], P, U))
        for i in range(3):
            for j in range(3):
                best = min(best, ee(P[[i, (i + 1) % 3]], U[[i, (i + 1) % 3]], Q[[j, (j + 1) % 3]], V[[j, (j + 1) % 3]]))
    else:
        best = ee(P, U, Q, V)
    return best


def test_pair_step_is_safe_and_tight():
    rng = np.random.default_rng(2)
    tested = 0
    for trial in range(200):
        m = 3 if trial % 2 == 0 else 2
        P = rng.standard_normal((m, 3)) * 0.5
        Q = rng.standard_normal((m, 3)) * 0.5 + np.array([0, 0, 1.5])
        U = 0.2 * rng.standard_normal((m, 3)) + np.array([0, 0, 1.0])
        V = 0.2 * rng.standard_normal((m, 3)) - np.array([0, 0, 1.0])
        tc = _pair_first_contact(P, U, Q, V, 2.0)
        t = O.safe_step_pair(P, U, Q, V, 2.0, EPS)
        if not np.isfinite(tc):
            assert t <= 2.0
            continue
        tested += 1
        assert t <= tc * (1 + 1e-9), (trial, t, tc)
        assert t >= 0.8 * tc, (trial, t, tc)
    assert tested >= 50


def _sheet(nx, ny, z, jitter, rng):
    xs, ys = np.meshgrid(np.linspace(0, 1, nx), np.linspace(0, 1, ny), indexing="ij")
    X = np.stack([xs.ravel(), ys.ravel(), z + jitter * rng.standard_normal(nx * ny)], 1)
    S = []
    for i in range(nx - 1):
        for j in range(ny - 1):
            a, b, c, d = i * ny + j, (i + 1) * ny + j, (i + 1) * ny + j + 1, i * ny + j + 1
            S += [[a, b, c], [a, c, d]]
    return X, np.array(S, np.int32)


def _mesh_first_contact(X, U, S, t_max):
    m = S.shape[1]
    best = np.inf
    for i in range(len(S)):
        for j in range(i + 1, len(S)):
            if set(S[i]) & set(S[j]):
                continue
            best = min(best, _pair_first_contact(X[S[i]], U[S[i]], X[S[j]], U[S[j]], t_max))
    return best


@pytest.mark.parametrize("kind", ["sheets", "polygons"])
def test_mesh_step_against_exact_first_contact(kind):
    rng = np.random.default_rng(3)
    if kind == "sheets":
        X1, S1 = _sheet(5, 5, 0.0, 0.01, rng)
        X2, S2 = _sheet(5, 5, 0.3, 0.01, rng)
    else:
        t = np.linspace(0, 2 * np.pi, 24, endpoint=False)
        X1 = np.stack([np.cos(t), np.sin(t), 0.01 * rng.standard_normal(24)], 1)
        X2 = np.stack([np.cos(t + 0.1), np.sin(t + 0.1), 0.3 + 0.01 * rng.standard_normal(24)], 1)
        E = np.stack([np.arange(24), (np.arange(24) + 1) % 24], 1)
        S1 = S2 = E.astype(np.int32)
    X = np.vstack([X1, X2]); S = np.vstack([S1, S2 + len(X1)]).astype(np.int32)
    U = np.zeros_like(X)
    U[: len(X1), 2] = 1.0 + 0.3 * rng.random(len(X1))
    U[len(X1):, 2] = -1.0 - 0.3 * rng.random(len(X2))
    t_max = 1.0
    tc = _mesh_first_contact(X, U, S, t_max)
    t, cand, tests = O.max_safe_step(X, U, S, t_max, EPS)
    print(f"{kind}: exact first contact {tc:.6f}, safe step {t:.6f}, {cand} candidate pairs, {tests} hull tests")
    assert np.isfinite(tc)
    assert t <= tc * (1 + 1e-9)
    assert t >= (1 - 2 * EPS) * tc
    # moving apart: nothing collides within t_max
    t2, _, _ = O.max_safe_step(X, -U, S, t_max, EPS)
    assert t2 == t_max
