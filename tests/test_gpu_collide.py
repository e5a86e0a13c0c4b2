"""MaximumSafeStepSize on the GPU (rp_max_safe_step) against the oracle.

* the simplex pairs the cluster-tree traversal hands to the bisection are exactly the non-adjacent
  pairs whose boxes over [0, t_max] meet (the oracle's brute-force count);
* the step equals the oracle's (same bisection per pair, GJK in R^4 here, an LP test there) and never
  exceeds the exact first contact of the approaching sheets and polygons;
* moving apart gives t_max; a random field on a sphere and on a torus matches the oracle.
"""

import numpy as np
import pytest

import oracle as O
import repulsor_amd as R
from repulsor_amd.workloads import WORKLOADS
from test_oracle_collision import _mesh_first_contact, _sheet

pytestmark = pytest.mark.gpu

EPS = 0.0625


def _approaching(kind, seed=3):
    rng = np.random.default_rng(seed)
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
    return X, S, U


@pytest.mark.parametrize("kind", ["sheets", "polygons"])
def test_safe_step_against_oracle_and_exact_contact(kind):
    X, S, U = _approaching(kind)
    M = R.SimplicialMesh(X, S)
    M.cluster_tree_settings.split_threshold = 4
    t, st = M.MaximumSafeStepSize(U, 1.0, EPS, return_stats=True)
    to, cand, tests = O.max_safe_step(X, U, S, 1.0, EPS)
    tc = _mesh_first_contact(X, U, S, 1.0)
    print(f"{kind}: gpu {t:.8f} oracle {to:.8f} exact contact {tc:.8f}; pairs {st['pairs']} (oracle {cand}), "
          f"hull tests {st['hull_tests']} (oracle {tests}), leaf pairs {st['leaf_pairs']}", flush=True)
    assert st["pairs"] == cand
    assert t == to
    assert t <= tc * (1 + 1e-9) and t >= (1 - 2 * EPS) * tc
    assert M.MaximumSafeStepSize(-U, 1.0, EPS) == 1.0


@pytest.mark.parametrize("name,scale,t_max", [("sphere_1280", 0.3, 1.0), ("torus_2400", 0.05, 2.0), ("trefoil_3000", 0.05, 1.0)])
def test_safe_step_random_field(name, scale, t_max):
    wl = WORKLOADS[name]
    M = wl.mesh()
    b = M.build()
    X = M.VertexCoordinates().copy()
    rng = np.random.default_rng(7)
    U = scale * rng.standard_normal(X.shape)
    t, st = M.MaximumSafeStepSize(U, t_max, EPS, return_stats=True)
    to, cand, tests = O.max_safe_step(X, U, M.simplices, t_max, EPS)
    print(f"{name}: gpu {t:.8f} oracle {to:.8f}; pairs {st['pairs']} (oracle {cand}), hull tests {st['hull_tests']} "
          f"(oracle {tests}), leaf pairs {st['leaf_pairs']}, T {b['T']}", flush=True)
    assert st["pairs"] == cand
    assert t == to
    assert 0.0 < t <= t_max


def _submesh_near_moving(X, U, S, t_max):
    """Simplices whose swept boxes over [0, t_max] meet the swept box of a moving simplex."""
    P0 = X[S]; P1 = (X + t_max * U)[S]
    lo = np.minimum(P0, P1).min(1); hi = np.maximum(P0, P1).max(1)
    moving = np.flatnonzero(np.any(U[S] != 0.0, axis=(1, 2)))
    keep = np.zeros(len(S), bool)
    for k in moving:
        keep |= np.all((lo <= hi[k]) & (hi >= lo[k]), axis=1)
    return np.flatnonzero(keep)


@pytest.mark.parametrize("name,t_max", [("sphere_81920", 1.0), ("trefoil_200000", 1.0), ("sphere_1310720", 0.5)])
def test_safe_step_localized_field(name, t_max):
    """A field that moves only a small patch: a simplex pair whose swept boxes do not meet a moving simplex is
    static and embedded, so its step is t_max, and the whole mesh's step is the oracle's step over the simplices
    near the patch (run with the full vertex array, so adjacency is unchanged)."""
    wl = WORKLOADS[name]
    M = wl.mesh()
    M.build()
    X = M.VertexCoordinates().copy()
    S = M.simplices
    rng = np.random.default_rng(11)
    c = X[rng.integers(len(X))]
    h = np.median(np.linalg.norm(X[S[:, 1]] - X[S[:, 0]], axis=1))
    near = np.linalg.norm(X - c, axis=1) < 6 * h
    U = np.zeros_like(X)
    U[near] = 3 * h * rng.standard_normal((near.sum(), 3))
    t, st = M.MaximumSafeStepSize(U, t_max, EPS, return_stats=True)
    sub = _submesh_near_moving(X, U, S, t_max)
    to, cand, _ = O.max_safe_step(X, U, S[sub], t_max, EPS)
    print(f"{name}: {near.sum()} moving vertices, {len(sub)} simplices near them; gpu {t:.8f} oracle {to:.8f}; "
          f"gpu pairs {st['pairs']} (near-patch pairs {cand}), leaf pairs {st['leaf_pairs']}", flush=True)
    assert to < t_max            # the patch folds onto itself within t_max
    assert t == to
    assert st["pairs"] >= cand
