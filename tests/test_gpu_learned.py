"""Learned preconditioner P = D^{1/2} (P0 + w diag(A)^-1) D^{1/2} on the GPU.

* vertex scalings: the FP32 path equals a numpy evaluation of the same MLP, root-path sums, vertex
  averages and 16 smoothing steps (relative 1e-5); the BF16 inference path agrees to 2e-2;
* P applied on the GPU equals D^{1/2} (P0 + w J) D^{1/2} composed from the oracle's operators (1e-9)
  and is symmetric;
* the training gradient of rp_pc_learn_step matches central finite differences of its loss;
* the solve converges with a model; a deliberately poor model is stopped by the residual guard at
  iteration 12, long before max_iter, and the solve falls back to P0 + w diag(A)^-1.
"""

import ctypes as C

import numpy as np
import pytest

import repulsor_amd as R
from repulsor_amd import learned_pc as LP
from repulsor_amd.workloads import WORKLOADS, random_field
from helpers import OracleSolve, oracle_energy_diff, rel

pytestmark = pytest.mark.gpu


def _mesh(name):
    wl = WORKLOADS[name]
    M = wl.mesh()
    M.build()
    return wl, M


def _features(M):
    b = M.build()
    F = LP.feature_count()
    f = np.zeros((b["C"], F), dtype=np.float32)
    R._check(R.H.rpm_learned_features(M._h, f.ctypes.data_as(C.POINTER(C.c_float))))
    return f


def _model(F, seed=3, scale=0.3):
    p = LP.init_params(F, seed=seed)
    rng = np.random.default_rng(seed + 1)
    H = LP.HIDDEN
    p[H * F + H: H * F + 2 * H] = (scale * rng.standard_normal(H)).astype(np.float32)   # w2 != 0
    return p


def _numpy_sqrt_d(M, params, feats):
    lam = np.log(LP.node_scalings(params, feats))
    tr = M.tree()
    parent = np.full(len(tr["C_left"]), -1)
    for c in range(len(parent)):
        if tr["C_left"][c] >= 0:
            parent[tr["C_left"][c]] = c
            parent[tr["C_right"][c]] = c
    path = lam.copy()
    for c in range(len(parent)):            # parents precede children in depth-first numbering
        if parent[c] >= 0:
            path[c] += path[parent[c]]
    leaf_of_prim = np.zeros(M.T, dtype=np.int64)
    for c in range(len(parent)):
        if tr["C_left"][c] < 0:
            leaf_of_prim[tr["C_begin"][c]:tr["C_end"][c]] = c
    td = M.tree_data()
    cols = np.asarray(td["pre_cols"])
    acc = np.zeros(M.V); cnt = np.zeros(M.V)
    np.add.at(acc, cols.ravel(), np.repeat(path[leaf_of_prim], cols.shape[1]))
    np.add.at(cnt, cols.ravel(), 1.0)
    logd = acc / cnt
    for _ in range(16):                     # smoothing: vertex mean of primitive means
        pm = logd[cols].mean(axis=1)
        acc[:] = 0.0
        np.add.at(acc, cols.ravel(), np.repeat(pm, cols.shape[1]))
        logd = acc / cnt
    return np.exp(0.5 * logd)


def test_learned_scaling_matches_numpy():
    wl, M = _mesh("torus_2400")
    feats = _features(M)
    params = _model(feats.shape[1])
    ref = _numpy_sqrt_d(M, params, feats)
    A = R.TangentPointMetric0(wl.q, wl.p)
    A.SetLearnedPreconditioner(params, bf16=False)
    d32 = A.LearnedScaling(M)
    assert rel(d32, ref) <= 1e-5
    A.SetLearnedPreconditioner(params, bf16=True)
    d16 = A.LearnedScaling(M)
    print(f"scaling rel err fp32 {rel(d32, ref):.2e} bf16 {rel(d16, ref):.2e}", flush=True)
    assert rel(d16, ref) <= 2e-2


def test_learned_apply_matches_oracle_composition():
    wl, M = _mesh("torus_2400")
    feats = _features(M)
    params = _model(feats.shape[1])
    A = R.TangentPointMetric0(wl.q, wl.p)
    A.SetLearnedPreconditioner(params, bf16=False)
    sd = A.LearnedScaling(M)
    X = random_field(M.V, seed=21); Y = random_field(M.V, seed=22)
    PX = np.zeros_like(X); A.MultiplyPreconditioner(M, 1.0, X, 0.0, PX)
    PY = np.zeros_like(Y); A.MultiplyPreconditioner(M, 1.0, Y, 0.0, PY)
    # weight in use = exp(learned log w) x the estimate lambda_max(P0 A) / 4 of the two-level form
    w = A.PreconditionerWeightInUse(M)
    plain = R.TangentPointMetric0(wl.q, wl.p)
    plain.SetPreconditionerDiagonalWeight(-1.0)
    plain.MultiplyPreconditioner(M, 1.0, X, 0.0, np.zeros_like(X))
    w_est = plain.PreconditionerWeightInUse(M)
    assert w_est > 0 and abs(w - w_est * np.exp(float(params[-1]))) <= 1e-6 * w
    S = OracleSolve(M, wl.q, wl.p)
    S.with_diagonal_term(w, S.diagonal())
    ref = sd[:, None] * S.P(np.ascontiguousarray(sd[:, None] * X))
    assert rel(PX, ref) <= 1e-9
    assert abs(np.sum(X * PY) - np.sum(PX * Y)) <= 1e-10 * np.linalg.norm(PX) * np.linalg.norm(Y)


def test_learned_and_plain_objects_share_a_mesh():
    """Switching the device between a learned model and the plain forms on one mesh leaves no stale
    state: each solve matches the same solve on a fresh mesh."""
    wl = WORKLOADS["torus_2400"]
    F = LP.feature_count()
    params = _model(F, scale=0.3)

    def solve(M, kind):
        A = R.TangentPointMetric0(wl.q, wl.p)
        if kind == "learned":
            A.SetLearnedPreconditioner(params, bf16=False)
        elif kind == "auto":
            A.SetPreconditionerDiagonalWeight(-1.0)
        B = R.TangentPointEnergy0(wl.q, wl.p).Differential(M)
        X = np.zeros_like(B)
        A.Solve(M, 1.0, B, 0.0, X, max_iter=400, tolerance=1e-6, raise_on_stagnation=False)
        return A.Solver_IterationCount(), A.PreconditionerWeightInUse(M), X

    fresh = {}
    for kind in ["p0", "auto", "learned"]:
        M = wl.mesh(); M.build()
        fresh[kind] = solve(M, kind)
    M = wl.mesh(); M.build()
    for kind in ["learned", "p0", "auto", "learned", "auto"]:
        it, w, X = solve(M, kind)
        print(f"{kind}: {it} iterations, weight {w:.4g} (fresh {fresh[kind][0]}, {fresh[kind][1]:.4g})", flush=True)
        assert it == fresh[kind][0] and w == fresh[kind][1]
        assert rel(X, fresh[kind][2]) <= 1e-12


def test_learn_step_gradient_finite_differences():
    wl, M = _mesh("torus_2400")
    F = LP.feature_count()
    params = _model(F, scale=0.5)
    E = R.TangentPointEnergy0(wl.q, wl.p)
    B = E.Differential(M)
    A = R.TangentPointMetric0(wl.q, wl.p)
    A.SetLearnedPreconditioner(params, bf16=False)
    X = np.zeros_like(B)
    A.Solve(M, 1.0, B, 0.0, X, max_iter=300, tolerance=1e-10, record=True, raise_on_stagnation=False)
    z = np.zeros_like(params)
    loss0, grad = A.PreconditionerLearnStep(M, params.copy(), z.copy(), z.copy(), 1, 0.0)

    def loss_at(p):
        return A.PreconditionerLearnStep(M, p.astype(np.float32), z.copy(), z.copy(), 1, 0.0)[0]

    def fd(direction, h):
        """Richardson-extrapolated central difference of the loss along `direction`. The steps are
        small because the loss is strongly curved: a shift of every log s_c by t scales P by about
        exp(2 t depth)."""
        d1 = (loss_at(params + h * direction) - loss_at(params - h * direction)) / (2 * h)
        d2 = (loss_at(params + 0.5 * h * direction) - loss_at(params - 0.5 * h * direction)) / h
        return (4 * d2 - d1) / 3

    # directional derivative along the gradient
    u = grad / np.linalg.norm(grad)
    dd = fd(u, 0.01)
    print(f"loss {loss0:.6e} |grad| {np.linalg.norm(grad):.6e} directional fd {dd:.6e}", flush=True)
    assert abs(dd - np.linalg.norm(grad)) <= 2e-2 * np.linalg.norm(grad)
    # single components: first-layer weight, hidden bias, output weight, output bias, log w
    H = LP.HIDDEN
    for q in [0, 22, H * F + 7, H * F + H + 11, H * F + 2 * H, params.size - 1]:
        e = np.zeros(params.size); e[q] = 1.0
        g = fd(e, 0.02)
        print(f"param {q}: grad {grad[q]:.6e} fd {g:.6e}", flush=True)
        assert abs(grad[q] - g) <= 5e-2 * max(abs(g), 0.05 * np.linalg.norm(grad)), (q, grad[q], g)


def test_learned_solve_and_fallback():
    wl, M = _mesh("sphere_1280")
    F = LP.feature_count()
    E = R.TangentPointEnergy0(wl.q, wl.p)
    B = E.Differential(M)
    S = OracleSolve(M, wl.q, wl.p)
    _, Bo, _, _ = oracle_energy_diff(M, wl.q, wl.p)
    A = R.TangentPointMetric0(wl.q, wl.p)
    A.SetLearnedPreconditioner(_model(F, scale=0.2))
    X = np.zeros_like(B)
    A.Solve(M, 1.0, B, 0.0, X, max_iter=300, tolerance=1e-6)
    assert not A.Solver_FellBack()
    true = np.linalg.norm(Bo - S.A(X), axis=0) / np.linalg.norm(Bo, axis=0)
    assert np.all(true <= 1.5e-6), true
    # a model with wildly varying scalings is stopped by the residual guard at iteration 12 (far below
    # max_iter) and the solve falls back to the two-level form
    plain = R.TangentPointMetric0(wl.q, wl.p)
    plain.SetPreconditionerDiagonalWeight(-1.0)
    plain.Solve(M, 1.0, B, 0.0, np.zeros_like(B), max_iter=300, tolerance=1e-6)
    n_plain = plain.Solver_IterationCount()
    bad = _model(F, scale=40.0)
    A.SetLearnedPreconditioner(bad)
    X2 = np.zeros_like(B)
    A.Solve(M, 1.0, B, 0.0, X2, max_iter=300, tolerance=1e-6, raise_on_stagnation=False)
    print(f"bad model: fell back {A.Solver_FellBack()}, iterations {A.Solver_IterationCount()} "
          f"(two-level alone {n_plain})", flush=True)
    assert A.Solver_FellBack()
    assert A.Solver_IterationCount() == 12 + n_plain
    true2 = np.linalg.norm(Bo - S.A(X2), axis=0) / np.linalg.norm(Bo, axis=0)
    assert np.all(true2 <= 1.5e-6), true2
    # with the guard off, the same model falls back only at max_iter
    A.SetLearnedPreconditionerGuard(0)
    X3 = np.zeros_like(B)
    A.Solve(M, 1.0, B, 0.0, X3, max_iter=40, tolerance=1e-6, raise_on_stagnation=False)
    assert A.Solver_FellBack() and A.Solver_IterationCount() >= 40
