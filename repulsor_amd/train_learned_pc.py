#!/usr/bin/env python3
"""Trains the learned preconditioner on seeded perturbed tori and spheres.

Training meshes alternate between a torus (R = 1, r ~ U[0.3, 0.5] on a 200 x 150 grid, normal jitter
sigma ~ U[0.01, 0.05]) and an icosphere of level 6 (81,920 triangles) with radial spherical-harmonic
noise (l <= lmax ~ U{4..16}, amplitude ~ U[0.01, 0.05]), all seeded. For each, the energy differential B
is computed, PCG with the current model solves A X = B to a tight tolerance while recording r_k and x_k,
and one Adam step is taken on the pairs (r_k, e_k = x_last - x_k) with the loss
min_alpha mean_k |alpha P r_k - e_k|_A^2 / |e_k|_A^2 (one minus the mean squared A-cosine of P r_k and
e_k after the best common scale; PCG does not see the scale of P), gradient by the library's HIP
kernels (rp_pc_learn_step). The trained model is written to repulsor_amd/models/learned_pc_trained.rplp
and the packaged one, its output layer shrunk toward the identity by learned_pc.PACKAGED_SHRINK, to
repulsor_amd/models/learned_pc.rplp.

Usage: python -m repulsor_amd.train_learned_pc [--meshes 192] [--lr 3e-3] [--start K] [--count N]
       [--checkpoint state.npz]
With --checkpoint, the parameters, Adam moments and step count are read from the file if it exists
and written back after every mesh, so that a long run can be split into several processes (meshes
start .. start+count-1 per process).
"""

from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import repulsor_amd as R  # noqa: E402
from repulsor_amd import learned_pc as LP  # noqa: E402


def training_mesh(k, nu, nv, seed0):
    """Mesh k: a perturbed torus for even k, a perturbed icosphere for odd k."""
    rng = np.random.default_rng(seed0 + k)
    if k % 2 == 0:
        sigma = float(rng.uniform(0.01, 0.05))
        r = float(rng.uniform(0.3, 0.5))
        X, S = R.make_torus(nu, nv, 1.0, r, sigma, seed0 + k)
        desc = f"torus r {r:.3f} sigma {sigma:.4f}"
    else:
        lmax = int(rng.integers(4, 17))
        amp = float(rng.uniform(0.01, 0.05))
        X, S = R.make_icosphere(6, lmax, amp, seed0 + k)
        desc = f"sphere lmax {lmax} amplitude {amp:.4f}"
    M = R.SimplicialMesh(X, S)
    M.cluster_tree_settings.split_threshold = 16
    M.cluster_tree_settings.device_build = True
    M.block_cluster_tree_settings.far_field_separation_parameter = 0.5
    M.adaptivity_settings.theta = 10.0
    M.build()
    return M, desc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--meshes", type=int, default=192)
    ap.add_argument("--grid", type=int, nargs=2, default=[200, 150], help="torus grid")
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--record-tol", type=float, default=1e-9)
    ap.add_argument("--out", default=LP.TRAINED_MODEL)
    ap.add_argument("--packaged", default=LP.DEFAULT_MODEL)
    ap.add_argument("--start", type=int, default=0)
    ap.add_argument("--count", type=int, default=-1)
    ap.add_argument("--checkpoint", default="")
    args = ap.parse_args()

    F = LP.feature_count()
    params = LP.init_params(F)
    m = np.zeros_like(params)
    v = np.zeros_like(params)
    if args.checkpoint and os.path.exists(args.checkpoint):
        st = np.load(args.checkpoint)
        params, m, v = st["params"].astype(np.float32), st["m"].astype(np.float32), st["v"].astype(np.float32)
        print(f"resumed from {args.checkpoint} after {int(st['step'])} steps", flush=True)
    q, p = 6.0, 12.0
    t0 = time.time()
    end = args.meshes if args.count < 0 else min(args.meshes, args.start + args.count)
    for k in range(args.start, end):
        M, desc = training_mesh(k, args.grid[0], args.grid[1], args.seed)
        E = R.TangentPointEnergy0(q, p)
        B = E.Differential(M)
        A = R.TangentPointMetric0(q, p)
        A.SetLearnedPreconditioner(params, bf16=False)
        A.SetLearnedPreconditionerGuard(0)          # record the model's own iterates, never a fallback's
        X = np.zeros_like(B)
        A.Solve(M, 1.0, B, 0.0, X, max_iter=300, tolerance=args.record_tol, record=True, raise_on_stagnation=False)
        it = A.Solver_IterationCount()
        loss, grad = A.PreconditionerLearnStep(M, params, m, v, k + 1, args.lr)
        del A, E, M
        print(f"mesh {k:3d} ({desc}): {it} PCG iterations to {args.record_tol:g}, loss {loss:.4e}, "
              f"|grad| {np.linalg.norm(grad):.3e}, w factor {np.exp(params[-1]):.3f}, {time.time() - t0:.1f} s", flush=True)
        if not np.all(np.isfinite(params)):
            raise RuntimeError("training diverged")
        if args.checkpoint:
            np.savez(args.checkpoint, params=params, m=m, v=v, step=k + 1)
    LP.save_model(args.out, params, F)
    LP.save_model(args.packaged, LP.shrink_output_layer(params, F, LP.PACKAGED_SHRINK), F)
    print(f"wrote {args.out} and {args.packaged}", flush=True)


if __name__ == "__main__":
    main()
