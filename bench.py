#!/usr/bin/env python3
"""Benchmark of the tangent-point hot path on one MI355X.

Metric: "energy+grad ms & metric-solve ms to 1e-6 residual, learned vs plain PC, 1 GPU".
Workload (default): subdivided icosahedron level 8 (655,362 vertices, 1,310,720 triangles),
radial spherical-harmonic noise l<=16 amplitude 0.03 (seed 5), q=6, p=12, far theta 0.5,
near theta 10, leaf size 16.

One step = TangentPointEnergy0 energy + differential (near and far kernels, derivative
collection, chain rule, vertex assembly) with all inputs resident in HBM and the output in a
device buffer, followed by the metric solve A X = dE (TangentPointMetric0::Solve, three
columns, PCG with the learned preconditioner D^1/2 (P0 + w diag(A)^-1) D^1/2 around the reference
P0 = H^-1 L_{2-s} H^-1, stopped at |dE - A X|_2 / |dE|_2 <= 1e-6 per column) from and into device
buffers. The same right-hand side is also solved with the two-level form, in FP64, and with the
reference P0 alone, and reported beside the headline. Tree/BCT construction, uploads, the per-state metric and
pseudo-Laplacian assembly (device) and the H^1 factorisation (host) are timed separately and
reported, never folded into the step.

Prints one JSON line: value = ms per step (lower is better).
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

# Flop counts per upper-triangle pair, counted from the reference kernels with
# add/sub/mul = 1, div = 1, integer powers as their multiplies (TP0_Kernel_NF.hpp:140-338 with
# loadT's barycenter and writeT's 16-wide accumulation; TP0_Kernel_FF.hpp:130-319 with writeT's
# 10-wide accumulation), for q=6, p=12 (integer exponents).
NF_FLOPS_VALUE_DIFF = 263
FF_FLOPS_VALUE_DIFF = 194
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector (equal to the FP64 matrix peak)
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

# HBM bytes per k_nf_spmm launch measured with rocprofv3 FETCH_SIZE + WRITE_SIZE, each scaled by
# the counter's reading for a copy kernel of known size (profiles/k_nf_spmm_sphere_1310720_traffic.md).
MEASURED_SPMM_TRAFFIC = {("sphere_1310720", "fp64"): 11.137e9, ("sphere_1310720", "fp32-copies"): 5.918e9}


def spmm_algorithmic_bytes(b, precision):
    """Compulsory bytes of one near-field SpMM launch: the metric values of every (row, column) pair
    of the full near pattern (2 per pair: 16 B in FP64, 8 B in FP32), the tile column list (column
    index, 4 B per tile column entry; masked slots hold 0, so the row masks are not read), and per
    primitive the input block (12 doubles), the diagonal (2 doubles), the post-operator weights
    (4 x 3 doubles) and the output (3 x 3 doubles), plus the far-field total of every leaf (12 doubles)."""
    val_bytes = 8 if precision == "mixed" else 16
    return (val_bytes * b["near_nnz"] + 4 * b["tile_cols"]
            + b["T"] * (96 + 16 + 96 + 72) + 96 * b["L"])


def _hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipDeviceSynchronize.argtypes = []
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return lib


def host_cpu_info():
    """Cores this process may use (its CPU share on the GPU box) and the CPU model."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, model


def cpu_baseline(M, wl, threads, solve_iters=0):
    """The oracle restatement (-O3 -ffast-math build) timed on full passes of the same workload, on
    `threads` host threads (rows split by cost, per-thread accumulators, as the reference's
    FMM_Traversor / AddReduce) and on one thread: energy + differential (near and far pair kernels,
    derivative collection, chain rule, assembly), one MultiplyMetric and one preconditioner application
    (pseudo-Laplacian product between two host H^1 solves), all complete, nothing sampled. The CPU step is
    energy+grad + iterations x (MultiplyMetric + preconditioner), with the iteration count of the GPU's
    solve (same method). Metric and pseudo-Laplacian assembly (per state, like the GPU's) are timed on
    `threads` threads and reported apart. The derivative collection/chain rule and the H^1 solves run
    on one thread in every column."""
    import oracle as O
    t_start = time.perf_counter()
    b = M.build()
    td = M.tree_data()
    tr = M.tree()
    n = M.dom_dim
    npat = M.pattern("near")
    fpat = M.pattern("far")
    P_near, C_far = O.c64(td["P_near"]), O.c64(td["C_far"])
    s_pl = 2.0 - (wl.p - n) / wl.q
    Xv = np.random.default_rng(11).standard_normal((M.V, 3))
    Yv = np.zeros_like(Xv)

    def timed(fn):
        t0 = time.perf_counter(); r = fn(); return time.perf_counter() - t0, r

    def eg(th):
        o = O.tp0_compute(n, P_near, C_far, npat, fpat, wl.q, wl.p, 3, impl=O.fast_mt("orc_tp0_compute", th))
        O.differential(n, M.VertexCoordinates(), M.simplices, tr, o["P_D_near_tree"], o["C_D_far"])
        return o["energy"]

    t_asm, mv = timed(lambda: O.tp0_compute(n, P_near, C_far, npat, fpat, wl.q, wl.p, 4, impl=O.fast_mt("orc_tp0_compute", threads)))
    t_plasm, pl = timed(lambda: O.pl_compute(n, P_near, C_far, npat, fpat, s_pl, impl=O.fast_mt("orc_pl_compute", threads)))
    M.h1_solve_host(Xv)                          # factorisation (per state, not timed)

    def mm(th):
        return O.multiply_metric(n, M.V, tr, td["pre_cols"], td["pre_vals"], npat, mv["NF"], mv["NF_diag"], fpat, mv["FF"],
                                 mv["FF_diag"], 1.0, Xv, 0.0, Yv, impl=O.fast_mt("orc_multiply_metric", th))

    def pc(th):
        Z, _ = M.h1_solve_host(Xv)
        W = np.zeros_like(Z)
        O.pl_multiply(n, M.V, tr, td["pre_cols"], td["pre_vals"], npat, pl, fpat, 1.0, Z, 0.0, W,
                      impl=O.fast_mt("orc_pl_multiply", th))
        return M.h1_solve_host(W)[0]

    cols = {}
    for th in sorted({threads, 1}, reverse=True):
        t_eg, _ = timed(lambda: eg(th))
        t_mm = min(timed(lambda: mm(th))[0] for _ in range(2))
        t_pc = min(timed(lambda: pc(th))[0] for _ in range(2))
        step = t_eg + solve_iters * (t_mm + t_pc)
        cols[th] = {"cores": th, "energy_grad_ms": round(t_eg * 1e3, 1), "multiply_metric_ms": round(t_mm * 1e3, 1),
                    "preconditioner_ms": round(t_pc * 1e3, 1), "step_ms": round(step * 1e3, 1),
                    "pairs_per_s": round((b["near_nnz"] + b["far_nnz"]) / 2 / t_eg, 1)}
    allc = cols[threads]
    out = {
        "value": allc["step_ms"], "unit": "ms (energy+grad + iterations x (MultiplyMetric + preconditioner), full passes)",
        "cores": threads, "kind": "port",
        "sample": (f"full: all {b['near_nnz'] // 2} near and {b['far_nnz'] // 2} far upper pairs, complete MultiplyMetric "
                   f"and preconditioner applications on the same mesh; solve = {solve_iters} iterations (the GPU solve's "
                   f"count) x (MultiplyMetric + preconditioner); oracle -O3 -ffast-math build, std::threads"),
        "energy_grad_ms": allc["energy_grad_ms"], "multiply_metric_ms": allc["multiply_metric_ms"],
        "preconditioner_ms": allc["preconditioner_ms"], "iterations": solve_iters,
        "one_core": cols[1],
        "metric_assembly_ms": round(t_asm * 1e3, 1), "pl_assembly_ms": round(t_plasm * 1e3, 1),
    }
    out["cpu_s"] = round(time.perf_counter() - t_start, 1)
    return out


def tree_pass_bytes(b, n):
    """Compulsory bytes of the parts of one MultiplyMetric outside the near-field SpMM (pre-operator,
    leaf and up-sweep sums, far-field SpMM, leaf path sums, vertex gather of the post-operator):
    pre_cols (4 B) and pre_vals (4 x 8 B) per simplex vertex, X (24 B per vertex), P_in written (96 B per
    primitive) and read by the leaf sums (96 B), 96 B per leaf written and 288 B per interior cluster for
    the up-sweep, far-field values (16 B) and columns (4 B) per far nnz with 8 + 96 + 16 + 96 B per cluster
    (row pointer, input, diagonal, output), 96 B read per cluster and 96 B written per leaf for the path
    sums, and the post-operator's per-slot inputs (24 B) and entry indices (4 B) per simplex vertex with
    8 B per vertex pointer and 24 B per output vertex."""
    T, V, C, L = b["T"], b["V"], b["C"], b["L"]
    s = n + 1
    return (T * s * (4 + 32) + 24 * V + 96 * T + 96 * T + 96 * L + 288 * (C - L)
            + 20 * b["far_nnz"] + C * (8 + 96 + 16 + 96) + 96 * C + 96 * L + T * s * (24 + 4) + V * (8 + 24))


def preconditioner_comparison(R, WORKLOADS, name, tol, max_iter, model, threads):
    """Metric solve A X = dE to |dE - A X|_2 / |dE|_2 <= tol (PCG, FP64 metric) on one workload with the
    reference preconditioner P0, the two-level P0 + w diag(A)^-1 and the learned model; each solve's
    residual is recomputed with the FP64 product. Times are device events of the solve (setup excluded)."""
    wl = WORKLOADS[name]
    M = wl.mesh(threads=threads)
    M.cluster_tree_settings.device_build = True
    b = M.build()
    G = R.TangentPointEnergy0(wl.q, wl.p).Differential(M)
    A64 = R.TangentPointMetric0(wl.q, wl.p)
    out = {"workload": name, "triangles": b["T"], "method": "pcg", "tolerance": tol}
    for label in ("reference_P0", "two_level", "learned"):
        S = R.TangentPointMetric0(wl.q, wl.p)
        if label == "two_level":
            S.SetPreconditionerDiagonalWeight(-1.0)
        if label == "learned":
            if not model or not os.path.exists(model):
                continue
            S.SetLearnedPreconditioner(path=model)
        Xs = np.zeros_like(G)
        S.Solve(M, 1.0, G, 0.0, Xs, max_iter=max_iter, tolerance=tol, raise_on_stagnation=False)   # setup and warm-up
        Xs[:] = 0.0
        S.Solve(M, 1.0, G, 0.0, Xs, max_iter=max_iter, tolerance=tol, raise_on_stagnation=False)
        ms = M.last_timing(3)
        AX = np.zeros_like(Xs)
        A64.MultiplyMetric(M, 1.0, Xs, 0.0, AX)
        res = np.linalg.norm(G - AX, axis=0) / np.linalg.norm(G, axis=0)
        out[label] = {"iterations": S.Solver_IterationCount(), "solve_ms": round(ms, 2),
                      "fp64_residual_max": float(res.max()), "converged": bool(np.all(res <= 1.001 * tol))}
        if label == "learned":
            out[label]["fell_back"] = S.Solver_FellBack()
            out[label]["model"] = os.path.relpath(model, ROOT)
        print(f"# {name} {label}: {out[label]}", file=sys.stderr, flush=True)
    if "learned" in out:
        out["learned_vs_P0_iterations"] = round(out["reference_P0"]["iterations"] / max(1, out["learned"]["iterations"]), 2)
        out["learned_vs_two_level_iterations"] = round(out["two_level"]["iterations"] / max(1, out["learned"]["iterations"]), 3)
    return out


def calibrate_copy(mb):
    lib = C.CDLL(os.path.join(ROOT, "repulsor_amd", "lib", "librepulsor_calib.so"))
    lib.rpc_copy.argtypes = [C.c_int64, C.c_int, C.POINTER(C.c_double)]
    ms = C.c_double(0)
    st = lib.rpc_copy(int(mb) * (1 << 20), 5, C.byref(ms))
    print(json.dumps({"calibration_copy_bytes": int(mb) * (1 << 20), "ms": ms.value, "status": st,
                      "GBs": 2 * int(mb) * (1 << 20) / (ms.value * 1e-3) / 1e9}), flush=True)
    return st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="sphere_1310720")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"],
                    help="energy + differential: FP64, or FP32 pair arithmetic with FP64 accumulation")
    ap.add_argument("--metric-precision", default="mixed-refined", choices=["fp64", "mixed", "mixed-refined"],
                    help="metric blocks of the solve (FP64 keeps the final residual the FP64 one; mixed-refined: FP64 "
                         "values plus FP32 copies for the CG iterations, convergence judged by the FP64 residual)")
    ap.add_argument("--learned", default="packaged",
                    help="learned preconditioner model of the timed solves: 'packaged' (repulsor_amd/models/"
                         "learned_pc.rplp, the default), a model path, or 'none' for the two-level form")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the comparison solves (two-level FP64, reference P0 FP64) and the FP64 energy timing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="host threads of the CPU baseline (0: this process's CPU share, at most 16)")
    ap.add_argument("--calibrate-copy", type=int, default=0, help="run only the copy calibration kernel (MiB)")
    ap.add_argument("--solver", default="pcg", choices=["pcg", "gmres"])
    ap.add_argument("--max-iter", type=int, default=3000)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-solve", action="store_true", help="time energy+grad only")
    ap.add_argument("--host-tree", action="store_true", help="build the cluster tree and block cluster tree on the host")
    ap.add_argument("--pc-precision", default="bf16", choices=["fp64", "mixed", "bf16"],
                    help="pseudo-Laplacian values inside the preconditioner: FP64, FP32, or FP32 arithmetic stored in "
                         "BF16 (the metric keeps --metric-precision; the solve is judged by the FP64 residual)")
    ap.add_argument("--pc-weight", type=float, default=-1.0,
                    help="diagonal weight w of P = P0 + w diag(A)^-1 (0: reference P0, <0: estimated per state)")
    ap.add_argument("--plain-pc-max-iter", type=int, default=4000,
                    help="iteration cap of the one reported solve with the reference P0 (0 skips it)")
    ap.add_argument("--pc-compare-workload", default="torus_153600",
                    help="workload of the preconditioner comparison (reference P0, two-level, learned); '' skips it")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not an implemented path).", flush=True)
        sys.exit(2)
    if args.calibrate_copy:
        sys.exit(calibrate_copy(args.calibrate_copy))

    import repulsor_amd as R
    from repulsor_amd.workloads import WORKLOADS

    wl = WORKLOADS[args.workload]
    threads = min(16, os.cpu_count() or 1)
    M = wl.mesh(threads=threads)
    M.cluster_tree_settings.device_build = not args.host_tree
    t0 = time.perf_counter()
    b = M.build()
    t_build = time.perf_counter() - t0
    build_kernel_ms = M.build_device_ms() if not args.host_tree else None
    print(f"# workload {wl.name}: T={b['T']} V={b['V']} C={b['C']} leaves={b['L']} near nnz={b['near_nnz']} "
          f"far nnz={b['far_nnz']} very near={b['verynear_nnz']} tree+BCT build {t_build:.2f}s "
          f"({'host' if args.host_tree else f'device, kernels {build_kernel_ms:.1f} ms'})", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    M.upload()
    t_upload = time.perf_counter() - t0

    from repulsor_amd import learned_pc as LP
    learned = {"packaged": LP.DEFAULT_MODEL, "none": "", "": ""}.get(args.learned, args.learned)
    E = R.TangentPointEnergy0(wl.q, wl.p, precision=args.precision)
    A = R.TangentPointMetric0(wl.q, wl.p, precision=args.metric_precision)
    A.SetPreconditionerDiagonalWeight(args.pc_weight)
    A.SetPreconditionerPrecision(args.pc_precision)
    if learned:
        A.SetLearnedPreconditioner(path=learned, bf16=True)

    hip = _hip()
    Yd = C.c_void_p()
    assert hip.hipMalloc(C.byref(Yd), M.V * 3 * 8) == 0
    Xd = C.c_void_p()
    assert hip.hipMalloc(C.byref(Xd), M.V * 3 * 8) == 0
    en = np.zeros(1)
    from repulsor_amd import H, _p, _check, RP_ESTAGNATION  # noqa: E402

    # per-state preparation, timed separately: metric blocks and pseudo-Laplacian on the device,
    # H^1 factorisation on the host (+ upload)
    t0 = time.perf_counter()
    A.assemble(M)
    t_assemble = time.perf_counter() - t0
    metric_assembly_dev_ms = M.last_timing(0)
    t0 = time.perf_counter()
    Z0 = np.zeros((M.V, 3))
    A.MultiplyPreconditioner(M, 1.0, np.ones((M.V, 3)), 0.0, Z0)
    t_pc_prep = time.perf_counter() - t0
    pl_assembly_dev_ms = M.last_timing(9)
    pc_setup_dev_ms = M.last_timing(10)
    pc_weight = A.PreconditionerWeightInUse(M)
    print(f"# metric assembly {metric_assembly_dev_ms:.2f} ms (device), pseudo-Laplacian {pl_assembly_dev_ms:.2f} ms (device), "
          f"preconditioner setup {pc_setup_dev_ms:.2f} ms (device, weight {pc_weight:.4g}), "
          f"H1 assembly+factor+upload+setup {t_pc_prep:.2f} s (wall)", file=sys.stderr, flush=True)

    it_c = (C.c_int * 2)()     # iteration count, 1 if the learned preconditioner fell back
    rel = np.zeros(3)
    solve_flags = 0 if args.solver == "pcg" else R.RP_SOLVE_GMRES
    dY = C.cast(Yd, C.POINTER(C.c_double))
    dX = C.cast(Xd, C.POINTER(C.c_double))

    fell_back = []
    converged = []

    def step():
        _check(H.rpm_energy_differential(E._h, M._h, 1.0, 0.0, dY, 3, _p(en), 1))
        t_eg = M.last_timing(0) + M.last_timing(1)
        tn, tf = M.last_timing(4), M.last_timing(5)
        if args.no_solve:
            return t_eg, tn, tf, 0.0
        st = H.rpm_metric_solve(A._h, M._h, 1.0, dY, 3, 0.0, dX, 3, 3, args.max_iter, args.tol, solve_flags, 1, it_c, _p(rel))
        if st not in (0, RP_ESTAGNATION):
            _check(st)
        fell_back.append(int(it_c[1]))
        converged.append(st == 0)
        return t_eg, tn, tf, M.last_timing(3)

    for _ in range(args.warmup):
        step()
    converged.clear()
    nf_ms, ff_ms, eg_ms, sv_ms = [], [], [], []
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t_eg, tn, tf, ts = step()
        nf_ms.append(tn); ff_ms.append(tf); eg_ms.append(t_eg); sv_ms.append(ts)
        print(f"# step: energy+grad {t_eg:.3f} ms, solve {ts:.1f} ms, {it_c[0]} iterations, rel {rel.max():.2e}",
              file=sys.stderr, flush=True)
    hip.hipDeviceSynchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    solve_iters = int(it_c[0])
    solve_rel = [float(x) for x in rel]     # the last timed solve's (the comparison solves reuse rel)
    stagnated_steps = sum(1 for c in converged if not c)
    if stagnated_steps:
        print(f"# WARNING: {stagnated_steps} of {len(converged)} timed solves stopped at max_iter without meeting the tolerance",
              file=sys.stderr, flush=True)
    print(f"# energy {en[0]:.15e}  step {ms_per_step:.3f} ms (device energy+grad {np.mean(eg_ms):.3f} ms: near {np.mean(nf_ms):.3f}, "
          f"far {np.mean(ff_ms):.3f}; solve {np.mean(sv_ms):.1f} ms in {solve_iters} iterations)", file=sys.stderr, flush=True)
    # operator timings (device events; host staging excluded)
    X = np.random.default_rng(11).standard_normal((M.V, 3))
    Y = np.zeros_like(X)
    A.MultiplyMetric(M, 1.0, X, 0.0, Y)
    mm_ms, spmm_ms, pcl_ms, h1l_ms, spmm32_ms = [], [], [], [], []
    # host-pointer (cold) calls: the step's two operations with host memory on both sides (PCIe included)
    Gh = np.zeros((M.V, 3))
    cold_eg, cold_sv = [], []
    if not args.no_solve:
        for _ in range(2):
            t0 = time.perf_counter()
            _check(H.rpm_energy_differential(E._h, M._h, 1.0, 0.0, _p(Gh), 3, _p(en), 0))
            cold_eg.append((time.perf_counter() - t0) * 1e3)
            Xc = np.zeros_like(Gh)
            t0 = time.perf_counter()
            stc = H.rpm_metric_solve(A._h, M._h, 1.0, _p(Gh), 3, 0.0, _p(Xc), 3, 3, args.max_iter, args.tol, solve_flags, 0, it_c, _p(rel))
            cold_sv.append((time.perf_counter() - t0) * 1e3)
            if stc not in (0, RP_ESTAGNATION):
                _check(stc)
    refined = args.metric_precision == "mixed-refined"
    for _ in range(max(3, args.steps)):
        A.MultiplyMetric(M, 1.0, X, 0.0, Y)
        mm_ms.append(M.last_timing(2)); spmm_ms.append(M.last_timing(6))
        if refined:
            A.MultiplyMetricInner(M, 1.0, X, 0.0, Y)
            spmm32_ms.append(M.last_timing(6))
        A.MultiplyPreconditioner(M, 1.0, X, 0.0, Y)
        pcl_ms.append(M.last_timing(7))
        M.h1_solve(X)
        h1l_ms.append(M.last_timing(8))
    pc_ms, h1_ms = float(np.mean(pcl_ms)), float(np.mean(h1l_ms))

    # the line-search bound that follows the solve in a gradient flow: MaximumSafeStepSize along the
    # descent direction -X (staged in HBM, device-pointer entry), max_time such that the fastest vertex
    # moves one mean edge length
    safe = None
    fp64_check = None
    if not args.no_solve:
        Xh = np.zeros((M.V, 3))
        assert hip.hipMemcpy(Xh.ctypes.data, Xd, M.V * 3 * 8, 2) == 0
        # the last solve's residual recomputed with the FP64 metric product (independent of the solver's
        # own bookkeeping; with --metric-precision mixed the product is the FP32-valued one)
        Bh = np.zeros((M.V, 3))
        assert hip.hipMemcpy(Bh.ctypes.data, Yd, M.V * 3 * 8, 2) == 0
        AX = np.zeros_like(Xh)
        A.MultiplyMetric(M, 1.0, Xh, 0.0, AX)
        fp64_check = [float(x) for x in np.linalg.norm(Bh - AX, axis=0) / np.linalg.norm(Bh, axis=0)]
        Dh = np.ascontiguousarray(-Xh)
        Dd = C.c_void_p()
        assert hip.hipMalloc(C.byref(Dd), M.V * 3 * 8) == 0
        assert hip.hipMemcpy(Dd, Dh.ctypes.data, M.V * 3 * 8, 1) == 0
        Vx = M.VertexCoordinates()
        Sx = M.simplices
        h_mean = float(np.mean(np.linalg.norm(Vx[Sx[:, 1]] - Vx[Sx[:, 0]], axis=1)))
        t_cap = h_mean / float(np.max(np.linalg.norm(Xh, axis=1)))
        ss_ms = []
        for _ in range(3):
            t0 = time.perf_counter()
            t_safe, ss = M.MaximumSafeStepSize(None, t_cap, 0.0625, return_stats=True, device_ptr=Dd.value)
            ss_ms.append((time.perf_counter() - t0) * 1e3)
        hip.hipFree(Dd)
        safe = {"direction": "-(metric solve)", "max_time": t_cap, "max_displacement": "one mean edge length",
                "t": t_safe, "ms": round(float(np.median(ss_ms)), 3), "pair_stage_ms": round(M.last_timing(12), 3),
                "leaf_pairs": ss["leaf_pairs"], "simplex_pairs": ss["pairs"], "hull_tests": ss["hull_tests"]}
        print(f"# MaximumSafeStepSize: {safe}", file=sys.stderr, flush=True)

    # comparison forms on the same right-hand side:
    #  - the two-level preconditioner at the headline's precisions (what the learned scalings buy),
    #  - FP64 energy + differential and FP64 metric with the headline's preconditioner (what the FP32 /
    #    mixed-refined arithmetic buys, preconditioner held fixed),
    #  - the two-level and the learned form in FP64,
    #  - the plain FP64 form (reference P0 alone; the P0 solve is capped to about a minute of iterations)
    plain = two_level = learned_cmp = two_level_same = eg64 = None
    headline_learned = bool(learned) and not any(fell_back)

    def compare_solve(precision, model, pc_precision=None):
        Ac = R.TangentPointMetric0(wl.q, wl.p, precision=precision)
        Ac.SetPreconditionerDiagonalWeight(args.pc_weight if args.pc_weight != 0 else -1.0)
        if pc_precision:
            Ac.SetPreconditionerPrecision(pc_precision)
        if model:
            Ac.SetLearnedPreconditioner(path=model, bf16=True)
        st = H.rpm_metric_solve(Ac._h, M._h, 1.0, dY, 3, 0.0, dX, 3, 3, args.max_iter, args.tol, solve_flags, 1, it_c, _p(rel))
        out = {"metric_precision": precision, "solve_ms": round(M.last_timing(3), 1), "iterations": int(it_c[0]),
               "converged": st == 0, "rel_residuals": [float(x) for x in rel]}
        if model:
            out.update({"model": os.path.relpath(model, ROOT), "inference": "bf16", "fell_back": bool(it_c[1])})
        return out

    if not args.no_solve and not args.no_compare:
        E64 = R.TangentPointEnergy0(wl.q, wl.p, precision="fp64")
        t64 = []
        for _ in range(3):
            _check(H.rpm_energy_differential(E64._h, M._h, 1.0, 0.0, dY, 3, _p(en), 1))
            t64.append(M.last_timing(0) + M.last_timing(1))
        eg64 = float(np.mean(t64))
        # every comparison solves against the FP64 differential of the same state
        if learned:
            two_level_same = compare_solve(args.metric_precision, None, args.pc_precision)
            two_level_same["step_ms"] = round(float(np.mean(eg_ms)) + two_level_same["solve_ms"], 1)
            print(f"# two-level P0 + w diag(A)^-1, headline precisions: {two_level_same}", file=sys.stderr, flush=True)
        two_level = compare_solve("fp64", None)
        two_level["energy_grad_fp64_ms"] = round(eg64, 3)
        two_level["step_ms"] = round(eg64 + two_level["solve_ms"], 1)
        print(f"# two-level P0 + w diag(A)^-1, FP64: {two_level}", file=sys.stderr, flush=True)
        model = learned or (LP.DEFAULT_MODEL if os.path.exists(LP.DEFAULT_MODEL) else "")
        if model:
            learned_cmp = compare_solve("fp64", model)
            learned_cmp["energy_grad_fp64_ms"] = round(eg64, 3)
            learned_cmp["step_ms"] = round(eg64 + learned_cmp["solve_ms"], 1)
            print(f"# learned D^1/2 (P0 + w diag(A)^-1) D^1/2, FP64: {learned_cmp}", file=sys.stderr, flush=True)
        if args.plain_pc_max_iter > 0:
            A0 = R.TangentPointMetric0(wl.q, wl.p, precision="fp64")
            A0.SetPreconditionerDiagonalWeight(0.0)
            per_it = float(np.mean(mm_ms)) + pc_ms
            cap = int(max(10, min(args.plain_pc_max_iter, 60000.0 / max(per_it, 1e-3))))
            st = H.rpm_metric_solve(A0._h, M._h, 1.0, dY, 3, 0.0, dX, 3, 3, cap, args.tol, solve_flags, 1, it_c, _p(rel))
            plain = {"solve_ms": round(M.last_timing(3), 1), "iterations": int(it_c[0]), "converged": st == 0,
                     "rel_residuals": [float(x) for x in rel], "iteration_cap": cap,
                     "energy_grad_fp64_ms": round(eg64, 3),
                     "step_ms": round(eg64 + M.last_timing(3), 1)}
            plain["step_ms_is_lower_bound"] = not plain["converged"]
            print(f"# plain FP64 form (reference P0 alone): {plain}", file=sys.stderr, flush=True)

    pc_compare = None
    if args.pc_compare_workload and not args.no_solve and not args.no_compare:
        try:
            pc_compare = preconditioner_comparison(R, WORKLOADS, args.pc_compare_workload, args.tol, args.max_iter,
                                                   LP.DEFAULT_MODEL, threads)
        except Exception as e:  # reported, not fatal
            pc_compare = {"error": str(e)}

    pairs_nf = b["near_nnz"] // 2
    pairs_ff = b["far_nnz"] // 2
    ff_s = float(np.mean(ff_ms)) * 1e-3
    ff_achieved = pairs_ff * FF_FLOPS_VALUE_DIFF / ff_s / 1e12 if ff_s > 0 else 0.0
    tree_ms = float(np.mean(mm_ms)) - float(np.mean(spmm_ms))
    tree_bytes = tree_pass_bytes(b, M.dom_dim)
    nf_flops = pairs_nf * NF_FLOPS_VALUE_DIFF
    nf_avg_s = float(np.mean(nf_ms)) * 1e-3
    achieved = nf_flops / nf_avg_s / 1e12
    peak = PEAK_FP32_TFLOPS if args.precision == "mixed" else PEAK_FP64_TFLOPS
    spmm_bytes = spmm_algorithmic_bytes(b, args.metric_precision)
    spmm_s = float(np.mean(spmm_ms)) * 1e-3
    spmm_gbs = spmm_bytes / spmm_s / 1e9
    traffic = MEASURED_SPMM_TRAFFIC.get((wl.name, "fp64" if refined else args.metric_precision))
    # the roofline line follows the SpMM the solve launches most: with mixed-refined the CG iterations
    # multiply with the FP32 copies (one FP64 product per solve replaces the residual)
    if refined:
        rl_kernel = ("k_nf_spmm_rows (near-field SpMM over the FP32 copies of the metric values, FP64 "
                     "accumulation, fused post-operator epilogue; one per CG iteration)")
        rl_bytes = spmm_algorithmic_bytes(b, "mixed")
        rl_s = float(np.mean(spmm32_ms)) * 1e-3
        rl_traffic = MEASURED_SPMM_TRAFFIC.get((wl.name, "fp32-copies"))
    else:
        rl_kernel = "k_nf_spmm (near-field metric block SpMM with fused post-operator epilogue, one per MultiplyMetric)"
        rl_bytes, rl_s, rl_traffic = spmm_bytes, spmm_s, traffic
    rl_gbs = rl_bytes / rl_s / 1e9
    result = {
        "metric": "energy+grad ms & metric-solve ms to 1e-6 residual, learned vs plain PC, 1 GPU",
        "value": round(ms_per_step, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64" if args.precision == "fp64" else "f32-pairs/f64-accumulate")
                 + ("" if args.metric_precision == args.precision else f", metric {args.metric_precision}"),
        "data": "synthetic",
        "config": {
            "workload": wl.name, "triangles": b["T"], "vertices": b["V"], "q": wl.q, "p": wl.p,
            "theta_far": wl.far_theta, "theta_near": wl.near_theta, "leaf_size": wl.leaf,
            "near_nnz": b["near_nnz"], "far_nnz": b["far_nnz"], "very_near_nnz": b["verynear_nnz"],
            "precision": args.precision, "metric_precision": args.metric_precision,
            "preconditioner_precision": args.pc_precision,
            "step": ("energy + differential (device-resident) + metric solve to |B - A X|/|B| <= "
                     f"{args.tol:g} per column ({args.solver})") if not args.no_solve
                    else "energy + differential (device-resident)",
            "preconditioner": ("learned D^1/2 (P0 + w diag(A)^-1) D^1/2 (BF16 inference), model "
                               + os.path.relpath(learned, ROOT)) if learned else
                              ("P0 + w diag(A)^-1" if args.pc_weight != 0 else "reference P0 = H^-1 L H^-1"),
        },
        "energy": float(en[0]),
        "energy_grad_ms": round(float(np.mean(eg_ms)), 3),
        "near_kernel_ms": round(float(np.mean(nf_ms)), 3),
        "far_kernel_ms": round(float(np.mean(ff_ms)), 3),
        "metric_assembly_ms": round(metric_assembly_dev_ms, 3),
        "metric_assembly_wall_s": round(t_assemble, 3),
        "pl_assembly_ms": round(pl_assembly_dev_ms, 3),
        "h1_factor_upload_s": round(t_pc_prep, 2),
        "next_state": None,
        "multiply_metric_ms": round(float(np.mean(mm_ms)), 3),
        "near_spmm_ms": round(float(np.mean(spmm_ms)), 3),
        "near_spmm_GBs_algorithmic": round(spmm_gbs, 1),
        "near_spmm_fp32_copies_ms": round(float(np.mean(spmm32_ms)), 3) if refined else None,
        "near_spmm_fp32_copies_GBs_algorithmic": round(spmm_algorithmic_bytes(b, "mixed") / (float(np.mean(spmm32_ms)) * 1e-3) / 1e9, 1)
                                                 if refined else None,
        "multiply_preconditioner_ms": round(pc_ms, 3),
        "h1_solve_ms": round(h1_ms, 3),
        "solve_ms": None if args.no_solve else round(float(np.mean(sv_ms)), 2),
        "solve_iterations": None if args.no_solve else solve_iters,
        "solve_rel_residuals": None if args.no_solve else solve_rel,
        "solve_rel_residuals_fp64_check": fp64_check,
        "solve_stagnated_steps": None if args.no_solve else stagnated_steps,
        "cold": None if args.no_solve else {
            "energy_grad_host_ms": round(float(np.median(cold_eg)), 2), "solve_host_ms": round(float(np.median(cold_sv)), 2),
            "note": "wall time of the host-pointer calls: the differential copied to host memory (D2H), the solve reading B "
                    "from and writing X to host memory (H2D + D2H)"},
        "preconditioner_comparison": pc_compare,
        "solver": args.solver,
        "pc_weight": pc_weight,
        "pc_setup_ms": round(pc_setup_dev_ms, 3),
        "solve_learned_fell_back": (not headline_learned) if learned else None,
        "solve_two_level_same_precision": two_level_same,
        "solve_two_level_fp64": two_level,
        "solve_learned_fp64": learned_cmp,
        "speedup_vs_fp64_same_preconditioner": (round((learned_cmp if learned else two_level)["step_ms"] / ms_per_step, 2)
                                                if (learned_cmp if learned else two_level) else None),
        "plain_fp64_form": plain,
        "speedup_vs_plain_fp64_form": (round(plain["step_ms"] / ms_per_step, 2) if plain else None),
        "max_safe_step": safe,
        "tree_bct_build_s": round(t_build, 3),
        "tree_bct_build": "host" if args.host_tree else "device",
        "tree_bct_build_kernel_ms": (round(build_kernel_ms, 2) if build_kernel_ms is not None else None),
        "upload_s": round(t_upload, 2),
        "pairs_per_s": round((pairs_nf + pairs_ff) / (float(np.mean(eg_ms)) * 1e-3), 1),
        "roofline": {
            "kernel": rl_kernel,
            "bound": "hbm",
            "achieved": round(rl_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(rl_gbs / PEAK_HBM_GBS, 4),
            "algorithmic_bytes": int(rl_bytes), "kernel_ms": round(rl_s * 1e3, 4),
            "traffic": rl_traffic,
            "traffic_source": "profiles/k_nf_spmm_sphere_1310720_traffic.md" if rl_traffic else None,
        },
        "roofline_fp64_spmm": {
            "kernel": "k_nf_spmm<double2> (FP64 metric values: MultiplyMetric, the residual replacement)",
            "bound": "hbm", "achieved": round(spmm_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(spmm_gbs / PEAK_HBM_GBS, 4), "algorithmic_bytes": int(spmm_bytes),
            "kernel_ms": round(spmm_s * 1e3, 4), "traffic": traffic,
        } if refined else None,
        "roofline_far_energy": {
            "kernel": "k_ff (far-field energy + differential, cluster pairs)", "bound": "mfma", "compute_unit": "valu",
            "achieved": round(ff_achieved, 3), "peak": peak, "unit": "TFLOP/s", "frac": round(ff_achieved / peak, 4),
            "flops_per_pair": FF_FLOPS_VALUE_DIFF, "pairs": pairs_ff, "kernel_ms": round(ff_s * 1e3, 4),
        },
        "tree_passes": {
            "kernels": "pre-operator, leaf and up-sweep sums, far-field SpMM, leaf path sums, post-operator vertex gather "
                       "(one MultiplyMetric minus its near-field SpMM)",
            "bound": "hbm", "ms": round(tree_ms, 4), "algorithmic_bytes": int(tree_bytes),
            "achieved": round(tree_bytes / (tree_ms * 1e-3) / 1e9, 1) if tree_ms > 0 else None, "peak": PEAK_HBM_GBS, "unit": "GB/s",
        },
        "roofline_near_energy": {
            "kernel": "k_nf<Q6P12,E,D> (near-field energy + differential)",
            "bound": "mfma", "compute_unit": "valu",
            "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "algorithmic_flops": nf_flops, "flops_per_pair": NF_FLOPS_VALUE_DIFF, "pairs": pairs_nf,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        cores, model = host_cpu_info()
        th = args.cpu_threads or min(16, cores)
        try:
            result["cpu_baseline"] = cpu_baseline(M, wl, th, 0 if args.no_solve else solve_iters)
            result["cpu_baseline"].update({"host_cores_available": cores, "cpu_model": model})
        except Exception as e:  # reported, not fatal
            result["cpu_baseline"] = {"error": str(e)}
    # per-state preparation after a move, what every further state of a gradient flow pays: the vertices
    # moved by a small random field (same topology), then the same preparation as for the first state; the
    # H^1 factor keeps its symbolic structure (numeric refactorisation, values-only upload). Runs last, so
    # that every measurement above (CPU baseline included) sees the first state.
    next_state = None
    if not args.no_solve:
        Xv = M.VertexCoordinates()
        M.SemiStaticUpdate(Xv + 1e-4 * np.random.default_rng(2).standard_normal(Xv.shape))
        t0 = time.perf_counter()
        A.assemble(M)
        t_asm2 = time.perf_counter() - t0
        t0 = time.perf_counter()
        A.MultiplyPreconditioner(M, 1.0, np.ones((M.V, 3)), 0.0, Z0)
        t_prep2 = time.perf_counter() - t0
        next_state = {"metric_assembly_wall_s": round(t_asm2, 3), "h1_factor_upload_s": round(t_prep2, 2),
                      "note": "after SemiStaticUpdate with a 1e-4 random move: metric assembly, then pseudo-Laplacian, "
                              "H1 numeric refactorisation + values-only upload and preconditioner setup"}
        print(f"# next state: metric assembly {t_asm2 * 1e3:.1f} ms (wall), H1 refactor+upload+setup {t_prep2:.2f} s (wall)",
              file=sys.stderr, flush=True)

    result["next_state"] = next_state
    hip.hipFree(Yd)
    hip.hipFree(Xd)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
