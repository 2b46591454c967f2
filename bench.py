"""Throughput of the single-device Fold3dPMM solve on one MI355X.

Workload (the reference's own headline configuration): the fold model
F|grad T| + a.grad T = 1 on the reference's implicit-surface input, 512^3
nodes, FP64, 8 iterations of the 6 directional sweeps (reference
Stencil-OpenACC/AN_ISO_1_OpenACC.c:1187-1237).  One step = one whole solve of a
fresh copy of the input field (the device-to-device reset copy is inside the
timed region).  Metric: G cell-updates/s, cell-updates = 8 * 6 * (n-1) * n^2.

The JSON line also carries the roofline of the dominant kernel (HBM-bound
algorithmic bytes: read + write of the target node, 16 B per cell-update in
FP64), an FP32 run of the same workload, and the CPU baseline (the oracle's
OpenMP restatement on this host's cores, on a bounded sample).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
for p in (REPO, os.path.join(REPO, "fgcs-openacc-advanced_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
# HBM bytes per plane launch of the split kernel (2 waves) at 512^3 FP64, from
# two rocprofv3 --pmc passes over the first iteration: FETCH_SIZE x 2.0 plus
# WRITE_SIZE x 1.0 KiB, the factors read off a plain device copy of known
# size (profiles/fold_split_f64_512_pmc_traffic.md)
PMC_TRAFFIC_512_F64 = (3454.1 * 2.0 + 898.2 * 1.0) * 1024
# The VALU view of the same launches, all from one profiled run over the first
# iteration (profiles/fold_split_f64_512_pmc_traffic.md): wave-instructions
# per plane launch (SQ_INSTS_VALU: 8192 waves of 436.0), the launch's average
# duration (kernel trace of the same command) and the clock implied by the
# FP64 FMA rate of an earlier run of the same passes.
PMC_VALU_INSTS_512_F64 = 3571925.6
PMC_VALU_LAUNCH_US_512_F64 = 12.867
PMC_VALU_CLOCK_GHZ = 1.955
# FP32 fold solve vs FP64 on the reference workload (8 iterations, up to 512^3)
FP32_BOUND_TEXT = ("signs and reached set equal; relative error < 1e-2 everywhere, above 1e-3 "
                   "at <= 0.5% of reached nodes, above 1e-4 at <= 10%, above 1e-5 at <= 25%")
# Traffic below L2 per launch of the other rows' dominant kernels, from
# rocprofv3 --pmc passes (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 KiB, the factors
# of the copy-kernel calibration above), keyed by row.  Each entry: (bytes per
# launch, the kernel's average launch in the kernel trace of the same command
# in us, which launches the passes counted, the profile file).
PMC_ROW_TRAFFIC = {
    "blob_speed_f64": ((3963.0 * 2.0 + 742.3) * 1024, 8.28,
                       "iteration 1 of the solve, 2 waves per row (ellip_split_kernel<double, "
                       "true, 2, 2>)", "profiles/ellip_split_f64_512_pmc.md"),
    "blob_speed_f32": ((2372.6 * 2.0 + 416.0) * 1024, 6.57,
                       "iteration 1 of the solve, 2 waves per row (ellip_split_kernel<float, "
                       "true, 2, 2>)", "profiles/ellip_split_f32_512_pmc.md"),
    "tilted_field": ((48087.4 * 2.0 + 2531.8) * 1024, 25.46,
                     "iteration 1 of the solve (ellip_metric_kernel<float, true, 1>), counted "
                     "on the kernel before its loads were issued ahead of the first wait (the "
                     "same bytes; 25.46 us per launch then, 22.9 us now)",
                     "profiles/metric_tilted_field_1280_f32_pmc.md"),
    "batch": ((79909.3 * 2.0 + 4453.3) * 1024, 49.41,
              "the 1914 launches of iteration 1 of one batch of all 96 instances (bytes per "
              "launch are their mean)", "profiles/batch_plane_96inst_f32_pmc.md"),
    "adjoint": ((25560.1 * 2.0 + 4285.3) * 1024, 14.39,
                "all 6144 launches of one gradient's adjoint solves",
                "profiles/adjoint_plane_256_f32_pmc.md"),
}


def roofline_row(updates, bytes_per_update, sweep_ms, launches, nodes_per_launch, kernel,
                 row=None):
    """HBM roofline of one row: algorithmic bytes (bytes_per_update per
    cell-update, the read and write of the target node plus the per-node
    parameters the workload names) over the device time of the row's sweeps
    (HIP events around them, so layout moves and launch gaps are included: a
    lower bound on the kernel's own rate), against the 8 TB/s peak."""
    achieved = updates * bytes_per_update / (sweep_ms / 1e3) / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "kernel": kernel, "bytes_per_update": bytes_per_update,
           "bytes_per_launch": bytes_per_update * nodes_per_launch,
           "kernel_us": sweep_ms * 1e3 / launches if launches else None, "traffic": None}
    if row in PMC_ROW_TRAFFIC:
        tb, us, scope, src = PMC_ROW_TRAFFIC[row]
        out.update({"traffic": tb, "traffic_scope": scope, "traffic_source": src,
                    "profiled_kernel_us": us,
                    "profiled_traffic_gbs": tb / (us * 1e-6) / 1e9,
                    "profiled_algorithmic_gbs": bytes_per_update * nodes_per_launch /
                                                (us * 1e-6) / 1e9})
    return out


A100_V2_512_GCUPS = 8 * 6 * 511 * 512 ** 2 / 3.038 / 1e9  # plots/perfect-2-gpu.py:20,24
# the reference's published single-A100 times (s) of its version 2 by grid side
A100_V2_S = dict(zip([128, 256, 384, 512, 640, 768, 896, 1024, 1152, 1280],
                     [0.182216, 0.607416, 1.502711, 3.037851, 5.495726, 8.727377, 13.411582,
                      19.195663, 26.060296, 35.163471]))  # plots/perfect-2-gpu.py:20,24


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cell_updates(n, iters):
    return iters * 6 * (n - 1) * n * n


def time_solves(solver, T0d, n, steps, warmup, dtype, torch, engine=0, iters=8, stats=None):
    """Wall time of `steps` solves (reset copy included), plus the summed
    HIP-event time of the sweep loops themselves."""
    T = torch.empty_like(T0d, dtype=dtype)
    for _ in range(warmup):
        T.copy_(T0d)
        solver.solve(T, n, max_iters=iters, engine=engine)
    torch.cuda.synchronize()
    sweep_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        T.copy_(T0d)
        st = solver.solve(T, n, max_iters=iters, engine=engine)
        sweep_ms += st.sweep_ms
        if stats is not None:
            stats.append(st)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3, T, sweep_ms


def ellip_run(solver, n, dtype, tol, torch, threads=0, cpu_full=False, groups=0):
    """The 512^3 tilted-metric, Gaussian-blob speed workload, to tolerance.
    With `threads`, the FP64 run also times the elliptic oracle on the host:
    its first iteration (bounded sample), or with `cpu_full` the whole solve
    to the same tolerance.  `groups` (1, 2 or 4 waves per node row) replaces
    the autotuner's pick, so that profiled runs all see the same kernel."""
    import fold3d
    from fold3d import workloads as wl
    W = wl.blob_speed(n, xp=torch, dtype=dtype, device="cuda")
    T0 = wl.point_source_start(W, xp=torch, dtype=dtype, device="cuda")
    if groups:
        solver.set_tuning(groups, fold3d.MODEL_ELLIP)
    else:
        groups, _ = solver.autotune(T0, n, spacing=W.h, model=W.model(), iterations=1)
    T = T0.clone()
    st = solver.solve(T, n, spacing=W.h, model=W.model(), max_iters=200, tol=tol)  # warm
    T.copy_(T0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = solver.solve(T, n, spacing=W.h, model=W.model(), max_iters=200, tol=tol)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    W_upd = st.cell_updates
    r = 8 if dtype == torch.float64 else 4
    out = {"iterations": st.iterations, "max_dT": st.max_dT, "sweep_ms": st.sweep_ms,
           "wall_ms": wall, "value": W_upd / (st.sweep_ms / 1e3) / 1e9,
           "unit": "Gcell-updates/s", "cell_updates": W_upd, "split_groups": groups,
           "roofline": roofline_row(W_upd, 3 * r, st.sweep_ms, st.iterations * 6 * (n - 1), n * n,
                                    f"ellip_split_kernel<{'double' if r == 8 else 'float'}>",
                                    f"blob_speed_{'f64' if r == 8 else 'f32'}")}
    if threads and dtype == torch.float64:
        # bounded sample: the first iteration, on the host (elliptic oracle)
        # and on the GPU from the same start field
        from oracle.oracle import Oracle
        o = Oracle()
        o.set_threads(threads)
        Tc = T0.cpu().numpy().copy()
        V = W.speed_field.cpu().numpy()
        t0 = time.perf_counter()
        if cpu_full:
            its, _ = o.ellip_sweeps(Tc, n, 200, tol, W.h, M=W.M, speed=W.speed, speed_field=V)
        else:
            its = 1
            o.ellip_sweeps(Tc, n, 1, 0.0, W.h, M=W.M, speed=W.speed, speed_field=V)
        dt = time.perf_counter() - t0
        T1 = T0.clone()
        if cpu_full:
            T1 = T  # the GPU's solve to the same tolerance
        else:
            solver.solve(T1, n, spacing=W.h, model=W.model(), max_iters=1)
        what = (f"the whole solve to {tol:g} ({its} iterations)" if cpu_full else
                "1 iteration (6 sweeps)")
        out["cpu_baseline"] = {
            "value": cell_updates(n, its) / dt / 1e9, "unit": "Gcell-updates/s",
            "cores": threads, "kind": "port",
            "sample": f"{what} of the same {n}^3 FP64 solve, {cell_updates(n, its):.3e} "
                      f"cell-updates in {dt:.1f} s",
            "gpu_same_bits": bool(np.array_equal(Tc.view(np.uint64),
                                                 T1.cpu().numpy().view(np.uint64)))}
        del T1
    return out, T


def batch_run(solver, torch, count=96, check=24, tuning=None):
    """96 instances, sides uniform in {64, ..., 320}: per-instance blob speed,
    tilt/axis-ratio fields and interior source, FP32 to max|dT| < 1e-6 —
    solved as one batch and as a loop of plain solves; then every instance
    again in FP64 (batches of `check`) against the FP32 batch's fields."""
    from fold3d import workloads as wl
    sides = wl.batch_sides(count)
    Ws, T0s = [], []
    for k, n in enumerate(sides):
        W = wl.batch_instance(n, 1000 + k, xp=torch, dtype=torch.float32, device="cuda")
        Ws.append(W)
        T0s.append(wl.point_source_start(W, xp=torch, dtype=torch.float32, device="cuda"))
    cells = sum(n ** 3 for n in sides)
    log(f"bench: batch of {count} instances, {cells / 1e9:.2f} G nodes")
    # warm-up on three instances (code load, graph path), then the schedule
    # autotuner; a schedule that is set is in force from the warm-up on
    if tuning:
        solver.set_batch_tuning(*tuning)
    warm = [t.clone() for t in T0s[:3]]
    solver.solve_batched(Ws[:3], warm, max_iters=2, tol=0.0)
    if tuning:
        tuned = solver.batch_tuning()
        log(f"bench: batch schedule (waves per row, blocks per CU, rows per pull) = {tuned} "
            "(set, no autotuning)")
    else:
        tuned, tune_ms = solver.autotune_batched(Ws, T0s, iterations=1)
        log(f"bench: batch autotuner picked (waves per row, blocks per CU, rows per pull) = "
            f"{tuned} ({tune_ms:.0f} ms/iteration)")
    # three batched solves from the same start fields; the median one is
    # reported (single solves have read 226-236 solves/s)
    Ts = [t.clone() for t in T0s]
    runs = []
    for _ in range(3):
        for t, t0_ in zip(Ts, T0s):
            t.copy_(t0_)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its, mds, dev_ms = solver.solve_batched(Ws, Ts, max_iters=200, tol=1e-6)
        torch.cuda.synchronize()
        runs.append((time.perf_counter() - t0, its, mds, dev_ms))
    batch_runs_s = sorted(r[0] for r in runs)
    batch_s, its, mds, dev_ms = sorted(runs, key=lambda r: r[0])[1]
    upd = sum(6 * (n - 1) * n * n * k for n, k in zip(sides, its))
    log(f"bench: batched {batch_s:.2f} s ({count / batch_s:.1f} solves/s), iterations "
        f"{min(its)}..{max(its)}")
    # looped plain solves of the same instances
    Ls = [t.clone() for t in T0s]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lits = []
    for W, T in zip(Ws, Ls):
        st = solver.solve(T, W.n, spacing=W.h, model=W.model(), max_iters=200, tol=1e-6)
        lits.append(st.iterations)
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    log(f"bench: looped {loop_s:.2f} s ({count / loop_s:.1f} solves/s)")
    # FP32 batch vs FP64 batches of the same instances, all of them, `check`
    # instances per FP64 batch (bounded memory); nodes at least 8 cells from
    # the source (T > 8 h / 3.6), relative error
    del Ls, T0s
    torch.cuda.empty_cache()
    worst, worst_k, reached = 0.0, -1, True
    t0 = time.perf_counter()
    for g0 in range(0, count, check):
        idx = list(range(g0, min(count, g0 + check)))
        W64 = [wl.batch_instance(sides[k], 1000 + k, xp=torch, dtype=torch.float64, device="cuda")
               for k in idx]
        T64 = [wl.point_source_start(W, xp=torch, dtype=torch.float64, device="cuda")
               for W in W64]
        solver.solve_batched(W64, T64, max_iters=200, tol=1e-10)
        for k, T in zip(idx, T64):
            a, b = T, Ts[k].double()
            reached &= bool(torch.equal(a < 1e15, b < 1e15))
            far = (a < 1e15) & (a > 8 * Ws[k].h / 3.6)
            e = float(((b - a).abs()[far] / a[far]).max())
            if e > worst:
                worst, worst_k = e, k
        del W64, T64
        torch.cuda.empty_cache()
    log(f"bench: FP32 batch vs FP64 batches, all {count} instances: max relative error "
        f"{worst:.2e} (instance {worst_k}, side {sides[worst_k]}), "
        f"{time.perf_counter() - t0:.1f} s")
    del Ws, Ts
    torch.cuda.empty_cache()
    return {"instances": count, "nodes": cells, "batched_s": batch_s,
            "batched_s_runs": batch_runs_s, "batched_device_ms": dev_ms,
            "batched_solves_per_s": count / batch_s, "looped_s": loop_s,
            "looped_solves_per_s": count / loop_s, "speedup": loop_s / batch_s,
            "iterations_min": min(its), "iterations_max": max(its),
            "batched_cell_updates": upd, "batched_gcups": upd / batch_s / 1e9,
            # 7 x 4 B per cell-update (T read and write; v, theta, phi, ra, rb),
            # over the batch's device time; a launch serves every instance that
            # has its plane step: with all instances live, sum (n-1) n^2 nodes
            # over the max(n) - 1 steps of a direction
            "roofline": roofline_row(upd, 28, dev_ms, batch_launches(sides, its),
                                     sum((n - 1) * n * n for n in sides) / (max(sides) - 1),
                                     "batch_plane_kernel<float>", "batch"),
            "batch_tuning": {"waves_per_row": tuned[0], "blocks_per_cu": tuned[1],
                             "rows_per_pull": tuned[2], "sub_batches": solver.batch_streams()},
            "fp32_max_rel_err_vs_fp64": {
                "instances": count, "value": worst, "worst_instance": worst_k,
                "worst_side": sides[worst_k], "reached_sets_equal": reached,
                "nodes": "reached, T > 8 h / 3.6 (8 cells from the source at speed 3.6)",
                "fp64": f"per-instance FP64 solves to 1e-10, batches of {check}"}}


def size_sweep(torch, sizes, steps=3):
    """The headline solve (fold model, reference input built on the device,
    FP64, 8 iterations, autotuned waves per row) at other grid sides: median
    sweep time of `steps` solves, against the reference's published A100 time
    for the same side."""
    import fold3d
    out = []
    solver = fold3d.Solver(0)
    try:
        for n in sizes:
            T0 = torch.empty((n + 2) ** 3, dtype=torch.float64, device="cuda")
            solver.initialiser_device(T0, n)
            groups, _ = solver.autotune(T0, n, iterations=1)
            T = T0.clone()
            solver.solve(T, n, max_iters=8)  # warm
            ms = []
            for _ in range(steps):
                T.copy_(T0)
                ms.append(solver.solve(T, n, max_iters=8).sweep_ms)
            med = sorted(ms)[len(ms) // 2]
            r = {"n": n, "groups": groups, "sweep_ms": med,
                 "value": cell_updates(n, 8) / (med / 1e3) / 1e9, "unit": "Gcell-updates/s"}
            if n in A100_V2_S:
                r["a100_v2_s"] = A100_V2_S[n]
                r["speedup_vs_a100_v2"] = A100_V2_S[n] / (med / 1e3)
            log(f"bench: {n}^3: {med:.1f} ms, {r['value']:.2f} Gcell-upd/s ({groups} waves per row)")
            out.append(r)
            del T0, T
            torch.cuda.empty_cache()
    finally:
        solver.close()
    return out


def batch_launches(sides, its):
    """Plane launches of a batch: one per (iteration, direction, plane step)
    while any instance still has that step (instances are sorted by size)."""
    total = 0
    for it in range(max(its)):
        live = [n for n, k in zip(sides, its) if k > it]
        total += 6 * (max(live) - 1)
    return total


def tilted_homogeneous_run(solver, torch, n=128, threads=0):
    """Homogeneous tilted medium, 128^3 FP64 on [0,1]^3, source at the centre
    (between nodes), to max|dT| < 1e-10; error against sqrt(x^T M^-1 x)/v."""
    from fold3d import workloads as wl
    W = wl.tilted_homogeneous(n)
    T0 = wl.point_source_start(W, xp=torch, dtype=torch.float64, device="cuda")
    T = T0.clone()
    solver.solve(T, n, spacing=W.h, model=W.model(), max_iters=200, tol=1e-10)
    T.copy_(T0)
    torch.cuda.synchronize()
    st = solver.solve(T, n, spacing=W.h, model=W.model(), max_iters=200, tol=1e-10)
    exact = wl.closed_form(W, xp=torch, device="cuda")
    reached = T < 1e15
    err = (T - exact).abs()[reached]
    far = reached & (exact > 0.25)
    out = {"n": n, "dtype": "f64", "iterations": st.iterations, "max_dT": st.max_dT,
           "sweep_ms": st.sweep_ms, "value": st.cell_updates / (st.sweep_ms / 1e3) / 1e9,
           "unit": "Gcell-updates/s", "max_abs_err_vs_closed_form": float(err.max()),
           "mean_abs_err_vs_closed_form": float(err.mean()),
           "max_rel_err_beyond_r0.25": float(((T - exact).abs()[far] / exact[far]).max())}
    if threads:  # the elliptic oracle on the host, same start field and stop rule
        from oracle.oracle import Oracle
        o = Oracle()
        o.set_threads(threads)
        Tc = T0.cpu().numpy().copy()
        t0 = time.perf_counter()
        it, _ = o.ellip_sweeps(Tc, n, 200, 1e-10, W.h, M=W.M, speed=W.speed)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cell_updates(n, it) / dt / 1e9, "unit": "Gcell-updates/s",
                               "cores": threads, "kind": "port", "s": dt, "iterations": it,
                               "same_bits_as_gpu": bool(np.array_equal(
                                   Tc.view(np.uint64), T.cpu().numpy().view(np.uint64)))}
    return out


def reference_small_run(torch, threads, n=128):
    """The reference's own workload at 128^3 (implicit-surface input, fold
    model, FP64, 8 iterations) on the GPU and, with `threads`, in the oracle on
    the host: time and the bits of the two results."""
    import fold3d
    T0 = fold3d.implicit_initialiser(n)
    T = torch.from_numpy(T0).cuda()
    solver = fold3d.Solver(0)  # a handle of its own, tuned for this size
    try:
        groups, _ = solver.autotune(T, n, iterations=1)
        solver.solve(T, n, max_iters=8)  # warm
        T.copy_(torch.from_numpy(T0))
        torch.cuda.synchronize()
        st = solver.solve(T, n, max_iters=8)
    finally:
        solver.close()
    out = {"n": n, "dtype": "f64", "iterations": 8, "sweep_ms": st.sweep_ms, "split_groups": groups,
           "value": cell_updates(n, 8) / (st.sweep_ms / 1e3) / 1e9, "unit": "Gcell-updates/s"}
    if threads:
        from oracle.oracle import Oracle
        o = Oracle()
        o.set_threads(threads)
        Tc = T0.copy()
        t0 = time.perf_counter()
        o.sweeps(Tc, n, 8)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cell_updates(n, 8) / dt / 1e9, "unit": "Gcell-updates/s",
                               "cores": threads, "kind": "port", "s": dt,
                               "same_bits_as_gpu": bool(np.array_equal(
                                   Tc.view(np.uint64), T.cpu().numpy().view(np.uint64)))}
    return out


def tilted_field_run(torch, n=1280):
    """Per-node tilt and axis-ratio fields with v = 1.5 + 2z + 5 % blobs,
    FP32, source near the top face, to max|dT| < 1e-6 (metric engine)."""
    import fold3d
    from fold3d import workloads as wl
    solver = fold3d.Solver(0)
    try:
        W = wl.tilted_field(n, xp=torch, dtype=torch.float32, device="cuda")
        T0 = wl.point_source_start(W, xp=torch, dtype=torch.float32, device="cuda")
        torch.cuda.empty_cache()
        groups, tune_ms = solver.autotune(T0, n, spacing=W.h, model=W.model(), iterations=1)
        log(f"bench: tilted field: autotuner picked {groups} waves per node row ({tune_ms:.0f} ms/iteration)")
        # three solves from the same start; the median one is reported (one
        # solve right after the autotuner has read up to 6 % slow)
        T = torch.empty_like(T0)
        runs = []
        for _ in range(3):
            T.copy_(T0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = solver.solve(T, n, spacing=W.h, model=W.model(), max_iters=200, tol=1e-6)
            torch.cuda.synchronize()
            runs.append((st.sweep_ms, time.perf_counter() - t0, st))
        runs.sort(key=lambda r: r[0])
        _, wall, st = runs[1]
        out = {"n": n, "dtype": "f32", "iterations": st.iterations, "max_dT": st.max_dT,
               "sweep_ms": st.sweep_ms, "sweep_ms_runs": [r[0] for r in runs], "wall_s": wall,
               "groups": groups,
               "value": st.cell_updates / (st.sweep_ms / 1e3) / 1e9, "unit": "Gcell-updates/s",
               "cell_updates": st.cell_updates,
               # 7 x 4 B: T read and write, v, theta, phi, ra, rb per node
               "roofline": roofline_row(st.cell_updates, 28, st.sweep_ms,
                                        st.iterations * 6 * (n - 1), n * n,
                                        "ellip_metric_kernel<float>", "tilted_field"),
               "torch_device_bytes": torch.cuda.max_memory_allocated(),
               # the metric engine's scratch: a T copy and the seven metric
               # fields in each of the two sweep layouts
               "solver_scratch_bytes": 15 * (n + 2) ** 3 * 4}
        del W, T0, T, runs
    finally:
        solver.close()
        torch.cuda.empty_cache()
    return out


def adjoint_run(solver, torch, n=256):
    """Tomography gradient on the 256^3 FP32 model: 16 sources on a 4x4 grid
    below the top face, 64x64 top-face receivers, observed times from the
    background plus 3 anomalies of +-5 %.  One gradient evaluation = the 16
    forward solves (one batch, checkpointed) + 16 adjoint solves."""
    from fold3d.tomography import Tomography, tomography_fields, top_face_survey
    f = tomography_fields(n, xp=torch, dtype=torch.float32, device="cuda")
    ft = tomography_fields(n, xp=torch, dtype=torch.float32, device="cuda", anomalies=True)
    S = top_face_survey(n, 4, 64)
    tomo = Tomography(solver, S, f["phi"], f["ra"], f["rb"], batch=16)
    tomo.observe(ft["slowness"], ft["theta"])       # also the warm-up
    del ft
    torch.cuda.synchronize()
    # looped plain forward solves of the 16 sources
    t0 = time.perf_counter()
    for k in range(len(S.sources)):
        tomo.solve_group([k], f["slowness"], f["theta"])
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    tomo.stats.clear()
    s0 = f["slowness"].clone().requires_grad_(True)
    t0_ = f["theta"].clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = tomo.misfit(s0, t0_)
    torch.cuda.synchronize()
    fwd_s = time.perf_counter() - t0
    J.backward()
    torch.cuda.synchronize()
    grad_s = time.perf_counter() - t0
    st = tomo.stats
    nsrc = len(S.sources)
    adj_ms = sum(st["adjoint_ms"])
    out = {"n": n, "dtype": "f32", "sources": nsrc, "receivers": S.nrec,
           "misfit": float(J.detach()), "gradient_s": grad_s,
           "gradient_evals_per_s": 1.0 / grad_s,
           "forward_batched_s": fwd_s, "forward_solves_per_s": nsrc / fwd_s,
           "forward_looped_solves_per_s": nsrc / loop_s,
           "adjoint_device_ms_total": adj_ms, "adjoint_solves_per_s": nsrc / (adj_ms / 1e3),
           "adjoint_iterations": [min(st["adjoint_iters"]), max(st["adjoint_iters"])],
           "forward_iterations": [min(st["forward_iters"]), max(st["forward_iters"])],
           "checkpoint_bytes": st["checkpoint_bytes"], "recomputed": st["recomputed"],
           "nodes_held_per_source": list(solver.last_held),
           # adjoint plane steps, per node visit: the in-edge mask (8 B), the
           # residual (4 B), the adjoint value read and written (8 B) and the
           # node's three stencil weights (12 B), each read once per pass by the
           # node its edge points to; the neighbours' adjoint values gathered
           # on top are taken as cache hits
           "roofline": roofline_row(nsrc * 6 * n ** 3 * float(np.mean(st["adjoint_iters"])), 32,
                                    adj_ms, 6 * n * max(st["adjoint_iters"]), nsrc * n * n,
                                    "adjoint_plane_kernel<float>", "adjoint"),
           "grad_norm_slowness": float(s0.grad.norm()), "grad_norm_theta": float(t0_.grad.norm())}
    del tomo, f, s0, t0_
    torch.cuda.empty_cache()
    return out


def _seeded_cells(g, srcs, H, count, seed):
    """`count` seeded cells whose gradient is above 1e-3 of the largest, at
    least 3 cells (max norm) from every source."""
    idx = np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max())
    i, j, k = idx % H, (idx // H) % H, idx // (H * H)
    far = np.ones(idx.size, bool)
    for s_ in srcs:
        si, sj, sk = s_ % H, (s_ // H) % H, s_ // (H * H)
        far &= np.maximum(np.maximum(abs(i - si), abs(j - sj)), abs(k - sk)) >= 3
    return np.random.default_rng(seed).choice(idx[far], count, replace=False)


def adjoint_fd_run(solver, torch, n=256, cells=32, threads=0):
    """The tomography gradient at its full configuration (256^3, 16 sources on
    a 4x4 grid, 64x64 top-face receivers, per-node speed and tilt), FP64,
    against central finite differences of the forward solve (relative step
    1e-3, and its Richardson combination with step 5e-4) at `cells` seeded
    cells per parameter; and, at 64^3 with the same survey layout, against the
    FP64 CPU adjoint (the oracle) on the same converged fields."""
    from fold3d.tomography import Tomography, tomography_fields, top_face_survey
    out = {"n": n, "dtype": "f64", "sources": 16, "receivers": 64 * 64, "cells_per_parameter": cells,
           "relative_steps": [1e-3, 5e-4]}
    t_start = time.perf_counter()
    f = tomography_fields(n, xp=torch, dtype=torch.float64, device="cuda")
    ft = tomography_fields(n, xp=torch, dtype=torch.float64, device="cuda", anomalies=True)
    S = top_face_survey(n, 4, 64)
    tomo = Tomography(solver, S, f["phi"], f["ra"], f["rb"], batch=16)
    tomo.observe(ft["slowness"], ft["theta"])
    del ft
    s0 = f["slowness"].clone().requires_grad_(True)
    t0 = f["theta"].clone().requires_grad_(True)
    J0 = tomo.misfit(s0, t0)
    J0.backward()
    out["misfit"] = float(J0.detach())
    out["nodes_held_per_source"] = list(solver.last_held)
    H = n + 2
    srcs = []
    for src in S.sources:
        i, j, k = [min(n, max(1, int(round(c / S.h)) + 1)) for c in src]
        srcs.append(i + H * (j + H * k))

    def J_of(name, c, delta):
        s_, t_ = f["slowness"].clone(), f["theta"].clone()
        (s_ if name == "slowness" else t_)[c] += delta
        with torch.no_grad():
            return float(tomo.forward_pass(s_, t_)[0])

    for name, grad in (("slowness", s0.grad), ("theta", t0.grad)):
        g = grad.cpu().numpy()
        sel = _seeded_cells(g, srcs, H, cells, seed=20261209)
        rows = []
        for q, c in enumerate(sel):
            base = float(f[name][c])
            d = 1e-3 * abs(base)
            d1 = (J_of(name, c, d) - J_of(name, c, -d)) / (2 * d)
            d2 = (J_of(name, c, d / 2) - J_of(name, c, -d / 2)) / d
            rows.append((g[c], d1, (4 * d2 - d1) / 3))
            if q % 8 == 7:
                log(f"bench: adjoint FD {name}: {q + 1} of {len(sel)} cells")
        R = np.array(rows)
        raw = np.abs(R[:, 0] - R[:, 1]) / np.abs(R[:, 1])
        ric = np.abs(R[:, 0] - R[:, 2]) / np.abs(R[:, 2])
        out[name] = {
            "raw_central_1e-3": {"median_rel": float(np.median(raw)), "max_rel": float(raw.max()),
                                 "frac_within_1e-4": float(np.mean(raw < 1e-4)),
                                 "frac_within_1e-3": float(np.mean(raw < 1e-3))},
            "richardson": {"median_rel": float(np.median(ric)), "max_rel": float(ric.max()),
                           "frac_within_1e-4": float(np.mean(ric < 1e-4)),
                           "frac_within_1e-3": float(np.mean(ric < 1e-3)),
                           "max_abs_over_max_grad": float(np.abs(R[:, 0] - R[:, 2]).max() /
                                                          np.abs(R[:, 2]).max())}}
        log(f"bench: adjoint FD {name}: raw central median {np.median(raw):.2e}, "
            f"Richardson median {np.median(ric):.2e}, within 1e-4: {np.mean(ric < 1e-4):.2f}")
    del tomo, f, s0, t0
    torch.cuda.empty_cache()
    # the same survey layout at 64^3 against the CPU adjoint
    if threads:
        from oracle.oracle import Oracle
        o = Oracle()
        o.set_threads(threads)
        m_ = 64
        f = tomography_fields(m_, xp=torch, dtype=torch.float64, device="cuda")
        ft = tomography_fields(m_, xp=torch, dtype=torch.float64, device="cuda", anomalies=True)
        S = top_face_survey(m_, 4, 64)
        tomo = Tomography(solver, S, f["phi"], f["ra"], f["rb"], batch=16)
        tomo.observe(ft["slowness"], ft["theta"])
        s0 = f["slowness"].clone().requires_grad_(True)
        t0 = f["theta"].clone().requires_grad_(True)
        _, res, ckpt = tomo.forward_pass(s0.detach(), t0.detach())
        tomo.misfit(s0, t0).backward()
        hf = {k: v.cpu().numpy() for k, v in f.items()}
        cs = np.zeros((m_ + 2) ** 3)
        ct = np.zeros((m_ + 2) ** 3)
        tc = time.perf_counter()
        for k in range(len(S.sources)):
            a, b, _ = o.ellip_adjoint(ckpt[k].cpu().numpy(), m_, S.h, S.rec_idx.astype(np.int64),
                                      res[k].cpu().numpy(), speed_field=1.0 / hf["slowness"],
                                      theta=hf["theta"], phi=hf["phi"], ra=hf["ra"], rb=hf["rb"])
            cs += a
            ct += b
        out["cpu_adjoint_64"] = {
            "n": m_, "sources": 16, "receivers": S.nrec, "cpu_s": time.perf_counter() - tc,
            "max_abs_diff_over_max_slowness": float(np.abs(s0.grad.cpu().numpy() - cs).max() /
                                                    np.abs(cs).max()),
            "max_abs_diff_over_max_theta": float(np.abs(t0.grad.cpu().numpy() - ct).max() /
                                                 np.abs(ct).max())}
        del tomo, f, ft, s0, t0, ckpt
        torch.cuda.empty_cache()
    out["check_s"] = time.perf_counter() - t_start
    return out


def host_cpu():
    """Model, sockets and physical cores of this host (from /proc/cpuinfo)."""
    model, cores, sockets = None, set(), set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                if key == "model name" and model is None:
                    model = val
                elif key == "physical id":
                    phys = val
                    sockets.add(val)
                elif key == "core id":
                    core = val
                elif not key and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    return {"model": model, "sockets": len(sockets) or None, "physical_cores": len(cores) or None,
            "logical_cpus": os.cpu_count()}


def cpu_baseline(T0_host, n, threads, iters=1):
    """The oracle (FP64 OpenMP restatement, bit-identical to the reference)
    on `iters` of the 8 iterations of the same 512^3 solve (one by default: a
    bounded sample; all 8 with --cpu-full)."""
    from oracle.oracle import Oracle
    o = Oracle()
    o.set_threads(threads)
    T = T0_host.copy()
    t0 = time.perf_counter()
    o.sweeps(T, n, iters)
    dt = time.perf_counter() - t0
    return cell_updates(n, iters) / dt / 1e9, dt, T


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-ellip", action="store_true")
    ap.add_argument("--no-ablation", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-small", "--no-c0", dest="no_small", action="store_true",
                    help="skip the 128^3 runs")
    ap.add_argument("--groups", type=int, default=0, choices=[0, 2, 4, 8],
                    help="waves per node row of the fold split kernel (0: autotune)")
    ap.add_argument("--ellip-groups", type=int, default=0, choices=[0, 1, 2, 4],
                    help="waves per node row of the elliptic split kernel in the blob-speed "
                         "rows (0: autotune)")
    ap.add_argument("--no-fixpt", action="store_true",
                    help="skip the solves to the exact fixed point (split and locked engines)")
    ap.add_argument("--field-n", "--c2-n", dest="field_n", type=int, default=1280,
                    help="side of the per-node tilt-field run (0: skip)")
    ap.add_argument("--iters", type=int, default=8, help="iterations per solve (the reference runs 8)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baselines over the whole solves (8 iterations of the headline, the "
                         "blob-speed FP64 solve to tolerance) instead of a one-iteration sample")
    ap.add_argument("--batch-tuning", default="",
                    help="schedule of the batched solve as 'waves,blocks_per_cu,rows_per_pull' "
                         "(default: the batch autotuner picks)")
    ap.add_argument("--batch-streams", type=int, default=0,
                    help="sub-batches the batched solve runs at once (default: the library's)")
    ap.add_argument("--sizes", default="",
                    help="comma-separated grid sides at which to time the headline solve as well "
                         "(device-built input; default: none)")
    ap.add_argument("--fd-cells", type=int, default=32,
                    help="seeded cells per parameter of the tomography gradient's finite-difference "
                         "check at 256^3 (0: skip the check)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this solver runs on one GPU only (asked for --gpus {args.gpus}); "
              "the solve does not shard across GPUs", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import fold3d

    n = args.n
    log(f"bench: host initializer {n}^3")
    t0 = time.perf_counter()
    T0 = fold3d.implicit_initialiser(n)
    log(f"bench: initializer took {time.perf_counter() - t0:.1f} s")
    T0d = torch.from_numpy(T0).to("cuda")
    solver = fold3d.Solver(0)
    host_init_s = time.perf_counter() - t0
    # the same field built on the device (same bits as the host initializer)
    Td = torch.empty_like(T0d)
    solver.initialiser_device(Td, n)  # warm-up (code load)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    n_exact = solver.initialiser_device(Td, n)
    torch.cuda.synchronize()
    dev_init_s = time.perf_counter() - t1
    init = {"host_s": host_init_s, "device_s": dev_init_s, "host_evaluated_nodes": n_exact,
            "same_bits": bool(torch.equal(Td.view(torch.int64), T0d.view(torch.int64)))}
    log(f"bench: device initializer {dev_init_s:.3f} s (host {host_init_s:.1f} s), "
        f"same bits: {init['same_bits']}")
    del Td

    if args.groups:
        solver.set_tuning(args.groups)
        groups = args.groups
        log(f"bench: {groups} waves per node row (set, no autotuning)")
    else:
        log("bench: autotuning waves per node row (one iteration each of 2, 4, 8)")
        groups, tune_ms = solver.autotune(T0d, n, iterations=1)
        log(f"bench: autotuner picked {groups} waves per node row ({tune_ms:.1f} ms/iteration)")

    iters = args.iters
    W = cell_updates(n, iters)
    launches = iters * 6 * (n - 1)
    log(f"bench: FP64 solve x{args.steps} (+{args.warmup} warmup), split engine (default)")
    head_stats = []
    ms, T, sweep_ms = time_solves(solver, T0d, n, args.steps, args.warmup, torch.float64, torch,
                                  iters=iters, stats=head_stats)
    ms_step = ms / args.steps
    sweep_ms_steps = sorted(st.sweep_ms for st in head_stats)
    value = W * args.steps / (ms / 1e3) / 1e9
    log(f"bench: FP64 {ms_step:.1f} ms/solve, {value:.2f} Gcell-upd/s")
    # average plane-kernel duration from the solve's own HIP events (the
    # four layout changes per iteration and inter-launch gaps are included,
    # so this is an upper bound on the kernel time)
    us = sweep_ms * 1e3 / (args.steps * launches)
    bytes_per_launch = 16.0 * n * n
    achieved = bytes_per_launch / (us * 1e-6) / 1e9

    # exact locking (a separate engine, same bits): time to the same result
    log("bench: FP64 solve, locked engine (exact locking of unchanged rows)")
    lst = []
    ms_l, T_l, _ = time_solves(solver, T0d, n, args.steps, args.warmup, torch.float64, torch,
                               engine=fold3d.ENGINE_LOCKED, iters=iters, stats=lst)
    locked = {"ms_per_step": ms_l / args.steps,
              "reference_equivalent_gcups": W * args.steps / (ms_l / 1e3) / 1e9,
              "computed_fraction": lst[-1].updates_computed / lst[-1].cell_updates,
              "same_bits_as_split": bool(torch.equal(T_l.view(torch.int64), T.view(torch.int64))),
              "note": "time to the same 8-iteration result; rows whose source rows did not "
                      "change since their last update in the same direction are not recomputed"}
    log(f"bench: locked {locked['ms_per_step']:.1f} ms/solve, "
        f"{locked['computed_fraction']:.3f} of the updates computed, same bits: "
        f"{locked['same_bits_as_split']}")
    del T_l

    # time to the exact fixed point (no node changes in a whole iteration),
    # split and locked engines: the reference's 8 iterations stop well short of it
    fixpt = None
    if not args.no_fixpt:
        log("bench: FP64 solves to the exact fixed point, split and locked engines")
        fixpt = {}
        res = {}
        for name, eng in (("split", fold3d.ENGINE_SPLIT), ("locked", fold3d.ENGINE_LOCKED),
                          ("auto", fold3d.ENGINE_AUTO)):
            Tf = T0d.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st_f = solver.solve(Tf, n, max_iters=1000, tol=1e-300, engine=eng)
            torch.cuda.synchronize()
            fixpt[name] = {"s": time.perf_counter() - t0, "iterations": st_f.iterations,
                           "computed_fraction": st_f.updates_computed / st_f.cell_updates}
            res[name] = Tf
        fixpt["same_bits"] = bool(torch.equal(res["split"].view(torch.int64),
                                              res["locked"].view(torch.int64)) and
                                  torch.equal(res["split"].view(torch.int64),
                                              res["auto"].view(torch.int64)))
        fixpt["locked_speedup"] = fixpt["split"]["s"] / fixpt["locked"]["s"]
        log(f"bench: fixed point after {fixpt['split']['iterations']} iterations: split "
            f"{fixpt['split']['s']:.2f} s, locked {fixpt['locked']['s']:.2f} s "
            f"({fixpt['locked']['computed_fraction']:.3f} of the updates computed), default "
            f"engine (locked from iteration 24) {fixpt['auto']['s']:.2f} s, same bits: "
            f"{fixpt['same_bits']}")
        del res, Tf

    engines = {}
    if not args.no_ablation:
        for name, eng in (("plane", fold3d.ENGINE_PLANE), ("hoisted", fold3d.ENGINE_HOISTED),
                          ("tiled", fold3d.ENGINE_TILED)):
            log(f"bench: FP64 solve, {name} engine")
            ms_e, _, sw_e = time_solves(solver, T0d, n, 1, 0, torch.float64, torch, engine=eng,
                                        iters=iters)
            engines[name] = {"ms_per_step": ms_e, "value": W / (ms_e / 1e3) / 1e9}
        engines["split"] = {"ms_per_step": ms_step, "value": value}

    fp32 = None
    if not args.no_fp32:
        log("bench: FP32 solve")
        ms32, T32, _ = time_solves(solver, T0d, n, args.steps, args.warmup, torch.float32, torch,
                                   iters=iters)
        ref = T.cpu().numpy()
        got = T32.double().cpu().numpy()
        reached = np.abs(ref) < 1e15
        rel = np.abs(got[reached] - ref[reached]) / np.abs(ref[reached])
        fr = {t: float(np.mean(rel > t)) for t in (1e-3, 1e-4, 1e-5)}
        fp32 = {"value": W * args.steps / (ms32 / 1e3) / 1e9, "unit": "Gcell-updates/s",
                "ms_per_step": ms32 / args.steps, "max_rel_err_vs_fp64": float(rel.max()),
                "frac_rel_err_above_1e-3": fr[1e-3], "frac_rel_err_above_1e-4": fr[1e-4],
                "frac_rel_err_above_1e-5": fr[1e-5],
                "signs_and_reached_equal": bool(np.array_equal(np.sign(got), np.sign(ref)) and
                                                np.array_equal(reached, np.abs(got) < 1e15)),
                "stated_bound": FP32_BOUND_TEXT,
                "within_stated_bound": bool(rel.max() < 1e-2 and fr[1e-3] <= 0.005 and
                                            fr[1e-4] <= 0.10 and fr[1e-5] <= 0.25)}
        log(f"bench: FP32 {fp32['value']:.2f} Gcell-upd/s, max rel err {rel.max():.2e}")
        del T32

    # host threads of the CPU baselines (0: none are run)
    cpu_threads = 0 if args.no_cpu else (
        int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1))

    ellip = None
    if not args.no_ellip:
        ellip = {}
        for name, dt, tol in (("fp64", torch.float64, 1e-10), ("fp32", torch.float32, 1e-6)):
            log(f"bench: elliptic blob-speed workload {n}^3 {name}, tol {tol}")
            r, Te = ellip_run(solver, n, dt, tol, torch, threads=cpu_threads,
                              cpu_full=args.cpu_full, groups=args.ellip_groups)
            ellip[name] = r
            if name == "fp64":
                ref64 = Te
            else:
                a = ref64.double()
                b = Te.double()
                reached = a < 1e15
                far = reached & (a > 8 / (n - 1) / 2.3)
                r["max_rel_err_vs_fp64"] = float(((b - a).abs()[far] / a[far]).max())
            log(f"bench: elliptic {name}: {r['iterations']} iterations, {r['value']:.2f} Gcell-upd/s")
            if "cpu_baseline" in r:
                cb = r["cpu_baseline"]
                log(f"bench: elliptic {name} CPU ({cb['cores']} threads; {cb['sample']}): "
                    f"{cb['value']:.3f} Gcell-upd/s, GPU same bits: {cb['gpu_same_bits']}")
            del Te
        del ref64

    homog = None
    ref_small = None
    if not args.no_small:
        log("bench: the reference workload at 128^3 FP64, 8 iterations")
        ref_small = reference_small_run(torch, cpu_threads)
        msg = (f"bench: reference workload 128^3: {ref_small['sweep_ms']:.1f} ms, "
               f"{ref_small['value']:.2f} Gcell-upd/s")
        if "cpu_baseline" in ref_small:
            cb = ref_small["cpu_baseline"]
            msg += (f"; CPU ({cb['cores']} threads) {cb['s']:.2f} s, same bits: "
                    f"{cb['same_bits_as_gpu']}")
        log(msg)
    if not args.no_small:
        log("bench: homogeneous tilted medium 128^3 FP64 vs closed form")
        homog = tilted_homogeneous_run(solver, torch, threads=cpu_threads)
        log(f"bench: tilted homogeneous 128^3: {homog['iterations']} iterations, max abs err {homog['max_abs_err_vs_closed_form']:.3e}")
        if "cpu_baseline" in homog:
            cb = homog["cpu_baseline"]
            log(f"bench: tilted homogeneous CPU ({cb['cores']} threads): {cb['s']:.2f} s, {cb['iterations']} "
                f"iterations, same bits as the GPU: {cb['same_bits_as_gpu']}")

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        ci = iters if args.cpu_full else 1
        log(f"bench: CPU baseline, {threads} threads, {ci} iteration(s) of {n}^3")
        v, dt, T_cpu = cpu_baseline(T0, n, threads, ci)
        # the same iterations on the GPU: the oracle's field is the checker
        T1 = T0d.clone()
        solver.solve(T1, n, max_iters=ci)
        same = bool(np.array_equal(T1.cpu().numpy().view(np.uint64), T_cpu.view(np.uint64)))
        cpu = {"value": v, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
               "sample": f"{ci} of the {iters} iterations (6 sweeps each) of the same {n}^3 "
                         f"FP64 solve, {cell_updates(n, ci):.3e} cell-updates in {dt:.1f} s",
               "gpu_same_bits": same, "host": host_cpu(),
               "threads_note": "the threads one GPU's jobs are granted on the GPU host "
                               "(OMP_NUM_THREADS there), not all of its physical cores"}
        log(f"bench: CPU {v:.3f} Gcell-upd/s; GPU field after those iterations bit-identical: {same}")
        del T1, T_cpu

    batch = None
    if not args.no_batch:
        bt = tuple(int(v) for v in args.batch_tuning.split(",")) if args.batch_tuning else None
        if bt is not None and len(bt) != 3:
            ap.error("--batch-tuning takes three integers: waves,blocks_per_cu,rows_per_pull")
        if args.batch_streams:
            solver.set_batch_streams(args.batch_streams)
        batch = batch_run(solver, torch, tuning=bt)

    adjoint = None
    if not args.no_adjoint:
        log("bench: tomography gradient, 256^3 FP32, 16 sources, 64x64 receivers")
        adjoint = adjoint_run(solver, torch)
        log(f"bench: gradient {adjoint['gradient_s']:.2f} s (forward {adjoint['forward_batched_s']:.2f} s, "
            f"adjoint {adjoint['adjoint_device_ms_total'] / 1e3:.2f} s)")

    adjoint_fd = None
    if not args.no_adjoint and args.fd_cells:
        log(f"bench: tomography gradient vs finite differences, 256^3 FP64, {args.fd_cells} cells "
            "per parameter; vs the CPU adjoint at 64^3")
        adjoint_fd = adjoint_fd_run(solver, torch, cells=args.fd_cells, threads=cpu_threads)
        log(f"bench: adjoint checks took {adjoint_fd['check_s']:.0f} s")

    field = None
    if args.field_n:
        del solver
        torch.cuda.empty_cache()
        log(f"bench: per-node tilt-field workload {args.field_n}^3 FP32")
        field = tilted_field_run(torch, args.field_n)
        log(f"bench: tilted field: {field['iterations']} iterations, {field['value']:.2f} Gcell-upd/s")

    sizes = None
    if args.sizes:
        log(f"bench: the headline solve at sides {args.sizes}")
        sizes = size_sweep(torch, [int(v) for v in args.sizes.split(",")])

    log("bench: measured ceilings of this GPU (FMA rate, device copy)")
    peaks = fold3d.measure_peaks()
    log(f"bench: {peaks}")
    valu = {"measured_fp64_tflops": peaks["fp64_tflops"], "published_fp64_tflops": 78.6,
            "measured_copy_gbs": peaks["copy_gbs"]}
    # VALU view of the same kernel, the bound that binds: wave-instructions per
    # plane launch over the issue rate of 256 CUs x 4 SIMDs at one
    # wave-instruction per 4 cycles.  Counts, launch time and clock come from
    # one profiled run of the first iteration (the counters cannot be taken
    # live); that iteration is the heaviest (most nodes change in it).
    valu_issue = None
    if n == 512 and groups == 2:
        clock_hz = PMC_VALU_CLOCK_GHZ * 1e9
        issue_peak = 256 * 4 * clock_hz / 4
        achieved_i = PMC_VALU_INSTS_512_F64 / (PMC_VALU_LAUNCH_US_512_F64 * 1e-6)
        valu_issue = {"wave_instructions_per_launch": PMC_VALU_INSTS_512_F64,
                      "launch_us": PMC_VALU_LAUNCH_US_512_F64, "achieved_per_s": achieved_i,
                      "peak_per_s": issue_peak, "frac": achieved_i / issue_peak,
                      "clock_ghz": PMC_VALU_CLOCK_GHZ, "iteration": 1,
                      "source": "profiles/fold_split_f64_512_pmc_traffic.md"}

    out = {
        "metric": "Gcell-updates/s at 512^3 (fold-model PMM solve, 8 iterations)",
        "value": value,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / A100_V2_512_GCUPS if n == 512 else None,
        "dtype": "f64",
        "data": "synthetic: the reference's implicit-surface initial field (Fold3dPMM.c)",
        "config": {"workload": f"fold-model PMM solve, {n}^3 nodes, FP64, {iters} iterations x 6 sweeps",
                   "n": n, "iterations": iters, "cell_updates_per_step": W,
                   "engine": "split", "split_groups": groups,
                   "sweep_ms_median": sweep_ms_steps[len(sweep_ms_steps) // 2],
                   "sweep_ms_min": sweep_ms_steps[0], "sweep_ms_max": sweep_ms_steps[-1]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": PMC_TRAFFIC_512_F64 if (n == 512 and groups == 2) else None,
                     "traffic_iteration": 1 if (n == 512 and groups == 2) else None,
                     "traffic_source": "profiles/fold_split_f64_512_pmc_traffic.md",
                     "kernel": "fold_split_kernel<double>", "kernel_us": us,
                     "bytes_per_launch": bytes_per_launch, "valu_issue": valu_issue},
        "engines_fp64": engines,
        "locked_engine_fp64": locked,
        "fixed_point_fp64": fixpt,
        "initializer": init,
        "fp32": fp32,
        "elliptic_blob_speed": ellip,
        "cpu_baseline": cpu,
        "batched": batch,
        "adjoint": adjoint,
        "adjoint_check": adjoint_fd,
        "reference_128": ref_small,
        "tilted_homogeneous_128": homog,
        "tilted_field": field,
        "sizes_fp64": sizes,
        "device_ceilings": valu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
