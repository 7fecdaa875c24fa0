#!/usr/bin/env python3
"""Benchmark of the flux integrator (ssol_solve) on one MI355X.

Workload: the 10,000-heliostat beam-down field (8x8-facet flat heliostats,
128 triangles each, km-scale layout; hyperboloid secondary; 3 mm thin glass
cover over a 64x64-cell receiver; atmosphere 0.1/km; Russian roulette on),
synthetic and seeded. A step is one ssol_solve of `--realisations` paths on
the scene already resident in HBM (the first warm-up solve flattens and
uploads it; later solves reuse it). The metric is Mpaths/s; the time to reach
0.1% relative standard error on absorbed_by_receivers is measured with
RNG-chained batches (4e6 realisations, then batches sized from the merged
standard error) merged like src/test_ssol_solver6.c:219-235.

Prints one JSON line. --gpus N > 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "solstice-solver_amd", "python")]

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per traced segment (FP64 path state):
#  trace kernel: reads the 56 B of path state it needs, writes a 32 B hit record
#  shade kernel: reads path state (96 B) and hit (32 B), writes path state
#                (96 B) and the compacted queue index (8 B)
BYTES_TRACE_FP64 = 56 + 32
BYTES_SHADE_FP64 = 96 + 32 + 96 + 8
BYTES_TRACE_FP32 = 40 + 16
BYTES_SHADE_FP32 = 64 + 16 + 64 + 8

# Work started within a few seconds of closing a context that held tens of
# GB of device memory ran up to 2x slower on MI355X (the kernels' own times
# unchanged, the wall time not); the legs below wait that long between
# contexts so that each one times its own work.
SETTLE_S = 4.0


def settle():
    time.sleep(SETTLE_S)


def merge(a, b):
    """Merge two (E, SE, N) estimates (src/test_ssol_solver6.c:219-235)."""
    (E1, SE1, N1), (E2, SE2, N2) = a, b
    sw = E1 * N1 + E2 * N2
    sw2 = (SE1 * SE1 * N1 + E1 * E1) * N1 + (SE2 * SE2 * N2 + E2 * E2) * N2
    N = N1 + N2
    E = sw / N
    V = max(sw2 / N - E * E, 0.0)
    return E, math.sqrt(V / N), N


def time_to_rel_se(lib, batch, target=1e-3, seed=7, max_paths=2_000_000_000, max_batch=100_000_000):
    """Wall time until absorbed_by_receivers reaches `target` relative
    standard error: a first solve of `batch` realisations, then solves sized
    from the merged estimate (the realisations its SE says are still needed,
    plus 5 %), each continuing the RNG state of the previous one and merged
    like src/test_ssol_solver6.c:219-235."""
    import ssol_ctypes as L
    t0 = time.perf_counter()
    rng = L.Rng(seed, 0)
    acc = None
    state = rng
    keep = []
    n = batch
    while True:
        est = lib.solve(n, rng=state)
        g = est.global_()["absorbed_by_receivers"]
        cur = (g[0], g[2], est.realisation_count)
        acc = cur if acc is None else merge(acc, cur)
        keep.append(est)          # the next RNG state is borrowed from it
        state = est.rng_state
        if len(keep) > 2:
            keep.pop(0)
        if acc[0] > 0 and acc[1] / acc[0] <= target:
            break
        if acc[2] >= max_paths:
            break
        rel = acc[1] / acc[0] if acc[0] > 0 else 1.0
        need = acc[2] * (rel / target) ** 2 - acc[2]
        n = int(min(max(1.05 * need, batch / 4), max_batch))
    return time.perf_counter() - t0, acc


def copy_bandwidth(nbytes=1 << 30, reps=5):
    """GB/s of a plain device copy (read + write bytes), the achievable HBM
    yardstick; one dispatch moves nbytes in and nbytes out."""
    import ctypes as C
    import ssol_ctypes as L
    hip = L.lib_hip()
    ctx = C.c_void_p()
    if hip.ssolh_create(0, C.byref(ctx)):
        raise RuntimeError("ssolh_create failed")
    gbs = C.c_double()
    hip.ssolh_copy_bandwidth.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
    rc = hip.ssolh_copy_bandwidth(ctx, nbytes, reps, C.byref(gbs))
    hip.ssolh_destroy(ctx)
    if rc:
        raise RuntimeError(f"ssolh_copy_bandwidth returned {rc}")
    return gbs.value


def load_traffic(kernel, workload):
    """HBM bytes per segment of `kernel` on `workload` from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, calibrated against a
    plain copy), or None when no such summary exists."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def other_forms(spec, B, ttt_batch, args):
    """The FP32 form and the graph-captured FP64 form on the same field:
    Mpaths/s, time to 0.1% rel. SE on absorbed, and the FP32 estimate's
    difference from an FP64 run of the same size (in combined SE)."""
    import ssol_ctypes as L
    out = {}
    ref = None
    for name, prec, graph in (("fp64_graph", "fp64", 16), ("fp32", "fp32", 0),
                              ("fp32_graph", "fp32", 16)):
        lib = L.Scene(spec, precision=prec, graph_iterations=graph)
        lib.solve(B, seed=200)  # warm-up (flatten, upload, allocation)
        t0 = time.perf_counter()
        est = lib.solve(B, seed=201)
        dt = time.perf_counter() - t0
        g = est.global_()["absorbed_by_receivers"]
        entry = {"value": round(B / dt / 1e6, 3), "unit": "Mpaths/s"}
        if prec == "fp64":
            ref = g
        elif ref is not None:
            entry["absorbed_rel_diff_vs_fp64"] = (g[0] - ref[0]) / ref[0]
            entry["absorbed_diff_in_combined_se"] = (g[0] - ref[0]) / math.hypot(g[2], ref[2])
        if not args.no_ttt:
            t, acc = time_to_rel_se(lib, batch=ttt_batch)
            entry["time_to_0.1pct_rel_se_s"] = round(t, 4)
        out[name] = entry
        print(f"[bench] {name}: {entry}", file=sys.stderr, flush=True)
        del est
        lib.close()
        settle()
    return out


def sweep(n_suns=512, log_min=5.0, log_max=7.0, seq_suns=64, graph_iterations=16):
    """The 100-paraboloid field under n_suns sun positions with seeded
    log-uniform realisation counts: one batched solve driven by the host
    loop, the same batch with graph_iterations wavefront iterations per HIP
    graph launch and, for comparison, the first seq_suns suns as separate
    ssol_solve calls on the same scene through the public API (each sun
    change re-flattens the scene, as any caller's per-sun solve would)."""
    import ctypes as C
    import ssol_ctypes as L
    import ssol_scene as S
    spec, dirs, counts = S.config4(n_suns=n_suns, log_min=log_min, log_max=log_max)
    n_total = int(sum(int(c) for c in counts))
    t_batch = {}
    for graph in (0, graph_iterations):
        lib = L.Scene(spec, graph_iterations=graph)
        lib.solve_sun_batch(dirs[:4], [10000] * 4, seed=1)  # warm-up
        t0 = time.perf_counter()
        res, got = lib.solve_sun_batch(dirs, counts, seed=2)
        t_batch[graph] = time.perf_counter() - t0
        lib.close()
        settle()
    # the widths and graph lengths picked by the autotuner on a probe batch
    # (every 8th sun, a sixteenth of its realisations), then the sweep
    lib = L.Scene(spec)
    tuned = lib.autotune_batch(dirs[::8], [max(1000, int(c) // 16) for c in counts[::8]])
    lib.solve_sun_batch(dirs[:4], [10000] * 4, seed=1)  # warm-up
    t0 = time.perf_counter()
    lib.solve_sun_batch(dirs, counts, seed=2)
    t_tuned = time.perf_counter() - t0
    lib.close()
    settle()
    # sequential: one solve per sun position on one scene
    sl = L.Scene(spec)
    sl.solve(10000, seed=1)  # warm-up
    t0 = time.perf_counter()
    n_seq = 0
    for j in range(seq_suns):
        L.call("ssol_sun_set_direction", sl.sun, (C.c_double * 3)(*dirs[j]))
        sl.solve(int(counts[j]), seed=3)
        n_seq += int(counts[j])
    t_seq = time.perf_counter() - t0
    sl.close()
    settle()
    rate = {g: n_total / t / 1e6 for g, t in t_batch.items()}
    seq_rate = n_seq / t_seq / 1e6
    out = {"suns": n_suns, "paths": n_total,
           "batched_s": round(t_batch[0], 4), "batched_Mpaths_s": round(rate[0], 2),
           "graph_iterations": graph_iterations,
           "batched_graph_s": round(t_batch[graph_iterations], 4),
           "batched_graph_Mpaths_s": round(rate[graph_iterations], 2),
           "sequential_sample_suns": seq_suns, "sequential_sample_paths": n_seq,
           "sequential_sample_s": round(t_seq, 4),
           "sequential_Mpaths_s": round(seq_rate, 2),
           "speedup_vs_sequential": round(rate[0] / seq_rate, 2),
           "graph_speedup_vs_sequential": round(rate[graph_iterations] / seq_rate, 2),
           "autotuned": {"path_slots": tuned.path_slots,
                         "graph_iterations": tuned.graph_iterations,
                         "candidates": tuned.candidates, "batched_s": round(t_tuned, 4),
                         "batched_Mpaths_s": round(n_total / t_tuned / 1e6, 2)}}
    print(f"[bench] sweep: {out}", file=sys.stderr, flush=True)
    return out


def adjoint_timing(n=20_000_000):
    """The 1,000-heliostat differentiable-aiming field: forward solve and
    adjoint (aim angles and slope errors of every heliostat) of the same
    realisations."""
    import ssol_ctypes as L
    import ssol_scene as S
    spec, aim, slope, axes = S.config5()
    lib = L.Scene(spec)
    lib.solve(1_000_000, seed=1)
    t0 = time.perf_counter()
    est = lib.solve(n, seed=2)
    t_fwd = time.perf_counter() - t0
    lib.solve_adjoint(1_000_000, spec.instances[:-1], axes, d_absorbed=1.0, seed=1)
    t0 = time.perf_counter()
    g = lib.solve_adjoint(n, spec.instances[:-1], axes, d_absorbed=1.0, seed=2)
    t_adj = time.perf_counter() - t0

    # one aiming step (new aim angles for every heliostat, forward and
    # adjoint): moving the instances of the resident scene, which rebuilds
    # only instance frames and the TLAS, against building the scene anew
    aim2 = aim + 1e-4
    spec2 = S.config2(0, n_helio=len(aim), radius=900.0, aim_perturb=aim2, slope_errors=slope)
    axes2 = S.aim_axes(0, len(aim), 900.0, None, aim2)
    steps = {}
    for kind in ("move", "rebuild"):
        t0 = time.perf_counter()
        if kind == "move":
            for a, b in zip(spec.instances[:-1], spec2.instances[:-1]):
                lib.set_transform(a, b.transform)
            sc, helios = lib, spec.instances[:-1]
        else:
            sc, helios = L.Scene(spec2), spec2.instances[:-1]
        sc.solve(n, seed=3)
        sc.solve_adjoint(n, helios, axes2, d_absorbed=1.0, seed=3)
        steps[kind] = time.perf_counter() - t0
        if kind == "rebuild":
            sc.close()
    lib.close()
    settle()
    out = {"heliostats": len(spec.instances) - 1, "paths": n, "forward_s": round(t_fwd, 4),
           "adjoint_s": round(t_adj, 4), "adjoint_Mpaths_s": round(n / t_adj / 1e6, 2),
           "grad_norm": float(np.linalg.norm(g)),
           "aim_step_s_moving_instances": round(steps["move"], 4),
           "aim_step_s_rebuilding_scene": round(steps["rebuild"], 4)}
    print(f"[bench] adjoint: {out}", file=sys.stderr, flush=True)
    return out


def host_cores():
    """Cores this process may use: the CPU affinity set, bounded by the
    cgroup CPU quota when one is set (a container's share of the host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) / int(period)))
    except (OSError, ValueError):
        pass
    return (min(n, quota) if quota else n), n, quota


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spec, seconds: float, threads: int):
    """The oracle (FP64 OpenMP restatement of the reference solver) on the
    host cores, on a bounded sample of the same workload: solves of 1e6
    realisations and more (each solve has a fixed cost of thread start-up and
    per-thread tally merge, which small solves would charge to the CPU), the
    last one sized to end near `seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ctypes as O
    o = O.Oracle(spec, track_sampled_x_receiver=False)
    o.solve(20_000, seed=99, nthreads=threads)  # warm-up (page-in, thread pool)
    n = 0
    k = 0
    chunk = 1_000_000
    t0 = time.perf_counter()
    while True:
        o.solve(chunk, seed=1000 + k, nthreads=threads)
        n += chunk
        k += 1
        dt = time.perf_counter() - t0
        left = seconds - dt
        if left <= 0.05 * seconds:
            break
        chunk = max(1_000_000, int(n / dt * left))
    dt = time.perf_counter() - t0
    return n / dt / 1e6, n, dt, o


def cpu_time_to_rel_se(o, batch, threads, target=1e-3, seed=7, max_seconds=60.0):
    """The progressive recipe of time_to_rel_se on the oracle: a first solve
    of `batch` realisations, then solves sized from the merged estimate,
    each continuing the RNG counter of the previous one, merged like
    src/test_ssol_solver6.c:219-235. Stops at the target or after
    max_seconds (then returns the time so far and the estimate reached)."""
    t0 = time.perf_counter()
    acc = None
    counter = 0
    n = batch
    while True:
        o.solve(n, seed=seed, counter=counter, nthreads=threads)
        counter += n
        g = o.global_()["absorbed_by_receivers"]
        cur = (g[0], g[2], o.realisation_count)
        acc = cur if acc is None else merge(acc, cur)
        if acc[0] > 0 and acc[1] / acc[0] <= target:
            return time.perf_counter() - t0, acc, True
        if time.perf_counter() - t0 > max_seconds:
            return time.perf_counter() - t0, acc, False
        rel = acc[1] / acc[0] if acc[0] > 0 else 1.0
        need = acc[2] * (rel / target) ** 2 - acc[2]
        n = int(max(1.05 * need, batch / 4))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--realisations", type=int, default=100_000_000)
    ap.add_argument("--config", choices=("c1", "c2", "c3"), default="c3")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    ap.add_argument("--path-slots", type=int, default=0)
    ap.add_argument("--graph", type=int, default=0,
                    help="wavefront iterations per HIP graph launch (0: host-driven)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-ttt-seconds", type=float, default=40.0,
                    help="bound on the CPU time-to-0.1%% run (0: skip it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-copy", action="store_true", help="skip the copy-bandwidth yardstick")
    ap.add_argument("--no-ttt", action="store_true", help="skip the time-to-0.1%% run")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the FP32 / graph-captured forms, the sweep and the adjoint")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py: the flux integrator runs on one GPU; --gpus {args.gpus} is not "
              "supported (across GPUs the realisations would shard by id with one "
              "all-reduce of the folded tallies).", file=sys.stderr)
        return 2

    import ssol_scene as S
    import ssol_ctypes as L

    spec = {"c1": S.config1, "c2": S.config2, "c3": S.config3}[args.config]()
    workload = {"c1": "config1_flat_mirror", "c2": "config2_100_paraboloids",
                "c3": "config3_10k_heliostat_beam_down"}[args.config]
    print(f"[bench] building {spec.name} ({len(spec.instances)} instances)", file=sys.stderr,
          flush=True)
    lib = L.Scene(spec, precision=args.precision, path_slots=args.path_slots,
                  graph_iterations=args.graph)
    B = args.realisations

    # the first solve is cold: it flattens the scene (BVHs, sampling tables),
    # creates the GPU context, uploads the scene and allocates the wavefront
    cold_s = None
    for w in range(args.warmup):
        t = time.perf_counter()
        est = lib.solve(B, seed=100 + w)
        if w == 0:
            cold_s = time.perf_counter() - t
        print(f"[bench] warmup {w}: {time.perf_counter() - t:.3f} s", file=sys.stderr, flush=True)

    times, segs, trace_ms, shade_ms, iters = [], 0, 0.0, 0.0, 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        t = time.perf_counter()
        est = lib.solve(B, seed=k)
        times.append(time.perf_counter() - t)
        st = est.stats
        segs += st.segments
        trace_ms += st.trace_ms
        shade_ms += st.shade_ms
        iters += st.iterations
        print(f"[bench] step {k}: {times[-1]:.3f} s, {B / times[-1] / 1e6:.1f} Mpaths/s, "
              f"{st.segments / B:.3f} segments/path, trace {st.trace_ms:.1f} ms, "
              f"shade {st.shade_ms:.1f} ms, {st.iterations} iterations", file=sys.stderr, flush=True)
    total = time.perf_counter() - t0
    value = args.steps * B / total / 1e6
    g = est.global_()

    # roofline of the dominant kernel, from in-stream HIP events (the
    # host-driven loop records them; a graph-captured run has none)
    fp32 = args.precision == "fp32"
    if trace_ms <= 0 and shade_ms <= 0:
        print("bench.py: no per-kernel times (graph mode); run with --graph 0 for the roofline",
              file=sys.stderr)
        trace_ms = shade_ms = float("nan")
    if not (shade_ms > trace_ms):
        kernel, ms, per_seg = "k_trace", trace_ms, (BYTES_TRACE_FP32 if fp32 else BYTES_TRACE_FP64)
    else:
        kernel, ms, per_seg = "k_shade", shade_ms, (BYTES_SHADE_FP32 if fp32 else BYTES_SHADE_FP64)
    achieved = per_seg * segs / (ms * 1e-3) / 1e9
    seg_per_launch = segs / iters
    if fp32:
        kernel += "32"   # the FP32 form's kernels: k_trace32, k_shade32
    tr = load_traffic(kernel, workload + ("_fp32" if fp32 else ""))
    traffic = None if tr is None else round(tr["hbm_bytes_per_segment"] * seg_per_launch)
    roofline = {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": traffic, "algorithmic_bytes_per_launch": round(per_seg * seg_per_launch),
                "bytes_per_segment": per_seg, "avg_launch_ms": round(ms / iters, 4),
                "segments_per_path": round(segs / (B * args.steps), 4),
                "trace_ms_per_step": round(trace_ms / args.steps, 3),
                "shade_ms_per_step": round(shade_ms / args.steps, 3)}
    if not args.no_copy:
        roofline["copy_gbs"] = round(copy_bandwidth(), 1)

    out = {
        "metric": "Mpaths/s + time to 0.1% rel. std-err on absorbed flux, 10k-heliostat field",
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if not fp32 else "f32",
        "data": "synthetic",
        "config": {"workload": workload, "realisations_per_step": B,
                   "instances": len(spec.instances), "precision": args.precision},
        "roofline": roofline,
        "absorbed_by_receivers": {"E": g["absorbed_by_receivers"][0],
                                  "SE": g["absorbed_by_receivers"][2]},
        "cold_first_solve_s": None if cold_s is None else round(cold_s, 4),
        "warm_solve_s": round(total / args.steps, 4),
    }

    if not args.no_ttt:
        dt, acc = time_to_rel_se(lib, batch=min(B, 4_000_000))
        out["time_to_0.1pct_rel_se_s"] = round(dt, 4)
        out["paths_to_0.1pct_rel_se"] = acc[2]
        print(f"[bench] 0.1% rel. SE on absorbed after {acc[2]} paths in {dt:.3f} s",
              file=sys.stderr, flush=True)

    # the default wavefront holds every realisation of a step in device
    # memory: free it before the other forms allocate theirs
    del est
    lib.close()

    if not args.no_forms and args.config == "c3":
        settle()
        out["forms"] = other_forms(spec, B, min(B, 4_000_000), args)
        out["sweep_512_suns"] = sweep()
        out["adjoint_1000_heliostats"] = adjoint_timing()

    if not args.no_cpu:
        threads, affinity, quota = host_cores()
        mps, n, dt, o = cpu_baseline(spec, args.cpu_seconds, threads)
        out["cpu_baseline"] = {"value": round(mps, 4), "unit": "Mpaths/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{n} paths of the same workload in {dt:.1f} s "
                                         "(oracle: FP64 OpenMP restatement of the reference "
                                         "solver with its own SAH BVH)",
                               "cpu_model": cpu_model(), "cpus_in_affinity_set": affinity,
                               "cgroup_cpu_quota": quota}
        if "time_to_0.1pct_rel_se_s" in out and args.cpu_ttt_seconds > 0:
            t, acc, reached = cpu_time_to_rel_se(o, 1_000_000, threads,
                                                 max_seconds=args.cpu_ttt_seconds)
            key = "time_to_0.1pct_rel_se_s" if reached else "time_to_0.1pct_rel_se_s_lower_bound"
            out["cpu_baseline"][key] = round(t, 2)
            out["cpu_baseline"]["paths_to_0.1pct_rel_se"] = acc[2]
            out["cpu_baseline"]["rel_se_reached"] = acc[1] / acc[0]
        print(f"[bench] CPU baseline: {mps:.4f} Mpaths/s on {threads} threads "
              f"({affinity} CPUs in the affinity set, quota {quota})", file=sys.stderr, flush=True)

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
