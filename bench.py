#!/usr/bin/env python3
"""Benchmark of the filament velocity + timestep path on one MI355X.

Headline workload (the batched ensemble): 256 independent periodic vortex
tangles, point counts log-uniform in [1k, 40k] (2.64 M points in all), each
a seeded random-ring tangle (radii U[0.01, 0.03] cm, centres uniform in a
periodic 0.1 cm box, orientations uniform on the sphere, delta = 5e-4 cm,
kappa = 9.97e-4, core 1.3e-8 cm, dt = 5e-6 s), seed 1238.  Synthetic data,
generated here.  Direct Biot-Savart over straight segments with the 26
periodic images, FP32 Kahan-compensated far field and FP64 near field,
reconnections and re-meshing on, 3rd-order Adams-Bashforth.

A "step" is one pass of the whole reference main-loop body for every tangle
of the ensemble on the device: ghosts -> velocity (one persistent direct-sum
launch for all 256 tangles, local term, mutual friction) -> AB update ->
point insertion -> brute close-point search -> reconnection -> point
removal -> periodic wrap, recorded once as a HIP graph and replayed every
step.  ``value`` = point-updates/s = sum over the timed steps of the points
moved, divided by the wall time of those steps (state resident in HBM).

Also reported: the same ensemble step with the plain FP64 kernel (the >= 3x
target), the FP32C velocity error against the FP64 kernel (every point) and
against the FP64 oracle (sampled targets of several tangles), the roofline
of the direct-sum kernel on the ensemble launch, the oracle's FP64 CPU
velocity on the sampled targets as the CPU baseline, and two more workloads:
64 rings (16.7k points, one tangle) and the 1M-point counterflow octree step.
"""
from __future__ import annotations

import argparse
import ctypes as C
import dataclasses
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "vortex point-updates/s (points×AB steps) + max rel. velocity err vs FP64"
FLOP_PER_PAIR = 43          # reference arithmetic per (target, segment, image) evaluation
PEAK_FP32 = 157.3           # TFLOP/s, MI355X vector FP32
PEAK_FP64 = 78.6            # TFLOP/s, MI355X vector FP64


# one autotuner cache per bench run: the first solver of a workload times
# the direct-sum variants, every later solver of the same size class reads
# the choice back, so all timings of a workload use the same variant
TUNE_CACHE = os.path.join(tempfile.gettempdir(), "nuvort_bench_tune_%d.txt" % os.getpid())


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_err_stats(u_ref: np.ndarray, u: np.ndarray) -> dict:
    """Per-point e_i = |u_i - u_ref_i| / |u_ref_i| over points with
    |u_ref_i| >= 1e-3 max|u_ref|; normwise = max |du| / max |u_ref|."""
    mag = np.linalg.norm(u_ref, axis=1)
    err = np.linalg.norm(u - u_ref, axis=1)
    sel = mag >= 1e-3 * mag.max()
    e = err[sel] / mag[sel]
    return {"max_rel": float(e.max()), "rms_rel": float(np.sqrt(np.mean(e ** 2))),
            "normwise_max": float(err.max() / mag.max()), "points": int(sel.sum())}


def workload():
    from nuvort_amd import tangles
    from nuvort_amd.runfile import RunParams
    st = tangles.random_rings(1236, 64, 0.01, 0.03, 0.1, 5e-4)
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.1, 0.1, 0.1), velocity="BS",
                  hip_precision="fp32c", recon_shots=1, overide_timestep_check=False)
    return st, p


def timed_steps(p, st, steps, warmup, blocks=1):
    """Warm up, then time `steps` steps as `blocks` consecutive nv_step calls
    (wall clock around each call; the library synchronises before it
    returns).  Returns totals and the per-block ms/step."""
    from nuvort_amd.engine import FilamentSolver
    s = FilamentSolver(p)
    s.set_state(st)
    tune = s.autotune(TUNE_CACHE)
    if warmup:
        s.step(warmup)
    s0 = s.stats()
    sizes = [steps // blocks + (1 if b < steps % blocks else 0) for b in range(blocks)]
    per_block = []
    t0 = time.perf_counter()
    for k in sizes:
        if k == 0:
            continue
        tb = time.perf_counter()
        s.step(k)
        per_block.append((time.perf_counter() - tb) / k * 1e3)
    t1 = time.perf_counter()
    s1 = s.stats()
    out = {"seconds": t1 - t0, "point_updates": s1["point_updates"] - s0["point_updates"],
           "block_ms_per_step": per_block,
           "pair_ms": s1["pair_ms_total"] - s0["pair_ms_total"],
           "pair_launches": s1["pair_launches"] - s0["pair_launches"],
           "pair_evals": s1["pair_evals_total"] - s0["pair_evals_total"],
           "walk_ms": s1["walk_ms_total"] - s0["walk_ms_total"],
           "walk_launches": s1["walk_launches"] - s0["walk_launches"],
           "walk_nodes": s1["walk_nodes_total"] - s0["walk_nodes_total"],
           "final_points": s.get_state().active, "tuning": tune}
    s.close()
    return out


def kernel_timing(p, st, reps):
    """Average duration of the direct-sum launch: `reps` velocity passes on
    the workload's state, each launch bracketed by HIP events on the
    library's stream (the recorded step carries no events)."""
    from nuvort_amd.engine import FilamentSolver
    s = FilamentSolver(p)
    s.set_state(st)
    s.autotune(TUNE_CACHE)
    s.ghostp()
    s.velocity()
    s0 = s.stats()
    for _ in range(reps):
        s.velocity()
    s1 = s.stats()
    s.close()
    launches = s1["pair_launches"] - s0["pair_launches"]
    return ((s1["pair_ms_total"] - s0["pair_ms_total"]) / launches * 1e-3,
            (s1["pair_evals_total"] - s0["pair_evals_total"]) / launches, launches)


def velocity_once(p, st, precision, tune=True):
    from nuvort_amd.engine import FilamentSolver
    q = p.__class__(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    q.hip_precision = precision
    s = FilamentSolver(q)
    s.set_state(st)
    if tune:
        s.autotune(TUNE_CACHE)
    s.ghostp()
    s.velocity()
    g = s.get_state()
    s.close()
    return g


def cpu_baseline_parallel(st, p, sample: int, threads: int):
    """The oracle's FP64 velocity (the reference's arithmetic, one OpenMP
    loop over targets) for the first `sample` points against all sources and
    the 27 images; returns seconds, the rows and the velocities."""
    from oracle import port
    L = port.lib()
    ps = port.PortState(st)
    pr = port.params_struct(p, threads)
    L.nvo_ghostp(ps.ref(), C.byref(pr))
    first = 1
    u = np.zeros((sample, 3))
    L.nvo_velocity(ps.ref(), C.byref(pr), None, first, 8, u.ctypes.data_as(port._dp))
    t0 = time.perf_counter()
    L.nvo_velocity(ps.ref(), C.byref(pr), None, first, sample, u.ctypes.data_as(port._dp))
    t1 = time.perf_counter()
    return t1 - t0, np.arange(sample), u


def dense(args):
    """800 rings, radii U[0.005, 0.04] cm, delta 2.5e-4 cm, periodic 0.1 cm
    box (seed 1237): 466,977 points.  Relaxed 50 steps with the FP32C step
    (dt 1.9e-6 s, below the 1.95e-6 s bound of src/initial.f90:169-180,
    reconnection and re-meshing every step), then one velocity pass each in
    FP32C and FP64 on the relaxed state (every point), and the oracle's FP64
    velocity on a sample of targets."""
    from nuvort_amd import tangles
    from nuvort_amd.engine import FilamentSolver
    from nuvort_amd.runfile import RunParams
    st = tangles.random_rings(1237, 800, 0.005, 0.04, 0.1, 2.5e-4)
    p = RunParams(dt=1.9e-6, delta=2.5e-4, corea=1.3e-8, box_size=(0.1, 0.1, 0.1), velocity="BS",
                  hip_precision="fp32c", recon_shots=1)
    log("dense: %d points" % st.active)
    s = FilamentSolver(p)
    s.set_state(st)
    tune = s.autotune(TUNE_CACHE)
    log("tuned: %s" % tune)
    blocks, t_all, pu_all = [], 0.0, 0
    for b in range(5):
        s0 = s.stats()
        t0 = time.perf_counter()
        s.step(10)
        dtb = time.perf_counter() - t0
        pu = s.stats()["point_updates"] - s0["point_updates"]
        blocks.append(dtb / 10 * 1e3)
        t_all += dtb
        pu_all += pu
        log("relax steps %d-%d: %.1f ms/step" % (10 * b + 1, 10 * b + 10, dtb / 10 * 1e3))
    relaxed = s.get_state()
    ev = len(s.recon_events())
    s.close()
    n = relaxed.active
    g32 = velocity_once(p, relaxed, "fp32c")
    log("fp32c velocity done")
    log("fp64 velocity (about 30 s at this size)")
    g64 = velocity_once(p, relaxed, "fp64", tune=False)
    log("fp64 velocity done")
    act = g64.infront != 0
    acc = {"fp32c_vs_fp64_gpu": rel_err_stats(g64.u_sup[act], g32.u_sup[act])}
    t_k, e_k, _ = kernel_timing(p, relaxed, 3)
    sample = 4096
    log("oracle sample: %d targets on %d threads" % (sample, args.cpu_threads))
    secs, idx, ucpu = cpu_baseline_parallel(relaxed, p, sample, args.cpu_threads)
    acc["fp32c_vs_fp64_oracle_sample"] = rel_err_stats(ucpu, g32.u[idx])
    acc["fp64_gpu_vs_fp64_oracle_sample"] = rel_err_stats(ucpu, g64.u[idx])
    line = {
        "metric": METRIC, "value": pu_all / t_all, "unit": "point-updates/s", "n_gpus": 1, "steps": 50,
        "warmup": 0, "ms_per_step": t_all / 50 * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32c+f64", "data": "synthetic",
        "config": {"workload": "dense tangle: 800 random rings, radii U[0.005,0.04] cm, periodic 0.1 cm box, "
                               "delta 2.5e-4 cm, direct Biot-Savart with 26 images, FP32C, relaxed 50 steps",
                   "points_initial": st.active, "points_relaxed": n, "seed": 1237, "dt_s": 1.9e-6,
                   "reconnections_during_relaxation": ev, "parallelism": "1 GPU"},
        "block_ms_per_step": blocks, "tuning": tune,
        "accuracy": acc, "max_rel_velocity_err_vs_fp64": acc["fp32c_vs_fp64_gpu"]["max_rel"],
        "kernel": {"name": "k_bs32c", "ms": t_k * 1e3, "pair_evals": int(e_k),
                   "achieved_tflops": FLOP_PER_PAIR * e_k / t_k / 1e12,
                   "frac": FLOP_PER_PAIR * e_k / t_k / 1e12 / PEAK_FP32},
        "oracle_sample": {"targets": sample, "seconds": secs, "threads": args.cpu_threads},
    }
    print(json.dumps(line), flush=True)
    if os.path.exists(TUNE_CACHE):
        os.remove(TUNE_CACHE)


def oracle_tree_velocity(st, p, ranges, threads):
    """The oracle's tree velocity (reference tree build and walk) for the
    rows in `ranges` [(first0, count), ...] (0-based first rows)."""
    from oracle import port
    L = port.lib()
    ps = port.PortState(st)
    pr = port.params_struct(p, threads)
    L.nvo_ghostp(ps.ref(), C.byref(pr))
    tr = port.PortTree()
    rc = L.nvo_tree_build(ps.ref(), C.byref(pr), tr.ref())
    if rc != 0:
        raise RuntimeError("oracle tree build failed: %d" % rc)
    out = []
    for first0, count in ranges:
        u = np.zeros((count, 3))
        L.nvo_velocity(ps.ref(), C.byref(pr), tr.ref(), first0 + 1, count, u.ctypes.data_as(port._dp))
        out.append(u)
    tr.free()
    return np.concatenate(out)


def tree_phases(p, st, steps):
    """Octree steps launched eagerly from state st: the wall time per step and
    the build, velocity walk and close search durations (HIP events), which
    recorded steps do not time by phase."""
    from nuvort_amd.engine import FilamentSolver
    q = p.__class__(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    q.hip_step_launch = "eager"
    s = FilamentSolver(q)
    s.set_state(st)
    s.step(1)
    out = []
    for _ in range(steps):
        a = s.stats()
        t0 = time.perf_counter()
        s.step(1)
        dt_ = time.perf_counter() - t0
        b = s.stats()
        out.append({"step_ms": dt_ * 1e3, "tree_build_ms": b["ms_tree_build"], "close_ms": b["ms_close"],
                    "walk_ms": b["walk_ms_total"] - a["walk_ms_total"],
                    "walk_nodes": b["walk_nodes_total"] - a["walk_nodes_total"]})
    s.close()
    return out


def tree1m(args):
    """Random rings (radii U[0.01, 0.03] cm, delta 5e-4 cm, seed 1239) until
    1M points in a periodic 0.2 cm box; Barnes-Hut theta 0.4 (octree rebuilt
    every step), uniform normal flow 0.5 cm/s along x with mutual friction
    alpha = 0.206, alpha' = 0.0083, dt 5e-6 s, reconnection and re-meshing
    every step (stale-tree close search).  Velocities of a sample of targets
    are checked against the oracle's direct FP64 sum (the tree's own
    approximation error) and, for the FP32 walk, against the FP64 walk."""
    from nuvort_amd import tangles
    from nuvort_amd.engine import FilamentSolver
    from nuvort_amd.runfile import RunParams
    st = tangles.random_rings(1239, 0, 0.01, 0.03, 0.2, 5e-4, target_points=1000000)
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.2, 0.2, 0.2), velocity="Tree", tree_theta=0.4,
                  normal_velocity="xflow", norm_vel_xflow=0.5, alpha=(0.206, 0.0083), recon_shots=1,
                  hip_precision=args.tree_precision)
    log("tree1m: %d points, walk %s" % (st.active, args.tree_precision))
    s = FilamentSolver(p)
    s.set_state(st)
    s.step(1)
    log("warm-up step done")
    per = []
    s0 = s.stats()
    t_all = 0.0
    for k in range(args.steps if args.steps < 100 else 5):
        t0 = time.perf_counter()
        s.step(1)
        dt_ = time.perf_counter() - t0
        t_all += dt_
        per.append(dt_ * 1e3)
        log("step: %.1f ms" % (dt_ * 1e3))
    s1 = s.stats()
    nsteps = len(per)
    pu = s1["point_updates"] - s0["point_updates"]
    launch = {"recorded_steps": s1["tree_replays"] - s0["tree_replays"],
              "redone_eagerly": s1["tree_redone"] - s0["tree_redone"]}
    cur = s.get_state()
    s.close()
    stats = tree_phases(p, cur, 3)
    for x in stats:
        log("eager step: %.1f ms (build %.1f, walk %.1f, close %.1f)" % (x["step_ms"], x["tree_build_ms"],
                                                                         x["walk_ms"], x["close_ms"]))
    # velocities on the current state: this walk, the FP64 walk, and the oracle's direct sum on a sample
    def tree_velocity(prec):
        q = p.__class__(**{k: getattr(p, k) for k in p.__dataclass_fields__})
        q.hip_precision = prec
        v = FilamentSolver(q)
        v.set_state(cur)
        v.ghostp()
        v.construct_tree()
        v.velocity()
        g = v.get_state()
        v.close()
        return g
    g = tree_velocity(args.tree_precision)
    g64 = tree_velocity("fp64") if args.tree_precision != "fp64" else g
    acc = {}
    if args.tree_precision != "fp64":
        act = g64.infront != 0
        acc["walk_vs_fp64_walk"] = rel_err_stats(g64.u_sup[act], g.u_sup[act])
    # the oracle's tree (the reference's algorithm) on 8 blocks of 512 rows
    n_all = cur.pcount
    ranges = [(int(b * (n_all - 512) / 7), 512) for b in range(8)]
    rows = np.concatenate([np.arange(a, a + c) for a, c in ranges])
    log("oracle tree velocity on %d targets" % len(rows))
    uot = oracle_tree_velocity(cur, p, ranges, args.cpu_threads)
    keep = cur.infront[rows] != 0
    acc["fp64_walk_vs_oracle_tree_sample"] = rel_err_stats(uot[keep], g64.u[rows][keep])
    if args.tree_precision != "fp64":
        acc["walk_vs_oracle_tree_sample"] = rel_err_stats(uot[keep], g.u[rows][keep])
    sample = 1024
    log("oracle direct sum on %d targets" % sample)
    pd = p.__class__(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    pd.velocity, pd.tree_theta = "BS", 0.0
    secs, idx, ucpu = cpu_baseline_parallel(cur, pd, sample, args.cpu_threads)
    # the tree's own approximation error (theta 0.4, the reference's monopole
    # nodes) against the direct sum
    acc["tree_vs_direct_fp64_oracle_sample"] = rel_err_stats(ucpu, g.u[idx])
    walk_nodes = np.mean([x["walk_nodes"] for x in stats])
    walk_ms = np.mean([x["walk_ms"] for x in stats])
    n = st.active
    rec = 48 if args.tree_precision == "fp32c" else 80
    # node records read, plus per target and image its FP64 position read
    # and its 3-double partial written
    walk_bytes = walk_nodes * rec + n * 27 * 48
    line = {
        "metric": METRIC, "value": pu / t_all, "unit": "point-updates/s", "n_gpus": 1, "steps": nsteps,
        "warmup": 1, "ms_per_step": t_all / nsteps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32c+f64" if args.tree_precision == "fp32c" else "f64", "data": "synthetic",
        "config": {"workload": "counterflow tangle: random rings to 1M points in a periodic 0.2 cm box, delta "
                               "5e-4 cm, Barnes-Hut theta 0.4 rebuilt every step, v_n = 0.5 cm/s along x, "
                               "alpha 0.206 / 0.0083, reconnection and re-meshing every step",
                   "points": n, "seed": 1239, "dt_s": 5e-6, "walk": args.tree_precision, "parallelism": "1 GPU"},
        "step_ms": per, "steps_launched": launch, "eager_steps_by_phase": stats, "accuracy": acc,
        "walk": {"ms": walk_ms, "node_reads": walk_nodes, "record_bytes": rec,
                 "algorithmic_GBps": walk_bytes / (walk_ms * 1e-3) / 1e9, "hbm_peak_GBps": 8000.0},
        "oracle_sample": {"targets": sample, "seconds": secs, "threads": args.cpu_threads},
    }
    print(json.dumps(line), flush=True)


def cpu_info(threads: int) -> dict:
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "threads_used": threads}


def calibration() -> dict | None:
    """The reference build's time over the restatement's on one input,
    measured in the development container (oracle/calibrate.py)."""
    path = os.path.join(ROOT, "oracle", "calibration.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def oracle_sample(sts, p, picks, per_tangle: int, threads: int):
    """The oracle's FP64 velocity (the reference's arithmetic: local term,
    direct sum over the 27 images) of `per_tangle` targets of each tangle in
    `picks`, spread over its rows.  Returns the seconds of the velocity
    calls, {tangle: (rows, u)} and the evaluated (target, source) count."""
    from oracle import port
    L = port.lib()
    out, secs, pairs = {}, 0.0, 0
    for k in picks:
        st = sts[k]
        ps = port.PortState(st)
        pr = port.params_struct(p, threads)
        L.nvo_ghostp(ps.ref(), C.byref(pr))
        n = st.pcount
        m = min(per_tangle, n)
        blocks = 4
        starts = [int(b * (n - m // blocks) / blocks) for b in range(blocks)]
        rows, us = [], []
        for a in starts:
            cnt = min(m // blocks, n - a)
            u = np.zeros((cnt, 3))
            t0 = time.perf_counter()
            L.nvo_velocity(ps.ref(), C.byref(pr), None, a + 1, cnt, u.ctypes.data_as(port._dp))
            secs += time.perf_counter() - t0
            rows.append(np.arange(a, a + cnt))
            us.append(u)
            pairs += cnt * st.active
        out[k] = (np.concatenate(rows), np.concatenate(us))
    return secs, out, pairs


def ensemble_workload():
    from nuvort_amd import tangles
    from nuvort_amd.runfile import RunParams
    sts = tangles.ensemble(1238, 256, 1000, 40000)
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.1, 0.1, 0.1), velocity="BS",
                  hip_precision="fp32c", recon_shots=1)
    return sts, p


def ensemble_steps(p, sts, precision, warmup, steps):
    """Warm-up steps, then `steps` timed steps of the whole batch (wall
    clock around the nv_step calls, each of which synchronises before it
    returns), split into up to 5 calls for the spread; and the run's
    reconnections as (step, tangle, i, j)."""
    from nuvort_amd.engine import FilamentSolver
    q = dataclasses.replace(p, hip_precision=precision)
    s = FilamentSolver(q)
    s.set_state_batch(sts)
    tune = s.autotune(TUNE_CACHE) if precision == "fp32c" else None
    if warmup:
        s.step(warmup)
    blocks = min(5, steps)
    sizes = [steps // blocks + (1 if b < steps % blocks else 0) for b in range(blocks)]
    s0 = s.stats()
    per = []
    t0 = time.perf_counter()
    for k in sizes:
        tb = time.perf_counter()
        s.step(k)
        per.append((time.perf_counter() - tb) / k * 1e3)
    sec = time.perf_counter() - t0
    s1 = s.stats()
    events = [(e[0], e[7], e[1], e[2]) for e in s.recon_events()]
    s.close()
    return {"seconds": sec, "point_updates": s1["point_updates"] - s0["point_updates"], "block_ms_per_step": per,
            "tuning": tune, "events": events}


def ensemble_velocity(p, sts, precision, reps):
    """One velocity pass of the batch (every point), then `reps` more whose
    direct-sum launches are timed with HIP events on the library's stream."""
    from nuvort_amd.engine import FilamentSolver
    q = dataclasses.replace(p, hip_precision=precision)
    s = FilamentSolver(q)
    s.set_state_batch(sts)
    if precision == "fp32c":
        s.autotune(TUNE_CACHE)
    s.ghostp()
    s.velocity()
    gs = s.get_state_batch()
    s0 = s.stats()
    for _ in range(reps):
        s.velocity()
    s1 = s.stats()
    s.close()
    launches = s1["pair_launches"] - s0["pair_launches"]
    kern = None
    if launches:
        kern = {"ms": (s1["pair_ms_total"] - s0["pair_ms_total"]) / launches,
                "pair_evals": (s1["pair_evals_total"] - s0["pair_evals_total"]) / launches, "launches": launches}
    return gs, kern


def trajectory(p, st, checkpoints=(10, 50, 100)):
    """FP32C against FP64 from the same start, before chaotic divergence:
    at each checkpoint, whether the links and the reconnection log (step, i, j)
    are identical, and the largest distance between a point's two positions
    (periodic minimum image) over max |u| x elapsed time."""
    from nuvort_amd.engine import FilamentSolver
    q = dataclasses.replace(p, hip_precision="fp64")
    a, b = FilamentSolver(p), FilamentSolver(q)
    a.set_state(st)
    b.set_state(st)
    box = np.asarray(p.box_size, dtype=np.float64)
    out, done, ea, eb = {}, 0, [], []
    for c in checkpoints:
        a.step(c - done)
        b.step(c - done)
        done = c
        ea += [e[:3] for e in a.recon_events()]
        eb += [e[:3] for e in b.recon_events()]
        sa, sb = a.get_state(), b.get_state()
        same = sa.pcount == sb.pcount and np.array_equal(sa.infront, sb.infront) and \
            np.array_equal(sa.behind, sb.behind)
        entry = {"links_identical": bool(same), "reconnections_identical": ea == eb, "reconnections": len(eb)}
        if same:
            act = sb.infront != 0
            d = sa.x[act] - sb.x[act]
            d -= box * np.round(d / box)
            umax = float(np.linalg.norm(sb.u[act], axis=1).max())
            entry["max_dx_over_umax_t"] = float(np.linalg.norm(d, axis=1).max() / (umax * c * p.dt))
        out[str(c)] = entry
    a.close()
    b.close()
    return out


def rings64_extra(steps):
    """64 rings in one tangle (16.7k points): FP32C and FP64 steps, and the
    FP32C trajectory against the FP64 one over 100 steps."""
    st, p = workload()
    head = timed_steps(p, st, steps, 3, blocks=min(5, steps))
    q = dataclasses.replace(p, hip_precision="fp64")
    k64 = max(5, steps // 4)
    r64 = timed_steps(q, st, k64, 2)
    t_k, e_k, _ = kernel_timing(p, st, 10)
    t64, e64, _ = kernel_timing(q, st, 3)
    v32 = head["point_updates"] / head["seconds"]
    v64 = r64["point_updates"] / r64["seconds"]
    return {"points": st.active, "seed": 1236, "fp32c": {"value": v32, "ms_per_step": head["seconds"] / steps * 1e3,
                                                         "steps": steps, "kernel_ms": t_k * 1e3,
                                                         "frac_fp32_peak": FLOP_PER_PAIR * e_k / t_k / 1e12 / PEAK_FP32},
            "fp64_plain": {"value": v64, "ms_per_step": r64["seconds"] / k64 * 1e3, "steps": k64,
                           "kernel_ms": t64 * 1e3,
                           "frac_fp64_peak": FLOP_PER_PAIR * e64 / t64 / 1e12 / PEAK_FP64},
            "fp32c_over_fp64_step": v32 / v64,
            "trajectory_fp32c_vs_fp64": trajectory(p, st),
            "note": "64 random rings (radii U[0.01,0.03] cm) in one periodic 0.1 cm tangle, delta 5e-4 cm"}


def tree1m_extra(steps, precision="fp32c"):
    """The 1M-point counterflow octree step (timing only)."""
    from nuvort_amd import tangles
    from nuvort_amd.engine import FilamentSolver
    from nuvort_amd.runfile import RunParams
    st = tangles.random_rings(1239, 0, 0.01, 0.03, 0.2, 5e-4, target_points=1000000)
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.2, 0.2, 0.2), velocity="Tree", tree_theta=0.4,
                  normal_velocity="xflow", norm_vel_xflow=0.5, alpha=(0.206, 0.0083), recon_shots=1,
                  hip_precision=precision)
    s = FilamentSolver(p)
    s.set_state(st)
    s.step(2)
    s0 = s.stats()
    t0 = time.perf_counter()
    s.step(steps)
    sec = time.perf_counter() - t0
    s1 = s.stats()
    cur = s.get_state()
    s.close()
    ph = tree_phases(p, cur, 2)
    walk = [x["walk_ms"] for x in ph]
    build = [x["tree_build_ms"] for x in ph]
    close = [x["close_ms"] for x in ph]
    nodes = float(np.mean([x["walk_nodes"] for x in ph]))
    n = st.active
    rec = 48 if precision == "fp32c" else 80
    wms = float(np.mean(walk))
    walk_bytes = nodes * rec + n * 27 * 48
    return {"points": n, "seed": 1239, "walk": precision, "value": (s1["point_updates"] - s0["point_updates"]) / sec,
            "ms_per_step": sec / steps * 1e3, "steps": steps, "warmup": 2, "walk_ms": wms,
            "tree_build_ms": float(np.mean(build)), "close_search_ms": float(np.mean(close)),
            "recorded_steps": s1["tree_replays"] - s0["tree_replays"],
            "phases": "walk, build and close-search times from 2 eager steps after the timed ones",
            "walk_node_reads": nodes, "walk_record_bytes": rec,
            "walk_algorithmic_GBps": walk_bytes / (wms * 1e-3) / 1e9, "hbm_peak_GBps": 8000.0,
            "note": "random rings to 1M points in a periodic 0.2 cm box, delta 5e-4 cm, Barnes-Hut theta 0.4 "
                    "rebuilt every step, v_n = 0.5 cm/s along x, alpha 0.206 / 0.0083, reconnection and "
                    "re-meshing every step"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=5, help="timed steps of the plain FP64 ensemble step")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    ap.add_argument("--cpu-per-tangle", type=int, default=2048, help="oracle targets per sampled tangle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the 64-ring and 1M-point octree workloads")
    ap.add_argument("--workload", choices=("ensemble256", "rings64", "dense460k", "tree1m"), default="ensemble256",
                    help="ensemble256: the headline batched step (default); rings64: 64 rings, 16.7k points; "
                         "dense460k: the dense 800-ring tangle (467k points), relaxed 50 steps, FP32C error "
                         "statistics against FP64; tree1m: 1M points in a 0.2 cm box, Barnes-Hut theta 0.4, "
                         "counterflow mutual friction")
    ap.add_argument("--precision", choices=("fp64", "fp32c"), default="fp32c",
                    help="rings64: direct-sum arithmetic of the timed steps")
    ap.add_argument("--tree-precision", choices=("fp64", "fp32c"), default="fp32c",
                    help="tree1m: walk arithmetic (FP32 far field + FP64 near field, or FP64)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times (the reference kernel "
                         "for scaling HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 5}))
        return

    from nuvort_amd import lib as nvlib
    nvlib.load()
    if args.workload == "dense460k":
        dense(args)
        return
    if args.workload == "tree1m":
        tree1m(args)
        return
    if args.workload == "rings64":
        rings64_main(args)
        return

    sts, p = ensemble_workload()
    npts = sum(g.active for g in sts)
    log("ensemble: %d tangles, %d points" % (len(sts), npts))
    head = ensemble_steps(p, sts, "fp32c", args.warmup, args.steps)
    ms_step = head["seconds"] / args.steps * 1e3
    value = head["point_updates"] / head["seconds"]
    blk = head["block_ms_per_step"]
    log("fp32c ensemble step: %.1f ms, %.4g point-updates/s (blocks %s)" % (
        ms_step, value, " ".join("%.1f" % b for b in blk)))
    do64 = args.fp64_steps > 0   # 0: no FP64 work (profiling runs of the FP32C kernel)
    if do64:
        r64 = ensemble_steps(p, sts, "fp64", 1, args.fp64_steps)
        v64 = r64["point_updates"] / r64["seconds"]
        log("fp64 ensemble step: %.1f ms (fp32c x%.2f)" % (r64["seconds"] / args.fp64_steps * 1e3, value / v64))

    # velocities of every point in both forms; kernel times for the roofline
    g32, k32 = ensemble_velocity(p, sts, "fp32c", 3)
    acc = {}
    if do64:
        g64, k64 = ensemble_velocity(p, sts, "fp64", 1)
        u32 = np.concatenate([g.u_sup[g.infront != 0] for g in g32])
        u64 = np.concatenate([g.u_sup[g.infront != 0] for g in g64])
        acc["fp32c_vs_fp64_gpu"] = rel_err_stats(u64, u32)
        # the reconnection pairs of the steps both runs took from the same start
        last = 1 + args.fp64_steps
        e32 = sorted(e for e in head["events"] if e[0] <= last)
        e64 = sorted(r64["events"])
        acc["reconnections_fp32c_vs_fp64"] = {"steps": last, "fp32c": len(e32), "fp64": len(e64),
                                              "identical": e32 == e64}
    t_kernel = k32["ms"] * 1e-3
    achieved = FLOP_PER_PAIR * k32["pair_evals"] / t_kernel / 1e12
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "direct_fp32c_ensemble256_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32, 4), "traffic": traffic,
                "kernel": "k_bs32c (direct sum of all 256 tangles, FP32 far field)",
                "kernel_ms": round(k32["ms"], 3), "pair_evals_per_launch": int(k32["pair_evals"]),
                "flop_per_pair": FLOP_PER_PAIR, "launches_timed": k32["launches"]}
    fp64 = None
    if do64:
        a64 = FLOP_PER_PAIR * k64["pair_evals"] / (k64["ms"] * 1e-3) / 1e12
        fp64 = {"value": v64, "ms_per_step": r64["seconds"] / args.fp64_steps * 1e3, "steps": args.fp64_steps,
                "warmup": 1, "kernel_ms": k64["ms"], "achieved_tflops": a64, "frac_fp64_peak": a64 / PEAK_FP64,
                "fp32c_over_fp64_step": value / v64, "fp32c_over_fp64_kernel": k64["ms"] / k32["ms"]}

    cpu = None
    if not args.no_cpu:
        sizes = np.array([g.active for g in sts])
        order = np.argsort(sizes)
        picks = [int(order[int(q * (len(order) - 1))]) for q in np.linspace(0.0, 1.0, 8)]
        log("oracle sample: %d targets of each of %d tangles on %d threads" % (args.cpu_per_tangle, len(picks),
                                                                             args.cpu_threads))
        secs, samp, pairs = oracle_sample(sts, p, picks, args.cpu_per_tangle, args.cpu_threads)
        ucpu = np.concatenate([samp[k][1] for k in picks])
        ug32 = np.concatenate([g32[k].u[samp[k][0]] for k in picks])
        ug64 = np.concatenate([g64[k].u[samp[k][0]] for k in picks]) if do64 else None
        acc["fp32c_vs_fp64_oracle_sample"] = rel_err_stats(ucpu, ug32)
        if do64:
            acc["fp64_gpu_vs_fp64_oracle_sample"] = rel_err_stats(ucpu, ug64)
        # a CPU step of the ensemble is its velocity pass (>99% at these
        # sizes): sum over tangles of N_k^2 (target, source) evaluations at
        # the sampled rate per (target, source)
        per_pair = secs / pairs
        cpu_step = per_pair * float(np.sum(sizes.astype(np.float64) ** 2))
        cal = calibration()
        cpu = {"value": npts / cpu_step, "unit": "point-updates/s", "cores": args.cpu_threads, "kind": "port",
               "sample": "FP64 velocity (local term + direct sum over the 27 images, the reference's arithmetic) "
                         "of %d targets from each of %d tangles of the same ensemble (sizes %s), oracle C "
                         "restatement with OpenMP, %.2f s; extrapolated to the whole ensemble step by its "
                         "sum of N_k^2 (the velocity is >99%% of a CPU step at these sizes)"
                         % (args.cpu_per_tangle, len(picks), [int(sizes[k]) for k in picks], secs),
               "host": cpu_info(args.cpu_threads),
               "calibration": None if cal is None else {
                   "reference_over_restatement_time": cal["reference_over_restatement_time"],
                   "measured_on": cal["cpu_model"], "threads": cal["threads"], "points": cal["points"],
                   "note": "the reference build itself (not shipped here) took this many times longer than the "
                           "restatement on the same input: the reference's CPU rate is this baseline divided "
                           "by it"}}
        log("cpu baseline: %.4g point-updates/s on %d threads" % (cpu["value"], args.cpu_threads))

    extra = {"step_ms_blocks": {"median": float(np.median(blk)), "min": float(min(blk)), "max": float(max(blk)),
                                "blocks": len(blk)},
             "fp64_plain": fp64}
    if not args.no_extra:
        log("64-ring tangle")
        extra["rings64"] = rings64_extra(40)
        log("1M-point octree step")
        extra["tree1m"] = tree1m_extra(10)
        log("tree1m: %.2f ms/step" % extra["tree1m"]["ms_per_step"])

    line = {
        "metric": METRIC, "value": value, "unit": "point-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32c+f64", "data": "synthetic",
        "config": {"workload": "batched ensemble: 256 independent periodic random-ring tangles, point counts "
                               "log-uniform in [1k, 40k], delta 5e-4 cm, direct Biot-Savart with 26 images, "
                               "FP32 Kahan far field + FP64 near field, AB3, reconnection and re-meshing every "
                               "step, one recorded HIP graph per step",
                   "tangles": len(sts), "points": npts, "pair_evals_per_step": int(k32["pair_evals"]),
                   "seed": 1238, "dt_s": 5e-6, "box_cm": 0.1, "near_radius_delta": p.hip_near_radius,
                   "parallelism": "1 GPU"},
        "accuracy": acc,
        "max_rel_velocity_err_vs_fp64": acc["fp32c_vs_fp64_gpu"]["max_rel"] if do64 else None,
        "roofline": roofline,
        "tuning": head["tuning"],
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)
    if os.path.exists(TUNE_CACHE):
        os.remove(TUNE_CACHE)


def rings64_main(args):
    """The 64-ring tangle as its own bench line (16.7k points)."""
    st, p = workload()
    p = dataclasses.replace(p, hip_precision=args.precision)
    n = st.active
    head = timed_steps(p, st, args.steps, args.warmup, blocks=min(5, args.steps))
    value = head["point_updates"] / head["seconds"]
    t_kernel, evals, nl = kernel_timing(p, st, max(10, min(args.steps, 50)))
    achieved = FLOP_PER_PAIR * evals / t_kernel / 1e12
    g64 = velocity_once(p, st, "fp64")
    g32 = velocity_once(p, st, "fp32c")
    act = g64.infront != 0
    peak = PEAK_FP32 if args.precision == "fp32c" else PEAK_FP64
    line = {"metric": METRIC, "value": value, "unit": "point-updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": head["seconds"] / args.steps * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32c+f64" if args.precision == "fp32c" else "f64",
            "data": "synthetic",
            "config": {"workload": "64 random vortex rings in a periodic 0.1 cm box, delta 5e-4 cm, direct "
                                   "Biot-Savart with 26 images, FP32C, AB3, reconnection and re-meshing",
                       "points": n, "seed": 1236, "precision": args.precision, "parallelism": "1 GPU"},
            "accuracy": {"fp32c_vs_fp64_gpu": rel_err_stats(g64.u_sup[act], g32.u_sup[act])},
            "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                         "frac": round(achieved / peak, 4), "traffic": None, "kernel_ms": t_kernel * 1e3,
                         "pair_evals_per_launch": int(evals), "launches_timed": nl},
            "tuning": head["tuning"], "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    if os.path.exists(TUNE_CACHE):
        os.remove(TUNE_CACHE)


if __name__ == "__main__":
    main()
