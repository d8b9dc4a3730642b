#!/usr/bin/env python3
"""Throughput of the two-scale WCSPH step on one MI355X.

Workload (the headline "1M + 4M two-scale" configuration): large dam break,
domain [0,4]x[0,1], 1414x707 = 999,698 low-res particles (dx = 0.8/707),
every low particle split 1 -> 4 at the first step (forced all-split: the
reference's split rule with the threshold left of the domain), giving
999,698 low + 3,998,792 high-res particles advanced per step, CFL-0.4
adaptive dt, FP32 with compensated sums, hipGraph-captured steps.

A step advances every low and high particle once: grid build (hash, stable
sort by cell, cell ranges, reorder), density + Tait, forces + symplectic
Euler, adaptivity + list compaction, FP64 reductions (dt, conservation).
value = particle-steps/s = steps * (n_low + n_high) / wall time, with the
state resident in HBM.

Beside the headline: sections for the plain (C1), two-scale (C2) and
natural-band large (C3) dam breaks, each with GPU and CPU particle-steps/s,
p50/p99 step latency and the error against the FP64 oracle; the batched
(C4) and served (C5) scene mixes against the same solver run
scene-at-a-time.

Prints ONE JSON line. --gpus must be 1 (one GPU per box).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

# the CPU legs' OpenMP threads stay on one core each, packed (read by the
# OpenMP runtime when it starts, before any parallel region)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 1M+4M two-scale; p50/p99 step latency, batched scenes"

# MI355X published peaks
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3          # published FP32 (vector, packed issue)
FP32_NONPACKED_TFLOPS = 78.6      # without packed-FP32 issue

# per-item flop weights of the neighbour kernels (candidate test, density
# pair, force pair, wall pair)
W_CAND, W_DENS, W_FORCE, W_WALL = 7, 28, 73, 20


def _log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def _ensure_built():
    from twoscalesplit_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        from twoscalesplit_amd import build
        build.build_lib()


def grid_bytes(n: int, cells: int, passes: int) -> float:
    """Algorithmic HBM bytes of one grid build by the radix-sort formula:
    hash (read pos, write key and index), P radix passes over key+value,
    range read, reorder gather + scatter of the 21-byte payload, cell table
    clear + writes."""
    S = 21
    return n * (16 + 16 * passes + 4 + 2 * S) + 16 * cells


def count_sort_bytes(n: int, cells: int) -> float:
    """Algorithmic HBM bytes of the counting-sort grid build: hash (list 4,
    position 8, key 4), scatter (key 4, list 4, slot 8, sorted key 4),
    reorder (slot 8, sorted key 4, gathered state + position + velocity 17;
    writes id 4, position 8, position + velocity 16, state 1, id -> index 4,
    scene 4); per cell: the count read twice (tile sums, ranges), the owner
    4, the range 8, the count reset 4."""
    return n * (16 + 20 + 66) + cells * 24


def grid_sort_path(cap_low: int, cells: int) -> str:
    """The grid build's sort as TSS_SORT_AUTO picks it (include/tss.h): the
    counting sort when the cells number at most 4 per key, else one
    workgroup up to 65536 keys, else the radix sort."""
    keys = 5 * max(cap_low, 1)
    if cells <= 4 * keys:
        return "count"
    return "small" if keys <= 65536 else "radix"


def _traffic(kernel: str) -> dict | None:
    """HBM bytes per dispatch of `kernel` from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, each scaled by the
    reading of a plain copy kernel of known size)."""
    path = os.path.join(ROOT, "profiles", "pmc_c3_forced.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 6 and parts[0] == "traffic" and parts[1] == kernel:
                return {"bytes": int(parts[2]),
                        "source": "profiles/pmc_c3_forced.txt"}
    return None


def host_cpu() -> dict:
    """The host's CPU as the CPU legs use it: model (lscpu), logical CPUs,
    the CPUs this process may run on, and the OpenMP threads of the legs
    (every CPU of the job's share: its affinity mask, bounded by the
    OMP_NUM_THREADS the job is given)."""
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True,
                             timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    aff = len(os.sched_getaffinity(0))
    threads = aff
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": aff,
            "omp_num_threads": omp or None, "threads": threads,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES")}


def cpu_baseline(snap: dict, sc, steps: int, cpu: dict) -> dict:
    """The oracle's FP32 transcription of the reference kernels (OpenMP)
    on the GPU state of the same workload, on every thread of the job's
    CPU share and (one step) on one."""
    from oracle import binding as ob
    threads = cpu["threads"]

    def rate(nsteps: int, nthreads: int) -> tuple[float, float, int]:
        o = ob.Scene(ob.params_from_config(sc.cfg), sc.cap_low, sc.bnd_xy)
        o.load_snapshot(snap)
        n = o.s.n_l + o.s.n_h
        t0 = time.perf_counter()
        for _ in range(nsteps):
            o.step(32, nthreads)
        dt = time.perf_counter() - t0
        return nsteps * n / dt, dt, n

    value, dt, n = rate(steps, threads)
    one, dt1, _ = rate(1, 1)
    return {"value": value, "unit": "particle-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the same {n}-particle state "
                      f"(taken from the GPU after warmup), oracle FP32 "
                      f"transcription of the reference kernels, OpenMP",
            "seconds": round(dt, 3),
            "value_1_thread": one, "seconds_1_thread": round(dt1, 3),
            **cpu}


def oracle_errors(pre: dict, post: dict, sc) -> dict:
    """CPU leg: the GPU's step from `pre` to `post` against the FP64 oracle
    re-seeded from `pre` -- max relative density error, pressure error
    scaled by max(|p|, B) and by max(|p|, 7 B 2^-23) (p50/p99/max), and the
    acceleration error on the GPU's density (scaled by max(|a|, sum of the
    term magnitudes))."""
    from oracle import binding as ob
    o = ob.Scene(ob.params_from_config(sc.cfg), sc.cap_low, sc.bnd_xy)
    o.load_snapshot(pre)
    g = o.grid()
    rl, pl, rh, ph = o.density64(g)
    c = pre["counts"]
    lists = (pre["LIST_L"][:c["n_low"]], pre["LIST_H"][:c["n_high"]])
    B = float(sc.cfg.tait_coefficient)
    er, ep, epr = [], [], []
    for ids, r64, p64, R, P in ((lists[0], rl, pl, "RHO_L", "P_L"),
                                (lists[1], rh, ph, "RHO_H", "P_H")):
        if len(ids) == 0:
            continue
        rg = post[R][ids].astype(np.float64)
        pg = post[P][ids].astype(np.float64)
        er.append(np.abs(rg - r64[ids]) / np.abs(r64[ids]))
        d = np.abs(pg - p64[ids])
        ep.append(d / np.maximum(np.abs(p64[ids]), B))
        epr.append(d / np.maximum(np.abs(p64[ids]), 7.0 * B * 2.0 ** -23))
    for k, v in (("rho_l", "RHO_L"), ("p_l", "P_L"), ("rho_h", "RHO_H"),
                 ("p_h", "P_H")):
        o.a[k][:] = post[v]
    al, ah, sl, sh = o.force64(g)
    ea = []
    for ids, a64, s, A in ((lists[0], al, sl, "ACC_L"),
                           (lists[1], ah, sh, "ACC_H")):
        if len(ids) == 0:
            continue
        d = np.linalg.norm(post[A][ids].astype(np.float64) - a64[ids], axis=1)
        ea.append(d / np.maximum(np.linalg.norm(a64[ids], axis=1), s[ids]))
    cat = lambda v: np.concatenate(v) if v else np.zeros(1)
    er, ep, epr, ea = cat(er), cat(ep), cat(epr), cat(ea)
    pct = lambda v: [float(np.percentile(v, 50)), float(np.percentile(v, 99)),
                     float(v.max())]
    return {"rho_rel_max": float(er.max()), "p_rel_B_max": float(ep.max()),
            "p_rel_floor_p50_p99_max": pct(epr),
            "acc_rel_max": float(ea.max()),
            "tolerance": "1e-5 on rho, p (B-scaled) and acc"}


def run_section(sc, steps: int, latency_steps: int) -> tuple[dict, tuple]:
    """One scene, alone, on the GPU: particle-steps/s over `steps`
    graph-replayed steps from the initial state (the state resident in
    HBM), the time-averaged particle counts, p50/p99 of synchronised single
    steps; returns the section and the (pre, post) states of one more step
    for the CPU leg's oracle check."""
    from twoscalesplit_amd import Solver
    sol = Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    try:
        sol.synchronize()
        c0 = sol.counts()
        t0 = time.perf_counter()
        sol.step_graph(steps)
        sol.synchronize()
        t1 = time.perf_counter()
        c1 = sol.counts()
        ps = c1["particle_steps"] - c0["particle_steps"]
        lat = []
        for _ in range(latency_steps):
            a = time.perf_counter()
            sol.step(1)
            sol.synchronize()
            lat.append((time.perf_counter() - a) * 1e3)
        pre = sol.snapshot()
        sol.step(1)
        post = sol.snapshot()
        cfin = sol.counts()
    finally:
        sol.close()
    out = {
        "workload": sc.name,
        "steps": steps,
        "value": round(ps / (t1 - t0), 1),
        "unit": "particle-steps/s",
        "ms_per_step": round((t1 - t0) / steps * 1e3, 4),
        "particles_per_step_mean": round(ps / steps, 1),
        "n_low": c1["n_low_last"],
        "n_high_mean": round(ps / steps - c1["n_low_last"], 1),
        "n_high_at_end": cfin["n_high_last"],
        "step_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                            "p99": round(float(np.percentile(lat, 99)), 4)},
    }
    return out, (pre, post)


def section_cpu_leg(out: dict, sc, states: tuple, cpu_steps: int, cpu: dict,
                    cpu_1thread_steps: int) -> None:
    """CPU leg of a section: the oracle check of the GPU's step, then the
    oracle's FP32 transcription of the reference kernels timed from the
    same initial state on the job's CPU threads (and a shorter run on
    one). Runs after every GPU measurement: the OpenMP runtime binds the
    calling thread when it starts."""
    out["oracle_error"] = oracle_errors(states[0], states[1], sc)
    from oracle import binding as ob

    def cpu_run(nsteps: int, threads: int) -> dict:
        o = ob.Scene.from_initial(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
        ps0 = 0
        t = time.perf_counter()
        for _ in range(nsteps):
            ps0 += o.s.n_l + o.s.n_h
            o.step(32, threads)
        dt = time.perf_counter() - t
        return {"value": round(ps0 / dt, 1), "steps": nsteps,
                "seconds": round(dt, 3), "threads": threads}
    out["cpu"] = cpu_run(cpu_steps, cpu["threads"])
    out["cpu_1_thread"] = cpu_run(cpu_1thread_steps, 1)
    out["speedup_vs_cpu"] = round(out["value"] / out["cpu"]["value"], 1)


def run_batched(args, rule: str) -> dict:
    """Batched scenes: LogU[2k,64k] two-scale dam breaks of random aspect and
    viscosity advanced together (one launch per phase, graph-replayed steps)
    against the same solver run scene-at-a-time."""
    from twoscalesplit_amd import Batch, Solver, scenes
    scs = scenes.c4_batched(n_scenes=args.batch_scenes,
                            steps=args.batch_steps, rule=rule)
    cap = sum(s.cap_low for s in scs)
    nb = sum(len(s.bnd_xy) for s in scs)
    cl = sum(s.cells[0] for s in scs)
    ch = sum(s.cells[1] for s in scs)
    _log("batched:", len(scs), "scenes,", cap, "low particles")
    b = Batch(len(scs), cap, nb, cl, ch)
    ids = [b.add_scene(s.cfg, s.low_xy, s.cap_low, s.bnd_xy) for s in scs]
    b.synchronize()
    # the autotuner's trials are the run's first steps
    t0 = time.perf_counter()
    tune = b.autotune(steps_per_trial=1, rounds=3)
    b.synchronize()
    t1 = time.perf_counter()
    rest = args.batch_steps - tune["steps_taken"]
    b.step_graph(rest)
    b.synchronize()
    t2 = time.perf_counter()
    cnt = [b.counts(i) for i in ids]
    ps_total = sum(c["particle_steps"] for c in cnt)
    n_low_total = sum(c["n_low_last"] for c in cnt)
    # post-tuning particle-steps: total minus the tuning steps' share
    lat = []
    for _ in range(args.latency_steps):
        a = time.perf_counter()
        b.step(1)
        b.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    cnt2 = [b.counts(i) for i in ids]
    ps_lat = sum(c["particle_steps"] for c in cnt2) - ps_total
    lat = np.array(lat)
    n_low = sum(c["n_low"] for c in cnt)
    n_high = sum(c["n_high"] for c in cnt)
    b.close()

    # scene-at-a-time: the same single-scene solver (graph-replayed steps),
    # one scene after the other, same step count
    _log("scene-at-a-time baseline")
    ps_one, t_one = 0, 0.0
    for s in scs:
        sol = Solver(s.cfg, s.low_xy, s.cap_low, s.bnd_xy)
        sol.synchronize()
        a = time.perf_counter()
        sol.step_graph(args.batch_steps)
        sol.synchronize()
        t_one += time.perf_counter() - a
        ps_one += sol.counts()["particle_steps"]
        sol.close()
    value = ps_total / (t2 - t0)
    base = ps_one / t_one
    return {
        "workload": f"{len(scs)} two-scale dam breaks, low counts LogU[2k,64k], "
                    f"aspect U(0.25,2), alpha U(0.02,0.1), high cap 4x low, "
                    f"{args.batch_steps} steps each, "
                    + ("surface-band injection/removal" if rule == "band" else
                       "reference split rule at mid-block"),
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "ms_per_step": round((t2 - t0) / args.batch_steps * 1e3, 4),
        "post_tuning_ms_per_step": round((t2 - t1) / max(rest, 1) * 1e3, 4),
        "autotune_steps": tune["steps_taken"],
        "tuning": tune["tuning"],
        "n_low": n_low,
        "n_high_mean": round(ps_total / args.batch_steps - n_low_total, 1),
        "n_high_at_end": n_high,
        "particle_steps": ps_total,
        "step_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                            "p99": round(float(np.percentile(lat, 99)), 4)},
        "latency_particle_steps_per_step": round(ps_lat / max(len(lat), 1)),
        "scene_at_a_time": {"value": round(base, 1),
                            "seconds": round(t_one, 3),
                            "particle_steps": ps_one},
        "speedup_vs_scene_at_a_time": round(value / base, 3),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--latency-steps", type=int, default=60)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--natural", action="store_true",
                    help="surface-band C3 instead of the forced all-split")
    ap.add_argument("--batch-scenes", type=int, default=256,
                    help="scenes of the batched dam-break mix (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="requests of the served run (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=64.0,
                    help="Poisson arrival rate of the served run (req/s)")
    ap.add_argument("--copy-probe-mb", type=int, default=512,
                    help="plain-copy bandwidth probe size (0: skip); also "
                         "the reference dispatch for HBM counter calibration")
    ap.add_argument("--sections", default="c1,c2,c3band",
                    help="comma list of per-config sections beside the "
                         "headline (c1, c2, c3band; empty: none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU "
                          "sharding is design-only (--gpus must be 1)"}))
        return 2
    _ensure_built()
    # the CPU share, read before the OpenMP runtime starts (it binds the
    # calling thread)
    cpu = host_cpu()
    from twoscalesplit_amd import Solver, scenes

    _log("scene")
    sc = scenes.c3_large(forced=not args.natural)
    sol = Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    # warmup: the native autotuner picks the neighbour kernels' launch
    # shapes on this workload (its trials are ordinary steps), then both
    # graph variants (plain and timed) are built
    _log("warmup + autotune")
    tune = sol.autotune(steps_per_trial=1, rounds=3)
    sol.step_graph(args.warmup)
    sol.enable_timing(True)
    sol.step_graph(max(4, args.warmup // 4))
    sol.synchronize()
    w0 = sol.work_counts()
    c0 = sol.counts()
    n0 = c0["n_low"] + c0["n_high"]

    # timed region: K graph-replayed steps; hipEvents inside the graphs
    # time every phase of every step on the solver's stream
    sol.enable_timing(True)
    _log("timed", args.steps, "steps of", n0, "particles")
    t0 = time.perf_counter()
    sol.step_graph(args.steps)
    sol.synchronize()
    t1 = time.perf_counter()
    ph = sol.phase_times()
    sol.enable_timing(False)
    c1 = sol.counts()
    w1 = sol.work_counts()
    elapsed = t1 - t0
    # particle-steps advanced in the timed region (device-side counter)
    psteps = c1["particle_steps"] - c0["particle_steps"]
    value = psteps / elapsed

    # the same replays without timing nodes (reported beside value)
    _log("untimed", args.steps)
    u0 = time.perf_counter()
    sol.step_graph(args.steps)
    sol.synchronize()
    untimed_ms = (time.perf_counter() - u0) / args.steps * 1e3

    # per-step latency (host wall clock, one synchronised step each)
    _log("latency", args.latency_steps)
    lat = []
    for _ in range(args.latency_steps):
        a = time.perf_counter()
        sol.step(1)
        sol.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    lat = np.array(lat)

    # algorithmic work per launch: mean of the counts before and after.
    # The force passes do not test their same-level candidates again: they
    # walk the density pass's recorded accept masks, testing only the
    # decoded candidates (the density's same-level pairs, counted here with
    # its cross-level ones as an upper bound) for the open ball.
    work = {k: (w0[k] + w1[k]) / 2.0 for k in w0}
    fl = {
        "density_low": W_CAND * (work["cand_ll"] + work["cand_lh"]) +
                       W_DENS * work["dens_pairs_l"],
        "density_high": W_CAND * (work["cand_hh"] + work["cand_hl_density"]) +
                        W_DENS * work["dens_pairs_h"],
        "force_low": W_CAND * (work["dens_pairs_l"] + work["cand_lh"]) +
                     W_FORCE * work["force_pairs_l"] + W_WALL * work["wall_l"],
        "force_high": W_CAND * (work["dens_pairs_h"] + work["cand_hl"]) +
                      W_FORCE * work["force_pairs_h"] + W_WALL * work["wall_h"],
    }
    tflops = {k: fl[k] / (ph[k] * 1e-3) / 1e12 for k in fl if ph[k] > 0}
    dom = max(fl, key=lambda k: ph[k])
    kernel_name = {"density_low": "k_density<false>",
                   "density_high": "k_density<true>",
                   "force_low": "k_force<false>",
                   "force_high": "k_force<true>"}[dom]
    n_now = c1["n_low"] + c1["n_high"]
    cells = c1["cells_low"] + c1["cells_high"]
    # 8-bit radix digits over the unified cell keys (the formula's P)
    passes = max(1, math.ceil(math.ceil(math.log2(max(cells, 2))) / 8))
    gb = grid_bytes(n_now, cells, passes)
    grid_gbs = gb / (ph["grid"] * 1e-3) / 1e9
    gpath = grid_sort_path(c1["cap_low"], cells)
    gpb = count_sort_bytes(n_now, cells) if gpath == "count" else gb
    traffic = _traffic(kernel_name)
    # every neighbour kernel and the grid build against its attainable
    # rate min(HBM peak x AI, FP32 peak), AI = algorithmic flops / bytes;
    # bytes per level (compulsory traffic: each particle's own data once,
    # the cell table, the other level's data once): density
    # N 17 + 8 C + N_other 9, force + integration N 49 + 8 C + N_other 25
    # + N_B 8; counter bytes from the committed PMC summary
    nl_, nh_ = c1["n_low_last"], c1["n_high_last"]
    cl_, ch_ = c1["cells_low"], c1["cells_high"]
    nb_ = c1["n_bnd"]
    alg_bytes = {"density_low": nl_ * 17 + 8 * cl_ + nh_ * 9,
                 "density_high": nh_ * 17 + 8 * ch_ + nl_ * 9,
                 "force_low": nl_ * 49 + 8 * cl_ + nh_ * 25 + nb_ * 8,
                 "force_high": nh_ * 49 + 8 * ch_ + nl_ * 25 + nb_ * 8}
    kern_names = {"density_low": "k_density<false>",
                  "density_high": "k_density<true>",
                  "force_low": "k_force<false>", "force_high": "k_force<true>"}
    kernels = {}
    for k in fl:
        ms = ph[k]
        ai = fl[k] / alg_bytes[k]
        att = min(HBM_PEAK_GBS * 1e9 * ai, FP32_PEAK_TFLOPS * 1e12)
        tr = _traffic(kern_names[k])
        kernels[k] = {
            "kernel": kern_names[k], "ms": round(ms, 4),
            "tflops": round(fl[k] / (ms * 1e-3) / 1e12, 3),
            "algorithmic_gbs": round(alg_bytes[k] / (ms * 1e-3) / 1e9, 1),
            "ai_flop_per_byte": round(ai, 1),
            "attainable_tflops": round(att / 1e12, 1),
            "frac_of_attainable": round(fl[k] / (ms * 1e-3) / att, 4),
            "counter_bytes": tr["bytes"] if tr else None,
            "counter_gbs": round(tr["bytes"] / (ms * 1e-3) / 1e9, 1) if tr else None,
        }
    # the grid build's counter bytes: the kernels of the sort path it takes
    grid_counter = None
    if gpath == "count":
        gtr = {kn: _traffic(kn) for kn in ("k_begin_hash", "k_cell_sums", "k_cell_ranges",
                                           "k_cell_scatter", "k_cell_reorder")}
        if all(gtr.values()):
            grid_counter = sum(v["bytes"] for v in gtr.values())
    elif gpath == "radix":
        gtr = {kn: _traffic(kn) for kn in ("k_begin_hash", "k_sort_hist", "k_sort_scan",
                                           "k_sort_scatter", "k_reorder")}
        if all(gtr.values()):
            grid_counter = (gtr["k_begin_hash"]["bytes"] + gtr["k_reorder"]["bytes"] +
                            passes * (gtr["k_sort_hist"]["bytes"] + gtr["k_sort_scan"]["bytes"] +
                                      gtr["k_sort_scatter"]["bytes"]))
    roofline = {
        "kernel": kernel_name,
        "bound": "valu",
        "achieved": round(tflops[dom], 3),
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(tflops[dom] / FP32_PEAK_TFLOPS, 4),
        "traffic": traffic["bytes"] if traffic else None,
        "frac_nonpacked_peak": round(tflops[dom] / FP32_NONPACKED_TFLOPS, 4),
        "algorithmic_flops_per_launch": int(fl[dom]),
        "avg_launch_ms": round(ph[dom], 4),
        "traffic_source": traffic["source"] if traffic else None,
    }
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": sc.name,
            "n_low": c1["n_low_last"],
            "n_high": c1["n_high_last"],
            "dt": "CFL 0.4 (FP64 max-reduction on the GPU)",
            "step": "hipGraph-captured pairs of steps, device-resident counts",
        },
        "untimed_graph_ms_per_step": round(untimed_ms, 4),
        "tuning": tune["tuning"],
        "step_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                            "p99": round(float(np.percentile(lat, 99)), 4)},
        "phase_ms": {k: round(v, 4) for k, v in ph.items()},
        "kernel_tflops": {k: round(v, 3) for k, v in tflops.items()},
        "roofline": roofline,
        "grid_roofline": {
            "kernel": {"count": "grid build (hash + counting sort by cell + ranges + reorder)",
                       "radix": "grid build (hash + radix sort + ranges + reorder)",
                       "small": "grid build (hash + one-workgroup radix sort + reorder)"}[gpath],
            "bound": "hbm",
            "achieved": round(grid_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(grid_gbs / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": int(gb),
            "radix_passes": passes,
            "sort_path": gpath,
            "path_bytes_per_launch": int(gpb),
            "path_frac": round(gpb / (ph["grid"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "counter_bytes": grid_counter,
            "counter_gbs": (round(grid_counter / (ph["grid"] * 1e-3) / 1e9, 1)
                            if grid_counter else None),
        },
        "kernels": kernels,
        "work": {k: int(v) for k, v in work.items()},
    }
    if args.copy_probe_mb > 0:
        import ctypes
        from twoscalesplit_amd import _native
        g = ctypes.c_double()
        _native.check(_native.lib().tss_bandwidth_probe(
            args.copy_probe_mb << 20, 3, ctypes.byref(g)), "bandwidth probe")
        out["copy_probe_gbs"] = round(g.value, 1)
    snap = None if args.no_cpu_baseline else sol.snapshot()
    sol.close()
    secs = [s for s in args.sections.split(",") if s]
    cpu_legs = []
    if secs:
        out["configs"] = {}
    for name in secs:
        # C1: 1000 fixed-dt steps on the GPU and on the CPU transcription;
        # C2: 2000 steps; C3 band: 500 GPU steps and a bounded CPU sample
        if name == "c1":
            sc1, st, cs, c1t = scenes.c1_plain(), 1000, 1000, 100
        elif name == "c2":
            sc1, st, cs, c1t = scenes.c2_two_scale(), 2000, 2000, 100
        elif name == "c3band":
            sc1, st, cs, c1t = scenes.c3_large(forced=False), 500, 10, 1
        else:
            raise SystemExit(f"unknown section {name!r}")
        _log("section", name, st, "steps")
        sec, states = run_section(sc1, st, args.latency_steps)
        out["configs"][name] = sec
        if not args.no_cpu_baseline:
            cpu_legs.append((sec, sc1, states, cs, c1t))
    if args.batch_scenes > 0:
        out["batched"] = run_batched(args, "band")
        out["batched_reference_rule"] = run_batched(args, "reference")
    if args.serve_requests > 0:
        from twoscalesplit_amd import serve
        _log("served:", args.serve_requests, "requests at",
             args.serve_rate, "/s")
        out["served"] = serve.run_rate(scenes.SEED, args.serve_rate,
                                       args.serve_requests, 8_000_000)
    for sec, sc1, states, cs, c1t in cpu_legs:
        _log("cpu leg", sec["workload"])
        section_cpu_leg(sec, sc1, states, cs, cpu, c1t)
    if snap is not None:
        _log("cpu baseline", args.cpu_steps, "steps,", cpu["threads"], "threads")
        out["cpu_baseline"] = cpu_baseline(snap, sc, args.cpu_steps, cpu)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out, default=float), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
