"""Served scenes: a Poisson stream of two-scale dam-break requests run by the
native continuous-batching loop (tss_serve), against the same solver run
scene-at-a-time.

    python -m twoscalesplit_amd.serve --rates 1 2 4 8 16 32 64 --requests 48

Each request is a dam break with low count LogU[4k, 256k] and a step count
U[500, 5000] (``scenes.c5_request``). Requests are admitted against an HBM
budget (default 200 GB) into one resident batch -- the serving pool, whose
capacities are set here -- and advance together one launch per phase.

The scene-at-a-time reference runs each request alone on a fresh
single-scene solver (graph-replayed steps) and queues them first-in
first-out against the same arrival times: its request latency is
completion - arrival with completion_i = max(arrival_i, completion_{i-1})
+ service_i, service_i measured.

Prints one JSON line per rate.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import _native as N
from . import scenes
from .solver import Solver

BUDGET_BYTES = 200 * 10**9


class Trace:
    """Requests (scenes + arrival times) of one Poisson stream. The scenes
    depend on the seed only, the arrival times on the seed and the rate."""

    def __init__(self, seed: int, rate: float, n: int, lo: int = 4000,
                 hi: int = 256000, smin: int = 500, smax: int = 5000,
                 scene_list=None):
        rng = np.random.Generator(np.random.PCG64([seed, int(rate * 1000)]))
        gaps = rng.exponential(1.0 / rate, n)
        self.arrival = np.cumsum(gaps) - gaps[0]
        self.scenes = scene_list if scene_list is not None else [
            scenes.c5_request(seed, i, lo, hi, smin, smax) for i in range(n)]
        self.rate = rate


def pool_desc(trace: Trace, max_resident_low: int, max_scenes: int = 256
              ) -> N.TssBatchDesc:
    """Serving-pool capacities: room for max_resident_low low particles (and
    at least the largest request), cells and walls in the trace's ratio."""
    sc = trace.scenes
    n_low = sum(s.cap_low for s in sc)
    cl = sum(s.cells[0] for s in sc) / n_low
    ch = sum(s.cells[1] for s in sc) / n_low
    nb = sum(len(s.bnd_xy) for s in sc) / n_low
    cap = max(max_resident_low, max(s.cap_low for s in sc))
    big = max(sc, key=lambda s: s.cap_low)
    return N.TssBatchDesc(
        max_scenes, cap,
        max(int(nb * cap * 1.5), len(big.bnd_xy)),
        max(int(cl * cap * 1.5), big.cells[0]),
        max(int(ch * cap * 1.5), big.cells[1]))


def serve(trace: Trace, desc: N.TssBatchDesc, budget: int = BUDGET_BYTES,
          outputs: tuple[str, ...] = (), device_gen: bool = False
          ) -> tuple[dict, list[dict], list]:
    """Run the trace through tss_serve. Returns (stats, per-request results,
    per-request output arrays). ``device_gen``: requests carry their scene's
    generator spec instead of position arrays, and the device generates the
    fluid block at admission."""
    n = len(trace.scenes)
    reqs = (N.TssRequest * n)()
    keep = []
    outs = []
    for i, s in enumerate(trace.scenes):
        low = np.ascontiguousarray(s.low_xy, np.float32)
        bnd = np.ascontiguousarray(s.bnd_xy, np.float32)
        keep += [low, bnd, s.cfg]
        r = reqs[i]
        r.cfg = C.pointer(s.cfg)
        r.low_xy = low.ctypes.data
        r.n_low = len(low)
        r.cap_low = s.cap_low
        r.bnd_xy = bnd.ctypes.data
        r.n_bnd = len(bnd)
        if device_gen:
            if s.gen is None:
                raise ValueError(f"{s.name} has no generator spec")
            r.gen = C.pointer(s.gen)
            r.low_xy = None
            r.bnd_xy = None
            r.n_bnd = 0
        r.steps = s.steps
        r.arrival_s = float(trace.arrival[i])
        o = {}
        for name in outputs:
            arr = _out_array(name, s)
            o[name] = arr
            r.out[N.FIELD[name]] = arr.ctypes.data
            r.out_bytes[N.FIELD[name]] = arr.nbytes
        outs.append(o)
    st = N.TssServeStats()
    N.check(N.lib().tss_serve(C.byref(desc), int(budget), reqs, n,
                              C.byref(st)), "tss_serve")
    per = [{"status": r.status, "arrival_s": r.arrival_s,
            "admit_s": r.admit_s, "finish_s": r.finish_s,
            "particle_steps": r.particle_steps} for r in reqs]
    return st.as_dict(), per, outs


def _out_array(name: str, s) -> np.ndarray:
    cap = s.cap_low
    shapes = {"POS_L": (cap, 2), "VEL_L": (cap, 2), "RHO_L": (cap,),
              "P_L": (cap,), "STATE_L": (cap,), "POS_H": (4 * cap, 2),
              "VEL_H": (4 * cap, 2), "RHO_H": (4 * cap,)}
    dt = np.uint8 if name.startswith("STATE") else np.float32
    return np.zeros(shapes[name], dt)


def service_times(scene_list) -> tuple[list[float], list[int]]:
    """Each request alone on a fresh single-scene solver (graph-replayed
    steps): its service time and particle-steps."""
    service, psteps = [], []
    for s in scene_list:
        sol = Solver(s.cfg, s.low_xy, s.cap_low, s.bnd_xy)
        sol.synchronize()
        a = time.perf_counter()
        sol.step_graph(s.steps)
        sol.synchronize()
        service.append(time.perf_counter() - a)
        psteps.append(sol.counts()["particle_steps"])
        sol.close()
    return service, psteps


def scene_at_a_time(trace: Trace, measured=None) -> dict:
    """Requests served one at a time, first-in first-out, against the
    trace's arrivals (service times measured alone)."""
    service, psteps = measured or service_times(trace.scenes)
    done, lat = 0.0, []
    for a, sv in zip(trace.arrival, service):
        done = max(done, a) + sv
        lat.append((done - a) * 1e3)
    return {"particle_steps": int(sum(psteps)),
            "busy_s": float(sum(service)),
            "wall_s": float(done),
            "particle_steps_per_s": float(sum(psteps) / done),
            "request_ms_p50": float(np.percentile(lat, 50)),
            "request_ms_p99": float(np.percentile(lat, 99))}


def run_rate(seed: int, rate: float, n: int, pool_low: int,
             baseline: bool = True, budget: int = BUDGET_BYTES,
             scene_list=None, measured=None, device_gen: bool = False) -> dict:
    tr = Trace(seed, rate, n, scene_list=scene_list)
    desc = pool_desc(tr, pool_low)
    st, per, _ = serve(tr, desc, budget, device_gen=device_gen)
    ok = sum(1 for p in per if p["status"] == 0)
    out = {
        "rate_per_s": rate, "requests": n, "served": ok,
        "admission": "device-generated fluid blocks" if device_gen else "uploaded positions",
        "particle_steps_per_s": round(st["particle_steps_per_s"], 1),
        "wall_s": round(st["wall_s"], 3),
        "busy_s": round(st["busy_s"], 3),
        "step_ms": {"p50": round(st["step_ms_p50"], 4),
                    "p99": round(st["step_ms_p99"], 4)},
        "request_ms": {"p50": round(st["request_ms_p50"], 1),
                       "p99": round(st["request_ms_p99"], 1)},
        "queue_ms": {"p50": round(st["queue_ms_p50"], 1),
                     "p99": round(st["queue_ms_p99"], 1)},
        "ticks": st["ticks"], "graph_ticks": st["graph_ticks"],
        "max_resident": st["max_resident"],
        "pool_bytes": st["batch_bytes"], "budget_bytes": st["budget_bytes"],
        "peak_admitted_bytes": st["peak_admitted_bytes"],
        "pool_low_capacity": desc.cap_low,
    }
    if baseline:
        b = scene_at_a_time(tr, measured)
        out["scene_at_a_time"] = {k: (round(v, 3) if isinstance(v, float)
                                      else v) for k, v in b.items()}
        out["speedup_vs_scene_at_a_time"] = round(
            st["particle_steps_per_s"] / b["particle_steps_per_s"], 3)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", type=float, nargs="+",
                    default=[1, 2, 4, 8, 16, 32, 64])
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--seed", type=int, default=scenes.SEED)
    ap.add_argument("--pool-low", type=int, default=8_000_000,
                    help="low-particle capacity of the serving pool")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--no-baseline", action="store_true")
    ap.add_argument("--device-gen", action="store_true",
                    help="requests carry generator specs; fluid blocks are "
                         "generated on the device at admission")
    args = ap.parse_args(argv)
    scene_list = Trace(args.seed, 1.0, args.requests).scenes
    measured = None
    if not args.no_baseline:
        print("[serve] scene-at-a-time service times", file=sys.stderr,
              flush=True)
        measured = service_times(scene_list)
    for r in args.rates:
        print(f"[serve] rate {r}/s, {args.requests} requests", file=sys.stderr,
              flush=True)
        res = run_rate(args.seed, r, args.requests, args.pool_low,
                       not args.no_baseline, int(args.budget_gb * 1e9),
                       scene_list, measured, args.device_gen)
        print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
