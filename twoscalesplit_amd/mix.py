"""Batched scene mixes from the command line: the 256-scene two-scale dam
breaks (low counts LogU[2k, 64k], aspect U(0.25, 2), alpha U(0.02, 0.1),
high cap 4x low) advanced together, one launch per phase, graph-replayed
steps; prints one JSON line with the mix's particle-steps/s and per-step
time.

    python -m twoscalesplit_amd.mix --scenes 256 --steps 300 --rule band
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from . import scenes
from .solver import Batch


def run(n_scenes: int, steps: int, rule: str, tune: bool = True) -> dict:
    scs = scenes.c4_batched(n_scenes=n_scenes, steps=steps, rule=rule)
    b = Batch(len(scs), sum(s.cap_low for s in scs), sum(len(s.bnd_xy) for s in scs),
              sum(s.cells[0] for s in scs), sum(s.cells[1] for s in scs))
    try:
        ids = [b.add_scene(s.cfg, s.low_xy, s.cap_low, s.bnd_xy) for s in scs]
        b.synchronize()
        t0 = time.perf_counter()
        taken = b.autotune(steps_per_trial=1, rounds=3)["steps_taken"] if tune else 0
        b.step_graph(max(steps - taken, 0))
        b.synchronize()
        t1 = time.perf_counter()
        cnt = [b.counts(i) for i in ids]
    finally:
        b.close()
    ps = sum(c["particle_steps"] for c in cnt)
    return {"scenes": len(scs), "rule": rule, "steps": steps,
            "n_low": sum(c["n_low_last"] for c in cnt),
            "n_high_at_end": sum(c["n_high_last"] for c in cnt),
            "particle_steps": ps,
            "particle_steps_per_s": round(ps / (t1 - t0), 1),
            "ms_per_step": round((t1 - t0) / steps * 1e3, 4)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--rule", choices=("band", "reference"), default="band")
    ap.add_argument("--no-tune", action="store_true")
    args = ap.parse_args(argv)
    print(json.dumps(run(args.scenes, args.steps, args.rule, not args.no_tune)),
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
