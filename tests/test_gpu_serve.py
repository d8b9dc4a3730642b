"""The native serving loop (tss_serve): requests arriving over time are
admitted into one resident batch, advanced together and retired; each
request's final state must equal the same scene run alone, bit for bit
(every per-particle sum is lane-local and scene-local), whatever else is
resident -- also when a tight budget makes requests wait.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from twoscalesplit_amd import Solver, scenes
from twoscalesplit_amd import _native as N
from twoscalesplit_amd import serve as S

pytestmark = pytest.mark.gpu

OUT = ("POS_L", "VEL_L", "RHO_L", "P_L", "STATE_L", "POS_H", "VEL_H")


def _trace(n=6, rate=200.0, seed=11):
    tr = S.Trace(seed, rate, n, lo=2000, hi=12000, smin=20, smax=70)
    # a promoted-group scene in the mix as well
    tr.scenes[1].cfg.promote_min_age = 2
    tr.scenes[1].cfg.promote_max_age = 5
    return tr


def _alone(s):
    sol = Solver(s.cfg, s.low_xy, s.cap_low, s.bnd_xy)
    try:
        sol.step(s.steps)
        return {k: sol.read(k) for k in OUT}, sol.counts()
    finally:
        sol.close()


def _check(tr, per, outs):
    for i, s in enumerate(tr.scenes):
        assert per[i]["status"] == 0, (i, per[i])
        want, cnt = _alone(s)
        assert per[i]["particle_steps"] == cnt["particle_steps"]
        for k in OUT:
            np.testing.assert_array_equal(outs[i][k], want[k],
                                          err_msg=f"request {i}: {k}")
        assert per[i]["admit_s"] >= per[i]["arrival_s"] - 1e-9
        assert per[i]["finish_s"] > per[i]["admit_s"]


def test_served_results_equal_alone():
    tr = _trace()
    print("[case] serve", len(tr.scenes), "requests",
          [s.n_low for s in tr.scenes], [s.steps for s in tr.scenes],
          flush=True)
    desc = S.pool_desc(tr, max_resident_low=200_000, max_scenes=8)
    st, per, outs = S.serve(tr, desc, outputs=OUT)
    print("[serve]", st, flush=True)
    assert st["max_resident"] >= 2          # requests overlapped
    assert st["ticks"] > 0 and st["graph_ticks"] > 0
    assert st["particle_steps"] == sum(p["particle_steps"] for p in per)
    _check(tr, per, outs)


def test_tight_budget_queues_and_rejects():
    tr = _trace(n=5, rate=1000.0, seed=12)
    big = scenes.c5_request(12, 99, lo=60000, hi=60001, smin=10, smax=11)
    tr.scenes.append(big)
    tr.arrival = np.append(tr.arrival, tr.arrival[-1])
    # budget: room for the two largest small requests, not for `big`
    sizes = [N.lib().tss_scene_bytes(ctypes.byref(s.cfg), s.cap_low,
                                     len(s.bnd_xy)) for s in tr.scenes]
    budget = sorted(sizes[:-1])[-1] + sorted(sizes[:-1])[-2]
    assert sizes[-1] > budget
    print("[case] serve with budget", budget, sizes, flush=True)
    desc = S.pool_desc(tr, max_resident_low=400_000, max_scenes=8)
    st, per, outs = S.serve(tr, desc, budget=budget, outputs=OUT)
    print("[serve]", st, flush=True)
    assert per[-1]["status"] != 0              # never fits: rejected
    assert st["peak_admitted_bytes"] <= budget
    assert st["queue_ms_p99"] > 0              # someone waited for room
    tr.scenes.pop()
    tr.arrival = tr.arrival[:-1]
    _check(tr, per[:-1], outs[:-1])
