"""Phase timing through the C ABI: hipEvents recorded inside the replayed
graphs (event-record nodes re-pointed per replay) and around eager steps.
Timing must not change results, and every phase must be timed.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

from twoscalesplit_amd import MODE_BAND, Solver, scenes
from twoscalesplit_amd.solver import PHASES

pytestmark = pytest.mark.gpu


def _fields(sol):
    snap = sol.snapshot()
    c = snap.pop("counts")
    return snap, c


def test_timed_graph_equals_plain_graph_and_times_every_phase():
    sc = scenes.small_two_scale(nx=32, ny=48, nxd=96, nyd=80, mode=MODE_BAND)
    print("[case] timing", sc.name, flush=True)
    a = Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    b = Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    try:
        a.step_graph(24)
        b.enable_timing(True)
        t0 = time.perf_counter()
        b.step_graph(24)
        b.synchronize()
        wall = (time.perf_counter() - t0) / 24 * 1e3
        ph = b.phase_times()
        sa, ca = _fields(a)
        sb, cb = _fields(b)
        assert ca == cb
        for k in sa:
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
        assert list(ph) == PHASES
        assert all(v > 0 for v in ph.values()), ph
        total = sum(ph.values())
        # device phases fit inside the host-observed time per step
        assert total <= wall * 1.05 + 0.05, (total, wall)
        # eager steps are timed as well; enable_timing resets the averages
        b.enable_timing(True)
        b.step(3)
        ph2 = b.phase_times()
        assert all(v > 0 for v in ph2.values()), ph2
        b.enable_timing(False)
        b.step(2)
        a.step(5)
        sa, _ = _fields(a)
        sb, _ = _fields(b)
        for k in sa:
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    finally:
        a.close()
        b.close()
