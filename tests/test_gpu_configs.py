"""The batched and served configurations at their real sizes, the binding
high-resolution cap, and host writes that break the list invariants.

* 256-scene batched mix (low counts LogU[2k, 64k]): scenes spread over the
  size range equal the same scenes run alone, bit for bit, and the largest
  scene's next step, taken from the mix, passes the oracle check.
* Served trace with request sizes LogU[4k, 256k] (steps truncated so the
  test takes seconds): every request's final state equals the scene run
  alone.
* C2 with a cap that binds on the first step: the admitted splits are the
  first candidates in sorted order (the oracle's rule), the cap is filled
  and never exceeded.
* Children pulled beyond their clamp radius: clamped bit-exactly as the
  oracle does.
* Host writes that break the list invariants are refused before a launch.

Each case prints its name before touching the GPU.
"""
from __future__ import annotations

import numpy as np
import pytest

from parity import check_step
from twoscalesplit_amd import Batch, Solver, scenes
from twoscalesplit_amd import _native as N
from twoscalesplit_amd import serve as S

pytestmark = pytest.mark.gpu

FIELDS = ("POS_L", "VEL_L", "ACC_L", "RHO_L", "P_L", "VIS_L", "STATE_L",
          "AGE_L", "POS_H", "VEL_H", "ACC_H", "RHO_H", "P_H", "STATE_H",
          "LIST_L", "LIST_H", "LIST_G", "DT")


def _say(*a):
    print("[case]", *a, flush=True)


def _solver(sc):
    return Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)


def test_c4_full_mix_equals_alone():
    scs = scenes.c4_batched(n_scenes=256, steps=10)
    cap = sum(s.cap_low for s in scs)
    _say("C4: 256 scenes,", cap, "low particles, 10 steps (4 eager + 6 graph)")
    b = Batch(len(scs), cap, sum(len(s.bnd_xy) for s in scs),
              sum(s.cells[0] for s in scs), sum(s.cells[1] for s in scs))
    try:
        ids = [b.add_scene(s.cfg, s.low_xy, s.cap_low, s.bnd_xy) for s in scs]
        b.step(4)
        b.step_graph(6)
        b.synchronize()
        order = np.argsort([s.n_low for s in scs])
        picks = [int(order[k]) for k in np.linspace(0, len(scs) - 1, 8).astype(int)]
        n_high = 0
        for i in picks:
            s = scs[i]
            _say("C4 scene", i, "n_low", s.n_low, "alone vs mix")
            alone = _solver(s)
            try:
                alone.step(10)
                for k in FIELDS:
                    np.testing.assert_array_equal(b.read(ids[i], k), alone.read(k),
                                                  err_msg=f"scene {i}: {k}")
                cb, _ = b.conservation(ids[i])
                ca, _ = alone.conservation()
                np.testing.assert_array_equal(cb, ca)
                n_high += alone.counts()["n_high"]
            finally:
                alone.close()
        assert n_high > 0, "the band rules must have split in these scenes"
        # the largest scene's next step from the mix's state, against the
        # oracle (pair sets, grid, density, force, adaptivity)
        big = picks[-1]
        _say("C4 largest scene", big, "n_low", scs[big].n_low, "oracle step")
        sol = Solver.from_snapshot(b.snapshot(ids[big]))
        try:
            rep = []
            check_step(sol, scs[big].cfg, scs[big].bnd_xy, rep)
            print(rep, flush=True)
        finally:
            sol.close()
    finally:
        b.close()


def test_c5_real_size_trace_equals_alone():
    tr = S.Trace(29, 64.0, 8, lo=4000, hi=256000, smin=20, smax=60)
    sizes = [s.n_low for s in tr.scenes]
    _say("C5 trace", sizes, [s.steps for s in tr.scenes])
    assert max(sizes) > 64000            # the upper part of LogU[4k, 256k]
    desc = S.pool_desc(tr, max_resident_low=sum(sizes), max_scenes=8)
    st, per, outs = S.serve(tr, desc, outputs=("POS_L", "VEL_L", "RHO_L",
                                                "P_L", "STATE_L", "POS_H",
                                                "VEL_H"))
    print("[serve]", st, flush=True)
    assert st["max_resident"] >= 2
    for i, s in enumerate(tr.scenes):
        assert per[i]["status"] == 0, (i, per[i])
        alone = _solver(s)
        try:
            alone.step(s.steps)
            assert per[i]["particle_steps"] == alone.counts()["particle_steps"]
            for k, v in outs[i].items():
                np.testing.assert_array_equal(v, alone.read(k),
                                              err_msg=f"request {i}: {k}")
        finally:
            alone.close()


def test_binding_cap_parity():
    sc = scenes.c2_two_scale(cap_high=2048)
    _say(sc.name, "cap_high 2048")
    sol = _solver(sc)
    try:
        rep = []
        check_step(sol, sc.cfg, sc.bnd_xy, rep)
        c = sol.counts()
        # more band candidates than the cap admits: the first 512 in sorted
        # order split, and the cap is full
        assert c["n_split"] == 512 and c["n_high"] == 2048, c
        for _ in range(3):
            sol.step(4)
            check_step(sol, sc.cfg, sc.bnd_xy, rep)
            assert sol.counts()["n_high"] <= 2048
        print(rep, flush=True)
    finally:
        sol.close()


def test_binding_clamp_parity():
    """Children pulled beyond the clamp radius of their parent: the step
    brings them back (R/WCSPHSolver.cu:1007-1022, FP32 r / dist), bit-exact
    with the oracle."""
    from twoscalesplit_amd import MODE_REFERENCE
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_REFERENCE, split_x=0.08)
    _say(sc.name, "children beyond the clamp radius")
    sol = _solver(sc)
    try:
        sol.step(2)
        groups = sol.read("LIST_G")
        assert len(groups) >= 8
        ph, pl, rho = sol.read("POS_H"), sol.read("POS_L"), sol.read("RHO_L")
        m = sc.cfg.mass_low
        for q, p in enumerate(groups[:8]):
            rmax = 0.75 * np.sqrt(m / (float(rho[p]) * np.pi))
            k = 4 * int(p) + (q % 4)
            d = np.float32([np.cos(q), np.sin(q)])
            ph[k] = (pl[p] + np.float32(1.5 * rmax) * d).astype(np.float32)
        sol.write("POS_H", ph)
        check_step(sol, sc.cfg, sc.bnd_xy, [])
    finally:
        sol.close()


def test_invalid_list_writes_are_refused():
    sc = scenes.small_two_scale(nx=16, ny=24)
    _say(sc.name, "host list writes")
    sol = _solver(sc)
    try:
        sol.step(2)
        good = {k: sol.read(k) for k in ("LIST_L", "LIST_H", "LIST_G")}
        cnt = sol.counts()
        bad = good["LIST_L"].copy()
        bad[1] = bad[0]                  # a repeated low id
        sol.write("LIST_L", bad)
        with pytest.raises(RuntimeError, match="invalid low list"):
            sol.step(1)
        assert sol.counts()["step"] == cnt["step"]
        # a listed child of an unlisted group
        sol.write("LIST_L", good["LIST_L"])
        if cnt["n_groups"] > 0:
            sol.set_counts(cnt["n_low"], cnt["n_high"], cnt["n_groups"] - 1)
            with pytest.raises(RuntimeError, match="child of an unlisted group"):
                sol.step_graph(2)
        # the valid state is accepted again and steps
        sol.set_counts(cnt["n_low"], cnt["n_high"], cnt["n_groups"])
        sol.step(1)
        assert sol.counts()["step"] == cnt["step"] + 1
    finally:
        sol.close()
