"""Launch shapes of the neighbour kernels (block size, LDS buffer depth,
XCD block order, persistent grid cap) and the native autotuner: no shape
may change a result, so every tuned run must equal the default run
bit-for-bit, and the steps the autotuner takes are ordinary steps.
"""
from __future__ import annotations

import numpy as np
import pytest

from twoscalesplit_amd import MODE_BAND, Batch, Solver, scenes
from twoscalesplit_amd import _native as N

pytestmark = pytest.mark.gpu

FIELDS = ("POS_L", "VEL_L", "RHO_L", "P_L", "ACC_L", "STATE_L", "AGE_L",
          "POS_H", "VEL_H", "RHO_H", "P_H", "ACC_H", "STATE_H", "LIST_L",
          "LIST_H", "LIST_G", "DT")


def _scene():
    sc = scenes.c2_two_scale()
    sc.cfg.promote_min_age = 2
    sc.cfg.promote_max_age = 6
    return sc


def _run(sc, tuning, steps):
    sol = Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    try:
        if tuning is not None:
            sol.set_tuning(tuning)
        sol.step(3)
        sol.step_graph(steps - 3)
        return {k: sol.read(k) for k in FIELDS}, sol.counts()
    finally:
        sol.close()


@pytest.mark.parametrize("shape", [
    (128, 8, 0, 0), (512, 32, 0, 0), (256, 16, 1, 0), (128, 32, 1, 96),
    (512, 8, 0, 64)])
def test_shapes_do_not_change_results(shape):
    sc = _scene()
    print("[case] tuning shape", shape, flush=True)
    want, cw = _run(sc, None, 16)
    got, cg = _run(sc, N.TssTuning.uniform(*shape), 16)
    assert cw == cg
    for k in FIELDS:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


def test_mixed_shapes_and_invalid_shape_rejected():
    sc = _scene()
    t = N.TssTuning.uniform()
    for i, (b, dp, x, g) in enumerate([(128, 32, 1, 0), (512, 8, 0, 32),
                                       (256, 8, 1, 128), (128, 16, 0, 0)]):
        t.block[i], t.depth[i], t.xcd_order[i], t.grid[i] = b, dp, x, g
    want, _ = _run(sc, None, 10)
    got, _ = _run(sc, t, 10)
    for k in FIELDS:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    sol = Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    try:
        bad = N.TssTuning.uniform(block=96)
        with pytest.raises(RuntimeError, match="block"):
            sol.set_tuning(bad)
        assert sol.get_tuning().block[0] == 256
    finally:
        sol.close()


def test_autotune_steps_are_ordinary_steps():
    scs = [scenes.small_two_scale(nx=24, ny=32, mode=MODE_BAND),
           scenes.c2_two_scale()]
    cap = sum(s.cap_low for s in scs)
    nb = sum(len(s.bnd_xy) for s in scs)
    cl = sum(s.cells[0] for s in scs)
    ch = sum(s.cells[1] for s in scs)
    print("[case] autotune batch of 2", flush=True)
    a = Batch(2, cap, nb, cl, ch)
    b = Batch(2, cap, nb, cl, ch)
    try:
        ia = [a.add_scene(s.cfg, s.low_xy, s.cap_low, s.bnd_xy) for s in scs]
        ib = [b.add_scene(s.cfg, s.low_xy, s.cap_low, s.bnd_xy) for s in scs]
        res = a.autotune(steps_per_trial=1, rounds=1)
        taken = res["steps_taken"]
        assert taken == 1 + 9 + 5
        assert len(res["trials"]) == taken - 1
        assert all(row[4 + k] > 0 for row in res["trials"] for k in range(4))
        a.step_graph(4)
        b.step(taken)
        b.step_graph(4)
        for sa, sb in zip(ia, ib):
            assert a.counts(sa) == b.counts(sb)
            for k in FIELDS:
                np.testing.assert_array_equal(a.read(sa, k), b.read(sb, k),
                                              err_msg=k)
        t = a.get_tuning().as_dict()
        print("[tuning]", t, flush=True)
        for k in t.values():
            assert k["block"] in (128, 256, 512) and k["depth"] in (8, 16, 32)
    finally:
        a.close()
        b.close()
