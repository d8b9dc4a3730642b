"""Relaxation of newly injected children (opt-in, relax_iters > 0) on the
GPU against the FP64 oracle, and its determinism:

* one step after a split step, from the same state: lists, states and ages
  bit-exact; densities within 1e-5; positions and velocities within the
  bound that the relaxation's own accelerations (1e-5 relative, like the
  force check) propagate over relax_iters sub-steps plus the main step;
* graph-replayed steps equal eager steps; a scene's results are the same
  alone and in a batch with a scene that does not relax, also after that
  other scene is retired.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import binding as ob
from twoscalesplit_amd import MODE_BAND, MODE_REFERENCE, Batch, Solver, scenes

pytestmark = pytest.mark.gpu

FIELDS = ("POS_L", "VEL_L", "RHO_L", "P_L", "STATE_L", "POS_H", "VEL_H",
          "RHO_H", "STATE_H", "LIST_L", "LIST_H", "LIST_G", "DT")


def _scene(iters=4):
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_REFERENCE,
                                split_x=0.06)
    sc.cfg.relax_iters = iters
    sc.cfg.relax_dt_frac = 0.1
    return sc


def test_relaxation_matches_oracle():
    sc = _scene()
    print("[case] relaxation parity", sc.name, flush=True)
    sol = Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    try:
        sol.step(1)                      # splits: children to relax
        pre = sol.snapshot()
        nsplit = pre["counts"]["n_split"]
        assert nsplit > 0
        sol.step(1)                      # relaxes them, then a full step
        post = sol.snapshot()
    finally:
        sol.close()
    o = ob.Scene(ob.params_from_config(sc.cfg), sc.cap_low, sc.bnd_xy)
    o.load_snapshot(pre)
    assert o.s.n_fresh == 4 * nsplit
    o.step(64, 0)
    nl, nh, ng = o.n
    c = post["counts"]
    assert (nl, nh, ng) == (c["n_low"], c["n_high"], c["n_groups"])
    np.testing.assert_array_equal(post["LIST_L"], o.a["list_l"][:nl])
    np.testing.assert_array_equal(post["LIST_H"], o.a["list_h"][:nh])
    np.testing.assert_array_equal(post["LIST_G"], o.a["list_g"][:ng])
    np.testing.assert_array_equal(post["STATE_L"], o.a["state_l"])
    lh = post["LIST_H"]
    np.testing.assert_array_equal(post["STATE_H"][lh], o.a["state_h"][lh])
    for R, key, ids in (("RHO_L", "rho_l", post["LIST_L"]), ("RHO_H", "rho_h", lh)):
        err = np.abs(post[R][ids] / o.a[key][ids] - 1)
        assert err.max() <= 1e-5, (R, err.max())
    dt = float(pre["DT"][0])
    for P, V, kp, kv, ids in (("POS_H", "VEL_H", "pos_h", "vel_h", lh),
                              ("POS_L", "VEL_L", "pos_l", "vel_l", post["LIST_L"])):
        dv = np.abs(post[V][ids].astype(np.float64) - o.a[kv][ids])
        dx = np.abs(post[P][ids].astype(np.float64) - o.a[kp][ids])
        vscale = np.abs(o.a[kv][ids]).max() + 1.0
        print("[relax]", P, "max |dx|", dx.max(), "max |dv|", dv.max(), flush=True)
        assert dv.max() <= 1e-4 * vscale
        assert dx.max() <= 2e-4 * vscale * dt + 4e-7
    # the relaxation moved the new children
    ref = Solver(_scene(0).cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    try:
        ref.step(2)
        moved = np.abs(ref.read("POS_H")[lh] - post["POS_H"][lh]).max()
    finally:
        ref.close()
    assert moved > 0


def test_relaxation_graph_and_batch_determinism():
    sc = _scene(3)
    other = scenes.small_two_scale(nx=20, ny=20, mode=MODE_BAND)
    print("[case] relaxation determinism", flush=True)
    a = Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    b = Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    scs = [other, sc]
    cap = sum(s.cap_low for s in scs)
    nb = sum(len(s.bnd_xy) for s in scs)
    cl = sum(s.cells[0] for s in scs)
    ch = sum(s.cells[1] for s in scs)
    bt = Batch(2, cap, nb, cl, ch)
    try:
        a.step(12)
        b.step(2)
        b.step_graph(10)
        ids = [bt.add_scene(s.cfg, s.low_xy, s.cap_low, s.bnd_xy) for s in scs]
        bt.step(2)
        bt.step_graph(6)
        bt.remove_scene(ids[0])
        bt.step(4)
        for k in FIELDS:
            wa = a.read(k)
            np.testing.assert_array_equal(b.read(k), wa, err_msg=f"graph {k}")
            np.testing.assert_array_equal(bt.read(ids[1], k), wa, err_msg=f"batch {k}")
    finally:
        a.close()
        b.close()
        bt.close()
