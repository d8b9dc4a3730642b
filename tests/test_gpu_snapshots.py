"""Scene snapshots on the GPU: a run snapshotted mid-way and resumed --
in a new solver, from a file, or in a batch that holds other scenes (with
children of its own awaiting relaxation) -- ends bit for bit where the
uninterrupted run ends, counters included; the snapshot itself reads back
as the scene's fields.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from twoscalesplit_amd import MODE_BAND, MODE_REFERENCE, Batch, Solver, scenes
from twoscalesplit_amd import snapshots as S

pytestmark = pytest.mark.gpu

FIELDS = ("POS_L", "VEL_L", "ACC_L", "RHO_L", "P_L", "STATE_L", "AGE_L",
          "VIS_L", "POS_H", "VEL_H", "ACC_H", "RHO_H", "P_H", "STATE_H",
          "VIS_H", "LIST_L", "LIST_H", "LIST_G", "DT")
COUNTS = ("n_low", "n_high", "n_groups", "step", "particle_steps",
          "n_split", "n_merge", "n_promote", "n_dissolve")


def _band():
    sc = scenes.c2_two_scale()
    sc.cfg.promote_min_age = 2
    sc.cfg.promote_max_age = 6
    return sc


def _relaxing():
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_REFERENCE,
                                split_x=0.06)
    sc.cfg.relax_iters = 3
    sc.cfg.relax_dt_frac = 0.1
    return sc


def _state(read, counts):
    cnt = counts()
    return {k: read(k, cnt) for k in FIELDS}, {k: cnt[k] for k in COUNTS}


def _same(a, b):
    fa, ca = a
    fb, cb = b
    assert ca == cb
    for k in FIELDS:
        np.testing.assert_array_equal(fb[k], fa[k], err_msg=k)


@pytest.mark.parametrize("which,first,rest", [("band", 12, 18),
                                              ("relax", 1, 7)])
def test_resume_equals_uninterrupted(which, first, rest, tmp_path):
    sc = _band() if which == "band" else _relaxing()
    print("[case] snapshot resume", which, sc.name, first, rest, flush=True)
    with Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy) as ref:
        ref.step(first + rest)
        want = _state(ref.read, ref.counts)
    with Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy) as a:
        a.step(first)
        cnt = a.counts()
        if which == "relax":
            assert cnt["n_split"] > 0          # children await relaxation
        snap = a.save_state()
        hd, arr = S.parse(snap)
        assert (hd.n_low, hd.n_high, hd.step) == (cnt["n_low"], cnt["n_high"],
                                                 cnt["step"])
        if which == "relax":
            assert hd.n_fresh == 4 * cnt["n_split"]
        for k in ("POS_L", "RHO_H", "STATE_L", "LIST_H", "LIST_G"):
            np.testing.assert_array_equal(arr[k], a.read(k, cnt), err_msg=k)
        path = os.path.join(tmp_path, "scene.tsss")
        a.save(path)
    for src in (snap, path):
        with Solver.from_snapshot(src) as b:
            b.step(2)
            b.step_graph(rest - 2)
            _same(want, _state(b.read, b.counts))


def test_restore_into_a_busy_batch():
    """The restored scene's lists are spliced into lists that hold other
    scenes, one of them with fresh children of its own."""
    sc = _relaxing()
    busy = [scenes.small_two_scale(nx=20, ny=28, mode=MODE_REFERENCE,
                                   split_x=0.05),
            scenes.small_two_scale(nx=16, ny=24, mode=MODE_BAND)]
    busy[0].cfg.relax_iters = 2
    busy[0].cfg.relax_dt_frac = 0.1
    print("[case] restore into a busy batch", flush=True)
    with Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy) as ref:
        ref.step(9)
        want = _state(ref.read, ref.counts)
    with Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy) as a:
        a.step(1)
        snap = a.save_state()
    others = []
    for s in busy:
        with Solver(s.cfg, s.low_xy, s.cap_low, s.bnd_xy) as o:
            o.step(2)
            others.append(_state(o.read, o.counts))
    allsc = busy + [sc]
    cap = sum(s.cap_low for s in allsc)
    nb = sum(len(s.bnd_xy) for s in allsc)
    cl = sum(s.cells[0] for s in allsc)
    ch = sum(s.cells[1] for s in allsc)
    with Batch(3, cap, nb, cl, ch) as bt:
        ids = [bt.add_scene(s.cfg, s.low_xy, s.cap_low, s.bnd_xy) for s in busy]
        bt.step(1)
        assert bt.counts(ids[0])["n_split"] > 0   # fresh children in the tail
        sid = bt.restore(snap)
        bt.step(1)
        for i, o in zip(ids, others):
            _same(o, _state(lambda k, c: bt.read(i, k, c),
                            lambda: bt.counts(i)))
        bt.step(3)
        bt.step_graph(4)
        _same(want, _state(lambda k, c: bt.read(sid, k, c),
                           lambda: bt.counts(sid)))
        # and the batch snapshot of that scene restores alone again
        again = bt.snapshot(sid)
    with Solver.from_snapshot(again) as c:
        got = _state(c.read, c.counts)
    _same(want, got)
