"""The grid build's sort paths -- the one-workgroup radix sort, the
multi-tile radix sort and the counting sort by cell -- give the same sorted
order and cell ranges, so the same steps bit for bit. The counting sort also
passes the oracle check (grid, ranges, lists and states bit-exact; density,
pressure and acceleration within tolerance) on scenes with particles that
change cell, entries appended to the lists (splits, merges), cells holding
more entries than a wave has lanes, and a batch whose cells have a gap left
by a removed scene.

Each case prints its name before touching the GPU.
"""
from __future__ import annotations

import numpy as np
import pytest

from parity import check_step
from twoscalesplit_amd import MODE_BAND, MODE_REFERENCE, Batch, Solver, scenes
from twoscalesplit_amd import _native as N

pytestmark = pytest.mark.gpu

PATHS = {"small": N.SORT_SMALL, "radix": N.SORT_RADIX, "count": N.SORT_COUNT}
FIELDS = ("POS_L", "VEL_L", "ACC_L", "RHO_L", "P_L", "STATE_L", "POS_H",
          "VEL_H", "RHO_H", "STATE_H", "LIST_L", "LIST_H", "LIST_G",
          "SORTED_L", "SORTED_H", "KEYS_L", "KEYS_H", "CELL_START_L",
          "CELL_END_L", "CELL_START_H", "CELL_END_H")


def _say(*a):
    print("[case]", *a, flush=True)


def _solver(sc, path):
    s = Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    s.set_grid_sort(path)
    return s


def _band_scene():
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_BAND)
    sc.cfg.promote_min_age = 1
    sc.cfg.promote_max_age = 3
    return sc


def _same(a, b, what):
    for k in FIELDS:
        np.testing.assert_array_equal(a.read(k), b.read(k), err_msg=f"{what}: {k}")


@pytest.mark.parametrize("name", ["c2", "band"])
def test_sort_paths_give_the_same_steps(name):
    sc = scenes.c2_two_scale() if name == "c2" else _band_scene()
    runs = {}
    try:
        for p, path in PATHS.items():
            _say(sc.name, "sort path", p, "8 eager + 8 graph steps")
            s = _solver(sc, path)
            runs[p] = s
            s.step(8)
            s.step_graph(8)
        ref = runs["small"]
        assert ref.counts()["n_high"] > 0
        for p in ("radix", "count"):
            assert runs[p].counts() == ref.counts(), p
            _same(runs[p], ref, p)
    finally:
        for s in runs.values():
            s.close()


def test_count_sort_oracle_checks():
    sc = _band_scene()
    _say(sc.name, "counting sort, oracle checks over 20 steps")
    s = _solver(sc, N.SORT_COUNT)
    try:
        rep = []
        for _ in range(5):
            check_step(s, sc.cfg, sc.bnd_xy, rep)
            s.step(4)
        assert sum(r["events"][0] for r in rep) > 0
        print(rep, flush=True)
    finally:
        s.close()


def test_count_sort_cells_wider_than_a_wave():
    """Half of the block squeezed to 0.45 dx spacing: about 80 particles per
    low cell, so a cell's sorted range spans more than one wave's slots."""
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_REFERENCE,
                                split_x=0.02)
    xy = sc.low_xy.astype(np.float64)
    left = xy[:, 0] < np.median(xy[:, 0])
    x0, y0 = xy[left, 0].min(), xy[left, 1].min()
    xy[left, 0] = x0 + 0.45 * (xy[left, 0] - x0)
    xy[left, 1] = y0 + 0.45 * (xy[left, 1] - y0)
    sc.low_xy = xy.astype(np.float32)
    _say(sc.name, "counting sort, cells of >64 entries")
    s = _solver(sc, N.SORT_COUNT)
    r = _solver(sc, N.SORT_RADIX)
    try:
        check_step(s, sc.cfg, sc.bnd_xy, [])
        # the cell table of that step's grid
        st, en = s.read("CELL_START_L"), s.read("CELL_END_L")
        assert int((en - st)[st >= 0].max()) > 64
        r.step(1)
        _same(s, r, "count vs radix")
    finally:
        s.close()
        r.close()


def test_count_sort_batch_with_gap():
    """Three scenes, the middle one removed after two steps: its cells are a
    gap below the last resident scene's cells. A fourth scene added later
    takes the gap. Every resident scene equals the same scene run alone."""
    scs = [scenes.small_two_scale(nx=16, ny=24, mode=MODE_BAND),
           scenes.small_two_scale(nx=20, ny=20, mode=MODE_REFERENCE, split_x=0.05),
           scenes.small_two_scale(nx=24, ny=16, mode=MODE_REFERENCE, split_x=0.08),
           scenes.small_two_scale(nx=12, ny=12, mode=MODE_BAND)]
    cap = sum(s.cap_low for s in scs)
    b = Batch(4, cap, sum(len(s.bnd_xy) for s in scs),
              sum(s.cells[0] for s in scs), sum(s.cells[1] for s in scs))
    keys = ("POS_L", "VEL_L", "RHO_L", "P_L", "STATE_L", "POS_H", "VEL_H",
            "RHO_H", "LIST_L", "LIST_H", "LIST_G")

    def check(pairs):
        for i, steps in pairs:
            _say("batch scene", i, "vs alone,", steps, "steps")
            a = _solver(scs[i], N.SORT_RADIX)
            try:
                a.step(steps)
                for k in keys:
                    np.testing.assert_array_equal(b.read(ids[i], k), a.read(k),
                                                  err_msg=f"scene {i}: {k}")
            finally:
                a.close()

    try:
        b.set_grid_sort(N.SORT_COUNT)
        _say("batch with a gap: 3 scenes, the middle one removed")
        ids = {i: b.add_scene(s.cfg, s.low_xy, s.cap_low, s.bnd_xy)
               for i, s in enumerate(scs[:3])}
        b.step(2)
        b.remove_scene(ids.pop(1))
        b.step(3)
        b.step_graph(4)
        b.synchronize()
        check(((0, 9), (2, 9)))
        _say("batch: a fourth scene in the gap")
        ids[3] = b.add_scene(scs[3].cfg, scs[3].low_xy, scs[3].cap_low, scs[3].bnd_xy)
        b.step(2)
        b.synchronize()
        check(((0, 11), (2, 11), (3, 2)))
    finally:
        b.close()
