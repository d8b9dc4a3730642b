"""GPU parity tests: the HIP path (through the C ABI) against the oracle.

Each case prints its name, scene and step before touching the GPU.
"""
from __future__ import annotations

import json
import math
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from parity import check_pairs, check_step
from twoscalesplit_amd import MODE_BAND, MODE_REFERENCE, Batch, Solver
from twoscalesplit_amd import scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _solver(sc):
    return Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)


def _run_checks(sc, steps_between, report):
    sol = _solver(sc)
    try:
        for k, adv in enumerate(steps_between):
            if adv:
                _say(sc.name, "advance", adv)
                sol.step(adv)
            _say(sc.name, "check step", sol.counts()["step"])
            check_step(sol, sc.cfg, sc.bnd_xy, report)
    finally:
        sol.close()


def test_c1_plain_parity():
    sc = scenes.c1_plain()
    rep = []
    _run_checks(sc, [0, 0, 20, 80], rep)
    print(json.dumps(rep))


def test_c2_two_scale_parity():
    sc = scenes.c2_two_scale()
    rep = []
    _run_checks(sc, [0, 0, 0, 30], rep)
    print(json.dumps(rep))
    assert rep[0]["events"][0] > 0, "the first step must split the surface band"


def test_reference_rule_parity():
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_REFERENCE,
                                split_x=0.08)
    rep = []
    _run_checks(sc, [0, 0, 0, 10, 10], rep)
    print(json.dumps(rep))
    assert sum(r["events"][0] for r in rep) > 0


def test_crowded_targets_parity():
    """Half of the block squeezed to 0.6 dx spacing: up to ~130 same-level
    neighbours within h there and ~50 in the other half, in the same waves
    (long and short candidate runs, full LDS buffers); split to populate
    the high level."""
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_REFERENCE,
                                split_x=0.02)
    xy = sc.low_xy.astype(np.float64)
    left = xy[:, 0] < np.median(xy[:, 0])
    x0, y0 = xy[left, 0].min(), xy[left, 1].min()
    xy[left, 0] = x0 + 0.6 * (xy[left, 0] - x0)
    xy[left, 1] = y0 + 0.6 * (xy[left, 1] - y0)
    sc.low_xy = xy.astype(np.float32)
    rep = []
    _run_checks(sc, [0, 0, 0], rep)
    print(json.dumps(rep))


def _four_span(o, R, dims, near):
    """A float32 coordinate whose +-R box spans four cells of size R."""
    from test_oracle_pins import _boundary_positions, _spans
    y = _boundary_positions(np.float32(o), np.float32(R), dims)
    n = _spans(y, np.float32(R), np.float32(o), dims)
    c = y[n == 4]
    return np.float32(c[np.argmin(np.abs(c - near))])


def test_four_row_box_parity():
    """Targets whose scan box spans 4 x 4 cells (FP32 rounding at two cell
    edges), on both levels: four rows of four cells are scanned."""
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_REFERENCE,
                                split_x=0.08)
    cfg = sc.cfg
    h, H = cfg.effective_radius, cfg.effective_radius_high
    xl = _four_span(cfg.domain_origin[0], h, cfg.domain_dims[0], 0.15)
    yl = _four_span(cfg.domain_origin[1], h, cfg.domain_dims[1], 0.15)
    low = sc.low_xy.copy()
    k = int(np.argmin(np.hypot(low[:, 0] - xl, low[:, 1] - yl)))
    low[k] = (xl, yl)
    sc.low_xy = low
    _say("four-row boxes: low particle", k, "at", float(xl), float(yl))
    sol = _solver(sc)
    try:
        check_step(sol, cfg, sc.bnd_xy, [])
        sol.step(1)
        xh = _four_span(cfg.domain_origin[0], H, cfg.domain_dims_high[0], 0.1)
        yh = _four_span(cfg.domain_origin[1], H, cfg.domain_dims_high[1], 0.1)
        lh = sol.read("LIST_H")
        ph = sol.read("POS_H")
        a = int(lh[np.argmin(np.hypot(ph[lh, 0] - xh, ph[lh, 1] - yh))])
        ph[a] = (xh, yh)
        sol.write("POS_H", ph)
        _say("four-row boxes: high particle", a, "at", float(xh), float(yh))
        check_step(sol, cfg, sc.bnd_xy, [])
    finally:
        sol.close()


def test_promote_merge_dissolve_parity():
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_BAND)
    sc.cfg.promote_min_age = 1
    sc.cfg.promote_max_age = 3
    sc.cfg.promote_fluc = 1.0
    rep = []
    _run_checks(sc, [0, 0, 0, 0, 0, 0, 10, 0, 0], rep)
    print(json.dumps(rep))
    ev = np.array([r["events"] for r in rep]).sum(axis=0)
    assert ev[0] > 0 and ev[1] > 0, ev


def test_work_counts_match_the_kernels_pairs():
    # tss_work_counts (the flop counts of the roofline) against the pairs the
    # density and force kernels themselves add (tss_pair_sets, their
    # recording variant) on the same states: density pairs (same + cross
    # level), force pairs and wall candidates per level, targets; the
    # density's radius-H cross-level box scans no more candidates than the
    # force's radius-h box and at least the pairs it adds
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_BAND)
    sc.cfg.promote_min_age = 1
    sc.cfg.promote_max_age = 3
    sc.cfg.promote_fluc = 1.0
    sol = _solver(sc)
    try:
        seen_hl = 0
        for adv in (1, 3, 4):
            sol.step(adv)
            _say(sc.name, "work counts at step", sol.counts()["step"])
            w = sol.work_counts()
            cl, _ = sol.pair_sets(0)
            ch, _ = sol.pair_sets(1)
            assert w["dens_pairs_l"] == int(cl[:, 0].sum()) + int(cl[:, 1].sum())
            assert w["dens_pairs_h"] == int(ch[:, 0].sum()) + int(ch[:, 1].sum())
            assert w["force_pairs_l"] == int(cl[:, 2].sum()) + int(cl[:, 3].sum())
            assert w["force_pairs_h"] == int(ch[:, 2].sum()) + int(ch[:, 3].sum())
            assert w["wall_l"] == int(cl[:, 4].sum())
            assert w["wall_h"] == int(ch[:, 4].sum())
            assert w["targets_l"] == len(cl) and w["targets_h"] == len(ch)
            assert int(ch[:, 1].sum()) <= w["cand_hl_density"] <= w["cand_hl"]
            seen_hl += int(ch[:, 1].sum())
        assert seen_hl > 0, "no cross-level density pair of a high target"
    finally:
        sol.close()


def test_wide_high_radius_parity():
    # H = 0.6 h: a high target's own pairs reach past 2r/h = 1, so the high
    # passes take the general blended kernel (H = h/2 scenes take W_h's
    # inner branch)
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_BAND, h_ratio=0.6)
    rep = []
    _run_checks(sc, [0, 0, 0, 5], rep)
    print(json.dumps(rep))
    assert rep[0]["events"][0] > 0, "the first step must split the surface band"


@pytest.mark.parametrize("h_ratio", [0.5, 0.6])
def test_tension_parity(h_ratio):
    # sigma != 0: the force passes add -sigma x_ij W_bar per pair (the
    # tension variants of the mask walk and of the scans), on both kernels
    # of the high level (inner branch at H = h/2, general at H = 0.6 h)
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_BAND, h_ratio=h_ratio)
    sc.cfg.tension_coefficient = 0.3
    rep = []
    _run_checks(sc, [0, 0, 0, 5], rep)
    print(json.dumps(rep))
    assert rep[0]["events"][0] > 0, "the first step must split the surface band"


def test_cfl_dt_parity():
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_BAND, dt_cfl=True)
    rep = []
    _run_checks(sc, [0, 0, 5], rep)
    print(json.dumps(rep))


def test_neighbour_sets_bit_exact(oracle):
    sc = scenes.c2_two_scale()
    _say(sc.name, "neighbour sets after 3 steps")
    sol = _solver(sc)
    try:
        sol.step(3)
        snap = sol.snapshot()
        o = oracle.Scene(oracle.params_from_config(sc.cfg), sc.cap_low,
                         sc.bnd_xy)
        o.load_snapshot(snap)
        g = o.grid()
        for level in (0, 1):
            c_gpu, s_gpu = sol.neighbors(level)
            c_orc, s_orc = o.neighbors(g, level)
            assert len(c_gpu) > 0
            np.testing.assert_array_equal(c_gpu, c_orc)
            np.testing.assert_array_equal(s_gpu, s_orc)
        # the sets the hot-path kernels themselves add (state weights on)
        nl, nh = check_pairs(sol, o, g)
        assert nl > 0 and nh > 0
    finally:
        sol.close()


def test_conservation_fp64_bound():
    sc = scenes.c2_two_scale()
    _say(sc.name, "conservation sums")
    sol = _solver(sc)
    try:
        sol.step(5)
        pre = sol.snapshot()
        sol.step(1)
        post = sol.snapshot()
        sums, bound = sol.conservation()
        sl = sol.read("SORTED_L")
        sh = sol.read("SORTED_H")
        c = sc.cfg
        terms = [[] for _ in range(5)]
        for ids, lvl in ((sl, 0), (sh, 1)):
            m = float(c.mass_high if lvl else c.mass_low)
            P = post["POS_H" if lvl else "POS_L"][ids].astype(np.float64)
            V = post["VEL_H" if lvl else "VEL_L"][ids].astype(np.float64)
            if lvl:
                real = pre["STATE_H"][ids] == 0
                P, V = P[real], V[real]
            terms[0] += [m] * len(P)
            terms[1] += list(m * V[:, 0])
            terms[2] += list(m * V[:, 1])
            terms[3] += list(0.5 * (m * (V[:, 0] ** 2 + V[:, 1] ** 2)))
            terms[4] += list((m * float(c.gravity)) * P[:, 1])
        for q in range(5):
            exact = math.fsum(terms[q])
            assert abs(sums[q] - exact) <= bound[q] + 1e-300, (q, sums[q],
                                                                exact, bound[q])
    finally:
        sol.close()


def test_graph_replay_equals_eager():
    sc = scenes.c2_two_scale()
    _say(sc.name, "graph replay vs eager, 10 steps")
    a, b = _solver(sc), _solver(sc)
    try:
        a.step(10)
        b.step_graph(10)
        sa, sb = a.snapshot(), b.snapshot()
        for k in sa:
            if k == "counts":
                assert sa[k] == sb[k]
            else:
                np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
        # replays interleaved with eager steps and reads stay identical
        a.step(4)
        b.step_graph(2)
        b.step(2)
        for k in ("POS_L", "POS_H", "RHO_L", "LIST_L", "LIST_H"):
            np.testing.assert_array_equal(a.read(k), b.read(k), err_msg=k)
    finally:
        a.close()
        b.close()


def _batch_for(scs):
    cap = sum(s.cap_low for s in scs)
    nb = sum(len(s.bnd_xy) for s in scs)
    cl = sum(s.cells[0] for s in scs)
    ch = sum(s.cells[1] for s in scs)
    return Batch(len(scs) + 1, cap + 100, nb + 10, cl + 10, ch + 10)


def test_batch_alone_equals_mix():
    # the batch holds a band-rule scene, so its reference-rule scenes run the
    # group pass with a lane per child; alone they run a lane per group
    scs = [scenes.small_two_scale(nx=16, ny=24, mode=MODE_BAND),
           scenes.c2_two_scale(),
           scenes.small_two_scale(nx=20, ny=12, mode=MODE_REFERENCE,
                                  split_x=0.05)]
    scs[0].cfg.promote_min_age = 1
    scs[0].cfg.promote_max_age = 3
    _say("batch of", [s.name for s in scs], "vs alone, 12 steps")
    b = _batch_for(scs)
    try:
        ids = [b.add_scene(s.cfg, s.low_xy, s.cap_low, s.bnd_xy) for s in scs]
        b.step(5)
        b.step_graph(7)
        for s, sid in zip(scs, ids):
            alone = _solver(s)
            alone.step(12)
            for k in ("POS_L", "VEL_L", "RHO_L", "P_L", "ACC_L", "STATE_L",
                      "POS_H", "VEL_H", "RHO_H", "STATE_H", "LIST_L",
                      "LIST_H", "LIST_G", "DT"):
                np.testing.assert_array_equal(b.read(sid, k), alone.read(k),
                                              err_msg=f"{s.name}:{k}")
            cb, _ = b.conservation(sid)
            ca, _ = alone.conservation()
            np.testing.assert_array_equal(cb, ca)
            alone.close()
        # retire the middle scene; the others keep advancing unchanged
        b.remove_scene(ids[1])
        b.step(3)
        for s, sid in ((scs[0], ids[0]), (scs[2], ids[2])):
            alone = _solver(s)
            alone.step(15)
            np.testing.assert_array_equal(b.read(sid, "POS_L"),
                                          alone.read("POS_L"))
            alone.close()
    finally:
        b.close()


def test_facade_demo_runs():
    exe = os.path.join(os.path.dirname(os.path.dirname(__file__)),
                       "twoscalesplit_amd", "lib", "facade_demo")
    _say("facade demo (reference class API), 2000 steps")
    state = os.path.join(tempfile.mkdtemp(), "demo.tsss")
    out = subprocess.run([exe, "2000", state], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    assert res["finite"] == 1 and res["n_low"] == 101 * 151
    # SaveState / LoadState continue the run bit for bit
    assert res["resume_equal"] == 1
    # the collapsing column crosses x = 0.5, where the reference rule splits
    assert res["n_high"] > 0 and res["n_high"] % 4 == 0
    assert 900.0 < res["rho_mean"] < 1100.0
