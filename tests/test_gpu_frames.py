"""Frame export on the GPU: every section of a frame equals the read-back
fields at the ids of the current lists, in list order, bit for bit -- for a
single scene, for each scene of a batch whose lists interleave several
scenes (also after one is retired), and at full size, where the selection
spans thousands of tiles. Frames written to a file read back unchanged.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from twoscalesplit_amd import MODE_BAND, MODE_REFERENCE, Batch, Solver, scenes
from twoscalesplit_amd import frames as F

pytestmark = pytest.mark.gpu


def _expect(read, cnt):
    """The frame a handle should produce, from its field read-backs."""
    ll, lh = read("LIST_L"), read("LIST_H")
    out = {}
    for tag, ids in (("L", ll), ("H", lh)):
        out[tag] = {"id": ids,
                    "pos": read("POS_" + tag)[ids],
                    "vel": read("VEL_" + tag)[ids],
                    "rho": read("RHO_" + tag)[ids],
                    "p": read("P_" + tag)[ids],
                    "vis": read("VIS_" + tag)[ids],
                    "state": read("STATE_" + tag)[ids]}
    return out, read("POS_B"), read("DT")[0]


def _check(fr, want, bnd, dt, cnt):
    assert fr.step == cnt["step"]
    assert np.float32(fr.dt) == dt
    assert (fr.n_low, fr.n_high) == (cnt["n_low"], cnt["n_high"])
    for lv, tag in ((fr.low, "L"), (fr.high, "H")):
        for k, v in want[tag].items():
            np.testing.assert_array_equal(lv[k], v, err_msg=f"{tag} {k}")
    np.testing.assert_array_equal(fr.bnd, bnd)


def test_solver_frame_equals_fields():
    sc = scenes.c2_two_scale()
    print("[case] frame", sc.name, flush=True)
    with Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy) as sol:
        sol.step(4)
        sol.step_graph(40)
        cnt = sol.counts()
        assert cnt["n_high"] > 0
        fr = sol.frame(F.ALL)
        want, bnd, dt = _expect(lambda k: sol.read(k, cnt), cnt)
        _check(fr, want, bnd, dt, cnt)
        # field subsets select sections, not values
        sub = sol.frame(F.RENDER)
        assert sorted(sub.low) == ["id", "pos", "vis"] and sub.bnd is None
        np.testing.assert_array_equal(sub.high["pos"], fr.high["pos"])


def test_batch_frames_with_interleaved_lists():
    scs = [scenes.small_two_scale(nx=24, ny=32, mode=MODE_BAND),
           scenes.small_two_scale(nx=20, ny=28, mode=MODE_REFERENCE,
                                  split_x=0.05),
           scenes.c2_two_scale()]
    cap = sum(s.cap_low for s in scs)
    nb = sum(len(s.bnd_xy) for s in scs)
    cl = sum(s.cells[0] for s in scs)
    ch = sum(s.cells[1] for s in scs)
    print("[case] batch frames", [s.name for s in scs], flush=True)
    with Batch(3, cap, nb, cl, ch) as bt:
        ids = [bt.add_scene(s.cfg, s.low_xy, s.cap_low, s.bnd_xy) for s in scs]
        bt.step(3)
        bt.step_graph(12)
        for sid in ids:
            cnt = bt.counts(sid)
            want, bnd, dt = _expect(lambda k: bt.read(sid, k, cnt), cnt)
            _check(bt.frame(sid, F.ALL), want, bnd, dt, cnt)
        bt.remove_scene(ids[1])
        bt.step(2)
        for sid in (ids[0], ids[2]):
            cnt = bt.counts(sid)
            want, bnd, dt = _expect(lambda k: bt.read(sid, k, cnt), cnt)
            _check(bt.frame(sid, F.ALL), want, bnd, dt, cnt)


def test_record_file_and_buffer_checks(tmp_path):
    sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_REFERENCE,
                                split_x=0.06)
    path = os.path.join(tmp_path, "run.tssf")
    print("[case] record frames", flush=True)
    with Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy) as sol:
        total = F.record(sol, path, n_frames=5, every=6)
        last = sol.frame(F.RENDER)
        import ctypes as C
        from twoscalesplit_amd import _native as N
        need = C.c_size_t()
        assert N.lib().tss_frame(sol._h, F.RENDER, None, 0, C.byref(need)) == 0
        small = np.zeros(need.value - 16, np.uint8)
        rc = N.lib().tss_frame(sol._h, F.RENDER, small.ctypes.data,
                               small.nbytes, C.byref(need))
        assert rc != 0 and "too small" in N.lib().tss_last_error().decode()
    got = list(F.read_frames(path))
    assert [g.step for g in got] == [0, 6, 12, 18, 24]
    assert os.path.getsize(path) == total
    for k in last.low:
        np.testing.assert_array_equal(got[-1].low[k], last.low[k])
        np.testing.assert_array_equal(got[-1].high[k], last.high[k])
    assert got[0].n_high == 0 and got[-1].n_high > 0


def test_full_size_frame():
    sc = scenes.c3_large(forced=True)
    print("[case] full-size frame", sc.n_low, flush=True)
    with Solver(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy) as sol:
        sol.step(2)
        cnt = sol.counts()
        fr = sol.frame(F.ID | F.POS | F.RHO | F.VIS | F.STATE)
        for tag, lv in (("L", fr.low), ("H", fr.high)):
            ids = sol.read("LIST_" + tag, cnt)
            np.testing.assert_array_equal(lv["id"], ids)
            for k in ("POS", "RHO", "VIS", "STATE"):
                np.testing.assert_array_equal(lv[k.lower()],
                                              sol.read(f"{k}_{tag}", cnt)[ids],
                                              err_msg=f"{tag} {k}")
        assert fr.n_high == 4 * sc.n_low
