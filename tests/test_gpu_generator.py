"""Scenes generated on the device (tss_batch_add_dam_break, and serving
requests that carry a generator spec): the fluid block a kernel writes must
equal the host generator's positions bit for bit, and so must every field
after stepping; the served results with device generation must equal those
with uploaded arrays.
"""
from __future__ import annotations

import numpy as np
import pytest

from twoscalesplit_amd import MODE_BAND, MODE_REFERENCE, Batch, scenes
from twoscalesplit_amd import serve as S

pytestmark = pytest.mark.gpu

FIELDS = ("POS_L", "VEL_L", "RHO_L", "P_L", "STATE_L", "AGE_L", "POS_H",
          "VEL_H", "RHO_H", "STATE_H", "LIST_L", "LIST_H", "LIST_G", "DT")


def _pair(scs):
    cap = sum(s.cap_low for s in scs)
    nb = sum(len(s.bnd_xy) for s in scs)
    cl = sum(s.cells[0] for s in scs)
    ch = sum(s.cells[1] for s in scs)
    return Batch(len(scs), cap, nb, cl, ch), Batch(len(scs), cap, nb, cl, ch)


@pytest.mark.parametrize("which", ["small_band", "small_reference", "c2",
                                   "c4_scene"])
def test_device_generated_scene_equals_uploaded(which):
    if which == "small_band":
        sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_BAND)
    elif which == "small_reference":
        sc = scenes.small_two_scale(nx=24, ny=32, mode=MODE_REFERENCE,
                                    split_x=0.06)
    elif which == "c2":
        sc = scenes.c2_two_scale()
    else:
        sc = scenes.c4_batched(n_scenes=3)[2]
    print("[case] device generator", which, sc.name, flush=True)
    other = scenes.small_two_scale(nx=20, ny=20, mode=MODE_BAND)
    a, b = _pair([other, sc])
    try:
        a.add_scene(other.cfg, other.low_xy, other.cap_low, other.bnd_xy)
        b.add_dam_break(other.cfg, other.gen, other.cap_low)
        ia = a.add_scene(sc.cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
        ib = b.add_dam_break(sc.cfg, sc.gen, sc.cap_low)
        pa, pb = a.read(ia, "POS_L"), b.read(ib, "POS_L")
        np.testing.assert_array_equal(pb[:sc.n_low], sc.low_xy)
        np.testing.assert_array_equal(pb, pa)
        a.step(3)
        b.step(3)
        a.step_graph(5)
        b.step_graph(5)
        assert a.counts(ia) == b.counts(ib)
        for k in FIELDS:
            np.testing.assert_array_equal(b.read(ib, k), a.read(ia, k),
                                          err_msg=k)
    finally:
        a.close()
        b.close()


def test_generator_spec_must_fit_capacity():
    sc = scenes.small_two_scale(nx=16, ny=16)
    a, _b = _pair([sc])
    _b.close()
    try:
        with pytest.raises(RuntimeError, match="cap_low"):
            a.add_dam_break(sc.cfg, sc.gen, sc.n_low - 1)
        sid = a.add_dam_break(sc.cfg, sc.gen, sc.cap_low)
        assert a.counts(sid)["n_low"] == sc.n_low
    finally:
        a.close()


def test_served_with_device_generation_equals_uploaded():
    tr = S.Trace(5, 300.0, 5, lo=2000, hi=10000, smin=15, smax=40)
    out = ("POS_L", "VEL_L", "RHO_L", "STATE_L", "POS_H", "VEL_H")
    print("[case] serve with device generation",
          [s.n_low for s in tr.scenes], flush=True)
    desc = S.pool_desc(tr, max_resident_low=120_000, max_scenes=8)
    st_a, per_a, outs_a = S.serve(tr, desc, outputs=out)
    st_b, per_b, outs_b = S.serve(tr, desc, outputs=out, device_gen=True)
    print("[serve]", st_a["particle_steps_per_s"], st_b["particle_steps_per_s"],
          flush=True)
    assert st_a["particle_steps"] == st_b["particle_steps"]
    for i in range(len(tr.scenes)):
        assert per_a[i]["status"] == per_b[i]["status"] == 0
        for k in out:
            np.testing.assert_array_equal(outs_b[i][k], outs_a[i][k],
                                          err_msg=f"request {i}: {k}")
