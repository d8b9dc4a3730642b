"""Write the golden step fixtures of tests/golden/ from the oracle.

    python tests/golden/make_golden.py

Each fixture is a small seeded scene advanced a few FP64 oracle steps, then
one more step recorded phase by phase (grid, neighbours, density, force,
integration + adaptivity + next lists). The reference ships no tests or
golden data, so these vectors come from the oracle; the oracle itself is
pinned by closed forms and brute force in tests/test_oracle_pins.py.

A fixture is kept only when every density-threshold decision of its
recorded step has a relative margin above MIN_MARGIN, so a HIP density
within the parity tolerance makes the same decisions.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from oracle import binding as ob  # noqa: E402
from twoscalesplit_amd import MODE_BAND, MODE_PLAIN, MODE_REFERENCE, scenes  # noqa: E402

import golden_io  # noqa: E402

MIN_MARGIN = 1e-4

SNAP_FIELDS = {"POS_L": "pos_l", "VEL_L": "vel_l", "ACC_L": "acc_l",
               "RHO_L": "rho_l", "P_L": "p_l", "STATE_L": "state_l",
               "AGE_L": "age_l", "POS_H": "pos_h", "VEL_H": "vel_h",
               "ACC_H": "acc_h", "RHO_H": "rho_h", "P_H": "p_h",
               "STATE_H": "state_h"}


def snapshot(o: ob.Scene) -> dict:
    a = o.a
    snap = {k: a[v].copy() for k, v in SNAP_FIELDS.items()}
    snap["LIST_L"] = a["list_l"].copy()
    snap["LIST_H"] = a["list_h"].copy()
    snap["LIST_G"] = a["list_g"].copy()
    snap["DT"] = np.array([o.s.dt], np.float32)
    snap["counts"] = np.array(o.n, np.int32)
    return snap


def record_step(o: ob.Scene, cfg) -> dict:
    """One oracle step, phase by phase (the sequence of orc_step)."""
    out = {}
    g = o.grid()
    for k in ("sorted_l", "keys_l", "sorted_h", "keys_h", "start_l", "end_l",
              "start_h", "end_h", "bsorted", "bstart", "bend"):
        out["grid." + k] = g[k].copy()
    for lvl in (0, 1):
        c, s = o.neighbors(g, lvl)
        out[f"nbr.count{lvl}"] = c.copy()
        out[f"nbr.csum{lvl}"] = s.copy()
    nl, nh, _ = o.n
    ll, lh = o.a["list_l"][:nl].copy(), o.a["list_h"][:nh].copy()
    rl, pl, rh, ph = o.density64(g)
    out["dens.rho_l"], out["dens.p_l"] = rl, pl
    out["dens.rho_h"], out["dens.p_h"] = rh, ph
    o.a["rho_l"][ll] = rl[ll].astype(np.float32)
    o.a["p_l"][ll] = pl[ll].astype(np.float32)
    o.a["rho_h"][lh] = rh[lh].astype(np.float32)
    o.a["p_h"][lh] = ph[lh].astype(np.float32)
    al, ah, sl, sh = o.force64(g)
    out["force.acc_l"], out["force.acc_h"] = al, ah
    out["force.scale_l"], out["force.scale_h"] = sl, sh
    acc_l = np.zeros_like(o.a["acc_l"])
    acc_h = np.zeros_like(o.a["acc_h"])
    acc_l[ll] = al[ll].astype(np.float32)
    acc_h[lh] = ah[lh].astype(np.float32)
    ppl, pph = o.a["pos_l"].copy(), o.a["pos_h"].copy()
    o.integrate(acc_l, acc_h)
    v2 = o.v2max()
    ad = o.adapt(g, ppl, pph)
    o.s.dt = o.next_dt(v2)
    o.s.step += 1
    for k in ("split_ids", "promote_ids", "merge_ids", "dissolve_ids"):
        out["adapt." + k] = ad[k].copy()
    out["adapt.surf_l"] = ad["surf_l"].copy()
    out["adapt.surf_h"] = ad["surf_h"].copy()
    # margin of the density-threshold decisions (surface tests)
    rho0 = np.float32(cfg.rest_density)
    t_l = np.float32(np.float32(cfg.surface_ratio) * rho0)
    t_h = np.float32(np.float32(cfg.surface_ratio_high) * rho0)
    m = [np.abs(rl[ll] / float(t_l) - 1.0).min()] if nl else []
    if nh:
        m.append(np.abs(rh[lh] / float(t_h) - 1.0).min())
    out["margin"] = np.array([min(m) if m else 1.0])
    post = snapshot(o)
    for k, v in post.items():
        out["post." + k] = v
    return out


def make(name: str, sc, warm: int, want_split: bool) -> None:
    cfg = sc.cfg
    o = ob.Scene.from_initial(cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    for _ in range(warm):
        o.step(64, 1)
    pre = snapshot(o)
    chk = ob.Scene.from_initial(cfg, sc.low_xy, sc.cap_low, sc.bnd_xy)
    chk.load_snapshot({**pre, "counts": dict(zip(golden_io.COUNT_KEYS,
                                                  map(int, pre["counts"])))})
    chk.step(64, 1)
    rec = record_step(o, cfg)
    # the phase-by-phase record is the oracle's own step
    assert chk.n == o.n
    for k, n in zip(("list_l", "list_h", "list_g"), o.n):
        np.testing.assert_array_equal(chk.a[k][:n], o.a[k][:n], err_msg=k)
    for k in ("pos_l", "vel_l", "pos_h", "state_l", "state_h", "age_l"):
        np.testing.assert_array_equal(chk.a[k], o.a[k], err_msg=k)
    assert chk.s.dt == o.s.dt
    margin = float(rec["margin"][0])
    assert margin > MIN_MARGIN, (name, margin)
    if want_split:
        assert len(rec["adapt.split_ids"]) > 0, name
    arrays = {"pre." + k: v for k, v in pre.items()}
    arrays.update(rec)
    arrays["init.low_xy"] = sc.low_xy
    arrays["init.bnd_xy"] = sc.bnd_xy
    meta = {"name": name, "warm_steps": warm, "cap_low": sc.cap_low,
            "margin": margin, "n": [int(x) for x in pre["counts"]],
            "n_split": int(len(rec["adapt.split_ids"])),
            "n_dissolve": int(len(rec["adapt.dissolve_ids"])),
            "n_merge": int(len(rec["adapt.merge_ids"])),
            "n_promote": int(len(rec["adapt.promote_ids"]))}
    golden_io.save(name, cfg, meta, arrays)
    print(name, meta)


def main() -> None:
    make("plain_12x16", scenes.small_two_scale(nx=12, ny=16, nxd=24, nyd=24,
                                               mode=MODE_PLAIN), 40, False)
    make("reference_rule_10x12",
         scenes.small_two_scale(nx=10, ny=12, nxd=24, nyd=20,
                                mode=MODE_REFERENCE, split_x=0.03), 0, True)
    # band rules with promotion: step 4 promotes, merges and dissolves
    # groups; step 12 splits and promotes
    for warm, name in ((3, "band_merge_16x24"), (11, "band_split_16x24")):
        band = scenes.small_two_scale(mode=MODE_BAND)
        band.cfg.promote_min_age = 2
        band.cfg.promote_max_age = 4
        make(name, band, warm, warm == 11)


if __name__ == "__main__":
    main()
