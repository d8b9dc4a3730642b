"""One-step parity of the HIP solver against the oracle (CPU restatement).

The oracle is re-seeded from the GPU state before the step; every phase is
then checked on the same inputs the GPU used:

* grid (hashes, stable sort order, cell ranges, boundary grid): bit-exact
* the pairs the step's own density and force kernels add (tss_pair_sets: the
  kernels' recording variant, same masks, scans and skips), per target and
  category, counts and id checksums: bit-exact against orc_pair_sets
* density (FP32, compensated) vs the FP64 oracle:
      |rho - rho64| / rho64                       <= TOL_RHO
      |p - p64| / max(|p64|, B)                   <= TOL_P
  (p/B + 1 = (rho/rho0)^7: the pressure is checked through the Tait ratio;
  dp/drho = 7 B (rho/rho0)^6 / rho0, so an FP32 density -- half an ulp,
  6e-8 -- moves p by ~7 B 6e-8 = 0.09 Pa however small |p| is. Reported
  beside it, not asserted: |p - p64| / max(|p64|, 7 B 2^-23), p50/p99/max)
* visual quantity (setQuantities, R/WCSPHSolver.cu:1030-1051):
  min(|rho - rho0|, 400) / 400 of the step's FP32 density, bit-exact
* force vs the FP64 oracle fed the GPU's rho and p:
      |a - a64| / max(|a64|, S)                   <= TOL_ACC
  S = sum of the magnitudes of the pair, wall and gravity terms (the
  conditioning scale of the cancelling sum)
* symplectic Euler, split / dissolve / promote / merge decisions, child
  placement and clamp, next lists, states, ages, dt: bit-exact given the
  GPU's accelerations and densities
"""
from __future__ import annotations

import numpy as np

from oracle import binding as ob

TOL_RHO = 1e-5
TOL_P = 1e-5
TOL_ACC = 1e-5


def _listed(snap, level):
    c = snap["counts"]
    if level == 0:
        return snap["LIST_L"][:c["n_low"]]
    return snap["LIST_H"][:c["n_high"]]


def check_pairs(sol, o, g) -> tuple[int, int]:
    """The hot-path pair sets of the current state against the oracle's
    (bit-exact); returns the number of listed targets of each level."""
    n = []
    for level in (0, 1):
        c_gpu, s_gpu = sol.pair_sets(level)
        c_orc, s_orc = o.pair_sets(g, level)
        np.testing.assert_array_equal(c_gpu, c_orc, err_msg=f"pair counts, level {level}")
        np.testing.assert_array_equal(s_gpu, s_orc, err_msg=f"pair checksums, level {level}")
        n.append(len(c_gpu))
    return n[0], n[1]


def check_step(sol, cfg, bnd_xy, report=None, pairs=True):
    """Advance `sol` one step and check it against the oracle (with pairs:
    first the kernels' pair sets on the pre-step state)."""
    pre = sol.snapshot()
    cap = pre["counts"]["cap_low"]
    if pairs and getattr(cfg, "relax_iters", 0) == 0:
        o0 = ob.Scene(ob.params_from_config(cfg), cap, bnd_xy)
        o0.load_snapshot(pre)
        check_pairs(sol, o0, o0.grid())
    sol.step(1)
    post = sol.snapshot()
    cnt = sol.counts()
    grid_fields = ["SORTED_L", "SORTED_H", "KEYS_L", "KEYS_H", "CELL_START_L",
                   "CELL_END_L", "CELL_START_H", "CELL_END_H", "BND_SORTED",
                   "BND_START", "BND_END"]
    gg = {k: sol.read(k, cnt) for k in grid_fields}

    o = ob.Scene(ob.params_from_config(cfg), cap, bnd_xy)
    o.load_snapshot(pre)
    g = o.grid()
    # ---- grid: bit-exact -------------------------------------------------
    np.testing.assert_array_equal(gg["SORTED_L"], g["sorted_l"])
    np.testing.assert_array_equal(gg["KEYS_L"], g["keys_l"])
    np.testing.assert_array_equal(gg["SORTED_H"], g["sorted_h"])
    np.testing.assert_array_equal(gg["KEYS_H"], g["keys_h"])
    np.testing.assert_array_equal(gg["CELL_START_L"], g["start_l"])
    np.testing.assert_array_equal(gg["CELL_END_L"], g["end_l"])
    np.testing.assert_array_equal(gg["CELL_START_H"], g["start_h"])
    np.testing.assert_array_equal(gg["CELL_END_H"], g["end_h"])
    np.testing.assert_array_equal(gg["BND_SORTED"], g["bsorted"])
    np.testing.assert_array_equal(gg["BND_START"], g["bstart"])
    np.testing.assert_array_equal(gg["BND_END"], g["bend"])

    # ---- density ---------------------------------------------------------
    rl, pl, rh, ph = o.density64(g)
    il, ih = _listed(pre, 0), _listed(pre, 1)
    B = float(cfg.tait_coefficient)
    err = {}
    for lvl, ids, r64, p64, R, P in ((0, il, rl, pl, "RHO_L", "P_L"),
                                     (1, ih, rh, ph, "RHO_H", "P_H")):
        if len(ids) == 0:
            continue
        rg = post[R][ids].astype(np.float64)
        pg = post[P][ids].astype(np.float64)
        er = np.abs(rg - r64[ids]) / np.abs(r64[ids])
        ep = np.abs(pg - p64[ids]) / np.maximum(np.abs(p64[ids]), B)
        err[f"rho{lvl}"] = float(er.max())
        err[f"p{lvl}"] = float(ep.max())
        epr = np.abs(pg - p64[ids]) / np.maximum(np.abs(p64[ids]), 7.0 * B * 2.0 ** -23)
        err[f"p_rel{lvl}"] = [float(np.percentile(epr, 50)),
                              float(np.percentile(epr, 99)), float(epr.max())]
        assert er.max() <= TOL_RHO, (lvl, er.max(), ids[er.argmax()])
        assert ep.max() <= TOL_P, (lvl, ep.max(), ids[ep.argmax()])

    # ---- visual quantity: bit-exact from the step's density -------------
    r0 = np.float32(cfg.rest_density)
    for ids, R, V in ((il, "RHO_L", "VIS_L"), (ih, "RHO_H", "VIS_H")):
        if len(ids) == 0:
            continue
        rho = post[R][ids].astype(np.float32)
        want = np.minimum(np.abs(rho - r0), np.float32(400.0)) / np.float32(400.0)
        np.testing.assert_array_equal(post[V][ids], want.astype(np.float32),
                                      err_msg=V)

    # ---- force (on the GPU's rho/p) --------------------------------------
    o.a["rho_l"][:] = post["RHO_L"]
    o.a["p_l"][:] = post["P_L"]
    o.a["rho_h"][:] = post["RHO_H"]
    o.a["p_h"][:] = post["P_H"]
    al, ah, sl, sh = o.force64(g)
    for lvl, ids, a64, sc, A in ((0, il, al, sl, "ACC_L"),
                                 (1, ih, ah, sh, "ACC_H")):
        if len(ids) == 0:
            continue
        ag = post[A][ids].astype(np.float64)
        d = np.linalg.norm(ag - a64[ids], axis=1)
        scale = np.maximum(np.linalg.norm(a64[ids], axis=1), sc[ids])
        ea = d / scale
        err[f"acc{lvl}"] = float(ea.max())
        assert ea.max() <= TOL_ACC, (lvl, ea.max(), ids[ea.argmax()])

    # ---- integrate + adapt + dt: bit-exact --------------------------------
    ppl = o.a["pos_l"].copy()
    pph = o.a["pos_h"].copy()
    o.integrate(post["ACC_L"], post["ACC_H"])
    v2 = o.v2max()
    out = o.adapt(g, ppl, pph)
    dt = o.next_dt(v2)
    ol, oh, og = o.lists()
    np.testing.assert_array_equal(post["LIST_L"], ol)
    np.testing.assert_array_equal(post["LIST_H"], oh)
    np.testing.assert_array_equal(post["LIST_G"], og)
    np.testing.assert_array_equal(post["STATE_L"], o.a["state_l"])
    np.testing.assert_array_equal(post["AGE_L"], o.a["age_l"])
    hl = np.concatenate([il, ol]).astype(np.int64)
    np.testing.assert_array_equal(post["POS_L"][hl], o.a["pos_l"][hl])
    np.testing.assert_array_equal(post["VEL_L"][hl], o.a["vel_l"][hl])
    hh = np.concatenate([ih, oh]).astype(np.int64)
    np.testing.assert_array_equal(post["STATE_H"][hh], o.a["state_h"][hh])
    np.testing.assert_array_equal(post["POS_H"][hh], o.a["pos_h"][hh])
    np.testing.assert_array_equal(post["VEL_H"][hh], o.a["vel_h"][hh])
    assert np.float32(post["DT"][0]) == np.float32(dt), (post["DT"][0], dt)
    assert cnt["n_split"] == len(out["split_ids"])
    assert cnt["n_merge"] == len(out["merge_ids"])
    assert cnt["n_dissolve"] == len(out["dissolve_ids"])
    assert cnt["n_promote"] == len(out["promote_ids"])
    if report is not None:
        err["events"] = (len(out["split_ids"]), len(out["promote_ids"]),
                         len(out["merge_ids"]), len(out["dissolve_ids"]))
        err["n"] = (len(il), len(ih))
        report.append(err)
    return err
