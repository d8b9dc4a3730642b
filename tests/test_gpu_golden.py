"""The HIP path against the committed golden step fixtures (tests/golden/).

Each fixture's stored state is written into a fresh solver, one step is
taken through the C ABI, and every phase is compared with the fixture:

* hashes, stable sort order, cell ranges, boundary grid: bit-exact
* density:  |rho - rho64| / rho64 <= 1e-5,  |p - p64| / max(|p64|, B) <= 1e-5
* accelerations: |a - a64| / max(|a64|, S) <= 1e-4, S the conditioning
  scale of the pair sum. The fixture's force used the FP64 density rounded
  to FP32, the HIP force uses its own FP32 density (within 1e-5); the Tait
  pressure multiplies a density difference by 7, so the looser bound
  covers that propagation (the same-input force check, with the oracle fed
  the HIP density, holds 1e-5 in test_gpu_parity.py).
* next lists, states, ages and dt: bit-exact (every threshold decision of
  a fixture has a density margin of at least 1e-4);
* positions and velocities: within dt (resp. dt^2) times the force bound.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import golden_io  # noqa: E402
from twoscalesplit_amd import Solver  # noqa: E402

pytestmark = pytest.mark.gpu

TOL_RHO = 1e-5
TOL_P = 1e-5
TOL_ACC = 1e-4

GRID = {"SORTED_L": "grid.sorted_l", "KEYS_L": "grid.keys_l",
        "SORTED_H": "grid.sorted_h", "KEYS_H": "grid.keys_h",
        "CELL_START_L": "grid.start_l", "CELL_END_L": "grid.end_l",
        "CELL_START_H": "grid.start_h", "CELL_END_H": "grid.end_h",
        "BND_SORTED": "grid.bsorted", "BND_START": "grid.bstart",
        "BND_END": "grid.bend"}


@pytest.mark.parametrize("name", golden_io.names())
def test_hip_step_matches_fixture(name):
    cfg, meta, a = golden_io.load(name)
    print("[case] golden", name, meta["n"], flush=True)
    snap = golden_io.snapshot_of(a)
    c = snap["counts"]
    for k, n in (("LIST_L", c["n_low"]), ("LIST_H", c["n_high"]),
                 ("LIST_G", c["n_groups"])):
        snap[k] = snap[k][:n]
    sol = Solver(cfg, a["init.low_xy"], meta["cap_low"], a["init.bnd_xy"])
    try:
        sol.restore(snap)
        sol.step(1)
        cnt = sol.counts()
        # ---- grid ------------------------------------------------------
        for f, k in GRID.items():
            np.testing.assert_array_equal(sol.read(f, cnt), a[k], err_msg=f)
        # ---- density ---------------------------------------------------
        ll = snap["LIST_L"]
        lh = snap["LIST_H"]
        B = float(cfg.tait_coefficient)
        rep = {}
        for ids, R, P, lvl in ((ll, "RHO_L", "P_L", "l"), (lh, "RHO_H", "P_H", "h")):
            if len(ids) == 0:
                continue
            rg = sol.read(R, cnt)[ids].astype(np.float64)
            pg = sol.read(P, cnt)[ids].astype(np.float64)
            r64, p64 = a["dens.rho_" + lvl][ids], a["dens.p_" + lvl][ids]
            er = np.abs(rg / r64 - 1)
            ep = np.abs(pg - p64) / np.maximum(np.abs(p64), B)
            rep["rho_" + lvl], rep["p_" + lvl] = er.max(), ep.max()
            assert er.max() <= TOL_RHO, (name, lvl, er.max())
            assert ep.max() <= TOL_P, (name, lvl, ep.max())
        # ---- force -----------------------------------------------------
        dt = float(snap["DT"][0])
        for ids, A, lvl in ((ll, "ACC_L", "l"), (lh, "ACC_H", "h")):
            if len(ids) == 0:
                continue
            ag = sol.read(A, cnt)[ids].astype(np.float64)
            a64 = a["force.acc_" + lvl][ids]
            S = a["force.scale_" + lvl][ids]
            den = np.maximum(np.linalg.norm(a64, axis=1), S)
            ea = np.linalg.norm(ag - a64, axis=1) / den
            rep["acc_" + lvl] = ea.max()
            assert ea.max() <= TOL_ACC, (name, lvl, ea.max())
            # integration of the moved particles
            dv = TOL_ACC * den * dt
            V = "VEL_" + lvl.upper()
            X = "POS_" + lvl.upper()
            vg = sol.read(V, cnt)[ids].astype(np.float64)
            xg = sol.read(X, cnt)[ids].astype(np.float64)
            vw = a["post." + V][ids].astype(np.float64)
            xw = a["post." + X][ids].astype(np.float64)
            ulp_v = 4 * np.spacing(np.abs(vw).astype(np.float32)).astype(np.float64)
            ulp_x = 4 * np.spacing(np.abs(xw).astype(np.float32)).astype(np.float64)
            assert np.all(np.abs(vg - vw) <= dv[:, None] + ulp_v), (name, V)
            assert np.all(np.abs(xg - xw) <= dv[:, None] * dt + ulp_x + 1e-6 * dt), (name, X)
        print("[golden]", name, {k: float(v) for k, v in rep.items()}, flush=True)
        # ---- adaptivity: next lists, states, ages, dt ------------------
        n_post = [int(x) for x in a["post.counts"]]
        assert [cnt["n_low"], cnt["n_high"], cnt["n_groups"]] == n_post
        for f, n in zip(("LIST_L", "LIST_H", "LIST_G"), n_post):
            np.testing.assert_array_equal(sol.read(f, cnt), a["post." + f][:n],
                                          err_msg=f)
        np.testing.assert_array_equal(sol.read("STATE_L", cnt), a["post.STATE_L"])
        np.testing.assert_array_equal(sol.read("AGE_L", cnt), a["post.AGE_L"])
        lh_post = a["post.LIST_H"][:n_post[1]]
        np.testing.assert_array_equal(sol.read("STATE_H", cnt)[lh_post],
                                      a["post.STATE_H"][lh_post])
        assert sol.read("DT", cnt)[0] == a["post.DT"][0]
        assert cnt["n_split"] == meta["n_split"]
        assert cnt["n_dissolve"] == meta["n_dissolve"]
        assert cnt["n_merge"] == meta["n_merge"]
        assert cnt["n_promote"] == meta["n_promote"]
        # new children: placed from the parent's density, within the force
        # bound of the fixture's placement
        pg = sol.read("POS_H", cnt)[lh_post].astype(np.float64)
        pw = a["post.POS_H"][lh_post].astype(np.float64)
        assert np.all(np.abs(pg - pw) <= 1e-6), (name, np.abs(pg - pw).max())
    finally:
        sol.close()
