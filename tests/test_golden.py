"""The committed golden step fixtures (tests/golden/) against the oracle, on
the CPU: the oracle re-run from each fixture's stored state reproduces
every recorded phase exactly, and the canonical-FP32 oracle stays inside
the density tolerance the HIP path is held to and makes the same
decisions (the fixtures keep a density margin of 1e-4 around every
threshold, ten times that tolerance).
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import golden_io  # noqa: E402
import make_golden  # noqa: E402
from oracle import binding as ob  # noqa: E402

NAMES = golden_io.names()
TOL_RHO = 1e-5


def _scene(name):
    cfg, meta, a = golden_io.load(name)
    o = ob.Scene.from_initial(cfg, a["init.low_xy"], meta["cap_low"],
                              a["init.bnd_xy"])
    o.load_snapshot(golden_io.snapshot_of(a))
    return cfg, meta, a, o


def test_fixtures_present():
    assert {"plain_12x16", "reference_rule_10x12", "band_merge_16x24",
            "band_split_16x24"} <= set(NAMES)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    cfg, meta, a, o = _scene(name)
    rec = make_golden.record_step(o, cfg)
    assert set(rec) <= set(a), set(rec) - set(a)
    n_post = a["post.counts"]
    for k, v in rec.items():
        want = a[k]
        if k in ("post.LIST_L", "post.LIST_H", "post.LIST_G"):
            n = int(n_post[("post.LIST_L", "post.LIST_H", "post.LIST_G").index(k)])
            v, want = v[:n], want[:n]
        np.testing.assert_array_equal(v, want, err_msg=f"{name}: {k}")
    ev = [len(a["adapt." + k]) for k in ("split_ids", "dissolve_ids",
                                          "merge_ids", "promote_ids")]
    assert ev == [meta["n_split"], meta["n_dissolve"], meta["n_merge"],
                  meta["n_promote"]]


@pytest.mark.parametrize("name", NAMES)
def test_fp32_oracle_within_tolerance_of_fixture(name):
    cfg, meta, a, o = _scene(name)
    nl, nh, _ = (int(x) for x in a["pre.counts"])
    ll, lh = a["pre.LIST_L"][:nl], a["pre.LIST_H"][:nh]
    o.step(32, 0)
    # densities of the listed particles (stored by the FP32 step)
    for ids, R, r64 in ((ll, "rho_l", a["dens.rho_l"]), (lh, "rho_h", a["dens.rho_h"])):
        if len(ids):
            err = np.abs(o.a[R][ids] / r64[ids] - 1)
            assert err.max() <= TOL_RHO, (name, R, err.max())
    # same decisions, hence the same next lists and states
    for k, lst in (("post.LIST_L", "list_l"), ("post.LIST_H", "list_h"),
                   ("post.LIST_G", "list_g")):
        n = int(a["post.counts"][("post.LIST_L", "post.LIST_H",
                                  "post.LIST_G").index(k)])
        np.testing.assert_array_equal(o.a[lst][:n], a[k][:n], err_msg=k)
    np.testing.assert_array_equal(o.a["state_l"], a["post.STATE_L"])
    np.testing.assert_array_equal(o.a["age_l"], a["post.AGE_L"])
