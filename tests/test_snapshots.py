"""Scene snapshots, host side (no GPU): the library's snapshot size equals
the layout the Python reader computes, and every malformed snapshot is
rejected with its reason before anything is allocated on a device (the
library validates sizes, ids, states, finiteness and the list invariants
first; none of these calls reaches the GPU).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from twoscalesplit_amd import _native as N
from twoscalesplit_amd import scenes
from twoscalesplit_amd import snapshots as S


@pytest.mark.parametrize("sizes", [(1, 0, 0, 0, 0), (7, 5, 3, 8, 2),
                                   (100, 33, 97, 400, 100), (16, 1, 16, 0, 0)])
def test_library_size_equals_layout(sizes):
    got = N.lib().tss_snapshot_size(*sizes)
    _, want = S.layout(*sizes)
    assert got == want and got % 16 == 0


def test_size_of_invalid_arguments_is_zero():
    lib = N.lib()
    assert lib.tss_snapshot_size(0, 0, 0, 0, 0) == 0
    assert lib.tss_snapshot_size(4, 0, 5, 0, 0) == 0        # n_low > cap
    assert lib.tss_snapshot_size(4, 0, 4, 17, 0) == 0       # n_high > 4 cap
    assert lib.tss_snapshot_size(4, -1, 4, 0, 0) == 0


def _valid():
    """A small consistent snapshot: 16 low particles, particle 3 split into
    four listed children (a transient group), particle 9 promoted (deleted
    parent, children listed, parent not in the low list)."""
    sc = scenes.small_two_scale(nx=4, ny=4)
    cap = 16
    hd = N.TssSnapshotHeader()
    hd.cfg = sc.cfg
    hd.cap_low, hd.n_bnd = cap, len(sc.bnd_xy)
    state_l = np.zeros(cap, np.uint8)
    state_l[3], state_l[9] = 1, 2
    ll = np.array([i for i in range(cap) if i != 9], np.int32)
    lg = np.array([3, 9], np.int32)
    lh = np.array([12, 13, 14, 15, 36, 37, 38, 39], np.int32)
    hd.n_low, hd.n_high, hd.n_groups = len(ll), len(lh), len(lg)
    hd.dt = 1e-4
    arrays = {"POS_B": sc.bnd_xy, "POS_L": sc.low_xy[:cap],
              "RHO_L": np.full(cap, 1000, np.float32), "STATE_L": state_l,
              "LIST_L": ll, "LIST_H": lh, "LIST_G": lg,
              "STATE_H": np.ones(4 * cap, np.uint8)}
    return hd, arrays


def _reject(buf, match):
    h = C.c_void_p()
    rc = N.lib().tss_create_from_snapshot(buf.ctypes.data, buf.nbytes,
                                          C.byref(h))
    msg = N.lib().tss_last_error().decode()
    assert rc == -1 and h.value is None, (rc, msg)
    assert match in msg, msg


def test_reader_round_trip():
    hd, arrays = _valid()
    buf = S.build(hd, arrays)
    assert buf.nbytes == N.lib().tss_snapshot_size(16, hd.n_bnd, 15, 8, 2)
    h2, got = S.parse(buf)
    assert (h2.n_low, h2.n_high, h2.n_groups) == (15, 8, 2)
    assert h2.cfg.effective_radius == hd.cfg.effective_radius
    for k, v in arrays.items():
        np.testing.assert_array_equal(got[k], v)
    assert bytes(buf[:4]) == b"TSSS"


@pytest.mark.parametrize("case,match", [
    ("short", "too short"),
    ("magic", "not a snapshot"),
    ("version", "version or layout"),
    ("counts", "counts out of range"),
    ("size", "size does not match"),
    ("nan_pos", "non-finite position"),
    ("nan_bnd", "non-finite boundary"),
    ("state", "unknown particle state"),
    ("dup", "invalid low list"),
    ("range", "invalid high list"),
    ("active_group", "group parent is active"),
    ("deleted_listed", "deleted particle listed"),
    ("orphan_child", "child of an unlisted group"),
    ("config", "Invalid configuration"),
])
def test_malformed_snapshots_are_rejected(case, match):
    hd, arrays = _valid()
    if case == "counts":
        hd.n_fresh = hd.n_high + 1
    elif case == "nan_pos":
        arrays["POS_L"] = arrays["POS_L"].copy()
        arrays["POS_L"][5, 1] = np.nan
    elif case == "nan_bnd":
        arrays["POS_B"] = arrays["POS_B"].copy()
        arrays["POS_B"][0, 0] = np.inf
    elif case == "state":
        arrays["STATE_H"] = arrays["STATE_H"].copy()
        arrays["STATE_H"][7] = 3
    elif case == "dup":
        arrays["LIST_L"] = arrays["LIST_L"].copy()
        arrays["LIST_L"][4] = arrays["LIST_L"][5]
    elif case == "range":
        arrays["LIST_H"] = arrays["LIST_H"].copy()
        arrays["LIST_H"][0] = 64
    elif case == "active_group":
        arrays["STATE_L"] = arrays["STATE_L"].copy()
        arrays["STATE_L"][3] = 0
    elif case == "deleted_listed":
        arrays["STATE_L"] = arrays["STATE_L"].copy()
        arrays["STATE_L"][5] = 2
    elif case == "orphan_child":
        arrays["LIST_H"] = arrays["LIST_H"].copy()
        arrays["LIST_H"][0] = 20
    elif case == "config":
        hd.cfg.effective_radius = -1.0
    buf = S.build(hd, arrays)
    if case == "short":
        buf = buf[:40].copy()
    elif case == "magic":
        buf[0] ^= 1
    elif case == "version":
        buf[4] = 9
    elif case == "size":
        buf = buf[:-16].copy()
    _reject(buf, match)
