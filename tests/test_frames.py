"""Frame export layout, host side (no GPU): the library's frame size equals
the layout the Python reader parses, a frame serialised field by field reads
back unchanged (odd counts, so every padding rule is exercised), and the
header is the documented 64 bytes starting with "TSSF".
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from twoscalesplit_amd import _native as N
from twoscalesplit_amd import frames as F


@pytest.mark.parametrize("fields", [F.RENDER, F.ALL, F.POS, F.STATE,
                                    F.ID | F.STATE | F.BND, 0, F.BND])
@pytest.mark.parametrize("counts", [(0, 0, 0), (1, 0, 3), (7, 28, 5),
                                    (1000, 4001, 77)])
def test_library_size_equals_layout(fields, counts):
    nl, nh, nb = counts
    got = N.lib().tss_frame_size(fields, nl, nh, nb)
    assert got == F.frame_size(fields, nl, nh, nb)
    # sections are 16-byte aligned: sizes are multiples of 16
    assert got % 16 == 0


def test_size_of_invalid_arguments_is_zero():
    lib = N.lib()
    assert lib.tss_frame_size(256, 1, 1, 1) == 0
    assert lib.tss_frame_size(F.ALL, -1, 0, 0) == 0
    # explicit layout: RENDER with 3 low, 5 high
    want = 64 + (16 + 32 + 16) + (32 + 48 + 32)
    assert lib.tss_frame_size(F.RENDER, 3, 5, 0) == want


def _synthetic(fields, nl, nh, nb, seed=1):
    rng = np.random.default_rng(seed)

    def level(n):
        return {"id": rng.integers(0, 1 << 20, n).astype(np.int32),
                "pos": rng.random((n, 2), np.float32),
                "vel": rng.standard_normal((n, 2)).astype(np.float32),
                "rho": rng.random(n, np.float32) + 999,
                "p": rng.standard_normal(n).astype(np.float32),
                "vis": rng.random(n, np.float32),
                "state": rng.integers(0, 3, n).astype(np.uint8)}

    keep = [name for bit, name, _, _ in F._SECTIONS if fields & bit]
    lo = {k: v for k, v in level(nl).items() if k in keep}
    hi = {k: v for k, v in level(nh).items() if k in keep}
    bnd = rng.random((nb, 2), np.float32) if fields & F.BND else None
    return F.Frame(12, 2.5e-4, fields, (0.0, 0.0), (4.0, 1.0), lo, hi, bnd)


@pytest.mark.parametrize("fields", [F.ALL, F.RENDER, F.STATE | F.BND, F.VEL])
def test_pack_parse_round_trip(fields, tmp_path):
    frs = [_synthetic(fields, 5, 13, 3, seed=s) for s in range(3)]
    blobs = [F.pack(fr) for fr in frs]
    for fr, b in zip(frs, blobs):
        assert len(b) == F.frame_size(fields, 5, 13,
                                      3 if fields & F.BND else 0)
        assert b[:4] == b"TSSF"
    path = os.path.join(tmp_path, "f.tssf")
    with open(path, "wb") as fh:
        fh.write(b"".join(blobs))
    got = list(F.read_frames(path))
    assert len(got) == 3
    for a, b in zip(frs, got):
        assert (a.step, a.fields) == (b.step, b.fields)
        assert np.float32(a.dt) == np.float32(b.dt)
        assert b.domain_end == (4.0, 1.0)
        for la, lb in ((a.low, b.low), (a.high, b.high)):
            assert sorted(la) == sorted(lb)
            for k in la:
                np.testing.assert_array_equal(lb[k], la[k])
        if fields & F.BND:
            np.testing.assert_array_equal(b.bnd, a.bnd)


def test_padding_bytes_are_zero():
    fr = _synthetic(F.ALL, 3, 1, 1)
    b = F.pack(fr)
    # low id section: 12 bytes + 4 pad
    assert b[64 + 12:64 + 16] == bytes(4)
    # low state section: 3 bytes + 13 pad, the last low section
    low = F.level_bytes(F.ALL, 3)
    assert b[64 + low - 13:64 + low] == bytes(13)


def test_parse_rejects_foreign_bytes():
    with pytest.raises(ValueError, match="not a frame"):
        F.parse(bytes(64))
