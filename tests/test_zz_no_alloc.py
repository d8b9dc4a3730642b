"""tss_batch_frame_noalloc: the frame export staged in the library's fixed
device array, called through the C ABI, gives the same bytes as
tss_batch_frame (which stages in a buffer it allocates), on consecutive
calls -- the second one smaller than the first, so stale contents of the
shared array would show -- and refuses a frame beyond the array's size.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from twoscalesplit_amd import Batch, scenes
from twoscalesplit_amd import _native as N
from twoscalesplit_amd import frames as F

pytestmark = pytest.mark.gpu

ALL = F.ALL
TSS_E_INVALID = -1
NOALLOC_BYTES = 64 << 20


def _say(*a):
    print("[case]", *a, flush=True)


def _frame(fn, b, sid, fields):
    need = C.c_size_t()
    N.check(fn(b._h, sid, fields, None, 0, C.byref(need)), "frame size")
    buf = np.full(need.value, 0xAB, np.uint8)
    got = C.c_size_t()
    rc = fn(b._h, sid, fields, buf.ctypes.data, buf.nbytes, C.byref(got))
    return rc, buf, got.value


def _batch(scs):
    b = Batch(len(scs), sum(s.cap_low for s in scs), sum(len(s.bnd_xy) for s in scs),
              sum(s.cells[0] for s in scs), sum(s.cells[1] for s in scs))
    ids = [b.add_scene(s.cfg, s.low_xy, s.cap_low, s.bnd_xy) for s in scs]
    return b, ids


def test_noalloc_frame_equals_frame():
    lib = N.lib()
    scs = [scenes.c2_two_scale(), scenes.small_two_scale(nx=16, ny=24)]
    _say("no-alloc frames: C2 + small scene, 3 steps")
    b, ids = _batch(scs)
    try:
        b.step(3)
        b.synchronize()
        # first a large frame (every field, walls), then a smaller one
        calls = [(ids[0], ALL), (ids[1], F.POS | F.RHO),
                 (ids[0], F.ID | F.STATE)]
        for sid, fields in calls:
            rc0, want, n0 = _frame(lib.tss_batch_frame, b, sid, fields)
            rc1, got, n1 = _frame(lib.tss_batch_frame_noalloc, b, sid, fields)
            assert rc0 == 0 and rc1 == 0, (rc0, rc1, lib.tss_last_error())
            assert n0 == n1 == want.nbytes
            np.testing.assert_array_equal(got, want)
        # the state stepped on: frames still agree
        b.step_graph(2)
        rc0, want, _ = _frame(lib.tss_batch_frame, b, ids[1], ALL)
        rc1, got, _ = _frame(lib.tss_batch_frame_noalloc, b, ids[1], ALL)
        assert rc0 == 0 and rc1 == 0
        np.testing.assert_array_equal(got, want)
    finally:
        b.close()


def test_noalloc_frame_refuses_oversize():
    lib = N.lib()
    sc = scenes.c3_large(forced=True)
    _say("no-alloc frame beyond its array: C3 forced after the split")
    b, ids = _batch([sc])
    try:
        b.step(1)
        need = C.c_size_t()
        N.check(lib.tss_batch_frame_noalloc(b._h, ids[0], ALL, None, 0,
                                            C.byref(need)), "size query")
        assert need.value > NOALLOC_BYTES
        buf = np.zeros(need.value, np.uint8)
        got = C.c_size_t()
        rc = lib.tss_batch_frame_noalloc(b._h, ids[0], ALL, buf.ctypes.data,
                                         buf.nbytes, C.byref(got))
        assert rc == TSS_E_INVALID
        assert b"no-allocation export holds at most" in lib.tss_last_error()
        # a frame that fits still works after the refusal
        rc = lib.tss_batch_frame_noalloc(b._h, ids[0], F.STATE,
                                         buf.ctypes.data, buf.nbytes,
                                         C.byref(got))
        assert rc == 0
    finally:
        b.close()
