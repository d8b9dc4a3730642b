"""Scene snapshots (tss_snapshot / tss_batch_restore) read and written
without the library: the header (``_native.TssSnapshotHeader``) padded to
16 bytes, then sections padded to 16 bytes -- boundary positions by id; low
arrays over cap_low ids (pos, vel, acc, rho, p, vis, state, age); high
arrays over 4 cap_low ids (pos, vel, acc, rho, p, vis, state); the low,
high and group lists (scene-local ids, in list order).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

MAGIC = 0x53535354
VERSION = 1

LOW = (("POS_L", "<f4", 2), ("VEL_L", "<f4", 2), ("ACC_L", "<f4", 2),
       ("RHO_L", "<f4", 1), ("P_L", "<f4", 1), ("VIS_L", "<f4", 1),
       ("STATE_L", "u1", 1), ("AGE_L", "<i4", 1))
HIGH = (("POS_H", "<f4", 2), ("VEL_H", "<f4", 2), ("ACC_H", "<f4", 2),
        ("RHO_H", "<f4", 1), ("P_H", "<f4", 1), ("VIS_H", "<f4", 1),
        ("STATE_H", "u1", 1))
LISTS = ("LIST_L", "LIST_H", "LIST_G")


def _pad(x: int) -> int:
    return (x + 15) & ~15


def layout(cap_low: int, n_bnd: int, n_low: int, n_high: int,
           n_groups: int) -> tuple[dict, int]:
    """Section name -> (offset, dtype, shape), and the total size."""
    out = {}
    off = _pad(C.sizeof(N.TssSnapshotHeader))

    def put(name, dt, shape):
        nonlocal off
        out[name] = (off, np.dtype(dt), shape)
        off += _pad(int(np.prod(shape)) * np.dtype(dt).itemsize)

    put("POS_B", "<f4", (n_bnd, 2))
    for name, dt, k in LOW:
        put(name, dt, (cap_low, k) if k > 1 else (cap_low,))
    for name, dt, k in HIGH:
        put(name, dt, (4 * cap_low, k) if k > 1 else (4 * cap_low,))
    for name, n in zip(LISTS, (n_low, n_high, n_groups)):
        put(name, "<i4", (n,))
    return out, off


def header(buf) -> N.TssSnapshotHeader:
    b = np.ascontiguousarray(buf, np.uint8)
    return N.TssSnapshotHeader.from_buffer_copy(
        b[:C.sizeof(N.TssSnapshotHeader)].tobytes())


def parse(buf) -> tuple[N.TssSnapshotHeader, dict]:
    """Header and arrays of a snapshot (views into ``buf``)."""
    b = np.ascontiguousarray(buf, np.uint8)
    hd = header(b)
    if hd.magic != MAGIC or hd.version != VERSION:
        raise ValueError("not a version-1 scene snapshot")
    secs, total = layout(hd.cap_low, hd.n_bnd, hd.n_low, hd.n_high,
                         hd.n_groups)
    if total != hd.total_bytes or total > b.nbytes:
        raise ValueError("snapshot size does not match its header")
    arrays = {}
    for name, (off, dt, shape) in secs.items():
        n = int(np.prod(shape)) * dt.itemsize
        arrays[name] = b[off:off + n].view(dt).reshape(shape)
    return hd, arrays


def build(hd: N.TssSnapshotHeader, arrays: dict) -> np.ndarray:
    """A snapshot from a header (sizes and counters) and its arrays; magic,
    version and byte counts are filled in."""
    secs, total = layout(hd.cap_low, hd.n_bnd, hd.n_low, hd.n_high,
                         hd.n_groups)
    hd.magic, hd.version = MAGIC, VERSION
    hd.header_bytes = C.sizeof(N.TssSnapshotHeader)
    hd.config_bytes = C.sizeof(N.TssConfig)
    hd.total_bytes = total
    out = np.zeros(total, np.uint8)
    raw = bytes(hd)
    out[:len(raw)] = np.frombuffer(raw, np.uint8)
    for name, (off, dt, shape) in secs.items():
        a = np.ascontiguousarray(arrays.get(name, np.zeros(shape, dt)), dt)
        if a.shape != tuple(shape):
            raise ValueError(f"{name}: shape {a.shape}, expected {shape}")
        out[off:off + a.nbytes] = a.reshape(-1).view(np.uint8)
    return out
