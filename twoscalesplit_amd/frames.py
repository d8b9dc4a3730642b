"""Headless frame export (tss_frame / tss_batch_frame): the particles a
frame of the reference's renderer draws -- every entry of the current low
and high lists with its position and visual quantity, and optionally the
other per-particle fields -- as self-describing binary frames that can be
written to a file one after another and read back without the library.

Frame layout (little-endian): a 64-byte header (``HEADER``), then the low
section, the high section and, with ``BND``, the boundary positions by id.
A level's section holds one array per requested field in bit order
(ID, POS, VEL, RHO, P, VIS, STATE), each padded with zeros to a multiple of
16 bytes.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

ID, POS, VEL, RHO, P, VIS, STATE, BND = 1, 2, 4, 8, 16, 32, 64, 128
ALL = 255
# the fields the reference's renderer consumes: index list, positions and
# the visual quantity
RENDER = ID | POS | VIS
MAGIC = 0x46535354
VERSION = 1

HEADER = np.dtype([
    ("magic", "<u4"), ("version", "<u4"), ("fields", "<u4"), ("step", "<i4"),
    ("n_low", "<i4"), ("n_high", "<i4"), ("n_bnd", "<i4"), ("dt", "<f4"),
    ("domain_origin", "<f4", (2,)), ("domain_end", "<f4", (2,)),
    ("reserved", "<u4", (4,)),
])
assert HEADER.itemsize == 64

# (bit, name, dtype, components) in section order
_SECTIONS = ((ID, "id", np.dtype("<i4"), 1), (POS, "pos", np.dtype("<f4"), 2),
             (VEL, "vel", np.dtype("<f4"), 2), (RHO, "rho", np.dtype("<f4"), 1),
             (P, "p", np.dtype("<f4"), 1), (VIS, "vis", np.dtype("<f4"), 1),
             (STATE, "state", np.dtype("u1"), 1))


def _pad(x: int) -> int:
    return (x + 15) & ~15


def level_bytes(fields: int, n: int) -> int:
    return sum(_pad(n * dt.itemsize * k) for bit, _, dt, k in _SECTIONS
               if fields & bit)


def frame_size(fields: int, n_low: int, n_high: int, n_bnd: int = 0) -> int:
    return (HEADER.itemsize + level_bytes(fields, n_low) +
            level_bytes(fields, n_high) +
            (_pad(8 * n_bnd) if fields & BND else 0))


@dataclass
class Frame:
    step: int
    dt: float
    fields: int
    domain_origin: tuple
    domain_end: tuple
    low: dict = field(default_factory=dict)
    high: dict = field(default_factory=dict)
    bnd: np.ndarray | None = None

    @property
    def n_low(self) -> int:
        return _count(self.low)

    @property
    def n_high(self) -> int:
        return _count(self.high)


def _count(level: dict) -> int:
    for v in level.values():
        return len(v)
    return 0


def parse(buf, offset: int = 0) -> tuple[Frame, int]:
    """The frame at ``offset`` of ``buf`` and the offset after it."""
    mv = memoryview(buf)
    hd = np.frombuffer(mv[offset:offset + HEADER.itemsize], HEADER)[0]
    if int(hd["magic"]) != MAGIC:
        raise ValueError(f"not a frame at offset {offset}")
    if int(hd["version"]) != VERSION:
        raise ValueError(f"frame version {int(hd['version'])} not supported")
    f = int(hd["fields"])
    nl, nh, nb = int(hd["n_low"]), int(hd["n_high"]), int(hd["n_bnd"])
    pos = offset + HEADER.itemsize
    levels = []
    for n in (nl, nh):
        lv = {}
        for bit, name, dt, k in _SECTIONS:
            if f & bit:
                nbytes = n * dt.itemsize * k
                a = np.frombuffer(mv[pos:pos + nbytes], dt)
                lv[name] = a.reshape(n, k) if k > 1 else a
                pos += _pad(nbytes)
        levels.append(lv)
    bnd = None
    if f & BND:
        bnd = np.frombuffer(mv[pos:pos + 8 * nb], "<f4").reshape(nb, 2)
        pos += _pad(8 * nb)
    fr = Frame(int(hd["step"]), float(hd["dt"]), f,
               tuple(float(x) for x in hd["domain_origin"]),
               tuple(float(x) for x in hd["domain_end"]), levels[0], levels[1],
               bnd)
    return fr, pos


def pack(fr: Frame) -> bytes:
    """Serialise a Frame in the library's layout (for writers that build
    frames themselves, and to check the layout)."""
    hd = np.zeros(1, HEADER)
    hd["magic"], hd["version"], hd["fields"] = MAGIC, VERSION, fr.fields
    hd["step"], hd["dt"] = fr.step, fr.dt
    hd["n_low"], hd["n_high"] = fr.n_low, fr.n_high
    nb = len(fr.bnd) if (fr.fields & BND) and fr.bnd is not None else 0
    hd["n_bnd"] = nb
    hd["domain_origin"], hd["domain_end"] = fr.domain_origin, fr.domain_end
    out = [hd.tobytes()]
    for lv in (fr.low, fr.high):
        for bit, name, dt, k in _SECTIONS:
            if fr.fields & bit:
                b = np.ascontiguousarray(lv[name], dt).tobytes()
                out.append(b + bytes(_pad(len(b)) - len(b)))
    if fr.fields & BND:
        b = np.ascontiguousarray(fr.bnd if nb else np.zeros((0, 2)),
                                 "<f4").tobytes()
        out.append(b + bytes(_pad(len(b)) - len(b)))
    return b"".join(out)


def read_frames(path: str):
    """Iterate over the frames of a file written by ``FrameWriter``."""
    with open(path, "rb") as fh:
        data = fh.read()
    off = 0
    while off < len(data):
        fr, off = parse(data, off)
        yield fr


class FrameWriter:
    """Appends frames to a file (a headless replacement of the reference's
    video capture of every sixth displayed step, R/main.cpp:93-97)."""

    def __init__(self, path: str, fields: int = RENDER):
        self.path = path
        self.fields = fields
        self._fh = open(path, "wb")
        self._buf = bytearray(0)
        self.frames = 0

    def capture(self, handle, scene: int | None = None) -> int:
        """Writes the current frame of a Solver (scene None) or of one scene
        of a Batch; returns its size in bytes."""
        need = C.c_size_t()
        lib = N.lib()
        if scene is None:
            N.check(lib.tss_frame(handle._h, self.fields, None, 0,
                                  C.byref(need)), "tss_frame")
        else:
            N.check(lib.tss_batch_frame(handle._h, scene, self.fields, None, 0,
                                        C.byref(need)), "tss_batch_frame")
        if len(self._buf) < need.value:
            self._buf = bytearray(need.value)
        ptr = C.addressof((C.c_char * len(self._buf)).from_buffer(self._buf))
        got = C.c_size_t()
        if scene is None:
            N.check(lib.tss_frame(handle._h, self.fields, ptr, len(self._buf),
                                  C.byref(got)), "tss_frame")
        else:
            N.check(lib.tss_batch_frame(handle._h, scene, self.fields, ptr,
                                        len(self._buf), C.byref(got)),
                    "tss_batch_frame")
        self._fh.write(memoryview(self._buf)[:got.value])
        self.frames += 1
        return got.value

    def close(self) -> None:
        if self._fh:
            self._fh.close()
            self._fh = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def grab(handle, fields: int = ALL, scene: int | None = None) -> Frame:
    """The current frame of a Solver (scene None) or of one Batch scene."""
    lib = N.lib()
    need = C.c_size_t()
    if scene is None:
        call = lambda p, n, out: lib.tss_frame(handle._h, fields, p, n, out)
    else:
        call = lambda p, n, out: lib.tss_batch_frame(handle._h, scene, fields,
                                                     p, n, out)
    N.check(call(None, 0, C.byref(need)), "frame size")
    buf = np.zeros(max(need.value, 1), np.uint8)
    got = C.c_size_t()
    N.check(call(buf.ctypes.data, buf.nbytes, C.byref(got)), "frame")
    fr, end = parse(buf[:got.value].tobytes())
    assert end == got.value
    return fr


def record(solver, path: str, n_frames: int, every: int = 6,
           fields: int = RENDER, graph: bool = True) -> int:
    """Advances a Solver, writing a frame before the first step and after
    every ``every`` steps; returns the bytes written."""
    total = 0
    with FrameWriter(path, fields) as w:
        total += w.capture(solver)
        for _ in range(n_frames - 1):
            (solver.step_graph if graph else solver.step)(every)
            total += w.capture(solver)
    assert os.path.getsize(path) == total
    return total
