"""Seeded dam-break scenes of the benchmark configurations.

C1 plain      : 64x128 low particles, dx = 1/160, domain [0,1]^2, fixed dt 1e-4
C2 two-scale  : C1 + high-res band (4 dx from the rho < 0.9 rho0 surface),
                cap 32768
C3 large      : 1414x707 low (dx = 0.8/707) in [0,4]x[0,1], CFL 0.4;
                ``forced=True`` splits every low particle at the first step
                (999,698 low + 3,998,792 high)
C4 batched    : 256 scenes, low counts LogU[2k, 64k], aspect U(0.25, 2),
                alpha U(0.02, 0.1), cap 4x low; band rules, or the reference's
                split rule at mid-block (sustained two-scale load)
C5 served     : requests with low counts LogU[4k, 256k], steps U[500, 5000]

Kernel support radius h = 4 dx ("support 2h, h = 2dx" of the smoothing
length), H = h/2, rho0 = 1000, m = rho0 dx^2, m_H = m/4, gamma 7,
c0 = 10 sqrt(2 g 0.8), B = rho0 c0^2 / 7, three wall layers at spacing dx.
Positions are jittered by U(-0.01 dx, 0.01 dx) from SplitMix64 keyed by
(seed, scene, particle index) inside libtss (tss_scene_dam_break).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .solver import config_init

SEED = 20130213
G = 9.81
RHO0 = 1000.0
BLOCK_H = 0.8


@dataclass
class Scene:
    name: str
    cfg: N.TssConfig
    low_xy: np.ndarray
    bnd_xy: np.ndarray
    cap_low: int
    steps: int
    meta: dict = field(default_factory=dict)
    gen: N.TssDamBreak | None = None    # the same scene as a device generator

    @property
    def n_low(self) -> int:
        return len(self.low_xy)

    @property
    def cells(self) -> tuple[int, int]:
        c = self.cfg
        return (c.domain_dims[0] * c.domain_dims[1],
                c.domain_dims_high[0] * c.domain_dims_high[1])


def spec(seed: int, scene: int, nx: int, ny: int, dx: float, nxd: int,
         nyd: int, layers: int = 3, jitter: float = 0.01) -> N.TssDamBreak:
    """The arguments of ``generate`` as a device generator spec."""
    return N.TssDamBreak(seed, scene, nx, ny, float(dx), float(jitter), nxd,
                         nyd, layers)


def generate(seed: int, scene: int, nx: int, ny: int, dx: float, nxd: int,
             nyd: int, layers: int = 3, jitter: float = 0.01
             ) -> tuple[np.ndarray, np.ndarray]:
    lib = N.lib()
    nb = lib.tss_canvas_count(nxd, nyd, layers)
    low = np.zeros((nx * ny, 2), np.float32)
    bnd = np.zeros((max(nb, 0), 2), np.float32)
    N.check(lib.tss_scene_dam_break(seed, scene, nx, ny, float(dx),
                                    float(jitter), nxd, nyd, layers,
                                    N.ptr(low), N.ptr(bnd)),
            "tss_scene_dam_break")
    return low, bnd


def _config(width: float, height: float, dx: float, two_scale: bool,
            alpha: float = 0.05, dt: float = 1e-4, cfl: bool = False,
            block_h: float = BLOCK_H, h_ratio: float = 0.5) -> N.TssConfig:
    h = np.float32(4.0 * dx)
    H = np.float32(h * np.float32(h_ratio)) if two_scale else h
    c0 = 10.0 * math.sqrt(2.0 * G * block_h)
    B = RHO0 * c0 * c0 / 7.0
    dt0 = dt
    if cfl:
        dt0 = min(dt, 0.4 * 0.5 * float(h) / c0)
    cfg = config_init(0.0, 0.0, width, height, float(h), float(H), RHO0, B,
                      c0, alpha, 0.0, dt0)
    m = np.float32(RHO0 * dx * dx)
    cfg.mass_low = float(m)
    cfg.mass_high = float(m / np.float32(4)) if two_scale else float(m)
    cfg.mass_boundary = float(m)
    cfg.gravity = G
    if cfl:
        cfg.dt_mode = N.DT_CFL
        cfg.cfl = 0.4
        cfg.dt_max = dt
    return cfg


def c1_plain(seed: int = SEED) -> Scene:
    dx = 1.0 / 160.0
    low, bnd = generate(seed, 0, 64, 128, dx, 160, 160)
    cfg = _config(1.0, 1.0, dx, two_scale=False)
    cfg.mode = N.MODE_PLAIN
    return Scene("C1 plain 2D dam break (8192 low)", cfg, low, bnd, len(low),
                 1000, {"dx": dx}, spec(seed, 0, 64, 128, dx, 160, 160))


def c2_two_scale(seed: int = SEED, cap_high: int = 32768) -> Scene:
    dx = 1.0 / 160.0
    low, bnd = generate(seed, 0, 64, 128, dx, 160, 160)
    cfg = _config(1.0, 1.0, dx, two_scale=True)
    cfg.mode = N.MODE_BAND
    cfg.band_radius = cfg.effective_radius
    cfg.merge_radius = float(np.float32(1.5) * np.float32(cfg.effective_radius))
    cfg.cap_high = cap_high
    return Scene("C2 two-scale dam break (8192 low + <=32768 high)", cfg, low,
                 bnd, len(low), 2000, {"dx": dx},
                 spec(seed, 0, 64, 128, dx, 160, 160))


def c3_large(seed: int = SEED, forced: bool = True) -> Scene:
    nx, ny = 1414, 707
    dx = 0.8 / 707.0
    nxd = int(round(4.0 / dx))
    nyd = int(math.ceil(1.0 / dx))
    low, bnd = generate(seed, 0, nx, ny, dx, nxd, nyd)
    cfg = _config(4.0, 1.0, dx, two_scale=True, cfl=True)
    cfg.cap_high = 4 * nx * ny
    if forced:
        # every active low particle splits at its first step (the
        # reference's split rule with a threshold left of the domain)
        cfg.mode = N.MODE_REFERENCE
        cfg.split_x = -1.0
    else:
        cfg.mode = N.MODE_BAND
        cfg.merge_radius = float(np.float32(1.5) *
                                 np.float32(cfg.effective_radius))
    name = ("C3 large two-scale dam break, forced all-split "
            f"({nx * ny} low + {4 * nx * ny} high)" if forced else
            f"C3 large two-scale dam break ({nx * ny} low, band)")
    return Scene(name, cfg, low, bnd, nx * ny, 500,
                 {"dx": dx, "forced": forced},
                 spec(seed, 0, nx, ny, dx, nxd, nyd))


def _scene_from_count(seed: int, idx: int, n_low: int, aspect: float,
                      alpha: float, steps: int) -> Scene:
    ny = max(4, int(round(math.sqrt(n_low / aspect))))
    nx = max(4, int(round(n_low / ny)))
    dx = BLOCK_H / ny
    width = 2.5 * nx * dx
    height = 1.25 * BLOCK_H
    nxd = int(round(width / dx))
    nyd = int(round(height / dx))
    low, bnd = generate(seed, idx, nx, ny, dx, nxd, nyd)
    cfg = _config(width, height, dx, two_scale=True, alpha=alpha, cfl=True)
    cfg.mode = N.MODE_BAND
    cfg.merge_radius = float(np.float32(1.5) * np.float32(cfg.effective_radius))
    cfg.cap_high = 4 * nx * ny
    return Scene(f"scene {idx} ({nx}x{ny})", cfg, low, bnd, nx * ny, steps,
                 {"dx": dx, "aspect": aspect, "alpha": alpha},
                 spec(seed, idx, nx, ny, dx, nxd, nyd))


def c4_batched(seed: int = SEED, n_scenes: int = 256, lo: int = 2000,
               hi: int = 64000, steps: int = 300,
               rule: str = "band") -> list[Scene]:
    """rule "band": the surface-band injection/removal rules; "reference":
    the reference's split rule (split once x > split_x, never merged) with
    split_x at the middle of each fluid block, so half of every scene runs
    two-scale from the first step on."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n_scenes):
        n = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
        aspect = float(rng.uniform(0.25, 2.0))
        alpha = float(rng.uniform(0.02, 0.1))
        sc = _scene_from_count(seed, i, n, aspect, alpha, steps)
        if rule == "reference":
            sc.cfg.mode = N.MODE_REFERENCE
            nx = int(round(sc.cfg.domain_end[0] / 2.5 / sc.meta["dx"]))
            sc.cfg.split_x = float(np.float32(0.5 * nx * sc.meta["dx"]))
        elif rule != "band":
            raise ValueError(f"unknown rule {rule!r}")
        out.append(sc)
    return out


def c5_request(seed: int, idx: int, lo: int = 4000, hi: int = 256000,
               smin: int = 500, smax: int = 5000) -> Scene:
    rng = np.random.Generator(np.random.PCG64([seed, idx]))
    n = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
    aspect = float(rng.uniform(0.25, 2.0))
    alpha = float(rng.uniform(0.02, 0.1))
    steps = int(rng.integers(smin, smax + 1))
    return _scene_from_count(seed, idx, n, aspect, alpha, steps)


def small_two_scale(seed: int = SEED, nx: int = 16, ny: int = 24,
                    mode: int = N.MODE_BAND, nxd: int = 48, nyd: int = 40,
                    dx: float = 1.0 / 160.0, split_x: float = 0.5,
                    dt_cfl: bool = False, h_ratio: float = 0.5) -> Scene:
    """Small scene for parity tests (oracle runs in seconds); h_ratio = H/h."""
    low, bnd = generate(seed, 7, nx, ny, dx, nxd, nyd)
    cfg = _config(nxd * dx, nyd * dx, dx, two_scale=mode != N.MODE_PLAIN,
                  cfl=dt_cfl, h_ratio=h_ratio)
    cfg.mode = mode
    cfg.split_x = split_x
    cfg.cap_high = 4 * nx * ny if mode != N.MODE_PLAIN else 0
    cfg.merge_radius = float(np.float32(1.5) * np.float32(cfg.effective_radius))
    return Scene(f"small {nx}x{ny} mode {mode}", cfg, low, bnd, nx * ny, 0,
                 {"dx": dx}, spec(seed, 7, nx, ny, dx, nxd, nyd))
