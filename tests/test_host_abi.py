"""Host-side checks of libtss.so that need no GPU: every symbol declared in
include/tss.h is exported, configuration defaults and errors follow
WCSPHConfig (R/WCSPHSolver.cu:2287-2332), and the seeded scene generator
matches an independent restatement plus published SplitMix64 outputs.
No call here launches a kernel.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import re

import numpy as np
import pytest
from scipy import integrate

from twoscalesplit_amd import _native as N
from twoscalesplit_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tss.h")


def _declared_functions() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(tss_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_symbols_exported():
    names = _declared_functions()
    assert len(names) >= 30, names
    lib = C.CDLL(N.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a prototype for each of them
    bound = N.lib()
    for n in names:
        assert getattr(bound, n).argtypes is not None or n in ("tss_version",), n


def test_version_string():
    v = N.lib().tss_version()
    assert isinstance(v, bytes) and len(v) > 0


# ----------------------------------------------------------- configuration --
def test_config_init_dims_and_defaults():
    cfg = N.TssConfig()
    rc = N.lib().tss_config_init(C.byref(cfg), 0.0, 0.0, 1.0, 1.0, 0.025,
                                 0.0125, 1000.0, 224228.6, 39.6, 0.05, 0.0, 1e-4)
    assert rc == N.TSS_OK
    # dims = (int)ceil(extent / radius) per level (R/WCSPHSolver.cu:2323-2331)
    assert list(cfg.domain_dims) == [40, 40]
    assert list(cfg.domain_dims_high) == [80, 80]
    assert cfg.mode == N.MODE_PLAIN and cfg.dt_mode == N.DT_FIXED
    assert cfg.gravity == pytest.approx(9.81)
    assert cfg.promote_max_age == 0          # promotion is opt-in
    assert cfg.dt_max == np.float32(1e-4)


def test_config_dims_round_up():
    cfg = N.TssConfig()
    assert N.lib().tss_config_init(C.byref(cfg), 0.0, 0.0, 4.0, 1.0, 0.0045,
                                   0.00225, 1000, 1, 1, 0, 0, 1e-4) == 0
    e = np.float32(4.0) / np.float32(0.0045)
    assert cfg.domain_dims[0] == math.ceil(e)
    assert cfg.domain_dims_high[0] == math.ceil(np.float32(4.0) / np.float32(0.00225))


@pytest.mark.parametrize("args", [
    (1.0, 0.0, 1.0, 1.0),   # xs >= xe
    (0.0, 2.0, 1.0, 1.0),   # ys >= ye
])
def test_config_init_rejects_empty_domain(args):
    cfg = N.TssConfig()
    rc = N.lib().tss_config_init(C.byref(cfg), *args, 0.025, 0.0125, 1000.0,
                                 1.0, 1.0, 0.0, 0.0, 1e-4)
    assert rc != N.TSS_OK
    assert b"Invalid configuration parameters" in N.lib().tss_last_error()


def test_surface_ratio_high_is_bulk_reading_of_high_estimator():
    """The high level sums the blended kernel (W_h + W_H)/2 only over r <= H,
    so its bulk reading is rho0 (1 + F)/2 with F = int_{r<=H} W_h dA; the
    default surface threshold scales 0.9 by it. F by quadrature of the
    oracle's kernel (an independent evaluation), and by a lattice sum."""
    from oracle import binding as ob
    L = ob.lib()
    h, H = 0.025, 0.0125
    F, _ = integrate.quad(lambda r: L.orc_kernel_w(r, h) * 2 * math.pi * r, 0, H,
                          epsabs=1e-14, epsrel=1e-13)
    cfg = N.TssConfig()
    assert N.lib().tss_config_init(C.byref(cfg), 0, 0, 1, 1, h, H, 1000, 1, 1,
                                   0, 0, 1e-4) == 0
    assert cfg.surface_ratio_high == pytest.approx(0.9 * (1 + F) / 2, rel=2e-7)
    # rest lattice of high particles (spacing H/2, mass rho0 (H/2)^2)
    dxh = H / 2
    k = np.arange(-8, 9)
    X, Y = np.meshgrid(k * dxh, k * dxh)
    r = np.hypot(X, Y).ravel()
    r = r[r <= H]
    est = sum((L.orc_kernel_w(x, h) + L.orc_kernel_w(x, H)) / 2 for x in r) * dxh * dxh
    assert est == pytest.approx((1 + F) / 2, rel=0.01)
    # equal radii: the estimator is the plain kernel sum, ratio 0.9
    assert N.lib().tss_config_init(C.byref(cfg), 0, 0, 1, 1, h, h, 1000, 1, 1,
                                   0, 0, 1e-4) == 0
    assert cfg.surface_ratio_high == pytest.approx(0.9, rel=1e-6)


# --------------------------------------------------------------- generator --
GOLD = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


def _mix(z: int) -> int:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014; Vigna's splitmix64.c)."""
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_splitmix64_published_outputs():
    # first outputs of splitmix64 seeded with 0
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F,
            0xF88BB8A8724C81EC]
    st = C.c_uint64(0)
    got = [N.lib().tss_splitmix64_next(C.byref(st)) for _ in want]
    assert got == want
    assert st.value == (4 * GOLD) & M64


def _dam_break_py(seed, scene, nx, ny, dx, jitter, nxd, nyd, layers):
    f32 = np.float32
    dx32 = f32(dx)
    amp = f32(f32(jitter) * dx32)
    low = np.zeros((nx * ny, 2), np.float32)
    for j in range(ny):
        for i in range(nx):
            idx = j * nx + i
            st = (_mix((seed + GOLD * (scene + 1)) & M64) + GOLD * (idx + 1)) & M64
            st = _mix(st)
            st = (st + GOLD) & M64
            ux = f32(_mix(st) >> 40) * f32(1.0 / 16777216.0)
            st = (st + GOLD) & M64
            uy = f32(_mix(st) >> 40) * f32(1.0 / 16777216.0)
            jx = f32(f32(f32(2.0) * ux) - f32(1.0)) * amp
            jy = f32(f32(f32(2.0) * uy) - f32(1.0)) * amp
            low[idx, 0] = f32(f32(f32(0.5) + f32(i)) * dx32) + jx
            low[idx, 1] = f32(f32(f32(0.5) + f32(j)) * dx32) + jy
    bnd = []
    for j in range(-layers, nyd + layers):
        for i in range(-layers, nxd + layers):
            if i < 0 or i >= nxd or j < 0 or j >= nyd:
                bnd.append((f32(f32(0.5) + f32(i)) * dx32, f32(f32(0.5) + f32(j)) * dx32))
    return low, np.array(bnd, np.float32).reshape(-1, 2)


@pytest.mark.parametrize("seed,scene", [(0, 0), (20130213, 7), (2**63 + 5, 255)])
def test_dam_break_generator_matches_restatement(seed, scene):
    nx, ny, dx, nxd, nyd, layers = 9, 6, 1.0 / 160.0, 14, 11, 3
    low, bnd = scenes.generate(seed, scene, nx, ny, dx, nxd, nyd, layers)
    wl, wb = _dam_break_py(seed, scene, nx, ny, dx, 0.01, nxd, nyd, layers)
    np.testing.assert_array_equal(low, wl)
    np.testing.assert_array_equal(bnd, wb)


def test_dam_break_scene_geometry():
    sc = scenes.c1_plain()
    dx = sc.meta["dx"]
    assert sc.n_low == 64 * 128
    assert len(sc.bnd_xy) == N.lib().tss_canvas_count(160, 160, 3) == 166 * 166 - 160 * 160
    i = np.tile(np.arange(64), 128)
    j = np.repeat(np.arange(128), 64)
    off = sc.low_xy - np.stack([(i + 0.5) * dx, (j + 0.5) * dx], axis=1)
    assert np.all(np.abs(off) <= 0.01 * dx * (1 + 1e-4))
    assert np.std(off) > 0.004 * dx          # jitter is present
    # walls: three layers outside [0,1]^2, on the lattice
    b = sc.bnd_xy
    outside = (b[:, 0] < 0) | (b[:, 0] > 1) | (b[:, 1] < 0) | (b[:, 1] > 1)
    assert np.all(outside)
    assert np.all(np.abs(b).max(axis=1) < 1 + 3 * dx + 1e-6)
    # a different scene index draws different jitter
    o2, _ = scenes.generate(scenes.SEED, 1, 64, 128, dx, 160, 160)
    assert not np.array_equal(o2, sc.low_xy)


def test_generator_rejects_bad_arguments():
    low = np.zeros((4, 2), np.float32)
    rc = N.lib().tss_scene_dam_break(0, 0, 2, 2, -1.0, 0.01, 4, 4, 1,
                                     N.ptr(low), None)
    assert rc != N.TSS_OK
    assert N.lib().tss_canvas_count(-1, 4, 1) == -1


def test_header_is_plain_c99():
    import shutil
    import subprocess
    import tempfile
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "h.c")
        with open(src, "w") as f:
            f.write('#include "tss.h"\nint main(void) { return 0; }\n')
        r = subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", "-pedantic",
                            "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                            src], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_tuning_struct_layout():
    t = N.TssTuning.uniform(128, 8, 1, 1024)
    assert C.sizeof(t) == 64
    d = t.as_dict()
    assert d["force_high"] == {"block": 128, "depth": 8, "xcd_order": 1, "grid": 1024}
