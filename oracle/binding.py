"""ctypes binding of the CPU restatement (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

MODE_PLAIN, MODE_REFERENCE, MODE_BAND = 0, 1, 2

_P = C.c_void_p


class OrcParams(C.Structure):
    _fields_ = [
        ("origin", C.c_float * 2), ("end", C.c_float * 2),
        ("dims_l", C.c_int * 2), ("dims_h", C.c_int * 2),
        ("h", C.c_float), ("H", C.c_float),
        ("rho0", C.c_float), ("B", C.c_float), ("c", C.c_float),
        ("alpha", C.c_float), ("sigma", C.c_float), ("g", C.c_float),
        ("m_l", C.c_float), ("m_h", C.c_float), ("m_b", C.c_float),
        ("mode", C.c_int), ("split_x", C.c_float), ("surf_ratio", C.c_float),
        ("surf_ratio_h", C.c_float),
        ("r_band", C.c_float), ("r_merge", C.c_float),
        ("promote_min_age", C.c_int), ("promote_max_age", C.c_int),
        ("promote_fluc", C.c_float), ("cap_h", C.c_int),
        ("dt_mode", C.c_int), ("cfl", C.c_float), ("dt_max", C.c_float),
        ("relax_iters", C.c_int), ("relax_frac", C.c_float),
    ]


class OrcState(C.Structure):
    _fields_ = [
        ("cap_l", C.c_int), ("n_b", C.c_int),
        ("pos_l", _P), ("vel_l", _P), ("acc_l", _P), ("rho_l", _P),
        ("p_l", _P), ("state_l", _P), ("age_l", _P),
        ("pos_h", _P), ("vel_h", _P), ("acc_h", _P), ("rho_h", _P),
        ("p_h", _P), ("state_h", _P), ("pos_b", _P),
        ("n_l", C.c_int), ("n_h", C.c_int), ("n_g", C.c_int),
        ("list_l", _P), ("list_h", _P), ("list_g", _P),
        ("dt", C.c_float), ("step", C.c_int), ("n_fresh", C.c_int),
    ]


class OrcGrid(C.Structure):
    _fields_ = [(n, _P) for n in (
        "sorted_l", "sorted_h", "keys_l", "keys_h", "start_l", "end_l",
        "start_h", "end_h", "bsorted", "bstart", "bend")]


class OrcAdaptOut(C.Structure):
    _fields_ = [("n_split", C.c_int), ("n_promote", C.c_int),
                ("n_merge", C.c_int), ("n_dissolve", C.c_int),
                ("split_ids", _P), ("promote_ids", _P), ("merge_ids", _P),
                ("dissolve_ids", _P), ("surf_l", _P), ("surf_h", _P)]


_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        PS, SS, GS = C.POINTER(OrcParams), C.POINTER(OrcState), C.POINTER(OrcGrid)
        sig = {
            "orc_params_init": (C.c_int, [PS] + [C.c_float] * 11),
            "orc_axis_cell": (C.c_int, [C.c_float, C.c_float, C.c_float,
                                        C.c_float, C.c_int]),
            "orc_hash": (C.c_int, [PS, C.c_int, C.c_float, C.c_float]),
            "orc_hash_list": (None, [PS, C.c_int, _P, _P, C.c_int, _P]),
            "orc_stable_sort": (None, [_P, _P, C.c_int, C.c_int, _P, _P]),
            "orc_cell_ranges": (None, [_P, C.c_int, C.c_int, _P, _P]),
            "orc_build_grid": (None, [PS, SS, GS]),
            "orc_build_boundary_grid": (None, [PS, SS, GS]),
            "orc_neighbors": (None, [PS, SS, GS, C.c_int, _P, _P]),
            "orc_pair_sets": (None, [PS, SS, GS, C.c_int, _P, _P]),
            "orc_density_f64": (None, [PS, SS, GS, _P, _P, _P, _P]),
            "orc_density_f32": (None, [PS, SS, GS, _P, _P, _P, _P]),
            "orc_force_f64": (None, [PS, SS, GS, _P, _P, _P, _P]),
            "orc_force_f32": (None, [PS, SS, GS, _P, _P, _P, _P]),
            "orc_integrate": (None, [PS, SS, _P, _P]),
            "orc_adapt": (None, [PS, SS, GS, _P, _P, C.POINTER(OrcAdaptOut)]),
            "orc_next_dt": (C.c_float, [PS, SS, C.c_double]),
            "orc_v2max": (C.c_double, [SS]),
            "orc_conservation": (None, [PS, SS, _P, _P]),
            "orc_step": (C.c_int, [PS, SS, C.c_int, C.c_int]),
            "orc_relax_f64": (None, [PS, SS, GS]),
            "orc_relax_f32": (None, [PS, SS, GS]),
            "orc_kernel_w": (C.c_double, [C.c_double, C.c_double]),
            "orc_kernel_gradf": (C.c_double, [C.c_double, C.c_double]),
            "orc_wall_gamma": (C.c_double, [C.c_double, C.c_double,
                                            C.c_double]),
            "orc_tait": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return C.c_void_p(a.ctypes.data)


def params_from_config(cfg) -> OrcParams:
    """orc_params with the same values as a libtss tss_config."""
    p = OrcParams()
    p.origin[:] = list(cfg.domain_origin)
    p.end[:] = list(cfg.domain_end)
    p.dims_l[:] = list(cfg.domain_dims)
    p.dims_h[:] = list(cfg.domain_dims_high)
    p.h, p.H = cfg.effective_radius, cfg.effective_radius_high
    p.rho0, p.B, p.c = cfg.rest_density, cfg.tait_coefficient, cfg.speed_sound
    p.alpha, p.sigma, p.g = cfg.alpha, cfg.tension_coefficient, cfg.gravity
    p.m_l, p.m_h, p.m_b = cfg.mass_low, cfg.mass_high, cfg.mass_boundary
    p.mode, p.split_x = cfg.mode, cfg.split_x
    p.surf_ratio = cfg.surface_ratio
    p.surf_ratio_h = cfg.surface_ratio_high
    p.r_band, p.r_merge = cfg.band_radius, cfg.merge_radius
    p.promote_min_age, p.promote_max_age = (cfg.promote_min_age,
                                            cfg.promote_max_age)
    p.promote_fluc = cfg.promote_fluc
    p.cap_h = cfg.cap_high
    p.dt_mode, p.cfl, p.dt_max = cfg.dt_mode, cfg.cfl, cfg.dt_max
    p.relax_iters = getattr(cfg, "relax_iters", 0)
    p.relax_frac = getattr(cfg, "relax_dt_frac", 0.1)
    return p


class Scene:
    """An oracle scene: caller-owned numpy arrays behind an orc_state."""

    def __init__(self, params: OrcParams, cap_l: int, pos_b: np.ndarray):
        self.p = params
        self.cap_l = cap_l
        cap_h = 4 * cap_l
        f32 = np.float32
        self.a = {
            "pos_l": np.zeros((cap_l, 2), f32), "vel_l": np.zeros((cap_l, 2), f32),
            "acc_l": np.zeros((cap_l, 2), f32), "rho_l": np.zeros(cap_l, f32),
            "p_l": np.zeros(cap_l, f32), "state_l": np.zeros(cap_l, np.uint8),
            "age_l": np.zeros(cap_l, np.int32),
            "pos_h": np.zeros((cap_h, 2), f32), "vel_h": np.zeros((cap_h, 2), f32),
            "acc_h": np.zeros((cap_h, 2), f32), "rho_h": np.zeros(cap_h, f32),
            "p_h": np.zeros(cap_h, f32), "state_h": np.ones(cap_h, np.uint8),
            "pos_b": np.ascontiguousarray(pos_b, dtype=f32).reshape(-1, 2),
            "list_l": np.zeros(cap_l, np.int32), "list_h": np.zeros(cap_h, np.int32),
            "list_g": np.zeros(cap_l, np.int32),
        }
        self.s = OrcState()
        self.s.cap_l = cap_l
        self.s.n_b = len(self.a["pos_b"])
        for k, v in self.a.items():
            setattr(self.s, k, ptr(v))
        self.s.n_l = self.s.n_h = self.s.n_g = 0
        self.s.dt = 0.0
        self.s.step = 0

    @classmethod
    def from_initial(cls, cfg, low_xy, cap_l, bnd_xy) -> "Scene":
        sc = cls(params_from_config(cfg), cap_l, bnd_xy)
        n = len(low_xy)
        sc.a["pos_l"][:n] = low_xy
        sc.a["list_l"][:n] = np.arange(n, dtype=np.int32)
        sc.s.n_l = n
        sc.s.dt = cfg.time_step
        return sc

    def load_snapshot(self, snap: dict) -> None:
        """Re-seed from a libtss Solver.snapshot() of the same scene."""
        m = {"pos_l": "POS_L", "vel_l": "VEL_L", "acc_l": "ACC_L",
             "rho_l": "RHO_L", "p_l": "P_L", "state_l": "STATE_L",
             "age_l": "AGE_L", "pos_h": "POS_H", "vel_h": "VEL_H",
             "acc_h": "ACC_H", "rho_h": "RHO_H", "p_h": "P_H",
             "state_h": "STATE_H"}
        for k, v in m.items():
            self.a[k][...] = snap[v]
        c = snap["counts"]
        for lst, key, n in (("list_l", "LIST_L", c["n_low"]),
                            ("list_h", "LIST_H", c["n_high"]),
                            ("list_g", "LIST_G", c["n_groups"])):
            self.a[lst][:n] = snap[key][:n]
        self.s.n_l, self.s.n_h, self.s.n_g = (c["n_low"], c["n_high"],
                                              c["n_groups"])
        self.s.dt = float(snap["DT"][0])
        # children of the last step's splits close the high list
        self.s.n_fresh = 4 * int(c.get("n_split", 0))

    @property
    def n(self) -> tuple[int, int, int]:
        return self.s.n_l, self.s.n_h, self.s.n_g

    def grid(self) -> dict:
        p = self.p
        ncl = p.dims_l[0] * p.dims_l[1]
        nch = p.dims_h[0] * p.dims_h[1]
        g = {
            "sorted_l": np.zeros(max(self.s.n_l, 1), np.int32),
            "keys_l": np.zeros(max(self.s.n_l, 1), np.int32),
            "sorted_h": np.zeros(max(self.s.n_h, 1), np.int32),
            "keys_h": np.zeros(max(self.s.n_h, 1), np.int32),
            "start_l": np.zeros(ncl, np.int32), "end_l": np.zeros(ncl, np.int32),
            "start_h": np.zeros(nch, np.int32), "end_h": np.zeros(nch, np.int32),
            "bsorted": np.zeros(max(self.s.n_b, 1), np.int32),
            "bstart": np.zeros(ncl, np.int32), "bend": np.zeros(ncl, np.int32),
        }
        G = OrcGrid()
        for k, v in g.items():
            setattr(G, k, ptr(v))
        lib().orc_build_boundary_grid(C.byref(p), C.byref(self.s), C.byref(G))
        lib().orc_build_grid(C.byref(p), C.byref(self.s), C.byref(G))
        g["_struct"] = G
        g["sorted_l"] = g["sorted_l"][:self.s.n_l]
        g["keys_l"] = g["keys_l"][:self.s.n_l]
        g["sorted_h"] = g["sorted_h"][:self.s.n_h]
        g["keys_h"] = g["keys_h"][:self.s.n_h]
        g["bsorted"] = g["bsorted"][:self.s.n_b]
        return g

    def neighbors(self, g: dict, level: int) -> tuple[np.ndarray, np.ndarray]:
        n = self.s.n_h if level else self.s.n_l
        c = np.zeros((max(n, 1), 5), np.uint32)
        s = np.zeros((max(n, 1), 5), np.uint32)
        lib().orc_neighbors(C.byref(self.p), C.byref(self.s),
                            C.byref(g["_struct"]), level, ptr(c), ptr(s))
        return c[:n], s[:n]

    def pair_sets(self, g: dict, level: int) -> tuple[np.ndarray, np.ndarray]:
        """Per listed target of a level: the pairs with a nonzero state
        weight, as orc_neighbors' five categories (counts, id checksums)."""
        n = self.s.n_h if level else self.s.n_l
        c = np.zeros((max(n, 1), 5), np.uint32)
        s = np.zeros((max(n, 1), 5), np.uint32)
        lib().orc_pair_sets(C.byref(self.p), C.byref(self.s),
                            C.byref(g["_struct"]), level, ptr(c), ptr(s))
        return c[:n], s[:n]

    def density64(self, g: dict):
        cl, ch = self.cap_l, 4 * self.cap_l
        out = [np.zeros(cl), np.zeros(cl), np.zeros(ch), np.zeros(ch)]
        lib().orc_density_f64(C.byref(self.p), C.byref(self.s),
                              C.byref(g["_struct"]), *[ptr(o) for o in out])
        return out

    def force64(self, g: dict):
        cl, ch = self.cap_l, 4 * self.cap_l
        al, ah = np.zeros((cl, 2)), np.zeros((ch, 2))
        sl, sh = np.zeros(cl), np.zeros(ch)
        lib().orc_force_f64(C.byref(self.p), C.byref(self.s),
                            C.byref(g["_struct"]), ptr(al), ptr(ah), ptr(sl),
                            ptr(sh))
        return al, ah, sl, sh

    def integrate(self, acc_l: np.ndarray, acc_h: np.ndarray) -> None:
        al = np.ascontiguousarray(acc_l, dtype=np.float32)
        ah = np.ascontiguousarray(acc_h, dtype=np.float32)
        lib().orc_integrate(C.byref(self.p), C.byref(self.s), ptr(al), ptr(ah))

    def adapt(self, g: dict, pre_pos_l: np.ndarray, pre_pos_h: np.ndarray
              ) -> dict:
        cl = self.cap_l
        out = {"split_ids": np.zeros(cl, np.int32),
               "promote_ids": np.zeros(cl, np.int32),
               "merge_ids": np.zeros(cl, np.int32),
               "dissolve_ids": np.zeros(cl, np.int32),
               "surf_l": np.zeros(cl, np.uint8),
               "surf_h": np.zeros(4 * cl, np.uint8)}
        ao = OrcAdaptOut()
        for k, v in out.items():
            setattr(ao, k, ptr(v))
        ppl = np.ascontiguousarray(pre_pos_l, dtype=np.float32)
        pph = np.ascontiguousarray(pre_pos_h, dtype=np.float32)
        lib().orc_adapt(C.byref(self.p), C.byref(self.s),
                        C.byref(g["_struct"]), ptr(ppl), ptr(pph),
                        C.byref(ao))
        out["split_ids"] = out["split_ids"][:ao.n_split]
        out["promote_ids"] = out["promote_ids"][:ao.n_promote]
        out["merge_ids"] = out["merge_ids"][:ao.n_merge]
        out["dissolve_ids"] = out["dissolve_ids"][:ao.n_dissolve]
        return out

    def v2max(self) -> float:
        return lib().orc_v2max(C.byref(self.s))

    def next_dt(self, v2max: float) -> float:
        return lib().orc_next_dt(C.byref(self.p), C.byref(self.s), v2max)

    def conservation(self) -> tuple[np.ndarray, np.ndarray]:
        out, ab = np.zeros(5), np.zeros(5)
        lib().orc_conservation(C.byref(self.p), C.byref(self.s), ptr(out),
                               ptr(ab))
        return out, ab

    def relax(self, g: dict, precision: int = 64) -> None:
        """Relaxation of the last n_fresh children of list_h on grid g."""
        fn = lib().orc_relax_f64 if precision == 64 else lib().orc_relax_f32
        fn(C.byref(self.p), C.byref(self.s), C.byref(g["_struct"]))

    def step(self, precision: int = 64, nthreads: int = 0) -> None:
        rc = lib().orc_step(C.byref(self.p), C.byref(self.s), precision,
                            nthreads)
        assert rc == 0

    def lists(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        return (self.a["list_l"][:self.s.n_l].copy(),
                self.a["list_h"][:self.s.n_h].copy(),
                self.a["list_g"][:self.s.n_g].copy())
