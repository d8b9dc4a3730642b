"""ctypes binding of libtss.so (include/tss.h).

The library is loaded from this package's ``lib/`` directory; a missing or
unloadable library raises at once -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib",
                        "libtss.so")

TSS_OK = 0
MODE_PLAIN, MODE_REFERENCE, MODE_BAND = 0, 1, 2
# grid-build sort paths (tss_set_grid_sort)
SORT_AUTO, SORT_RADIX, SORT_COUNT, SORT_SMALL = 0, 1, 2, 3
DT_FIXED, DT_CFL = 0, 1

FIELDS = [
    "POS_L", "VEL_L", "ACC_L", "RHO_L", "P_L", "STATE_L", "AGE_L", "VIS_L",
    "POS_H", "VEL_H", "ACC_H", "RHO_H", "P_H", "STATE_H", "VIS_H",
    "LIST_L", "LIST_H", "LIST_G", "SORTED_L", "SORTED_H", "KEYS_L", "KEYS_H",
    "CELL_START_L", "CELL_END_L", "CELL_START_H", "CELL_END_H",
    "BND_SORTED", "BND_START", "BND_END", "POS_B", "DT",
]
FIELD = {name: i for i, name in enumerate(FIELDS)}


class TssConfig(C.Structure):
    _fields_ = [
        ("domain_origin", C.c_float * 2),
        ("domain_end", C.c_float * 2),
        ("domain_dims", C.c_int * 2),
        ("domain_dims_high", C.c_int * 2),
        ("effective_radius", C.c_float),
        ("effective_radius_high", C.c_float),
        ("rest_density", C.c_float),
        ("tait_coefficient", C.c_float),
        ("speed_sound", C.c_float),
        ("alpha", C.c_float),
        ("tension_coefficient", C.c_float),
        ("time_step", C.c_float),
        ("gravity", C.c_float),
        ("mass_low", C.c_float),
        ("mass_high", C.c_float),
        ("mass_boundary", C.c_float),
        ("mode", C.c_int),
        ("split_x", C.c_float),
        ("surface_ratio", C.c_float),
        ("surface_ratio_high", C.c_float),
        ("band_radius", C.c_float),
        ("merge_radius", C.c_float),
        ("promote_min_age", C.c_int),
        ("promote_max_age", C.c_int),
        ("promote_fluc", C.c_float),
        ("cap_high", C.c_int),
        ("dt_mode", C.c_int),
        ("cfl", C.c_float),
        ("dt_max", C.c_float),
        ("relax_iters", C.c_int),
        ("relax_dt_frac", C.c_float),
    ]


class TssCounts(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "n_low", "n_high", "n_groups", "n_low_last", "n_high_last", "n_split",
        "n_promote", "n_merge", "n_dissolve", "step", "cap_low", "n_bnd",
        "cells_low",
        "cells_high")] + [("particle_steps", C.c_longlong)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class TssTuning(C.Structure):
    """Launch shapes of the neighbour kernels [density low, density high,
    force low, force high]."""
    _fields_ = [("block", C.c_int * 4), ("depth", C.c_int * 4),
                ("xcd_order", C.c_int * 4), ("grid", C.c_int * 4)]

    KERNELS = ("density_low", "density_high", "force_low", "force_high")

    def as_dict(self) -> dict:
        return {k: {"block": self.block[i], "depth": self.depth[i],
                    "xcd_order": self.xcd_order[i], "grid": self.grid[i]}
                for i, k in enumerate(self.KERNELS)}

    @classmethod
    def uniform(cls, block=256, depth=16, xcd_order=0, grid=0) -> "TssTuning":
        t = cls()
        for i in range(4):
            t.block[i], t.depth[i] = block, depth
            t.xcd_order[i], t.grid[i] = xcd_order, grid
        return t


class TssDamBreak(C.Structure):
    """Seeded dam-break scene that the device generates (tss_dam_break)."""
    _fields_ = [
        ("seed", C.c_uint64), ("scene", C.c_uint64),
        ("nx", C.c_int), ("ny", C.c_int),
        ("dx", C.c_float), ("jitter", C.c_float),
        ("nxd", C.c_int), ("nyd", C.c_int), ("layers", C.c_int),
    ]


class TssSnapshotHeader(C.Structure):
    """tss_snapshot_header (natural C alignment, like the library's)."""
    _fields_ = [
        ("magic", C.c_uint32), ("version", C.c_uint32),
        ("header_bytes", C.c_uint32), ("config_bytes", C.c_uint32),
        ("total_bytes", C.c_uint64),
        ("cfg", TssConfig),
        ("cap_low", C.c_int32), ("n_bnd", C.c_int32),
        ("n_low", C.c_int32), ("n_high", C.c_int32), ("n_groups", C.c_int32),
        ("n_fresh", C.c_int32),
        ("step", C.c_int32), ("n_low_last", C.c_int32),
        ("n_high_last", C.c_int32),
        ("n_split", C.c_int32), ("n_promote", C.c_int32),
        ("n_merge", C.c_int32), ("n_dissolve", C.c_int32),
        ("n_cand", C.c_int32),
        ("particle_steps", C.c_int64),
        ("dt", C.c_float), ("reserved0", C.c_int32),
        ("conservation", C.c_double * 10),
    ]


class TssRequest(C.Structure):
    """One scene request of tss_serve."""
    _fields_ = [
        ("cfg", C.POINTER(TssConfig)),
        ("low_xy", C.c_void_p), ("n_low", C.c_int), ("cap_low", C.c_int),
        ("bnd_xy", C.c_void_p), ("n_bnd", C.c_int),
        ("gen", C.POINTER(TssDamBreak)),
        ("steps", C.c_int),
        ("arrival_s", C.c_double),
        ("out", C.c_void_p * len(FIELDS)),
        ("out_bytes", C.c_size_t * len(FIELDS)),
        ("status", C.c_int),
        ("admit_s", C.c_double), ("finish_s", C.c_double),
        ("particle_steps", C.c_longlong),
    ]


class TssServeStats(C.Structure):
    _fields_ = [
        ("wall_s", C.c_double), ("busy_s", C.c_double),
        ("particle_steps", C.c_longlong),
        ("particle_steps_per_s", C.c_double),
        ("ticks", C.c_int), ("graph_ticks", C.c_int),
        ("max_resident", C.c_int),
        ("step_ms_p50", C.c_double), ("step_ms_p99", C.c_double),
        ("request_ms_p50", C.c_double), ("request_ms_p99", C.c_double),
        ("queue_ms_p50", C.c_double), ("queue_ms_p99", C.c_double),
        ("batch_bytes", C.c_longlong), ("budget_bytes", C.c_longlong),
        ("peak_admitted_bytes", C.c_longlong),
    ]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class TssBatchDesc(C.Structure):
    _fields_ = [("max_scenes", C.c_int), ("cap_low", C.c_longlong),
                ("cap_bnd", C.c_longlong), ("cells_low", C.c_longlong),
                ("cells_high", C.c_longlong)]


_P = C.c_void_p
_F = C.c_float
_I = C.c_int
_SZ = C.c_size_t

# name -> (restype, argtypes)
PROTOTYPES = {
    "tss_config_init": (_I, [C.POINTER(TssConfig), _F, _F, _F, _F, _F, _F, _F,
                             _F, _F, _F, _F, _F]),
    "tss_create": (_I, [C.POINTER(TssConfig), _P, _I, _I, _P, _I,
                        C.POINTER(_P)]),
    "tss_step": (_I, [_P, _I]),
    "tss_step_graph": (_I, [_P, _I]),
    "tss_counts": (_I, [_P, C.POINTER(TssCounts)]),
    "tss_read": (_I, [_P, _I, _P, _SZ]),
    "tss_write": (_I, [_P, _I, _P, _SZ]),
    "tss_set_counts": (_I, [_P, _I, _I, _I]),
    "tss_device_ptr": (_I, [_P, _I, C.POINTER(_P)]),
    "tss_synchronize": (_I, [_P]),
    "tss_conservation": (_I, [_P, _P, _P]),
    "tss_neighbors": (_I, [_P, _I, _P, _P, _SZ]),
    "tss_pair_sets": (_I, [_P, _I, _P, _P, _SZ]),
    "tss_enable_timing": (_I, [_P, _I]),
    "tss_phase_times": (_I, [_P, _P]),
    "tss_work_counts": (_I, [_P, _P]),
    "tss_set_grid_sort": (_I, [_P, _I]),
    "tss_batch_set_grid_sort": (_I, [_P, _I]),
    "tss_get_tuning": (_I, [_P, C.POINTER(TssTuning)]),
    "tss_set_tuning": (_I, [_P, C.POINTER(TssTuning)]),
    "tss_autotune": (_I, [_P, _I, _I, C.POINTER(TssTuning), C.POINTER(_I), _P,
                          _I, C.POINTER(_I)]),
    "tss_destroy": (None, [_P]),
    "tss_batch_memory_bytes": (_I, [C.POINTER(TssBatchDesc),
                                    C.POINTER(C.c_longlong)]),
    "tss_batch_create": (_I, [C.POINTER(TssBatchDesc), C.POINTER(_P)]),
    "tss_batch_add_scene": (_I, [_P, C.POINTER(TssConfig), _P, _I, _I, _P, _I,
                                 C.POINTER(_I)]),
    "tss_batch_add_dam_break": (_I, [_P, C.POINTER(TssConfig),
                                     C.POINTER(TssDamBreak), _I,
                                     C.POINTER(_I)]),
    "tss_batch_remove_scene": (_I, [_P, _I]),
    "tss_batch_step": (_I, [_P, _I]),
    "tss_batch_step_graph": (_I, [_P, _I]),
    "tss_batch_counts": (_I, [_P, _I, C.POINTER(TssCounts)]),
    "tss_batch_read": (_I, [_P, _I, _I, _P, _SZ]),
    "tss_frame_size": (_SZ, [C.c_uint32, _I, _I, _I]),
    "tss_snapshot_size": (_SZ, [_I, _I, _I, _I, _I]),
    "tss_snapshot": (_I, [_P, _P, _SZ, C.POINTER(_SZ)]),
    "tss_batch_snapshot": (_I, [_P, _I, _P, _SZ, C.POINTER(_SZ)]),
    "tss_create_from_snapshot": (_I, [_P, _SZ, C.POINTER(_P)]),
    "tss_batch_restore": (_I, [_P, _P, _SZ, C.POINTER(_I)]),
    "tss_frame": (_I, [_P, C.c_uint32, _P, _SZ, C.POINTER(_SZ)]),
    "tss_batch_frame": (_I, [_P, _I, C.c_uint32, _P, _SZ, C.POINTER(_SZ)]),
    "tss_batch_frame_noalloc": (_I, [_P, _I, C.c_uint32, _P, _SZ, C.POINTER(_SZ)]),
    "tss_batch_conservation": (_I, [_P, _I, _P, _P]),
    "tss_batch_synchronize": (_I, [_P]),
    "tss_batch_num_scenes": (_I, [_P, C.POINTER(_I)]),
    "tss_batch_get_tuning": (_I, [_P, C.POINTER(TssTuning)]),
    "tss_batch_set_tuning": (_I, [_P, C.POINTER(TssTuning)]),
    "tss_batch_autotune": (_I, [_P, _I, _I, C.POINTER(TssTuning),
                                C.POINTER(_I), _P, _I, C.POINTER(_I)]),
    "tss_batch_destroy": (None, [_P]),
    "tss_serve": (_I, [C.POINTER(TssBatchDesc), C.c_longlong,
                       C.POINTER(TssRequest), _I, C.POINTER(TssServeStats)]),
    "tss_scene_bytes": (C.c_longlong, [C.POINTER(TssConfig), _I, _I]),
    "tss_canvas_count": (_I, [_I, _I, _I]),
    "tss_scene_dam_break": (_I, [C.c_uint64, C.c_uint64, _I, _I, _F, _F, _I,
                                 _I, _I, _P, _P]),
    "tss_splitmix64_next": (C.c_uint64, [C.POINTER(C.c_uint64)]),
    "tss_bandwidth_probe": (_I, [_SZ, _I, C.POINTER(C.c_double)]),
    "tss_gather_probe": (_I, [_SZ, _I, _I, C.POINTER(C.c_double)]),
    "tss_last_error": (C.c_char_p, []),
    "tss_version": (C.c_char_p, []),
}

_lib = None


def lib() -> C.CDLL:
    """Load libtss.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libtss.so not built at {LIB_PATH}; run "
                "`python -m twoscalesplit_amd.build`")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class TssError(RuntimeError):
    def __init__(self, code: int, what: str):
        msg = lib().tss_last_error().decode(errors="replace")
        super().__init__(f"{what} failed ({code}): {msg}")
        self.code = code


def check(rc: int, what: str) -> None:
    if rc != TSS_OK:
        raise TssError(rc, what)


def ptr(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)
