"""Python handles over libtss: a single-scene ``Solver`` (the reference's
WCSPHSolver, TwoScale2D/WCSPHSolver.h:37-63) and a multi-scene ``Batch``.

Arrays cross the boundary as numpy host arrays; the device state stays
resident in the library between calls.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

_DTYPES = {
    "POS": (np.float32, 2), "VEL": (np.float32, 2), "ACC": (np.float32, 2),
    "RHO": (np.float32, 1), "P": (np.float32, 1), "VIS": (np.float32, 1),
    "STATE": (np.uint8, 1), "AGE": (np.int32, 1),
}


# tss_phase_times order (include/tss.h)
PHASES = ["grid", "density_low", "density_high", "force_low", "force_high",
          "reduce", "adapt", "finalize"]
# tss_work_counts order
WORK = ["cand_ll", "cand_lh", "cand_hh", "cand_hl", "dens_pairs_l",
        "dens_pairs_h", "force_pairs_l", "force_pairs_h", "wall_l", "wall_h",
        "targets_l", "targets_h", "cand_hl_density"]


def _field_shape(name: str, cnt: dict) -> tuple[np.dtype, tuple]:
    """dtype and shape of a field for the current counts."""
    if name in ("DT",):
        return np.float32, (1,)
    if name == "POS_B":
        return np.float32, (cnt["n_bnd"], 2)
    if name == "BND_SORTED":
        return np.int32, (cnt["n_bnd"],)
    if name in ("BND_START", "BND_END", "CELL_START_L", "CELL_END_L"):
        return np.int32, (cnt["cells_low"],)
    if name in ("CELL_START_H", "CELL_END_H"):
        return np.int32, (cnt["cells_high"],)
    if name == "LIST_L":
        return np.int32, (cnt["n_low"],)
    if name == "LIST_H":
        return np.int32, (cnt["n_high"],)
    if name == "LIST_G":
        return np.int32, (cnt["n_groups"],)
    if name in ("SORTED_L", "KEYS_L"):
        return np.int32, (cnt["n_low_last"],)
    if name in ("SORTED_H", "KEYS_H"):
        return np.int32, (cnt["n_high_last"],)
    base, lvl = name.rsplit("_", 1)
    dt, comp = _DTYPES[base]
    n = cnt["cap_low"] * (4 if lvl == "H" else 1)
    return dt, ((n, comp) if comp > 1 else (n,))


def config_init(xs, ys, xe, ye, h, H, rho0, B, c, alpha, sigma, dt
                ) -> N.TssConfig:
    """WCSPHConfig's constructor (TwoScale2D/WCSPHSolver.cu:2287-2332)."""
    cfg = N.TssConfig()
    N.check(N.lib().tss_config_init(C.byref(cfg), xs, ys, xe, ye, h, H, rho0,
                                    B, c, alpha, sigma, dt),
            "tss_config_init")
    return cfg


# tss_snapshot_header: magic, version, header_bytes, config_bytes (4 x u32),
# total_bytes (u64), then the configuration
_CFG_OFFSET = 24


def _snapshot_buffer(snap) -> np.ndarray:
    if isinstance(snap, str):
        return np.fromfile(snap, np.uint8)
    if isinstance(snap, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(snap), np.uint8).copy()
    return np.ascontiguousarray(snap, np.uint8)


def _take_snapshot(call) -> np.ndarray:
    need = C.c_size_t()
    N.check(call(None, 0, C.byref(need)), "snapshot size")
    buf = np.zeros(need.value, np.uint8)
    got = C.c_size_t()
    N.check(call(buf.ctypes.data, buf.nbytes, C.byref(got)), "snapshot")
    assert got.value == buf.nbytes
    return buf


class Solver:
    """One scene resident on the GPU."""

    def __init__(self, cfg: N.TssConfig, low_xy: np.ndarray, cap_low: int,
                 bnd_xy: np.ndarray):
        lib = N.lib()
        low = np.ascontiguousarray(low_xy, dtype=np.float32).reshape(-1, 2)
        bnd = np.ascontiguousarray(bnd_xy, dtype=np.float32).reshape(-1, 2)
        h = C.c_void_p()
        N.check(lib.tss_create(C.byref(cfg), N.ptr(low), len(low), int(cap_low),
                               N.ptr(bnd), len(bnd), C.byref(h)), "tss_create")
        self._h = h
        self.cfg = cfg

    @classmethod
    def from_snapshot(cls, snap) -> "Solver":
        """A solver holding the scene of a snapshot (bytes, a uint8 array or
        a file path written by ``save``); stepping it continues the
        snapshotted run bit for bit."""
        buf = _snapshot_buffer(snap)
        h = C.c_void_p()
        N.check(N.lib().tss_create_from_snapshot(buf.ctypes.data, buf.nbytes,
                                                 C.byref(h)),
                "tss_create_from_snapshot")
        self = cls.__new__(cls)
        self._h = h
        self.cfg = N.TssConfig.from_buffer_copy(
            buf[_CFG_OFFSET:_CFG_OFFSET + C.sizeof(N.TssConfig)].tobytes())
        return self

    def save_state(self) -> np.ndarray:
        """Snapshot of the scene (tss_snapshot) as a uint8 array."""
        return _take_snapshot(lambda p, n, out: N.lib().tss_snapshot(
            self._h, p, n, out))

    def save(self, path: str) -> int:
        """Writes the snapshot to a file; returns its size."""
        s = self.save_state()
        s.tofile(path)
        return s.nbytes

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            N.lib().tss_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- stepping ----------------------------------------------------------
    def step(self, n: int = 1) -> None:
        N.check(N.lib().tss_step(self._h, n), "tss_step")

    def step_graph(self, n: int = 1) -> None:
        N.check(N.lib().tss_step_graph(self._h, n), "tss_step_graph")

    def synchronize(self) -> None:
        N.check(N.lib().tss_synchronize(self._h), "tss_synchronize")

    # -- state -------------------------------------------------------------
    def counts(self) -> dict:
        c = N.TssCounts()
        N.check(N.lib().tss_counts(self._h, C.byref(c)), "tss_counts")
        return c.as_dict()

    def read(self, name: str, cnt: dict | None = None) -> np.ndarray:
        cnt = cnt or self.counts()
        dt, shape = _field_shape(name, cnt)
        out = np.zeros(shape, dtype=dt)
        N.check(N.lib().tss_read(self._h, N.FIELD[name], N.ptr(out),
                                 out.nbytes), f"tss_read({name})")
        return out

    def write(self, name: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        N.check(N.lib().tss_write(self._h, N.FIELD[name], N.ptr(a), a.nbytes),
                f"tss_write({name})")

    def set_counts(self, n_low: int, n_high: int, n_groups: int) -> None:
        N.check(N.lib().tss_set_counts(self._h, n_low, n_high, n_groups),
                "tss_set_counts")

    def device_ptr(self, name: str) -> int:
        p = C.c_void_p()
        N.check(N.lib().tss_device_ptr(self._h, N.FIELD[name], C.byref(p)),
                "tss_device_ptr")
        return p.value

    def conservation(self) -> tuple[np.ndarray, np.ndarray]:
        out = np.zeros(5, np.float64)
        err = np.zeros(5, np.float64)
        N.check(N.lib().tss_conservation(self._h, N.ptr(out), N.ptr(err)),
                "tss_conservation")
        return out, err

    def neighbors(self, level: int) -> tuple[np.ndarray, np.ndarray]:
        cnt = self.counts()
        n = cnt["n_high"] if level else cnt["n_low"]
        c = np.zeros((n, 5), np.uint32)
        s = np.zeros((n, 5), np.uint32)
        N.check(N.lib().tss_neighbors(self._h, level, N.ptr(c), N.ptr(s), n),
                "tss_neighbors")
        return c, s

    def pair_sets(self, level: int) -> tuple[np.ndarray, np.ndarray]:
        """The pairs the step's density and force kernels add, per listed
        particle of a level (tss_pair_sets; the neighbors() categories)."""
        cnt = self.counts()
        n = cnt["n_high"] if level else cnt["n_low"]
        c = np.zeros((n, 5), np.uint32)
        s = np.zeros((n, 5), np.uint32)
        N.check(N.lib().tss_pair_sets(self._h, level, N.ptr(c), N.ptr(s), n),
                "tss_pair_sets")
        return c, s

    def enable_timing(self, on: bool = True) -> None:
        N.check(N.lib().tss_enable_timing(self._h, 1 if on else 0),
                "tss_enable_timing")

    def phase_times(self) -> dict:
        """Mean device ms per phase over the steps since enable_timing."""
        out = np.zeros(len(PHASES), np.float32)
        N.check(N.lib().tss_phase_times(self._h, N.ptr(out)),
                "tss_phase_times")
        return {k: float(v) for k, v in zip(PHASES, out)}

    def work_counts(self) -> dict:
        """Exact candidate / pair / wall / target counts of the density and
        force kernels on the current state, per target level."""
        out = np.zeros(len(WORK), np.uint64)
        N.check(N.lib().tss_work_counts(self._h, N.ptr(out)),
                "tss_work_counts")
        return {k: int(v) for k, v in zip(WORK, out)}

    # -- launch shapes -------------------------------------------------------
    def get_tuning(self) -> N.TssTuning:
        t = N.TssTuning()
        N.check(N.lib().tss_get_tuning(self._h, C.byref(t)), "get_tuning")
        return t

    def set_tuning(self, t: N.TssTuning) -> None:
        N.check(N.lib().tss_set_tuning(self._h, C.byref(t)), "set_tuning")

    def set_grid_sort(self, path: int) -> None:
        """Sort of the grid build: N.SORT_AUTO, SORT_RADIX, SORT_COUNT or
        SORT_SMALL (tss_set_grid_sort); results do not depend on it."""
        N.check(N.lib().tss_set_grid_sort(self._h, path), "set_grid_sort")

    def autotune(self, steps_per_trial: int = 1, rounds: int = 2) -> dict:
        """Native autotuner (advances the scene by real steps)."""
        t = N.TssTuning()
        taken = C.c_int()
        log = np.zeros((64, 8), np.float32)
        nlog = C.c_int()
        N.check(N.lib().tss_autotune(self._h, steps_per_trial, rounds,
                                   C.byref(t), C.byref(taken), N.ptr(log), 64,
                                   C.byref(nlog)), "autotune")
        return {"tuning": t.as_dict(), "steps_taken": taken.value,
                "trials": log[:nlog.value].tolist(), "_struct": t}

    def frame(self, fields: int = 255):
        """The current frame (frames.Frame): list-ordered particles of both
        levels with the requested fields (frames.ID | frames.POS | ...)."""
        from . import frames
        return frames.grab(self, fields)

    def snapshot(self) -> dict:
        """Every by-id field, the lists and dt (synchronising)."""
        cnt = self.counts()
        names = ["POS_L", "VEL_L", "ACC_L", "RHO_L", "P_L", "VIS_L", "STATE_L",
                 "AGE_L", "POS_H", "VEL_H", "ACC_H", "RHO_H", "P_H", "VIS_H",
                 "STATE_H", "LIST_L", "LIST_H", "LIST_G", "DT"]
        snap = {n: self.read(n, cnt) for n in names}
        snap["counts"] = cnt
        return snap

    def restore(self, snap: dict) -> None:
        """Re-seed the solver with a snapshot taken from a solver of the
        same configuration."""
        for n in ["POS_L", "VEL_L", "ACC_L", "RHO_L", "P_L", "STATE_L",
                  "AGE_L", "POS_H", "VEL_H", "ACC_H", "RHO_H", "P_H",
                  "STATE_H", "LIST_L", "LIST_H", "LIST_G", "DT"]:
            self.write(n, snap[n])
        c = snap["counts"]
        self.set_counts(c["n_low"], c["n_high"], c["n_groups"])


class Batch:
    """Many independent scenes advanced by one launch per phase."""

    def __init__(self, max_scenes: int, cap_low: int, cap_bnd: int,
                 cells_low: int, cells_high: int):
        self.desc = N.TssBatchDesc(max_scenes, cap_low, cap_bnd, cells_low,
                                   cells_high)
        h = C.c_void_p()
        N.check(N.lib().tss_batch_create(C.byref(self.desc), C.byref(h)),
                "tss_batch_create")
        self._h = h

    @staticmethod
    def memory_bytes(max_scenes, cap_low, cap_bnd, cells_low, cells_high
                     ) -> int:
        d = N.TssBatchDesc(max_scenes, cap_low, cap_bnd, cells_low, cells_high)
        b = C.c_longlong()
        N.check(N.lib().tss_batch_memory_bytes(C.byref(d), C.byref(b)),
                "tss_batch_memory_bytes")
        return b.value

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            N.lib().tss_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_scene(self, cfg: N.TssConfig, low_xy: np.ndarray, cap_low: int,
                  bnd_xy: np.ndarray) -> int:
        low = np.ascontiguousarray(low_xy, dtype=np.float32).reshape(-1, 2)
        bnd = np.ascontiguousarray(bnd_xy, dtype=np.float32).reshape(-1, 2)
        sid = C.c_int()
        N.check(N.lib().tss_batch_add_scene(self._h, C.byref(cfg), N.ptr(low),
                                            len(low), int(cap_low), N.ptr(bnd),
                                            len(bnd), C.byref(sid)),
                "tss_batch_add_scene")
        return sid.value

    def add_dam_break(self, cfg: N.TssConfig, gen: N.TssDamBreak,
                      cap_low: int) -> int:
        """A seeded dam-break scene whose fluid block the device generates
        (positions bit for bit those of ``scenes.generate``)."""
        sid = C.c_int()
        N.check(N.lib().tss_batch_add_dam_break(self._h, C.byref(cfg),
                                                C.byref(gen), int(cap_low),
                                                C.byref(sid)),
                "tss_batch_add_dam_break")
        return sid.value

    def remove_scene(self, sid: int) -> None:
        N.check(N.lib().tss_batch_remove_scene(self._h, sid),
                "tss_batch_remove_scene")

    def step(self, n: int = 1) -> None:
        N.check(N.lib().tss_batch_step(self._h, n), "tss_batch_step")

    def step_graph(self, n: int = 1) -> None:
        N.check(N.lib().tss_batch_step_graph(self._h, n),
                "tss_batch_step_graph")

    def synchronize(self) -> None:
        N.check(N.lib().tss_batch_synchronize(self._h), "tss_batch_synchronize")

    def counts(self, sid: int) -> dict:
        c = N.TssCounts()
        N.check(N.lib().tss_batch_counts(self._h, sid, C.byref(c)),
                "tss_batch_counts")
        return c.as_dict()

    def read(self, sid: int, name: str, cnt: dict | None = None
             ) -> np.ndarray:
        cnt = cnt or self.counts(sid)
        dt, shape = _field_shape(name, cnt)
        out = np.zeros(shape, dtype=dt)
        N.check(N.lib().tss_batch_read(self._h, sid, N.FIELD[name],
                                       N.ptr(out), out.nbytes),
                f"tss_batch_read({name})")
        return out

    def snapshot(self, sid: int) -> np.ndarray:
        """Snapshot of one scene (tss_batch_snapshot) as a uint8 array."""
        return _take_snapshot(lambda p, n, out: N.lib().tss_batch_snapshot(
            self._h, sid, p, n, out))

    def restore(self, snap) -> int:
        """Adds the scene of a snapshot to the batch; returns its id."""
        buf = _snapshot_buffer(snap)
        sid = C.c_int()
        N.check(N.lib().tss_batch_restore(self._h, buf.ctypes.data, buf.nbytes,
                                          C.byref(sid)), "tss_batch_restore")
        return sid.value

    def frame(self, sid: int, fields: int = 255):
        """The current frame of one scene (frames.Frame)."""
        from . import frames
        return frames.grab(self, fields, sid)

    def conservation(self, sid: int) -> tuple[np.ndarray, np.ndarray]:
        out = np.zeros(5, np.float64)
        err = np.zeros(5, np.float64)
        N.check(N.lib().tss_batch_conservation(self._h, sid, N.ptr(out),
                                               N.ptr(err)),
                "tss_batch_conservation")
        return out, err

    def num_scenes(self) -> int:
        n = C.c_int()
        N.check(N.lib().tss_batch_num_scenes(self._h, C.byref(n)),
                "tss_batch_num_scenes")
        return n.value

    def get_tuning(self) -> N.TssTuning:
        t = N.TssTuning()
        N.check(N.lib().tss_batch_get_tuning(self._h, C.byref(t)),
                "tss_batch_get_tuning")
        return t

    def set_grid_sort(self, path: int) -> None:
        """Sort of the grid build (tss_batch_set_grid_sort)."""
        N.check(N.lib().tss_batch_set_grid_sort(self._h, path), "set_grid_sort")

    def set_tuning(self, t: N.TssTuning) -> None:
        N.check(N.lib().tss_batch_set_tuning(self._h, C.byref(t)),
                "tss_batch_set_tuning")

    def autotune(self, steps_per_trial: int = 1, rounds: int = 2) -> dict:
        """Native autotuner (advances every scene by real steps)."""
        t = N.TssTuning()
        taken = C.c_int()
        log = np.zeros((64, 8), np.float32)
        nlog = C.c_int()
        N.check(N.lib().tss_batch_autotune(self._h, steps_per_trial, rounds,
                                           C.byref(t), C.byref(taken),
                                           N.ptr(log), 64, C.byref(nlog)),
                "tss_batch_autotune")
        return {"tuning": t.as_dict(), "steps_taken": taken.value,
                "trials": log[:nlog.value].tolist(), "_struct": t}
