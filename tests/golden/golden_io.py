"""Reading and writing of the golden step fixtures (plain numpy .npz, no
pickles, plus a JSON manifest with the configuration).

A fixture holds one step of a small scene:
  * ``pre.*``   -- the state before the step, in the field names of
                   ``Solver.snapshot()`` (POS_L ... LIST_G, DT) plus counts;
  * ``grid.*``  -- hashes, stable sort order, cell ranges (bit-exact);
  * ``nbr.*``   -- neighbour counts and id checksums per target and class;
  * ``dens.*``  -- FP64 density and pressure by id;
  * ``force.*`` -- FP64 accelerations by id and their conditioning scale;
  * ``post.*``  -- the state after one FP64 oracle step (lists, states,
                   ages, positions, velocities, dt) and the event counts.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

COUNT_KEYS = ("n_low", "n_high", "n_groups")


def config_to_dict(cfg) -> dict:
    out = {}
    for name, typ in cfg._fields_:
        v = getattr(cfg, name)
        if isinstance(v, C.Array):
            v = list(v)
        out[name] = v
    return out


def config_from_dict(d: dict):
    """Fields a fixture predates keep the values of a zeroed config (every
    optional control off)."""
    from twoscalesplit_amd import _native as N
    cfg = N.TssConfig()
    for name, typ in cfg._fields_:
        if name not in d:
            continue
        v = d[name]
        if isinstance(v, list):
            arr = getattr(cfg, name)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(cfg, name, v)
    return cfg


def save(name: str, cfg, meta: dict, arrays: dict) -> None:
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump({"config": config_to_dict(cfg), "meta": meta}, f, indent=1,
                  sort_keys=True)


def load(name: str):
    with open(os.path.join(HERE, name + ".json")) as f:
        man = json.load(f)
    with np.load(os.path.join(HERE, name + ".npz")) as z:
        arrays = {k: z[k] for k in z.files}
    return config_from_dict(man["config"]), man["meta"], arrays


def snapshot_of(arrays: dict, prefix: str = "pre.") -> dict:
    """The pre-step state as a ``Solver.snapshot()``-style dict."""
    snap = {k[len(prefix):]: v for k, v in arrays.items() if k.startswith(prefix)}
    c = snap.pop("counts")
    snap["counts"] = {k: int(c[i]) for i, k in enumerate(COUNT_KEYS)}
    return snap


def names() -> list[str]:
    return sorted(f[:-5] for f in os.listdir(HERE) if f.endswith(".json"))
