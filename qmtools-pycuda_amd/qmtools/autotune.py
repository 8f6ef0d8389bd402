"""Autotuner for the density contraction's tile shapes.

The library times every candidate tile shape on a slab of the caller's grid
(`qm_autotune_density`) and keeps the fastest in its context.  This module
runs it once per (device, precision, padded basis size) and caches the choice
in a JSON file (default `~/.cache/qmtools/tune.json`, or $QMTOOLS_TUNE_CACHE),
so later processes only apply the stored value.
"""
from __future__ import annotations

import ctypes
import json
import os

from . import _device
from . import _native as nat

_KNOBS = {"split_variant": nat.QM_TUNE_SPLIT_VARIANT, "f64_variant": nat.QM_TUNE_F64_VARIANT,
          "jacobi_zs": nat.QM_TUNE_JACOBI_ZS, "jacobi_pairs": nat.QM_TUNE_JACOBI_PAIRS,
          "tile_skip": nat.QM_TUNE_TILE_SKIP}
_PREC = {"fp64": nat.QM_PREC_FP64, "mixed": nat.QM_PREC_F16X3, "f16": nat.QM_PREC_F16X3,
         "bf16": nat.QM_PREC_BF16X3}


def cache_path() -> str:
    return os.environ.get("QMTOOLS_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "qmtools", "tune.json"))


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _save(path, table):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(table, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def get(knob: str) -> int:
    v = ctypes.c_int(0)
    nat.call("qm_tune_get", _KNOBS[knob], ctypes.byref(v))
    return v.value


def set(knob: str, value: int) -> None:  # noqa: A001 - mirrors the C-ABI name
    nat.call("qm_tune_set", _KNOBS[knob], int(value))


def key_for(molecule, precision: str) -> str:
    t = _device.torch()
    name = t.cuda.get_device_name(_device.device_index()) if t.cuda.is_available() else "cpu"
    nbf_pad = (molecule.nbf + 63) // 64 * 64
    return f"{name}|{precision}|nbf_pad={nbf_pad}"


def autotune_density(molecule, grid, precision="mixed", use_cache=True, path=None):
    """Choose the contraction tile shape for this molecule size; returns {knob: value, 'ms': t}."""
    from .qmtools import device_molecule
    path = path or cache_path()
    key = key_for(molecule, precision)
    table = _load(path) if use_cache else {}
    if key in table:
        for k, v in table[key].items():
            if k in _KNOBS:
                set(k, v)
        return table[key]
    dm = device_molecule(molecule)
    best = ctypes.c_double(0.0)
    nat.call("qm_autotune_density", dm.handle, nat.stream(), ctypes.byref(nat.geom(grid)), _PREC[precision],
             ctypes.byref(best))
    relevant = ["f64_variant"] if precision == "fp64" else ["split_variant"]
    entry = {k: get(k) for k in relevant}
    entry["ms"] = best.value
    if use_cache:
        table = _load(path)
        table[key] = entry
        _save(path, table)
    return entry
