"""Launch tuning of the persistent kernel, stored per mechanism.

``tuning/<mech>.json`` holds the launch parameters the autotuner
(``autotune.py``) selected on an MI355X and, under ``profiles``, HBM traffic
per launch measured with rocprofv3 for named workloads (the FETCH_SIZE and
WRITE_SIZE counters scaled by their reading for a copy kernel of known size;
the summaries are under profiles/ at the repository root).  Without a file
the library's compiled-in defaults apply.
"""
from __future__ import annotations

import json

from . import PKG_DIR
from . import lib as _lib

TUNE_DIR = PKG_DIR / "tuning"
LAUNCH_KEYS = ("block", "blocks_per_cu", "cells_per_wave", "grain", "variant")


def load(mech: str = "fullchem") -> dict:
    p = TUNE_DIR / f"{mech}.json"
    return json.loads(p.read_text()) if p.exists() else {}


def apply(lib, mech: str = "fullchem") -> dict:
    """Apply the stored launch parameters; returns the parameters in effect."""
    t = load(mech).get("launch", {})
    if t:
        _lib.set_tuning(lib, **{k: int(t[k]) for k in LAUNCH_KEYS if k in t})
    return _lib.get_tuning(lib)


def profile_traffic(name: str, mech: str = "fullchem") -> dict | None:
    return load(mech).get("profiles", {}).get(name)
