"""Autotuner of the engine's launch shapes.

The engine exposes three launch choices that change timing but never a result (each work unit,
voxel and tile is computed the same way whatever block runs it):

* ``superpose_grid``: work-groups launched per superposition radius class (the kernels walk
  their unit list with a grid stride);
* ``transfer_shape``: the transfer's 256-voxel brick, 0..4 = (8,8,4) (16,16,1) (16,8,2)
  (32,8,1) (8,4,8) along (cross, cross, depth) axes of the dose grid;
* ``one_stream``: the radius classes on the engine stream only (1) or also on side streams (0).

``tune`` measures the workload the caller gives it (one beam into a device dose buffer) under each
candidate, one knob at a time in the order above (coordinate descent, the others held at their
current best), and leaves the engine at the fastest setting. Each candidate is timed as the median
wall time of ``reps`` beams (queue + wait) after one warm-up beam. Results can be cached in a JSON
file keyed by a workload signature (device name, grid and beam sizes).
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

CANDIDATES = {
    "superpose_grid": (1024, 2048, 4096, 8192, 16384),
    "transfer_shape": (0, 1, 2, 3, 4),
    "one_stream": (0, 1),
}


def _timer(engine, beam, dose, reps: int):
    def run() -> float:
        engine.beam_dose(beam, dose)
        engine.sync()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            engine.beam_dose(beam, dose)
            engine.sync()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))
    return run


def tune(engine, beam: dict, dose, reps: int = 7, candidates: dict | None = None, measure=None,
         min_gain: float = 0.02) -> dict:
    """Pick the fastest launch shapes for ``beam`` on ``engine`` (left applied). A value replaces
    the knob's current one only when it is at least ``min_gain`` (fraction) faster, so timing
    noise does not move the setting. Returns {"best": {knob: value}, "ms": {knob: {value: median
    ms}}, "default_ms", "best_ms"}. ``measure`` (a callable returning ms) replaces the
    wall-clock timer, for tests."""
    cands = dict(CANDIDATES if candidates is None else candidates)
    current = engine.tuning()
    if "one_stream" in cands:
        # an engine without side streams (batch lanes, server slots) cannot leave one-stream mode
        try:
            engine.set_tuning(one_stream=0)
        except RuntimeError:
            cands.pop("one_stream")
        engine.set_tuning(one_stream=current["one_stream"])
    run = measure or _timer(engine, beam, dose, reps)
    default_ms = run()
    best = dict(current)
    best_ms = default_ms
    table = {}
    for knob, values in cands.items():
        table[knob] = {}
        for v in values:
            engine.set_tuning(**{knob: v})
            ms = run()
            table[knob][v] = ms
        v_best = min(table[knob], key=table[knob].get)
        if best[knob] in table[knob] and table[knob][v_best] > (1.0 - min_gain) * table[knob][best[knob]]:
            v_best = best[knob]
        engine.set_tuning(**{knob: v_best})
        best[knob] = v_best
        best_ms = table[knob][v_best]
    return {"best": best, "ms": table, "default_ms": default_ms, "best_ms": best_ms}


def signature(case: dict, beam: dict, device_name: str = "") -> str:
    """Cache key of a workload: device, CT and dose grids, layers and spots."""
    n_spots = int(np.asarray(beam.get("spot_weights", [])).size)
    return "|".join([device_name, "x".join(map(str, case["ct_dims"])), "x".join(map(str, case["dose_dims"])),
                     f"L{len(beam['energies'])}", f"S{n_spots}"])


def load(path: str, key: str) -> dict | None:
    try:
        with open(path) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def save(path: str, key: str, best: dict) -> None:
    data = {}
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        pass
    data[key] = best
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)
