"""Launch-shape autotuner for the wavefront backend.

Sweeps the two knobs of ``ahip_set_launch`` — paths per wavefront pass (queue capacity) and the
pixel tile one 64-lane wavefront covers (4x16, 8x8, 16x4) — on a workload, times each setting with
the library's own HIP-event render time, and caches the best per device name in
``anacapa_amd/tuning.json``.  Rendering results do not depend on the knobs (only the grouping of
per-pixel float sums does), so tuning never changes the image beyond rounding.
"""
from __future__ import annotations

import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuning.json")

PATH_CHOICES = (1 << 20, 1 << 21, 1 << 22, 1 << 23)
TILE_CHOICES = (4, 8, 16)


def load(device: str, workload: str):
    try:
        with open(CACHE) as f:
            return json.load(f).get(device, {}).get(workload)
    except (OSError, ValueError):
        return None


def save(device: str, workload: str, entry: dict):
    data = {}
    try:
        with open(CACHE) as f:
            data = json.load(f)
    except (OSError, ValueError):
        pass
    data.setdefault(device, {})[workload] = entry
    with open(CACHE, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def tune(be, run, workload: str, paths=PATH_CHOICES, tiles=TILE_CHOICES, repeats=2, log=print) -> dict:
    """`run()` renders the workload once on backend `be`; returns the best setting (and stores it)."""
    results = []
    for p in paths:
        for t in tiles:
            be.set_launch(max_paths=p, tile=t)
            run()                                   # warm: allocation at this capacity
            best = float("inf")
            for _ in range(repeats):
                t0 = time.perf_counter()
                run()
                best = min(best, time.perf_counter() - t0)
            results.append({"max_paths": p, "tile": t, "seconds": best})
            log(f"[autotune] {workload}: max_paths={p} tile={t}x{64 // t}: {best * 1e3:.1f} ms")
    win = min(results, key=lambda r: r["seconds"])
    entry = {"max_paths": win["max_paths"], "tile": win["tile"], "seconds": win["seconds"], "sweep": results}
    save(be.name(), workload, entry)
    be.set_launch(max_paths=win["max_paths"], tile=win["tile"])
    return entry
