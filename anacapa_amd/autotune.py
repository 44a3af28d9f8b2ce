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

PATH_CHOICES = (1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26)
TILE_CHOICES = (4, 8, 16)


def device_key(name: str) -> str:
    """Cache key of a device: the architecture and CU count the library reports in parentheses
    (the marketing name in front of it depends on a lookup table the runtime may not find)."""
    i, j = name.rfind("("), name.rfind(")")
    return name[i + 1:j] if 0 <= i < j else name


def load(device: str, workload: str):
    device = device_key(device)
    try:
        with open(CACHE) as f:
            return json.load(f).get(device, {}).get(workload)
    except (OSError, ValueError):
        return None


def save(device: str, workload: str, entry: dict):
    device = device_key(device)
    data = {}
    try:
        with open(CACHE) as f:
            data = json.load(f)
    except (OSError, ValueError):
        pass
    data.setdefault(device, {})[workload] = entry
    with open(CACHE, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def tune(be, run, workload: str, paths=PATH_CHOICES, tiles=TILE_CHOICES, repeats=2, log=print,
         warm=None) -> dict:
    """`run()` renders the workload once on backend `be`; returns the best setting (and stores it).
    `warm()` (default: `run`) is called once per setting first so buffer growth is not timed."""
    results = []
    for p in paths:
        for t in tiles:
            be.set_launch(max_paths=p, tile=t)
            (warm or run)()
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


def plain_key(w, h, spp, depth):
    return f"plain{w}x{h}x{spp}d{depth}"


def batch_key(views, seed, depth):
    return f"batch{views}s{seed}d{depth}"


def main(argv=None):
    """Tune the 1M-triangle GGX-spheres workloads on GPU 0: the plain 1920x1080x16 frame and the
    batched 64-view set (seeded orbit views of mixed resolution and spp)."""
    import argparse
    import sys
    from . import scenes
    from .integrator import HipPathIntegrator
    ap = argparse.ArgumentParser()
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--repeats", type=int, default=1)
    ap.add_argument("--only", choices=("plain", "batch"), default=None)
    a = ap.parse_args(argv)
    out = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    sc = scenes.ggx_spheres(max_depth=a.depth)
    be = HipPathIntegrator(0)
    if not be.is_valid():
        raise SystemExit(f"HIP backend invalid: {be.last_error()}")
    be.prepare(sc)
    res = {}
    if a.only in (None, "plain"):
        def frame():
            be.clear_accum()
            be.render_frame(0, sc.spp)
        frame()
        e = tune(be, frame, plain_key(sc.width, sc.height, sc.spp, a.depth), repeats=max(2, a.repeats), log=out)
        res["plain"] = {k: e[k] for k in ("max_paths", "tile", "seconds")}
    if a.only in (None, "batch"):
        specs = scenes.batch_views(sc, a.views, a.seed)
        cams = [be.pinhole(v.frm, v.at, (0.0, 1.0, 0.0), sc.vfov, v.width, v.height) for v in specs]
        counts = [v.spp for v in specs]
        starts = [0] * len(cams)
        big = max(range(len(cams)), key=lambda i: cams[i].width * cams[i].height)

        def run():
            be.render_batch(cams, starts, counts, a.depth)

        def warm():
            be.render_batch([cams[big]], [0], [8], a.depth)

        run()                                       # first call sizes the shared film buffer
        e = tune(be, run, batch_key(a.views, a.seed, a.depth), repeats=a.repeats, log=out, warm=warm)
        res["batch"] = {k: e[k] for k in ("max_paths", "tile", "seconds")}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
