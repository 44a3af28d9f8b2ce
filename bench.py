#!/usr/bin/env python3
"""Benchmark: Msamples/s of the MI355X wavefront Path+NEE backend.

Workload (one "step" = one full frame): the 1M-triangle GGX-spheres scene — 1024 displaced
geodesic spheres (980 triangles each) with per-sphere standard_surface materials on a 2-triangle
ground, lit by a procedural 2048x1024 HDRI dome with 2D importance sampling — rendered at
1920x1080, 16 samples per pixel, max depth 8, Russian roulette from bounce 2, Halton seed 0.
Scene data is synthetic (procedural, seed 0) and resident in HBM before timing starts.

Prints one JSON line with the metric, a roofline object for the dominant kernel family (BVH
traversal: extend / shadow / MIS) and a CPU baseline (the oracle's port of the reference CPU
integrator, multi-threaded on the host, on a bounded row sample of the same frame).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


DEFAULT_MAX_PATHS = 1 << 25   # ahip_set_launch default: one 1080p x 16 spp frame per pass


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--max-paths", type=int, default=0, help="paths per wavefront pass (0 = library default)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample duration")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the untimed node/triangle counting pass")
    ap.add_argument("--batch-views", type=int, default=64,
                    help="views of the batched-form comparison (0 = skip it)")
    ap.add_argument("--batch-seed", type=int, default=0)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward timing")
    ap.add_argument("--no-configs", action="store_true", help="skip the Cornell and hair frames")
    ap.add_argument("--no-bdpt", action="store_true", help="skip the BDPT Cornell frame")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run one elementwise copy of this many MiB (read + write), the known-size kernel "
                         "that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build renders on one GPU; multi-GPU sharding (split rows or sample "
                                   "ranges, one RCCL reduce of the accumulation buffer) is designed but not "
                                   "built, so --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    from anacapa_amd import scenes
    from anacapa_amd.integrator import HipPathIntegrator

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only; the library synchronises its own stream
        sync = lambda: None  # noqa: E731

    if args.copy_probe_mb:
        import torch
        n = args.copy_probe_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)   # one elementwise kernel: reads and writes copy_probe_mb MiB
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {args.copy_probe_mb} MiB read + {args.copy_probe_mb} MiB written")
        del src, dst

    t0 = time.time()
    sc = scenes.ggx_spheres(width=args.width, height=args.height, spp=args.spp, max_depth=args.depth)
    log(f"[bench] scene {sc.num_triangles} triangles generated in {time.time() - t0:.1f}s")
    be = HipPathIntegrator(0)
    if not be.is_valid():
        raise SystemExit(f"HIP backend invalid: {be.last_error()}")
    t0 = time.time()
    be.prepare(sc)
    from anacapa_amd import autotune
    launch = autotune.load(be.name(), autotune.plain_key(args.width, args.height, args.spp, args.depth))
    if args.max_paths:
        launch = {"max_paths": args.max_paths, "tile": 8}
    if launch:
        be.set_launch(max_paths=launch["max_paths"], tile=launch["tile"])
    plain_launch = launch and {"max_paths": launch["max_paths"], "tile": launch["tile"]}
    nodes, ntris, depth = be.bvh_info()
    log(f"[bench] prepare (SAH-12 build + upload) {time.time() - t0:.1f}s: {nodes} nodes, depth {depth}, "
        f"device {be.name()}")

    W, H, S = args.width, args.height, args.spp
    for i in range(args.warmup):
        be.clear_accum()
        be.render_frame(0, S)
        log(f"[bench] warmup {i + 1}/{args.warmup}: {be.stats()['render_ms']:.1f} ms")
    be.reset_stats()
    sync()
    t_start = time.perf_counter()
    for i in range(args.steps):
        be.clear_accum()
        be.render_frame(0, S)
        log(f"[bench] step {i + 1}/{args.steps}: {be.stats()['render_ms']:.1f} ms")
    sync()
    elapsed = time.perf_counter() - t_start
    st = be.stats()
    img = be.download_accum()
    assert np.isfinite(img).all() and (img[..., 3] == S).all()

    samples = W * H * S * args.steps
    msps = samples / elapsed / 1e6
    rays = st["rays_closest"] + st["rays_shadow"]
    mrays = rays / elapsed / 1e6
    trace_ms = st["trace_ms"]
    launches = max(1, st["trace_launches"])

    roofline = None
    if not args.no_traffic:
        roofline = traversal_roofline(be, args, S, trace_ms, launches, W, H)

    batched = None
    if args.batch_views:
        batched = batched_form(be, sc, args, log)

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_form(be, log)

    others = None
    if not args.no_configs:
        others = other_configs(be, log)

    bdpt = None
    if not args.no_bdpt:
        bdpt = bdpt_form(be, args, log)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sc, args)

    out = {
        "metric": "Msamples/s @1080p 1M-tri Path+NEE",
        "value": round(msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural scene, seed 0)",
        "config": {"workload": f"ggx_spheres {sc.num_triangles} tris, {W}x{H}, {S} spp, depth {args.depth}, HDRI dome",
                   "mode": "plain per-frame", "launch": plain_launch},
        "mrays_per_s": round(mrays, 2),
        "trace_ms_per_frame": round(trace_ms / args.steps, 3),
        "shade_ms_per_frame": round(st["shade_ms"] / args.steps, 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adjoint,
        "other_configs": others,
        "bdpt": bdpt,
    }
    print(json.dumps(out), flush=True)


def _profile(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        return json.load(f), "profiles/" + name


def traversal_roofline(be, args, S, trace_ms, launches, W, H):
    """Roofline of the dominant kernel family, BVH traversal (the extend, shadow and MIS launches).

    achieved: algorithmic bytes per launch / mean launch time (HIP events on the library's stream
    over the timed frames).  Per unit: 64 B per binary BVH node visited, 64 B per triangle tested,
    176 B of queue I/O per ray (a 32-B ray read, a 16-B hit write, 64-B path state read and
    written).  The node and triangle counts come from ahip_count_traffic, an untimed re-run of the
    same deterministic frame with counter-enabled traversal kernels; the kernels walk two-level
    records that each stand for a binary node's two children, so binary visits = 2 x record
    visits.  The node and triangle set (~85 MB) stays in the 256 MB Infinity Cache, so this rate
    can pass the HBM peak: what HBM itself carries is `traffic` (rocprofv3 FETCH_SIZE + WRITE_SIZE
    of this workload, copy-kernel calibrated), and the binding resource is VALU issue (`valu`:
    busy x lane utilization from the SQ counters, time-weighted over the three kernels); both come
    from profiles of this same workload committed under profiles/, named in `source`."""
    ct = be.count_traffic(0, S)
    nodes, tris = ct["nodes_visited"], ct["prims_tested"]
    rays = ct["rays_closest"] + ct["rays_shadow"]
    frame_bytes = 64.0 * 2 * nodes + 64.0 * tris + 176.0 * rays
    record_bytes = 112.0 * nodes + 48.0 * tris + 48.0 * rays   # what the kernels' records take
    per_launch = frame_bytes * args.steps / launches
    avg_launch_s = trace_ms / 1e3 / launches
    achieved = per_launch / avg_launch_s / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
           "kernel": "traversal (extend+shadow+mis launches)",
           "per_unit": "64 B per binary node visited, 64 B per triangle tested, 176 B queue I/O per ray",
           "bytes_per_launch": round(per_launch), "avg_launch_ms": round(avg_launch_s * 1e3, 4),
           "record_bytes_per_launch": round(record_bytes * args.steps / launches),
           "record_gbs": round(record_bytes * args.steps / launches / avg_launch_s / 1e9, 1),
           "node_records_per_frame": int(nodes), "binary_nodes_per_frame": int(2 * nodes),
           "tris_per_frame": int(tris), "rays_per_frame": int(rays)}
    full = (W, H, S, args.depth) == (1920, 1080, 16, 8)
    prof, src = _profile("traversal_ggx1080p16_traffic.json")
    if full and prof and prof.get("family_traffic_bytes_per_dispatch"):
        traffic = float(prof["family_traffic_bytes_per_dispatch"])
        out["traffic"] = round(traffic)
        out["hbm_measured"] = {"gbs": round(traffic / avg_launch_s / 1e9, 1),
                               "frac": round(traffic / avg_launch_s / 1e9 / HBM_PEAK_GBS, 4),
                               "traffic_over_algorithmic": round(traffic / per_launch, 3), "source": src}
    prof, src = _profile("traversal_ggx1080p16_valu.json")
    if full and prof and prof.get("sq") and prof.get("kernels"):
        per, wsum, acc = {}, 0.0, 0.0
        for name, m in prof["sq"].items():
            base = name.split("::")[-1]
            if not any(base.startswith(k) for k in ("k_extend", "k_shadow", "k_mis")):
                continue
            t = prof["kernels"].get(name, {}).get("total_ms", 0.0)
            f = min(m["valu_busy"] / 100.0, 1.0) * m["valu_util"] / 100.0
            per[base.split("<")[0]] = {"busy": round(m["valu_busy"] / 100.0, 3),
                                       "lane_utilization": round(m["valu_util"] / 100.0, 3), "frac": round(f, 3)}
            wsum += t
            acc += t * f
        if wsum > 0:
            out["valu"] = {"frac": round(acc / wsum, 3), "per_kernel": per,
                           "definition": "VALU busy (issue share of SIMD cycles, capped at 1) x lanes active per "
                                         "VALU instruction, weighted by kernel time",
                           "source": src}
    return out


def batch_views(be, sc, n, seed):
    """Cameras and spp of the batched workload: `n` seeded orbit views, resolutions from
    {256^2, 512^2, 1280x720, 1920x1080}, spp from {1, 4, 16, 64}."""
    from anacapa_amd import scenes
    specs = scenes.batch_views(sc, n, seed)
    cams = [be.pinhole(v.frm, v.at, (0.0, 1.0, 0.0), sc.vfov, v.width, v.height) for v in specs]
    return cams, [v.spp for v in specs]


def batched_form(be, sc, args, log):
    """Batched persistent form: the views rendered by one ahip_render_batch call (shared passes,
    one captured HIP graph replayed per pass, launch shape from the autotuner cache) against the
    same views rendered one after another by the plain per-frame path.  Both produce the same
    samples of the same paths, so their converged RMSE is equal; the largest per-pixel difference
    (float summation order only) is reported."""
    import numpy as np
    from anacapa_amd import autotune
    cams, counts = batch_views(be, sc, args.batch_views, args.batch_seed)
    starts = [0] * len(cams)
    samples = sum(c.width * c.height * n for c, n in zip(cams, counts))
    tuned = autotune.load(be.name(), autotune.batch_key(args.batch_views, args.batch_seed, args.depth))
    if tuned:
        be.set_launch(max_paths=tuned["max_paths"], tile=tuned["tile"])
    be.render_batch(cams[:2], starts[:2], [1, 1], args.depth)         # warm: buffers at this capacity
    be.reset_stats()
    t0 = time.perf_counter()
    outs = be.render_batch(cams, starts, counts, args.depth)
    t_batch = time.perf_counter() - t0
    st = be.stats()
    log(f"[bench] batched: {len(cams)} views, {samples / 1e6:.1f} Msamples in {t_batch * 1e3:.1f} ms, "
        f"{st['passes']} passes ({st['graph_passes']} graph replays)")
    plain = autotune.load(be.name(), autotune.plain_key(args.width, args.height, args.spp, args.depth))
    # sequential plain renders at the plain workload's tuned launch, else the library default
    be.set_launch(max_paths=plain["max_paths"] if plain else DEFAULT_MAX_PATHS, tile=plain["tile"] if plain else 8)
    t_seq, worst = 0.0, 0.0
    # occupancy of the sequential renders: every bounce's launches, timed with HIP events, against the
    # paths they carry; a launch with fewer continuation paths than the GPU holds resident threads
    # (CUs x 2048) leaves lanes idle, the only time packing several views into one pass can recover
    resident = 2048 * int(be.name().split(" CUs")[0].split()[-1]) if " CUs" in be.name() else 524288
    t_bounce = t_under = 0.0
    under_launches = launches = 0
    per_bounce = {}
    for cam, n, out in zip(cams, counts, outs):
        t0 = time.perf_counter()
        be.set_camera(cam)
        be.clear_accum()
        be.reset_stats()
        be.render_frame(0, n)
        rays, ms = be.bounce_stats()
        ref = be.download_accum()
        t_seq += time.perf_counter() - t0
        passes = max(1, int(be.stats()["passes"]))
        for b in range(len(ms)):
            launches += passes
            t_bounce += ms[b]
            pb = per_bounce.setdefault(b, {"bounce": b, "launches": 0, "paths": 0, "ms": 0.0,
                                           "underfilled_launches": 0, "underfilled_ms": 0.0})
            pb["launches"] += passes
            pb["paths"] += int(rays[b, 0])
            pb["ms"] += float(ms[b])
            if rays[b, 0] / passes < resident:
                t_under += ms[b]
                under_launches += passes
                pb["underfilled_launches"] += passes
                pb["underfilled_ms"] += float(ms[b])
        worst = max(worst, float(np.max(np.abs(ref[..., :3] - out[..., :3]) / (np.abs(ref[..., :3]) + 1.0))))
    be.set_camera(be.pinhole(sc.cam_from, sc.cam_at, sc.cam_up, sc.vfov, sc.width, sc.height))
    log(f"[bench] sequential: {t_seq * 1e3:.1f} ms (max rel diff {worst:.2e}); bounce launches under "
        f"{resident} paths: {under_launches} of {launches}, {t_under:.1f} of {t_bounce:.1f} ms")
    # the same call again: the captured pass graph is reused (no capture, no instantiation)
    be.reset_stats()
    t0 = time.perf_counter()
    be.render_batch(cams, starts, counts, args.depth)
    t_again = time.perf_counter() - t0
    captures = int(be.stats()["graph_captures"])
    log(f"[bench] batched again: {t_again * 1e3:.1f} ms, {captures} graph captures")
    return {"views": len(cams), "samples": samples, "depth": args.depth,
            "msps_batched": round(samples / t_batch / 1e6, 3),
            "msps_sequential": round(samples / t_seq / 1e6, 3),
            "speedup": round(t_seq / t_batch, 3), "passes": int(st["passes"]),
            "graph_passes": int(st["graph_passes"]), "max_rel_diff_vs_sequential": worst,
            "repeat_call_ms": round(t_again * 1e3, 1), "repeat_call_graph_captures": captures,
            "occupancy": {"resident_threads": resident, "sequential_bounce_launches": launches,
                          "underfilled_launches": under_launches, "sequential_bounce_ms": round(t_bounce, 2),
                          "underfilled_ms": round(t_under, 2),
                          "underfilled_share": round(t_under / max(t_bounce, 1e-9), 4),
                          "bound_on_speedup": round(t_seq * 1e3 / max(t_seq * 1e3 - t_under, 1e-9), 3),
                          "per_bounce": [dict(v, ms=round(v["ms"], 2), underfilled_ms=round(v["underfilled_ms"], 2))
                                         for _, v in sorted(per_bounce.items())],
                          "definition": "bounce launches (traversal + shade) of the 64 sequential renders whose "
                                        "continuation paths per pass are fewer than the GPU's resident threads"},
            "launch": tuned and {"max_paths": tuned["max_paths"], "tile": tuned["tile"]},
            "timing": "host wall clock incl. download of every view's film (comparison loop excluded)"}


def adjoint_form(be, log):
    """The inverse-rendering workload: 1M-triangle spheres with per-triangle albedo and a 1024^2 roughness texture,
    512x512, 8 spp, depth 4.  Times (HIP events) the plain wavefront forward render, the replay
    estimator's image and the adjoint pass at the perturbed initial parameters, with the image
    L2 loss against the ground-truth render as dL/dI."""
    import numpy as np
    from anacapa_amd import scenes
    sc = scenes.inverse_scene()
    be.prepare(sc)
    be.set_eval_params()
    t = be.replay(0, sc.spp)
    target = t[..., :3] / t[..., 3:4]
    alb0, tex0 = scenes.perturbed_params(sc)
    sc.tri_albedo, sc.textures = alb0, [tex0]
    be.prepare(sc)
    be.set_eval_params()

    def best(fn, reps=3):
        ms = []
        for _ in range(reps):
            fn()
            ms.append(be.stats()["render_ms"])
        return min(ms)

    def fwd():
        be.clear_accum()
        be.render_frame(0, sc.spp)

    fwd()
    fwd_ms = best(fwd)
    img = be.replay(0, sc.spp)
    I = img[..., :3] / img[..., 3:4]
    rep_ms = best(lambda: be.replay(0, sc.spp))
    dldi = (2.0 * (I - target)).astype(np.float32)
    adj_ms = best(lambda: be.adjoint(0, sc.spp, dldi, tex0.shape))
    da, dt = be.adjoint(0, sc.spp, dldi, tex0.shape)
    log(f"[bench] adjoint: forward {fwd_ms:.2f} ms, replay {rep_ms:.2f} ms, adjoint {adj_ms:.2f} ms")
    return {"workload": f"inverse_scene {sc.num_triangles} tris, 512x512, {sc.spp} spp, depth {sc.max_depth}, "
                        f"1024^2 roughness texture",
            "forward_ms": round(fwd_ms, 3), "replay_ms": round(rep_ms, 3), "adjoint_ms": round(adj_ms, 3),
            "adjoint_over_forward": round(adj_ms / fwd_ms, 3),
            "nonzero_albedo_grads": int((da != 0).sum()), "nonzero_texel_grads": int((dt != 0).sum()),
            "timing": "HIP events around the kernels of one call (best of 3)"}


def other_configs(be, log):
    """The other workloads on the same backend, one warm-up and two timed frames each (HIP events
    around the frame):
      cornell: the Cornell box (36 triangles, 512x512, 64 spp, depth 8);
      ggx_tilted_ground: the 1080p x 16 spp GGX workload with its ground turned by 0.05 degrees,
        so the ground is hit (flat, it sits alone in a box of no height that the reference's slab
        test never enters, and the timed workload is the spheres over the sky);
      hair: the hair sphere (200k Bezier strands = 800k segments, 1024x1024, 32 spp, depth 16) as the
        reference renders it: Kajiya-Kay hair, ribbon depth 6 (the parity mode);
      hair_config3: the same frame as the hair workload specifies it: Marschner R/TT/TRT hair
        (beta_m = beta_n = 0.3, alpha = 2 degrees, IOR 1.55, eumelanin sigma_a), ribbon depth 5."""
    from anacapa_amd import _abi as A
    from anacapa_amd import scenes
    out = {}
    be.set_launch(max_paths=DEFAULT_MAX_PATHS, tile=8)   # the library default, whatever ran before
    hair = scenes.hair_sphere()
    runs = (("cornell", scenes.cornell(), 0, 0, ""),
            ("ggx_tilted_ground", scenes.ggx_spheres(ground_tilt_deg=0.05), 0, 0, ", ground tilted 0.05 deg"),
            ("hair", hair, 0, 0, ", Kajiya-Kay, ribbon depth 6"),
            ("hair_config3", hair, A.AHIP_FLAG_HAIR_MARSCHNER, 5, ", Marschner R/TT/TRT, ribbon depth 5"))
    for name, sc, flags, cdepth, note in runs:
        be.prepare(sc, flags=flags, curve_depth=cdepth)
        be.clear_accum()
        be.render_frame(0, sc.spp)
        ms = []
        for _ in range(2):
            be.reset_stats()
            be.clear_accum()
            be.render_frame(0, sc.spp)
            ms.append(be.stats()["render_ms"])
        st = be.stats()
        t = min(ms)
        samples = sc.width * sc.height * sc.spp
        rays = st["rays_closest"] + st["rays_shadow"]
        out[name] = {"workload": f"{sc.name}: {sc.num_triangles} tris, {sc.num_strands} strands, "
                                 f"{sc.width}x{sc.height}, {sc.spp} spp, depth {sc.max_depth}{note}",
                     "ms_per_frame": round(t, 3), "msamples_per_s": round(samples / t / 1e3, 2),
                     "mrays_per_s": round(rays / t / 1e3, 2)}
        log(f"[bench] {name}: {t:.1f} ms/frame, {samples / t / 1e3:.1f} Msamples/s")
    return out


def bdpt_form(be, args, log):
    """The bidirectional path tracer on the Cornell box at the configuration of the reference's published
    BDPT timing (400x400, 64 spp, depth 8: ~1168 ms on an Apple M3 Pro with 11 threads, README.md:444-448),
    best of two timed frames after a warm-up, and the reference's own BDPTIntegrator (oracle/_ref) on
    the host threads through RenderSession's tile loop for the same frame, when that build is present."""
    from anacapa_amd import _abi as A
    from anacapa_amd import scenes
    sc = scenes.cornell(width=400, height=400, spp=64)
    be.prepare(sc, integrator=A.AHIP_INTEGRATOR_BDPT, firefly_clamp=10.0)
    be.clear_accum()
    be.render_frame(0, sc.spp)
    ms = []
    for _ in range(2):
        be.reset_stats()
        be.clear_accum()
        be.render_frame(0, sc.spp)
        ms.append(be.stats()["render_ms"])
    t = min(ms)
    samples = sc.width * sc.height * sc.spp
    out = {"workload": "cornell: 36 tris, 400x400, 64 spp, depth 8, firefly clamp 10 (BDPTIntegrator)",
           "ms_per_frame": round(t, 3), "msamples_per_s": round(samples / t / 1e3, 2),
           "published_cpu": {"ms_per_frame": 1168, "host": "Apple M3 Pro, 11 threads", "source": "README.md:444-448"}}
    log(f"[bench] bdpt: {t:.1f} ms/frame, {samples / t / 1e3:.1f} Msamples/s")
    # the same integrator on the 1M-triangle workload frame under its sky (one warm-up, one timed)
    wl = scenes.ggx_spheres()
    be.prepare(wl, integrator=A.AHIP_INTEGRATOR_BDPT, firefly_clamp=10.0)
    be.clear_accum()
    be.render_frame(0, wl.spp)
    be.reset_stats()
    be.clear_accum()
    be.render_frame(0, wl.spp)
    tw = be.stats()["render_ms"]
    ws = wl.width * wl.height * wl.spp
    out["workload_1m"] = {"workload": f"ggx_spheres: {wl.num_triangles} tris, {wl.width}x{wl.height}, {wl.spp} spp, "
                                     f"depth {wl.max_depth}, the 2048x1024 sky (dome) light, BDPTIntegrator",
                          "ms_per_frame": round(tw, 2), "msamples_per_s": round(ws / tw / 1e3, 2)}
    log(f"[bench] bdpt workload: {tw:.1f} ms/frame, {ws / tw / 1e3:.1f} Msamples/s")
    be.set_params(sc.max_depth, sc.min_depth, sc.seed)   # back to the path tracer
    if args.no_cpu_baseline:
        return out
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        from oracle_bindings import Reference, reference_available
        if not reference_available():
            return out
        rs = Reference().scene(sc)
    except Exception as e:  # noqa: BLE001
        log(f"[bench] bdpt cpu reference unavailable: {e}")
        return out
    threads = min(host_info()["hardware_concurrency"], 16)
    cspp = 16   # a quarter of the frame's samples keeps the CPU leg to seconds
    _, sec, used = rs.render_frame_bdpt(cspp, 10.0, threads=threads)
    csamples = sc.width * sc.height * cspp
    out["cpu_reference"] = {"msamples_per_s": round(csamples / sec / 1e6, 3),
                            "ms_per_frame": round(sec * 1e3 * sc.spp / cspp, 1), "threads": used, "kind": "reference",
                            "sample": f"the same frame at {cspp} spp ({csamples} samples, {sec:.2f}s), ms_per_frame "
                                      f"scaled to {sc.spp} spp",
                            "build": "oracle/_ref: BDPTIntegrator.cpp + MISWeight.cpp at the reference's Release flags"}
    log(f"[bench] bdpt cpu reference: {csamples / sec / 1e6:.2f} Msamples/s on {used} threads")
    return out


def host_info():
    """CPU model and hardware thread count of the host the baseline runs on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "hardware_concurrency": os.cpu_count() or 1}


def cpu_baseline(sc, args):
    """The CPU baseline on the same box's host cores: the reference's own code on the same workload.

    value (kind "reference"): RenderSession's non-adaptive tile loop (RenderSession.cpp:511-547:
    64x64 tiles on the reference ThreadPool, HaltonSampler clones, TileBuffers) over the reference's
    BVHBackend, PathIntegrator and ShadowRay.h, compiled where they lie at the reference's Release
    flags (oracle/_ref), on the same 1080p frame at full spp -- the whole frame when it fits
    --cpu-seconds, else evenly spaced 64-row tile rows.  The reference's StandardSurfaceMaterial and
    DomeLight reach OpenImageIO-only code (evalTOV, DomeLight::loadImage), so the materials are a
    test class restating StandardSurface's methods for constant parameters and the dome forwards to
    the reference DomeLight's own header methods (oracle/ref_driver.cpp); the timing window is the
    tile loop only (the BVH build excluded, RenderSession.cpp:479, 665-667).
    port: this repository's restatement (oracle/liboracle_release.so, the same Release flags) on the
    same rows and threads, for comparison.  reference_check: both on the Cornell box, which the
    reference renders with its own Lambert / area-light classes."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        from oracle_bindings import Oracle, Reference, reference_available
        orc = Oracle(release=True)
    except Exception as e:  # noqa: BLE001
        log(f"[bench] cpu baseline unavailable: {e}")
        return None
    host = host_info()
    # the GPU box grants a 16-thread CPU share per GPU whatever the host reports
    threads = min(host["hardware_concurrency"], 16)
    rows = list(range(0, args.height, 64))                     # the frame's tile rows
    out = {"unit": "Msamples/s", "cores": threads, "host": host}
    if reference_available():
        rs = Reference().scene(sc)
        _, sec1, used = rs.render_frame(args.spp, threads=threads, row0=rows[len(rows) // 2],
                                        row1=rows[len(rows) // 2] + 64)
        n = int(max(1, min(len(rows), args.cpu_seconds / max(sec1, 1e-3))))
        # the centre tile row of each of n equal strata of the frame (the sky rows are cheaper)
        pick = sorted({rows[min(len(rows) - 1, int((i + 0.5) * len(rows) / n))] for i in range(n)})
        sec, nrows = 0.0, 0
        for y in pick:
            _, s1, used = rs.render_frame(args.spp, threads=threads, row0=y, row1=y + 64)
            sec += s1
            nrows += min(64, args.height - y)
        samples = args.width * nrows * args.spp
        what = ("the whole frame" if len(pick) == len(rows) else
                f"{len(pick)} evenly spaced 64-row tile rows of the frame")
        log(f"[bench] cpu baseline (reference code): {what}, {samples} samples in {sec:.1f}s on {used} threads "
            f"({host['cpu_model']}, {host['hardware_concurrency']} hardware threads)")
        out.update({"value": round(samples / sec / 1e6, 4), "kind": "reference",
                    "sample": f"{what} ({args.width}x{args.height} at {args.spp} spp: {samples} samples, {sec:.1f}s)",
                    "build": "oracle/_ref: BVHBackend.cpp + PathIntegrator.cpp at the reference's Release flags, "
                             "RenderSession's tile loop on its ThreadPool; StandardSurface and DomeLight through "
                             "test classes (their OpenImageIO-only code is not buildable here)"})
        port_rows = pick
    else:
        port_rows = [rows[min(len(rows) - 1, int((i + 0.5) * len(rows) / 4))] for i in range(4)]
    os_ = orc.scene(sc)
    psec, prow = 0.0, 0
    for y in port_rows:
        _, s1 = os_.render_frame(args.spp, y, min(y + 64, args.height), threads=threads)
        psec += s1
        prow += min(64, args.height - y)
    psamples = args.width * prow * args.spp
    port = {"value": round(psamples / psec / 1e6, 4), "kind": "port", "cores": threads,
            "sample": f"{len(port_rows)} 64-row tile rows ({psamples} samples, {psec:.1f}s)",
            "build": "oracle/liboracle_release.so (the reference's Release flags)"}
    log(f"[bench] cpu port: {psamples} samples in {psec:.1f}s on {threads} threads")
    if "value" not in out:
        out.update(port)
    else:
        out["port"] = port
    out["reference_check"] = reference_check(orc, threads)
    return out


def reference_check(orc, threads):
    """The Cornell box (36 triangles, 512x512, depth 8) at 16 spp — a quarter of its 64-spp
    workload, to keep the check to seconds — rendered by the reference's own code (oracle/_ref,
    RenderSession.cpp:511-547: 64x64 tiles on its ThreadPool) and by the port, both timed around
    the tile loop only, median of 3."""
    from anacapa_amd import scenes
    from oracle_bindings import Reference, reference_available
    if not reference_available():
        return None
    sc = scenes.cornell()
    rs = Reference().scene(sc)
    os_ = orc.scene(sc)
    spp = 16
    ref_s = sorted(rs.render_frame(spp, threads=threads)[1] for _ in range(3))[1]
    port_s = sorted(os_.render_frame(spp, threads=threads)[1] for _ in range(3))[1]
    samples = sc.width * sc.height * spp
    log(f"[bench] cpu reference check (cornell): reference {ref_s:.2f}s, port {port_s:.2f}s on {threads} threads")
    return {"workload": "cornell 36 tris, 512x512, 16 spp, depth 8", "threads": threads,
            "reference_msps": round(samples / ref_s / 1e6, 3), "port_msps": round(samples / port_s / 1e6, 3),
            "port_over_reference": round(ref_s / port_s, 3)}


if __name__ == "__main__":
    main()
