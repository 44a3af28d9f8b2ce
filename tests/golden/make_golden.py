"""Write the golden fixtures from the reference's own code (oracle/_ref/libanacapa_ref.so).

The reference build compiles the reference's BVHBackend.cpp, PathIntegrator.cpp and the
header-only PCGRng / HaltonSampler / Camera / Lambertian / AreaLight / Distribution1D / GGX
helpers where they lie in the reference checkout (oracle/Makefile), so every value here is
what the reference computes.  Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from anacapa_amd import scenes  # noqa: E402
from oracle_bindings import Reference  # noqa: E402
from rays import random_rays, camera_rays  # noqa: E402


def main():
    R = Reference()
    out = {}
    # PCG32: published pcg32 demo seed (42, 54) plus a few (seed, seq=1) streams
    out["pcg_42_54"] = R.pcg(42, 54, 32)
    seeds = np.array([0, 1, 12345, 0xdeadbeefcafe, 2**63 + 7], np.uint64)
    out["pcg_seeds"] = seeds
    out["pcg_streams"] = np.stack([R.pcg(int(s), 1, 16) for s in seeds])
    # Halton: 256 random (pixel, index) triples x 176 dimensions (past the 128-dim PCG fallback)
    rng = np.random.default_rng(2024)
    trip = np.stack([rng.integers(0, 4096, 256), rng.integers(0, 4096, 256), rng.integers(0, 1 << 20, 256)], 1)
    trip[:8, 2] = [0, 1, 2, 3, 255, 256, 65535, (1 << 32) - 1]
    out["halton_trip"] = trip.astype(np.uint32)
    out["halton_vals"] = np.stack([R.halton_seq(0, int(a), int(b), int(c), 176) for a, b, c in trip])
    out["halton_vals_seed7"] = np.stack([R.halton_seq(7, int(a), int(b), int(c), 40) for a, b, c in trip[:32]])
    out["halton_vals_2d"] = np.stack([R.halton_seq2d(0, int(a), int(b), int(c), 20) for a, b, c in trip[:32]])
    # camera
    sc = scenes.cornell()
    cam = R.pinhole(sc.cam_from, sc.cam_at, sc.cam_up, sc.vfov, sc.width, sc.height)
    out["cornell_camera"] = np.frombuffer(bytes(cam), np.uint8).copy()
    g = scenes.ggx_spheres.__defaults__  # noqa: F841 (documentation of the defaults used below)
    px = rng.integers(0, 512, 512).astype(np.uint32)
    py = rng.integers(0, 512, 512).astype(np.uint32)
    jx = rng.uniform(0, 1, 512).astype(np.float32)
    jy = rng.uniform(0, 1, 512).astype(np.float32)
    out["camray_in"] = np.stack([px, py, jx.view(np.uint32), jy.view(np.uint32)], 1)
    out["camray_out"] = R.generate_rays(cam, px, py, jx, jy)
    # Cornell: BVH, hits, radiance
    rs = R.scene(sc)
    nodes, prims = rs.bvh()
    out["cornell_nodes"], out["cornell_prims"] = nodes, prims
    rays = np.concatenate([random_rays(sc, 3000, seed=5), camera_rays(cam, 1096, seed=6)])
    out["cornell_rays"] = rays
    h = rs.trace(rays)
    out["cornell_hit_t"], out["cornell_hit_mesh"], out["cornell_hit_prim"] = h["t"], h["mesh_id"], h["prim_id"]
    out["cornell_occ"] = rs.occluded(rays)
    out["cornell_tile_200_220_48x40_s0_n8"] = rs.render_tile(200, 220, 48, 40, 0, 8)
    out["cornell_tile_0_0_32x32_s3_n5_d3"] = rs.render_tile(0, 0, 32, 32, 3, 5, max_depth=3)
    # denoising AOVs of the same tiles (TileBuffer albedo / normals after one renderTile)
    for key, args, kw in (("cornell_tile_200_220_48x40_s0_n8", (200, 220, 48, 40, 0, 8), {}),
                          ("cornell_tile_0_0_32x32_s3_n5_d3", (0, 0, 32, 32, 3, 5), {"max_depth": 3})):
        acc, alb, nrm = rs.render_tile_aovs(*args, **kw)
        assert np.array_equal(acc, out[key])
        out[key + "_albedo"], out[key + "_normal"] = alb, nrm
    # triangle soup: topology and hits under heavy overlap
    soup = scenes.triangle_soup()
    out["soup_pos_sum"] = np.float64(soup.positions.astype(np.float64).sum())
    ss = R.scene(soup)
    n2, p2 = ss.bvh()
    out["soup_nodes"], out["soup_prims"] = n2, p2
    r2 = random_rays(soup, 4000, seed=9)
    out["soup_rays"] = r2
    h2 = ss.trace(r2)
    out["soup_hit_t"], out["soup_hit_mesh"], out["soup_hit_prim"] = h2["t"], h2["mesh_id"], h2["prim_id"]
    out["soup_occ"] = ss.occluded(r2)
    # free GGX helpers and Distribution1D
    a = rng.uniform(-0.2, 1.2, 2000).astype(np.float32)
    b = rng.uniform(1e-5, 1.0, 2000).astype(np.float32)
    c = rng.uniform(-0.2, 1.0, 2000).astype(np.float32)
    ior = rng.uniform(1.0, 2.5, 2000).astype(np.float32)
    out["ggx_in"] = np.stack([a, b, c, ior], 1)
    out["ggx_D"] = R.ggx_helper(0, a, b)
    out["ggx_G1"] = R.ggx_helper(1, a, b)
    out["ggx_schlick"] = R.ggx_helper(2, a, b)
    out["ggx_fresnel"] = R.ggx_helper(3, np.abs(a), ior)
    out["ggx_pdf"] = R.ggx_helper(4, a, b, c)
    out["ggx_f0"] = R.ggx_helper(5, ior)
    u = rng.uniform(0, 1, 2000).astype(np.float32)
    out["ggx_half"] = R.ggx_helper(6, u, np.abs(c).astype(np.float32), b)
    vals = np.abs(rng.normal(size=333)).astype(np.float32)
    vals[10:20] = 0.0
    us = np.concatenate([rng.uniform(0, 1, 3000), [0.0, 1.0, 0.5]]).astype(np.float32)
    idx, pdf, ur, cdf, integ = R.distribution1d(vals, us)
    out["d1_vals"], out["d1_us"], out["d1_idx"], out["d1_pdf"], out["d1_ur"], out["d1_cdf"] = vals, us, idx, pdf, ur, cdf
    out["d1_int"] = np.float32(integ)
    np.savez_compressed(os.path.join(HERE, "reference_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "reference_vectors.npz"), sum(v.nbytes for v in out.values()), "bytes raw")


if __name__ == "__main__":
    main()
