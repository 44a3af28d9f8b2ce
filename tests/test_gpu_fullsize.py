"""The benchmark workload at its full size: 1,003,522 triangles, 1920x1080, 16 spp, depth 8.

Checked directly against the oracle where the oracle is fast at this size (the SAH-12 tree of the
1M-triangle scene and 200k traced rays take it a few seconds), and through properties that do not
depend on size for the whole frame: a frame equals the four quadrants rendered as tiles, bit for
bit (each pixel sums its own samples in sample order whatever the region); splitting the sample
range adds up to the one-range frame (sums to float rounding, weights exactly); every sum finite.
"""
import numpy as np
import pytest

from anacapa_amd import scenes
from rays import camera_rays, random_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big():
    sc = scenes.ggx_spheres()
    print(f"[fullsize] {sc.num_triangles} triangles, {sc.width}x{sc.height} x {sc.spp} spp", flush=True)
    return sc


def test_full_scene_tree_and_hits_vs_oracle(gpu, oracle, big):
    gpu.prepare(big)
    os_ = oracle.scene(big)
    n_g, p_g = gpu.bvh_dump()
    n_o, p_o = os_.bvh()
    assert len(n_g) > 600_000
    assert np.array_equal(n_g, n_o) and np.array_equal(p_g, p_o)
    rays = np.concatenate([random_rays(big, 150_000, seed=31), camera_rays(os_.cam, 50_000, seed=32)])
    print(f"[fullsize] tracing {len(rays)} rays", flush=True)
    g, o = gpu.trace_rays(rays), os_.trace(rays)
    for k in ("mesh_id", "prim_id"):
        assert np.array_equal(g[k], o[k]), k
    hit = g["mesh_id"] != 0xFFFFFFFF
    assert hit.mean() > 0.3
    for k in ("t", "u", "v"):
        assert np.array_equal(g[k][hit].view(np.uint32), o[k][hit].view(np.uint32)), k
    assert np.array_equal(gpu.occluded_rays(rays), os_.occluded(rays))


def test_full_frame_equals_quadrant_tiles(gpu, big):
    gpu.prepare(big)
    gpu.set_launch(1 << 25, 0)      # the library default: the whole frame in one pass
    gpu.clear_accum()
    gpu.render_frame(0, 16)
    frame = gpu.download_accum()
    assert np.isfinite(frame).all() and np.all(frame[..., 3] == 16)
    W, H = big.width, big.height
    for x0, y0 in ((0, 0), (W // 2, 0), (0, H // 2), (W // 2, H // 2)):
        tile = gpu.render_tile(x0, y0, W // 2, H // 2, 0, 16)
        assert np.array_equal(tile, frame[y0:y0 + H // 2, x0:x0 + W // 2]), (x0, y0)
    # the sample range split in two adds up to the same sums
    gpu.clear_accum()
    gpu.render_frame(0, 7)
    gpu.render_frame(7, 9)
    split = gpu.download_accum()
    assert np.array_equal(split[..., 3], frame[..., 3])
    np.testing.assert_allclose(split[..., :3], frame[..., :3], rtol=2e-5, atol=1e-4)
    np.testing.assert_allclose(split[..., 4], frame[..., 4], rtol=2e-5, atol=1e-3)
