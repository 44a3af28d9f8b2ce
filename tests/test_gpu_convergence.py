"""Converged images: the HIP path's image over an independent range of sample indices agrees with
the oracle's 4096-spp render of the same pixels within the Monte Carlo 3-sigma band.

Both sides accumulate, per pixel, the radiance sum and the sum of squared luminance
(PathIntegrator.cpp:36-53), so each gives a per-pixel luminance mean and variance; the GPU uses
sample indices [4096, 8192), the CPU [0, 4096), so the two estimates share no samples.  Halton
points stratify better than independent samples, which makes the band conservative.
"""
import numpy as np
import pytest

from anacapa_amd import scenes

pytestmark = pytest.mark.gpu

N = 4096


def _lum(rgb):   # luminance, ACEScg weights (include/anacapa/core/Types.h:245-248)
    return 0.2722287 * rgb[..., 0] + 0.6740818 * rgb[..., 1] + 0.0536895 * rgb[..., 2]


def _stats(acc):
    a = acc.astype(np.float64)
    w = a[..., 3]
    mean = _lum(a[..., :3] / w[..., None])
    var = np.maximum(a[..., 4] / w - mean * mean, 0.0)
    return mean, var, w


@pytest.mark.parametrize("name", ["cornell", "small_ggx"])
def test_converged_image_within_3_sigma_of_cpu_4096spp(gpu, oracle, name):
    if name == "cornell":
        sc, (x0, y0, w, h) = scenes.cornell(width=64, height=64), (24, 24, 16, 16)
    else:
        sc, (x0, y0, w, h) = scenes.small_ggx(), (40, 24, 16, 16)
    gpu.prepare(sc)
    g = gpu.render_tile(x0, y0, w, h, N, N)
    c = oracle.scene(sc).render_tile(x0, y0, w, h, 0, N)
    mg, vg, wg = _stats(g)
    mc, vc, wc = _stats(c)
    assert (wg == N).all() and (wc == N).all()
    sigma = np.sqrt(vg / N + vc / N) + 1e-7 * np.maximum(mc, 1e-3)
    z = np.abs(mg - mc) / sigma
    assert (z <= 3.0).mean() >= 0.99, f"{(z > 3).sum()} of {z.size} pixels outside 3 sigma (max {z.max():.2f})"
    assert z.max() < 6.0
    # the tile mean, whose sigma is the per-pixel sigmas combined
    zt = abs(mg.mean() - mc.mean()) / (np.sqrt((vg / N + vc / N).sum()) / z.size + 1e-9)
    assert zt <= 3.0, zt
