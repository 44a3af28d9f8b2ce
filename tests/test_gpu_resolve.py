"""Device film resolve (ahip_resolve) against the reference formulas, and the PNG / EXR written
from it.

rgb and variance are restated in float32 numpy from the downloaded accumulation (PixelAccumulator::
resolve, Film.h:46-55; Film::varianceAt, Film.cpp:80-90; each product and sum rounded) and must
match bit for bit.  The 8-bit sRGB pixels follow Film::writePNG (Film.cpp:233-256); powf on the
device and in numpy may differ by an ulp, which can move a pixel by one code value, so at most 1
and on under 0.5% of channels.  The AOV means match the tile AOVs of the same samples.
"""
import numpy as np
import pytest

from anacapa_amd import scenes
from anacapa_amd.integrator import write_exr, write_png

pytestmark = pytest.mark.gpu


def test_resolve_matches_film_formulas(gpu, tmp_path):
    f = np.float32
    sc = scenes.cornell(width=64, height=48)
    gpu.prepare(sc)
    gpu.clear_accum()
    gpu.set_aovs(True)
    try:
        gpu.render_frame(0, 8)
        acc = gpu.download_accum()
        res = gpu.resolve(exposure=0.5, variance=True, aovs=True)
    finally:
        gpu.set_aovs(False)
    r, g, b, w, lsq = (acc[..., k] for k in range(5))
    inv = np.where(w > 0, f(1) / np.where(w > 0, w, f(1)), f(0)).astype(f)
    rgb = np.stack([r * inv, g * inv, b * inv], -1).astype(f)
    assert np.array_equal(res["rgb"].view(np.uint32), rgb.view(np.uint32))
    mean = ((r * f(0.2126) + g * f(0.7152)).astype(f) + b * f(0.0722)).astype(f) * inv
    var = np.where(w > 0, np.maximum(f(0), (lsq * inv).astype(f) - (mean * mean).astype(f)), f(0)).astype(f)
    assert np.array_equal(res["variance"].view(np.uint32), var.view(np.uint32))
    x = np.clip(rgb * f(np.float32(2.0) ** f(0.5)), f(0), f(1)).astype(f)
    enc = np.where(x <= f(0.0031308), f(12.92) * x, (f(1.055) * np.power(x, f(1) / f(2.4))).astype(f) - f(0.055))
    want = (enc.astype(f) * f(255) + f(0.5)).astype(np.uint8)
    d = np.abs(res["srgb8"].astype(int) - want.astype(int))
    assert d.max() <= 1 and (d > 0).mean() < 0.005, (d.max(), (d > 0).mean())
    # AOV means: the tile AOVs of the same samples
    _, alb, nrm = gpu.render_tile_aovs(0, 0, 64, 48, 0, 8)
    np.testing.assert_allclose(res["albedo"], alb, rtol=0, atol=1e-6)
    np.testing.assert_allclose(res["normal"], nrm, rtol=0, atol=1e-6)
    # the written files hold the resolved buffers
    from PIL import Image
    write_png(tmp_path / "f.png", res["srgb8"])
    with Image.open(tmp_path / "f.png") as im:
        assert np.array_equal(np.asarray(im), res["srgb8"])
    write_exr(tmp_path / "f.exr", {"": res["rgb"], "albedo.": res["albedo"], "normals.": res["normal"]})
    assert (tmp_path / "f.exr").stat().st_size > 64 * 48 * 9 * 4


def test_resolve_refuses_aovs_started_late(gpu):
    """AOV sums begun after radiance samples were accumulated cover fewer samples than the weight:
    the AOV means are refused until the next clear."""
    from anacapa_amd.integrator import BackendError
    sc = scenes.cornell(width=32, height=32)
    gpu.prepare(sc)
    gpu.clear_accum()
    gpu.render_frame(0, 2)
    gpu.set_aovs(True)
    try:
        gpu.render_frame(2, 2)
        with pytest.raises(BackendError):
            gpu.resolve(aovs=True)
        assert "rgb" in gpu.resolve(aovs=False)
        gpu.clear_accum()
        gpu.render_frame(0, 2)
        out = gpu.resolve(aovs=True)
        assert out["albedo"].max() > 0
    finally:
        gpu.set_aovs(False)
