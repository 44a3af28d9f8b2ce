"""Adaptive two-pass render on the device (ahip_render_adaptive) against the CPU tile path's rules.

RenderSession's adaptive mode (RenderSession.cpp:549-653) renders a base pass, scores each tile by
the mean of Film::varianceAt (Film.cpp:80-90) and hands out the remaining samples half uniformly,
half by log(1 + score).  Checked here:
  - the tile scores equal the reference formula restated in float32 numpy on the downloaded base
    pass (each product and sum rounded, row-major order), bit for bit;
  - the allocation equals the reference's arithmetic on those scores;
  - the final accumulation equals, per tile, the oracle's renders of [0, base) and
    [base, base + extra) within the radiance bar, with exact per-pixel sample counts;
  - with AOVs on, the AOV sums cover every sample of every pixel.
"""
import math

import numpy as np
import pytest

from anacapa_amd import scenes

pytestmark = pytest.mark.gpu


def _scores(acc, tile):
    """Film::varianceAt per pixel, summed per tile in row-major order, / pixel count (float32)."""
    f = np.float32
    H, W = acc.shape[:2]
    r, g, b, w, lsq = (acc[..., k].astype(f) for k in range(5))
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.where(w > 0, f(1) / w, f(0)).astype(f)
        mean = ((r * f(0.2126) + g * f(0.7152)).astype(f) + b * f(0.0722)).astype(f) * inv
        var = np.where(w > 0, np.maximum(f(0), (lsq * inv).astype(f) - (mean * mean).astype(f)), f(0)).astype(f)
    ty, tx = -(-H // tile), -(-W // tile)
    out = np.zeros((ty, tx), f)
    for j in range(ty):
        for i in range(tx):
            blk = var[j * tile:(j + 1) * tile, i * tile:(i + 1) * tile]
            s = f(0)
            for v in blk.reshape(-1):
                s = f(s + v)
            out[j, i] = f(s / f(blk.size))
    return out


def _allocation(scores, extra):
    """RenderSession.cpp:606-640 on the scores; also returns where the rounded value sits within
    1e-3 of a half (there an ulp of logf may move it either way)."""
    f = np.float32
    logw = [f(np.log(np.float64(f(1) + v))) for v in scores.reshape(-1)]
    total = f(0)
    for v in logw:
        total = f(total + v)
    n = len(logw)
    uniform, bonus_half = extra // 2, extra - extra // 2
    out, near = [], []
    for v in logw:
        x = float(f(f(v / total) * f(bonus_half * n))) if total > f(1e-6) else 0.0
        out.append(min(uniform + int(round(x)), extra))
        near.append(abs(x - math.floor(x) - 0.5) < 1e-3)
    return np.array(out, np.uint32).reshape(scores.shape), np.array(near).reshape(scores.shape)


def test_adaptive_scores_allocation_and_image(gpu, oracle):
    sc = scenes.cornell(width=96, height=80)
    tile, total = 32, 24
    gpu.prepare(sc)
    gpu.clear_accum()
    gpu.render_frame(0, 16)                                  # the base pass alone, for the restatement
    base_acc = gpu.download_accum()
    gpu.clear_accum()
    extra, var = gpu.render_adaptive(total, tile=tile)       # base 16 = min(32, max(16, 24 // 16))
    print(f"[adaptive] extra per tile {extra.reshape(-1).tolist()}", flush=True)
    assert np.array_equal(var.view(np.uint32), _scores(base_acc, tile).view(np.uint32))
    want, near = _allocation(var, total - 16)
    assert np.array_equal(extra[~near], want[~near])
    assert np.all(np.abs(extra.astype(np.int64) - want)[near] <= 1)
    assert extra.min() >= (total - 16) // 2 and extra.max() <= total - 16 and extra.max() > extra.min()
    acc = gpu.download_accum()
    os_ = oracle.scene(sc)
    ty, tx = extra.shape
    for j in range(ty):
        for i in range(tx):
            x0, y0 = i * tile, j * tile
            w, h = min(tile, 96 - x0), min(tile, 80 - y0)
            ref = os_.render_tile(x0, y0, w, h, 0, 16) + os_.render_tile(x0, y0, w, h, 16, int(extra[j, i]))
            got = acc[y0:y0 + h, x0:x0 + w]
            assert np.all(got[..., 3] == 16 + extra[j, i])
            g, r = got[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
            err = np.abs(g - r).max(-1) / np.maximum(np.abs(r).max(-1), 1e-3 * max(r.mean(), 1e-6))
            print(f"[radiance] adaptive tile {j},{i}: {(err > 1e-3).sum()} of {err.size} pixels beyond 1e-3", flush=True)
            assert (err <= 1e-3).all(), (j, i, (err > 1e-3).sum())     # measured: 0 beyond 1e-3
            assert abs(g.mean() / r.mean() - 1) < 1e-3


def test_adaptive_with_aovs_and_explicit_base(gpu):
    sc = scenes.cornell(width=64, height=48)
    gpu.prepare(sc)
    gpu.clear_accum()
    gpu.set_aovs(True)
    try:
        extra, _ = gpu.render_adaptive(12, tile=16, base_spp=4)
        acc = gpu.download_accum()
        aov = gpu.download_aovs()
    finally:
        gpu.set_aovs(False)
    per_pixel = np.kron(extra, np.ones((16, 16), np.uint32))[:48, :64] + 4
    assert np.array_equal(acc[..., 3], per_pixel.astype(np.float32))
    # every sample's first-hit normal is unit length or zero (miss): the normal sums never exceed
    # the sample count and reach it inside flat walls; reflectance sums stay under count * 0.73
    nlen = np.linalg.norm(aov[..., 3:], axis=-1)
    assert np.all(nlen <= per_pixel * (1 + 1e-5))
    assert (np.abs(nlen / per_pixel - 1) < 1e-4).mean() > 0.5
    assert np.all(aov[..., :3] <= per_pixel[..., None] * 0.73 * (1 + 1e-5))
    # no extra samples: the base pass alone
    gpu.clear_accum()
    extra, var = gpu.render_adaptive(4, tile=16, base_spp=4)
    assert not extra.any() and not var.any()
    assert np.all(gpu.download_accum()[..., 3] == 4)


def test_adaptive_and_resolve_argument_errors(gpu):
    from anacapa_amd.integrator import BackendError
    from anacapa_amd import _abi as A
    sc = scenes.cornell(width=32, height=32)
    gpu.prepare(sc)
    with pytest.raises(BackendError) as e:
        gpu.render_adaptive(8, tile=0)
    assert e.value.status == A.AHIP_E_INVALID_ARG
    with pytest.raises(BackendError) as e:
        gpu.render_adaptive(0, tile=16)
    assert e.value.status == A.AHIP_E_INVALID_ARG
    gpu.set_aovs(False)
    gpu.clear_accum()
    gpu.render_frame(0, 1)
    with pytest.raises(BackendError) as e:                      # AOV means need ahip_set_aovs first
        gpu.resolve(aovs=True)
    assert e.value.status == A.AHIP_E_INVALID_ARG
