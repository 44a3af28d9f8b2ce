"""The device BDPT (bdpt.hip) against the oracle's restatement of BDPTIntegrator (itself held to the
reference's own build by tests/test_oracle_bdpt.py), on the Cornell box with one and with two lights,
and with StandardSurface blocks and a smooth mirror.

Per sample (each sample index rendered on its own on both sides): every sample within 2e-4
relative except samples the oracle flags as hanging on the last bits of a float: a connection whose
visibility changes within an ulp of the hit distance (Scene::connVisible), or a subpath through a
smooth StandardSurface, whose throughput is NaN or finite as the rebuilt half-vector's z rounds to 1
or not (Scene::smooth).  Measured on the MI355X (Lambert): 20 of 16384 samples beyond 1e-2, all
flagged, the largest error of an unflagged sample 1.3e-4.  Sample weights exact, AOVs as the
reference records them (first-vertex albedo within 1e-6, normals within 1e-6)."""
import numpy as np
import pytest

from anacapa_amd import _abi as A
from anacapa_amd import scenes

pytestmark = pytest.mark.gpu
NS = 4
BDPT = A.AHIP_INTEGRATOR_BDPT


@pytest.mark.parametrize("two_lights,glossy", [(False, False), (True, False), (True, True), (False, "glass"),
                                               (False, "glass_texture")])
def test_bdpt_tile_vs_oracle(gpu, oracle, two_lights, glossy):
    """glossy: standard_surface gold metal and coated dielectric blocks and a mirror panel (a delta
    BSDF, whose connections the reference skips; measured: 64 samples beyond 2e-4, all flagged).
    "glass": a smooth and a rough tinted glass block and two panels of partial opacity, whose
    connections go through k_bdpt_conn_trace's stepping instance (measured: 33 samples beyond 2e-4,
    30 flagged, the other three within 7.3e-4).  "glass_texture": the first panel's opacity from a
    texture, read at the camera subpath's hits and at the connections' stepping hits (no reference
    pin: the reference's side takes constant opacity only; measured: 33 beyond 2e-4, 30 flagged, the
    others within 8.6e-4)."""
    if glossy in ("glass", "glass_texture"):
        sc = scenes.cornell_glass(width=64, height=64, alpha_texture=glossy == "glass_texture")
    else:
        sc = scenes.cornell(width=64, height=64, second_light=two_lights, glossy=glossy)
    os_ = oracle.scene(sc)
    gpu.prepare(sc, integrator=BDPT)
    print(f"[bdpt] {sc.name} 64x64 x {NS} spp, two lights {two_lights}, glossy {glossy}", flush=True)
    got, alb, nrm = gpu.render_tile_aovs(0, 0, 64, 64, 0, NS)
    ref, oalb, onrm, bd = os_.render_tile_bdpt(0, 0, 64, 64, 0, NS, border=True)
    assert np.array_equal(got[..., 3], ref[..., 3])
    assert abs(got[..., :3].astype(np.float64).mean() / ref[..., :3].astype(np.float64).mean() - 1) < 2e-4
    np.testing.assert_allclose(alb, oalb, rtol=1e-6, atol=0)
    np.testing.assert_allclose(nrm, onrm, rtol=0, atol=1e-6)
    g = np.stack([gpu.render_tile(0, 0, 64, 64, s, 1)[..., :3] for s in range(NS)]).astype(np.float64)
    r = np.stack([os_.render_tile_bdpt(0, 0, 64, 64, s, 1)[0][..., :3] for s in range(NS)]).astype(np.float64)
    np.testing.assert_allclose(g.sum(0), got[..., :3], rtol=1e-5, atol=1e-6)   # the samples add up to the tile
    border = np.moveaxis(bd, -1, 0) > 0
    scale = max(np.abs(r).mean(), 1e-6)
    e = np.abs(g - r).max(-1) / np.maximum(np.abs(r).max(-1), 1e-3 * scale)
    off = e > 2e-4
    print(f"[bdpt] {int(off.sum())} of {e.size} samples beyond 2e-4, {int((off & border).sum())} flagged "
          f"({int(border.sum())} flagged in all); largest unflagged {e[~border].max():.3g}", flush=True)
    if glossy in ("glass", "glass_texture"):   # the oracle against the reference leaves two unflagged within 6e-4
        assert e[~border].max() <= 1e-3
        assert (off & ~border).sum() <= 5e-4 * e.size
    else:
        assert not (off & ~border).any()
    assert off.sum() <= (5e-3 if glossy else 3e-3) * e.size


def test_bdpt_frame_equals_tile_and_clamp(gpu, oracle):
    """ahip_render_frame with the BDPT integrator adds the same sums as one tile over the film, and the
    firefly clamp parameter reaches the kernels: a clamp of 0.05 (low enough to cap many s, t >= 2
    strategies) and 0 (off), each against the oracle's."""
    sc = scenes.cornell(width=96, height=80)
    os_ = oracle.scene(sc)
    for clamp in (0.05, 0.0):
        gpu.prepare(sc, integrator=BDPT, firefly_clamp=clamp)
        tile = gpu.render_tile(0, 0, 96, 80, 0, 2)
        gpu.clear_accum()
        gpu.render_frame(0, 2)
        frame = gpu.download_accum()
        assert np.array_equal(frame, tile), clamp
        ref = os_.render_tile_bdpt(0, 0, 96, 80, 0, 2, clamp)[0]
        assert abs(tile[..., :3].astype(np.float64).mean() / ref[..., :3].astype(np.float64).mean() - 1) < 1e-3, clamp
        if clamp:
            clamped = tile
    assert clamped[..., :3].mean() < 0.99 * tile[..., :3].mean()


def test_bdpt_refusals(gpu):
    """BDPT is refused for scenes it does not take (at ahip_set_params once a scene is prepared, and
    at render time), and by the path-tracer-only entry points."""
    sc = scenes.cornell(width=16, height=16)
    sc.env_radiance = (0.1, 0.1, 0.1)   # a constant environment radiance is no light of BDPT's
    with pytest.raises(Exception):
        gpu.prepare(sc, integrator=BDPT)
    sun = scenes.with_sun(scenes.cornell(width=16, height=16))
    sun.lights[-1].scene_radius = -1.0   # invalid for any integrator
    with pytest.raises(Exception):
        gpu.prepare(sun)
    textured = scenes.cornell(width=16, height=16, glossy=True)
    textured.materials[4].roughness_texture = 0
    textured.textures = [np.full((4, 4), 0.3, np.float32)]
    with pytest.raises(Exception):
        gpu.prepare(textured, integrator=BDPT)
    gpu.prepare(sc)
    with pytest.raises(Exception):
        gpu.set_params(sc.max_depth, sc.min_depth, sc.seed, integrator=BDPT)
    gpu.render_tile(0, 0, 8, 8, 0, 1)   # the path tracer still serves the scene
    sc = scenes.cornell(width=32, height=32)
    gpu.prepare(sc, integrator=BDPT)
    gpu.render_tile(0, 0, 8, 8, 0, 1)
    cam = gpu.camera
    with pytest.raises(Exception):
        gpu.render_batch([cam], [0], [1], sc.max_depth)
    with pytest.raises(Exception):
        gpu.set_params(sc.max_depth, sc.min_depth, sc.seed, integrator=7)
    with pytest.raises(Exception):
        gpu.set_params(sc.max_depth, sc.min_depth, sc.seed, integrator=BDPT, firefly_clamp=-1.0)
    gpu.set_params(sc.max_depth, sc.min_depth, sc.seed)   # back to the path tracer
    gpu.render_batch([cam], [0], [1], sc.max_depth)


def test_bdpt_cornell_frame_bands_vs_oracle(gpu, oracle):
    """The Cornell frame at its configured size (512x512 x 64 spp, depth 8): two 4-row bands against
    the oracle's render of the same rows, sample weights exact and band means within 1e-3."""
    sc = scenes.cornell()
    os_ = oracle.scene(sc)
    gpu.prepare(sc, integrator=BDPT)
    gpu.clear_accum()
    gpu.render_frame(0, sc.spp)
    frame = gpu.download_accum()
    for y0 in (128, 384):
        print(f"[bdpt] cornell 512x512 x 64 spp, rows {y0}..{y0 + 3}", flush=True)
        ref = os_.render_tile_bdpt(0, y0, 512, 4, 0, sc.spp)[0]
        got = frame[y0:y0 + 4]
        assert np.array_equal(got[..., 3], ref[..., 3])
        m, mr = got[..., :3].astype(np.float64).mean(), ref[..., :3].astype(np.float64).mean()
        print(f"[bdpt] band means {m:.6f} / {mr:.6f}", flush=True)
        assert abs(m / mr - 1) < 1e-3


@pytest.mark.parametrize("sun", [False, True])
def test_bdpt_adapter_in_reference_render_loop(sun):
    """The reference-side adapter in BDPT mode (HipPathIntegrator(device, maxDepth, BDPT, clamp), the
    drop-in for RenderSession's BDPTIntegrator, RenderSession.cpp:459-461) in the reference's tile
    loop, against the reference's own BDPTIntegrator in the same loop on the same SceneView: sample
    counts exact, AOV sums equal, image mean within 2e-4, and the pixels beyond 1e-3 no more than the
    per-sample divergence allows (samples whose connection visibility hangs on the last bits of a hit
    distance, 0.2%: at 8 spp about 1.6% of pixels; measured 0.6%, means within 6e-6).  Pixels whose
    samples hit the ceiling and the light quad, whose normals face opposite ways, have a zero mean
    normal but a record: the adapter tells records by a nonzero albedo or normal.  A scene the
    device BDPT does not take leaves the adapter unprepared.  sun: a directional light through the
    box's open side besides the quad, whose scene sphere the adapter reads back from
    DirectionalLight::sampleLe (to ~1e-7 of the radius)."""
    from adapter_harness import AdapterHarness, available
    from oracle_bindings import Oracle
    if not available():
        pytest.fail("oracle/_ref/libadapter_harness.so not built")
    sc = scenes.cornell(width=200, height=136)
    if sun:
        sc = scenes.with_sun(sc)
    cam = Oracle().pinhole(sc.cam_from, sc.cam_at, sc.cam_up, sc.vfov, sc.width, sc.height)
    h = AdapterHarness(sc, cam, sc.max_depth, integrator=BDPT, clamp=10.0)
    st = h.status()
    print(f"[bdpt-adapter] {st}", flush=True)
    assert st["valid"] and st["prepared"], st
    got, ga, gn = h.render("adapter", 8, threads=8)
    ref, ra, rn = h.render("reference_bdpt", 8, threads=8)
    assert np.array_equal(got[..., 3], ref[..., 3]) and (got[..., 3] == 8).all()
    g, r = got[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    err = np.abs(g - r).max(-1) / np.maximum(np.abs(r).max(-1), 1e-3 * max(r.mean(), 1e-6))
    print(f"[bdpt-adapter] cornell sun={sun} {sc.width}x{sc.height} x 8 spp: {(err > 1e-3).sum()} of {err.size} pixels "
          f"beyond 1e-3; means {g.mean():.6f} / {r.mean():.6f}", flush=True)
    assert (err > 1e-3).mean() < 0.02
    assert abs(g.mean() / r.mean() - 1) < 2e-4
    np.testing.assert_array_equal(ga, ra)
    np.testing.assert_allclose(gn, rn, rtol=0, atol=1e-6)
    lit = scenes.cornell(width=16, height=16)
    lit.env_radiance = (0.2, 0.2, 0.2)   # a constant background: an environment, which BDPT refuses
    h2 = AdapterHarness(lit, Oracle().pinhole(lit.cam_from, lit.cam_at, lit.cam_up, lit.vfov, 16, 16),
                        lit.max_depth, integrator=BDPT)
    st2 = h2.status()
    assert st2["valid"] and not st2["prepared"], st2


@pytest.mark.parametrize("name", ["dome", "dome_sun", "cornell_sun"])
def test_bdpt_infinite_lights_vs_oracle(gpu, oracle, name):
    """The dome and directional lights in BDPT (the oracle is held to the reference's BDPTIntegrator
    on the same three scenes by tests/test_oracle_bdpt.py): environment vertices, light subpaths from
    the sky's and the sun's emission disks, the (1, t) strategy along the light subpath's direction.
    Per sample against the oracle: sample counts and AOVs equal (normals within 1e-4: the FMA build
    interpolates sphere normals differently), means within 1e-4, at most 0.5% of samples beyond 2e-4
    and every one within 2e-2 — the oracle built uncontracted differs from the reference's own build
    by up to 6e-3 on these scenes, the device's contraction pattern is a third one."""
    from test_oracle_bdpt import INFINITE_SCENES
    sc = INFINITE_SCENES[name]()
    os_ = oracle.scene(sc)
    gpu.prepare(sc, integrator=BDPT)
    print(f"[bdpt] {name} 64x64 x {NS} spp", flush=True)
    got, alb, nrm = gpu.render_tile_aovs(0, 0, 64, 64, 0, NS)
    ref, oalb, onrm = os_.render_tile_bdpt(0, 0, 64, 64, 0, NS)
    assert np.array_equal(got[..., 3], ref[..., 3])
    np.testing.assert_allclose(alb, oalb, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(nrm, onrm, rtol=0, atol=1e-4)
    gm, rm = got[..., :3].astype(np.float64).mean(), ref[..., :3].astype(np.float64).mean()
    g = np.stack([gpu.render_tile(0, 0, 64, 64, s, 1)[..., :3] for s in range(NS)]).astype(np.float64)
    r = np.stack([os_.render_tile_bdpt(0, 0, 64, 64, s, 1)[0][..., :3] for s in range(NS)]).astype(np.float64)
    e = np.abs(g - r).max(-1) / np.maximum(np.abs(r).max(-1), 1e-3 * max(np.abs(r).mean(), 1e-6))
    print(f"[bdpt] {name}: means {gm:.7g} / {rm:.7g}; {int((e > 2e-4).sum())} of {e.size} samples beyond "
          f"2e-4, largest {e.max():.3g}", flush=True)
    assert abs(gm / rm - 1) < 1e-4
    assert (e > 2e-4).sum() <= 5e-3 * e.size
    assert e.max() <= 2e-2
