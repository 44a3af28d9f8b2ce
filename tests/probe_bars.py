"""The bars for material-probe comparisons (sample / evaluate rows in the layout of
ahip_material_probe), shared by the oracle's tests against the reference vectors and the GPU's
tests against both.  The bars and why each is what it is are stated in test_oracle_shading.py."""
import numpy as np


def rel(a, b, floor=1e-30):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)


_rel = rel


def probe_close(got, ref, rows, hair, what="", sharp=False, sigma_rtol=0.0, floor_rel=1e-6, max_flips=0):
    """The probe bars of the module docstring; returns the number of rows compared.  sharp: the
    material has a GGX lobe of roughness < 0.18 (alpha^2 < 1e-3), whose sampled f and pdf are
    each held only within 0.15 (and 2e-2 on three rows in four); their ratio keeps the full bar.
    sigma_rtol: hair effectiveSigmaA is bit-exact between CPU builds (glibc pow / log); the device's
    powf / logf differ by an ulp, so device comparisons pass 1e-5.  floor_rel: values below this
    fraction of the rows' largest are compared absolutely (glass across the surface: 1 - F near the
    critical angle cancels in float, so its far-tail values keep few significant digits).
    max_flips: rows allowed to differ in sample validity, left out of the sampled-value bars (glass:
    its reflect-or-refract choice uComponent < F sits where F climbs steeply to 1 near total internal
    reflection, and the refracted branch's TIR test sin^2 t >= 1 on its edge, so the last bits of F
    can decide the branch)."""
    wo, n, t = rows[:, 9:12], rows[:, 0:3], rows[:, 3:6]
    ok = np.abs((wo * n).sum(1)) >= 1e-4
    if hair:
        ok &= np.abs((wo * t).sum(1)) <= 1 - 1e-6
    g, r = got[ok], ref[ok]
    floor = floor_rel * max(np.abs(r[:, 0:4]).max(), 1e-30)
    e_eval = _rel(g[:, 0:4], r[:, 0:4], floor).max(1)
    assert e_eval.max() <= 2e-3, f"{what}: evaluate differs by {e_eval.max():.3g}"
    if hair:
        if sigma_rtol == 0.0:
            assert np.array_equal(g[:, 13:16], r[:, 13:16]), f"{what}: effectiveSigmaA"
        else:
            assert rel(g[:, 13:16], r[:, 13:16]).max() <= sigma_rtol, f"{what}: effectiveSigmaA"
        near_axis = lambda x: np.abs((x[:, 5:8] * t[ok]).sum(1)) > 1 - 1e-5
        degenerate = near_axis(g) | near_axis(r)
    else:
        degenerate = np.zeros(len(g), bool)
    vg, vr = g[:, 12] > 0, r[:, 12] > 0
    flips = (vg != vr) & ~degenerate
    if flips.any():
        print(f"[probe] {what}: {int(flips.sum())} of {len(vg)} rows differ in sample validity", flush=True)
    assert flips.sum() <= max_flips, f"{what}: sample validity ({int(flips.sum())} rows)"
    both = vg & vr & ~degenerate
    assert np.abs(g[both, 5:8] - r[both, 5:8]).max(initial=0) <= 5e-4, f"{what}: sampled direction"
    if not both.any():
        return int(ok.sum())
    sfloor = floor_rel * max(np.abs(r[both, 8:12]).max(), 1e-30)
    e_s = _rel(g[both, 8:12], r[both, 8:12], sfloor).max(1)
    ratio = lambda x: x[:, 8:11] / x[:, 11:12]
    e_r = _rel(ratio(g[both]), ratio(r[both])).max(1)
    share = 0.75 if sharp else 0.99
    assert (e_s <= 2e-2).mean() >= share and e_s.max() <= 0.15, \
        f"{what}: sampled f/pdf, {(e_s > 2e-2).sum()} of {len(e_s)} rows beyond 2e-2, max {e_s.max():.3g}"
    assert (e_r <= 2e-2).mean() >= 0.99 and e_r.max() <= 0.15, \
        f"{what}: f/pdf ratio, {(e_r > 2e-2).sum()} of {len(e_r)} rows beyond 2e-2, max {e_r.max():.3g}"
    return int(ok.sum())
