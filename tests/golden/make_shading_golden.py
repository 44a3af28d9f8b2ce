"""Write tests/golden/shading_vectors.npz from the reference's own shading code.

oracle/_ref/libanacapa_ref.so (built by oracle/Makefile from the reference checkout) runs:
  * MarschnerHairMaterial as the reference runs it (Kajiya-Kay), and its R/TT/TRT code with the two
    Kajiya-Kay early returns removed (ref_hair_live.cpp), plus the free M_p / N_p / A_p helpers and
    the fiber geometry;
  * StandardSurfaceMaterial's lobe helpers (sampleGGX, sampleDiffuse, evalGGX, evalCombined) under
    the opaque branch of sample / evaluate for constant parameters;
  * DomeLight's distribution build, sampling, pdf and bilinear lookup over given pixels.
Every input row and every output value is stored, so the tests need neither the reference nor
this script.  Run from the repository root:  python tests/golden/make_shading_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from anacapa_amd import scenes  # noqa: E402
from anacapa_amd import _abi as A  # noqa: E402
from oracle_bindings import Reference, probe_inputs  # noqa: E402

MARSCHNER = A.AHIP_FLAG_HAIR_MARSCHNER


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def hair_materials():
    """Config-3 hair (beta 0.3 / 0.3, 2 degrees, IOR 1.55, eumelanin sigma_a) and two others that
    reach the other branches of M_p (v > 0.1 for every lobe; a sharp R lobe)."""
    return [
        scenes.hair(sigma_a=tuple(float(x) for x in scenes.eumelanin_sigma_a(0.8)[0]), eta=1.55, beta_m=0.3, beta_n=0.3, alpha=2.0),
        scenes.hair(sigma_a=(0.40, 0.65, 1.35), eta=1.4, beta_m=0.8, beta_n=0.9, alpha=4.0),
        scenes.hair(sigma_a=(0.05, 0.1, 0.3), eta=1.6, beta_m=0.08, beta_n=0.15, alpha=1.0),
    ]


def surface_materials(rng):
    """Opaque standard_surface cases whose mixtures differ: pure metal, pure diffuse, dielectric
    specular, coated, near-mirror roughness, and config-2-like random draws."""
    S = scenes.standard_surface
    mats = [
        S((0.9, 0.6, 0.2), metalness=1.0, roughness=0.3, coat=0.0),
        S((0.5, 0.7, 0.3), metalness=0.0, roughness=0.6, specular=0.0, coat=0.0),
        S((0.2, 0.3, 0.8), metalness=0.0, roughness=0.25, specular=1.0, coat=0.0),
        S((0.7, 0.2, 0.2), metalness=0.0, roughness=0.8, specular=0.5, coat=0.5),
        S((0.8, 0.8, 0.8), metalness=1.0, roughness=0.05, coat=0.3),
        S((0.3, 0.3, 0.3), metalness=0.0, roughness=1.0, specular=1.0, coat=0.0, ior=2.2),
    ]
    for _ in range(4):
        mats.append(S(tuple(rng.uniform(0, 1, 3)), metalness=float(rng.uniform() < 0.3),
                      roughness=float(rng.uniform(0.05, 1.0)), coat=float(rng.uniform(0, 0.5))))
    return mats


def curve_rows(rng, n):
    nn = unit(rng, n)
    r = unit(rng, n)
    t = r - nn * np.sum(r * nn, 1, keepdims=True)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    k = n // 5
    t[:k] = unit(rng, k)                                  # tangents not perpendicular to the normal
    wo = unit(rng, n)
    wi = unit(rng, n)
    # wo exactly along the fiber and within 1e-4 of it (woPerp degenerate), taken from rows whose
    # tangent is not perpendicular to the normal so that the front-face test stays well defined
    wo[:16] = t[:16] * rng.choice([-1.0, 1.0], (16, 1))
    wo[16:32] = t[16:32] + 1e-4 * unit(rng, 16)
    wo[16:32] /= np.linalg.norm(wo[16:32], axis=1, keepdims=True)
    col = np.ones((n, 3))
    col[n // 3:2 * n // 3] = rng.uniform(0.05, 0.97, (2 * n // 3 - n // 3, 3))   # per-strand colour
    col[2 * n // 3:] = rng.uniform(0.0, 0.1, (n - 2 * n // 3, 3))                # near black
    u = rng.uniform(0, 1, (n, 2))
    u[:8, 0] = 0.0                                         # Box-Muller at u1 = 0
    uc = rng.uniform(0, 1, n)
    return probe_inputs(nn, t, col, wo, wi, u, uc, 1.0)


def surface_rows(rng, n):
    nn = unit(rng, n)
    wo = unit(rng, n)
    wo = np.where(np.sum(wo * nn, 1, keepdims=True) < 0, -wo, wo)   # sample() needs wo above n
    wo[:n // 10] = unit(rng, n // 10)                                  # and some below (rejected)
    wi = unit(rng, n)
    u = rng.uniform(0, 1, (n, 2))
    uc = rng.uniform(0, 1, n)
    return probe_inputs(nn, np.zeros((n, 3)), np.ones((n, 3)), wo, wi, u, uc, 0.0)


def main():
    R = Reference()
    rng = np.random.default_rng(20261204)
    out = {}

    # ---- hair: constants, probes in both modes, lobe sums, helpers, geometry
    hm = hair_materials()
    out["hair_mats"] = np.stack([np.frombuffer(bytes(m), np.uint8) for m in hm])
    out["hair_consts"] = np.stack([R.hair_consts(m) for m in hm])
    rows = curve_rows(rng, 2000)
    out["hair_rows"] = rows
    out["hair_kk"] = np.stack([R.probe(m, rows, 0) for m in hm])
    out["hair_marschner"] = np.stack([R.probe(m, rows, MARSCHNER) for m in hm])
    g_in = np.concatenate([rows[:, 9:12], rows[:, 12:15], rows[:, 3:6], rows[:, 0:3]], 1)
    out["hair_geom_in"] = g_in
    out["hair_geom"] = R.hair_geom(g_in)
    geo = out["hair_geom"]
    sig = rng.uniform(0.0, 3.0, (len(rows), 3)).astype(np.float32)
    l_in = np.concatenate([geo, sig], 1)
    out["hair_lobes_in"] = l_in
    out["hair_lobes"] = np.stack([R.hair_lobes(m, l_in) for m in hm])
    n = 2000
    cosI, cosR = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    sinI = np.sqrt(1 - cosI ** 2) * rng.choice([-1, 1], n)
    sinR = np.sqrt(1 - cosR ** 2) * rng.choice([-1, 1], n)
    v = rng.choice([1e-6, 0.004, 0.02, 0.0846, 0.1, 0.1001, 0.34, 1.0, 4.0], n)
    mp_in = np.stack([cosI, sinI, cosR, sinR, v], 1).astype(np.float32)
    out["mh_mp_in"], out["mh_mp"] = mp_in, R.mh_fn(0, *mp_in.T)
    np_in = np.stack([rng.uniform(-7, 7, n), rng.integers(0, 3, n), rng.uniform(0.02, 1.5, n),
                      rng.uniform(-1.5, 1.5, n), rng.uniform(-1.0, 1.0, n)], 1).astype(np.float32)
    out["mh_np_in"], out["mh_np"] = np_in, R.mh_fn(1, *np_in.T)
    tl_in = np.stack([rng.uniform(-3.5, 3.5, n), rng.uniform(0.02, 1.5, n)], 1).astype(np.float32)
    out["mh_tl_in"], out["mh_tl"] = tl_in, R.mh_fn(2, *tl_in.T)
    st_in = np.stack([rng.uniform(0, 1, n), rng.uniform(0.02, 1.5, n)], 1).astype(np.float32)
    st_in[:4, 0] = [0.0, 1.0, 0.5, 1e-7]
    out["mh_stl_in"], out["mh_stl"] = st_in, R.mh_fn(3, *st_in.T)
    ga_in = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    ga_in[:4, 0] = [0.0, 1e-7, 1.0, 0.5]
    out["mh_gauss_in"], out["mh_gauss"] = ga_in, R.mh_fn(4, *ga_in.T)
    x = np.concatenate([rng.uniform(0, 30, n), [0.0, 11.999, 12.0, 12.001]]).astype(np.float32)
    out["mh_i0_in"], out["mh_logi0"], out["mh_i0"] = x, R.mh_fn(5, x), R.mh_fn(6, x)
    fr_in = np.stack([rng.uniform(0, 1, n), rng.uniform(1.0, 2.0, n)], 1).astype(np.float32)
    out["mh_fr_in"], out["mh_fr"] = fr_in, R.mh_fn(7, *fr_in.T)
    ph_in = np.stack([rng.integers(0, 3, n), rng.uniform(-1.6, 1.6, n), rng.uniform(-1.0, 1.0, n)], 1).astype(np.float32)
    out["mh_phi_in"], out["mh_phi"] = ph_in, R.mh_fn(8, *ph_in.T)
    ap_in = np.stack([rng.uniform(0.01, 1, n), rng.uniform(1.2, 1.8, n), rng.uniform(-1, 1, n),
                      rng.uniform(0, 3, n), rng.uniform(0, 3, n), rng.uniform(0, 3, n)], 1).astype(np.float32)
    out["mh_ap_in"], out["mh_ap"] = ap_in, R.mh_ap(ap_in)

    # ---- standard_surface: probes, one GGX lobe
    sm = surface_materials(rng)
    out["ss_mats"] = np.stack([np.frombuffer(bytes(m), np.uint8) for m in sm])
    srows = surface_rows(rng, 1500)
    out["ss_rows"] = srows
    out["ss_probe"] = np.stack([R.probe(m, srows) for m in sm])
    m = 2000
    wo_l, wi_l = unit(rng, m), unit(rng, m)
    wo_l[:, 2], wi_l[:, 2] = np.abs(wo_l[:, 2]), np.abs(wi_l[:, 2])
    wi_l[: m // 10, 2] *= -1
    a2 = (np.maximum(1e-4, rng.uniform(0, 1, m) ** 2) ** 2).astype(np.float32)
    f0 = rng.uniform(0, 1, (m, 3))
    ggx_in = np.concatenate([wo_l, wi_l, a2[:, None], f0], 1).astype(np.float32)
    out["ss_ggx_in"], out["ss_ggx"] = ggx_in, R.ss_eval_ggx(ggx_in)

    # ---- dome: small maps kept whole, plus the 2048x1024 sun HDRI through a pixel checksum
    def dome_case(key, rgb, rot=None, intensity=1.0, ns=2000):
        d = R.dome(rgb, intensity, rot)
        u = rng.uniform(0, 1, (ns, 2)).astype(np.float32)
        u[:6] = [[0, 0], [1 - 2 ** -24, 1 - 2 ** -24], [0.5, 0.5], [0, 1 - 2 ** -24], [1 - 2 ** -24, 0], [0.25, 0.75]]
        dirs = unit(rng, ns).astype(np.float32)
        dirs[:6] = [[0, 1, 0], [0, -1, 0], [1, 0, 0], [0, 0, 1], [-1, 0, 0], [0, 0, -1]]
        out[key + "_u"], out[key + "_sample"] = u, d.sample(u)
        out[key + "_dirs"], out[key + "_eval"] = dirs, d.eval(dirs)
        out[key + "_rot"] = np.asarray(np.eye(3) if rot is None else rot, np.float32)
        out[key + "_intensity"] = np.float32(intensity)
        return d

    px22 = np.array([[[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]], [[7.0, 8.0, 9.0], [10.0, 11.0, 12.0]]], np.float32)
    out["dome2x2_px"] = px22
    dome_case("dome2x2", px22)
    hot = np.full((32, 64, 3), 1e-3, np.float32)
    hot[9, 41] = [50.0, 40.0, 30.0]                       # one hot texel
    out["domehot_px"] = hot
    dome_case("domehot", hot)
    sky = scenes.procedural_hdri(96, 48, sun_peak=5e4, sun_sigma_deg=4.0).reshape(48, 96, 3).astype(np.float32)
    out["domesky_px"] = sky
    d = dome_case("domesky", sky)
    out["domesky_tables"] = np.concatenate([np.concatenate(R.dome_tables(d)[:2]),
                                            R.dome_tables(d)[2].ravel(), R.dome_tables(d)[3].ravel(),
                                            [R.dome_tables(d)[4]]]).astype(np.float32)
    c, s_ = np.cos(0.7), np.sin(0.7)
    rot = np.array([[c, 0, -s_], [0, 1, 0], [s_, 0, c]], np.float32)   # a turn about +Y
    dome_case("domerot", sky, rot=rot, intensity=2.5)
    big = scenes.procedural_hdri().reshape(1024, 2048, 3).astype(np.float32)
    out["domebig_sha"] = np.frombuffer(hashlib.sha256(big.tobytes()).digest(), np.uint8).copy()
    dome_case("domebig", big, ns=4000)

    dst = os.path.join(HERE, "shading_vectors.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main()
