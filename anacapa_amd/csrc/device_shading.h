// device_shading.h — materials, lights and the shading frame on gfx950.
//
// Lambert / Emissive: src/shading/Lambertian.h:12-99.
// StandardSurface (opaque layered branch: coat GGX + metal GGX + dielectric GGX + diffuse,
// sample one lobe then report the full-mixture f and pdf): src/shading/StandardSurface.h:20-92,
// 152-169, 231-241, 351-405, 410-503, 516-661.
// AreaLight: src/shading/lights/AreaLight.h:123-180.  DomeLight (equirect, 2D piecewise-constant
// importance sampling, bilinear lookup): src/shading/lights/DomeLight.h:19-55, 117-133, 228-364.
// ShadingContext: include/anacapa/shading/ShadingContext.h:28-59.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/anacapa_hip.h"
#include "device_math.h"
#include "device_types.h"

namespace ahip {

struct Frame {
    V3 n, ng, t, bt;
    V3 color;      // per-strand colour (curves); white for surfaces
    bool front;
};
__device__ __forceinline__ Frame makeFrame(V3 n, V3 ng, V3 rayDir) {
    Frame f;
    f.front = dot(-rayDir, ng) > 0.f;
    f.n = f.front ? n : -n;
    f.ng = f.front ? ng : -ng;
    onb(f.n, f.t, f.bt);
    f.color = V3{1.f, 1.f, 1.f};
    return f;
}
// Curve hit: tangent frame around the fiber (ShadingContext.h:39-44).
__device__ __forceinline__ Frame makeCurveFrame(V3 n, V3 tang, V3 rayDir, V3 color) {
    Frame f;
    f.front = dot(-rayDir, n) > 0.f;
    f.n = f.front ? n : -n;
    f.ng = f.n;
    f.t = tang;
    f.bt = safeNormalize(cross(f.n, tang));
    f.color = color;
    return f;
}

// MarschnerHairMaterial::effectiveSigmaA (MarschnerHair.h:430-450)
__device__ __forceinline__ V3 hairSigmaA(const DMat& m, V3 c) {
    if (c.x > 0.98f && c.y > 0.98f && c.z > 0.98f) return ld3(m.hairSigma);
    float lr = powf(fmaxf(c.x, 0.f), 2.2f), lg = powf(fmaxf(c.y, 0.f), 2.2f), lb = powf(fmaxf(c.z, 0.f), 2.2f);
    if (lr < 0.01f && lg < 0.01f && lb < 0.01f) return ld3(m.hairSigma);
    return V3{-logf(fmaxf(lr, 0.001f)), -logf(fmaxf(lg, 0.001f)), -logf(fmaxf(lb, 0.001f))};
}

// Kajiya-Kay, the hair BSDF the reference evaluates at run time (MarschnerHair.h:240, 395-422).
__device__ __forceinline__ V3 hairKK(const DMat& m, const struct Frame& fr, V3 wo, V3 wi, float& pdf) {
    V3 sa = hairSigmaA(m, fr.color);
    V3 f{expf(-sa.x) * kInvPi, expf(-sa.y) * kInvPi, expf(-sa.z) * kInvPi};
    float dTL = dot(fr.t, wi), dTV = dot(fr.t, wo);
    float sTL = sqrtf(fmaxf(0.f, 1.f - dTL * dTL)), sTV = sqrtf(fmaxf(0.f, 1.f - dTV * dTV));
    float spec = dTL * dTV + sTL * sTV;
    if (spec > 0.f) {
        float s = 0.15f * (powf(spec, 20.f) * kInvPi);
        f = f + V3{s, s, s};
    }
    pdf = fmaxf(0.f, dot(wi, fr.n)) * kInvPi;
    return f;
}
// ---- Marschner R / TT / TRT, the model MarschnerHairMaterial carries below its Kajiya-Kay returns
// (MarschnerHair.h:49-185 helpers, :243-257 evaluate, :284-375 sample, :490-556 lobes and pdf),
// selected by AHIP_FLAG_HAIR_MARSCHNER.  Per-material constants (variances, logistic scale, tilts)
// are computed on the host as the constructor computes them.
namespace mh {
__device__ __forceinline__ float I0(float x) {                                  // :62-71
    float sum = 0.f, x2i = 1.f, denom = 1.f;
    const float xx = x * x;
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        if (i > 0) {
            x2i *= xx;
            denom *= (float)i * (float)i * 4.f;
        }
        sum += x2i / denom;
    }
    return sum;
}
__device__ __forceinline__ float logI0(float x) {                               // :73-77
    if (x > 12.f) return x + 0.5f * (-logf(k2Pi) + logf(1.f / x) + 1.f / (8.f * x));
    return logf(I0(x));
}
__device__ __forceinline__ float Mp(float cosI, float sinI, float cosR, float sinR, float v) {   // :88-98
    if (v < 1e-5f) v = 1e-5f;
    float a = cosI * cosR / v, b = sinI * sinR / v;
    if (v <= 0.1f) return expf(logI0(a) - b - 1.f / v + 0.6931472f + logf(0.5f / v));
    return expf(-b) * I0(a) / (2.f * v * sinhf(1.f / v));
}
__device__ __forceinline__ float logisticCDF(float x, float s) { return 1.f / (1.f + expf(-x / s)); }
__device__ __forceinline__ float trimmedLogistic(float x, float s) {            // :103-116
    float e = expf(-fabsf(x) / s);
    return (e / (s * (1.f + e) * (1.f + e))) / (logisticCDF(kPi, s) - logisticCDF(-kPi, s));
}
__device__ __forceinline__ float sampleTrimmedLogistic(float u, float s) {      // :119-124
    float a = logisticCDF(-kPi, s), b = logisticCDF(kPi, s);
    float x = -s * logf(1.f / (a + u * (b - a)) - 1.f);
    return fminf(fmaxf(x, -kPi), kPi);
}
__device__ __forceinline__ float Phi(int p, float gO, float gT) {               // :129-131
    return 2.f * (float)p * gT - 2.f * gO + (float)p * kPi;
}
__device__ __forceinline__ float Np(float phi, int p, float s, float gO, float gT) {   // :133-142
    float x = phi - Phi(p, gO, gT);
    while (x > kPi) x -= k2Pi;
    while (x < -kPi) x += k2Pi;
    return trimmedLogistic(x, s);
}
__device__ __forceinline__ float clampH(float h) { return fminf(fmaxf(h, -1.f + 1e-5f), 1.f - 1e-5f); }
__device__ __forceinline__ float fresnel(float cosI, float eta) {              // :46-55
    float s2i = fmaxf(0.f, 1.f - cosI * cosI);
    float s2t = s2i / (eta * eta);
    if (s2t >= 1.f) return 1.f;
    float ct = sqrtf(1.f - s2t);
    float rs = (cosI - eta * ct) / (cosI + eta * ct);
    float rp = (eta * cosI - ct) / (eta * cosI + ct);
    return 0.5f * (rs * rs + rp * rp);
}
// A_p of R, TT, TRT (:146-176)
__device__ __forceinline__ void Ap(float cosO, float eta, float h, V3 sa, V3 ap[3]) {
    float s2 = fmaxf(0.f, 1.f - cosO * cosO);
    float etaP = sqrtf(fmaxf(0.f, eta * eta - s2)) / cosO;
    float sgt = clampH(h / (etaP > 1e-5f ? etaP : 1e-5f));
    float cgt = sqrtf(fmaxf(0.f, 1.f - sgt * sgt));
    V3 T{expf(-sa.x * 2.f * cgt), expf(-sa.y * 2.f * cgt), expf(-sa.z * 2.f * cgt)};
    float cgo = sqrtf(fmaxf(0.f, 1.f - h * h));
    float fr = fresnel(fmaxf(0.f, cosO * cgo), eta);
    ap[0] = V3{fr, fr, fr};
    ap[1] = T * ((1.f - fr) * (1.f - fr));
    ap[2] = ap[1] * T * fr;
}
__device__ __forceinline__ float lum(V3 s) { return 0.2722287f * s.x + 0.6740818f * s.y + 0.0536895f * s.z; }
// azimuths of the entry point and of the refracted ray for impact parameter h (:498-502)
__device__ __forceinline__ void gammas(float cosO, float eta, float h, float& gO, float& gT) {
    float etaP = sqrtf(fmaxf(0.f, eta * eta - (1.f - cosO * cosO))) / cosO;
    gO = asinf(clampH(h));
    gT = asinf(clampH(h / (etaP > 1e-5f ? etaP : 1e-5f)));
}
// sum over p of M_p A_p N_p / max(1/4, cos^2 theta_d) (:490-522) and the lobe mixture pdf
// (:528-556), sharing A_p and the azimuths
__device__ __forceinline__ V3 lobes(const DMat& m, float sO, float cO, float sI, float cI, float phi, float h, V3 sa,
                                    float& pdf) {
    float cd = sqrtf(fmaxf(0.f, 0.5f * (1.f + cO * cI + sO * sI)));
    float denom = fmaxf(0.25f, cd * cd);
    float gO, gT;
    gammas(cO, m.hairEta, h, gO, gT);
    V3 ap[3];
    Ap(cO, m.hairEta, h, sa, ap);
    float w0 = lum(ap[0]), w1 = lum(ap[1]), w2 = lum(ap[2]);
    float total = w0 + w1 + w2;
    V3 f{0.f, 0.f, 0.f};
    float pd = 0.f;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        float s2a, c2a;
        sincosf(2.f * m.hairAlpha[p], &s2a, &c2a);
        float so = sO * c2a + cO * s2a;
        float co = sqrtf(fmaxf(0.f, 1.f - so * so));
        float mp = Mp(cI, sI, co, so, m.hairV[p]);
        float np = Np(phi, p, m.hairS, gO, gT);
        f = f + ap[p] * (mp * np);
        float wp = p == 0 ? w0 : (p == 1 ? w1 : w2);
        pd += (wp / total) * (mp * cI) * np;
    }
    pdf = total < 1e-8f ? 0.f : fmaxf(0.f, pd);
    return f * (1.f / denom);
}
__device__ __forceinline__ void longitudinal(V3 d, V3 t, float& s, float& c) {   // :456-460
    s = dot(d, t);
    c = sqrtf(fmaxf(0.f, 1.f - s * s));
}
__device__ __forceinline__ float azimuth(V3 wo, V3 wi, V3 t, float sO, float sI) {   // :463-475
    V3 po = wo - t * sO, pi = wi - t * sI;
    float lo = len(po), li = len(pi);
    if (lo < 1e-5f || li < 1e-5f) return 0.f;
    po = po * (1.f / lo);
    pi = pi * (1.f / li);
    float c = fminf(fmaxf(dot(po, pi), -1.f), 1.f);
    float s = dot(cross(pi, po), t);
    return atan2f(s, c);
}
__device__ __forceinline__ float impact(V3 wo, V3 t, V3 n, float sO) {          // :478-483
    V3 po = wo - t * sO;
    float l = len(po);
    if (l < 1e-5f) return 0.f;
    return clampH(dot(po * (1.f / l), n));
}
}  // namespace mh

// evaluate (:243-257): f without cosine and the forward pdf
__device__ __forceinline__ V3 hairMarschner(const DMat& m, const Frame& fr, V3 wo, V3 wi, float& pdf) {
    float sO, cO, sI, cI;
    mh::longitudinal(wo, fr.t, sO, cO);
    mh::longitudinal(wi, fr.t, sI, cI);
    float phi = mh::azimuth(wo, wi, fr.t, sO, sI);
    float h = mh::impact(wo, fr.t, fr.n, sO);
    return mh::lobes(m, sO, cO, sI, cI, phi, h, hairSigmaA(m, fr.color), pdf);
}

__device__ __forceinline__ V3 toLocal(const Frame& f, V3 v) { return V3{dot(v, f.t), dot(v, f.bt), dot(v, f.n)}; }
__device__ __forceinline__ V3 toWorld(const Frame& f, V3 v) { return f.t * v.x + f.bt * v.y + f.n * v.z; }

struct BSample {
    V3 f, wi;
    float pdf;
    bool delta;
};
__device__ __forceinline__ bool bsValid(const BSample& s) { return s.pdf > 0.f && !isBlack(s.f); }

// ---- GGX helpers (StandardSurface.h:20-76)
__device__ __forceinline__ float dGGX(float cosH, float a2) {
    if (cosH <= 0.f) return 0.f;
    float den = cosH * cosH * (a2 - 1.f) + 1.f;
    return a2 / (kPi * den * den);
}
__device__ __forceinline__ float g1Smith(float c, float a2) {
    if (c <= 0.f) return 0.f;
    return 2.f * c / (c + sqrtf(a2 + (1.f - a2) * c * c));
}
__device__ __forceinline__ float g2Smith(float co, float ci, float a2) {
    return g1Smith(fabsf(co), a2) * g1Smith(fabsf(ci), a2);
}
__device__ __forceinline__ V3 sampleHalf(V2 u, float a2) {
    float phi = 2.f * kPi * u.x;
    float c2 = (1.f - u.y) / stdmax(1.f + (a2 - 1.f) * u.y, 1e-10f);
    float ch = sqrtf(stdmax(0.f, c2)), sh = sqrtf(stdmax(0.f, 1.f - c2));
    float sp, cp;
    sincosf(phi, &sp, &cp);
    return V3{sh * cp, sh * sp, ch};
}
__device__ __forceinline__ float pdfReflect(float cosH, float a2, float dotVH) {
    if (cosH <= 0.f || dotVH <= 0.f) return 0.f;
    return dGGX(cosH, a2) * cosH / (4.f * dotVH);
}
__device__ __forceinline__ float schlickScalar(float c, float f0) {
    float x = 1.f - fabsf(c);
    float x5 = x * x * x * x * x;
    return f0 + (1.f - f0) * x5;
}
__device__ __forceinline__ V3 schlickColor(float c, V3 f0) {
    float x = 1.f - fabsf(c);
    float x5 = x * x * x * x * x;
    return f0 + (V3{1.f, 1.f, 1.f} - f0) * x5;
}

// Bilinear lookup with periodic wrap, v flipped as Texture.cpp:67-68 does.  Uncontracted, so the
// value is the oracle's bit for bit: an opacity decides the alpha test's branch by comparison
// (opacity <= 0, sample >= opacity), where one ulp can change which samples a path draws.
__device__ __forceinline__ float texBilinear(const DTexture& t, V2 uv) {
#pragma clang fp contract(off)
    float s = uv.x, tt = 1.f - uv.y;
    float fx = s * (float)t.w - 0.5f, fy = tt * (float)t.h - 0.5f;
    float x0f = floorf(fx), y0f = floorf(fy);
    float ax = fx - x0f, ay = fy - y0f;
    int x0 = (int)x0f, y0 = (int)y0f;
    int W = (int)t.w, H = (int)t.h;
    int xa = x0 % W; xa = xa < 0 ? xa + W : xa;
    int xb = (x0 + 1) % W; xb = xb < 0 ? xb + W : xb;
    int ya = y0 % H; ya = ya < 0 ? ya + H : ya;
    int yb = (y0 + 1) % H; yb = yb < 0 ? yb + H : yb;
    float p00 = t.px[(size_t)ya * t.w + xa], p01 = t.px[(size_t)ya * t.w + xb];
    float p10 = t.px[(size_t)yb * t.w + xa], p11 = t.px[(size_t)yb * t.w + xb];
    return (1.f - ay) * ((1.f - ax) * p00 + ax * p01) + ay * ((1.f - ax) * p10 + ax * p11);
}

// Layer weights at wo (StandardSurface.h:355-371); false when they sum to zero.
__device__ __forceinline__ bool layerWeights(const DMat& m, float woz, float& wc, float& wm, float& ws, float& wd) {
    float coatF = schlickScalar(woz, m.coatF0), specF = schlickScalar(woz, m.f0d);
    wc = m.coat * coatF;
    wm = m.metal * (1.f - wc);
    ws = m.spec * (1.f - m.metal) * specF * (1.f - wc);
    wd = m.base * (1.f - m.metal) * (1.f - m.spec * specF) * (1.f - wc);
    float sum = wc + wm + ws + wd;
    if (!(sum > 0.f)) return false;
    float inv = 1.f / sum;
    wc *= inv; wm *= inv; ws *= inv; wd *= inv;
    return true;
}

__device__ __forceinline__ V3 evalLobe(V3 wo, V3 wi, float a2, V3 f0, float& pf) {   // StandardSurface.h:578-603
    pf = 0.f;
    if (wo.z <= 0.f || wi.z <= 0.f) return V3{0.f, 0.f, 0.f};
    V3 wh = safeNormalize(wo + wi);
    if (wh.z <= 0.f) return V3{0.f, 0.f, 0.f};
    float dVH = stdmax(0.f, dot(wo, wh));
    V3 F = schlickColor(dVH, f0);
    float D = dGGX(wh.z, a2), G = g2Smith(wo.z, wi.z, a2);
    float co = fabsf(wo.z), ci = fabsf(wi.z);
    if (co < 1e-6f || ci < 1e-6f) return V3{0.f, 0.f, 0.f};
    pf = pdfReflect(wh.z, a2, dVH);
    return F * (D * G / (4.f * co * ci));
}

__device__ __forceinline__ V3 evalMixture(const DMat& m, V3 wo, V3 wi, float wc, float wm, float ws, float wd,
                                          float a2, V3 base, float& pdf) {              // StandardSurface.h:608-661
    V3 f{0.f, 0.f, 0.f};
    pdf = 0.f;
    if (!(wo.z * wi.z > 0.f)) return f;
    float pf;
    if (wc > 0.f) { V3 c = evalLobe(wo, wi, m.coatA2, V3{m.coatF0, m.coatF0, m.coatF0}, pf); f = f + c * wc; pdf += pf * wc; }
    if (wm > 0.f) { V3 c = evalLobe(wo, wi, a2, base * m.base, pf); f = f + c * wm; pdf += pf * wm; }
    if (ws > 0.f) { V3 c = evalLobe(wo, wi, a2, ld3(m.specColor) * (m.spec * m.f0d), pf); f = f + c * ws; pdf += pf * ws; }
    if (wd > 0.f) {
        float ci = fabsf(wi.z);
        f = f + base * m.base * kInvPi * wd;
        pdf += ci * kInvPi * wd;
    }
    return f;
}

__device__ __forceinline__ float matRoughness(const DScene& sc, const DMat& m, V2 uv) {
    return m.roughTex >= 0 ? texBilinear(sc.textures[m.roughTex], uv) : m.rough;
}

// Exact Fresnel reflectance of a dielectric interface, 1 under total internal reflection
// (StandardSurface.h:84-92); eta = n_transmitted / n_incident.
__device__ __forceinline__ float fresnelDielectric(float cosI, float eta) {
    float sin2I = stdmax(0.f, 1.f - cosI * cosI);
    float sin2T = sin2I / (eta * eta);
    if (sin2T >= 1.f) return 1.f;
    float cosT = sqrtf(1.f - sin2T);
    float rs = (cosI - eta * cosT) / (cosI + eta * cosT);
    float rp = (eta * cosI - cosT) / (eta * cosI + cosT);
    return 0.5f * (rs * rs + rp * rp);
}

// IMaterial::evalOpacity (StandardSurface.h:197-201; every other material is opaque).
__device__ __forceinline__ float matOpacity(const DScene& sc, const DMat& m, V2 uv) {
    if (!m.alphaMask) return 1.f;
    return m.opacityTex >= 0 ? texBilinear(sc.textures[m.opacityTex], uv) : m.opacity;
}

// IMaterial::transmittanceColor, the tint a shadow ray takes through a surface (StandardSurface.h:
// 207-225; black = opaque, every other material): 1 - opacity for an alpha mask, base colour x
// transmission x (1 - Fresnel at normal incidence) for glass.
__device__ __forceinline__ V3 matTransmittance(const DScene& sc, const DMat& m, V3 base, V2 uv) {
    if (m.type != AHIP_MAT_STANDARD) return V3{0.f, 0.f, 0.f};
    if (m.alphaMask) {
        float t = 1.f - matOpacity(sc, m, uv);
        if (t <= 0.f) return V3{0.f, 0.f, 0.f};
        return V3{t, t, t};
    }
    if (!m.trans) return V3{0.f, 0.f, 0.f};
    return base * m.trScale;
}

// The dielectric glass branch of StandardSurface::sample (StandardSurface.h:248-347): smooth glass
// (roughness < 0.001) reflects or refracts by the exact Fresnel term (a delta sample), rough glass
// samples a GGX half-vector and then reflects (Fresnel-weighted GGX) or refracts (Walter et al.
// 2007 microfacet transmission, no eta^2).  wol.z > 0; eta = IOR entering, 1 / IOR leaving.
__device__ __noinline__ BSample glassSample(const Frame& fr, V3 wol, float rough, float eta, V3 base, V2 u, float uc) {
#pragma clang fp contract(off)
    BSample s{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, 0.f, false};
    const float cosI = wol.z;
    if (rough < 0.001f) {
        const float Fr = fresnelDielectric(cosI, eta);
        s.delta = true;
        if (uc < Fr) {
            s.wi = toWorld(fr, V3{-wol.x, -wol.y, wol.z});
            s.f = base * Fr;
            s.pdf = Fr;
        } else {
            float sin2T = stdmax(0.f, 1.f - cosI * cosI) / (eta * eta);
            float cosT = sqrtf(stdmax(0.f, 1.f - sin2T));
            s.wi = toWorld(fr, V3{-wol.x / eta, -wol.y / eta, -cosT});
            s.f = base * (1.f - Fr);
            s.pdf = 1.f - Fr;
        }
        return s;
    }
    const float alpha = stdmax(1e-4f, rough * rough), a2 = alpha * alpha;
    const V3 wh = sampleHalf(u, a2);
    if (wh.z <= 0.f) return s;
    const float cosIH = stdmax(0.f, dot(wol, wh));
    if (cosIH <= 0.f) return s;
    const float Fr = fresnelDielectric(cosIH, eta);
    if (uc < Fr) {
        V3 wil = wh * (2.f * cosIH) - wol;
        if (wil.z <= 0.f) return s;
        float D = dGGX(wh.z, a2), G2 = g2Smith(wol.z, wil.z, a2);
        float k = Fr * (D * G2 / (4.f * wol.z));
        s.wi = toWorld(fr, wil);
        s.f = V3{k * base.x, k * base.y, k * base.z};
        s.pdf = Fr * pdfReflect(wh.z, a2, cosIH);
        return s;
    }
    float sin2T = stdmax(0.f, 1.f - cosIH * cosIH) / (eta * eta);
    if (sin2T >= 1.f) return s;                                              // total internal reflection
    float cosTH = sqrtf(1.f - sin2T);
    V3 wil = wol * (-1.f / eta) + wh * (cosIH / eta - cosTH);
    if (wil.z >= 0.f) return s;
    const float cosO = wol.z, cosIt = fabsf(wil.z);
    const float D = dGGX(wh.z, a2), G2 = g2Smith(cosO, cosIt, a2);
    const float absTH = fabsf(dot(wil, wh));
    const float denom = cosIH + eta * absTH;
    if (denom < 1e-6f) return s;
    const float jac = eta * eta * absTH / (denom * denom);
    s.wi = toWorld(fr, wil);
    s.f = base * ((1.f - Fr) * D * G2 * cosIH * absTH / (cosO * denom * denom));
    s.pdf = (1.f - Fr) * (D * wh.z) * jac;
    return s;
}

// The glass branches of StandardSurface::evaluate (StandardSurface.h:410-474): across the surface
// the rough transmission lobe (smooth glass has no area density there), on the same side the
// Fresnel-weighted GGX reflection alone.  f without cosine; the forward pdf.
__device__ __noinline__ V3 glassEval(const DMat& m, bool front, V3 wol, V3 wil, float rough, V3 base, float& pdf) {
#pragma clang fp contract(off)
    pdf = 0.f;
    const V3 zero{0.f, 0.f, 0.f};
    const float eta = front ? m.eta : 1.f / m.eta;
    if (!(wol.z * wil.z > 0.f)) {
        if (!m.trans || rough < 0.001f) return zero;
        float a2 = rough * rough;
        a2 *= a2;
        V3 wh = safeNormalize(wol + wil * eta);
        if (wh.z < 0.f) wh = -wh;
        const float cosIH = stdmax(0.f, dot(wol, wh)), cosTH = fabsf(dot(wil, wh));
        const float Fr = fresnelDielectric(cosIH, eta);
        const float D = dGGX(wh.z, a2), G2 = g2Smith(wol.z, fabsf(wil.z), a2);
        const float denom = cosIH + eta * cosTH;
        if (denom < 1e-6f) return zero;
        const float k = (1.f - Fr) * D * G2 * cosIH * cosTH / (wol.z * fabsf(wil.z) * denom * denom);
        const float jac = eta * eta * cosTH / (denom * denom);
        pdf = (1.f - Fr) * D * wh.z * jac;
        return base * k;
    }
    if (rough < 0.001f) return zero;
    const float alpha = stdmax(1e-4f, rough * rough), a2 = alpha * alpha;
    const V3 wh = safeNormalize(wol + wil);
    if (wh.z < 0.f) return zero;
    const float cosIH = stdmax(0.f, dot(wol, wh));
    const float Fr = fresnelDielectric(cosIH, eta);
    const float D = dGGX(wh.z, a2), G2 = g2Smith(wol.z, wil.z, a2);
    pdf = Fr > 0.f ? Fr * pdfReflect(wh.z, a2, cosIH) : 0.f;
    return base * (Fr * D * G2 / (4.f * wol.z * wil.z));
}

// What a hit contributes to its material: the base colour (the material's, or the hit
// triangle's own when the scene carries per-triangle albedo; hair keeps its own) and the
// roughness looked up at the hit's uv (evalTOV).
struct MatIn {
    V3 base;
    float r;
};
__device__ __forceinline__ MatIn matInputs(const DScene& sc, const DMat& m, V2 uv, uint32_t commit) {
    MatIn mi;
    if (sc.triAlbedo && commit != ~0u && m.type != AHIP_MAT_HAIR) {
        const float* a = sc.triAlbedo + 3u * (size_t)commit;
        mi.base = V3{a[0], a[1], a[2]};
    } else {
        mi.base = ld3(m.albedo);
    }
    mi.r = m.type == AHIP_MAT_STANDARD ? matRoughness(sc, m, uv) : 0.f;
    return mi;
}

// IMaterial::reflectance, the denoising albedo: Lambertian.h:76-78 (also EmissiveMaterial, which
// inherits it), StandardSurface.h:189-195, MarschnerHair.h:224-232.
__device__ __forceinline__ V3 matReflectance(const DMat& m, const Frame& fr, const MatIn& mi) {
#pragma clang fp contract(off)
    if (m.type == AHIP_MAT_STANDARD) {
        float diff = (1.f - m.metal) * (1.f - m.spec * 0.5f);
        float k = diff + m.metal;
        return V3{mi.base.x * m.base * k, mi.base.y * m.base * k, mi.base.z * m.base * k};
    }
    if (m.type == AHIP_MAT_HAIR) {
        V3 sa = hairSigmaA(m, fr.color);
        return V3{expf(-sa.x), expf(-sa.y), expf(-sa.z)};
    }
    return mi.base;
}

// sample below the Kajiya-Kay return (MarschnerHair.h:284-375): a lobe by the luminance of A_p,
// theta_i from a Gaussian around the tilted mirror angle, phi from the trimmed logistic around the
// lobe's specular azimuth; f and pdf are the lobe sums at the resulting direction (f * |cos|).
__device__ __noinline__ BSample hairMarschnerSample(const DMat& m, const Frame& fr, V3 wo, V2 u, float uc) {
    BSample bs{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, 0.f, false};
    float sO, cO;
    mh::longitudinal(wo, fr.t, sO, cO);
    float h = mh::impact(wo, fr.t, fr.n, sO);
    V3 sa = hairSigmaA(m, fr.color);
    V3 ap[3];
    mh::Ap(cO, m.hairEta, h, sa, ap);
    float w0 = mh::lum(ap[0]), w1 = mh::lum(ap[1]), w2 = mh::lum(ap[2]);
    float total = w0 + w1 + w2;
    if (total < 1e-8f) return bs;
    float inv = 1.f / total;
    float cdf0 = w0 * inv, cdf1 = cdf0 + w1 * inv, cdf2 = 1.f;
    int lobe;
    float ur;
    if (uc < cdf0) { lobe = 0; ur = uc / cdf0; }
    else if (uc < cdf1) { lobe = 1; ur = (uc - cdf0) / (cdf1 - cdf0); }
    else { lobe = 2; ur = (uc - cdf1) / (cdf2 - cdf1); }
    float thetaO = asinf(mh::clampH(sO));
    float z = sqrtf(-2.f * logf(fmaxf(u.x, 1e-6f))) * cosf(k2Pi * u.y);      // Box-Muller (:181-184)
    float thetaI = thetaO + m.hairAlpha[lobe] + z * sqrtf(m.hairV[lobe]);
    thetaI = fminf(fmaxf(thetaI, -kPi * 0.5f), kPi * 0.5f);
    float sI, cI;
    sincosf(thetaI, &sI, &cI);
    float gO, gT;
    mh::gammas(cO, m.hairEta, h, gO, gT);
    float phiS = mh::sampleTrimmedLogistic(ur, m.hairS) + mh::Phi(lobe, gO, gT);
    V3 woPerp = safeNormalize(wo - fr.t * sO);
    V3 ctPerp = cross(fr.t, woPerp);
    float sp, cp;
    sincosf(phiS, &sp, &cp);
    V3 wi = safeNormalize(fr.t * sI + woPerp * (cI * cp) + ctPerp * (cI * sp));
    float sIa, cIa;
    mh::longitudinal(wi, fr.t, sIa, cIa);
    float phiA = mh::azimuth(wo, wi, fr.t, sO, sIa);
    float pd;
    V3 f = mh::lobes(m, sO, cO, sIa, cIa, phiA, h, sa, pd);
    if (pd < 1e-8f) return bs;
    bs.wi = wi;
    bs.f = f * fabsf(dot(wi, fr.n));
    bs.pdf = pd;
    return bs;
}

// IMaterial::sample.  HAIRM: hair evaluates the Marschner model instead of Kajiya-Kay.  GLASS: the
// glass branch is compiled in (instances for scenes without glass leave it out).
template <bool HAIRM = false, bool GLASS = false>
__device__ BSample matSample(const DScene& sc, const DMat& m, const Frame& fr, const MatIn& mi, V3 wo, V2 u,
                             float uc) {
    BSample s{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, 0.f, false};
    if (HAIRM && m.type == AHIP_MAT_HAIR) return hairMarschnerSample(m, fr, wo, u, uc);
    if (m.type == AHIP_MAT_HAIR) {                                       // MarschnerHair.h:267-282
        V3 kt, kb;
        onb(fr.n, kt, kb);
        float kphi = 2.f * kPi * u.x, kcos = sqrtf(u.y), ksin = sqrtf(1.f - u.y);
        float sp, cp;
        sincosf(kphi, &sp, &cp);
        V3 wi = safeNormalize(kt * (ksin * cp) + kb * (ksin * sp) + fr.n * kcos);
        float pdf;
        V3 f = hairKK(m, fr, wo, wi, pdf);
        s.wi = wi;
        s.f = f * kcos;
        s.pdf = kcos * kInvPi;
        return s;
    }
    if (m.type != AHIP_MAT_STANDARD) {                                   // Lambertian.h:36-51
        float r = sqrtf(u.x), phi = 2.f * kPi * u.y;
        float sp, cp;
        sincosf(phi, &sp, &cp);
        V3 wl{r * cp, r * sp, sqrtf(stdmax(0.f, 1.f - u.x))};
        if (toLocal(fr, wo).z < 0.f) wl.z = -wl.z;
        s.wi = toWorld(fr, wl);
        float c = fabsf(wl.z);
        s.pdf = stdmax(0.f, c) * kInvPi;
        s.f = mi.base * (kInvPi * c);
        return s;
    }
    V3 wol = toLocal(fr, wo);
    if (wol.z <= 0.f) return s;
    float r = mi.r;
    if (GLASS && m.glass) return glassSample(fr, wol, r, fr.front ? m.eta : 1.f / m.eta, mi.base, u, uc);
    float alpha = stdmax(1e-4f, r * r), a2 = alpha * alpha;
    float wc, wm, ws, wd;
    if (!layerWeights(m, wol.z, wc, wm, ws, wd)) return s;
    V3 base = mi.base;
    BSample res{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, 0.f, false};
    if (uc < wc + wm + ws) {
        float a2x;
        V3 f0;
        if (uc < wc) { a2x = m.coatA2; f0 = V3{m.coatF0, m.coatF0, m.coatF0}; }
        else if (uc < wc + wm) { a2x = a2; f0 = base * m.base; }
        else { a2x = a2; f0 = ld3(m.specColor) * (m.spec * m.f0d); }
        V3 wh = sampleHalf(u, a2x);                                        // StandardSurface.h:516-549
        if (wh.z <= 0.f) return s;
        float dVH = dot(wol, wh);
        if (dVH <= 0.f) return s;
        V3 wl = wh * (2.f * dVH) - wol;
        if (wl.z <= 0.f) return s;
        V3 F = schlickColor(dVH, f0);
        float D = dGGX(wh.z, a2x), G = g2Smith(wol.z, wl.z, a2x);
        float co = fabsf(wol.z), ci = fabsf(wl.z);
        if (co < 1e-6f || ci < 1e-6f) return s;
        res.wi = toWorld(fr, wl);
        res.f = F * (D * G / (4.f * co)) * ci;
        res.pdf = pdfReflect(wh.z, a2x, dVH);
    } else {                                                               // StandardSurface.h:554-573
        float phi = 2.f * kPi * u.x, rr = sqrtf(u.y), z = sqrtf(stdmax(0.f, 1.f - u.y));
        float sp, cp;
        sincosf(phi, &sp, &cp);
        V3 wl{rr * cp, rr * sp, z};
        if (wol.z < 0.f) wl.z = -wl.z;
        float ci = fabsf(wl.z);
        res.wi = toWorld(fr, wl);
        res.f = base * m.base * (kInvPi * ci);
        res.pdf = ci * kInvPi;
    }
    if (!bsValid(res)) return s;
    V3 wil = toLocal(fr, res.wi);
    float pdf;
    V3 fc = evalMixture(m, wol, wil, wc, wm, ws, wd, a2, base, pdf);
    res.f = fc * fabsf(wil.z);
    res.pdf = pdf;
    res.delta = m.specularFlag != 0u;
    return res;
}

// IMaterial::evaluate: f without cosine, forward pdf.  HAIRM and GLASS as for matSample.
template <bool HAIRM = false, bool GLASS = false>
__device__ V3 matEval(const DScene& sc, const DMat& m, const Frame& fr, const MatIn& mi, V3 wo, V3 wi, float& pdf) {
    if (m.type == AHIP_MAT_HAIR) return HAIRM ? hairMarschner(m, fr, wo, wi, pdf) : hairKK(m, fr, wo, wi, pdf);
    V3 wil = toLocal(fr, wi), wol = toLocal(fr, wo);
    pdf = 0.f;
    if (GLASS && m.type == AHIP_MAT_STANDARD && (m.glass || m.trans)) {
        // transmission across the surface (m.trans), Fresnel-weighted reflection (m.glass)
        if (!(wil.z * wol.z > 0.f) || m.glass) return glassEval(m, fr.front, wol, wil, mi.r, mi.base, pdf);
    }
    if (!(wil.z * wol.z > 0.f)) return V3{0.f, 0.f, 0.f};
    if (m.type != AHIP_MAT_STANDARD) {
        pdf = stdmax(0.f, fabsf(wil.z)) * kInvPi;
        return mi.base * kInvPi;
    }
    float r = mi.r;
    float alpha = stdmax(1e-4f, r * r), a2 = alpha * alpha;
    float wc, wm, ws, wd;
    if (!layerWeights(m, wol.z, wc, wm, ws, wd)) return V3{0.f, 0.f, 0.f};
    return evalMixture(m, wol, wil, wc, wm, ws, wd, a2, mi.base, pdf);
}

// matEval's f for a Lambert or standard_surface hit at base colour `base` and roughness `r`, with
// its derivatives: dRho = df_c/dbase_c (the same for every channel: f is linear in the base
// colour, channel by channel, through the diffuse lobe and the metal lobe's Schlick F0) and
// dR = df/dr.  Only D*G of the metal and specular lobes depends on r:
//   dD/da2 = (den - 2 a2 c^2) / (pi den^3),  den = c^2 (a2 - 1) + 1
//   dG1/da2 = -x (1 - x^2) / (s (x + s)^2),   s = sqrt(a2 + (1 - a2) x^2)
//   a2 = alpha^2 = r^4 while r^2 > 1e-4 (alpha = max(1e-4, r^2)), so da2/dr = 4 r^3, else 0.
// The lobe arithmetic runs in double: in float, den = c^2 (a2 - 1) + 1 cancels to a few bits near
// a glint (c -> 1, a2 << 1), which makes f a staircase in r with steps of ~1e-3 relative; the
// replay estimator evaluates f here so that its image is smooth in the evaluation parameters.
// The branch decisions (which lobes apply) are the float ones of evalMixture.
__device__ V3 matEvalDiff(const DMat& m, const Frame& fr, V3 base, float r, V3 wo, V3 wi, float& dRho, V3& dR) {
    dRho = 0.f;
    dR = V3{0.f, 0.f, 0.f};
    V3 wil = toLocal(fr, wi), wol = toLocal(fr, wo);
    if (!(wil.z * wol.z > 0.f)) return V3{0.f, 0.f, 0.f};
    if (m.type != AHIP_MAT_STANDARD) {
        dRho = kInvPi;
        return base * kInvPi;
    }
    float wc, wm, ws, wd;
    if (!layerWeights(m, wol.z, wc, wm, ws, wd)) return V3{0.f, 0.f, 0.f};
    const double kPiD = 3.14159265358979323846;
    double f[3] = {0.0, 0.0, 0.0}, g[3] = {0.0, 0.0, 0.0}, rho = 0.0;
    const double b[3] = {base.x, base.y, base.z};
    const double sc[3] = {m.specColor[0], m.specColor[1], m.specColor[2]};
    if (wol.z > 0.f && wil.z > 0.f) {
        V3 wh = safeNormalize(wol + wil);
        const float co = fabsf(wol.z), ci = fabsf(wil.z);
        if (wh.z > 0.f && co >= 1e-6f && ci >= 1e-6f) {
            const double dVH = stdmax(0.f, dot(wol, wh)), x = 1.0 - dVH, x5 = x * x * x * x * x;
            const double c2 = (double)wh.z * wh.z, inv4 = 1.0 / (4.0 * (double)co * (double)ci);
            // D*G at a2 and its derivative with respect to a2
            auto dg = [&](double a2, double& dDG) {
                double den = c2 * (a2 - 1.0) + 1.0;
                double D = a2 / (kPiD * den * den), dD = (den - 2.0 * a2 * c2) / (kPiD * den * den * den);
                auto g1 = [&](double xx, double& dg1) {
                    double s = sqrt(a2 + (1.0 - a2) * xx * xx), q = xx + s;
                    dg1 = -xx * (1.0 - xx * xx) / (s * q * q);
                    return 2.0 * xx / q;
                };
                double dGo, dGi;
                double Go = g1(co, dGo), Gi = g1(ci, dGi);
                dDG = dD * Go * Gi + D * (dGo * Gi + Go * dGi);
                return D * Go * Gi;
            };
            if (wc > 0.f) {
                double unused;
                double k = wc * (m.coatF0 + (1.0 - m.coatF0) * x5) * dg(m.coatA2, unused) * inv4;
                for (int c = 0; c < 3; ++c) f[c] += k;
            }
            if (wm > 0.f || ws > 0.f) {
                const double rd = r, alpha = fmax(1e-4, rd * rd), a2 = alpha * alpha;
                const double da2 = rd * rd > 1e-4 ? 4.0 * rd * rd * rd : 0.0;
                double dDG;
                const double DG = dg(a2, dDG);
                for (int c = 0; c < 3; ++c) {
                    if (wm > 0.f) {
                        double f0 = b[c] * m.base, F = f0 + (1.0 - f0) * x5;
                        f[c] += wm * F * DG * inv4;
                        g[c] += wm * F * dDG * da2 * inv4;
                    }
                    if (ws > 0.f) {
                        double f0 = sc[c] * ((double)m.spec * m.f0d), F = f0 + (1.0 - f0) * x5;
                        f[c] += ws * F * DG * inv4;
                        g[c] += ws * F * dDG * da2 * inv4;
                    }
                }
                if (wm > 0.f) rho += wm * m.base * (1.0 - x5) * DG * inv4;
            }
        }
    }
    if (wd > 0.f) {
        for (int c = 0; c < 3; ++c) f[c] += b[c] * m.base * (1.0 / kPiD) * wd;
        rho += m.base * (1.0 / kPiD) * wd;
    }
    dRho = (float)rho;
    dR = V3{(float)g[0], (float)g[1], (float)g[2]};
    return V3{(float)f[0], (float)f[1], (float)f[2]};
}

// ---------------------------------------------------------------- dome light
// First index i of cdf[0..n] with cdf[i] >= u (the lower-bound search of Distribution1D::sample),
// through a guide table: guide[k] = that index for u = k / M (M a power of two).  The search is
// monotone in u, so for u in [k/M, (k+1)/M) its result lies in [guide[k], guide[k+1]] and the
// restricted search returns the same index in one or two steps instead of log2(n).  u * M is
// exact and u < 1.
__device__ __forceinline__ uint32_t lowerBoundGuided(const float* cdf, const uint32_t* guide, uint32_t M, float u) {
    uint32_t k = (uint32_t)(u * (float)M);
    k = k < M ? k : M - 1u;
    uint32_t lo = guide[k], count = guide[k + 1u] - lo;
    while (count > 0u) {
        uint32_t step = count >> 1;
        if (cdf[lo + step] < u) { lo += step + 1u; count -= step + 1u; }
        else count = step;
    }
    return lo;
}

// Distribution1D::sample (DomeLight.h:40-54) over func[n], cdf[n+1].
__device__ __forceinline__ uint32_t dist1dSample(const float* func, const float* cdf, const uint32_t* guide,
                                                 uint32_t M, float integral, uint32_t n, float u, float& pdf,
                                                 float& ur) {
    int idx = (int)lowerBoundGuided(cdf, guide, M, u) - 1;
    idx = idx < 0 ? 0 : (idx > (int)n - 1 ? (int)n - 1 : idx);
    float bw = cdf[idx + 1] - cdf[idx];
    ur = (bw > 1e-12f) ? (u - cdf[idx]) / bw : 0.f;
    ur = stdmax(0.f, stdmin(ur, 1.f - 1e-7f));
    pdf = (integral > 0.f) ? func[idx] / integral : 0.f;
    return (uint32_t)idx;
}

__device__ __forceinline__ V3 domeToEnv(const DDome& d, V3 w) {
    return V3{d.rot[0] * w.x + d.rot[1] * w.y + d.rot[2] * w.z, d.rot[3] * w.x + d.rot[4] * w.y + d.rot[5] * w.z,
              d.rot[6] * w.x + d.rot[7] * w.y + d.rot[8] * w.z};
}
__device__ __forceinline__ V3 domeFromEnv(const DDome& d, V3 v) {
    return V3{d.rot[0] * v.x + d.rot[3] * v.y + d.rot[6] * v.z, d.rot[1] * v.x + d.rot[4] * v.y + d.rot[7] * v.z,
              d.rot[2] * v.x + d.rot[5] * v.y + d.rot[8] * v.z};
}

__device__ __forceinline__ void domeAngles(const DDome& d, V3 w, float& theta, float& phi) {
    V3 l = domeToEnv(d, w);
    theta = acosf(stdmax(-1.f, stdmin(1.f, l.y)));
    phi = atan2f(l.x, l.z);
    if (phi < 0.f) phi += k2Pi;
}

// DomeLight::evalEnvmap (DomeLight.h:326-364), times intensity.
__device__ V3 domeEval(const DDome& d, V3 w) {
    float theta, phi;
    domeAngles(d, w, theta, phi);
    float uu = phi / k2Pi, v = theta / kPi;
    float fc = uu * (float)d.W - 0.5f, fr = v * (float)d.H - 0.5f;
    int c0 = (int)floorf(fc), r0 = (int)floorf(fr);
    float tc = fc - (float)c0, tr = fr - (float)r0;
    int W = (int)d.W, H = (int)d.H;
    uint32_t c0u = (uint32_t)((c0 % W + W) % W), c1 = (uint32_t)(((c0 + 1) % W + W) % W);
    uint32_t r0u = (uint32_t)max(0, min(r0, H - 1)), r1u = (uint32_t)max(0, min(r0 + 1, H - 1));
    const float4* tex = reinterpret_cast<const float4*>(d.px);
    float4 p00 = tex[(size_t)r0u * d.W + c0u], p01 = tex[(size_t)r0u * d.W + c1];
    float4 p10 = tex[(size_t)r1u * d.W + c0u], p11 = tex[(size_t)r1u * d.W + c1];
    V3 s;
    s.x = stdmax(0.f, (1.f - tr) * ((1.f - tc) * p00.x + tc * p01.x) + tr * ((1.f - tc) * p10.x + tc * p11.x));
    s.y = stdmax(0.f, (1.f - tr) * ((1.f - tc) * p00.y + tc * p01.y) + tr * ((1.f - tc) * p10.y + tc * p11.y));
    s.z = stdmax(0.f, (1.f - tr) * ((1.f - tc) * p00.z + tc * p01.z) + tr * ((1.f - tc) * p10.z + tc * p11.z));
    return s * d.intensity;
}

// DomeLight::evalPdf (DomeLight.h:300-321)
__device__ float domePdf(const DDome& d, V3 w) {
    float theta, phi;
    domeAngles(d, w, theta, phi);
    float uu = phi / k2Pi, v = theta / kPi;
    uint32_t col = min((uint32_t)(uu * (float)d.W), d.W - 1u);
    uint32_t row = min((uint32_t)(v * (float)d.H), d.H - 1u);
    float st = stdmax(sinf(theta), 1e-6f);
    float pr = d.margInt > 0.f ? d.margFunc[row] / d.margInt : 0.f;
    float ci = d.condInt[row];
    float pc = ci > 0.f ? d.condFunc[(size_t)row * d.W + col] / ci : 0.f;
    return (pr * pc * (float)(d.W * d.H)) / (k2Pi * kPi * st);
}

// DomeLight::sampleDirection (DomeLight.h:264-295)
__device__ V3 domeSampleDir(const DDome& d, V2 u, float& pdf) {
    float pr, pc, ur, uc;
    uint32_t row = dist1dSample(d.margFunc, d.margCdf, d.margGuide, d.margM, d.margInt, d.H, u.y, pr, ur);
    uint32_t col = dist1dSample(d.condFunc + (size_t)row * d.W, d.condCdf + (size_t)row * (d.W + 1u),
                                d.condGuide + (size_t)row * (d.condM + 1u), d.condM, d.condInt[row], d.W, u.x, pc,
                                uc);
    float v = ((float)row + ur) / (float)d.H;
    float uu = ((float)col + uc) / (float)d.W;
    float theta = v * kPi, phi = uu * k2Pi;
    float st, ct, sp, cp;
    sincosf(theta, &st, &ct);
    sincosf(phi, &sp, &cp);
    V3 dir = domeFromEnv(d, V3{st * sp, ct, st * cp});
    float sa = stdmax(st, 1e-6f);
    pdf = (pr * pc * (float)(d.W * d.H)) / (k2Pi * kPi * sa);
    if (pdf <= 0.f) pdf = 0.f;
    return dir;
}

struct LSample {
    V3 Li, wi;
    float pdf, dist;
};

// ILight::sample
__device__ LSample lightSample(const DScene& sc, const DLight& L, V3 p, V2 u) {
    LSample s{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, 0.f, 0.f};
    if (L.type == AHIP_LIGHT_DIRECTIONAL) {
        // DirectionalLight::sample (DirectionalLight.h:50-76): a delta sample of pdf 1 toward the
        // light, jittered within the cone when it has an angular radius
        s.Li = ld3(L.Le);
        s.pdf = 1.f;
        s.dist = 1e10f;
        const V3 dl = ld3(L.dir);
        if (L.cosCone >= 1.f - 1e-6f) {
            s.wi = dl;
        } else {
            const float cosT = 1.f - u.x * (1.f - L.cosCone);
            const float sinT = sqrtf(stdmax(0.f, 1.f - cosT * cosT));
            const float phi = 2.f * kPi * u.y;
            float sp, cp;
            sincosf(phi, &sp, &cp);
            V3 t, bt;
            onb(dl, t, bt);
            s.wi = safeNormalize(t * (sinT * cp) + bt * (sinT * sp) + dl * cosT);
        }
        return s;
    }
    if (L.type == AHIP_LIGHT_DOME) {
        float pdf;
        V3 wi = domeSampleDir(sc.dome, u, pdf);
        if (pdf <= 0.f) return s;
        s.Li = domeEval(sc.dome, wi);
        s.wi = wi;
        s.pdf = pdf;
        s.dist = 1e10f;
        return s;
    }
    V3 pos = ld3(L.center) + ld3(L.uA) * (2.f * u.x - 1.f) + ld3(L.vA) * (2.f * u.y - 1.f);
    V3 tl = pos - p;
    float d2 = tl.x * tl.x + tl.y * tl.y + tl.z * tl.z;
    float dist = sqrtf(d2);
    V3 wi = tl * (1.f / dist);
    float cosL = dot(-wi, ld3(L.normal));
    if (cosL <= 0.f) return s;
    s.Li = ld3(L.Le);
    s.wi = wi;
    s.pdf = d2 / (cosL * L.area);
    s.dist = dist * (1.f - 1e-4f);
    return s;
}

// ILight::pdf (0 for the delta directional light, DirectionalLight.h:78-80)
__device__ float lightPdf(const DScene& sc, const DLight& L, V3 p, V3 wi) {
    if (L.type == AHIP_LIGHT_DIRECTIONAL) return 0.f;
    if (L.type == AHIP_LIGHT_DOME) return domePdf(sc.dome, wi);
    V3 nn = -ld3(L.normal);
    float den = dot(wi, nn);
    if (den < 1e-6f) return 0.f;
    float t = dot(ld3(L.center) - p, nn) / den;
    if (t <= 0.f) return 0.f;
    V3 off = (p + wi * t) - ld3(L.center);
    float up = dot(off, ld3(L.uA)) / (L.uLen * L.uLen);
    float vp = dot(off, ld3(L.vA)) / (L.vLen * L.vLen);
    if (fabsf(up) > 1.f || fabsf(vp) > 1.f) return 0.f;
    return t * t / (den * L.area);
}

__device__ __forceinline__ V3 envLe(const DScene& sc, V3 d) {
    return sc.hasEnv ? domeEval(sc.dome, d) : ld3(sc.envRadiance);
}

}  // namespace ahip
