// oracle.cpp — CPU restatement of the reference Path+NEE hot path.
//
// ORACLE / TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and bench.py's
// cpu_baseline leg as the checker, never by the product library.
//
// A plain, single-ray restatement of the reference algorithms, each function citing the
// reference file:line it follows (paths relative to the reference checkout).  Compiled
// with -ffp-contract=off so it is a fixed IEEE program; the only fused multiply-adds are
// the explicit ones the reference's own -mfma build produces where bit-exactness depends
// on them (Halton accumulation).  Scope: triangle meshes, Lambert / Emissive / opaque
// StandardSurface materials, quad AreaLight and equirect DomeLight (constant textures plus
// a single-channel roughness texture with this build's own bilinear lookup).
//
// Scope also covers curves (the ray-space ribbon test) and the hair material in both of its
// forms: the Kajiya-Kay BSDF the reference runs, and the Marschner R/TT/TRT model it carries
// below its Kajiya-Kay returns (render flag AHIP_FLAG_HAIR_MARSCHNER).
//
// Parity pins (see tests/): PCG vs the published pcg32 known answer and vs the reference
// build; Halton, camera rays, BVH topology, closest hits and Cornell-box radiance vs the
// reference build in oracle/_ref; GGX helpers and Distribution1D vs the reference's
// header-inline functions; the StandardSurface mixture (sample / evaluate), DomeLight sampling,
// pdf and lookup, and the hair material in both forms vs the reference's own code in oracle/_ref
// (tests/golden/shading_vectors.npz); curve subdivision against the analytic Bezier.
// Parity unpinned: the roughness-texture lookup (reference uses OIIO's TextureSystem,
// src/shading/Texture.cpp:35-83, which is not available here).

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

#include "oracle.h"

namespace orc {

struct SI {                                                                           // IAccelerationStructure.h:12-30
    V3 p, n, ng, dpdu, dpdv;
    V2 uv;
    float t = 0.f;
    uint32_t meshID = ~0u, primID = ~0u, tri = ~0u;
    uint32_t strandID = ~0u;
    bool isCurve = false;
    V3 color{1.f, 1.f, 1.f};
    bool hasAlb = false;   // per-triangle base colour (ahip_scene_desc::tri_albedo)
    V3 alb;
};
static void fillSI(const Tri& tr, float t, float u, float v, SI& si) {               // BVHBackend.cpp:285-317
    float w = 1.f - u - v;
    si.t = t;
    si.p = tr.v0 + tr.e1 * u + tr.e2 * v;
    si.ng = tr.n;
    si.n = safeNormalize(tr.sn0 * w + tr.sn1 * u + tr.sn2 * v);
    si.dpdu = tr.e1;
    si.dpdv = tr.e2;
    if (dot(si.n, si.ng) < 0.f) si.n = -si.n;
    si.uv.x = tr.uv0.x * w + tr.uv1.x * u + tr.uv2.x * v;
    si.uv.y = tr.uv0.y * w + tr.uv1.y * u + tr.uv2.y * v;
    si.meshID = tr.meshID;
    si.primID = tr.primID;
}

// ---------------------------------------------------------------- shading frame (ShadingContext.h:28-59)
struct Ctx {
    V3 p, n, ng, t, bt;
    V2 uv;
    V3 color{1.f, 1.f, 1.f};
    bool front = true;
    bool hasAlb = false;
    V3 alb;
    Ctx(const SI& si, V3 rayDir) {
        p = si.p; ng = si.ng; uv = si.uv; color = si.color;
        hasAlb = si.hasAlb; alb = si.alb;
        front = dot(-rayDir, ng) > 0.f;
        n = front ? si.n : -si.n;
        ng = front ? si.ng : -si.ng;
        if (si.isCurve) {          // fiber tangent frame (ShadingContext.h:39-44)
            t = si.dpdu;
            bt = safeNormalize(cross(n, t));
        } else {
            onb(n, t, bt);
        }
    }
    V3 toLocal(V3 v) const { return {dot(v, t), dot(v, bt), dot(v, n)}; }
    V3 toWorld(V3 v) const { return t * v.x + bt * v.y + n * v.z; }
};

// ---------------------------------------------------------------- textures (this build's lookup)
struct Texture {
    uint32_t w = 0, h = 0;
    std::vector<float> px;
    // Bilinear, periodic wrap in u and v, V flipped as Texture.cpp:67-68 does before the lookup.
    float sample(V2 uv) const {
        float s = uv.x, tt = 1.f - uv.y;
        float fx = s * (float)w - 0.5f, fy = tt * (float)h - 0.5f;
        float x0f = std::floor(fx), y0f = std::floor(fy);
        float ax = fx - x0f, ay = fy - y0f;
        int x0 = (int)x0f, y0 = (int)y0f;
        auto wrap = [](int i, int n) { int r = i % n; return r < 0 ? r + n : r; };
        int xa = wrap(x0, (int)w), xb = wrap(x0 + 1, (int)w), ya = wrap(y0, (int)h), yb = wrap(y0 + 1, (int)h);
        float p00 = px[(size_t)ya * w + xa], p01 = px[(size_t)ya * w + xb];
        float p10 = px[(size_t)yb * w + xa], p11 = px[(size_t)yb * w + xb];
        return (1.f - ay) * ((1.f - ax) * p00 + ax * p01) + ay * ((1.f - ax) * p10 + ax * p11);
    }
    // The four texels sample() blends and their weights (for scattering a derivative).
    void taps(V2 uv, uint32_t idx[4], float wt[4], float& axo, float& ayo) const {
        float s = uv.x, tt = 1.f - uv.y;
        float fx = s * (float)w - 0.5f, fy = tt * (float)h - 0.5f;
        float x0f = std::floor(fx), y0f = std::floor(fy);
        float ax = fx - x0f, ay = fy - y0f;
        int x0 = (int)x0f, y0 = (int)y0f;
        auto wrap = [](int i, int n) { int r = i % n; return r < 0 ? r + n : r; };
        int xa = wrap(x0, (int)w), xb = wrap(x0 + 1, (int)w), ya = wrap(y0, (int)h), yb = wrap(y0 + 1, (int)h);
        idx[0] = ya * w + xa; idx[1] = ya * w + xb; idx[2] = yb * w + xa; idx[3] = yb * w + xb;
        wt[0] = (1.f - ay) * (1.f - ax); wt[1] = (1.f - ay) * ax; wt[2] = ay * (1.f - ax); wt[3] = ay * ax;
        axo = ax; ayo = ay;
    }
};

// Forward-mode dual number (value, derivative) in double for the roughness derivative of the
// GGX lobes.
struct Dual {
    double v, d;
};
static inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
static inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
static inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
static inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
static inline Dual dsqrt(Dual a) { double r = std::sqrt(a.v); return {r, a.d / (2.0 * r)}; }
static inline Dual dconst(double c) { return {c, 0.0}; }

// ---------------------------------------------------------------- materials
struct BSDFSample { V3 f, wi; float pdf = 0.f; bool delta = false;
                    bool valid() const { return pdf > 0.f && !isBlack(f); } };        // IMaterial.h:26-40
struct BSDFEval { V3 f; float pdf = 0.f, pdfRev = 0.f; };

static inline V3 cosineHemisphere(V2 u) {                                             // Lambertian.h:12-20
    float r = std::sqrt(u.x), phi = 2.f * kPi * u.y;
    return {r * std::cos(phi), r * std::sin(phi), std::sqrt(fmaxs(0.f, 1.f - u.x))};
}
static inline float cosPdf(float c) { return fmaxs(0.f, c) * kInvPi; }               // Lambertian.h:22-24

// GGX helpers, StandardSurface.h:20-76
static inline float D_GGX(float cosH, float a2) {
    if (cosH <= 0.f) return 0.f;
    float den = cosH * cosH * (a2 - 1.f) + 1.f;
    return a2 / (kPi * den * den);
}
static inline float G1(float absCos, float a2) {
    if (absCos <= 0.f) return 0.f;
    float c2 = absCos * absCos;
    return 2.f * absCos / (absCos + std::sqrt(a2 + (1.f - a2) * c2));
}
static inline float G2(float co, float ci, float a2) { return G1(std::abs(co), a2) * G1(std::abs(ci), a2); }
static inline V3 sampleGGXh(V2 u, float a2) {
    float phi = 2.f * kPi * u.x;
    float c2 = (1.f - u.y) / fmaxs(1.f + (a2 - 1.f) * u.y, 1e-10f);
    float ch = std::sqrt(fmaxs(0.f, c2)), sh = std::sqrt(fmaxs(0.f, 1.f - c2));
    return {sh * std::cos(phi), sh * std::sin(phi), ch};
}
static inline float pdfGGXr(float cosH, float a2, float dotVH) {
    if (cosH <= 0.f || dotVH <= 0.f) return 0.f;
    return D_GGX(cosH, a2) * cosH / (4.f * dotVH);
}
static inline float F0ior(float ior) { float r = (1.f - ior) / (1.f + ior); return r * r; }
static inline float schlickD(float c, float f0) {
    float x = 1.f - std::abs(c);
    float x5 = x * x * x * x * x;
    return f0 + (1.f - f0) * x5;
}
static inline V3 schlickC(float c, V3 f0) {
    float x = 1.f - std::abs(c);
    float x5 = x * x * x * x * x;
    return f0 + (V3{1.f, 1.f, 1.f} - f0) * x5;
}
static inline float fresnelDielectric(float cosI, float eta) {                        // StandardSurface.h:84-92
    float s2i = fmaxs(0.f, 1.f - cosI * cosI);
    float s2t = s2i / (eta * eta);
    if (s2t >= 1.f) return 1.f;
    float ct = std::sqrt(1.f - s2t);
    float rs = (cosI - eta * ct) / (cosI + eta * ct);
    float rp = (eta * cosI - ct) / (eta * cosI + ct);
    return 0.5f * (rs * rs + rp * rp);
}

// ---------------------------------------------------------------- Marschner hair (MarschnerHair.h)
// The R / TT / TRT model the reference's MarschnerHairMaterial carries but does not run (its
// evaluate and sample return Kajiya-Kay first): longitudinal M_p, azimuthal N_p and attenuation
// A_p with the PBRT-v4 curve fits, as the reference states them.  Evaluated when the render flags
// ask for AHIP_FLAG_HAIR_MARSCHNER.
namespace mh {
static float I0(float x) {                                                            // :62-71
    float sum = 0.f, x2i = 1.f, denom = 1.f;
    for (int i = 0; i < 12; ++i) {
        if (i > 0) {
            x2i *= x * x;
            denom *= (float)i * (float)i * 4.f;
        }
        sum += x2i / denom;
    }
    return sum;
}
static float logI0(float x) {                                                         // :73-77
    if (x > 12.f) return x + 0.5f * (-std::log(k2Pi) + std::log(1.f / x) + 1.f / (8.f * x));
    return std::log(I0(x));
}
static float Mp(float cosI, float sinI, float cosR, float sinR, float v) {            // :88-98
    if (v < 1e-5f) v = 1e-5f;
    float a = cosI * cosR / v, b = sinI * sinR / v;
    if (v <= 0.1f) return std::exp(logI0(a) - b - 1.f / v + 0.6931472f + std::log(0.5f / v));
    return std::exp(-b) * I0(a) / (2.f * v * std::sinh(1.f / v));
}
static float logistic(float x, float s) {                                             // :103-106
    float e = std::exp(-std::abs(x) / s);
    return e / (s * (1.f + e) * (1.f + e));
}
static float logisticCDF(float x, float s) { return 1.f / (1.f + std::exp(-x / s)); }  // :108-110
static float trimmedLogistic(float x, float s) {                                      // :113-116
    return logistic(x, s) / (logisticCDF(kPi, s) - logisticCDF(-kPi, s));
}
static float sampleTrimmedLogistic(float u, float s) {                                // :119-124
    float a = logisticCDF(-kPi, s), b = logisticCDF(kPi, s);
    float x = -s * std::log(1.f / (a + u * (b - a)) - 1.f);
    return std::clamp(x, -kPi, kPi);
}
static float Phi(int p, float gO, float gT) {                                         // :129-131
    return 2.f * (float)p * gT - 2.f * gO + (float)p * kPi;
}
static float wrapPhi(float x) {                                                       // :133-137
    while (x > kPi) x -= k2Pi;
    while (x < -kPi) x += k2Pi;
    return x;
}
static float Np(float phi, int p, float s, float gO, float gT) {                      // :140-142
    return trimmedLogistic(wrapPhi(phi - Phi(p, gO, gT)), s);
}
static float clampH(float h) { return std::clamp(h, -1.f + 1e-5f, 1.f - 1e-5f); }
// sin of the refracted azimuth gamma_t for impact parameter h at the modified index of cos(theta_o)
static float sinGammaT(float cosO, float eta, float h) {
    float etaP = std::sqrt(fmaxs(0.f, eta * eta - (1.f - cosO * cosO))) / cosO;
    return clampH(h / (etaP > 1e-5f ? etaP : 1e-5f));
}
static void Ap(float cosO, float eta, float h, V3 sa, V3 ap[3]) {                      // :146-176
    float s2 = fmaxs(0.f, 1.f - cosO * cosO);
    float etaP = std::sqrt(fmaxs(0.f, eta * eta - s2)) / cosO;
    float sgt = clampH(h / (etaP > 1e-5f ? etaP : 1e-5f));
    float cgt = std::sqrt(fmaxs(0.f, 1.f - sgt * sgt));
    V3 T{std::exp(-sa.x * 2.f * cgt), std::exp(-sa.y * 2.f * cgt), std::exp(-sa.z * 2.f * cgt)};
    float cgo = std::sqrt(fmaxs(0.f, 1.f - h * h));
    float fr = fresnelDielectric(fmaxs(0.f, cosO * cgo), eta);
    ap[0] = V3{fr, fr, fr};
    ap[1] = T * ((1.f - fr) * (1.f - fr));
    ap[2] = ap[1] * T * fr;
}
static float sampleGaussian(float u1, float u2) {                                     // :181-184
    u1 = fmaxs(u1, 1e-6f);
    return std::sqrt(-2.f * std::log(u1)) * std::cos(k2Pi * u2);
}
static void longitudinal(V3 d, V3 t, float& s, float& c) {                            // :456-460
    s = dot(d, t);
    c = std::sqrt(fmaxs(0.f, 1.f - s * s));
}
static float azimuth(V3 wo, V3 wi, V3 t, float sO, float sI) {                        // :463-475
    V3 po = wo - t * sO, pi = wi - t * sI;
    float lo = length(po), li = length(pi);
    if (lo < 1e-5f || li < 1e-5f) return 0.f;
    po = po * (1.f / lo);
    pi = pi * (1.f / li);
    float c = std::clamp(dot(po, pi), -1.f, 1.f);
    float s = dot(cross(pi, po), t);
    return std::atan2(s, c);
}
static float impact(V3 wo, V3 t, V3 n, float sO) {                                    // :478-483
    V3 po = wo - t * sO;
    float l = length(po);
    if (l < 1e-5f) return 0.f;
    return clampH(dot(po * (1.f / l), n));
}
// Per-material constants of the constructor (:200-216): variances v[p], logistic scale s and the
// cuticle tilts alpha[p] of the R, TT and TRT lobes.
struct Consts {
    float v[3], s, alpha[3];
    void init(float betaM, float betaN, float alphaDeg) {
        float bm = std::clamp(betaM, 1e-3f, 1.f);
        float v0 = 0.726f * bm + 0.812f * bm * bm + 3.7f * std::pow(bm, 20.f);
        v[0] = v0 * v0;
        v[1] = v[0] * 0.25f;
        v[2] = v[0] * 4.0f;
        float bn = std::clamp(betaN, 1e-3f, 1.f);
        s = 0.626657069f * (0.265f * bn + 1.194f * bn * bn + 5.372f * std::pow(bn, 22.f));
        alpha[0] = -alphaDeg * kPi / 180.f;
        alpha[1] = alphaDeg * kPi / 180.f * 0.5f;
        alpha[2] = -alphaDeg * kPi / 180.f * 1.5f;
    }
    // sum over p of M_p A_p N_p / max(1/4, cos^2 theta_d) (:490-522)
    V3 lobes(float eta, float sO, float cO, float sI, float cI, float phi, float h, V3 sa) const {
        float cd = std::sqrt(fmaxs(0.f, 0.5f * (1.f + cO * cI + sO * sI)));
        float denom = fmaxs(0.25f, cd * cd);
        float gO = std::asin(clampH(h)), gT = std::asin(sinGammaT(cO, eta, h));
        V3 ap[3];
        Ap(cO, eta, h, sa, ap);
        V3 f;
        for (int p = 0; p < 3; ++p) {
            float so = sO * std::cos(2.f * alpha[p]) + cO * std::sin(2.f * alpha[p]);
            float co = std::sqrt(fmaxs(0.f, 1.f - so * so));
            float m = Mp(cI, sI, co, so, v[p]);
            float nn = Np(phi, p, s, gO, gT);
            f = f + ap[p] * (m * nn);
        }
        return f * (1.f / denom);
    }
    // lobe-energy mixture of M_p cos(theta_i) N_p (:528-556)
    float pdf(float eta, float sO, float cO, float sI, float cI, float phi, float h, V3 sa) const {
        float gO = std::asin(clampH(h)), gT = std::asin(sinGammaT(cO, eta, h));
        V3 ap[3];
        Ap(cO, eta, h, sa, ap);
        float w[3] = {luminance(ap[0]), luminance(ap[1]), luminance(ap[2])};
        float total = w[0] + w[1] + w[2];
        if (total < 1e-8f) return 0.f;
        float pd = 0.f;
        for (int p = 0; p < 3; ++p) {
            float so = sO * std::cos(2.f * alpha[p]) + cO * std::sin(2.f * alpha[p]);
            float co = std::sqrt(fmaxs(0.f, 1.f - so * so));
            float m = Mp(cI, sI, co, so, v[p]) * cI;
            float nn = Np(phi, p, s, gO, gT);
            pd += (w[p] / total) * m * nn;
        }
        return fmaxs(0.f, pd);
    }
};
}  // namespace mh

struct Material {
    ahip_material m;
    const Texture* roughTex = nullptr;
    const Texture* opacityTex = nullptr;
    // derived (StandardSurface.h:152-160)
    float coatA2 = 0.f, f0d = 0.f, coatF0 = 0.f;
    bool marschner = false;   // hair: R/TT/TRT instead of the reference's runtime Kajiya-Kay
    mh::Consts hc{};
    void init() {
        float ca = fmaxs(1e-4f, m.coat_roughness * m.coat_roughness);
        coatA2 = ca * ca;
        f0d = F0ior(m.specular_ior);
        coatF0 = F0ior(1.5f);
        hc.init(m.hair_beta_m, m.hair_beta_n, m.hair_alpha);
    }
    V3 albedo() const { return {m.albedo[0], m.albedo[1], m.albedo[2]}; }
    V3 baseOf(const Ctx& c) const { return c.hasAlb ? c.alb : albedo(); }
    float rough(V2 uv) const { return roughTex ? roughTex->sample(uv) : m.roughness; }  // evalTOV
    bool isDelta() const {                                                            // StandardSurface.h:162-169
        if (m.type != AHIP_MAT_STANDARD) return false;
        float r = m.roughness, mm = m.metalness, s = m.specular;
        if (r < 0.001f && m.transmission > 0.001f && mm < 0.001f) return true;
        return r < 0.001f && (mm > 0.999f || s > 0.f);
    }
    bool flagsSpecular() const { return m.type == AHIP_MAT_STANDARD && m.roughness < 0.001f; }
    // the dielectric branch of sample and of same-side evaluate (StandardSurface.h:248, 457)
    bool glass() const { return m.type == AHIP_MAT_STANDARD && m.transmission > 0.001f && m.metalness < 0.001f; }
    // evaluate across the surface (:412-421) and transmittanceColor (:219-221) pass these tests
    bool trans() const { return m.type == AHIP_MAT_STANDARD && !(m.transmission < 0.001f) && !(m.metalness > 0.001f); }
    // IMaterial::evalOpacity (StandardSurface.h:197-201; 1 for every other material)
    float opacity(V2 uv) const {
        if (m.type != AHIP_MAT_STANDARD || !m.alpha_mask) return 1.f;
        return opacityTex ? opacityTex->sample(uv) : m.opacity;
    }
    // IMaterial::transmittanceColor (StandardSurface.h:207-225; black for every other material)
    V3 transmittance(const Ctx& c) const {
        if (m.type != AHIP_MAT_STANDARD) return {};
        if (m.alpha_mask) {
            float t = 1.f - opacity(c.uv);
            if (t <= 0.f) return {};
            return {t, t, t};
        }
        if (!trans()) return {};
        float Fr = fresnelDielectric(1.f, m.specular_ior);
        return baseOf(c) * (m.transmission * (1.f - Fr));
    }
    // StandardSurface::sample, glass branch (StandardSurface.h:248-347)
    BSDFSample glassSample(const Ctx& c, V3 wo_l, V2 u, float uc) const {
        BSDFSample s;
        V3 base = baseOf(c);
        float r = rough(c.uv);
        float eta = c.front ? m.specular_ior : (1.f / m.specular_ior);
        float cosI = wo_l.z;
        if (r < 0.001f) {                                                             // :257-288
            float Fr = fresnelDielectric(cosI, eta);
            s.delta = true;
            if (uc < Fr) {
                s.wi = c.toWorld(V3{-wo_l.x, -wo_l.y, wo_l.z});
                s.f = base * Fr;
                s.pdf = Fr;
            } else {
                float sin2T = fmaxs(0.f, 1.f - cosI * cosI) / (eta * eta);
                float cosT = std::sqrt(fmaxs(0.f, 1.f - sin2T));
                s.wi = c.toWorld(V3{-wo_l.x / eta, -wo_l.y / eta, -cosT});
                s.f = base * (1.f - Fr);
                s.pdf = 1.f - Fr;
            }
            return s;
        }
        float alpha = fmaxs(1e-4f, r * r), a2 = alpha * alpha;                        // :290-346
        V3 wh = sampleGGXh(u, a2);
        if (wh.z <= 0.f) return {};
        float cosIH = fmaxs(0.f, dot(wo_l, wh));
        if (cosIH <= 0.f) return {};
        float Fr = fresnelDielectric(cosIH, eta);
        if (uc < Fr) {
            V3 wil = wh * (2.f * cosIH) - wo_l;
            if (wil.z <= 0.f) return {};
            float D = D_GGX(wh.z, a2), G = G2(wo_l.z, wil.z, a2);
            V3 f0{Fr, Fr, Fr};
            V3 bsdf = f0 * (D * G / (4.f * wo_l.z));
            s.wi = c.toWorld(wil);
            s.f = bsdf * base;
            s.pdf = Fr * pdfGGXr(wh.z, a2, cosIH);
            return s;
        }
        float sin2T = fmaxs(0.f, 1.f - cosIH * cosIH) / (eta * eta);
        if (sin2T >= 1.f) return {};
        float cosTH = std::sqrt(1.f - sin2T);
        V3 wil = wo_l * (-1.f / eta) + wh * (cosIH / eta - cosTH);
        if (wil.z >= 0.f) return {};
        float cosO = wo_l.z, cosIt = std::abs(wil.z);
        float D = D_GGX(wh.z, a2), G = G2(cosO, cosIt, a2);
        float denom = cosIH + eta * std::abs(dot(wil, wh));
        if (denom < 1e-6f) return {};
        float absTH = std::abs(dot(wil, wh));
        float pdfWh = D * wh.z;
        float jac = eta * eta * absTH / (denom * denom);
        s.wi = c.toWorld(wil);
        s.f = base * ((1.f - Fr) * D * G * cosIH * absTH / (cosO * denom * denom));
        s.pdf = (1.f - Fr) * pdfWh * jac;
        return s;
    }
    // StandardSurface::evaluate, glass branches (StandardSurface.h:410-474)
    BSDFEval glassEval(const Ctx& c, V3 wol, V3 wil) const {
        BSDFEval e;
        float r = rough(c.uv);
        float eta = c.front ? m.specular_ior : (1.f / m.specular_ior);
        V3 base = baseOf(c);
        if (!(wol.z * wil.z > 0.f)) {                                                 // :412-444
            if (!trans() || r < 0.001f) return e;
            float a2 = r * r;
            a2 *= a2;
            V3 wh = safeNormalize(wol + wil * eta);
            if (wh.z < 0.f) wh = -wh;
            float cosIH = fmaxs(0.f, dot(wol, wh)), cosTH = std::abs(dot(wil, wh));
            float Fr = fresnelDielectric(cosIH, eta);
            float D = D_GGX(wh.z, a2), G = G2(wol.z, std::abs(wil.z), a2);
            float denom = cosIH + eta * cosTH;
            if (denom < 1e-6f) return e;
            float k = (1.f - Fr) * D * G * cosIH * cosTH / (wol.z * std::abs(wil.z) * denom * denom);
            float jac = eta * eta * cosTH / (denom * denom);
            e.f = base * k;
            e.pdf = (1.f - Fr) * D * wh.z * jac;
            e.pdfRev = e.pdf;
            return e;
        }
        if (r < 0.001f) return e;                                                     // :457-474
        float alpha = fmaxs(1e-4f, r * r), a2 = alpha * alpha;
        V3 wh = safeNormalize(wol + wil);
        if (wh.z < 0.f) return e;
        float cosIH = fmaxs(0.f, dot(wol, wh));
        float Fr = fresnelDielectric(cosIH, eta);
        float D = D_GGX(wh.z, a2), G = G2(wol.z, wil.z, a2);
        float pdf = Fr > 0.f ? Fr * pdfGGXr(wh.z, a2, cosIH) : 0.f;
        e.f = base * (Fr * D * G / (4.f * wol.z * wil.z));
        e.pdf = pdf;
        e.pdfRev = pdf;
        return e;
    }
    // IMaterial::reflectance, the first-hit denoising albedo: LambertianMaterial::m_albedo
    // (Lambertian.h:76-78, inherited by EmissiveMaterial), StandardSurface.h:189-195,
    // MarschnerHair.h:224-232 (exp(-effectiveSigmaA)).
    V3 reflectance(const Ctx& c) const;
    V3 Le() const {
        if (m.type == AHIP_MAT_EMISSIVE) return {m.emission[0], m.emission[1], m.emission[2]};  // Lambertian.h:93-95
        if (m.type == AHIP_MAT_STANDARD && m.emission_weight > 0.f)                         // StandardSurface.h:183-187
            return V3{m.emission[0], m.emission[1], m.emission[2]} * m.emission_weight;
        return {};
    }

    // layer weights, StandardSurface.h:355-371
    void weights(float woz, float metal, float spec, float& wc, float& wm, float& ws, float& wd, bool& ok) const {
        float coatF = schlickD(woz, coatF0), specF = schlickD(woz, f0d);
        wc = m.coat * coatF;
        wm = metal * (1.f - wc);
        ws = spec * (1.f - metal) * specF * (1.f - wc);
        wd = m.base * (1.f - metal) * (1.f - spec * specF) * (1.f - wc);
        float sum = wc + wm + ws + wd;
        ok = sum > 0.f;
        if (!ok) return;
        float inv = 1.f / sum;
        wc *= inv; wm *= inv; ws *= inv; wd *= inv;
    }
    // evalGGX, StandardSurface.h:578-603
    static V3 evalGGX(V3 wo, V3 wi, float a2, V3 f0, float& pf) {
        pf = 0.f;
        if (wo.z <= 0.f || wi.z <= 0.f) return {};
        V3 wh = safeNormalize(wo + wi);
        if (wh.z <= 0.f) return {};
        float dVH = fmaxs(0.f, dot(wo, wh));
        V3 F = schlickC(dVH, f0);
        float D = D_GGX(wh.z, a2), G = G2(wo.z, wi.z, a2);
        float co = std::abs(wo.z), ci = std::abs(wi.z);
        if (co < 1e-6f || ci < 1e-6f) return {};
        V3 bsdf = F * (D * G / (4.f * co * ci));
        pf = pdfGGXr(wh.z, a2, dVH);
        return bsdf;
    }
    // evalCombined, StandardSurface.h:608-661
    V3 evalCombined(V3 wo, V3 wi, float wc, float wm, float ws, float wd, V3 base, float spec, float a2,
                    float& pdf) const {
        V3 f;
        pdf = 0.f;
        if (!(wo.z * wi.z > 0.f)) return {};
        float pf;
        if (wc > 0.f) { V3 fc = evalGGX(wo, wi, coatA2, V3{coatF0, coatF0, coatF0}, pf); f = f + fc * wc; pdf += pf * wc; }
        if (wm > 0.f) { V3 fm = evalGGX(wo, wi, a2, base * m.base, pf); f = f + fm * wm; pdf += pf * wm; }
        if (ws > 0.f) {
            V3 f0 = V3{m.specular_color[0], m.specular_color[1], m.specular_color[2]} * (spec * f0d);
            V3 fs = evalGGX(wo, wi, a2, f0, pf);
            f = f + fs * ws; pdf += pf * ws;
        }
        if (wd > 0.f) {
            float ci = std::abs(wi.z);
            V3 fd = base * m.base * kInvPi;
            f = f + fd * wd; pdf += ci * kInvPi * wd;
        }
        return f;
    }

    // MarschnerHairMaterial::effectiveSigmaA (MarschnerHair.h:430-450)
    V3 hairSigmaA(const Ctx& c) const {
        V3 col = c.color;
        if (col.x > 0.98f && col.y > 0.98f && col.z > 0.98f) return {m.hair_sigma_a[0], m.hair_sigma_a[1], m.hair_sigma_a[2]};
        auto lin = [](float s) { return std::pow(fmaxs(s, 0.0f), 2.2f); };
        float lr = lin(col.x), lg = lin(col.y), lb = lin(col.z);
        if (lr < 0.01f && lg < 0.01f && lb < 0.01f) return {m.hair_sigma_a[0], m.hair_sigma_a[1], m.hair_sigma_a[2]};
        return {-std::log(fmaxs(lr, 0.001f)), -std::log(fmaxs(lg, 0.001f)), -std::log(fmaxs(lb, 0.001f))};
    }
    // Kajiya-Kay, the hair BSDF the reference evaluates at run time (MarschnerHair.h:240, 395-422)
    BSDFEval kkEval(const Ctx& c, V3 wo, V3 wi) const {
        V3 sa = hairSigmaA(c);
        V3 alb{std::exp(-sa.x), std::exp(-sa.y), std::exp(-sa.z)};
        float dTL = dot(c.t, wi), dTV = dot(c.t, wo);
        float sTL = std::sqrt(fmaxs(0.f, 1.f - dTL * dTL)), sTV = std::sqrt(fmaxs(0.f, 1.f - dTV * dTV));
        BSDFEval e;
        e.f = alb * kInvPi;
        float spec = dTL * dTV + sTL * sTV;
        if (spec > 0.f) e.f = e.f + V3{0.15f, 0.15f, 0.15f} * (std::pow(spec, 20.f) * kInvPi);
        e.pdf = fmaxs(0.f, dot(wi, c.n)) * kInvPi;
        return e;
    }

    // MarschnerHairMaterial::evaluate below its Kajiya-Kay return (MarschnerHair.h:243-257)
    BSDFEval marschnerEval(const Ctx& c, V3 wo, V3 wi) const {
        float sO, cO, sI, cI;
        mh::longitudinal(wo, c.t, sO, cO);
        mh::longitudinal(wi, c.t, sI, cI);
        float phi = mh::azimuth(wo, wi, c.t, sO, sI);
        float h = mh::impact(wo, c.t, c.n, sO);
        V3 sa = hairSigmaA(c);
        BSDFEval e;
        e.f = hc.lobes(m.hair_eta, sO, cO, sI, cI, phi, h, sa);
        e.pdf = hc.pdf(m.hair_eta, sO, cO, sI, cI, phi, h, sa);
        e.pdfRev = hc.pdf(m.hair_eta, sI, cI, sO, cO, -phi, h, sa);
        return e;
    }
    // MarschnerHairMaterial::sample below its Kajiya-Kay return (MarschnerHair.h:284-375): pick a
    // lobe by the luminance of A_p, theta_i from a Gaussian around the tilted mirror angle, phi from
    // the trimmed logistic around the lobe's specular azimuth; f and pdf are the lobe sums at the
    // resulting direction.
    BSDFSample marschnerSample(const Ctx& c, V3 wo, V2 u, float uc) const {
        BSDFSample bs;
        float sO, cO;
        mh::longitudinal(wo, c.t, sO, cO);
        float h = mh::impact(wo, c.t, c.n, sO);
        V3 sa = hairSigmaA(c);
        V3 ap[3];
        mh::Ap(cO, m.hair_eta, h, sa, ap);
        float w[3] = {luminance(ap[0]), luminance(ap[1]), luminance(ap[2])};
        float total = w[0] + w[1] + w[2];
        if (total < 1e-8f) return bs;
        float inv = 1.f / total;
        float cdf0 = w[0] * inv, cdf1 = cdf0 + w[1] * inv, cdf2 = 1.f;
        int lobe = 0;
        float ur = uc;
        if (uc < cdf0) {
            ur = uc / cdf0;
        } else if (uc < cdf1) {
            lobe = 1;
            ur = (uc - cdf0) / (cdf1 - cdf0);
        } else {
            lobe = 2;
            ur = (uc - cdf1) / (cdf2 - cdf1);
        }
        float thetaO = std::asin(mh::clampH(sO));
        float z = mh::sampleGaussian(u.x, u.y);
        float thetaI = thetaO + hc.alpha[lobe] + z * std::sqrt(hc.v[lobe]);
        thetaI = std::clamp(thetaI, -kPi * 0.5f, kPi * 0.5f);
        float sI = std::sin(thetaI), cI = std::cos(thetaI);
        float gO = std::asin(mh::clampH(h)), gT = std::asin(mh::sinGammaT(cO, m.hair_eta, h));
        float phiS = mh::sampleTrimmedLogistic(ur, hc.s) + mh::Phi(lobe, gO, gT);
        V3 woPerp = safeNormalize(wo - c.t * sO);
        if (lengthSq(woPerp) < 1e-8f) {
            V3 unused;
            onb(c.t, woPerp, unused);
        }
        V3 ctPerp = cross(c.t, woPerp);
        V3 wi = c.t * sI + woPerp * (cI * std::cos(phiS)) + ctPerp * (cI * std::sin(phiS));
        wi = safeNormalize(wi);
        float sIa, cIa;
        mh::longitudinal(wi, c.t, sIa, cIa);
        float phiA = mh::azimuth(wo, wi, c.t, sO, sIa);
        V3 f = hc.lobes(m.hair_eta, sO, cO, sIa, cIa, phiA, h, sa);
        float pd = hc.pdf(m.hair_eta, sO, cO, sIa, cIa, phiA, h, sa);
        if (pd < 1e-8f) return bs;
        bs.wi = wi;
        bs.f = f * std::abs(dot(wi, c.n));
        bs.pdf = pd;
        return bs;
    }

    BSDFSample sample(const Ctx& c, V3 wo, V2 u, float uc) const {
        BSDFSample s;
        if (m.type == AHIP_MAT_HAIR && marschner) return marschnerSample(c, wo, u, uc);
        if (m.type == AHIP_MAT_HAIR) {                                                // MarschnerHair.h:267-282
            V3 kt, kb;
            onb(c.n, kt, kb);
            float kphi = 2.f * kPi * u.x, kcos = std::sqrt(u.y), ksin = std::sqrt(1.f - u.y);
            V3 wi = safeNormalize(kt * (ksin * std::cos(kphi)) + kb * (ksin * std::sin(kphi)) + c.n * kcos);
            BSDFEval e = kkEval(c, wo, wi);
            s.wi = wi;
            s.f = e.f * kcos;
            s.pdf = kcos * kInvPi;
            return s;
        }
        if (m.type != AHIP_MAT_STANDARD) {                                            // Lambertian.h:36-51
            V3 wl = cosineHemisphere(u);
            if (c.toLocal(wo).z < 0.f) wl.z = -wl.z;
            s.wi = c.toWorld(wl);
            s.pdf = cosPdf(std::abs(wl.z));
            s.f = baseOf(c) * (kInvPi * std::abs(wl.z));
            return s;
        }
        // StandardSurface::sample, opaque branch (StandardSurface.h:231-241, 351-405)
        V3 wo_l = c.toLocal(wo);
        if (wo_l.z <= 0.f) return {};
        if (glass()) return glassSample(c, wo_l, u, uc);
        V3 base = baseOf(c);
        float metal = m.metalness, spec = m.specular, r = rough(c.uv);
        float alpha = fmaxs(1e-4f, r * r), a2 = alpha * alpha;
        float wc, wm, ws, wd;
        bool ok;
        weights(wo_l.z, metal, spec, wc, wm, ws, wd, ok);
        if (!ok) return {};
        BSDFSample res;
        auto sGGX = [&](float a2x, V3 f0) {                                           // StandardSurface.h:516-549
            BSDFSample q;
            V3 wh = sampleGGXh(u, a2x);
            if (wh.z <= 0.f) return q;
            float dVH = dot(wo_l, wh);
            if (dVH <= 0.f) return q;
            V3 wl = wh * (2.f * dVH) - wo_l;
            if (wl.z <= 0.f) return q;
            V3 F = schlickC(dVH, f0);
            float D = D_GGX(wh.z, a2x), G = G2(wo_l.z, wl.z, a2x);
            float co = std::abs(wo_l.z), ci = std::abs(wl.z);
            if (co < 1e-6f || ci < 1e-6f) return q;
            V3 bsdf = F * (D * G / (4.f * co));
            q.wi = c.toWorld(wl);
            q.f = bsdf * ci;
            q.pdf = pdfGGXr(wh.z, a2x, dVH);
            return q;
        };
        if (uc < wc) res = sGGX(coatA2, V3{coatF0, coatF0, coatF0});
        else if (uc < wc + wm) res = sGGX(a2, base * m.base);
        else if (uc < wc + wm + ws)
            res = sGGX(a2, V3{m.specular_color[0], m.specular_color[1], m.specular_color[2]} * (spec * f0d));
        else {                                                                        // StandardSurface.h:554-573
            float phi = 2.f * kPi * u.x, rr = std::sqrt(u.y), z = std::sqrt(fmaxs(0.f, 1.f - u.y));
            V3 wl{rr * std::cos(phi), rr * std::sin(phi), z};
            if (wo_l.z < 0.f) wl.z = -wl.z;
            float ci = std::abs(wl.z);
            res.wi = c.toWorld(wl);
            res.f = base * m.base * (kInvPi * ci);
            res.pdf = ci * kInvPi;
        }
        if (!res.valid()) return {};
        float pdf;
        V3 wil = c.toLocal(res.wi);
        V3 fc = evalCombined(wo_l, wil, wc, wm, ws, wd, base, spec, a2, pdf);
        res.f = fc * std::abs(c.toLocal(res.wi).z);
        res.pdf = pdf;
        res.delta = flagsSpecular();                                                  // result.flags = flags()
        return res;
    }

    BSDFEval evaluate(const Ctx& c, V3 wo, V3 wi) const {
        if (m.type == AHIP_MAT_HAIR) return marschner ? marschnerEval(c, wo, wi) : kkEval(c, wo, wi);
        BSDFEval e;
        V3 wil = c.toLocal(wi), wol = c.toLocal(wo);
        if (m.type != AHIP_MAT_STANDARD) {                                            // Lambertian.h:53-66
            if (!(wil.z * wol.z > 0.f)) return e;
            e.f = baseOf(c) * kInvPi;
            e.pdf = cosPdf(std::abs(wil.z));
            e.pdfRev = cosPdf(std::abs(wol.z));
            return e;
        }
        // StandardSurface::evaluate: glass across the surface or on the same side, else the
        // opaque branch (StandardSurface.h:410-417, 448-503)
        if ((!(wol.z * wil.z > 0.f) && trans()) || ((wol.z * wil.z > 0.f) && glass())) return glassEval(c, wol, wil);
        if (!(wol.z * wil.z > 0.f)) return e;
        V3 base = baseOf(c);
        float metal = m.metalness, spec = m.specular, r = rough(c.uv);
        float alpha = fmaxs(1e-4f, r * r), a2 = alpha * alpha;
        float wc, wm, ws, wd;
        bool ok;
        weights(wol.z, metal, spec, wc, wm, ws, wd, ok);
        if (!ok) return e;
        e.f = evalCombined(wol, wil, wc, wm, ws, wd, base, spec, a2, e.pdf);
        return e;
    }

    // f(wo, wi) (no cosine) at base colour `base` and roughness `r`, with dRho = df_c/dbase_c
    // (one value for every channel: f is linear in the base colour, channel by channel) and
    // dR = df/dr.  Lambert and standard_surface only.  The GGX lobes are evaluated in double
    // (forward-mode dual numbers carry d/dr): in float, D's denominator c^2 (a2 - 1) + 1 cancels
    // near a glint and f becomes a staircase in r.  Which lobes apply is decided as in
    // evalCombined.
    V3 evalDiff(const Ctx& c, V3 wo, V3 wi, V3 base, float r, float& dRho, V3& dR) const {
        dRho = 0.f;
        dR = {};
        V3 wil = c.toLocal(wi), wol = c.toLocal(wo);
        if (!(wil.z * wol.z > 0.f)) return {};
        if (m.type != AHIP_MAT_STANDARD) {
            dRho = kInvPi;
            return base * kInvPi;
        }
        float metal = m.metalness, spec = m.specular;
        float wc, wm, ws, wd;
        bool ok;
        weights(wol.z, metal, spec, wc, wm, ws, wd, ok);
        if (!ok) return {};
        // f as a function of (base colour b, roughness rr): dual part = d/dr
        auto eval = [&](const double b[3], Dual rr, Dual out[3]) {
            for (int k = 0; k < 3; ++k) out[k] = dconst(0.0);
            if (wol.z > 0.f && wil.z > 0.f) {
                V3 wh = safeNormalize(wol + wil);
                float co = std::abs(wol.z), ci = std::abs(wil.z);
                if (wh.z > 0.f && co >= 1e-6f && ci >= 1e-6f) {
                    double dVH = fmaxs(0.f, dot(wol, wh)), x = 1.0 - dVH, x5 = x * x * x * x * x;
                    auto DG = [&](Dual a2) {
                        Dual ch = dconst(wh.z);
                        Dual den = ch * ch * (a2 - dconst(1.0)) + dconst(1.0);
                        Dual D = a2 / (dconst(3.14159265358979323846) * den * den);
                        auto g1 = [&](double cc) {
                            Dual C = dconst(cc);
                            return dconst(2.0) * C / (C + dsqrt(a2 + (dconst(1.0) - a2) * C * C));
                        };
                        return D * g1(co) * g1(ci) / dconst(4.0 * (double)co * (double)ci);
                    };
                    if (wc > 0.f) {
                        Dual v = DG(dconst(coatA2)) * dconst(wc * (coatF0 + (1.0 - coatF0) * x5));
                        for (int k = 0; k < 3; ++k) out[k] = out[k] + v;
                    }
                    if (wm > 0.f || ws > 0.f) {
                        Dual al = (rr.v * rr.v > 1e-4) ? rr * rr : dconst(1e-4);
                        Dual dg = DG(al * al);
                        const double sc[3] = {m.specular_color[0], m.specular_color[1], m.specular_color[2]};
                        for (int k = 0; k < 3; ++k) {
                            if (wm > 0.f) {
                                double f0 = b[k] * m.base;
                                out[k] = out[k] + dg * dconst(wm * (f0 + (1.0 - f0) * x5));
                            }
                            if (ws > 0.f) {
                                double f0 = sc[k] * ((double)spec * f0d);
                                out[k] = out[k] + dg * dconst(ws * (f0 + (1.0 - f0) * x5));
                            }
                        }
                    }
                }
            }
            if (wd > 0.f)
                for (int k = 0; k < 3; ++k) out[k] = out[k] + dconst(b[k] * m.base * (1.0 / 3.14159265358979323846) * wd);
        };
        const double bb[3] = {base.x, base.y, base.z}, one[3] = {1.0, 1.0, 1.0}, zero[3] = {0.0, 0.0, 0.0};
        Dual fv[3], f1[3], f0v[3];
        eval(bb, Dual{r, 1.0}, fv);
        eval(one, dconst(r), f1);
        eval(zero, dconst(r), f0v);
        dRho = (float)(f1[0].v - f0v[0].v);
        dR = V3{(float)fv[0].d, (float)fv[1].d, (float)fv[2].d};
        return V3{(float)fv[0].v, (float)fv[1].v, (float)fv[2].v};
    }
};

// ---------------------------------------------------------------- lights
V3 Material::reflectance(const Ctx& c) const {
    if (m.type == AHIP_MAT_STANDARD) {
        float diff = (1.f - m.metalness) * (1.f - m.specular * 0.5f);
        return baseOf(c) * m.base * (diff + m.metalness);
    }
    if (m.type == AHIP_MAT_HAIR) {
        V3 sa = hairSigmaA(c);
        return {std::exp(-sa.x), std::exp(-sa.y), std::exp(-sa.z)};
    }
    return baseOf(c);
}

struct LightSample { V3 Li, wi; float pdf = 0.f, dist = 0.f; };

struct Dist1D {                                                                       // DomeLight.h:19-55
    std::vector<float> func, cdf;
    float integral = 0.f;
    uint32_t n = 0;
    void build(const float* v, uint32_t cnt) {
        n = cnt;
        func.assign(v, v + cnt);
        cdf.resize(cnt + 1);
        cdf[0] = 0.f;
        for (uint32_t i = 0; i < cnt; ++i) cdf[i + 1] = cdf[i] + func[i];
        integral = cdf[cnt];
        if (integral > 0.f) {
            float inv = 1.f / integral;
            for (auto& c : cdf) c *= inv;
        }
    }
    uint32_t sample(float u, float& pdf, float& ur) const {
        auto it = std::lower_bound(cdf.begin(), cdf.end(), u);
        int idx = (int)(it - cdf.begin()) - 1;
        idx = std::max(0, std::min(idx, (int)n - 1));
        float bw = cdf[idx + 1] - cdf[idx];
        ur = (bw > 1e-12f) ? (u - cdf[idx]) / bw : 0.f;
        ur = fmaxs(0.f, fmins(ur, 1.f - 1e-7f));
        pdf = (integral > 0.f) ? func[idx] / integral : 0.f;
        return (uint32_t)idx;
    }
};

struct Dome {                                                                         // DomeLight.h:78-388
    uint32_t W = 0, H = 0;
    std::vector<float> px;
    std::vector<Dist1D> cond;
    Dist1D marg;
    float intensity = 1.f;
    V3 rot[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

    void build() {                                                                    // :228-259
        if (W == 0 || H == 0) { W = H = 1; px = {1.f, 1.f, 1.f}; }
        std::vector<float> rowSums(H), rowF(W);
        cond.resize(H);
        for (uint32_t r = 0; r < H; ++r) {
            float theta = kPi * ((float)r + 0.5f) / (float)H;
            float st = std::sin(theta);
            for (uint32_t c = 0; c < W; ++c) {
                const float* p = &px[((size_t)r * W + c) * 3];
                float lum = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
                rowF[c] = lum * st;
            }
            cond[r].build(rowF.data(), W);
            rowSums[r] = cond[r].integral;
        }
        marg.build(rowSums.data(), H);
    }
    V3 toEnv(V3 w) const { return {dot(rot[0], w), dot(rot[1], w), dot(rot[2], w)}; }
    V3 fromEnv(V3 v) const {
        return {rot[0].x * v.x + rot[1].x * v.y + rot[2].x * v.z, rot[0].y * v.x + rot[1].y * v.y + rot[2].y * v.z,
                rot[0].z * v.x + rot[1].z * v.y + rot[2].z * v.z};
    }
    V3 sampleDir(V2 u, float& pdf) const {                                            // :264-295
        float pr, pc, ur, uc;
        uint32_t row = marg.sample(u.y, pr, ur);
        uint32_t col = cond[row].sample(u.x, pc, uc);
        float v = ((float)row + ur) / (float)H;
        float uu = ((float)col + uc) / (float)W;
        float theta = v * kPi, phi = uu * k2Pi;
        float st = std::sin(theta), ct = std::cos(theta);
        V3 e{st * std::sin(phi), ct, st * std::cos(phi)};
        V3 dir = fromEnv(e);
        float sa = fmaxs(st, 1e-6f);
        pdf = (pr * pc * (float)(W * H)) / (k2Pi * kPi * sa);
        if (pdf <= 0.f) pdf = 0.f;
        return dir;
    }
    float evalPdf(V3 w) const {                                                       // :300-321
        V3 l = toEnv(w);
        float theta = std::acos(fmaxs(-1.f, fmins(1.f, l.y)));
        float phi = std::atan2(l.x, l.z);
        if (phi < 0.f) phi += k2Pi;
        float uu = phi / k2Pi, v = theta / kPi;
        uint32_t col = std::min((uint32_t)(uu * (float)W), W - 1u);
        uint32_t row = std::min((uint32_t)(v * (float)H), H - 1u);
        float st = fmaxs(std::sin(theta), 1e-6f);
        float pr = marg.integral > 0.f ? marg.func[row] / marg.integral : 0.f;
        float pc = cond[row].integral > 0.f ? cond[row].func[col] / cond[row].integral : 0.f;
        return (pr * pc * (float)(W * H)) / (k2Pi * kPi * st);
    }
    V3 eval(V3 w) const {                                                             // :326-364
        V3 l = toEnv(w);
        float theta = std::acos(fmaxs(-1.f, fmins(1.f, l.y)));
        float phi = std::atan2(l.x, l.z);
        if (phi < 0.f) phi += k2Pi;
        float uu = phi / k2Pi, v = theta / kPi;
        float fc = uu * (float)W - 0.5f, fr = v * (float)H - 0.5f;
        int c0 = (int)std::floor(fc), r0 = (int)std::floor(fr);
        float tc = fc - (float)c0, tr = fr - (float)r0;
        auto cc = [&](int c) { return (uint32_t)((c % (int)W + (int)W) % (int)W); };
        auto cr = [&](int r) { return (uint32_t)std::max(0, std::min(r, (int)H - 1)); };
        uint32_t c1 = cc(c0 + 1), c0u = cc(c0), r1u = cr(r0 + 1), r0u = cr(r0);
        const float* p00 = &px[((size_t)r0u * W + c0u) * 3];
        const float* p01 = &px[((size_t)r0u * W + c1) * 3];
        const float* p10 = &px[((size_t)r1u * W + c0u) * 3];
        const float* p11 = &px[((size_t)r1u * W + c1) * 3];
        V3 s;
        for (int ch = 0; ch < 3; ++ch) {
            s[ch] = (1.f - tr) * ((1.f - tc) * p00[ch] + tc * p01[ch]) + tr * ((1.f - tc) * p10[ch] + tc * p11[ch]);
            s[ch] = fmaxs(0.f, s[ch]);
        }
        return s;
    }
};

struct Light {
    uint32_t type = 0;
    V3 center, uA, vA, Le, normal;
    float area = 1.f;
    const Dome* dome = nullptr;
    V3 dir;                    // directional: unit dirToLight
    float cosCone = 1.f;       // directional: cos of the angular radius
    V3 sceneC;                 // dome, directional: the scene sphere of sampleLe
    float sceneR = 0.f;
    bool isDelta() const { return type == AHIP_LIGHT_DIRECTIONAL; }                   // DirectionalLight.h:145
    void init() {                                                                     // AreaLight.h:123-131
        normal = safeNormalize(cross(uA, vA));
        area = length(cross(uA, vA)) * 4.f;
    }
    LightSample sample(V3 p, V2 u) const {
        LightSample s;
        if (type == AHIP_LIGHT_DIRECTIONAL) {                                         // DirectionalLight.h:50-76
            s.dist = 1e10f;
            s.Li = Le;
            s.pdf = 1.f;
            if (cosCone >= 1.f - 1e-6f) {
                s.wi = dir;
            } else {
                float cosT = 1.f - u.x * (1.f - cosCone);
                float sinT = std::sqrt(fmaxs(0.f, 1.f - cosT * cosT));
                float phi = 2.f * kPi * u.y;
                V3 t, bt;
                onb(dir, t, bt);
                s.wi = safeNormalize(t * (sinT * std::cos(phi)) + bt * (sinT * std::sin(phi)) + dir * cosT);
            }
            return s;
        }
        if (type == AHIP_LIGHT_DOME) {                                                // DomeLight.h:117-129
            float pdf;
            V3 wi = dome->sampleDir(u, pdf);
            if (pdf <= 0.f) return s;
            s.Li = dome->eval(wi) * dome->intensity;
            s.wi = wi; s.pdf = pdf; s.dist = 1e10f;
            return s;
        }
        V3 pos = center + uA * (2.f * u.x - 1.f) + vA * (2.f * u.y - 1.f);            // AreaLight.h:133-157
        V3 tl = pos - p;
        float d2 = lengthSq(tl), d = std::sqrt(d2);
        V3 wi = tl * (1.f / d);
        float cosL = dot(-wi, normal);
        if (cosL <= 0.f) return s;
        s.Li = Le; s.wi = wi; s.pdf = d2 / (cosL * area); s.dist = d * (1.f - 1e-4f);
        return s;
    }
    float pdf(V3 p, V3 wi) const {
        if (type == AHIP_LIGHT_DIRECTIONAL) return 0.f;                               // DirectionalLight.h:78-80
        if (type == AHIP_LIGHT_DOME) return dome->evalPdf(wi);
        float den = dot(wi, -normal);                                                 // AreaLight.h:159-180
        if (den < 1e-6f) return 0.f;
        float t = dot(center - p, -normal) / den;
        if (t <= 0.f) return 0.f;
        V3 hp = p + wi * t, off = hp - center;
        float ul = length(uA), vl = length(vA);
        float up = dot(off, uA) / (ul * ul), vp = dot(off, vA) / (vl * vl);
        if (std::abs(up) > 1.f || std::abs(vp) > 1.f) return 0.f;
        return t * t / (den * area);
    }
};

// ---------------------------------------------------------------- scene + integrator
// ---------------------------------------------------------------- curve intersection (CurveBrute.cpp:25-234)
struct RayFrame {
    V3 o, x, y, z;
    V3 project(V3 p) const {
        V3 q = p - o;
        return {dot(q, x), dot(q, y), dot(q, z)};
    }
};
static V3 bezierTangent(const V3 P[4], float t) {                                     // CurveBrute.cpp:59-64
    float mt = 1.f - t;
    return (P[1] - P[0]) * (3.f * mt * mt) + (P[2] - P[1]) * (6.f * mt * t) + (P[3] - P[2]) * (3.f * t * t);
}
static float min4(float a, float b, float c, float d) { return fmins(fmins(fmins(a, b), c), d); }
static float max4(float a, float b, float c, float d) { return fmaxs(fmaxs(fmaxs(a, b), c), d); }

// Recursive ribbon test (CurveBrute.cpp:85-176): de Casteljau midpoint splits to `depth` 0,
// then the closest point of the leaf chord to the ray in ray space against the interpolated
// half width; only strictly nearer hits are accepted.
static bool subdivide(const RayFrame& f, const V3 P[4], float hw0, float hw3, float tOff, float tScale, float tMin,
                      float& best, int depth, float& outT, V3& outTang) {
    V3 c0 = f.project(P[0]), c1 = f.project(P[1]), c2 = f.project(P[2]), c3 = f.project(P[3]);
    float mh = fmaxs(hw0, hw3);
    float zMin = min4(c0.z, c1.z, c2.z, c3.z), zMax = max4(c0.z, c1.z, c2.z, c3.z);
    if (zMin - mh > best) return false;
    if (zMax + mh < tMin) return false;
    float xMin = min4(c0.x, c1.x, c2.x, c3.x), xMax = max4(c0.x, c1.x, c2.x, c3.x);
    float yMin = min4(c0.y, c1.y, c2.y, c3.y), yMax = max4(c0.y, c1.y, c2.y, c3.y);
    float cx = (0.f < xMin) ? xMin : (0.f > xMax ? xMax : 0.f);
    float cy = (0.f < yMin) ? yMin : (0.f > yMax ? yMax : 0.f);
    if (cx * cx + cy * cy > mh * mh) return false;
    if (depth == 0) {
        float ax = c0.x, ay = c0.y, dx = c3.x - ax, dy = c3.y - ay;
        float l2 = dx * dx + dy * dy;
        float tl = l2 < 1e-12f ? 0.f : fmaxs(0.f, fmins(1.f, (-ax * dx - ay * dy) / l2));
        float px = ax + dx * tl, py = ay + dy * tl;
        float d2 = px * px + py * py;
        float hw = hw0 * (1.f - tl) + hw3 * tl;
        if (d2 > hw * hw) return false;
        float tr = c0.z + (c3.z - c0.z) * tl;
        if (tr < tMin || tr >= best) return false;
        V3 tg = bezierTangent(P, tl);
        if (lengthSq(tg) < 1e-12f) tg = P[3] - P[0];
        tg = safeNormalize(tg);
        best = tr;
        outT = tOff + tl * tScale;
        outTang = tg;
        return true;
    }
    V3 q01 = (P[0] + P[1]) * 0.5f, q12 = (P[1] + P[2]) * 0.5f, q23 = (P[2] + P[3]) * 0.5f;
    V3 r01 = (q01 + q12) * 0.5f, r12 = (q12 + q23) * 0.5f, m = (r01 + r12) * 0.5f;
    V3 L[4] = {P[0], q01, r01, m}, R[4] = {m, r12, q23, P[3]};
    float hm = (hw0 + hw3) * 0.5f, ts = tScale * 0.5f;
    bool hit = false;
    float tt;
    V3 tg;
    if (subdivide(f, L, hw0, hm, tOff, ts, tMin, best, depth - 1, tt, tg)) { outT = tt; outTang = tg; hit = true; }
    if (subdivide(f, R, hm, hw3, tOff + ts, ts, tMin, best, depth - 1, tt, tg)) { outT = tt; outTang = tg; hit = true; }
    return hit;
}

// CurveBrute::aabbHit (CurveBrute.cpp:478-498)
static bool curveBox(const CNode& n, const Ray& r, float maxT) {
    float tmin = r.tMin, tmax = maxT;
    for (int i = 0; i < 3; ++i) {
        float d = r.d[i];
        if (std::abs(d) < 1e-9f) {
            if (r.o[i] < n.bmin[i] || r.o[i] > n.bmax[i]) return false;
            continue;
        }
        float inv = 1.f / d;
        float t0 = (n.bmin[i] - r.o[i]) * inv, t1 = (n.bmax[i] - r.o[i]) * inv;
        if (inv < 0.f) std::swap(t0, t1);
        tmin = fmaxs(tmin, t0);
        tmax = fmins(tmax, t1);
        if (tmax < tmin) return false;
    }
    return true;
}

struct Scene {
    Bvh bvh;
    Curves curves;
    std::vector<Material> mats;
    std::vector<int32_t> meshMat;
    std::vector<Light> lights;
    std::vector<Texture> textures;
    Dome dome;
    bool hasEnv = false;
    V3 envRad;
    Camera cam;
    uint32_t maxDepth = 8, minDepth = 2;
    uint64_t seed = 0;
    bool fixDoubleCos = false;
    std::vector<float> triAlbedo;           // [ntris*3] commit order, or empty
    // evaluation parameters of the detached replay estimator (empty = the sampling values)
    std::vector<float> albE;
    std::vector<float> texE;
    uint32_t diffTex = 0;

    const Material* matOf(uint32_t mesh) const {
        if (mesh >= meshMat.size()) return nullptr;
        int32_t m = meshMat[mesh];
        return m >= 0 ? &mats[m] : nullptr;
    }
    // CurveBrute::trace (CurveBrute.cpp:398-440): triangles first, then the curve BVH with the
    // triangle hit as the bound; children pushed right then left, boxes tested when popped.
    bool traceSI(const Ray& r, SI& si, uint64_t* nodeCount = nullptr, uint64_t* primCount = nullptr) const {
        float t = 0.f, u = 0.f, v = 0.f;
        uint32_t h = bvh.trace(r, t, u, v, nodeCount, primCount);
        bool hit = h != ~0u;
        if (hit) {
            fillSI(bvh.tris[h], t, u, v, si);
            si.tri = h;
            if (!triAlbedo.empty()) {
                si.hasAlb = true;
                si.alb = V3{triAlbedo[3 * h], triAlbedo[3 * h + 1], triAlbedo[3 * h + 2]};
            }
        }
        if (curves.nodes.empty()) return hit;
        float best = hit ? si.t : r.tMax;
        RayFrame f;
        f.o = r.o;
        f.z = r.d;
        onb(f.z, f.x, f.y);
        uint32_t st[64];
        int top = 0;
        st[top++] = 0;
        while (top > 0) {
            const CNode& n = curves.nodes[st[--top]];
            if (nodeCount) ++*nodeCount;
            if (!curveBox(n, r, best)) continue;
            if (n.rc & 0x80000000u) {
                for (uint32_t i = n.lp; i < n.lp + (n.rc & 0x7fffffffu); ++i) {
                    const Seg& s = curves.segs[i];
                    float tSeg = 0.f;
                    V3 tang;
                    if (primCount) ++*primCount;
                    if (!subdivide(f, s.P, s.hw0, s.hw3, s.tOff, s.tScale, r.tMin, best, 6, tSeg, tang)) continue;
                    // CurveBrute.cpp:218-233
                    V3 nd = -r.d;
                    V3 nrm = safeNormalize(nd - tang * dot(nd, tang), f.x);
                    si = SI{};
                    si.p = r.o + r.d * best;
                    si.n = nrm;
                    si.ng = nrm;
                    si.dpdu = tang;
                    si.dpdv = cross(tang, nrm);
                    si.uv = V2{0.f, tSeg};
                    si.t = best;
                    si.meshID = curves.strandMesh[s.strand];
                    si.primID = s.seg;
                    si.strandID = s.strand;
                    si.isCurve = true;
                    si.color = curves.strandColor[s.strand];
                    hit = true;
                }
            } else {
                st[top++] = n.rc;
                st[top++] = n.lp;
            }
        }
        return hit;
    }
    V3 envLe(V3 d) const { return hasEnv ? dome.eval(d) * dome.intensity : envRad; }

    // shadowTransmittance (integrator/ShadowRay.h:26-61): step through up to 8 transmissive
    // surfaces, multiplying their tints; black at an opaque one.
    V3 shadowTr(Ray ray) const {
        V3 T{1.f, 1.f, 1.f};
        for (int i = 0; i < 8; ++i) {
            SI si;
            if (!traceSI(ray, si)) break;
            const Material* mat = matOf(si.meshID);
            if (!mat) return {};
            Ctx ctx(si, ray.d);
            V3 tint = mat->transmittance(ctx);
            if (isBlack(tint)) return {};
            T = T * tint;
            if (isBlack(T)) return {};
            float remaining = ray.tMax - si.t;
            if (remaining <= 1e-4f) break;
            ray = spawnRay(si.p, si.ng, ray.d);
            ray.tMax = remaining - 1e-4f;
        }
        return T;
    }

    static float power(float f, float g) { f = 1.f * f; g = 1.f * g; return (f * f) / (f * f + g * g); }  // PathIntegrator.h:55-59

    // PathIntegrator::estimateDirect (PathIntegrator.cpp:166-265)
    V3 estimateDirect(const SI& si, const Material& mat, V3 wo, const Light& light, Halton& smp) const {
        V3 Ld;
        Ctx ctx(si, -wo);
        {
            V2 u = smp.get2D();
            LightSample ls = light.sample(si.p, u);
            if (ls.pdf > 0.f && !isBlack(ls.Li)) {
                BSDFEval be = mat.evaluate(ctx, wo, ls.wi);
                if (!isBlack(be.f)) {
                    Ray sr = spawnRayTo(si.p, si.n, si.p + ls.wi * ls.dist);
                    V3 Tr = shadowTr(sr);
                    if (!isBlack(Tr)) {
                        float w = light.isDelta() ? 1.f : power(ls.pdf, be.pdf);
                        // hair: sin of the angle to the fiber (PathIntegrator.cpp:189-193)
                        float cosI = si.isCurve ? std::sqrt(fmaxs(0.f, 1.f - dot(ls.wi, ctx.t) * dot(ls.wi, ctx.t)))
                                                : absDot(ls.wi, si.n);
                        Ld = Ld + be.f * ls.Li * Tr * cosI * w / ls.pdf;
                    }
                }
            }
        }
        if (!light.isDelta()) {
            // mat.sample(ctx, wo, get2D(), get1D()): g++ evaluates the arguments right to left.
            float uc = smp.get1D();
            V2 u = smp.get2D();
            BSDFSample bs = mat.sample(ctx, wo, u, uc);
            if (bs.valid()) {
                float lp = light.pdf(si.p, bs.wi);
                if (lp > 0.f) {
                    float w = bs.delta ? 1.f : power(bs.pdf, lp);
                    Ray sr = spawnRay(si.p, si.ng, bs.wi);
                    sr.tMax = 1e10f;
                    V3 Tr{1.f, 1.f, 1.f};
                    SI hs;
                    bool hit = false;
                    for (int step = 0; step < 8; ++step) {                           // :217-236
                        hit = traceSI(sr, hs);
                        if (!hit) break;
                        const Material* m = matOf(hs.meshID);
                        if (!m) { Tr = {}; break; }
                        Ctx hctx(hs, sr.d);
                        V3 tint = m->transmittance(hctx);
                        if (isBlack(tint)) break;   // opaque / emitter
                        Tr = Tr * tint;
                        if (isBlack(Tr)) break;
                        float remaining = sr.tMax - hs.t;
                        if (remaining <= 1e-4f) { hit = false; break; }
                        Ray nr = spawnRay(hs.p, hs.ng, sr.d);
                        nr.tMax = remaining - 1e-4f;
                        sr = nr;
                    }
                    V3 Li;
                    if (!hit) Li = envLe(bs.wi);
                    else if (const Material* hm = matOf(hs.meshID)) Li = hm->Le();
                    if (!isBlack(Li) && !isBlack(Tr)) {
                        float cosI = fixDoubleCos ? 1.f : absDot(bs.wi, si.n);
                        Ld = Ld + bs.f * Li * Tr * cosI * w / bs.pdf;
                    }
                }
            }
        }
        return Ld;
    }

    // PathIntegrator::Li (PathIntegrator.cpp:71-164)
    // outAlb / outN (optional): first-hit reflectance and shading normal, left untouched when the
    // camera ray misses or hits a mesh without material (PathIntegrator.cpp:121-126).
    V3 Li(Ray r, Halton& smp, uint32_t* rays = nullptr, V3* outAlb = nullptr, V3* outN = nullptr) const {
        V3 L, beta{1.f, 1.f, 1.f};
        bool specular = false, firstHit = true;
        for (uint32_t bounce = 0; bounce <= maxDepth; ++bounce) {
            SI si;
            if (rays) ++*rays;
            if (!traceSI(r, si)) {
                if (bounce == 0 || specular) L = L + beta * envLe(r.d);
                break;
            }
            const Material* mat = matOf(si.meshID);
            if (!mat) break;
            V3 wo = -r.d;
            Ctx ctx(si, r.d);
            {   // alpha test (PathIntegrator.cpp:104-118); a sample is drawn only for partial opacity
                float op = mat->opacity(ctx.uv);
                bool pass = op <= 0.f || (op < 1.f && smp.get1D() >= op);
                if (pass) {
                    r = spawnRay(si.p, si.ng, r.d);
                    continue;
                }
            }
            if (firstHit && outAlb) {
                *outAlb = mat->reflectance(ctx);
                *outN = si.n;
            }
            firstHit = false;
            V3 Le = mat->Le();
            if (!isBlack(Le) && (bounce == 0 || specular)) L = L + beta * Le;
            if (!mat->isDelta() && !lights.empty()) {
                uint32_t li = (uint32_t)(smp.get1D() * (float)lights.size());
                li = std::min(li, (uint32_t)(lights.size() - 1));
                V3 Ld = estimateDirect(si, *mat, wo, lights[li], smp);
                L = L + beta * Ld * (float)lights.size();
            }
            float uc = smp.get1D();
            V2 u = smp.get2D();
            BSDFSample bs = mat->sample(ctx, wo, u, uc);
            if (!bs.valid()) break;
            specular = bs.delta;
            beta = beta * (bs.f / bs.pdf);
            r = spawnRay(si.p, si.n, bs.wi);
            if (bounce >= minDepth) {
                float q = 1.f - fmins(maxComponent(beta), 0.95f);
                if (smp.get1D() < q) break;
                beta = beta * (1.f / (1.f - q));
            }
        }
        return L;
    }

    // ------------------------------------------------------------ detached replay estimator
    // The bounce loop of Li with every random decision, pdf, MIS weight and Russian-roulette
    // probability taken at the sampling parameters (the scene) and every BSDF value f at the
    // evaluation parameters (albE / texE).  With `g` set it also returns the gradient of
    // sum_c delta_c * L_c with respect to the evaluation parameters, by leave-one-out throughput
    // products: for the continuation weight w_j of vertex j, dL/dw_j = sum over later emission
    // terms E_k of (prod_{i<k, i!=j} w_i) E_k, accumulated in one forward sweep.
    struct Grad {
        V3 delta;                               // dLoss/dL of this path (per channel)
        std::vector<double>* dAlb = nullptr;    // [ntris*3]
        std::vector<double>* dTex = nullptr;    // [texels]
    };
    struct EvalPt {                             // evaluation-parameter view of a hit
        bool ok = false;
        uint32_t tri = ~0u;
        V3 base;
        float r = 0.f;
        bool texR = false;
        uint32_t taps[4] = {0, 0, 0, 0};
        float tw[4] = {0.f, 0.f, 0.f, 0.f};
    };
    EvalPt evalPoint(const SI& si, const Ctx& ctx, const Material& m) const {
        EvalPt e;
        if (si.isCurve || m.m.type == AHIP_MAT_HAIR) return e;
        e.ok = true;
        e.tri = si.tri;
        e.base = m.baseOf(ctx);
        if (!albE.empty() && si.tri != ~0u)
            e.base = V3{albE[3 * si.tri], albE[3 * si.tri + 1], albE[3 * si.tri + 2]};
        e.r = m.rough(ctx.uv);
        if (m.m.type == AHIP_MAT_STANDARD && m.m.roughness_texture >= 0 && (uint32_t)m.m.roughness_texture == diffTex) {
            const Texture& T = textures[diffTex];
            e.texR = true;
            float ax, ay;
            T.taps(ctx.uv, e.taps, e.tw, ax, ay);
            if (!texE.empty()) {
                float p[4];
                for (int k = 0; k < 4; ++k) p[k] = texE[e.taps[k]];
                // the blend of Texture::sample over the evaluation texels
                e.r = (1.f - ay) * ((1.f - ax) * p[0] + ax * p[1]) + ay * ((1.f - ax) * p[2] + ax * p[3]);
            }
        }
        return e;
    }
    // One differentiable vertex j of a path: leave-one-out throughput texcl = prod_{i != j} w_i
    // over the weights so far, acc = sum of later emission terms seen through texcl (= dL/dw_j),
    // the continuation weight's derivatives (sRho = dw_c/dbase_c, gR = dw/dr) and the summed
    // coefficients of the vertex's own NEE terms (neeA per channel for the base colour, neeR per
    // channel for the roughness).
    struct Vtx {
        EvalPt e;
        V3 texcl, acc;
        float sRho = 0.f;
        V3 gR, neeA, neeR;
    };
    // d(sum_c delta_c L_c) of one vertex, scattered into g
    // A term whose float derivative overflowed (inf, or inf * 0) is dropped, as the forward drops
    // a non-finite radiance sample.
    void scatter(const Grad& g, const Vtx& v) const {
        V3 ka = g.delta * (v.acc * v.sRho + v.neeA);
        V3 kr = g.delta * (v.acc * v.gR + v.neeR);
        if (g.dAlb && v.e.tri != ~0u && isFinite(ka)) {
            (*g.dAlb)[3 * v.e.tri] += ka.x;
            (*g.dAlb)[3 * v.e.tri + 1] += ka.y;
            (*g.dAlb)[3 * v.e.tri + 2] += ka.z;
        }
        if (g.dTex && v.e.texR && isFinite(kr)) {
            double dr = (double)kr.x + (double)kr.y + (double)kr.z;
            for (int q = 0; q < 4; ++q) (*g.dTex)[v.e.taps[q]] += dr * v.e.tw[q];
        }
    }

    V3 LiReplay(Ray r, Halton& smp, const Grad* g) const {
        std::vector<Vtx> vs;
        V3 L, betaS{1.f, 1.f, 1.f}, betaE{1.f, 1.f, 1.f};
        const bool evalSet = !albE.empty() || !texE.empty();
        auto emit = [&](V3 E, float k) {        // emission term E * k (before the throughput)
            L = L + betaE * E * k;
            for (Vtx& v : vs) v.acc = v.acc + v.texcl * (E * k);
        };
        bool specular = false;
        for (uint32_t bounce = 0; bounce <= maxDepth; ++bounce) {
            SI si;
            if (!traceSI(r, si)) {
                if (bounce == 0 || specular) emit(envLe(r.d), 1.f);
                break;
            }
            const Material* mat = matOf(si.meshID);
            if (!mat) break;
            V3 wo = -r.d;
            Ctx ctx(si, r.d);
            Vtx cur;
            cur.e = evalPoint(si, ctx, *mat);
            cur.texcl = betaE;
            const EvalPt& ep = cur.e;
            V3 Le = mat->Le();
            if (!isBlack(Le) && (bounce == 0 || specular)) emit(Le, 1.f);
            if (!mat->isDelta() && !lights.empty()) {
                float nL = (float)lights.size();
                uint32_t li = (uint32_t)(smp.get1D() * nL);
                li = std::min(li, (uint32_t)(lights.size() - 1));
                const Light& light = lights[li];
                V3 Ld;
                {
                    V2 u = smp.get2D();
                    LightSample ls = light.sample(si.p, u);
                    if (ls.pdf > 0.f && !isBlack(ls.Li)) {
                        BSDFEval be = mat->evaluate(ctx, wo, ls.wi);          // sampling params: pdf
                        if (!isBlack(be.f)) {
                            Ray sr = spawnRayTo(si.p, si.n, si.p + ls.wi * ls.dist);
                            V3 Tr = shadowTr(sr);
                            if (!isBlack(Tr)) {
                                float w = power(ls.pdf, be.pdf);
                                float cosI = si.isCurve ? std::sqrt(fmaxs(0.f, 1.f - dot(ls.wi, ctx.t) * dot(ls.wi, ctx.t)))
                                                        : absDot(ls.wi, si.n);
                                V3 fe = be.f;
                                if (ep.ok) {
                                    float dr;
                                    V3 gr;
                                    V3 fd = mat->evalDiff(ctx, wo, ls.wi, ep.base, ep.r, dr, gr);
                                    if (evalSet) fe = fd;
                                    V3 c = betaE * (ls.Li * Tr * cosI * w / ls.pdf) * nL;
                                    cur.neeA = cur.neeA + c * dr;
                                    cur.neeR = cur.neeR + c * gr;
                                }
                                Ld = Ld + fe * ls.Li * Tr * cosI * w / ls.pdf;
                            }
                        }
                    }
                }
                {
                    float uc = smp.get1D();
                    V2 u = smp.get2D();
                    BSDFSample bs = mat->sample(ctx, wo, u, uc);
                    if (bs.valid()) {
                        float lp = light.pdf(si.p, bs.wi);
                        if (lp > 0.f) {
                            float w = bs.delta ? 1.f : power(bs.pdf, lp);
                            Ray sr = spawnRay(si.p, si.ng, bs.wi);
                            sr.tMax = 1e10f;
                            V3 Tr{1.f, 1.f, 1.f};
                            SI hs;
                            bool hit = traceSI(sr, hs);
                            if (hit && !matOf(hs.meshID)) Tr = {};
                            V3 Li;
                            if (!hit) Li = envLe(bs.wi);
                            else if (const Material* hm = matOf(hs.meshID)) Li = hm->Le();
                            if (!isBlack(Li) && !isBlack(Tr)) {
                                float cosI = fixDoubleCos ? 1.f : absDot(bs.wi, si.n);
                                V3 fe = bs.f;      // includes |cos| at the sampling parameters
                                if (ep.ok) {
                                    float cw = std::abs(ctx.toLocal(bs.wi).z);
                                    float dr;
                                    V3 gr;
                                    V3 fd = mat->evalDiff(ctx, wo, bs.wi, ep.base, ep.r, dr, gr) * cw;
                                    if (evalSet) fe = fd;
                                    V3 c = betaE * (Li * Tr * cosI * w / bs.pdf) * nL;
                                    cur.neeA = cur.neeA + c * (dr * cw);
                                    cur.neeR = cur.neeR + c * (gr * cw);
                                }
                                Ld = Ld + fe * Li * Tr * cosI * w / bs.pdf;
                            }
                        }
                    }
                }
                emit(Ld, nL);
            }
            float uc = smp.get1D();
            V2 u = smp.get2D();
            BSDFSample bs = mat->sample(ctx, wo, u, uc);
            if (!bs.valid()) {
                if (ep.ok) vs.push_back(cur);
                break;
            }
            specular = bs.delta;
            V3 wS = bs.f / bs.pdf, wE = wS;
            if (ep.ok) {
                float cw = std::abs(ctx.toLocal(bs.wi).z);
                float dr;
                V3 gr;
                V3 fe = mat->evalDiff(ctx, wo, bs.wi, ep.base, ep.r, dr, gr);
                if (evalSet) wE = fe * cw / bs.pdf;
                cur.sRho = dr * cw / bs.pdf;
                cur.gR = gr * (cw / bs.pdf);
            }
            for (Vtx& o : vs) o.texcl = o.texcl * wE;
            if (ep.ok) vs.push_back(cur);
            betaS = betaS * wS;
            betaE = betaE * wE;
            r = spawnRay(si.p, si.n, bs.wi);
            if (bounce >= minDepth) {
                float q = 1.f - fmins(maxComponent(betaS), 0.95f);
                if (smp.get1D() < q) break;
                float k = 1.f / (1.f - q);
                betaS = betaS * k;
                betaE = betaE * k;
                for (Vtx& o : vs) o.texcl = o.texcl * k;
            }
        }
        if (g && isFinite(L))
            for (const Vtx& v : vs) scatter(*g, v);
        return L;
    }

    // PathIntegrator::renderTile (PathIntegrator.cpp:8-69): out[w*h*5] = r,g,b (pre-weighted), w, lumSq
    // With albedo / normal [w*h*3] set, also the tile's AOVs (PathIntegrator.cpp:28-31, 40-47,
    // 58-66): mean first-hit reflectance, mean first-hit normal scaled to unit length.
    void renderTile(uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t s0, uint32_t ns, float* out,
                    float* albedo = nullptr, float* normal = nullptr) const {
        Halton smp;
        smp.seed = seed;
        for (uint32_t ty = 0; ty < h; ++ty)
            for (uint32_t tx = 0; tx < w; ++tx) {
                uint32_t px = x0 + tx, py = y0 + ty;
                V3 acc, accA, accN;
                uint32_t aovCount = 0;
                float lsq = 0.f;
                for (uint32_t s = 0; s < ns; ++s) {
                    smp.start(px, py, s0 + s);
                    V2 jit = smp.get2D();
                    smp.get2D();   // lens
                    smp.get1D();   // time
                    Ray ray = generateRay(cam, px, py, jit.x, jit.y);
                    V3 a, nrm;
                    V3 L = Li(ray, smp, nullptr, albedo ? &a : nullptr, &nrm);
                    if (isFinite(L)) {
                        acc = acc + L;
                        float lum = luminance(L);
                        lsq += lum * lum;
                    }
                    accA = accA + a;
                    accN = accN + nrm;
                    ++aovCount;
                }
                if (albedo && aovCount > 0) {
                    float invN = 1.f / (float)aovCount;
                    V3 ma = accA * invN, mn = accN * invN;
                    float len = length(mn);
                    if (len > 1e-6f) mn = mn * (1.f / len);
                    float* A = albedo + 3 * (ty * w + tx);
                    float* N = normal + 3 * (ty * w + tx);
                    A[0] = ma.x; A[1] = ma.y; A[2] = ma.z;
                    N[0] = mn.x; N[1] = mn.y; N[2] = mn.z;
                }
                float inv = 1.f / (float)ns, wt = (float)ns;
                float* o = out + 5 * (ty * w + tx);
                o[0] = acc.x * inv * wt;
                o[1] = acc.y * inv * wt;
                o[2] = acc.z * inv * wt;
                o[3] = wt;
                o[4] = lsq;
            }
    }

    // ------------------------------------------------------------ BDPT (BDPTIntegrator.cpp, MISWeight.cpp)
    // Scenes of triangles with Lambert / Emissive / StandardSurface materials without a roughness texture, lit by quad
    // area lights, the dome and directional lights: the scope the device implements.
    struct PVert {                                                                    // PathVertex.h:28-60
        V3 p, n, wo, beta, Le;
        float pdfFwd = 0.f, pdfRev = 0.f, minRough = 1.f;
        uint32_t flags = 0;
        const Material* mat = nullptr;
        const Light* light = nullptr;
    };
    static constexpr uint32_t kVCamera = 0, kVSurface = 1, kVLight = 2, kVDelta = 4, kVInf = 8;
    struct PPath {
        std::vector<PVert> v;
        uint32_t count = 0;
        float lightPdf = 1.f;
        uint32_t type(uint32_t i) const { return v[i].flags & 3u; }
        bool delta(uint32_t i) const { return (v[i].flags & kVDelta) != 0u; }
        bool full() const { return count >= v.size(); }
    };
    // LightSampler::build / sample / pdf (LightSampler.h:28-91): Vose's alias table over power
    struct AliasTable {
        std::vector<float> prob, w;
        std::vector<uint32_t> alias;
        void build(const std::vector<Light>& L) {
            uint32_t n = (uint32_t)L.size();
            if (n == 0) return;
            std::vector<float> wt(n);
            float total = 0.f;
            for (uint32_t i = 0; i < n; ++i) {
                float pw;
                if (L[i].type == AHIP_LIGHT_DOME) {                                   // DomeLight.h:176-179, 255-257
                    const Dome& D = *L[i].dome;
                    float envIntegral = D.marg.integral * k2Pi * kPi / (float)(D.W * D.H);
                    pw = envIntegral * 4.f * kPi * D.intensity;
                } else if (L[i].type == AHIP_LIGHT_DIRECTIONAL) {                     // DirectionalLight.h:41, 136-139
                    float diskArea = kPi * L[i].sceneR * L[i].sceneR;
                    pw = luminance(L[i].Le) * diskArea * kPi;
                } else {
                    pw = luminance(L[i].Le) * L[i].area * kPi;                         // AreaLight.h:114-117
                }
                wt[i] = fmaxs(0.f, pw);
                total += wt[i];
            }
            if (total <= 0.f) {
                for (auto& x : wt) x = 1.f;
                total = (float)n;
            }
            float inv = 1.f / total;
            for (auto& x : wt) x *= inv;
            prob.resize(n);
            alias.assign(n, 0u);
            std::vector<uint32_t> small, large;
            for (uint32_t i = 0; i < n; ++i) {
                prob[i] = wt[i] * (float)n;
                (prob[i] < 1.f ? small : large).push_back(i);
            }
            while (!small.empty() && !large.empty()) {
                uint32_t s = small.back(); small.pop_back();
                uint32_t l = large.back(); large.pop_back();
                alias[s] = l;
                prob[l] = prob[l] + prob[s] - 1.f;
                (prob[l] < 1.f ? small : large).push_back(l);
            }
            for (uint32_t i : small) prob[i] = 1.f;
            for (uint32_t i : large) prob[i] = 1.f;
            w = wt;
        }
        uint32_t sample(float u) const {
            uint32_t n = (uint32_t)prob.size();
            uint32_t i = std::min((uint32_t)(u * (float)n), n - 1u);
            float uf = u * (float)n - (float)i;
            return uf < prob[i] ? i : alias[i];
        }
    };
    static float toArea(float pdfSA, V3 prev, V3 next, V3 nextN) {                 // PathVertex.h:147-156
        V3 d = next - prev;
        float d2 = lengthSq(d);
        if (d2 < 1e-10f) return 0.f;
        return pdfSA * std::abs(dot(safeNormalize(d), nextN)) / d2;
    }
    // BDPTIntegrator::evalVertex (BDPTIntegrator.cpp:40-63)
    static V3 evalVertex(const PPath& P, uint32_t i, V3 toNext, float& fwd, float& rev) {
        fwd = rev = 0.f;
        if (P.type(i) != kVSurface || !P.v[i].mat) return {};
        SI si;
        si.p = P.v[i].p;
        si.n = si.ng = P.v[i].n;
        Ctx ctx(si, -P.v[i].wo);
        BSDFEval be = P.v[i].mat->evaluate(ctx, P.v[i].wo, toNext);
        fwd = be.pdf;
        rev = be.pdfRev;
        return be.f * std::abs(dot(toNext, P.v[i].n));
    }
    // the shared bounce loop of traceCameraSubpath / traceLightSubpath (BDPTIntegrator.cpp:97-190,
    // 240-300); `light` selects the light subpath's differences (no emission recorded, no
    // environment vertex)
    void walk(Ray ray, V3 beta, float pdfFwd, Halton& smp, PPath& P, bool light) const {
        for (uint32_t depth = 0; depth < maxDepth && !P.full(); ++depth) {
            SI si;
            if (!traceSI(ray, si)) {
                // the camera subpath's environment vertex (BDPTIntegrator.cpp:100-118)
                if (!light && hasEnv && !P.full()) {
                    PVert& V = P.v[P.count++];
                    V = PVert{};
                    V.p = ray.o + ray.d * 1e6f;
                    V.n = -ray.d;
                    V.wo = -ray.d;
                    V.beta = beta;
                    V.Le = dome.eval(ray.d) * dome.intensity;
                    V.flags = kVLight | kVInf;
                }
                break;
            }
            const Material* mat = matOf(si.meshID);
            if (!mat) break;
            V3 wo = -ray.d;
            Ctx ctx(si, ray.d);
            if (!light) {
                // the camera subpath's alpha test (BDPTIntegrator.cpp:130-143): a pass-through is
                // not a bounce; a sample is drawn only for a partial opacity
                float op = mat->opacity(ctx.uv);
                if (op <= 0.f || (op < 1.f && smp.get1D() >= op)) {
                    ray = spawnRay(si.p, si.ng, ray.d);
                    --depth;
                    continue;
                }
            }
            float pdfFwdArea = toArea(pdfFwd, P.v[P.count - 1].p, si.p, si.n);
            uint32_t vi = P.count++;
            PVert& V = P.v[vi];
            V = PVert{};
            V.p = si.p;
            V.n = si.n;
            V.wo = wo;
            V.beta = beta;
            V.Le = light ? V3{} : mat->Le();
            V.pdfFwd = pdfFwdArea;
            V.flags = kVSurface | (mat->isDelta() ? kVDelta : 0u);
            V.mat = mat;
            // IMaterial::roughness() (IMaterial.h:112: 1; StandardSurface.h:171: the roughness value),
            // a delta vertex counting as 1 (BDPTIntegrator.cpp:163-172)
            const float rough = mat->m.type == AHIP_MAT_STANDARD ? mat->m.roughness : 1.f;
            V.minRough = fmins(vi > 0 ? P.v[vi - 1].minRough : 1.f, mat->isDelta() ? 1.f : rough);
            if (vi > 0) {
                float dummy, revSA;
                evalVertex(P, vi, -wo, dummy, revSA);
                P.v[vi - 1].pdfRev = toArea(revSA, si.p, P.v[vi - 1].p, P.v[vi - 1].n);
            }
            if (depth >= 2) {
                float q = 1.f - fmins(maxComponent(beta), 0.95f);
                if (smp.get1D() < q) break;
                beta = beta * (1.f / (1.f - q));
            }
            // mat->sample(ctx, wo, get2D(), get1D()): g++ evaluates the arguments right to left
            float uc = smp.get1D();
            V2 u = smp.get2D();
            BSDFSample bs = mat->sample(ctx, wo, u, uc);
            if (mat->isDelta() && !mat->glass()) ++smooth;   // smooth glass samples the exact Fresnel
            if (!bs.valid()) break;
            beta = beta * (bs.f / bs.pdf);
            pdfFwd = bs.pdf;
            ray = spawnRay(si.p, si.n, bs.wi);
        }
    }
    // traceCameraSubpath (BDPTIntegrator.cpp:70-194) with the pinhole's area pdf 1
    void cameraSubpath(const Ray& primary, Halton& smp, PPath& P) const {
        P.count = 0;
        PVert& V = P.v[P.count++];
        V = PVert{};
        V.p = primary.o;
        V.n = -primary.d;
        V.wo = -primary.d;
        V.beta = {1.f, 1.f, 1.f};
        V.pdfFwd = 1.f;
        V.flags = kVCamera;
        walk(primary, {1.f, 1.f, 1.f}, 1.f, smp, P, false);
    }
    // ILight::sampleLe: AreaLight.h:81-105, DomeLight.h:142-168, DirectionalLight.h:88-125.  pdfPos 0:
    // no sample.
    struct LeSmp {
        V3 Le, pos, n, dir;
        float pdfPos = 0.f, pdfDir = 0.f;
    };
    static LeSmp sampleLe(const Light& L, V2 uPos, V2 uDir) {
        LeSmp s;
        if (L.type == AHIP_LIGHT_DOME) {
            float pdfDir;
            V3 envDir = L.dome->sampleDir(uDir, pdfDir);
            if (pdfDir <= 0.f) return s;
            V3 t, bt;
            onb(-envDir, t, bt);
            float r = L.sceneR * std::sqrt(uPos.x);
            float phi = k2Pi * uPos.y;
            V3 dc = L.sceneC + envDir * L.sceneR;
            s.pos = dc + t * (r * std::cos(phi)) + bt * (r * std::sin(phi));
            float diskArea = kPi * L.sceneR * L.sceneR;
            s.Le = L.dome->eval(envDir) * L.dome->intensity;
            s.n = -envDir;
            s.dir = -envDir;
            s.pdfPos = 1.f / diskArea;
            s.pdfDir = pdfDir;
            return s;
        }
        if (L.type == AHIP_LIGHT_DIRECTIONAL) {
            float r = L.sceneR * std::sqrt(uPos.x);
            float phi = 2.f * kPi * uPos.y;
            V3 t, bt;
            onb(L.dir, t, bt);
            V3 dc = L.sceneC + L.dir * (L.sceneR * 100.f);
            s.pos = dc + t * (r * std::cos(phi)) + bt * (r * std::sin(phi));
            V3 dir;
            if (L.cosCone >= 1.f - 1e-6f) {
                dir = -L.dir;
            } else {
                float cosT = 1.f - uDir.x * (1.f - L.cosCone);
                float sinT = std::sqrt(fmaxs(0.f, 1.f - cosT * cosT));
                float p = 2.f * kPi * uDir.y;
                V3 nd = -L.dir, dt, dbt;
                onb(nd, dt, dbt);
                dir = safeNormalize(dt * (sinT * std::cos(p)) + dbt * (sinT * std::sin(p)) + nd * cosT);
            }
            float diskArea = kPi * L.sceneR * L.sceneR;
            s.Le = L.Le;
            s.n = dir;
            s.dir = dir;
            s.pdfPos = 1.f / diskArea;
            s.pdfDir = 1.f;
            return s;
        }
        s.pos = L.center + L.uA * (2.f * uPos.x - 1.f) + L.vA * (2.f * uPos.y - 1.f);
        float phi = 2.f * kPi * uDir.x;
        float cosT = std::sqrt(uDir.y), sinT = std::sqrt(1.f - uDir.y);
        V3 t, bt;
        onb(L.normal, t, bt);
        s.dir = t * (sinT * std::cos(phi)) + bt * (sinT * std::sin(phi)) + L.normal * cosT;
        s.n = L.normal;
        s.Le = L.Le;
        s.pdfPos = 1.f / L.area;
        s.pdfDir = cosT * kInvPi;
        return s;
    }
    // traceLightSubpath (BDPTIntegrator.cpp:199-303)
    void lightSubpath(uint32_t li, float lpdf, Halton& smp, PPath& P) const {
        P.count = 0;
        P.lightPdf = lpdf;
        if (lpdf <= 0.f) return;
        const Light& L = lights[li];
        // light->sampleLe(get2D(), get2D()): g++ evaluates the arguments right to left
        V2 uDir = smp.get2D();
        V2 uPos = smp.get2D();
        LeSmp ls = sampleLe(L, uPos, uDir);
        if (ls.pdfPos <= 0.f || ls.pdfDir <= 0.f) return;
        V3 beta = ls.Le * std::abs(dot(ls.dir, ls.n)) / (lpdf * ls.pdfPos * ls.pdfDir);
        PVert& V = P.v[P.count++];
        V = PVert{};
        V.p = ls.pos;
        V.n = ls.n;
        V.wo = ls.dir;
        V.beta = ls.Le / (lpdf * ls.pdfPos);
        V.Le = ls.Le;
        V.pdfFwd = lpdf * ls.pdfPos;
        V.flags = kVLight | (L.isDelta() ? kVDelta : 0u) | (L.type != AHIP_LIGHT_AREA ? kVInf : 0u);
        V.light = &L;
        Ray ray{ls.pos, ls.dir};
        ray.tMin = 1e-4f;
        walk(ray, beta, ls.pdfDir, smp, P, true);
    }
    // bdptMISWeight (MISWeight.cpp:14-72), in double as there
    static float misWeight(const PPath& lp, const PPath& cp, uint32_t s, uint32_t t) {
        if (s + t == 2) return 1.f;
        double sum = 0.0, ri = 1.0;
        for (int i = (int)s - 1; i >= 0; --i) {
            float num = i < (int)lp.count ? lp.v[i].pdfRev : 0.f;
            float den = i < (int)lp.count ? lp.v[i].pdfFwd : 0.f;
            if (den == 0.f) break;
            ri *= (double)num / (double)den;
            bool prevDelta = i > 0 && lp.delta(i - 1);
            bool curDelta = i < (int)lp.count && lp.delta(i);
            if (!prevDelta && !curDelta) sum += ri * ri;
        }
        ri = 1.0;
        for (int i = (int)t - 1; i >= 0; --i) {
            float num = i < (int)cp.count ? cp.v[i].pdfRev : 0.f;
            float den = i < (int)cp.count ? cp.v[i].pdfFwd : 0.f;
            if (den == 0.f) break;
            ri *= (double)num / (double)den;
            bool prevDelta = i > 0 && cp.delta(i - 1);
            bool curDelta = i < (int)cp.count && cp.delta(i);
            if (!prevDelta && !curDelta) sum += ri * ri;
        }
        return (float)(1.0 / (sum + 1.0));
    }
    // shadowTransmittance of a connection (opaque scenes: visible or not).  A connection to a light
    // point ends 1e-4 short of a surface that is itself geometry (the emitter's own mesh), so its
    // verdict can hang on the last bits of the hit distance; `border` counts the verdicts that change
    // between tMax * (1 - 4e-7) and tMax * (1 + 4e-7), where builds that round differently can
    // disagree.
    mutable uint32_t border = 0;
    // shadowTransmittance of a connection (ShadowRay.h:26-61), flagging it (border) when the verdict
    // changes between tMax (1 -/+ 4e-7)
    V3 connTr(const Ray& r) const {
        V3 T = shadowTr(r);
        Ray a = r, b = r;
        a.tMax = r.tMax * (1.f - 4e-7f);
        b.tMax = r.tMax * (1.f + 4e-7f);
        if (isBlack(shadowTr(a)) != isBlack(shadowTr(b))) ++border;
        return T;
    }
    // A smooth StandardSurface (roughness < 0.001: isDelta) is sampled as GGX at alpha = 1e-4,
    // alpha^2 = 1e-8, where alpha^2 - 1 rounds to -1 in float: the sampled half-vector is the normal
    // itself, and evalCombined's D_GGX of the half-vector rebuilt from (wo, wi) is infinite when
    // that vector's z rounds to exactly 1 (cos^2 * (alpha^2 - 1) + 1 = 0) and finite when it rounds
    // to 1 - 2^-24.  An infinite D makes f / pdf = inf / inf = NaN, and the rest of the subpath
    // then adds nothing (the non-finite contributions are dropped).  Which way the last bit falls
    // depends on how the shading-frame transforms of wo and wi were rounded, so builds that
    // contract differently disagree on those samples; `smooth` counts such sampled vertices.
    mutable uint32_t smooth = 0;
    // BDPTIntegrator::connect (BDPTIntegrator.cpp:308-459) without film splats (t = 1 is unused there)
    V3 connect(const PPath& lp, uint32_t s, const PPath& cp, uint32_t t) const {
        if (t == 0) return {};
        if (s == 0) {
            if (t < 2) return {};
            uint32_t last = t - 1;
            if (cp.type(last) == kVSurface) return cp.v[last].beta * cp.v[last].Le;
            // light-type camera vertices are environment vertices, counted right after the camera
            if ((cp.v[last].flags & kVInf) && t == 2) return cp.v[last].beta * cp.v[last].Le;
            return {};
        }
        if (s == 1 && (lp.v[0].flags & kVInf)) {
            // the dome or a directional light: along the light subpath's own direction
            // (BDPTIntegrator.cpp:357-365, 395-402)
            uint32_t ct = t - 1;
            if (cp.delta(ct) || !cp.v[ct].mat) return {};
            const Light* light = lp.v[0].light;
            if (!light) return {};
            V3 wi = safeNormalize(-lp.v[0].wo);
            Ray sr = spawnRayTo(cp.v[ct].p, cp.v[ct].n, cp.v[ct].p + wi * 1e10f);
            const V3 Tr1 = connTr(sr);
            if (isBlack(Tr1)) return {};
            SI si;
            si.p = cp.v[ct].p;
            si.n = si.ng = cp.v[ct].n;
            Ctx ctx(si, -cp.v[ct].wo);
            BSDFEval be = cp.v[ct].mat->evaluate(ctx, cp.v[ct].wo, wi);
            if (isBlack(be.f)) return {};
            V3 Le;
            float pdfDir;
            if (light->type == AHIP_LIGHT_DOME) {                                          // DomeLight.h:131-133, 171-173
                Le = light->dome->eval(wi) * light->dome->intensity;
                pdfDir = light->dome->evalPdf(wi);
            } else {                                                                   // DirectionalLight.h:78-80, 131-134
                Le = dot(wi, light->dir) > light->cosCone - 1e-4f ? light->Le : V3{};
                pdfDir = 0.f;
            }
            if (isBlack(Le)) return {};
            float cosI = std::abs(dot(wi, cp.v[ct].n));
            if (pdfDir <= 0.f) pdfDir = 1.f;
            return cp.v[ct].beta * be.f * cosI * Le * Tr1 / (lp.lightPdf * pdfDir);
        }
        if (s == 1) {
            uint32_t ct = t - 1;
            if (cp.delta(ct) || !cp.v[ct].mat) return {};
            const Light* light = lp.v[0].light;
            if (!light) return {};
            V3 toL = lp.v[0].p - cp.v[ct].p;
            float dist = length(toL);
            if (dist < 1e-6f) return {};
            V3 wi = toL * (1.f / dist);
            Ray sr = spawnRayTo(cp.v[ct].p, cp.v[ct].n, cp.v[ct].p + wi * dist);
            const V3 Tr1 = connTr(sr);
            if (isBlack(Tr1)) return {};
            SI si;
            si.p = cp.v[ct].p;
            si.n = si.ng = cp.v[ct].n;
            Ctx ctx(si, -cp.v[ct].wo);
            BSDFEval be = cp.v[ct].mat->evaluate(ctx, cp.v[ct].wo, wi);
            if (isBlack(be.f)) return {};
            V3 Le = dot(wi, lp.v[0].n) < 0.f ? light->Le : V3{};                     // AreaLight.h:110-112
            if (isBlack(Le)) return {};
            float cosI = std::abs(dot(wi, cp.v[ct].n));
            float cosL = std::abs(dot(-wi, lp.v[0].n));
            float d2 = dist * dist;
            return cp.v[ct].beta * be.f * cosI * Le * cosL * Tr1 / (d2 * lp.v[0].pdfFwd);
        }
        if (t == 1) return {};
        uint32_t ls = s - 1, ct = t - 1;
        if (lp.delta(ls) || cp.delta(ct)) return {};
        if (!lp.v[ls].mat || !cp.v[ct].mat) return {};
        if (lp.v[ls].minRough < 0.08f || cp.v[ct].minRough < 0.08f) return {};
        Ray sr = spawnRayTo(lp.v[ls].p, lp.v[ls].n, cp.v[ct].p);
        const V3 Tr2 = connTr(sr);
        if (isBlack(Tr2)) return {};
        V3 d = cp.v[ct].p - lp.v[ls].p;
        float dist = length(d);
        if (dist < 1e-6f) return {};
        V3 wi = d * (1.f / dist);
        float fL1, rL1, fC1, rC1;
        V3 fL = evalVertex(lp, ls, wi, fL1, rL1);
        if (isBlack(fL)) return {};
        V3 fC = evalVertex(cp, ct, -wi, fC1, rC1);
        if (isBlack(fC)) return {};
        float invD2 = 1.f / (dist * dist);
        return lp.v[ls].beta * fL * invD2 * fC * cp.v[ct].beta * Tr2;
    }
    // BDPTIntegrator::renderTile (BDPTIntegrator.cpp:464-581): out[w*h*5] as renderTile above
    // border (optional, [w*h*ns]): bit 0 where a connection's visibility of the sample hangs on the
    // last bits of its hit distance (connTr), bit 1 where a subpath sampled a smooth
    // StandardSurface (`smooth`).
    void renderTileBDPT(uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t s0, uint32_t ns, float clamp,
                        float* out, float* albedo = nullptr, float* normal = nullptr,
                        uint8_t* borderOut = nullptr) const {
        AliasTable ls;
        ls.build(lights);
        Halton smp;
        smp.seed = seed;
        PPath cp, lp;
        cp.v.resize(maxDepth + 2);
        lp.v.resize(maxDepth + 2);
        for (uint32_t ty = 0; ty < h; ++ty)
            for (uint32_t tx = 0; tx < w; ++tx) {
                uint32_t px = x0 + tx, py = y0 + ty;
                V3 pixL, accA, accN;
                uint32_t aovCount = 0;
                float lsq = 0.f;
                for (uint32_t s = 0; s < ns; ++s) {
                    smp.start(px, py, s0 + s);
                    V2 jit = smp.get2D();
                    smp.get2D();   // lens
                    smp.get1D();   // time
                    Ray primary = generateRay(cam, px, py, jit.x, jit.y);
                    border = smooth = 0;
                    cameraSubpath(primary, smp, cp);
                    if (cp.count >= 2 && cp.type(1) == kVSurface && cp.v[1].mat) {
                        SI si;
                        si.p = cp.v[1].p;
                        si.n = si.ng = cp.v[1].n;
                        Ctx ctx(si, -cp.v[1].wo);
                        accA = accA + cp.v[1].mat->reflectance(ctx);
                        accN = accN + cp.v[1].n;
                        ++aovCount;
                    }
                    uint32_t li = 0;
                    if (!lights.empty()) li = ls.sample(smp.get1D());
                    lightSubpath(li, lights.empty() ? 0.f : ls.w[li], smp, lp);
                    V3 sL;
                    for (uint32_t t = 1; t <= cp.count; ++t)
                        for (uint32_t s2 = 0; s2 <= lp.count; ++s2) {
                            if (s2 + t < 2) continue;
                            V3 C = connect(lp, s2, cp, t);
                            if (isBlack(C) || !isFinite(C)) continue;
                            float wgt = misWeight(lp, cp, s2, t);
                            V3 c = C * wgt;
                            if (!isFinite(c)) continue;
                            if (clamp > 0.f && s2 >= 2 && t >= 2) {
                                float lum = luminance(c);
                                if (lum > clamp) c = c * (clamp / lum);
                            }
                            sL = sL + c;
                        }
                    pixL = pixL + sL;
                    float lum = luminance(sL);
                    lsq += lum * lum;
                    if (borderOut) borderOut[(ty * w + tx) * ns + s] = (border ? 1u : 0u) | (smooth ? 2u : 0u);
                }
                float inv = 1.f / (float)ns, wt = (float)ns;
                float* o = out + 5 * (ty * w + tx);
                o[0] = pixL.x * inv * wt;
                o[1] = pixL.y * inv * wt;
                o[2] = pixL.z * inv * wt;
                o[3] = wt;
                o[4] = lsq;
                if (albedo) {
                    V3 ma, mn;
                    if (aovCount > 0) {
                        float invN = 1.f / (float)aovCount;
                        ma = accA * invN;
                        mn = accN * invN;
                        float len = length(mn);
                        if (len > 1e-6f) mn = mn * (1.f / len);
                    }
                    float* A = albedo + 3 * (ty * w + tx);
                    float* N = normal + 3 * (ty * w + tx);
                    A[0] = ma.x; A[1] = ma.y; A[2] = ma.z;
                    N[0] = mn.x; N[1] = mn.y; N[2] = mn.z;
                }
            }
    }
};

}  // namespace orc

using namespace orc;

extern "C" {

int orc_pcg_u32(uint64_t seed, uint64_t seq, uint32_t n, uint32_t* out) {
    Pcg r(seed, seq);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next();
    return 0;
}

float orc_radical_inverse(uint32_t base, uint64_t n) { return radicalInverse(base, n); }

int orc_halton(uint64_t seed, uint32_t n, const uint32_t* px, const uint32_t* py, const uint32_t* idx,
               const uint32_t* dim, float* out) {
    Halton h;
    h.seed = seed;
    for (uint32_t i = 0; i < n; ++i) {
        h.start(px[i], py[i], idx[i]);
        if (dim[i] < 128) out[i] = h.scrambled(dim[i], idx[i]);
        else {
            float v = 0.f;
            for (uint32_t k = 128; k <= dim[i]; ++k) v = h.rng.nextFloat();
            out[i] = v;
        }
    }
    return 0;
}

int orc_halton_seq(uint64_t seed, uint32_t px, uint32_t py, uint32_t idx, uint32_t nd, float* out) {
    Halton h;
    h.seed = seed;
    h.start(px, py, idx);
    for (uint32_t d = 0; d < nd; ++d) out[d] = h.get1D();
    return 0;
}

int orc_make_pinhole(const float* f, const float* a, const float* u, float vfov, uint32_t w, uint32_t h,
                     ahip_camera* out) {
    Camera c = makePinhole(V3{f[0], f[1], f[2]}, V3{a[0], a[1], a[2]}, V3{u[0], u[1], u[2]}, vfov, w, h);
    for (int i = 0; i < 3; ++i) {
        out->origin[i] = c.origin[i]; out->lower_left[i] = c.llc[i];
        out->horizontal[i] = c.hor[i]; out->vertical[i] = c.ver[i];
    }
    out->width = w; out->height = h;
    return 0;
}

static Camera toCam(const ahip_camera* c) {
    Camera k;
    k.origin = {c->origin[0], c->origin[1], c->origin[2]};
    k.llc = {c->lower_left[0], c->lower_left[1], c->lower_left[2]};
    k.hor = {c->horizontal[0], c->horizontal[1], c->horizontal[2]};
    k.ver = {c->vertical[0], c->vertical[1], c->vertical[2]};
    k.w = c->width; k.h = c->height;
    return k;
}

int orc_generate_rays(const ahip_camera* c, uint32_t n, const uint32_t* px, const uint32_t* py, const float* jx,
                      const float* jy, float* rays) {
    Camera k = toCam(c);
    for (uint32_t i = 0; i < n; ++i) {
        Ray r = generateRay(k, px[i], py[i], jx[i], jy[i]);
        for (int j = 0; j < 3; ++j) { rays[6 * i + j] = r.o[j]; rays[6 * i + 3 + j] = r.d[j]; }
    }
    return 0;
}

void* orc_scene_create(const ahip_scene_desc* d) {
    auto s = std::make_unique<Scene>();
    s->bvh.commit(d);
    if (d->num_strands) {
        if (!d->strand_cv_offset || !d->curve_cvs || !(d->strand_width || d->strand_widths) || !d->strand_mesh_index ||
            (d->strand_widths && !d->strand_width_offset))
            return nullptr;
        s->curves.build(d);
    }
    for (uint32_t i = 0; i < d->num_textures; ++i) {
        Texture t;
        t.w = d->texture_width[i]; t.h = d->texture_height[i];
        t.px.assign(d->texture_pixels[i], d->texture_pixels[i] + (size_t)t.w * t.h);
        s->textures.push_back(std::move(t));
    }
    for (uint32_t i = 0; i < d->num_materials; ++i) {
        Material m;
        m.m = d->materials[i];
        if (m.m.roughness_texture >= 0) {
            if ((uint32_t)m.m.roughness_texture >= d->num_textures) return nullptr;
            m.roughTex = &s->textures[m.m.roughness_texture];
        }
        if (m.m.type == AHIP_MAT_STANDARD && m.m.alpha_mask && m.m.opacity_texture >= 0) {
            if ((uint32_t)m.m.opacity_texture >= d->num_textures) return nullptr;
            m.opacityTex = &s->textures[m.m.opacity_texture];
        }
        m.init();
        s->mats.push_back(m);
    }
    s->meshMat.assign(d->mesh_material, d->mesh_material + d->num_meshes);
    if (d->tri_albedo) s->triAlbedo.assign(d->tri_albedo, d->tri_albedo + 3 * (size_t)s->bvh.tris.size());
    if (d->env_light >= 0) {
        const ahip_light& l = d->lights[d->env_light];
        s->dome.W = d->env_width; s->dome.H = d->env_height;
        s->dome.px.assign(d->env_pixels, d->env_pixels + (size_t)d->env_width * d->env_height * 3);
        s->dome.intensity = l.intensity;
        for (int r = 0; r < 3; ++r) s->dome.rot[r] = V3{l.rot[3 * r], l.rot[3 * r + 1], l.rot[3 * r + 2]};
        s->dome.build();
        s->hasEnv = true;
    }
    s->envRad = V3{d->env_radiance[0], d->env_radiance[1], d->env_radiance[2]};
    for (uint32_t i = 0; i < d->num_lights; ++i) {
        const ahip_light& l = d->lights[i];
        Light L;
        L.type = l.type;
        if (l.type == AHIP_LIGHT_AREA) {
            L.center = {l.center[0], l.center[1], l.center[2]};
            L.uA = {l.u_axis[0], l.u_axis[1], l.u_axis[2]};
            L.vA = {l.v_axis[0], l.v_axis[1], l.v_axis[2]};
            L.Le = {l.Le[0], l.Le[1], l.Le[2]};
            L.init();
        } else if (l.type == AHIP_LIGHT_DIRECTIONAL) {                                // DirectionalLight.h:33-45
            L.dir = safeNormalize(V3{l.direction[0], l.direction[1], l.direction[2]});
            L.Le = {l.Le[0], l.Le[1], l.Le[2]};
            L.cosCone = std::cos(l.half_angle * kPi / 180.f);
        } else {
            if ((int32_t)i != d->env_light) return nullptr;
            L.dome = &s->dome;
        }
        if (l.type != AHIP_LIGHT_AREA) {
            L.sceneC = V3{l.scene_center[0], l.scene_center[1], l.scene_center[2]};
            L.sceneR = l.scene_radius;
        }
        s->lights.push_back(L);
    }
    return s.release();
}

void orc_scene_destroy(void* h) { delete static_cast<Scene*>(h); }

int orc_set_render(void* h, const ahip_camera* cam, uint32_t maxDepth, uint32_t minDepth, uint64_t seed,
                   uint32_t flags) {
    auto* s = static_cast<Scene*>(h);
    if (cam) s->cam = toCam(cam);
    s->maxDepth = maxDepth; s->minDepth = minDepth; s->seed = seed;
    s->fixDoubleCos = (flags & AHIP_FLAG_FIX_DOUBLE_COSINE) != 0;
    for (Material& m : s->mats) m.marschner = (flags & AHIP_FLAG_HAIR_MARSCHNER) != 0;
    return 0;
}

int orc_bvh_sizes(void* h, uint32_t* nn, uint32_t* np) {
    auto* s = static_cast<Scene*>(h);
    *nn = (uint32_t)s->bvh.nodes.size();
    *np = (uint32_t)s->bvh.prims.size();
    return 0;
}
int orc_bvh_dump(void* h, uint32_t* nodes32, uint32_t* prims) {
    auto* s = static_cast<Scene*>(h);
    std::memcpy(nodes32, s->bvh.nodes.data(), s->bvh.nodes.size() * sizeof(Node));
    std::memcpy(prims, s->bvh.prims.data(), s->bvh.prims.size() * 4);
    return 0;
}

// rays[n*8]; hits as ahip_hit; optional per-ray node / triangle visit counts.
int orc_trace(void* h, const float* rays, uint32_t n, ahip_hit* hits, uint32_t* nodeCount, uint32_t* triCount) {
    auto* s = static_cast<Scene*>(h);
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = rays + 8 * i;
        Ray r;
        r.o = {q[0], q[1], q[2]}; r.d = {q[3], q[4], q[5]}; r.tMin = q[6]; r.tMax = q[7];
        float t = 0.f, u = 0.f, v = 0.f;
        uint64_t nv = 0, tv = 0;
        uint32_t k = s->bvh.trace(r, t, u, v, &nv, &tv);
        ahip_hit& o = hits[i];
        if (!s->curves.nodes.empty()) {
            SI si;
            nv = tv = 0;
            bool hit = s->traceSI(r, si, &nv, &tv);
            if (hit && si.isCurve) {
                o = {si.t, 0.f, si.uv.y, si.meshID, si.primID, 0x80000000u | si.strandID};
                if (nodeCount) nodeCount[i] = (uint32_t)nv;
                if (triCount) triCount[i] = (uint32_t)tv;
                continue;
            }
        }
        if (nodeCount) nodeCount[i] = (uint32_t)nv;
        if (triCount) triCount[i] = (uint32_t)tv;
        if (k == ~0u) { o = {0.f, 0.f, 0.f, ~0u, ~0u, ~0u}; continue; }
        o = {t, u, v, s->bvh.tris[k].meshID, s->bvh.tris[k].primID, k};
    }
    return 0;
}

// Curves only, without the curve BVH: every segment in leaf order through the same ribbon test
// (tree-independence check of the curve BVH).  out[n*4] = t (or 0), strand v, strand, segment
// (bits; ~0u on a miss).
int orc_curve_brute(void* h, const float* rays, uint32_t n, float* out) {
    auto* s = static_cast<Scene*>(h);
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = rays + 8 * i;
        Ray r;
        r.o = {q[0], q[1], q[2]}; r.d = {q[3], q[4], q[5]}; r.tMin = q[6]; r.tMax = q[7];
        RayFrame f;
        f.o = r.o; f.z = r.d;
        onb(f.z, f.x, f.y);
        float best = r.tMax, tSeg = 0.f;
        V3 tang;
        uint32_t hs = ~0u, hg = ~0u;
        for (const Seg& sg : s->curves.segs) {
            float tt;
            V3 tg;
            if (subdivide(f, sg.P, sg.hw0, sg.hw3, sg.tOff, sg.tScale, r.tMin, best, 6, tt, tg)) {
                tSeg = tt; tang = tg; hs = sg.strand; hg = sg.seg;
            }
        }
        out[4 * i] = hs == ~0u ? 0.f : best;
        out[4 * i + 1] = tSeg;
        std::memcpy(&out[4 * i + 2], &hs, 4);
        std::memcpy(&out[4 * i + 3], &hg, 4);
    }
    return 0;
}

int orc_occluded(void* h, const float* rays, uint32_t n, uint8_t* out) {
    auto* s = static_cast<Scene*>(h);
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = rays + 8 * i;
        Ray r;
        r.o = {q[0], q[1], q[2]}; r.d = {q[3], q[4], q[5]}; r.tMin = q[6]; r.tMax = q[7];
        if (s->curves.nodes.empty()) {
            out[i] = s->bvh.occluded(r) ? 1 : 0;
        } else {
            SI si;   // shadowTransmittance decides by a closest-hit trace (ShadowRay.h:26-61)
            out[i] = s->traceSI(r, si) ? 1 : 0;
        }
    }
    return 0;
}

int orc_render_tile_aovs(void* h, uint32_t x0, uint32_t y0, uint32_t w, uint32_t hh, uint32_t s0, uint32_t ns,
                         float* out, float* albedo, float* normal) {
    static_cast<Scene*>(h)->renderTile(x0, y0, w, hh, s0, ns, out, albedo, normal);
    return 0;
}

// BDPTIntegrator::renderTile (BDPTIntegrator.cpp:464-581) with firefly clamp `clamp` (0 = off):
// out[w*h*5] as orc_render_tile; albedo / normal [w*h*3] and the per-sample flags border [w*h*ns]
// (bit 0: Scene::connTr's borderline visibility, bit 1: Scene::smooth) may be null.  Scenes of triangles with
// Lambert / Emissive / StandardSurface materials (no roughness texture) lit by area lights, the dome and directional
// lights with a scene sphere only (returns 1 otherwise).
int orc_render_tile_bdpt(void* h, uint32_t x0, uint32_t y0, uint32_t w, uint32_t hh, uint32_t s0, uint32_t ns,
                         float clamp, float* out, float* albedo, float* normal, uint8_t* border) {
    const Scene* s = static_cast<Scene*>(h);
    if (!s->curves.nodes.empty() || !isBlack(s->envRad)) return 1;
    for (const Light& L : s->lights)
        if (L.type != AHIP_LIGHT_AREA && !(L.sceneR > 0.f)) return 1;
    if (!s->triAlbedo.empty()) return 1;
    for (const Material& m : s->mats)
        if (m.m.type != AHIP_MAT_LAMBERT && m.m.type != AHIP_MAT_EMISSIVE &&
            !(m.m.type == AHIP_MAT_STANDARD && !m.roughTex))
            return 1;
    s->renderTileBDPT(x0, y0, w, hh, s0, ns, clamp, out, albedo, normal, border);
    return 0;
}

int orc_render_tile(void* h, uint32_t x0, uint32_t y0, uint32_t w, uint32_t hh, uint32_t s0, uint32_t ns, float* out) {
    static_cast<Scene*>(h)->renderTile(x0, y0, w, hh, s0, ns, out);
    return 0;
}

// Evaluation parameters of the detached replay estimator (NULL = the sampling values).
int orc_set_eval(void* h, const float* triAlbedo, uint32_t texture, const float* texels) {
    auto* s = static_cast<Scene*>(h);
    s->diffTex = texture;
    if (triAlbedo) s->albE.assign(triAlbedo, triAlbedo + 3 * s->bvh.tris.size());
    else s->albE.clear();
    if (texels && texture < s->textures.size())
        s->texE.assign(texels, texels + (size_t)s->textures[texture].w * s->textures[texture].h);
    else s->texE.clear();
    return 0;
}

// Image of the detached replay estimator over a tile, laid out like orc_render_tile.
int orc_replay_tile(void* h, uint32_t x0, uint32_t y0, uint32_t w, uint32_t hh, uint32_t s0, uint32_t ns, float* out) {
    auto* s = static_cast<Scene*>(h);
    Halton smp;
    smp.seed = s->seed;
    for (uint32_t ty = 0; ty < hh; ++ty)
        for (uint32_t tx = 0; tx < w; ++tx) {
            uint32_t px = x0 + tx, py = y0 + ty;
            V3 acc;
            float lsq = 0.f;
            for (uint32_t k = 0; k < ns; ++k) {
                smp.start(px, py, s0 + k);
                V2 jit = smp.get2D();
                smp.get2D();
                smp.get1D();
                V3 L = s->LiReplay(generateRay(s->cam, px, py, jit.x, jit.y), smp, nullptr);
                if (isFinite(L)) {
                    acc = acc + L;
                    float lum = luminance(L);
                    lsq += lum * lum;
                }
            }
            float* o = out + 5 * (ty * w + tx);
            o[0] = acc.x; o[1] = acc.y; o[2] = acc.z; o[3] = (float)ns; o[4] = lsq;
        }
    return 0;
}

// Gradient of sum(dLdI * I), I = per-pixel mean of the detached estimator over samples
// [s0, s0+ns) of the whole film, w.r.t. the evaluation parameters: dAlb[ntris*3], dTex[texels]
// (either may be NULL), accumulated in double and returned as float.
int orc_adjoint(void* h, uint32_t s0, uint32_t ns, const float* dLdI, float* dAlb, float* dTex) {
    auto* s = static_cast<Scene*>(h);
    std::vector<double> ga(dAlb ? 3 * s->bvh.tris.size() : 0), gt;
    if (dTex && s->diffTex < s->textures.size()) gt.resize((size_t)s->textures[s->diffTex].w * s->textures[s->diffTex].h);
    Scene::Grad g;
    g.dAlb = dAlb ? &ga : nullptr;
    g.dTex = gt.empty() ? nullptr : &gt;
    Halton smp;
    smp.seed = s->seed;
    const uint32_t W = s->cam.w, H = s->cam.h;
    for (uint32_t py = 0; py < H; ++py)
        for (uint32_t px = 0; px < W; ++px) {
            const float* d = dLdI + 3 * ((size_t)py * W + px);
            g.delta = V3{d[0], d[1], d[2]} * (1.f / (float)ns);
            for (uint32_t k = 0; k < ns; ++k) {
                smp.start(px, py, s0 + k);
                V2 jit = smp.get2D();
                smp.get2D();
                smp.get1D();
                s->LiReplay(generateRay(s->cam, px, py, jit.x, jit.y), smp, &g);
            }
        }
    for (size_t i = 0; i < ga.size(); ++i) dAlb[i] = (float)ga[i];
    if (dTex)
        for (size_t i = 0; i < gt.size(); ++i) dTex[i] = (float)gt[i];
    return 0;
}

// Per-sample radiance for a list of (px, py, sampleIndex): out[n*3]; non-finite kept as is.
int orc_render_samples(void* h, uint32_t n, const uint32_t* px, const uint32_t* py, const uint32_t* idx, float* out,
                       uint32_t* raysOut) {
    auto* s = static_cast<Scene*>(h);
    Halton smp;
    smp.seed = s->seed;
    for (uint32_t i = 0; i < n; ++i) {
        smp.start(px[i], py[i], idx[i]);
        V2 jit = smp.get2D();
        smp.get2D();
        smp.get1D();
        Ray ray = generateRay(s->cam, px[i], py[i], jit.x, jit.y);
        uint32_t rc = 0;
        V3 L = s->Li(ray, smp, &rc);
        out[3 * i] = L.x; out[3 * i + 1] = L.y; out[3 * i + 2] = L.z;
        if (raysOut) raysOut[i] = rc;
    }
    return 0;
}

// Multi-threaded whole-frame render over 64x64 tiles (the RenderSession tile loop,
// RenderSession.cpp:511-527) for rows [row0, row1): out[W*H*5]. *seconds = wall time.
int orc_render_frame(void* h, uint32_t row0, uint32_t row1, uint32_t spp, uint32_t nthreads, float* out,
                     double* seconds) {
    auto* s = static_cast<Scene*>(h);
    const uint32_t W = s->cam.w, H = s->cam.h, ts = 64;
    row1 = std::min(row1, H);
    std::vector<uint32_t> tx, ty;
    for (uint32_t y = row0; y < row1; y += ts)
        for (uint32_t x = 0; x < W; x += ts) { tx.push_back(x); ty.push_back(y); }
    if (nthreads == 0) nthreads = std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint32_t> next{0};
    auto t0 = std::chrono::steady_clock::now();
    auto work = [&] {
        std::vector<float> buf(ts * ts * 5);
        for (;;) {
            uint32_t i = next.fetch_add(1);
            if (i >= tx.size()) break;
            uint32_t w = std::min(ts, W - tx[i]), hh = std::min(ts, row1 - ty[i]);
            s->renderTile(tx[i], ty[i], w, hh, 0, spp, buf.data());
            for (uint32_t yy = 0; yy < hh; ++yy)
                std::memcpy(out + 5 * ((size_t)(ty[i] + yy) * W + tx[i]), buf.data() + 5 * yy * w, sizeof(float) * 5 * w);
        }
    };
    std::vector<std::thread> th;
    for (uint32_t i = 0; i < nthreads; ++i) th.emplace_back(work);
    for (auto& t : th) t.join();
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return 0;
}

// Material sample / evaluate probes on a flat +Z frame for the invariant tests.
// in[n*6]: wo (world), u.x, u.y, uc; out[n*9]: wi, f, pdf, valid, evalPdf.
int orc_material_probe(const ahip_material* mat, uint32_t n, const float* in, float* out) {
    Material m;
    m.m = *mat;
    m.init();
    SI si;
    si.n = {0, 0, 1}; si.ng = {0, 0, 1};
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 6 * i;
        V3 wo{q[0], q[1], q[2]};
        Ctx c(si, -wo);
        BSDFSample bs = m.sample(c, wo, V2{q[3], q[4]}, q[5]);
        BSDFEval be = m.evaluate(c, wo, bs.wi);
        float* o = out + 9 * i;
        o[0] = bs.wi.x; o[1] = bs.wi.y; o[2] = bs.wi.z;
        o[3] = bs.f.x; o[4] = bs.f.y; o[5] = bs.f.z;
        o[6] = bs.pdf; o[7] = bs.valid() ? 1.f : 0.f; o[8] = be.pdf;
    }
    return 0;
}

// evaluate(wo, wi) on a flat +Z frame: in[n*6] = wo, wi; out[n*4] = f, pdf.
int orc_material_eval(const ahip_material* mat, uint32_t n, const float* in, float* out) {
    Material m;
    m.m = *mat;
    m.init();
    SI si;
    si.n = {0, 0, 1}; si.ng = {0, 0, 1};
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 6 * i;
        V3 wo{q[0], q[1], q[2]}, wi{q[3], q[4], q[5]};
        Ctx c(si, -wo);
        BSDFEval be = m.evaluate(c, wo, wi);
        out[4 * i] = be.f.x; out[4 * i + 1] = be.f.y; out[4 * i + 2] = be.f.z; out[4 * i + 3] = be.pdf;
    }
    return 0;
}

int orc_ggx_helper(int which, uint32_t n, const float* a, const float* b, const float* c, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        switch (which) {
            case 0: out[i] = D_GGX(a[i], b[i]); break;
            case 1: out[i] = G1(a[i], b[i]); break;
            case 2: out[i] = schlickD(a[i], b[i]); break;
            case 3: out[i] = fresnelDielectric(a[i], b[i]); break;
            case 4: out[i] = pdfGGXr(a[i], b[i], c[i]); break;
            case 5: out[i] = F0ior(a[i]); break;
            case 6: {
                V3 w = sampleGGXh(V2{a[i], b[i]}, c[i]);
                out[3 * i] = w.x; out[3 * i + 1] = w.y; out[3 * i + 2] = w.z;
                break;
            }
            default: return 1;
        }
    }
    return 0;
}

int orc_distribution1d(const float* values, uint32_t count, uint32_t ns, const float* us, uint32_t* idx, float* pdf,
                       float* ur, float* cdfOut, float* integral) {
    Dist1D d;
    d.build(values, count);
    for (uint32_t k = 0; k < ns; ++k) idx[k] = d.sample(us[k], pdf[k], ur[k]);
    if (cdfOut) std::memcpy(cdfOut, d.cdf.data(), sizeof(float) * (count + 1));
    if (integral) *integral = d.integral;
    return 0;
}

// Dome probes: sample directions for u[n*2] -> out[n*5] = dir, pdf, evalPdf(dir);
// and evaluate radiance for dirs[n*3] -> rgb[n*3].
int orc_dome_sample(void* h, uint32_t n, const float* u, float* out) {
    auto* s = static_cast<Scene*>(h);
    if (!s->hasEnv) return 1;
    for (uint32_t i = 0; i < n; ++i) {
        float pdf;
        V3 d = s->dome.sampleDir(V2{u[2 * i], u[2 * i + 1]}, pdf);
        out[5 * i] = d.x; out[5 * i + 1] = d.y; out[5 * i + 2] = d.z;
        out[5 * i + 3] = pdf; out[5 * i + 4] = s->dome.evalPdf(d);
    }
    return 0;
}
int orc_dome_eval(void* h, uint32_t n, const float* dirs, float* rgb, float* pdf) {
    auto* s = static_cast<Scene*>(h);
    if (!s->hasEnv) return 1;
    for (uint32_t i = 0; i < n; ++i) {
        V3 d{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
        V3 c = s->dome.eval(d) * s->dome.intensity;
        rgb[3 * i] = c.x; rgb[3 * i + 1] = c.y; rgb[3 * i + 2] = c.z;
        if (pdf) pdf[i] = s->dome.evalPdf(d);
    }
    return 0;
}

// Material probe, the layout every probe of this repository shares:
// in[n*19]  = normal n, tangent t, strand colour, wo, wi, u.x, u.y, uComponent, curve (0 / 1)
// out[n*16] = evaluate(wo, wi): f rgb, pdf, pdfRev | sample(wo, u, uc): wi, f rgb, pdf, valid |
//             effectiveSigmaA rgb (hair; zero otherwise)
// The shading context is the integrator's for a hit seen along -wo (a curve hit takes t as its
// fiber tangent).  flags: AHIP_FLAG_HAIR_MARSCHNER selects the R/TT/TRT hair model.
int orc_probe(const ahip_material* mat, uint32_t flags, uint32_t n, const float* in, float* out) {
    Material m;
    m.m = *mat;
    m.init();
    m.marschner = (flags & AHIP_FLAG_HAIR_MARSCHNER) != 0;
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 19 * i;
        SI si;
        si.n = V3{q[0], q[1], q[2]};
        si.ng = si.n;
        si.dpdu = V3{q[3], q[4], q[5]};
        si.color = V3{q[6], q[7], q[8]};
        si.isCurve = q[18] != 0.f;
        V3 wo{q[9], q[10], q[11]}, wi{q[12], q[13], q[14]};
        Ctx c(si, -wo);
        BSDFEval e = m.evaluate(c, wo, wi);
        BSDFSample bs = m.sample(c, wo, V2{q[15], q[16]}, q[17]);
        // slots 13..15: hair effectiveSigmaA, else transmittanceColor
        V3 sa = m.m.type == AHIP_MAT_HAIR ? m.hairSigmaA(c) : m.transmittance(c);
        float* o = out + 16 * i;
        o[0] = e.f.x; o[1] = e.f.y; o[2] = e.f.z; o[3] = e.pdf; o[4] = e.pdfRev;
        o[5] = bs.wi.x; o[6] = bs.wi.y; o[7] = bs.wi.z;
        o[8] = bs.f.x; o[9] = bs.f.y; o[10] = bs.f.z; o[11] = bs.pdf; o[12] = bs.valid() ? 1.f : 0.f;
        o[13] = sa.x; o[14] = sa.y; o[15] = sa.z;
    }
    return 0;
}

// The Marschner pieces one by one, in the layouts of the reference driver's probes:
// constants v[3], s, alpha[3]; the free helpers by index (see orc_mh_fn); A_p; the lobe sums and
// pdf; the fiber geometry.
int orc_hair_consts(const ahip_material* mat, float* out) {
    mh::Consts c;
    c.init(mat->hair_beta_m, mat->hair_beta_n, mat->hair_alpha);
    for (int p = 0; p < 3; ++p) out[p] = c.v[p];
    out[3] = c.s;
    for (int p = 0; p < 3; ++p) out[4 + p] = c.alpha[p];
    return 0;
}
// 0 Mp(a,b,c,d,e)  1 Np(a,(int)b,c,d,e)  2 trimmedLogistic(a,b)  3 sampleTrimmedLogistic(a,b)
// 4 sampleGaussian(a,b)  5 logI0(a)  6 I0(a)  7 fresnelDielectric(a,b)  8 Phi((int)a,b,c)
int orc_mh_fn(int which, uint32_t n, const float* a, const float* b, const float* c, const float* d, const float* e,
              float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        switch (which) {
            case 0: out[i] = mh::Mp(a[i], b[i], c[i], d[i], e[i]); break;
            case 1: out[i] = mh::Np(a[i], (int)b[i], c[i], d[i], e[i]); break;
            case 2: out[i] = mh::trimmedLogistic(a[i], b[i]); break;
            case 3: out[i] = mh::sampleTrimmedLogistic(a[i], b[i]); break;
            case 4: out[i] = mh::sampleGaussian(a[i], b[i]); break;
            case 5: out[i] = mh::logI0(a[i]); break;
            case 6: out[i] = mh::I0(a[i]); break;
            case 7: out[i] = fresnelDielectric(a[i], b[i]); break;
            case 8: out[i] = mh::Phi((int)a[i], b[i], c[i]); break;
            default: return 1;
        }
    }
    return 0;
}
int orc_mh_ap(uint32_t n, const float* in, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 6 * i;
        V3 ap[3];
        mh::Ap(q[0], q[1], q[2], V3{q[3], q[4], q[5]}, ap);
        for (int p = 0; p < 3; ++p)
            for (int k = 0; k < 3; ++k) out[9 * i + 3 * p + k] = ap[p][k];
    }
    return 0;
}
int orc_hair_lobes(const ahip_material* mat, uint32_t n, const float* in, float* out) {
    mh::Consts c;
    c.init(mat->hair_beta_m, mat->hair_beta_n, mat->hair_alpha);
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 9 * i;
        V3 sa{q[6], q[7], q[8]};
        V3 f = c.lobes(mat->hair_eta, q[0], q[1], q[2], q[3], q[4], q[5], sa);
        out[4 * i] = f.x; out[4 * i + 1] = f.y; out[4 * i + 2] = f.z;
        out[4 * i + 3] = c.pdf(mat->hair_eta, q[0], q[1], q[2], q[3], q[4], q[5], sa);
    }
    return 0;
}
int orc_hair_geom(uint32_t n, const float* in, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 12 * i;
        V3 wo{q[0], q[1], q[2]}, wi{q[3], q[4], q[5]}, t{q[6], q[7], q[8]}, nn{q[9], q[10], q[11]};
        float sO, cO, sI, cI;
        mh::longitudinal(wo, t, sO, cO);
        mh::longitudinal(wi, t, sI, cI);
        float* o = out + 6 * i;
        o[0] = sO; o[1] = cO; o[2] = sI; o[3] = cI;
        o[4] = mh::azimuth(wo, wi, t, sO, sI);
        o[5] = mh::impact(wo, t, nn, sO);
    }
    return 0;
}

// A dome over an RGB f32 equirect image, outside any scene (the reference driver's layout):
// create / destroy; sample u[n*2] -> out[n*7] = wi, pdf, Li rgb; eval dirs[n*3] -> out[n*4] =
// Le rgb, pdf.
void* orc_dome_create(uint32_t w, uint32_t h, const float* rgb, float intensity, const float* rot) {
    auto* d = new Dome();
    d->W = w;
    d->H = h;
    d->px.assign(rgb, rgb + (size_t)w * h * 3);
    d->intensity = intensity;
    for (int r = 0; r < 3; ++r) d->rot[r] = V3{rot[3 * r], rot[3 * r + 1], rot[3 * r + 2]};
    d->build();
    return d;
}
void orc_dome_destroy(void* p) { delete static_cast<Dome*>(p); }
int orc_dome_sample_raw(void* p, uint32_t n, const float* u, float* out) {
    const Dome& d = *static_cast<Dome*>(p);
    for (uint32_t i = 0; i < n; ++i) {
        float pdf;
        V3 wi = d.sampleDir(V2{u[2 * i], u[2 * i + 1]}, pdf);
        V3 li = pdf > 0.f ? d.eval(wi) * d.intensity : V3{};
        if (!(pdf > 0.f)) wi = V3{};
        float* o = out + 7 * i;
        o[0] = wi.x; o[1] = wi.y; o[2] = wi.z; o[3] = pdf; o[4] = li.x; o[5] = li.y; o[6] = li.z;
    }
    return 0;
}
int orc_dome_eval_raw(void* p, uint32_t n, const float* dirs, float* out) {
    const Dome& d = *static_cast<Dome*>(p);
    for (uint32_t i = 0; i < n; ++i) {
        V3 w{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
        V3 le = d.eval(w) * d.intensity;
        out[4 * i] = le.x; out[4 * i + 1] = le.y; out[4 * i + 2] = le.z;
        out[4 * i + 3] = d.evalPdf(w);
    }
    return 0;
}

// One GGX lobe (StandardSurface.h:578-603), the reference driver's layout:
// in[n*10] = woLocal, wiLocal, alpha2, f0 rgb; out[n*5] = f rgb, pdfFwd, 0.
int orc_eval_ggx(uint32_t n, const float* in, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 10 * i;
        float pf;
        V3 f = Material::evalGGX(V3{q[0], q[1], q[2]}, V3{q[3], q[4], q[5]}, q[6], V3{q[7], q[8], q[9]}, pf);
        out[5 * i] = f.x; out[5 * i + 1] = f.y; out[5 * i + 2] = f.z; out[5 * i + 3] = pf; out[5 * i + 4] = 0.f;
    }
    return 0;
}

}  // extern "C"
