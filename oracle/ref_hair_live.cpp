// ref_hair_live.cpp — test-infrastructure driver around the reference's Marschner hair code.
//
// ORACLE / TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ with the
// reference's Release flags; only tests/ and tests/golden/ load it.
//
// The reference's MarschnerHairMaterial (src/shading/MarschnerHair.h) evaluates Kajiya-Kay at run
// time: evaluate() returns kajiyaKayEval on its first line (:240) and sample() returns a
// cosine-sampled Kajiya-Kay sample (:266-282), so its R/TT/TRT code (:243-257, :284-375,
// :490-556) never runs.  This translation unit includes MarschnerHairLive.h, a copy of that header
// generated at build time in which exactly those 18 lines are deleted and the class is renamed
// MarschnerHairLive; every line that remains is the reference's.  Its private members are read
// through an access macro applied to that header only.  The Kajiya-Kay material itself is driven
// from ref_driver.cpp through the original header.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <anacapa/core/Types.h>
#include <anacapa/shading/IMaterial.h>
#include <anacapa/shading/ShadingContext.h>

#define private public
#include "MarschnerHairLive.h"
#undef private

#include "../include/anacapa_hip.h"

using namespace anacapa;

namespace {

Vec3f v3(const float* p) { return Vec3f{p[0], p[1], p[2]}; }

MarschnerHairLive::Params liveParams(const ahip_material& m) {
    MarschnerHairLive::Params p;
    p.eta = m.hair_eta;
    p.sigma_a = v3(m.hair_sigma_a);
    p.beta_m = m.hair_beta_m;
    p.beta_n = m.hair_beta_n;
    p.alpha = m.hair_alpha;
    return p;
}

}  // namespace

extern "C" {

// Material probe on a curve hit, the layout every probe of this repository shares:
// in[n*19]  = ribbon normal n, fiber tangent t, strand colour, wo, wi, u.x, u.y, uComponent, curve
// out[n*16] = evaluate(wo, wi): f rgb, pdf, pdfRev | sample(wo, u, uc): wi, f rgb, pdf, valid |
//             effectiveSigmaA rgb
// The ShadingContext is the one PathIntegrator builds for a curve hit seen along -wo
// (ShadingContext.h:28-45, isCurve = true, dpdu = tangent).
int ref_hair_live_probe(const ahip_material* mat, uint32_t n, const float* in, float* out) {
    MarschnerHairLive m(liveParams(*mat));
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 19 * i;
        SurfaceInteraction si;
        si.n = v3(q);
        si.ng = si.n;
        si.dpdu = v3(q + 3);
        si.color = v3(q + 6);
        si.isCurve = true;
        Vec3f wo = v3(q + 9), wi = v3(q + 12);
        ShadingContext ctx(si, -wo);
        BSDFEval e = m.evaluate(ctx, wo, wi);
        BSDFSample s = m.sample(ctx, wo, Vec2f{q[15], q[16]}, q[17]);
        Spectrum sa = m.effectiveSigmaA(ctx);
        float* o = out + 16 * i;
        o[0] = e.f.x; o[1] = e.f.y; o[2] = e.f.z; o[3] = e.pdf; o[4] = e.pdfRev;
        o[5] = s.wi.x; o[6] = s.wi.y; o[7] = s.wi.z;
        o[8] = s.f.x; o[9] = s.f.y; o[10] = s.f.z; o[11] = s.pdf; o[12] = s.isValid() ? 1.f : 0.f;
        o[13] = sa.x; o[14] = sa.y; o[15] = sa.z;
    }
    return 0;
}

// Constructor-derived constants (MarschnerHair.h:200-216): v[R,TT,TRT], s, alpha[R,TT,TRT].
int ref_hair_consts(const ahip_material* mat, float* out) {
    MarschnerHairLive m(liveParams(*mat));
    for (int p = 0; p < 3; ++p) out[p] = m.m_v[p];
    out[3] = m.m_s;
    for (int p = 0; p < 3; ++p) out[4 + p] = m.m_alpha[p];
    return 0;
}

// Free helpers (MarschnerHair.h:49-185), elementwise:
// 0 mh_Mp(a=cosI, b=sinI, c=cosR, d=sinR, e=v)   1 mh_Np(phi=a, p=(int)b, s=c, gammaO=d, gammaT=e)
// 2 mh_trimmedLogistic(a, b)   3 mh_sampleTrimmedLogistic(u=a, s=b)   4 mh_sampleGaussian(a, b)
// 5 mh_logI0(a)   6 mh_I0(a)   7 mh_fresnelDielectric(a, b)   8 mh_Phi(p=(int)a, b, c)
int ref_mh_fn(int which, uint32_t n, const float* a, const float* b, const float* c, const float* d,
              const float* e, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        switch (which) {
            case 0: out[i] = mh_Mp(a[i], b[i], c[i], d[i], e[i]); break;
            case 1: out[i] = mh_Np(a[i], (int)b[i], c[i], d[i], e[i]); break;
            case 2: out[i] = mh_trimmedLogistic(a[i], b[i]); break;
            case 3: out[i] = mh_sampleTrimmedLogistic(a[i], b[i]); break;
            case 4: out[i] = mh_sampleGaussian(a[i], b[i]); break;
            case 5: out[i] = mh_logI0(a[i]); break;
            case 6: out[i] = mh_I0(a[i]); break;
            case 7: out[i] = mh_fresnelDielectric(a[i], b[i]); break;
            case 8: out[i] = mh_Phi((int)a[i], b[i], c[i]); break;
            default: return 1;
        }
    }
    return 0;
}

// mh_Ap (MarschnerHair.h:146-176): in[n*6] = cosThetaO, eta, h, sigma_a rgb; out[n*9] = A_R, A_TT, A_TRT.
int ref_mh_ap(uint32_t n, const float* in, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 6 * i;
        auto ap = mh_Ap(q[0], q[1], q[2], v3(q + 3));
        for (int p = 0; p < 3; ++p) {
            out[9 * i + 3 * p + 0] = ap[p].x;
            out[9 * i + 3 * p + 1] = ap[p].y;
            out[9 * i + 3 * p + 2] = ap[p].z;
        }
    }
    return 0;
}

// evalLobes / evalPdf (MarschnerHair.h:490-556): in[n*9] = sinO, cosO, sinI, cosI, phi, h, sigma_a rgb;
// out[n*4] = f rgb, pdf.
int ref_hair_lobes(const ahip_material* mat, uint32_t n, const float* in, float* out) {
    MarschnerHairLive m(liveParams(*mat));
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 9 * i;
        Spectrum f = m.evalLobes(q[0], q[1], q[2], q[3], q[4], q[5], v3(q + 6));
        float pdf = m.evalPdf(q[0], q[1], q[2], q[3], q[4], q[5], v3(q + 6));
        out[4 * i] = f.x; out[4 * i + 1] = f.y; out[4 * i + 2] = f.z; out[4 * i + 3] = pdf;
    }
    return 0;
}

// Fiber geometry (MarschnerHair.h:456-484): in[n*12] = wo, wi, t, n; out[n*6] = sinO, cosO, sinI,
// cosI (hairLongitudinal), phi (hairPhi), h (hairImpactParam).
int ref_hair_geom(uint32_t n, const float* in, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 12 * i;
        Vec3f wo = v3(q), wi = v3(q + 3), t = v3(q + 6), nn = v3(q + 9);
        float sO, cO, sI, cI;
        MarschnerHairLive::hairLongitudinal(wo, t, sO, cO);
        MarschnerHairLive::hairLongitudinal(wi, t, sI, cI);
        float* o = out + 6 * i;
        o[0] = sO; o[1] = cO; o[2] = sI; o[3] = cI;
        o[4] = MarschnerHairLive::hairPhi(wo, wi, t, sO, sI);
        o[5] = MarschnerHairLive::hairImpactParam(wo, t, nn, sO);
    }
    return 0;
}

}  // extern "C"
