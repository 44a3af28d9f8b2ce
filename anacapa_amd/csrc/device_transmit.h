// device_transmit.h — stepping through transmissive surfaces on gfx950, shared by the path
// tracer's shadow and BSDF-sampled NEE rays and BDPT's connections.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.h"
#include "device_shading.h"
#include "device_traverse.h"
#include "device_types.h"

namespace ahip {

// The hit's texture coordinates (BVHBackend.cpp:285-317), uncontracted so that a texture lookup
// at them is the oracle's bit for bit.
__device__ __forceinline__ V2 interpUV(const DTriShade& S, float u, float v, float w) {
#pragma clang fp contract(off)
    return V2{S.uv0[0] * w + S.uv1[0] * u + S.uv2[0] * v, S.uv0[1] * w + S.uv1[1] * u + S.uv2[1] * v};
}

// In scenes with glass or alpha-masked materials a shadow ray and a BSDF-sampled NEE ray step
// through transmissive surfaces, closest hit by closest hit, up to 8 of them, multiplying their
// transmittanceColor tints: shadowTransmittance (ShadowRay.h:26-61) and the step loop of
// estimateDirect's BSDF strategy (PathIntegrator.cpp:217-236).  Each step re-spawns the ray from
// the hit point off the geometric normal with the remaining distance less 1e-4.
struct StepHit {
    bool hit;
    float t;
    V3 p, ng;
    V2 uv;
    int32_t mi;          // material index, -1 = none
    uint32_t commit;     // triangle in commit order (per-triangle albedo), ~0u for a curve
};
template <bool CURVES>
__device__ __forceinline__ StepHit stepClosest(const DScene& sc, const TRay& r, TStack& S, TraceCounters& tc,
                                               uint32_t* overflow) {
    StepHit sh;
    float t = 0.f, u = 0.f, v = 0.f;
    uint32_t h = traverse<false, false>(sc, r, S, t, u, v, tc, overflow);
    if (CURVES && sc.hasCurves) {
        float best = h != ~0u ? t : r.tMax;
        CurveHit ch;
        if (traverseCurves<false, false>(sc, r, S, best, ch, tc, overflow)) {
            // the curve hit's point and facing normal (CurveBrute.cpp:218-233)
            V3 dir = r.d, fx, fy;
            onbNC(dir, fx, fy);
            V3 nd = -dir, tang = ch.tang;
            V3 nrm = nd - tang * dot(nd, tang);
            float ln = len(nrm);
            nrm = ln > 1e-8f ? nrm * (1.f / ln) : fx;
            const uint32_t strand = sc.segIds[2u * ch.rec];
            const uint32_t mesh = __float_as_uint(sc.strandInfo[4u * (size_t)strand + 3u]);
            sh.hit = true;
            sh.t = best;
            sh.p = r.o + dir * best;
            sh.ng = nrm;
            sh.uv = V2{0.f, ch.tSeg};
            sh.mi = mesh < sc.numMeshes ? sc.meshMat[mesh] : -1;
            sh.commit = ~0u;
            return sh;
        }
    }
    sh.hit = h != ~0u;
    if (!sh.hit) return sh;
    const DTri* T = sc.tris + h;
    const DTriShade* SH = sc.shade + h;
    const float w = 1.f - u - v;
    sh.t = t;
    sh.p = ld3(T->v0) + ld3(T->e1) * u + ld3(T->e2) * v;
    sh.ng = ld3(T->n);
    sh.uv = interpUV(*SH, u, v, w);
    sh.mi = T->matIndex;
    sh.commit = T->commitIndex;
    return sh;
}
// transmittanceColor of a step's surface (black: opaque, or no material)
__device__ __forceinline__ V3 stepTint(const DScene& sc, const StepHit& h) {
    if (h.mi < 0) return V3{0.f, 0.f, 0.f};
    const DMat& m = sc.mats[h.mi];
    return matTransmittance(sc, m, matInputs(sc, m, h.uv, h.commit).base, h.uv);
}
__device__ __forceinline__ V3 stepOrigin(const StepHit& h, V3 d) {   // spawnRay(si.p, si.ng, d)
    V3 off = h.ng * 1e-4f;
    if (dot(d, h.ng) < 0.f) off = -off;
    return h.p + off;
}

// shadowTransmittance (ShadowRay.h:26-61) of a ray from o along d up to tMax: the product of the
// tints of up to 8 transmissive surfaces; false when an opaque one (or a surface without a
// material) blocks it or the product turns black.
template <bool CURVES>
__device__ __forceinline__ bool shadowTransmittance(const DScene& sc, V3 o, V3 d, float tMax, TStack& S,
                                                    TraceCounters& tc, uint32_t* overflow, V3& T) {
    T = V3{1.f, 1.f, 1.f};
    for (int step = 0; step < 8; ++step) {
        const StepHit h = stepClosest<CURVES>(sc, makeTRay(o, d, 1e-4f, tMax), S, tc, overflow);
        if (!h.hit) break;
        const V3 tint = stepTint(sc, h);
        if (isBlack(tint)) return false;
        T = T * tint;
        if (isBlack(T)) return false;
        const float remaining = tMax - h.t;
        if (remaining <= 1e-4f) break;
        o = stepOrigin(h, d);
        tMax = remaining - 1e-4f;
    }
    return true;
}

}  // namespace ahip
