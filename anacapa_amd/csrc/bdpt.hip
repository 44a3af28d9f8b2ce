// bdpt.hip — the bidirectional path tracer (BDPTIntegrator, src/integrator/BDPTIntegrator.cpp and
// MISWeight.cpp) on gfx950, for scenes of triangles with Lambert / Emissive / untextured opaque
// StandardSurface materials lit by quad area lights (anything else is refused on the host).
//
// Kernels over a chunk of (pixel, sample) slots:
//   k_bdpt_paths        — one slot per lane: camera subpath (BDPTIntegrator.cpp:70-194) and light
//                         subpath (:199-303), vertices written to a slot-major vertex store;
//   k_bdpt_conn_gen     — one slot per lane: every strategy (s, t) of the sample (:308-459,
//                         :521-557), the reference's connections (s = 0 emission, s = 1 light
//                         sampling, s, t >= 2 vertex pairs; t = 1 is not a strategy there) with
//                         bdptMISWeight in double (MISWeight.cpp:14-72) and the firefly clamp of
//                         s, t >= 2, into the slot's term list; the shadow rays into one queue;
//   k_bdpt_conn_trace   — one shadow ray per lane: any-hit visibility of the queue;
//   k_bdpt_conn_resolve — one slot per lane: the visible terms summed in the reference's order;
//   k_bdpt_accum        — one pixel per lane, its samples in order: the radiance sum, luminance
//                         squares and the first-hit AOVs, added to the output as renderTile's
//                         TileBuffer::add does (:559-580).
// Tracing the shadow rays from a queue keeps the lanes of a wave on rays, whatever the number of
// strategies of each slot (a slot's strategies vary from 1 to (maxDepth + 2)^2).
// Sampler dimensions follow the reference's call order (g++ evaluates sample(ctx, wo, get2D(),
// get1D()) and sampleLe(get2D(), get2D()) right to left).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/anacapa_hip.h"
#include "device_math.h"
#include "device_sampling.h"
#include "device_shading.h"
#include "device_transmit.h"
#include "device_traverse.h"
#include "device_types.h"
#include "kernels.h"

namespace ahip {

__constant__ HaltonTables c_haltonB;

namespace {

constexpr uint32_t kVCamera = 0u, kVSurface = 1u, kVLight = 2u, kVDelta = 4u, kVInf = 8u;   // PathVertex.h:24-29
constexpr uint32_t kBdptBS = 128u, kBdptK = 8u;
// term states (the w slot of a term): counted, waiting for its visibility, waiting and clamped then
constexpr float kTermPending = 0.f, kTermCounted = 1.f, kTermClamp = 2.f;

__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ V3 v3(float4 a) { return V3{a.x, a.y, a.z}; }

// vertex store: field f of vertex v of subpath `sub` (0 camera, 1 light) of slot i
//   f0: position | pdfFwd   f1: normal | pdfRev   f2: wo | flags   f3: beta | min roughness
//   f4: Le | material index (surface) or light index (light endpoint)
__device__ __forceinline__ size_t vAt(const BdptArgs& a, uint32_t sub, uint32_t v, uint32_t f, uint32_t i) {
    return ((size_t)((sub * a.maxV + v) * 5u + f)) * a.cap + i;
}

__device__ __forceinline__ float toArea(float pdfSA, V3 prev, V3 next, V3 nextN) {   // PathVertex.h:147-156
#pragma clang fp contract(off)
    V3 d = next - prev;
    float d2 = dot(d, d);
    if (d2 < 1e-10f) return 0.f;
    return pdfSA * fabsf(dot(safeNormalize(d), nextN)) / d2;
}

// The untextured material inputs of a vertex (the vertex keeps no uv: BDPTIntegrator::evalVertex
// rebuilds its ShadingContext from the position and normal only).
__device__ __forceinline__ MatIn vertexInputs(const DMat& m) { return MatIn{ld3(m.albedo), m.rough}; }

// BDPTIntegrator::evalVertex (BDPTIntegrator.cpp:40-63) at a surface vertex of material m, normal
// n, outgoing wo: f(wo, wi) |cos wi| and the forward pdf of IMaterial::evaluate.  The reverse pdf is
// read by the subpaths at wi = -wo only, where LambertianMaterial::evaluate (Lambertian.h:53-66) and
// the opaque StandardSurface (StandardSurface.h:410-417) reject the pair as opposite hemispheres and
// give 0, and rough glass answers with its transmission lobe, whose reverse pdf is its forward one
// (StandardSurface.h:444, 475); at other pairs it is formed for Lambert and glass only.
template <bool GLASS>
__device__ __forceinline__ V3 evalVertex(const DScene& sc, int32_t mi, V3 n, V3 wo, V3 wi, float& fwd, float& rev) {
    fwd = rev = 0.f;
    if (mi < 0) return V3{0.f, 0.f, 0.f};
    const DMat& m = sc.mats[mi];
    const Frame fr = makeFrame(n, n, -wo);
    const V3 f = matEval<false, GLASS>(sc, m, fr, vertexInputs(m), wo, wi, fwd);
    if (isBlack(f)) return f;
    if (m.type != AHIP_MAT_STANDARD) {
        rev = stdmax(0.f, fabsf(toLocal(fr, wo).z)) * kInvPi;
    } else {
        const bool across = !(toLocal(fr, wo).z * toLocal(fr, wi).z > 0.f);
        if (GLASS && ((across && m.trans) || (!across && m.glass))) rev = fwd;
    }
    return f * fabsf(dot(wi, n));
}

__device__ __forceinline__ float hs(const BdptArgs& a, uint32_t px, uint32_t py, uint32_t s, uint32_t& dim) {
    return haltonSample(c_haltonB, a.seed, px, py, s, dim++);
}

// The bounce loop shared by the two subpaths (BDPTIntegrator.cpp:97-190 / 240-300): from vertex
// `count - 1` at prevP along `r`, recording surface vertices until a miss, roulette or an invalid
// sample.  `light`: the light subpath (no emission recorded).  Returns the vertex count.
template <bool TRANSP, bool INF>
__device__ uint32_t walk(const DScene& sc, const BdptArgs& a, const TStack& S, uint32_t i, uint32_t sub,
                         uint32_t count, V3 prevP, TRay r, V3 beta, float pdfFwd, uint32_t px, uint32_t py,
                         uint32_t sidx, uint32_t& dim, bool light, TraceCounters& tc) {
    for (uint32_t depth = 0; depth < a.maxDepth && count < a.maxV; ++depth) {
        float t = 0.f, u = 0.f, v = 0.f;
        const uint32_t h = traverse<false, false>(sc, r, S, t, u, v, tc, a.overflow);
        if (h == ~0u) {
            // the camera subpath's environment vertex (BDPTIntegrator.cpp:100-118): the escaped
            // direction's radiance, an area pdf of 0
            if (INF && !light && sc.hasEnv && count < a.maxV) {
                const uint32_t vi = count++;
                a.verts[vAt(a, sub, vi, 0, i)] = f4(r.o + r.d * 1e6f, 0.f);
                a.verts[vAt(a, sub, vi, 1, i)] = f4(-r.d, 0.f);
                a.verts[vAt(a, sub, vi, 2, i)] = f4(-r.d, __uint_as_float(kVLight | kVInf));
                a.verts[vAt(a, sub, vi, 3, i)] = f4(beta, 1.f);
                a.verts[vAt(a, sub, vi, 4, i)] = f4(domeEval(sc.dome, r.d), __int_as_float(-1));
            }
            break;
        }
        // fillSurfaceInteraction (BVHBackend.cpp:285-317)
        const DTri* T = sc.tris + h;
        const float w = 1.f - u - v;
        const V3 v0 = ld3(T->v0), e1 = ld3(T->e1), e2 = ld3(T->e2), ng = ld3(T->n);
        const DTriShade* SH = sc.shade + h;
        const V3 p = v0 + e1 * u + e2 * v;
        V3 n = safeNormalize(ld3(SH->sn0) * w + ld3(SH->sn1) * u + ld3(SH->sn2) * v);
        if (dot(n, ng) < 0.f) n = -n;
        const V2 uv = interpUV(*SH, u, v, w);
        const int32_t mi = T->matIndex;
        if (mi < 0) break;
        const DMat& m = sc.mats[mi];
        const V3 dir = r.d, wo = -dir;
        if (TRANSP && !light && sc.hasAlpha) {
            // the camera subpath's alpha test (BDPTIntegrator.cpp:130-143): a pass-through is not a
            // bounce; a sample is drawn only for a partial opacity
            const float op = matOpacity(sc, m, uv);
            if (op <= 0.f || (op < 1.f && hs(a, px, py, sidx, dim) >= op)) {
                V3 off = ng * 1e-4f;                              // spawnRay(si.p, si.ng, dir)
                if (dot(dir, ng) < 0.f) off = -off;
                r = makeTRay(p + off, dir, 1e-4f, __int_as_float(0x7f800000));
                --depth;
                continue;
            }
        }
        const Frame fr = makeFrame(n, ng, dir);
        const float pdfFwdArea = toArea(pdfFwd, prevP, p, n);
        const uint32_t vi = count++;
        a.verts[vAt(a, sub, vi, 0, i)] = f4(p, pdfFwdArea);
        a.verts[vAt(a, sub, vi, 1, i)] = f4(n, 0.f);
        // IMaterial::roughness() (IMaterial.h:112: 1; StandardSurface.h:171: the roughness value), a
        // delta vertex counting as 1 (BDPTIntegrator.cpp:163-172)
        const float rough = m.isDelta ? 1.f : (m.type == AHIP_MAT_STANDARD ? m.rough : 1.f);
        const float minRough = vi > 0u ? stdmin(a.verts[vAt(a, sub, vi - 1u, 3, i)].w, rough) : rough;
        a.verts[vAt(a, sub, vi, 2, i)] = f4(wo, __uint_as_float(kVSurface | (m.isDelta ? kVDelta : 0u)));
        a.verts[vAt(a, sub, vi, 3, i)] = f4(beta, minRough);
        a.verts[vAt(a, sub, vi, 4, i)] = f4((!light && m.hasLe) ? ld3(m.Le) : V3{0.f, 0.f, 0.f}, __int_as_float(mi));
        if (vi > 0u) {   // the previous vertex's reverse pdf, from this one
            float dummy, revSA;
            evalVertex<TRANSP>(sc, mi, n, wo, -wo, dummy, revSA);
            const float4 pv = a.verts[vAt(a, sub, vi - 1u, 0, i)], pn = a.verts[vAt(a, sub, vi - 1u, 1, i)];
            a.verts[vAt(a, sub, vi - 1u, 1, i)] = make_float4(pn.x, pn.y, pn.z, toArea(revSA, p, v3(pv), v3(pn)));
        }
        if (depth >= 2u) {
            const float q = 1.f - stdmin(maxc(beta), 0.95f);
            if (hs(a, px, py, sidx, dim) < q) break;
            beta = beta * (1.f / (1.f - q));
        }
        const float uc = hs(a, px, py, sidx, dim);
        V2 ub;
        ub.x = hs(a, px, py, sidx, dim);
        ub.y = hs(a, px, py, sidx, dim);
        const MatIn mIn = matInputs(sc, m, uv, T->commitIndex);
        const BSample bs = matSample<false, TRANSP>(sc, m, fr, mIn, wo, ub, uc);
        if (!bsValid(bs)) break;
        beta = beta * (bs.f / bs.pdf);
        pdfFwd = bs.pdf;
        // spawnRay (Types.h:269-275)
        V3 off = n * 1e-4f;
        if (dot(bs.wi, n) < 0.f) off = -off;
        r = makeTRay(p + off, bs.wi, 1e-4f, __int_as_float(0x7f800000));
        prevP = p;
    }
    return count;
}

}  // namespace

namespace {

struct LeSample {
    V3 Le, pos, n, dir;
    float pdfPos, pdfDir;
};

// ILight::sampleLe: AreaLight.h:81-105 (cosine-weighted from a point of the quad), DomeLight.h:142-168
// (a direction from the sky's distribution, the origin on a disk of the scene sphere's radius facing
// the scene at its far side), DirectionalLight.h:88-125 (a disk 100 radii toward the light, the
// direction jittered within the cone).  A zero pdfPos marks no sample.
template <bool INF>
__device__ LeSample lightSampleLe(const DScene& sc, const DLight& L, V2 uPos, V2 uDir) {
    LeSample s{};
    if (INF && L.type == AHIP_LIGHT_DOME) {
        float pdfDir;
        const V3 envDir = domeSampleDir(sc.dome, uDir, pdfDir);
        if (pdfDir <= 0.f) return s;
        V3 t, bt;
        onb(-envDir, t, bt);
        const float R = L.sceneR, r = R * sqrtf(uPos.x), phi = 2.f * kPi * uPos.y;
        float sp, cp;
        sincosf(phi, &sp, &cp);
        const V3 dc = ld3(L.sceneC) + envDir * R;
        s.pos = dc + t * (r * cp) + bt * (r * sp);
        s.Le = domeEval(sc.dome, envDir);
        s.n = -envDir;
        s.dir = -envDir;
        s.pdfPos = 1.f / (kPi * R * R);
        s.pdfDir = pdfDir;
        return s;
    }
    if (INF && L.type == AHIP_LIGHT_DIRECTIONAL) {
        const float R = L.sceneR, r = R * sqrtf(uPos.x), phi = 2.f * kPi * uPos.y;
        float sp, cp;
        sincosf(phi, &sp, &cp);
        const V3 dl = ld3(L.dir);
        V3 t, bt;
        onb(dl, t, bt);
        const V3 dc = ld3(L.sceneC) + dl * (R * 100.f);
        s.pos = dc + t * (r * cp) + bt * (r * sp);
        V3 dir;
        if (L.cosCone >= 1.f - 1e-6f) {
            dir = -dl;
        } else {
            const float cosT = 1.f - uDir.x * (1.f - L.cosCone);
            const float sinT = sqrtf(stdmax(0.f, 1.f - cosT * cosT));
            float sq, cq;
            sincosf(2.f * kPi * uDir.y, &sq, &cq);
            const V3 nd = -dl;
            V3 dt, dbt;
            onb(nd, dt, dbt);
            dir = safeNormalize(dt * (sinT * cq) + dbt * (sinT * sq) + nd * cosT);
        }
        s.Le = ld3(L.Le);
        s.n = dir;
        s.dir = dir;
        s.pdfPos = 1.f / (kPi * R * R);
        s.pdfDir = 1.f;
        return s;
    }
    const V3 ln = ld3(L.normal);
    s.pos = ld3(L.center) + ld3(L.uA) * (2.f * uPos.x - 1.f) + ld3(L.vA) * (2.f * uPos.y - 1.f);
    const float phi = 2.f * kPi * uDir.x;
    const float cosT = sqrtf(uDir.y), sinT = sqrtf(1.f - uDir.y);
    float sp, cp;
    sincosf(phi, &sp, &cp);
    V3 t, bt;
    onb(ln, t, bt);
    s.dir = t * (sinT * cp) + bt * (sinT * sp) + ln * cosT;
    s.n = ln;
    s.Le = ld3(L.Le);
    s.pdfPos = 1.f / L.area;
    s.pdfDir = cosT * kInvPi;
    return s;
}

}  // namespace

// TRANSP: a scene with glass or alpha masks; INF: a scene with the dome or a directional light
template <bool TRANSP, bool INF>
__global__ __launch_bounds__(kBdptBS) void k_bdpt_paths(DScene sc, BdptArgs a) {
    __shared__ uint64_t lstk[kBdptK * kBdptBS];
    TStack S{lstk, a.ovfRef, a.ovfT, threadIdx.x, kBdptBS, kBdptK, blockIdx.x * kBdptBS + threadIdx.x,
             gridDim.x * kBdptBS, a.stackCap};
    TraceCounters tc{0, 0};
    for (uint32_t i = blockIdx.x * kBdptBS + threadIdx.x; i < a.slots; i += gridDim.x * kBdptBS) {
        const uint32_t pix = a.pix0 + i / a.ns, sidx = a.s0 + i % a.ns;
        const uint32_t px = a.x0 + pix % a.w, py = a.y0 + pix / a.w;
        uint32_t dim = 0u;
        // camera sample: jitter, lens, time (BDPTIntegrator.cpp:493-497); Camera::generateRay
        const float jx = hs(a, px, py, sidx, dim), jy = hs(a, px, py, sidx, dim);
        dim += 3u;
        const DCamera& cam = a.cam;
        const float fs = ((float)px + jx) / (float)cam.w;
        const float ft = ((float)(cam.h - 1u - py) + jy) / (float)cam.h;
        const V3 o = ld3(cam.origin), d = normalize(ld3(cam.llc) + ld3(cam.hor) * fs + ld3(cam.ver) * ft - o);
        // camera endpoint (BDPTIntegrator.cpp:80-92), pinhole area pdf 1
        a.verts[vAt(a, 0, 0, 0, i)] = f4(o, 1.f);
        a.verts[vAt(a, 0, 0, 1, i)] = f4(-d, 0.f);
        a.verts[vAt(a, 0, 0, 2, i)] = f4(-d, __uint_as_float(kVCamera));
        a.verts[vAt(a, 0, 0, 3, i)] = make_float4(1.f, 1.f, 1.f, 1.f);
        a.verts[vAt(a, 0, 0, 4, i)] = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
        const uint32_t nc = walk<TRANSP, INF>(sc, a, S, i, 0u, 1u, o, makeTRay(o, d, 1e-4f, __int_as_float(0x7f800000)),
                                 V3{1.f, 1.f, 1.f}, 1.f, px, py, sidx, dim, false, tc);
        // first-hit AOVs (BDPTIntegrator.cpp:502-512): IMaterial::reflectance at the first vertex
        float4 alb = make_float4(0.f, 0.f, 0.f, 0.f), nrm = make_float4(0.f, 0.f, 0.f, 0.f);
        if (nc >= 2u && (__float_as_uint(a.verts[vAt(a, 0, 1, 2, i)].w) & 3u) == kVSurface) {
            const int32_t mi = __float_as_int(a.verts[vAt(a, 0, 1, 4, i)].w);
            const DMat& m = sc.mats[mi];
            const V3 n1 = v3(a.verts[vAt(a, 0, 1, 1, i)]);
            const Frame fr = makeFrame(n1, n1, -v3(a.verts[vAt(a, 0, 1, 2, i)]));
            alb = f4(matReflectance(m, fr, vertexInputs(m)), 1.f);
            nrm = f4(n1, 0.f);
        }
        a.aovA[i] = alb;
        a.aovN[i] = nrm;
        // light subpath (BDPTIntegrator.cpp:514-519, 199-303)
        uint32_t nl = 0u;
        if (sc.numLights > 0u) {
            // LightSampler::sample (LightSampler.h:72-80)
            const float ul = hs(a, px, py, sidx, dim);
            const uint32_t n = sc.numLights;
            const uint32_t k = min((uint32_t)(ul * (float)n), n - 1u);
            const float uf = ul * (float)n - (float)k;
            const uint32_t li = uf < a.aliasProb[k] ? k : a.alias[k];
            const float lpdf = a.lightW[li];
            const DLight& L = sc.lights[li];
            const V2 uDir{hs(a, px, py, sidx, dim), hs(a, px, py, sidx, dim)};
            const V2 uPos{hs(a, px, py, sidx, dim), hs(a, px, py, sidx, dim)};
            if (lpdf > 0.f) {
                LeSample ls = lightSampleLe<INF>(sc, L, uPos, uDir);
                if (ls.pdfPos > 0.f && ls.pdfDir > 0.f) {
                    const V3 beta = ls.Le * fabsf(dot(ls.dir, ls.n)) / (lpdf * ls.pdfPos * ls.pdfDir);
                    const uint32_t fl = kVLight | (L.type == AHIP_LIGHT_DIRECTIONAL ? kVDelta : 0u) |
                                        (L.type != AHIP_LIGHT_AREA ? kVInf : 0u);
                    a.verts[vAt(a, 1, 0, 0, i)] = f4(ls.pos, lpdf * ls.pdfPos);
                    a.verts[vAt(a, 1, 0, 1, i)] = f4(ls.n, 0.f);
                    a.verts[vAt(a, 1, 0, 2, i)] = f4(ls.dir, __uint_as_float(fl));
                    a.verts[vAt(a, 1, 0, 3, i)] = f4(ls.Le / (lpdf * ls.pdfPos), 1.f);
                    a.verts[vAt(a, 1, 0, 4, i)] = f4(ls.Le, __int_as_float((int32_t)li));
                    nl = walk<TRANSP, INF>(sc, a, S, i, 1u, 1u, ls.pos, makeTRay(ls.pos, ls.dir, 1e-4f, __int_as_float(0x7f800000)),
                              beta, ls.pdfDir, px, py, sidx, dim, true, tc);
                }
            }
        }
        a.counts[i] = nc | (nl << 16);
    }
}

namespace {

// bdptMISWeight (MISWeight.cpp:14-72): pdfs as stored on the vertices, ratios in double
__device__ float misWeight(const BdptArgs& a, uint32_t i, uint32_t nl, uint32_t nc, uint32_t s, uint32_t t) {
    if (s + t == 2u) return 1.f;
    double sum = 0.0;
    for (uint32_t side = 0; side < 2u; ++side) {
        const uint32_t sub = side == 0u ? 1u : 0u, cnt = side == 0u ? nl : nc, k = side == 0u ? s : t;
        double ri = 1.0;
        for (int j = (int)k - 1; j >= 0; --j) {
            const bool in = j < (int)cnt;
            const float num = in ? a.verts[vAt(a, sub, (uint32_t)j, 1, i)].w : 0.f;
            const float den = in ? a.verts[vAt(a, sub, (uint32_t)j, 0, i)].w : 0.f;
            if (den == 0.f) break;
            ri *= (double)num / (double)den;
            // num / den of two floats with den != 0 is finite in double, so once the ratio product is
            // 0 every later term is 0 (a NaN product is kept to the end, as the reference's loop does)
            if (ri == 0.0) break;
            const bool prevDelta = j > 0 && (__float_as_uint(a.verts[vAt(a, sub, (uint32_t)j - 1u, 2, i)].w) & kVDelta);
            const bool curDelta = in && (__float_as_uint(a.verts[vAt(a, sub, (uint32_t)j, 2, i)].w) & kVDelta);
            if (!prevDelta && !curDelta) sum += ri * ri;
        }
    }
    return (float)(1.0 / (sum + 1.0));
}

}  // namespace

// Every strategy (s, t) of the slot in the reference's order (t outer, s inner), without its
// visibility: the weighted, clamped contribution goes to the slot's term list (term k of slot i at
// terms[k * cap + i]); a strategy that needs a visibility test leaves its term at weight 0 and
// appends its shadow ray (spawnRayTo, Types.h:277-284) to the connection queue.
template <bool TRANSP, bool INF>
__global__ __launch_bounds__(kBdptBS) void k_bdpt_conn_gen(DScene sc, BdptArgs a) {
    const uint32_t lane = __lane_id();
    for (uint32_t base0 = blockIdx.x * kBdptBS; base0 < a.slots; base0 += gridDim.x * kBdptBS) {
        const uint32_t i = base0 + threadIdx.x;
        const bool valid = i < a.slots;
        const uint32_t cnt = valid ? a.counts[i] : 0u, nc = cnt & 0xffffu, nl = cnt >> 16;
        uint32_t k = 0u, nr = 0u;
        for (uint32_t t = 1; t <= nc; ++t) {
            const uint32_t ct = t - 1u;
            const float4 c0 = a.verts[vAt(a, 0, ct, 0, i)], c1 = a.verts[vAt(a, 0, ct, 1, i)];
            const float4 c2 = a.verts[vAt(a, 0, ct, 2, i)], c3 = a.verts[vAt(a, 0, ct, 3, i)];
            const float4 c4 = a.verts[vAt(a, 0, ct, 4, i)];
            const uint32_t cflags = __float_as_uint(c2.w);
            const bool cSurf = (cflags & 3u) == kVSurface;
            const int32_t cm = cSurf ? __float_as_int(c4.w) : -1;
            const V3 cp = v3(c0), cn = v3(c1), cwo = v3(c2), cbeta = v3(c3);
            for (uint32_t s = 0; s <= nl; ++s) {
                if (s + t < 2u) continue;
                V3 C{0.f, 0.f, 0.f};
                bool ray = false;
                V3 from{0.f, 0.f, 0.f}, fromN{0.f, 0.f, 0.f}, to{0.f, 0.f, 0.f};
                if (s == 0u) {
                    // (0, t): emission at the last camera vertex (BDPTIntegrator.cpp:318-343); an
                    // environment vertex only straight after the camera (t = 2)
                    if (t >= 2u && (cSurf || (t == 2u && (cflags & kVInf)))) C = cbeta * v3(c4);
                } else if (INF && s == 1u && (__float_as_uint(a.verts[vAt(a, 1, 0, 2, i)].w) & kVInf)) {
                    // (1, t) from the dome or a directional light: along the light subpath's own
                    // direction, a shadow ray 1e10 long, Le / (light pdf x directional pdf)
                    // (:357-365, :395-402; DirectionalLight.h:78-80: pdf 0, taken as 1)
                    if (!(cflags & kVDelta) && cm >= 0) {
                        const float4 l2 = a.verts[vAt(a, 1, 0, 2, i)], l4 = a.verts[vAt(a, 1, 0, 4, i)];
                        const int32_t li = __float_as_int(l4.w);
                        const DLight& L = sc.lights[li];
                        const V3 wi = safeNormalize(-v3(l2));
                        float pdf;
                        const DMat& m = sc.mats[cm];
                        const Frame fr = makeFrame(cn, cn, -cwo);
                        const V3 f = matEval<false, TRANSP>(sc, m, fr, vertexInputs(m), cwo, wi, pdf);
                        V3 Le{0.f, 0.f, 0.f};
                        float pdfDir = 1.f;
                        if (L.type == AHIP_LIGHT_DOME) {
                            Le = domeEval(sc.dome, wi);
                            pdfDir = domePdf(sc.dome, wi);
                            if (pdfDir <= 0.f) pdfDir = 1.f;
                        } else if (dot(wi, ld3(L.dir)) > L.cosCone - 1e-4f) {   // DirectionalLight::Le (:131-134)
                            Le = ld3(L.Le);
                        }
                        if (!isBlack(f) && !isBlack(Le)) {
                            const float cosI = fabsf(dot(wi, cn));
                            C = cbeta * f * cosI * Le / (a.lightW[li] * pdfDir);
                            ray = true;
                            from = cp;
                            fromN = cn;
                            to = cp + wi * 1e10f;
                        }
                    }
                } else if (s == 1u) {
                    // (1, t): the light endpoint toward camera vertex t-1 (:345-410)
                    if (!(cflags & kVDelta) && cm >= 0) {
                        const float4 l0 = a.verts[vAt(a, 1, 0, 0, i)], l1 = a.verts[vAt(a, 1, 0, 1, i)];
                        const float4 l4 = a.verts[vAt(a, 1, 0, 4, i)];
                        const V3 toL = v3(l0) - cp;
                        const float dist = len(toL);
                        if (dist >= 1e-6f) {
                            const V3 wi = toL * (1.f / dist);
                            float pdf;
                            const DMat& m = sc.mats[cm];
                            const Frame fr = makeFrame(cn, cn, -cwo);
                            const V3 f = matEval<false, TRANSP>(sc, m, fr, vertexInputs(m), cwo, wi, pdf);
                            const V3 ln = v3(l1);
                            const V3 Le = dot(wi, ln) < 0.f ? v3(l4) : V3{0.f, 0.f, 0.f};   // AreaLight.h:110-112
                            if (!isBlack(f) && !isBlack(Le)) {
                                const float cosI = fabsf(dot(wi, cn)), cosL = fabsf(dot(-wi, ln));
                                C = cbeta * f * cosI * Le * cosL / (dist * dist * l0.w);
                                ray = true;
                                from = cp;
                                fromN = cn;
                                to = cp + wi * dist;
                            }
                        }
                    }
                } else if (t >= 2u) {
                    // (s, t >= 2): connect light vertex s-1 and camera vertex t-1 (:419-459)
                    const uint32_t ls = s - 1u;
                    const float4 l0 = a.verts[vAt(a, 1, ls, 0, i)], l1 = a.verts[vAt(a, 1, ls, 1, i)];
                    const float4 l2 = a.verts[vAt(a, 1, ls, 2, i)], l3 = a.verts[vAt(a, 1, ls, 3, i)];
                    const float4 l4 = a.verts[vAt(a, 1, ls, 4, i)];
                    const uint32_t lflags = __float_as_uint(l2.w);
                    const bool lSurf = (lflags & 3u) == kVSurface;
                    const int32_t lm = lSurf ? __float_as_int(l4.w) : -1;
                    if (!(lflags & kVDelta) && !(cflags & kVDelta) && lm >= 0 && cm >= 0 && l3.w >= 0.08f &&
                        c3.w >= 0.08f) {
                        const V3 d = cp - v3(l0);
                        const float dist = len(d);
                        if (dist >= 1e-6f) {
                            const V3 wi = d * (1.f / dist);
                            float f1, r1, f2, r2;
                            const V3 fL = evalVertex<TRANSP>(sc, lm, v3(l1), v3(l2), wi, f1, r1);
                            const V3 fC = isBlack(fL) ? V3{0.f, 0.f, 0.f} : evalVertex<TRANSP>(sc, cm, cn, cwo, -wi, f2, r2);
                            if (!isBlack(fC)) {
                                C = v3(l3) * fL * (1.f / (dist * dist)) * fC * cbeta;
                                ray = true;
                                from = v3(l0);
                                fromN = v3(l1);
                                to = cp;
                            }
                        }
                    }
                }
                bool keep = !isBlack(C) && finite3(C);
                V3 c{0.f, 0.f, 0.f};
                if (keep) {
                    c = C * misWeight(a, i, nl, nc, s, t);
                    keep = finite3(c);
                }
                if (keep) {
                    // a term without a visibility test counts at once (s = 0: never clamped); one with
                    // a test counts once its ray is traced, which multiplies in the connection's
                    // transmittance and then applies the firefly clamp of s, t >= 2 (kTermClamp)
                    const uint32_t idx = k * a.cap + i;
                    a.terms[idx] = f4(c, !ray ? kTermCounted : (s >= 2u && t >= 2u) ? kTermClamp : kTermPending);
                    if (ray) {   // staged in the slot's own rows until the wave's queue space is known
                        // spawnRayTo(from, fromN, to) (Types.h:277-284)
                        V3 dir = to - from;
                        const float dist = len(dir);
                        dir = dir * (1.f / dist);
                        V3 off = fromN * 1e-4f;
                        if (dot(dir, fromN) < 0.f) off = -off;
                        a.stageO[nr * a.cap + i] = f4(from + off, __uint_as_float(idx));
                        a.stageD[nr * a.cap + i] = f4(dir, dist * (1.f - 1e-4f));
                        ++nr;
                    }
                    ++k;
                }
            }
        }
        if (valid) a.nTerms[i] = k;
        // the wave's shadow rays go to the queue at a wave prefix of the slots' counts: one atomic
        // per wave
        uint32_t incl = nr;
        for (uint32_t off = 1; off < 64u; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off);
            if (lane >= off) incl += y;
        }
        const uint32_t total = __shfl(incl, 63);
        uint32_t qbase = 0u;
        if (lane == 0u && total) qbase = atomicAdd(a.connCount, total);
        qbase = __shfl(qbase, 0) + incl - nr;
        for (uint32_t r = 0; r < nr; ++r) {
            a.connO[qbase + r] = a.stageO[r * a.cap + i];
            a.connD[qbase + r] = a.stageD[r * a.cap + i];
        }
    }
}

// The connection queue's visibility, shadowTransmittance (ShadowRay.h:26-61): in an opaque scene any
// hit in [1e-4, tMax) blocks the connection; with glass or alpha masks (TRANSP) the ray steps through
// transmissive surfaces and their tint multiplies the term.  A visible term then gets the firefly
// clamp its strategy takes (BDPTIntegrator.cpp:547-556) and counts.
template <bool TRANSP>
__global__ __launch_bounds__(kBdptBS) void k_bdpt_conn_trace(DScene sc, BdptArgs a) {
    __shared__ uint64_t lstk[kBdptK * kBdptBS];
    TStack S{lstk, a.ovfRef, a.ovfT, threadIdx.x, kBdptBS, kBdptK, blockIdx.x * kBdptBS + threadIdx.x,
             gridDim.x * kBdptBS, a.stackCap};
    TraceCounters tc{0, 0};
    const uint32_t count = *a.connCount;
    for (uint32_t j = blockIdx.x * kBdptBS + threadIdx.x; j < count; j += gridDim.x * kBdptBS) {
        const float4 o = a.connO[j], d = a.connD[j];
        V3 T{1.f, 1.f, 1.f};
        bool visible;
        if constexpr (TRANSP) {
            visible = shadowTransmittance<false>(sc, v3(o), v3(d), d.w, S, tc, a.overflow, T);
        } else {
            float t, u, v;
            visible = traverse<true, false>(sc, makeTRay(v3(o), v3(d), 1e-4f, d.w), S, t, u, v, tc, a.overflow) == ~0u;
        }
        if (!visible) continue;
        const uint32_t idx = __float_as_uint(o.w);
        const float4 x = a.terms[idx];
        V3 c = TRANSP ? v3(x) * T : v3(x);
        if (x.w == kTermClamp && a.clamp > 0.f) {
            const float l = luminance(c);
            if (l > a.clamp) c = c * (a.clamp / l);
        }
        a.terms[idx] = f4(c, kTermCounted);
    }
}

// The slot's radiance: its visible terms in the reference's order (BDPTIntegrator.cpp:523-557).
__global__ __launch_bounds__(256) void k_bdpt_conn_resolve(BdptArgs a) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.slots; i += gridDim.x * blockDim.x) {
        const uint32_t n = a.nTerms[i];
        V3 sum{0.f, 0.f, 0.f};
        for (uint32_t k = 0; k < n; ++k) {
            const float4 x = a.terms[k * a.cap + i];
            if (x.w == kTermCounted) sum = sum + v3(x);
        }
        a.sampleL[i] = f4(sum, 0.f);
    }
}

// Per pixel of the chunk: its samples in order (BDPTIntegrator.cpp:521-580); the tile output gets
// pixelL / spp * spp, the weight spp, the luminance squares and the AOV sums.
__global__ __launch_bounds__(256) void k_bdpt_accum(BdptArgs a) {
    const uint32_t npix = a.slots / a.ns;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < npix; k += gridDim.x * blockDim.x) {
        const uint32_t pix = a.pix0 + k;
        const uint32_t lx = pix % a.w, ly = pix / a.w;
        V3 L{0.f, 0.f, 0.f}, A{0.f, 0.f, 0.f}, N{0.f, 0.f, 0.f};
        float lsq = 0.f, na = 0.f;
        for (uint32_t s = 0; s < a.ns; ++s) {
            const uint32_t i = k * a.ns + s;
            const V3 x = v3(a.sampleL[i]);
            L = L + x;
            const float l = luminance(x);
            lsq += l * l;
            const float4 al = a.aovA[i];
            if (al.w > 0.f) {
                A = A + v3(al);
                N = N + v3(a.aovN[i]);
                na += 1.f;
            }
        }
        const float inv = 1.f / (float)a.ns, wt = (float)a.ns;
        const size_t o = (size_t)(a.outY0 + ly) * a.outStride + (a.outX0 + lx);
        const float4 prev = a.outAcc[o];
        a.outAcc[o] = make_float4(prev.x + L.x * inv * wt, prev.y + L.y * inv * wt, prev.z + L.z * inv * wt,
                                  prev.w + wt);
        a.outLsq[o] += lsq;
        if (a.outAov) {   // mean reflectance and mean normal of the samples that hit a surface
            float* q = a.outAov + 6u * o;
            if (na > 0.f) {
                const float invN = 1.f / na;
                const V3 ma = A * invN;
                V3 mn = N * invN;
                const float l = len(mn);
                if (l > 1e-6f) mn = mn * (1.f / l);
                q[0] += ma.x; q[1] += ma.y; q[2] += ma.z;
                q[3] += mn.x; q[4] += mn.y; q[5] += mn.z;
            }
        }
    }
}

void setHaltonTablesBdpt(const HaltonTables& t, hipStream_t s) {
    hipMemcpyToSymbolAsync(HIP_SYMBOL(c_haltonB), &t, sizeof(t), 0, hipMemcpyHostToDevice, s);
}

__global__ void k_bdpt_zero(uint32_t* p) { *p = 0u; }

void launchBdpt(const DScene& sc, const BdptArgs& a, uint32_t blocks, uint32_t traceBlocks, hipStream_t s) {
    hipLaunchKernelGGL(k_bdpt_zero, dim3(1), dim3(1), 0, s, a.connCount);
    const bool inf = sc.hasEnv || sc.hasDirectional;
    if (sc.hasTransparent && inf) {
        hipLaunchKernelGGL((k_bdpt_paths<true, true>), dim3(blocks), dim3(kBdptBS), 0, s, sc, a);
        hipLaunchKernelGGL((k_bdpt_conn_gen<true, true>), dim3(blocks), dim3(kBdptBS), 0, s, sc, a);
    } else if (sc.hasTransparent) {
        hipLaunchKernelGGL((k_bdpt_paths<true, false>), dim3(blocks), dim3(kBdptBS), 0, s, sc, a);
        hipLaunchKernelGGL((k_bdpt_conn_gen<true, false>), dim3(blocks), dim3(kBdptBS), 0, s, sc, a);
    } else if (inf) {
        hipLaunchKernelGGL((k_bdpt_paths<false, true>), dim3(blocks), dim3(kBdptBS), 0, s, sc, a);
        hipLaunchKernelGGL((k_bdpt_conn_gen<false, true>), dim3(blocks), dim3(kBdptBS), 0, s, sc, a);
    } else {
        hipLaunchKernelGGL((k_bdpt_paths<false, false>), dim3(blocks), dim3(kBdptBS), 0, s, sc, a);
        hipLaunchKernelGGL((k_bdpt_conn_gen<false, false>), dim3(blocks), dim3(kBdptBS), 0, s, sc, a);
    }
    if (sc.hasTransparent) hipLaunchKernelGGL(k_bdpt_conn_trace<true>, dim3(traceBlocks), dim3(kBdptBS), 0, s, sc, a);
    else hipLaunchKernelGGL(k_bdpt_conn_trace<false>, dim3(traceBlocks), dim3(kBdptBS), 0, s, sc, a);
    hipLaunchKernelGGL(k_bdpt_conn_resolve, dim3((a.slots + 255u) / 256u), dim3(256), 0, s, a);
    const uint32_t npix = a.slots / a.ns;
    hipLaunchKernelGGL(k_bdpt_accum, dim3((npix + 255u) / 256u), dim3(256), 0, s, a);
}

}  // namespace ahip
