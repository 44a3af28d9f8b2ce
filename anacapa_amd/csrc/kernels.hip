// kernels.hip — gfx950 kernels of the wavefront Path+NEE integrator.
//
// One camera sample is one path.  A render is a sequence of passes over up to `cap` paths;
// each pass runs, for bounce = 0..maxDepth:
//   extend  — closest hit of every live path's ray (queue compacted by the previous shade)
//   shade   — PathIntegrator::Li's per-bounce body (src/integrator/PathIntegrator.cpp:80-161):
//             environment on a miss, emission, uniform light pick, the two NEE strategies of
//             estimateDirect (PathIntegrator.cpp:166-265) emitted as queued rays, continuation
//             BSDF sample and Russian roulette; survivors are appended to the next queue
//   shadow  — any-hit of the NEE light-sample rays; unoccluded ones add their contribution
//   mis     — the BSDF-sampled NEE rays: closest hit (or any hit when no surface emits) and
//             the MIS-weighted emitter / environment radiance
// then `accumulate` folds the pass's per-path radiance into the pixel sums exactly as
// PathIntegrator::renderTile does (finite samples only, luminance^2, weight = sample count).
//
// Queues are appended with one atomic per wavefront: a 64-lane ballot, a popcount prefix and a
// broadcast of the base slot.  Every kernel is a block-uniform grid-stride loop whose trip
// count comes from a device counter, so launch shapes are static and the host never reads a
// queue length (the same launches can be replayed from a HIP graph).
#include <hip/hip_runtime.h>
#include <cstdlib>
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

__constant__ HaltonTables c_halton;

// Counter slots per bounce: [0] continuation rays entering this bounce, [1] shadow rays,
// [2] MIS rays, [3] stack-overflow flag.
__device__ __forceinline__ uint32_t* ctr(const WaveQueues& q, uint32_t bounce, uint32_t k) {
    return q.counters + bounce * 4u + k;
}

// Block-level append to NQ queues: ballots per wave, a prefix over the block's waves in LDS and
// one atomic per queue per block (a wave-level atomic per queue made the shared counters the
// bottleneck at tens of millions of paths per pass).  Every thread of the block must call it.
template <uint32_t NQ, uint32_t WAVES>
__device__ __forceinline__ void blockAppend(uint32_t* const* ctrs, const bool* want, uint32_t* slot,
                                            uint32_t* lds /* [NQ * (WAVES + 1)] */) {
    const uint32_t wave = threadIdx.x >> 6, lane = __lane_id();
    const uint64_t below = lane == 0u ? 0ull : (~0ull >> (64u - lane));
    uint32_t rank[NQ];
#pragma unroll
    for (uint32_t k = 0; k < NQ; ++k) {
        uint64_t mask = __ballot(want[k] ? 1 : 0);
        rank[k] = (uint32_t)__popcll(mask & below);
        if (lane == 0u) lds[k * WAVES + wave] = (uint32_t)__popcll(mask);
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        const uint32_t k = threadIdx.x;
        uint32_t total = 0u;
        for (uint32_t w = 0; w < WAVES; ++w) {
            uint32_t c = lds[k * WAVES + w];
            lds[k * WAVES + w] = total;
            total += c;
        }
        lds[NQ * WAVES + k] = total ? atomicAdd(ctrs[k], total) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < NQ; ++k) slot[k] = lds[NQ * WAVES + k] + lds[k * WAVES + wave] + rank[k];
    __syncthreads();   // the LDS is reused by the next call
}

// Path flags in the dimension word of a queued ray: a padding slot outside its view; the last
// bounce sampled a delta lobe; the first-hit AOVs are written (the first hit an alpha mask does
// not pass through, PathIntegrator.cpp:104-126).
constexpr uint32_t kDeadPath = 1u << 30;
constexpr uint32_t kSpecularPath = 1u << 31;
constexpr uint32_t kAovDone = 1u << 29;
constexpr uint32_t kPathFlags = kDeadPath | kSpecularPath | kAovDone;

__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ V3 v3(float4 a) { return V3{a.x, a.y, a.z}; }

// Slot of a view -> pixel: 64 consecutive slots cover one tileW x tileH pixel tile, so a 64-lane
// wavefront traces one coherent tile at bounce 0.
__device__ __forceinline__ bool slotPixel(const ViewDesc& v, uint32_t slot, uint32_t& lx, uint32_t& ly) {
    uint32_t tile = slot >> 6, lane = slot & 63u;
    uint32_t tx = tile % v.tilesX, ty = tile / v.tilesX;
    lx = tx * v.tileW + lane % v.tileW;
    ly = ty * v.tileH + lane / v.tileW;
    return lx < v.w && ly < v.h;
}

// Segment of the pass holding `key` (largest k with base(k) <= key).
template <bool SLOTS>
__device__ __forceinline__ uint32_t findSeg(const PassTable& T, uint32_t key) {
    uint32_t lo = 0, hi = T.numSegs;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        uint32_t b = SLOTS ? T.seg[mid].slotBase : T.seg[mid].pathBase;
        if (b <= key) lo = mid;
        else hi = mid;
    }
    return lo;
}

// --------------------------------------------------------------------------- utility kernels
__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f32(float* p, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_halton_eval(uint64_t seed, const uint32_t* px, const uint32_t* py, const uint32_t* idx,
                              const uint32_t* dim, uint32_t n, float* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = haltonSample(c_halton, seed, px[i], py[i], idx[i], dim[i]);
}

// --------------------------------------------------------------------------- generate
// Camera rays: HaltonSampler::startPixelSample + get2D (jitter) + get2D (lens) + get1D (time)
// and Camera::generateRay for a pinhole (PathIntegrator.cpp:34-38; IIntegrator.h:90-108).
__global__ __launch_bounds__(256) void k_generate(RenderConsts rc, WaveQueues q) {
    const PassTable& T = q.tables[*q.passIdx];
    const uint32_t total = T.totalPaths;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
        uint32_t p0 = base + threadIdx.x;
        uint32_t px = 0, py = 0, s = 0;
        bool valid = false;
        const ViewDesc* V = nullptr;
        if (p0 < total) {
            const PassSeg& sg = T.seg[findSeg<false>(T, p0)];
            V = q.views + sg.view;
            uint32_t local = p0 - sg.pathBase, lx, ly;
            valid = slotPixel(*V, local % V->slots, lx, ly);
            px = V->x0 + lx;
            py = V->y0 + ly;
            s = sg.sampleStart + local / V->slots;
        }
        V3 o{0.f, 0.f, 0.f}, d{0.f, 0.f, 0.f};
        if (valid) {
            const DCamera& cam = V->cam;
            float jx = haltonSample(c_halton, rc.seed, px, py, s, 0u);
            float jy = haltonSample(c_halton, rc.seed, px, py, s, 1u);
            float fs = ((float)px + jx) / (float)cam.w;
            float ft = ((float)(cam.h - 1u - py) + jy) / (float)cam.h;
            V3 target = ld3(cam.llc) + ld3(cam.hor) * fs + ld3(cam.ver) * ft;
            o = ld3(cam.origin);
            d = normalize(target - o);
            q.pix[p0] = make_uint2(px | (py << 16), s);
        }
        // bounce-0 queue slot = path id (no compaction: padding slots are rare and marked dead)
        if (p0 < total) {
            q.L[p0] = make_float4(0.f, 0.f, 0.f, 0.f);
            q.ro[0][p0] = f4(o, __uint_as_float(p0));
            q.rd[0][p0] = f4(d, __uint_as_float(valid ? 5u : kDeadPath));   // dims 0..4: camera sample
            q.beta[0][p0] = make_float4(1.f, 1.f, 1.f, 0.f);
            if (q.aovA) {   // a camera ray that hits nothing reports zero AOVs
                q.aovA[p0] = make_float4(0.f, 0.f, 0.f, 0.f);
                q.aovN[p0] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
    }
    if (blockIdx.x == 0u && threadIdx.x == 0u) *ctr(q, 0, 0) = total;
}

// --------------------------------------------------------------------------- extend
// CURVES = false compiles the curve traversal out (triangle-only scenes keep fewer registers and
// more waves in flight); the host picks the instance from DScene::hasCurves.
template <uint32_t BS, uint32_t K, bool CURVES>
__global__ __launch_bounds__(BS) void k_extend(DScene sc, WaveQueues q, uint32_t bounce, uint32_t cur) {
    __shared__ uint64_t lstk[K * BS];
    TStack S{lstk, q.ovfRef, q.ovfT, threadIdx.x, BS, K, blockIdx.x * BS + threadIdx.x, gridDim.x * BS,
             q.stackCap};
    const uint32_t count = *ctr(q, bounce, 0);
    TraceCounters tc{0, 0};
    for (uint32_t base = blockIdx.x * BS; base < count; base += gridDim.x * BS) {
        uint32_t i = base + threadIdx.x;
        if (i >= count) break;
        float4 a = q.ro[cur][i], b = q.rd[cur][i];
        if (__float_as_uint(b.w) & kDeadPath) {
            q.hit[i] = make_float4(0.f, 0.f, 0.f, __uint_as_float(~0u));
            continue;
        }
        TRay r = makeTRay(v3(a), v3(b), 1e-4f, __int_as_float(0x7f800000));
        float t = 0.f, u = 0.f, v = 0.f;
        uint32_t h = traverse<false, false>(sc, r, S, t, u, v, tc, ctr(q, bounce, 3));
        if (CURVES && sc.hasCurves) {
            float best = h != ~0u ? t : r.tMax;
            CurveHit ch;
            if (traverseCurves<false, false>(sc, r, S, best, ch, tc, ctr(q, bounce, 3))) {
                h = kCurveHit | ch.rec;
                t = best;
                u = ch.tSeg;
                v = 0.f;
                q.hitAux[i] = f4(ch.tang, 0.f);
            }
        }
        q.hit[i] = make_float4(t, u, v, __uint_as_float(h));
    }
}

// --------------------------------------------------------------------------- shade
// HAIRM: hair materials evaluate the Marschner model (AHIP_FLAG_HAIR_MARSCHNER); TRANSP: the scene
// has glass or alpha-masked materials (the alpha test, the first-hit AOV flag and the glass BSDF);
// the instances without them carry none of that code.
template <bool CURVES, bool HAIRM, bool TRANSP>
__global__ __launch_bounds__(256, 4) void k_shade(DScene sc, RenderConsts pd, WaveQueues q, uint32_t bounce,
                                               uint32_t cur) {
    __shared__ uint32_t lds[3 * (4 + 1)];
    uint32_t* const ctrs[3] = {ctr(q, bounce + 1u, 0), ctr(q, bounce, 1), ctr(q, bounce, 2)};
    const uint32_t count = *ctr(q, bounce, 0);
    const uint32_t nxt = cur ^ 1u;
    const float nL = (float)sc.numLights;
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += gridDim.x * blockDim.x) {
        uint32_t i = base + threadIdx.x;
        bool active = i < count && !(__float_as_uint(q.rd[cur][i].w) & kDeadPath);
        bool cont = false, wantShadow = false, wantMis = false;
        uint32_t p0 = 0, dim = 0;
        V3 beta{0.f, 0.f, 0.f}, rayO{0.f, 0.f, 0.f}, rayD{0.f, 0.f, 0.f};
        V3 shO{0.f, 0.f, 0.f}, shD{0.f, 0.f, 0.f}, shC{0.f, 0.f, 0.f};
        float shTmax = 0.f;
        V3 miO{0.f, 0.f, 0.f}, miD{0.f, 0.f, 0.f}, miC{0.f, 0.f, 0.f};
        bool specularNext = false, aovDone = false;
        if (active) {
            float4 ra = q.ro[cur][i], rb = q.rd[cur][i], bb = q.beta[cur][i], hh = q.hit[i];
            p0 = __float_as_uint(ra.w);
            uint32_t meta = __float_as_uint(rb.w);
            dim = meta & ~kPathFlags;
            bool specular = (meta & kSpecularPath) != 0u;
            aovDone = TRANSP ? (meta & kAovDone) != 0u : bounce != 0u;
            beta = v3(bb);
            V3 dir = v3(rb);
            uint32_t px, py, sidx;
            uint2 pp = q.pix[p0];
            px = pp.x & 0xffffu;
            py = pp.x >> 16;
            sidx = pp.y;
            uint32_t tri = __float_as_uint(hh.w);
            V3 Ladd{0.f, 0.f, 0.f};
            if (tri == ~0u) {
                if (bounce == 0u || specular) Ladd = beta * envLe(sc, dir);   // PathIntegrator.cpp:83-91
            } else {
                V3 p, n, ng;
                V2 uv;
                uint32_t meshID, commit = ~0u;
                int32_t mi = -1;
                bool isCurve = CURVES && (tri & kCurveHit) != 0u;
                Frame fr;
                if (!isCurve) {
                    // fillSurfaceInteraction (BVHBackend.cpp:285-317)
                    const DTri* T = sc.tris + tri;
                    float u = hh.y, v = hh.z, w = 1.f - u - v;
                    V3 v0 = ld3(T->v0), e1 = ld3(T->e1), e2 = ld3(T->e2);
                    ng = ld3(T->n);
                    meshID = T->meshID;
                    commit = T->commitIndex;
                    mi = T->matIndex;   // sc.meshMat[meshID], resolved at prepare time
                    const DTriShade* SH = sc.shade + tri;
                    p = v0 + e1 * u + e2 * v;
                    n = safeNormalize(ld3(SH->sn0) * w + ld3(SH->sn1) * u + ld3(SH->sn2) * v);
                    if (dot(n, ng) < 0.f) n = -n;
                    uv = interpUV(*SH, u, v, w);
                    fr = makeFrame(n, ng, dir);
                } else {
                    // curve hit record (CurveBrute.cpp:218-233)
                    uint32_t rec = tri & ~kCurveHit;
                    uint32_t strand = sc.segIds[2u * rec];
                    const float* info = sc.strandInfo + 4u * (size_t)strand;
                    V3 tang = v3(q.hitAux[i]);
                    V3 fx, fy;
                    onbNC(dir, fx, fy);
                    V3 nd = -dir;
                    V3 nrm = nd - tang * dot(nd, tang);
                    float ln = len(nrm);
                    nrm = ln > 1e-8f ? nrm * (1.f / ln) : fx;
                    p = v3(ra) + dir * hh.x;
                    n = nrm;
                    ng = nrm;
                    uv = V2{0.f, hh.y};
                    meshID = __float_as_uint(info[3]);
                    fr = makeCurveFrame(n, tang, dir, V3{info[0], info[1], info[2]});
                    mi = meshID < sc.numMeshes ? sc.meshMat[meshID] : -1;
                }
                bool passThrough = false;
                if (TRANSP && mi >= 0 && sc.hasAlpha) {
                    // alpha test (PathIntegrator.cpp:104-118): a cut-out point, or a partly opaque one
                    // whose opacity is below a sample (drawn only then), continues the ray past the
                    // surface unshaded, in the same bounce state
                    const DMat& m = sc.mats[mi];
                    const float op = matOpacity(sc, m, uv);
                    passThrough = op <= 0.f ||
                                  (op < 1.f && haltonSample(c_halton, pd.seed, px, py, sidx, dim++) >= op);
                    if (passThrough && bounce < pd.maxDepth) {
                        V3 off = ng * 1e-4f;                     // spawnRay(si.p, si.ng, r.direction)
                        if (dot(dir, ng) < 0.f) off = -off;
                        rayO = p + off;
                        rayD = dir;
                        specularNext = specular;
                        cont = true;
                    }
                }
                if (mi >= 0 && !passThrough) {
                    const DMat& m = sc.mats[mi];
                    const MatIn mIn = matInputs(sc, m, uv, commit);
                    V3 wo = -dir;
                    if (!aovDone && q.aovA) {   // first-hit AOVs (PathIntegrator.cpp:121-126)
                        q.aovA[p0] = f4(matReflectance(m, fr, mIn), 0.f);
                        q.aovN[p0] = f4(n, 0.f);
                    }
                    aovDone = true;
                    if (m.hasLe && (bounce == 0u || specular)) Ladd = beta * ld3(m.Le);   // :128-132
                    if (!m.isDelta && sc.numLights > 0u) {                                // :134-144
                        float ul = haltonSample(c_halton, pd.seed, px, py, sidx, dim++);
                        uint32_t li = (uint32_t)(ul * nL);
                        li = li < sc.numLights - 1u ? li : sc.numLights - 1u;
                        const DLight& L = sc.lights[li];
                        const Frame& fe = fr;   // estimateDirect's ShadingContext(si, -wo): the same frame
                        // light-sampling strategy (PathIntegrator.cpp:176-198)
                        V2 u2;
                        u2.x = haltonSample(c_halton, pd.seed, px, py, sidx, dim++);
                        u2.y = haltonSample(c_halton, pd.seed, px, py, sidx, dim++);
                        LSample ls = lightSample(sc, L, p, u2);
                        if (ls.pdf > 0.f && !isBlack(ls.Li)) {
                            float bpdf;
                            V3 bf = matEval<HAIRM, TRANSP>(sc, m, fe, mIn, wo, ls.wi, bpdf);
                            if (!isBlack(bf)) {
                                // spawnRayTo(si.p, si.n, si.p + wi*dist) (Types.h:277-284)
                                V3 target = p + ls.wi * ls.dist;
                                V3 sd = target - p;
                                float dist = len(sd);
                                sd = sd * (1.f / dist);
                                V3 off = n * 1e-4f;
                                if (dot(sd, n) < 0.f) off = -off;
                                // a delta light's sample has weight 1 (PathIntegrator.cpp:185-187)
                                float f = ls.pdf, g = bpdf;
                                float wgt = L.type == AHIP_LIGHT_DIRECTIONAL ? 1.f : (f * f) / (f * f + g * g);
                                // hair: sin of the angle to the fiber (PathIntegrator.cpp:189-193)
                                float dt = dot(ls.wi, fe.t);
                                float cosI = isCurve ? sqrtf(fmaxf(0.f, 1.f - dt * dt)) : fabsf(dot(ls.wi, n));
                                V3 Ld = bf * ls.Li * cosI * wgt / ls.pdf;
                                shO = p + off;
                                shD = sd;
                                shTmax = dist * (1.f - 1e-4f);
                                shC = beta * Ld * nL;
                                wantShadow = true;
                            }
                        }
                        // BSDF-sampling strategy (PathIntegrator.cpp:200-262), none for a delta
                        // light; g++ evaluates sample(ctx, wo, get2D(), get1D()) right to left:
                        // uComponent first.
                        if (L.type != AHIP_LIGHT_DIRECTIONAL) {
                            float uc = haltonSample(c_halton, pd.seed, px, py, sidx, dim++);
                            V2 ub;
                            ub.x = haltonSample(c_halton, pd.seed, px, py, sidx, dim++);
                            ub.y = haltonSample(c_halton, pd.seed, px, py, sidx, dim++);
                            BSample bs = matSample<HAIRM, TRANSP>(sc, m, fe, mIn, wo, ub, uc);
                            if (bsValid(bs)) {
                                float lp = lightPdf(sc, L, p, bs.wi);
                                if (lp > 0.f) {
                                    float f = bs.pdf, g = lp;
                                    float wgt = bs.delta ? 1.f : (f * f) / (f * f + g * g);
                                    V3 off = ng * 1e-4f;
                                    if (dot(bs.wi, ng) < 0.f) off = -off;
                                    float cosI = (pd.flags & AHIP_FLAG_FIX_DOUBLE_COSINE) ? 1.f : fabsf(dot(bs.wi, n));
                                    miO = p + off;
                                    miD = bs.wi;
                                    miC = beta * (bs.f * cosI * wgt / bs.pdf) * nL;
                                    wantMis = true;
                                }
                            }
                        }
                    }
                    if (bounce < pd.maxDepth) {
                        // continuation (PathIntegrator.cpp:146-160)
                        float uc = haltonSample(c_halton, pd.seed, px, py, sidx, dim++);
                        V2 ub;
                        ub.x = haltonSample(c_halton, pd.seed, px, py, sidx, dim++);
                        ub.y = haltonSample(c_halton, pd.seed, px, py, sidx, dim++);
                        BSample bs = matSample<HAIRM, TRANSP>(sc, m, fr, mIn, wo, ub, uc);
                        if (bsValid(bs)) {
                            specularNext = bs.delta;
                            beta = beta * (bs.f / bs.pdf);
                            V3 off = n * 1e-4f;
                            if (dot(bs.wi, n) < 0.f) off = -off;
                            rayO = p + off;
                            rayD = bs.wi;
                            cont = true;
                            if (bounce >= pd.minDepth) {
                                float qq = 1.f - stdmin(maxc(beta), 0.95f);
                                if (haltonSample(c_halton, pd.seed, px, py, sidx, dim++) < qq) cont = false;
                                else beta = beta * (1.f / (1.f - qq));
                            }
                        }
                    }
                }
            }
            if (!isBlack(Ladd)) {
                float4 l = q.L[p0];
                q.L[p0] = make_float4(l.x + Ladd.x, l.y + Ladd.y, l.z + Ladd.z, 0.f);
            }
        }
        const bool want[3] = {cont, wantShadow, wantMis};
        uint32_t slots[3];
        blockAppend<3, 4>(ctrs, want, slots, lds);
        const uint32_t sc0 = slots[0], sc1 = slots[1], sc2 = slots[2];
        if (cont) {
            q.ro[nxt][sc0] = f4(rayO, __uint_as_float(p0));
            q.rd[nxt][sc0] = f4(rayD, __uint_as_float(dim | (specularNext ? kSpecularPath : 0u) |
                                                      (TRANSP && aovDone ? kAovDone : 0u)));
            q.beta[nxt][sc0] = f4(beta, 0.f);
        }
        if (wantShadow) {
            q.so[sc1] = f4(shO, __uint_as_float(p0));
            q.sd[sc1] = f4(shD, shTmax);
            q.sc[sc1] = f4(shC, 0.f);
        }
        if (wantMis) {
            q.mo[sc2] = f4(miO, __uint_as_float(p0));
            q.md[sc2] = f4(miD, 1e10f);
            q.mc[sc2] = f4(miC, 0.f);
        }
    }
}

// --------------------------------------------------------------------------- queries, curve scenes
// Through dense hair a ray's curve-tree walk reaches a hundred or more segment records whose boxes
// it crosses before its hit, most of them rejected by the first ray-space hull test and a few
// subdivided ten or more frames deep, so one ray per lane leaves most lanes idle.  In curve scenes
// the three queries of a bounce (the paths' closest hit, the shadow rays' any hit, the
// BSDF-sampled rays' closest or any hit) run in two kernels:
//   k_curve_tri  — the triangle tree, one ray per lane, as CurveBrute::trace does first; the hit
//                  (or tMax) bounds the curve part, and rays whose query is settled are marked;
//   k_curve_pool — one wave holds up to 64 rays, one per lane (the ray's "slot").  Each lane
//                  walks its ray's curve tree, nearer child first, and puts the records of the
//                  leaves it reaches into the wave's pool; every lane takes records from the pool,
//                  whichever ray they belong to, and advances one subdivision frame per step, taking
//                  the next record when its frame walk ends.  A ray's best hit so far is a 64-bit
//                  key (t, record) in LDS that every frame test and node test is bounded by; when a
//                  ray's walk is over and none of its records is pending, its lane writes the result
//                  and takes the next ray.
// The result is the reference's (CurveBrute.cpp:398-440).  Closest hit: the reference's walk
// visits records in increasing record order (the builder lays leaves out depth-first, left before
// right) and keeps a hit only if strictly nearer, so it returns the least (t, record) over the
// accepted chords, and within that record the first chord (in its recursion order) at that t.
// Box, hull and pop culls are conservative (a culled box or sub-curve holds no chord nearer than
// the bound it was culled with), so testing records in another order under any bound no smaller
// than the final one finds the same least key: a record below the key's record is tested against
// the next float above the key's t (an equal t still wins), any other record strictly below it.
// Any hit: a record's test under tMax does not depend on the order of tests, and no cull uses
// anything but tMax before the first hit.
enum : uint32_t { kQPath = 0u, kQShadow = 1u, kQMis = 2u };   // counter slot of each queue

template <uint32_t Q>
__device__ __forceinline__ TRay curveQueryRay(const WaveQueues& q, uint32_t cur, uint32_t i, bool& dead) {
    float4 a, b;
    float tMax;
    if (Q == kQPath) {
        a = q.ro[cur][i];
        b = q.rd[cur][i];
        tMax = __int_as_float(0x7f800000);
        dead = (__float_as_uint(b.w) & kDeadPath) != 0u;
    } else if (Q == kQShadow) {
        a = q.so[i];
        b = q.sd[i];
        tMax = b.w;
        dead = false;
    } else {
        a = q.mo[i];
        b = q.md[i];
        tMax = b.w;
        dead = false;
    }
    return makeTRay(v3(a), v3(b), kCurveTMin, tMax);
}
__device__ __forceinline__ float hitBound(const float4& h, float tMax) {
    return __float_as_uint(h.w) != ~0u ? h.x : tMax;
}

// The triangle part of the three queries; q.cState[i] = 1 when the ray's curve walk is to run.
template <uint32_t BS, uint32_t K, uint32_t Q, bool ANY>
__global__ __launch_bounds__(BS) void k_curve_tri(DScene sc, WaveQueues q, uint32_t bounce, uint32_t cur) {
    __shared__ uint64_t lstk[K * BS];
    TStack S{lstk, q.ovfRef, q.ovfT, threadIdx.x, BS, K, blockIdx.x * BS + threadIdx.x, gridDim.x * BS,
             q.stackCap};
    const uint32_t count = *ctr(q, bounce, Q);
    TraceCounters tc{0, 0};
    for (uint32_t base = blockIdx.x * BS; base < count; base += gridDim.x * BS) {
        const uint32_t i = base + threadIdx.x;
        if (i >= count) break;
        bool dead;
        const TRay r = curveQueryRay<Q>(q, cur, i, dead);
        if (dead) {
            q.hit[i] = make_float4(0.f, 0.f, 0.f, __uint_as_float(~0u));
            q.cState[i] = 0u;
            continue;
        }
        float t = 0.f, u = 0.f, v = 0.f;
        const uint32_t h = traverse<ANY, false>(sc, r, S, t, u, v, tc, ctr(q, bounce, 3));
        q.hit[i] = make_float4(t, u, v, __uint_as_float(h));
        q.cState[i] = (ANY && h != ~0u) ? 0u : 1u;   // occluded by a triangle: no curve walk
    }
}

__device__ __forceinline__ uint32_t laneRank(uint64_t mask) {   // set bits of `mask` below this lane
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
template <typename T>
__device__ __forceinline__ T ldsLoad(T* p) {
    return __atomic_load_n(p, __ATOMIC_RELAXED);
}

// The subdivision of a pooled record without a private stack.  Pending right halves wait in LDS,
// one slot per level (the recursion keeps at most one per level; the leaf level's, taken right
// after its left sibling, in registers), each as its first three control points and its first
// half width: a pending half's last control point and last half width are
// the first ones of the next shallower pending half (or the segment's own), because a half
// shares its end with its parent and descending right keeps the parent's end.  Strand parameters
// are formed only for an accepted chord, by the reference's sequence of halvings and additions
// along the path of left / right turns (CurveBrute.cpp:150-172).
constexpr uint32_t kSubComps = 10;   // P0, P1, P2, hw0 of a pending half
typedef __attribute__((address_space(3))) float lds_f32;
struct PoolSeg {
    V3 P[4];          // the frame being tested
    float hw0, hw3;
    uint32_t level;   // splits from the segment
    uint32_t path;    // bit i - 1 set: the right half was taken at level i
    uint32_t pend;    // bit i set: a right half of level i waits in LDS
    bool leaf;        // a right half of the leaf level waits in registers (lP*, lHw0)
    V3 lP0, lP1, lP2;
    float lHw0;
};
struct SegRoot {
    V3 P3;
    float hw3, tOff, tScale;
};
__device__ __forceinline__ uint32_t subIdx(uint32_t level, uint32_t comp, uint32_t lane) {
    return ((level - 1u) * kSubComps + comp) * 64u + lane;
}

// One frame of the pooled subdivision: the test of segmentStep; a frame that passes and is not a
// leaf leaves its right half in LDS and becomes its left half; otherwise the deepest pending half
// is taken.  Returns true when no frame is left.
__device__ __forceinline__ bool poolSegStep(PoolSeg& s, const SegRoot& rt, lds_f32* stk, uint32_t lane, V3 o, V3 d,
                                            V3 fx, V3 fy, float best, uint32_t depth, bool& acc, float& tr,
                                            float& tSeg, V3& tang) {
#pragma clang fp contract(off)
    acc = false;
    {
        const V3 c0 = projNC(s.P[0], o, fx, fy, d), c1 = projNC(s.P[1], o, fx, fy, d);
        const V3 c2 = projNC(s.P[2], o, fx, fy, d), c3 = projNC(s.P[3], o, fx, fy, d);
        const float mh = fmaxf(s.hw0, s.hw3);
        const float zMin = fminf(fminf(fminf(c0.z, c1.z), c2.z), c3.z);
        const float zMax = fmaxf(fmaxf(fmaxf(c0.z, c1.z), c2.z), c3.z);
        if (zMin - mh > best) goto next;
        if (zMax + mh < kCurveTMin) goto next;
        const float xMin = fminf(fminf(fminf(c0.x, c1.x), c2.x), c3.x), xMax = fmaxf(fmaxf(fmaxf(c0.x, c1.x), c2.x), c3.x);
        const float yMin = fminf(fminf(fminf(c0.y, c1.y), c2.y), c3.y), yMax = fmaxf(fmaxf(fmaxf(c0.y, c1.y), c2.y), c3.y);
        const float cx = (0.f < xMin) ? xMin : (0.f > xMax ? xMax : 0.f);
        const float cy = (0.f < yMin) ? yMin : (0.f > yMax ? yMax : 0.f);
        if (cx * cx + cy * cy > mh * mh) goto next;
        if (s.level == depth) {
            const float ax = c0.x, ay = c0.y, dx = c3.x - ax, dy = c3.y - ay;
            const float l2 = dx * dx + dy * dy;
            const float tl = l2 < 1e-12f ? 0.f : stdmax(0.f, stdmin(1.f, (-ax * dx - ay * dy) / l2));
            const float px = ax + dx * tl, py = ay + dy * tl;
            const float hw = s.hw0 * (1.f - tl) + s.hw3 * tl;
            if (px * px + py * py > hw * hw) goto next;
            const float t = c0.z + (c3.z - c0.z) * tl;
            if (t < kCurveTMin || t >= best) goto next;
            // bezierTangent (CurveBrute.cpp:59-64)
            const float mt = 1.f - tl;
            const float a0 = 3.f * mt * mt, a1 = 6.f * mt * tl, a2 = 3.f * tl * tl;
            const V3 d10{s.P[1].x - s.P[0].x, s.P[1].y - s.P[0].y, s.P[1].z - s.P[0].z};
            const V3 d21{s.P[2].x - s.P[1].x, s.P[2].y - s.P[1].y, s.P[2].z - s.P[1].z};
            const V3 d32{s.P[3].x - s.P[2].x, s.P[3].y - s.P[2].y, s.P[3].z - s.P[2].z};
            V3 tg{d10.x * a0 + d21.x * a1 + d32.x * a2, d10.y * a0 + d21.y * a1 + d32.y * a2,
                  d10.z * a0 + d21.z * a1 + d32.z * a2};
            if (tg.x * tg.x + tg.y * tg.y + tg.z * tg.z < 1e-12f)
                tg = V3{s.P[3].x - s.P[0].x, s.P[3].y - s.P[0].y, s.P[3].z - s.P[0].z};
            const float l = sqrtf(tg.x * tg.x + tg.y * tg.y + tg.z * tg.z);
            if (l > 1e-8f) {
                const float il = 1.f / l;
                tg = V3{tg.x * il, tg.y * il, tg.z * il};
            } else {
                tg = V3{0.f, 0.f, 1.f};
            }
            // the leaf's parameter offset and width: tOff + tScale/2 at each right turn, in order
            float off = rt.tOff, sc = rt.tScale;
            for (uint32_t i = 0; i < s.level; ++i) {
                sc = sc * 0.5f;
                if ((s.path >> i) & 1u) off = off + sc;
            }
            acc = true;
            tr = t;
            tSeg = off + tl * sc;
            tang = tg;
            goto next;
        }
        const V3 q01 = midNC(s.P[0], s.P[1]), q12 = midNC(s.P[1], s.P[2]), q23 = midNC(s.P[2], s.P[3]);
        const V3 r01 = midNC(q01, q12), r12 = midNC(q12, q23), m = midNC(r01, r12);
        const float hm = (s.hw0 + s.hw3) * 0.5f;
        const uint32_t lv = s.level + 1u;
        if (lv == depth) {   // its left half is a leaf: the right half is taken next
            s.leaf = true;
            s.lP0 = m;
            s.lP1 = r12;
            s.lP2 = q23;
            s.lHw0 = hm;
        } else {
        stk[subIdx(lv, 0, lane)] = m.x;
        stk[subIdx(lv, 1, lane)] = m.y;
        stk[subIdx(lv, 2, lane)] = m.z;
        stk[subIdx(lv, 3, lane)] = r12.x;
        stk[subIdx(lv, 4, lane)] = r12.y;
        stk[subIdx(lv, 5, lane)] = r12.z;
        stk[subIdx(lv, 6, lane)] = q23.x;
        stk[subIdx(lv, 7, lane)] = q23.y;
        stk[subIdx(lv, 8, lane)] = q23.z;
        stk[subIdx(lv, 9, lane)] = hm;
        s.pend |= 1u << lv;
        }
        s.P[1] = q01;
        s.P[2] = r01;
        s.P[3] = m;
        s.hw3 = hm;
        s.level = lv;
        s.path &= ~(1u << (lv - 1u));
        return false;
    }
next:
    uint32_t k;
    if (s.leaf) {
        k = depth;
        s.P[0] = s.lP0;
        s.P[1] = s.lP1;
        s.P[2] = s.lP2;
        s.hw0 = s.lHw0;
        s.leaf = false;
    } else {
        if (s.pend == 0u) return true;
        k = 31u - (uint32_t)__clz(s.pend);
        s.P[0] = V3{stk[subIdx(k, 0, lane)], stk[subIdx(k, 1, lane)], stk[subIdx(k, 2, lane)]};
        s.P[1] = V3{stk[subIdx(k, 3, lane)], stk[subIdx(k, 4, lane)], stk[subIdx(k, 5, lane)]};
        s.P[2] = V3{stk[subIdx(k, 6, lane)], stk[subIdx(k, 7, lane)], stk[subIdx(k, 8, lane)]};
        s.hw0 = stk[subIdx(k, 9, lane)];
        s.pend &= ~(1u << k);
    }
    if (s.pend != 0u) {
        const uint32_t k2 = 31u - (uint32_t)__clz(s.pend);
        s.P[3] = V3{stk[subIdx(k2, 0, lane)], stk[subIdx(k2, 1, lane)], stk[subIdx(k2, 2, lane)]};
        s.hw3 = stk[subIdx(k2, 9, lane)];
    } else {
        s.P[3] = rt.P3;
        s.hw3 = rt.hw3;
    }
    s.level = k;
    s.path = (s.path & ((1u << (k - 1u)) - 1u)) | (1u << (k - 1u));
    return false;
}

// One wave per block (the pool, the slots and the keys are the wave's own, so no barrier is
// needed: a wave's LDS accesses take effect in program order).  K: LDS entries of each lane's walk
// stack, deeper entries in q.cwRef / q.cwT.
template <uint32_t K, uint32_t Q, bool ANY>
__global__ __launch_bounds__(64) void k_curve_pool(DScene sc, WaveQueues q, uint32_t bounce, uint32_t cur) {
    __shared__ uint64_t lstk[K * 64];
    __shared__ float4 sO[64], sD[64];          // slot rays
    __shared__ float4 sRes[64];                // tangent, strand parameter of the slot's key
    __shared__ unsigned long long sKey[64];    // t bits << 32 | record; any hit: 0 once occluded
    __shared__ uint32_t sOut[64];              // the slot's records in the pool or being tested
    __shared__ uint32_t sPool[kPoolCap];       // slot << kPoolSlotShift | record
    __shared__ float sSub[(kMaxCurveDepth - 1u) * kSubComps * 64];   // pending right halves above the leaf level
    lds_f32* const stk = (lds_f32*)sSub;
    const uint32_t lane = threadIdx.x;
    const TStack S{lstk, q.cwRef, q.cwT, lane, 64u, K, blockIdx.x * 64u + lane, gridDim.x * 64u, q.cwCap};
    const uint32_t count = *ctr(q, bounce, Q);
    uint32_t* const overflow = ctr(q, bounce, 3);
    const float inf = __int_as_float(0x7f800000);
    const uint32_t depth = sc.curveDepth;

    // this lane's slot
    uint32_t ray = ~0u, wsp = 0u;
    bool walking = false;
    V3 wo{0.f, 0.f, 0.f}, wd{0.f, 0.f, 1.f}, winv{0.f, 0.f, 0.f};
    unsigned long long initKey = 0ull;
    // the record this lane tests
    bool busy = false;
    uint32_t tSlot = 0u, tRec = 0u;
    V3 to{0.f, 0.f, 0.f}, td{0.f, 0.f, 1.f}, fx{0.f, 0.f, 0.f}, fy{0.f, 0.f, 0.f};
    PoolSeg ps;
    SegRoot rt;
    float own = inf;
    // wave-uniform: rays blockIdx.x + gridDim.x * k for k < nextK have been taken; the pool holds
    // `avail` records from `head` on
    uint32_t nextK = 0u, head = 0u, avail = 0u;

    for (uint32_t it = 0;; ++it) {
        // 1. a slot whose walk is over and whose records are all tested hands back its result
        if (ray != ~0u && !walking && ldsLoad(&sOut[lane]) == 0u) {
            const unsigned long long key = ldsLoad(&sKey[lane]);
            if (key != initKey) {
                if (ANY) {
                    q.hit[ray] = make_float4(0.f, 0.f, 0.f, __uint_as_float(kCurveHit));
                } else {
                    const float4 res = sRes[lane];
                    q.hit[ray] = make_float4(__uint_as_float((uint32_t)(key >> 32)), res.w, 0.f,
                                             __uint_as_float(kCurveHit | (uint32_t)key));
                    q.hitAux[ray] = make_float4(res.x, res.y, res.z, 0.f);
                }
            }
            ray = ~0u;
        }
        // empty slots take the next rays whose curve walk is to run
        while (true) {
            const uint64_t want = __ballot(ray == ~0u);
            const uint64_t first = (uint64_t)blockIdx.x + (uint64_t)gridDim.x * nextK;
            if (want == 0ull || first >= count) break;
            if (ray == ~0u) {
                const uint64_t i = first + (uint64_t)gridDim.x * laneRank(want);
                if (i < count && q.cState[i] != 0u) {
                    bool dead;
                    const TRay r = curveQueryRay<Q>(q, cur, (uint32_t)i, dead);
                    const float b0 = hitBound(q.hit[i], r.tMax);
                    ray = (uint32_t)i;
                    initKey = (unsigned long long)__float_as_uint(b0) << 32;
                    sO[lane] = f4(r.o, 0.f);
                    sD[lane] = f4(r.d, 0.f);
                    sRes[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
                    sKey[lane] = initKey;
                    sOut[lane] = 0u;
                    wo = r.o;
                    wd = r.d;
                    winv = V3{1.f / r.d.x, 1.f / r.d.y, 1.f / r.d.z};   // used only where |d| >= 1e-9
                    wsp = 0u;
                    float te;
                    const CNode root = loadCNode(sc.cnodes, 0u);
                    if (curveBoxT(root.bmin, root.bmax, wo, wd, winv, kCurveTMin, b0, te))
                        spush(S, wsp, (root.rc & 0x80000000u) ? kCWLeaf | ((root.rc - 1u) & 3u) << 29 | root.lp : 0u, te);
                    walking = wsp > 0u;
                }
            }
            nextK += (uint32_t)__popcll(want);
        }
        // 2. when the pool runs low, every walking lane takes one step of its walk
        if (avail < kPoolLow) {
            uint32_t n = 0u, lp = 0u;
            if (walking) {
                const unsigned long long key = ldsLoad(&sKey[lane]);
                if (ANY && key != initKey) {   // occluded: the walk is over
                    walking = false;
                } else {
                    const float bound = __uint_as_float((uint32_t)(key >> 32));
                    uint32_t ref;
                    float te;
                    spop(S, wsp, ref, te);
                    if (te > bound) {
                        // nearer hits were found since the entry was pushed
                    } else if (ref & kCWLeaf) {
                        n = ((ref >> 29) & 3u) + 1u;
                        lp = ref & ((1u << 29) - 1u);
                    } else if (wsp + 2u > S.cap) {
                        *overflow = 1u;
                        wsp = 0u;
                    } else {
                        const CWide w = loadCWide(sc.cwide, ref);
                        float ta = 0.f, tb = 0.f;
                        const bool ha = curveBoxT(w.a, w.a + 3, wo, wd, winv, kCurveTMin, bound, ta);
                        const bool hb = curveBoxT(w.b, w.b + 3, wo, wd, winv, kCurveTMin, bound, tb);
                        if (ha && hb) {   // the nearer child is taken next (the left one on a tie)
                            if (tb < ta) {
                                spush(S, wsp, w.refA, ta);
                                spush(S, wsp, w.refB, tb);
                            } else {
                                spush(S, wsp, w.refB, tb);
                                spush(S, wsp, w.refA, ta);
                            }
                        } else if (ha) {
                            spush(S, wsp, w.refA, ta);
                        } else if (hb) {
                            spush(S, wsp, w.refB, tb);
                        }
                    }
                    walking = wsp > 0u;
                }
            }
            // the leaf's records go to the pool at a wave prefix of the counts (at most 7 each)
            const uint64_t b0 = __ballot((n & 1u) != 0u), b1 = __ballot((n & 2u) != 0u), b2 = __ballot((n & 4u) != 0u);
            const uint32_t pre = laneRank(b0) + 2u * laneRank(b1) + 4u * laneRank(b2);
            const uint32_t tot = (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1) + 4u * (uint32_t)__popcll(b2);
            for (uint32_t k = 0; k < n; ++k)
                sPool[(head + avail + pre + k) & (kPoolCap - 1u)] = (lane << kPoolSlotShift) | (lp + k);
            if (n) atomicAdd(&sOut[lane], n);
            avail += tot;
        }
        // 3. lanes without a record take the next ones of the pool
        {
            const uint64_t idle = __ballot(!busy);
            const uint32_t take = min((uint32_t)__popcll(idle), avail);
            if (!busy) {
                const uint32_t rank = laneRank(idle);
                if (rank < take) {
                    const uint32_t item = ldsLoad(&sPool[(head + rank) & (kPoolCap - 1u)]);
                    tSlot = item >> kPoolSlotShift;
                    tRec = item & ((1u << kPoolSlotShift) - 1u);
                    to = v3(sO[tSlot]);
                    td = v3(sD[tSlot]);
                    onbNC(td, fx, fy);
                    const DSeg& g = sc.segs[tRec];
                    ps.P[0] = V3{g.P[0], g.P[1], g.P[2]};
                    ps.P[1] = V3{g.P[3], g.P[4], g.P[5]};
                    ps.P[2] = V3{g.P[6], g.P[7], g.P[8]};
                    ps.P[3] = V3{g.P[9], g.P[10], g.P[11]};
                    ps.hw0 = g.hw0;
                    ps.hw3 = g.hw3;
                    ps.level = ps.path = ps.pend = 0u;
                    ps.leaf = false;
                    rt = SegRoot{ps.P[3], g.hw3, g.tOff, g.tScale};
                    own = inf;
                    busy = true;
                }
            }
            head += take;
            avail -= take;
        }
        // 4. one subdivision frame per busy lane
        if (busy) {
            const unsigned long long key = ldsLoad(&sKey[tSlot]);
            bool stop;
            if (ANY && key == 0ull) {
                stop = true;
            } else {
                const float kt = __uint_as_float((uint32_t)(key >> 32));
                const float thr = (tRec < (uint32_t)key && kt < inf) ? __uint_as_float(__float_as_uint(kt) + 1u) : kt;
                bool acc;
                float t, ts;
                V3 tg;
                stop = poolSegStep(ps, rt, stk, lane, to, td, fx, fy, fminf(own, thr), depth, acc, t, ts, tg);
                if (acc) {
                    own = t;
                    const unsigned long long mine = ANY ? 0ull : (((unsigned long long)__float_as_uint(t) << 32) | tRec);
                    atomicMin(&sKey[tSlot], mine);
                    // the lanes of this step that hold the slot's least key write its tangent
                    if (!ANY && ldsLoad(&sKey[tSlot]) == mine) sRes[tSlot] = f4(tg, ts);
                    if (ANY) stop = true;
                }
            }
            if (stop) {
                busy = false;
                atomicSub(&sOut[tSlot], 1u);
            }
        }
        const bool rest = (uint64_t)blockIdx.x + (uint64_t)gridDim.x * nextK < count;
        if (__ballot(ray != ~0u || busy) == 0ull && avail == 0u && !rest) break;
        if (it >= kPoolIterCap) {   // never reached by a scene the tree and the pool can hold
            *overflow = 1u;
            break;
        }
    }
}

// The shadow and BSDF-sampled queues' results, as k_shadow and k_mis apply them.
__global__ __launch_bounds__(256) void k_shadow_resolve(WaveQueues q, uint32_t bounce) {
    const uint32_t count = *ctr(q, bounce, kQShadow);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        if (__float_as_uint(q.hit[i].w) != ~0u) continue;
        const uint32_t p0 = __float_as_uint(q.so[i].w);
        const float4 c = q.sc[i], l = q.L[p0];
        q.L[p0] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.f);
    }
}
template <bool EMIT>
__global__ __launch_bounds__(256) void k_mis_resolve(DScene sc, WaveQueues q, uint32_t bounce) {
    const uint32_t count = *ctr(q, bounce, kQMis);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t h = __float_as_uint(q.hit[i].w);
        const V3 d = v3(q.md[i]);
        V3 Li{0.f, 0.f, 0.f};
        if (h == ~0u) Li = envLe(sc, d);
        else if (EMIT) {
            const uint32_t mesh = (h & kCurveHit)
                ? __float_as_uint(sc.strandInfo[4u * (size_t)sc.segIds[2u * (h & ~kCurveHit)] + 3u])
                : sc.tris[h].meshID;
            const int32_t mi = mesh < sc.numMeshes ? sc.meshMat[mesh] : -1;
            if (mi >= 0 && sc.mats[mi].hasLe) Li = ld3(sc.mats[mi].Le);
        }
        if (!isBlack(Li)) {
            const uint32_t p0 = __float_as_uint(q.mo[i].w);
            const float4 c = q.mc[i], l = q.L[p0];
            const V3 add = v3(c) * Li;
            q.L[p0] = make_float4(l.x + add.x, l.y + add.y, l.z + add.z, 0.f);
        }
    }
}

// --------------------------------------------------------------------------- shadow (any hit)
template <uint32_t BS, uint32_t K, bool CURVES>
__global__ __launch_bounds__(BS) void k_shadow(DScene sc, WaveQueues q, uint32_t bounce) {
    __shared__ uint64_t lstk[K * BS];
    TStack S{lstk, q.ovfRef, q.ovfT, threadIdx.x, BS, K, blockIdx.x * BS + threadIdx.x, gridDim.x * BS,
             q.stackCap};
    const uint32_t count = *ctr(q, bounce, 1);
    TraceCounters tc{0, 0};
    for (uint32_t base = blockIdx.x * BS; base < count; base += gridDim.x * BS) {
        uint32_t i = base + threadIdx.x;
        if (i >= count) break;
        float4 a = q.so[i], b = q.sd[i];
        TRay r = makeTRay(v3(a), v3(b), 1e-4f, b.w);
        float t, u, v;
        uint32_t h = traverse<true, false>(sc, r, S, t, u, v, tc, ctr(q, bounce, 3));
        if (CURVES && h == ~0u && sc.hasCurves) {
            float best = r.tMax;
            CurveHit ch;
            if (traverseCurves<true, false>(sc, r, S, best, ch, tc, ctr(q, bounce, 3))) h = kCurveHit;
        }
        if (h == ~0u) {
            uint32_t p0 = __float_as_uint(a.w);
            float4 c = q.sc[i], l = q.L[p0];
            q.L[p0] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.f);
        }
    }
}

// --------------------------------------------------------------------------- MIS (BSDF-sampled NEE)
// EMIT = sc.hasEmitters, as a template parameter so the any-hit instance (no emitting surface)
// keeps only the registers of the any-hit traversal.
template <uint32_t BS, uint32_t K, bool CURVES, bool EMIT>
__global__ __launch_bounds__(BS) void k_mis(DScene sc, WaveQueues q, uint32_t bounce) {
    __shared__ uint64_t lstk[K * BS];
    TStack S{lstk, q.ovfRef, q.ovfT, threadIdx.x, BS, K, blockIdx.x * BS + threadIdx.x, gridDim.x * BS,
             q.stackCap};
    const uint32_t count = *ctr(q, bounce, 2);
    TraceCounters tc{0, 0};
    for (uint32_t base = blockIdx.x * BS; base < count; base += gridDim.x * BS) {
        uint32_t i = base + threadIdx.x;
        if (i >= count) break;
        float4 a = q.mo[i], b = q.md[i];
        V3 d = v3(b);
        TRay r = makeTRay(v3(a), d, 1e-4f, b.w);
        float t, u, v;
        V3 Li{0.f, 0.f, 0.f};
        if constexpr (EMIT) {
            uint32_t h = traverse<false, false>(sc, r, S, t, u, v, tc, ctr(q, bounce, 3));
            uint32_t mesh = h != ~0u ? sc.tris[h].meshID : ~0u;
            if (CURVES && sc.hasCurves) {
                float best = h != ~0u ? t : r.tMax;
                CurveHit ch;
                if (traverseCurves<false, false>(sc, r, S, best, ch, tc, ctr(q, bounce, 3))) {
                    h = kCurveHit | ch.rec;
                    mesh = __float_as_uint(sc.strandInfo[4u * (size_t)sc.segIds[2u * ch.rec] + 3u]);
                }
            }
            if (h == ~0u) Li = envLe(sc, d);
            else {
                int32_t mi = mesh < sc.numMeshes ? sc.meshMat[mesh] : -1;
                if (mi >= 0 && sc.mats[mi].hasLe) Li = ld3(sc.mats[mi].Le);
            }
        } else {
            // no surface emits: only an escaping ray carries radiance, so any hit decides
            uint32_t h = traverse<true, false>(sc, r, S, t, u, v, tc, ctr(q, bounce, 3));
            if (CURVES && h == ~0u && sc.hasCurves) {
                float best = r.tMax;
                CurveHit ch;
                if (traverseCurves<true, false>(sc, r, S, best, ch, tc, ctr(q, bounce, 3))) h = kCurveHit;
            }
            if (h == ~0u) Li = envLe(sc, d);
        }
        if (!isBlack(Li)) {
            uint32_t p0 = __float_as_uint(a.w);
            float4 c = q.mc[i], l = q.L[p0];
            V3 add = v3(c) * Li;
            q.L[p0] = make_float4(l.x + add.x, l.y + add.y, l.z + add.z, 0.f);
        }
    }
}

// --------------------------------------------------------------------------- transparent scenes
// (StepHit, stepClosest, shadowTransmittance: device_transmit.h)
template <uint32_t BS, uint32_t K, bool CURVES>
__global__ __launch_bounds__(BS) void k_shadow_tr(DScene sc, WaveQueues q, uint32_t bounce) {
    __shared__ uint64_t lstk[K * BS];
    TStack S{lstk, q.ovfRef, q.ovfT, threadIdx.x, BS, K, blockIdx.x * BS + threadIdx.x, gridDim.x * BS,
             q.stackCap};
    const uint32_t count = *ctr(q, bounce, 1);
    TraceCounters tc{0, 0};
    for (uint32_t base = blockIdx.x * BS; base < count; base += gridDim.x * BS) {
        uint32_t i = base + threadIdx.x;
        if (i >= count) break;
        const float4 a = q.so[i], b = q.sd[i];
        V3 T;
        if (shadowTransmittance<CURVES>(sc, v3(a), v3(b), b.w, S, tc, ctr(q, bounce, 3), T)) {
            const uint32_t p0 = __float_as_uint(a.w);
            const V3 add = v3(q.sc[i]) * T;
            const float4 l = q.L[p0];
            q.L[p0] = make_float4(l.x + add.x, l.y + add.y, l.z + add.z, 0.f);
        }
    }
}

template <uint32_t BS, uint32_t K, bool CURVES>
__global__ __launch_bounds__(BS) void k_mis_tr(DScene sc, WaveQueues q, uint32_t bounce) {
    __shared__ uint64_t lstk[K * BS];
    TStack S{lstk, q.ovfRef, q.ovfT, threadIdx.x, BS, K, blockIdx.x * BS + threadIdx.x, gridDim.x * BS,
             q.stackCap};
    const uint32_t count = *ctr(q, bounce, 2);
    TraceCounters tc{0, 0};
    for (uint32_t base = blockIdx.x * BS; base < count; base += gridDim.x * BS) {
        uint32_t i = base + threadIdx.x;
        if (i >= count) break;
        const float4 a = q.mo[i], b = q.md[i];
        V3 o = v3(a);
        const V3 d = v3(b);
        float tMax = b.w;
        V3 Tr{1.f, 1.f, 1.f};
        StepHit h;
        h.hit = false;
        for (int step = 0; step < 8; ++step) {                   // PathIntegrator.cpp:217-236
            h = stepClosest<CURVES>(sc, makeTRay(o, d, 1e-4f, tMax), S, tc, ctr(q, bounce, 3));
            if (!h.hit) break;
            if (h.mi < 0) { Tr = V3{0.f, 0.f, 0.f}; break; }
            const V3 tint = stepTint(sc, h);
            if (isBlack(tint)) break;                            // opaque or emitter: stop here
            Tr = Tr * tint;
            if (isBlack(Tr)) break;
            const float remaining = tMax - h.t;
            if (remaining <= 1e-4f) { h.hit = false; break; }
            o = stepOrigin(h, d);
            tMax = remaining - 1e-4f;
        }
        V3 Li{0.f, 0.f, 0.f};
        if (!h.hit) Li = envLe(sc, d);
        else if (h.mi >= 0 && sc.mats[h.mi].hasLe) Li = ld3(sc.mats[h.mi].Le);
        if (!isBlack(Li) && !isBlack(Tr)) {
            const uint32_t p0 = __float_as_uint(a.w);
            const V3 add = v3(q.mc[i]) * Li * Tr;
            const float4 l = q.L[p0];
            q.L[p0] = make_float4(l.x + add.x, l.y + add.y, l.z + add.z, 0.f);
        }
    }
}

// --------------------------------------------------------------------------- accumulate
// PathIntegrator::renderTile's per-pixel sums (PathIntegrator.cpp:27-57) for this pass.
__global__ __launch_bounds__(256) void k_accumulate(WaveQueues q) {
    const PassTable& T = q.tables[*q.passIdx];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < T.totalSlots; k += gridDim.x * blockDim.x) {
        const PassSeg& sg = T.seg[findSeg<true>(T, k)];
        const ViewDesc& V = q.views[sg.view];
        uint32_t slot = k - sg.slotBase, lx, ly;
        if (!slotPixel(V, slot, lx, ly)) continue;
        V3 acc{0.f, 0.f, 0.f};
        float lsq = 0.f;
        for (uint32_t j = 0; j < sg.samples; ++j) {
            float4 l = q.L[sg.pathBase + j * V.slots + slot];
            V3 L = v3(l);
            if (finite3(L)) {
                acc = acc + L;
                float lum = luminance(L);
                lsq += lum * lum;
            }
        }
        size_t o = (size_t)ly * V.stride + lx;
        float4 a = V.accum[o];
        V.accum[o] = make_float4(a.x + acc.x, a.y + acc.y, a.z + acc.z, a.w + (float)sg.samples);
        V.accumLumSq[o] += lsq;
        if (V.accumAov) {
            // first-hit reflectance and normal of every sample (PathIntegrator.cpp:45-47), in
            // sample order
            V3 sa{0.f, 0.f, 0.f}, sn{0.f, 0.f, 0.f};
            for (uint32_t j = 0; j < sg.samples; ++j) {
                sa = sa + v3(q.aovA[sg.pathBase + j * V.slots + slot]);
                sn = sn + v3(q.aovN[sg.pathBase + j * V.slots + slot]);
            }
            float* A = V.accumAov + 6u * o;
            A[0] += sa.x; A[1] += sa.y; A[2] += sa.z;
            A[3] += sn.x; A[4] += sn.y; A[5] += sn.z;
        }
    }
}

// --------------------------------------------------------------------------- adaptive sampling
// Per-tile score of the adaptive pass (RenderSession.cpp:592-600): Film::varianceAt (Film.cpp:80-90,
// Rec.709 weights on the mean, the ACEScg luminance-squared sum) summed over the tile's pixels in
// row-major order and divided by the pixel count.  One thread per tile keeps that sequential float
// order, and every product and sum is rounded on its own (no contraction), so the host can restate
// the score exactly.
__global__ __launch_bounds__(64) void k_tile_variance(const float4* acc, const float* lsq, uint32_t W, uint32_t H,
                                                      uint32_t ts, uint32_t tilesX, uint32_t nTiles, float* out) {
#pragma clang fp contract(off)
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nTiles) return;
    const uint32_t x0 = (t % tilesX) * ts, y0 = (t / tilesX) * ts;
    const uint32_t w = min(ts, W - x0), h = min(ts, H - y0);
    float sum = 0.f;
    for (uint32_t y = y0; y < y0 + h; ++y) {
        const float4* row = acc + (size_t)y * W;
        const float* lrow = lsq + (size_t)y * W;
        for (uint32_t x = x0; x < x0 + w; ++x) {
            const float4 a = row[x];
            float v = 0.f;
            if (a.w > 0.f) {
                const float invW = 1.f / a.w;
                const float meanL = (a.x * 0.2126f + a.y * 0.7152f + a.z * 0.0722f) * invW;
                const float eLsq = lrow[x] * invW;
                v = fmaxf(0.f, eLsq - meanL * meanL);
            }
            sum = sum + v;
        }
    }
    out[t] = sum / (float)(w * h);
}

// Film resolve of the accumulation (PixelAccumulator::resolve, Film.h:46-55; Film::varianceAt,
// Film.cpp:80-90; writePNG's exposure, clamp and sRGB encoding, Film.cpp:233-256; the AOV means of
// PathIntegrator::renderTile, PathIntegrator.cpp:58-66) in one pass over the buffers, each output
// optional.  Contraction is off: every product and sum rounds as written.
__global__ __launch_bounds__(256) void k_resolve(const float4* acc, const float* lsq, const float* aov, size_t n,
                                                 float evScale, uint8_t* srgb, float* rgb, float* var, float* alb,
                                                 float* nrm) {
#pragma clang fp contract(off)
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float4 a = acc[i];
        const bool has = a.w > 0.f;
        const float invW = has ? 1.f / a.w : 0.f;
        const float c[3] = {a.x * invW, a.y * invW, a.z * invW};
        if (rgb)
            for (int k = 0; k < 3; ++k) rgb[3 * i + k] = c[k];
        if (var) {
            float v = 0.f;
            if (has) {
                const float meanL = (a.x * 0.2126f + a.y * 0.7152f + a.z * 0.0722f) * invW;
                v = fmaxf(0.f, lsq[i] * invW - meanL * meanL);
            }
            var[i] = v;
        }
        if (srgb)
            for (int k = 0; k < 3; ++k) {
                const float x = fmaxf(0.f, fminf(1.f, c[k] * evScale));
                const float e = x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.f / 2.4f) - 0.055f;
                srgb[3 * i + k] = (uint8_t)(e * 255.f + 0.5f);
            }
        if (aov && (alb || nrm)) {
            const float* s = aov + 6 * i;
            float m[6];
            for (int k = 0; k < 6; ++k) m[k] = s[k] * invW;
            if (alb)
                for (int k = 0; k < 3; ++k) alb[3 * i + k] = m[k];
            if (nrm) {
                const float len = sqrtf(m[3] * m[3] + m[4] * m[4] + m[5] * m[5]);
                if (len > 1e-6f) {
                    const float il = 1.f / len;
                    m[3] = m[3] * il; m[4] = m[4] * il; m[5] = m[5] * il;
                }
                for (int k = 0; k < 3; ++k) nrm[3 * i + k] = m[3 + k];
            }
        }
    }
}

// --------------------------------------------------------------------------- material probe
// IMaterial::evaluate and sample of one material on independent rows (the layout of
// ahip_material_probe): the frame is the integrator's for a hit seen along -wo, a curve row takes
// its tangent as the fiber direction.  out: f rgb, pdf, 0 (pdfRev is BDPT's), sampled wi, f rgb,
// pdf, valid, hair effectiveSigmaA rgb.
template <bool HAIRM>
__global__ __launch_bounds__(256) void k_material_probe(DMat m, uint32_t n, const float* in, float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* q = in + 19u * (size_t)i;
    V3 nn{q[0], q[1], q[2]}, t{q[3], q[4], q[5]}, col{q[6], q[7], q[8]};
    V3 wo{q[9], q[10], q[11]}, wi{q[12], q[13], q[14]};
    V2 u{q[15], q[16]};
    const float uc = q[17];
    Frame fr = q[18] != 0.f ? makeCurveFrame(nn, t, -wo, col) : makeFrame(nn, nn, -wo);
    DScene sc{};
    const MatIn mi = matInputs(sc, m, V2{0.f, 0.f}, ~0u);
    float pdf;
    V3 f = matEval<HAIRM, true>(sc, m, fr, mi, wo, wi, pdf);
    BSample bs = matSample<HAIRM, true>(sc, m, fr, mi, wo, u, uc);
    DScene none{};   // the probe's materials read no texture
    V3 sa = m.type == AHIP_MAT_HAIR ? hairSigmaA(m, fr.color) : matTransmittance(none, m, mi.base, V2{0.f, 0.f});
    float* o = out + 16u * (size_t)i;
    const bool valid = bsValid(bs);
    o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = pdf; o[4] = 0.f;
    o[5] = bs.wi.x; o[6] = bs.wi.y; o[7] = bs.wi.z;
    o[8] = bs.f.x; o[9] = bs.f.y; o[10] = bs.f.z; o[11] = bs.pdf; o[12] = valid ? 1.f : 0.f;
    o[13] = sa.x; o[14] = sa.y; o[15] = sa.z;
}

// --------------------------------------------------------------------------- ray-batch entry points
template <uint32_t BS, uint32_t K, bool ANY>
__global__ __launch_bounds__(BS) void k_trace_batch(DScene sc, const float* rays, uint32_t n, ahip_hit* hits,
                                                    uint8_t* occ, uint32_t* ovfRef, float* ovfT, uint32_t cap,
                                                    uint32_t* overflow, uint32_t* counts) {
    __shared__ uint64_t lstk[K * BS];
    TStack S{lstk, ovfRef, ovfT, threadIdx.x, BS, K, blockIdx.x * BS + threadIdx.x, gridDim.x * BS, cap};
    for (uint32_t base = blockIdx.x * BS; base < n; base += gridDim.x * BS) {
        uint32_t i = base + threadIdx.x;
        if (i >= n) break;
        const float* q = rays + 8 * (size_t)i;
        TRay r = makeTRay(V3{q[0], q[1], q[2]}, V3{q[3], q[4], q[5]}, q[6], q[7]);
        float t = 0.f, u = 0.f, v = 0.f;
        TraceCounters tc{0, 0};
        uint32_t h;
        if (counts) h = traverse<ANY, true>(sc, r, S, t, u, v, tc, overflow);
        else h = traverse<ANY, false>(sc, r, S, t, u, v, tc, overflow);
        CurveHit ch;
        bool curve = false;
        if (sc.hasCurves && !(ANY && h != ~0u)) {
            float best = h != ~0u ? t : r.tMax;
            if (counts) curve = traverseCurves<ANY, true>(sc, r, S, best, ch, tc, overflow);
            else curve = traverseCurves<ANY, false>(sc, r, S, best, ch, tc, overflow);
            if (curve) t = best;
        }
        if (counts) { counts[2 * i] = tc.nodes; counts[2 * i + 1] = tc.tris; }
        if (ANY) {
            occ[i] = (h != ~0u || curve) ? 1 : 0;
        } else {
            ahip_hit o;
            if (curve) {
                uint32_t strand = sc.segIds[2u * ch.rec];
                uint32_t mesh = __float_as_uint(sc.strandInfo[4u * (size_t)strand + 3u]);
                o = ahip_hit{t, 0.f, ch.tSeg, mesh, sc.segIds[2u * ch.rec + 1u], kCurveHit | strand};
            } else if (h == ~0u) {
                o = ahip_hit{0.f, 0.f, 0.f, ~0u, ~0u, ~0u};
            } else {
                o = ahip_hit{t, u, v, sc.tris[h].meshID, sc.tris[h].primID, sc.tris[h].commitIndex};
            }
            hits[i] = o;
        }
    }
}

// --------------------------------------------------------------------------- traffic counting
// Counter-enabled copies of extend / shadow / mis: per-ray node and triangle visits summed with
// one atomic per wavefront.  Used by a separate, untimed pass.
template <uint32_t BS, uint32_t K, int KIND, bool CURVES>
__global__ __launch_bounds__(BS) void k_count(DScene sc, WaveQueues q, uint32_t bounce, uint32_t cur,
                                              unsigned long long* sums) {
    __shared__ uint64_t lstk[K * BS];
    TStack S{lstk, q.ovfRef, q.ovfT, threadIdx.x, BS, K, blockIdx.x * BS + threadIdx.x, gridDim.x * BS,
             q.stackCap};
    const uint32_t count = *ctr(q, bounce, KIND);
    for (uint32_t base = blockIdx.x * BS; base < count; base += gridDim.x * BS) {
        uint32_t i = base + threadIdx.x;
        TraceCounters tc{0, 0};
        if (i < count) {
            float4 a, b;
            if (KIND == 0) {
                a = q.ro[cur][i];
                b = q.rd[cur][i];
                b.w = (__float_as_uint(b.w) & kDeadPath) ? -1.f : __int_as_float(0x7f800000);
            }
            else if (KIND == 1) { a = q.so[i]; b = q.sd[i]; }
            else { a = q.mo[i]; b = q.md[i]; }
            TRay r = makeTRay(v3(a), v3(b), 1e-4f, b.w);
            float t, u, v;
            bool any = KIND == 1 || (KIND == 2 && !sc.hasEmitters);
            uint32_t h;
            if (any) h = traverse<true, true>(sc, r, S, t, u, v, tc, ctr(q, bounce, 3));
            else h = traverse<false, true>(sc, r, S, t, u, v, tc, ctr(q, bounce, 3));
            if (CURVES && sc.hasCurves && !(any && h != ~0u)) {
                float best = h != ~0u ? t : r.tMax;
                CurveHit ch;
                if (any) traverseCurves<true, true>(sc, r, S, best, ch, tc, ctr(q, bounce, 3));
                else traverseCurves<false, true>(sc, r, S, best, ch, tc, ctr(q, bounce, 3));
            }
        }
        unsigned long long nodes = tc.nodes, trisT = tc.tris;
        for (int off = 32; off > 0; off >>= 1) {
            nodes += __shfl_down(nodes, off);
            trisT += __shfl_down(trisT, off);
        }
        if (__lane_id() == 0u) {
            atomicAdd(&sums[0], nodes);
            atomicAdd(&sums[1], trisT);
        }
    }
}

// End of a pass: ray totals ([0] continuation, [1] shadow, [2] MIS rays, [3] stack overflows) and
// the device-side pass index advanced, so the next replay of the same launches runs the next pass.
// totals[0..3]: the pass's continuation, shadow and MIS rays and stack overflows summed over the
// bounces; totals[8 + 4 b + k]: the same per bounce b (the occupancy of each bounce's launches).
__global__ void k_pass_end(uint32_t* counters, uint32_t maxDepth, unsigned long long* totals, uint32_t* passIdx) {
    if (blockIdx.x != 0u) return;
    if (threadIdx.x < 4u) {
        unsigned long long s = 0ull;
        for (uint32_t b = 0; b <= maxDepth; ++b) s += counters[b * 4u + threadIdx.x];
        totals[threadIdx.x] += s;
    }
    for (uint32_t i = threadIdx.x; i < (maxDepth + 1u) * 4u; i += blockDim.x) totals[8u + i] += counters[i];
    __syncthreads();
    // counters of the next pass start at zero (a kernel, not a memset, so it can live in a graph)
    for (uint32_t i = threadIdx.x; i < (maxDepth + 2u) * 4u; i += blockDim.x) counters[i] = 0u;
    if (threadIdx.x == 0u) *passIdx += 1u;
}

// --------------------------------------------------------------------------- host-side launchers
void setHaltonTables(const HaltonTables& t, hipStream_t s) {
    hipMemcpyToSymbolAsync(HIP_SYMBOL(c_halton), &t, sizeof(t), 0, hipMemcpyHostToDevice, s);
}

constexpr uint32_t kTraceBS = 128;
constexpr uint32_t kStackK = 8;
// curve scenes walk a deep median-split curve BVH on the same stack: a deeper LDS part pays there
// (hair frame 4% faster at 12 entries, while the triangle-only kernels lose occupancy past 8-10)
constexpr uint32_t kStackKCurves = 12;

uint32_t lds_stack_entries() { return kStackK; }

void launchFillU32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s) {
    uint32_t blocks = (n + 255u) / 256u;
    if (blocks > 1024u) blocks = 1024u;
    if (blocks == 0u) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, n, v);
}
void launchTileVariance(const float4* acc, const float* lsq, uint32_t W, uint32_t H, uint32_t tile, float* out,
                        hipStream_t s) {
    const uint32_t tilesX = (W + tile - 1u) / tile, tilesY = (H + tile - 1u) / tile, n = tilesX * tilesY;
    if (n == 0u) return;
    hipLaunchKernelGGL(k_tile_variance, dim3((n + 63u) / 64u), dim3(64), 0, s, acc, lsq, W, H, tile, tilesX, n, out);
}
void launchResolve(const float4* acc, const float* lsq, const float* aov, size_t n, float evScale, uint8_t* srgb,
                   float* rgb, float* var, float* alb, float* nrm, hipStream_t s) {
    size_t blocks = std::min<size_t>((n + 255u) / 256u, 8192u);
    if (blocks == 0u) return;
    hipLaunchKernelGGL(k_resolve, dim3((uint32_t)blocks), dim3(256), 0, s, acc, lsq, aov, n, evScale, srgb, rgb, var,
                       alb, nrm);
}
void launchFillF32(float* p, size_t n, float v, hipStream_t s) {
    size_t blocks = (n + 255u) / 256u;
    if (blocks > 4096u) blocks = 4096u;
    if (blocks == 0u) return;
    hipLaunchKernelGGL(k_fill_f32, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, v);
}
void launchHalton(uint64_t seed, const uint32_t* px, const uint32_t* py, const uint32_t* idx, const uint32_t* dim,
                  uint32_t n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_halton_eval, dim3((n + 255u) / 256u), dim3(256), 0, s, seed, px, py, idx, dim, n, out);
}
void launchPassEnd(const WaveQueues& q, uint32_t maxDepth, unsigned long long* totals, hipStream_t s) {
    hipLaunchKernelGGL(k_pass_end, dim3(1), dim3(256), 0, s, q.counters, maxDepth, totals, q.passIdx);
}
uint32_t trace_block_size() { return kTraceBS; }
uint32_t curve_pool_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_curve_pool<kPoolStackK, kQPath, false>, 64, 0) != hipSuccess ||
        n <= 0)
        return 1u;
    return (uint32_t)n;
}
void launchGenerate(const RenderConsts& rc, const WaveQueues& q, uint32_t blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_generate, dim3(blocks), dim3(256), 0, s, rc, q);
}
template <uint32_t Q, bool ANY>
void launchCurveQuery(const DScene& sc, const WaveQueues& q, uint32_t bounce, uint32_t cur, uint32_t blocks,
                      hipStream_t s) {
    hipLaunchKernelGGL((k_curve_tri<kTraceBS, kStackK, Q, ANY>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce, cur);
    hipLaunchKernelGGL((k_curve_pool<kPoolStackK, Q, ANY>), dim3(q.poolBlocks), dim3(64), 0, s, sc, q, bounce, cur);
}
// The pooled curve queries need their buffers (q.cState); AHIP_DEBUG_WHOLE_RAY=1 selects the
// whole-ray kernels instead, for comparisons.
bool curveRounds(const DScene& sc, const WaveQueues& q) {
    const char* dbg = getenv("AHIP_DEBUG_WHOLE_RAY");
    return sc.hasCurves && q.cState && !(dbg && dbg[0] == '1');
}

void launchExtend(const DScene& sc, const WaveQueues& q, uint32_t bounce, uint32_t cur, uint32_t blocks,
                  hipStream_t s) {
    if (curveRounds(sc, q))
        launchCurveQuery<kQPath, false>(sc, q, bounce, cur, blocks, s);
    else if (sc.hasCurves)
        hipLaunchKernelGGL((k_extend<kTraceBS, kStackKCurves, true>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce, cur);
    else
        hipLaunchKernelGGL((k_extend<kTraceBS, kStackK, false>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce, cur);
}
void launchShade(const DScene& sc, const RenderConsts& rc, const WaveQueues& q, uint32_t bounce, uint32_t cur,
                 uint32_t blocks, hipStream_t s) {
    const bool hairm = sc.hasHair && (rc.flags & AHIP_FLAG_HAIR_MARSCHNER) != 0u;
    const dim3 g(blocks), b(256);
    if (sc.hasTransparent) {
        if (sc.hasCurves) {
            if (hairm) hipLaunchKernelGGL((k_shade<true, true, true>), g, b, 0, s, sc, rc, q, bounce, cur);
            else hipLaunchKernelGGL((k_shade<true, false, true>), g, b, 0, s, sc, rc, q, bounce, cur);
        } else {
            if (hairm) hipLaunchKernelGGL((k_shade<false, true, true>), g, b, 0, s, sc, rc, q, bounce, cur);
            else hipLaunchKernelGGL((k_shade<false, false, true>), g, b, 0, s, sc, rc, q, bounce, cur);
        }
    } else if (sc.hasCurves) {
        if (hairm) hipLaunchKernelGGL((k_shade<true, true, false>), g, b, 0, s, sc, rc, q, bounce, cur);
        else hipLaunchKernelGGL((k_shade<true, false, false>), g, b, 0, s, sc, rc, q, bounce, cur);
    } else {
        if (hairm) hipLaunchKernelGGL((k_shade<false, true, false>), g, b, 0, s, sc, rc, q, bounce, cur);
        else hipLaunchKernelGGL((k_shade<false, false, false>), g, b, 0, s, sc, rc, q, bounce, cur);
    }
}
void launchShadow(const DScene& sc, const WaveQueues& q, uint32_t bounce, uint32_t blocks, hipStream_t s) {
    if (sc.hasTransparent) {
        if (sc.hasCurves)
            hipLaunchKernelGGL((k_shadow_tr<kTraceBS, kStackKCurves, true>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce);
        else
            hipLaunchKernelGGL((k_shadow_tr<kTraceBS, kStackK, false>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce);
    } else if (curveRounds(sc, q)) {
        launchCurveQuery<kQShadow, true>(sc, q, bounce, 0u, blocks, s);
        hipLaunchKernelGGL(k_shadow_resolve, dim3(blocks), dim3(256), 0, s, q, bounce);
    } else if (sc.hasCurves)
        hipLaunchKernelGGL((k_shadow<kTraceBS, kStackKCurves, true>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce);
    else
        hipLaunchKernelGGL((k_shadow<kTraceBS, kStackK, false>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce);
}
void launchMis(const DScene& sc, const WaveQueues& q, uint32_t bounce, uint32_t blocks, hipStream_t s) {
    if (sc.hasTransparent) {
        if (sc.hasCurves)
            hipLaunchKernelGGL((k_mis_tr<kTraceBS, kStackKCurves, true>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce);
        else
            hipLaunchKernelGGL((k_mis_tr<kTraceBS, kStackK, false>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce);
    } else if (curveRounds(sc, q)) {
        if (sc.hasEmitters) {
            launchCurveQuery<kQMis, false>(sc, q, bounce, 0u, blocks, s);
            hipLaunchKernelGGL((k_mis_resolve<true>), dim3(blocks), dim3(256), 0, s, sc, q, bounce);
        } else {
            launchCurveQuery<kQMis, true>(sc, q, bounce, 0u, blocks, s);
            hipLaunchKernelGGL((k_mis_resolve<false>), dim3(blocks), dim3(256), 0, s, sc, q, bounce);
        }
    } else if (sc.hasCurves) {
        if (sc.hasEmitters)
            hipLaunchKernelGGL((k_mis<kTraceBS, kStackKCurves, true, true>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce);
        else
            hipLaunchKernelGGL((k_mis<kTraceBS, kStackKCurves, true, false>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce);
    } else {
        if (sc.hasEmitters)
            hipLaunchKernelGGL((k_mis<kTraceBS, kStackK, false, true>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce);
        else
            hipLaunchKernelGGL((k_mis<kTraceBS, kStackK, false, false>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce);
    }
}
void launchAccumulate(const WaveQueues& q, uint32_t blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_accumulate, dim3(blocks), dim3(256), 0, s, q);
}
void launchMaterialProbe(const DMat& m, bool marschner, uint32_t n, const float* in, float* out, hipStream_t s) {
    const uint32_t blocks = (n + 255u) / 256u;
    if (marschner) hipLaunchKernelGGL(k_material_probe<true>, dim3(blocks), dim3(256), 0, s, m, n, in, out);
    else hipLaunchKernelGGL(k_material_probe<false>, dim3(blocks), dim3(256), 0, s, m, n, in, out);
}
void launchTraceBatch(const DScene& sc, const float* rays, uint32_t n, ahip_hit* hits, uint8_t* occ, bool any,
                      uint32_t* ovfRef, float* ovfT, uint32_t cap, uint32_t blocks, uint32_t* overflow,
                      uint32_t* counts, hipStream_t s) {
    if (any)
        hipLaunchKernelGGL((k_trace_batch<kTraceBS, kStackK, true>), dim3(blocks), dim3(kTraceBS), 0, s, sc, rays, n,
                           hits, occ, ovfRef, ovfT, cap, overflow, counts);
    else
        hipLaunchKernelGGL((k_trace_batch<kTraceBS, kStackK, false>), dim3(blocks), dim3(kTraceBS), 0, s, sc, rays, n,
                           hits, occ, ovfRef, ovfT, cap, overflow, counts);
}
template <int KIND>
static void countLaunch(const DScene& sc, const WaveQueues& q, uint32_t bounce, uint32_t cur, uint32_t blocks,
                        unsigned long long* sums, hipStream_t s) {
    if (sc.hasCurves)
        hipLaunchKernelGGL((k_count<kTraceBS, kStackKCurves, KIND, true>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce,
                           cur, sums);
    else
        hipLaunchKernelGGL((k_count<kTraceBS, kStackK, KIND, false>), dim3(blocks), dim3(kTraceBS), 0, s, sc, q, bounce,
                           cur, sums);
}
void launchCount(const DScene& sc, const WaveQueues& q, uint32_t bounce, uint32_t cur, int kind, uint32_t blocks,
                 unsigned long long* sums, hipStream_t s) {
    if (kind == 0) countLaunch<0>(sc, q, bounce, cur, blocks, sums, s);
    else if (kind == 1) countLaunch<1>(sc, q, bounce, cur, blocks, sums, s);
    else countLaunch<2>(sc, q, bounce, cur, blocks, sums, s);
}

}  // namespace ahip
