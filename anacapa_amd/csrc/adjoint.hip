// adjoint.hip — the detached path-replay estimator on gfx950: its image, and the reverse-mode
// gradient of sum(dL/dI * I) with respect to per-triangle albedo and one roughness texture.
//
// The reference has no differentiable path.  The estimator replays the reference bounce loop
// (PathIntegrator::Li / estimateDirect, PathIntegrator.cpp:71-265) from the same Halton
// indices the forward kernels use, taking every random decision, pdf, MIS weight and
// Russian-roulette probability at the sampling parameters (the prepared scene) and every BSDF
// value f at the evaluation parameters.  The image is therefore smooth in the evaluation
// parameters and equals the ordinary render when both sets agree.
//
// Gradient: with continuation weights w_j = f_j cos / pdf_j and emission terms E_k (emitter hits,
// NEE light samples, MIS BSDF samples), L = sum_k (prod_{i<k} w_i) E_k, so
//   dL/dw_j = sum_{k>j} (prod_{i<k, i!=j} w_i) E_k.
// Each differentiable vertex keeps its leave-one-out product and that sum in a record that is
// updated as the path goes (no division by w_j, no second pass, nothing stored between kernels),
// plus the derivatives of its own NEE terms.  When the path ends with a finite radiance (the
// forward drops non-finite samples) the records are scattered with atomics.
//
// Launch: one thread per (pixel, sample); 128-thread blocks with the LDS short stack of the
// forward traversal; vertex records in a structure-of-arrays scratch indexed by thread, so a
// wave's record accesses are coalesced.
#include <hip/hip_runtime.h>

#include "device_math.h"
#include "device_sampling.h"
#include "device_shading.h"
#include "device_traverse.h"
#include "device_types.h"
#include "kernels.h"

namespace ahip {

__constant__ HaltonTables c_haltonR;

constexpr uint32_t kReplayBS = 128;
constexpr uint32_t kReplayK = 16;

namespace {

// Record fields (kRecFields floats per vertex).
enum : uint32_t {
    F_TRI = 0,      // commit index (bits)
    F_TAP = 1,      // first texel xa | ya << 16 (bits), ~0 = no texture parameter
    F_AX = 2,       // bilinear fractions
    F_AY = 3,
    F_TC = 4,       // leave-one-out throughput (3)
    F_ACC = 7,      // dL/dw (3)
    F_SRHO = 10,    // dw_c/dbase_c
    F_U = 11,       // delta * dw/dr (3)
    F_KA = 14,      // delta * (own NEE terms' d/dbase) (3)
    F_KR = 17,      // sum_c delta_c * (own NEE terms' d/dr)_c
};

struct Rec {
    float* base;
    uint32_t maxV, stride, tid;
    __device__ float& at(uint32_t f, uint32_t v) const { return base[((size_t)f * maxV + v) * stride + tid]; }
    __device__ V3 get3(uint32_t f, uint32_t v) const { return V3{at(f, v), at(f + 1, v), at(f + 2, v)}; }
    __device__ void set3(uint32_t f, uint32_t v, V3 x) const {
        at(f, v) = x.x;
        at(f + 1, v) = x.y;
        at(f + 2, v) = x.z;
    }
};

// Bilinear taps of texBilinear (same arithmetic): first texel and fractions.
__device__ __forceinline__ void texTaps(const DTexture& t, V2 uv, uint32_t& xa, uint32_t& ya, float& ax, float& ay) {
    float s = uv.x, tt = 1.f - uv.y;
    float fx = s * (float)t.w - 0.5f, fy = tt * (float)t.h - 0.5f;
    float x0f = floorf(fx), y0f = floorf(fy);
    ax = fx - x0f;
    ay = fy - y0f;
    int x0 = (int)x0f, y0 = (int)y0f;
    int W = (int)t.w, H = (int)t.h;
    int a = x0 % W;
    int b = y0 % H;
    xa = (uint32_t)(a < 0 ? a + W : a);
    ya = (uint32_t)(b < 0 ? b + H : b);
}
__device__ __forceinline__ void tapIdx(uint32_t W, uint32_t H, uint32_t xa, uint32_t ya, uint32_t idx[4]) {
    uint32_t xb = xa + 1u == W ? 0u : xa + 1u, yb = ya + 1u == H ? 0u : ya + 1u;
    idx[0] = ya * W + xa;
    idx[1] = ya * W + xb;
    idx[2] = yb * W + xa;
    idx[3] = yb * W + xb;
}

struct EvalPt {
    bool ok, texR;
    uint32_t commit, xa, ya;
    float ax, ay;
    V3 base;
    float r;
};

}  // namespace

template <bool ADJ>
__global__ __launch_bounds__(kReplayBS, 3) void k_replay(DScene sc, ReplayArgs a) {
    __shared__ uint64_t lstk[kReplayK * kReplayBS];
    const uint32_t gtid = blockIdx.x * kReplayBS + threadIdx.x;
    TStack S{lstk, a.ovfRef, a.ovfT, threadIdx.x, kReplayBS, kReplayK, gtid, a.nthreads, a.stackCap};
    Rec R{a.rec, a.maxV, a.nthreads, gtid};
    const uint32_t total = a.slots * a.ns;
    const bool evalSet = a.albE != nullptr || a.texE != nullptr;
    const float nL = (float)sc.numLights;
    TraceCounters tc{0, 0};
    for (uint32_t idx = gtid; idx < total; idx += a.nthreads) {
        const uint32_t sl = idx / a.slots, slot = idx - sl * a.slots;
        const uint32_t tile = slot >> 6, lane = slot & 63u;
        const uint32_t px = (tile % a.tilesX) * 8u + (lane & 7u), py = (tile / a.tilesX) * 8u + (lane >> 3);
        if (px >= a.w || py >= a.h) continue;
        const uint32_t s = a.s0 + sl;
        const size_t pix = (size_t)py * a.w + px;
        V3 delta{0.f, 0.f, 0.f};
        if (ADJ) {
            const float inv = 1.f / (float)a.ns;
            delta = V3{a.dLdI[3 * pix], a.dLdI[3 * pix + 1], a.dLdI[3 * pix + 2]} * inv;
        }
        // camera ray (k_generate)
        const DCamera& cam = a.cam;
        float jx = haltonSample(c_haltonR, a.seed, px, py, s, 0u);
        float jy = haltonSample(c_haltonR, a.seed, px, py, s, 1u);
        float fs = ((float)px + jx) / (float)cam.w;
        float ft = ((float)(cam.h - 1u - py) + jy) / (float)cam.h;
        V3 target = ld3(cam.llc) + ld3(cam.hor) * fs + ld3(cam.ver) * ft;
        V3 rayO = ld3(cam.origin);
        V3 dir = normalize(target - rayO);
        uint32_t dim = 5u;
        V3 L{0.f, 0.f, 0.f}, betaS{1.f, 1.f, 1.f}, betaE{1.f, 1.f, 1.f};
        uint32_t nv = 0;
        bool specular = false;
        auto emit = [&](V3 E) {    // emission term seen through the leave-one-out products
            if (!ADJ) return;
            for (uint32_t j = 0; j < nv; ++j) R.set3(F_ACC, j, R.get3(F_ACC, j) + R.get3(F_TC, j) * E);
        };
        for (uint32_t bounce = 0; bounce <= a.maxDepth; ++bounce) {
            // closest hit (k_extend)
            TRay ray = makeTRay(rayO, dir, 1e-4f, __int_as_float(0x7f800000));
            float th = 0.f, tu = 0.f, tv = 0.f;
            uint32_t tri = traverse<false, false>(sc, ray, S, th, tu, tv, tc, a.overflow);
            if (tri == ~0u) {   // PathIntegrator.cpp:83-91
                if (bounce == 0u || specular) {
                    V3 Le = envLe(sc, dir);
                    L = L + betaE * Le;
                    emit(Le);
                }
                break;
            }
            // fillSurfaceInteraction (k_shade)
            const DTri* T = sc.tris + tri;
            float u = tu, v = tv, w = 1.f - u - v;
            V3 v0 = ld3(T->v0), e1 = ld3(T->e1), e2 = ld3(T->e2);
            V3 ng = ld3(T->n);
            const uint32_t commit = T->commitIndex;
            const DTriShade* SH = sc.shade + tri;
            V3 p = v0 + e1 * u + e2 * v;
            V3 n = safeNormalize(ld3(SH->sn0) * w + ld3(SH->sn1) * u + ld3(SH->sn2) * v);
            if (dot(n, ng) < 0.f) n = -n;
            V2 uv{SH->uv0[0] * w + SH->uv1[0] * u + SH->uv2[0] * v, SH->uv0[1] * w + SH->uv1[1] * u + SH->uv2[1] * v};
            Frame fr = makeFrame(n, ng, dir);
            const int32_t mi = T->matIndex;   // sc.meshMat[meshID], resolved at prepare time
            if (mi < 0) break;
            const DMat& m = sc.mats[mi];
            const MatIn mIn = matInputs(sc, m, uv, commit);
            // evaluation-parameter view of the hit
            EvalPt ep;
            ep.ok = m.type != AHIP_MAT_HAIR;
            ep.commit = commit;
            ep.base = (a.albE && ep.ok) ? V3{a.albE[3u * (size_t)commit], a.albE[3u * (size_t)commit + 1u],
                                               a.albE[3u * (size_t)commit + 2u]}
                                        : mIn.base;
            ep.r = mIn.r;
            ep.texR = false;
            ep.xa = ep.ya = 0u;
            ep.ax = ep.ay = 0.f;
            if (m.type == AHIP_MAT_STANDARD && m.roughTex >= 0 && (uint32_t)m.roughTex == a.diffTex) {
                const DTexture& tx = sc.textures[m.roughTex];
                ep.texR = true;
                texTaps(tx, uv, ep.xa, ep.ya, ep.ax, ep.ay);
                if (a.texE) {
                    uint32_t ti[4];
                    tapIdx(tx.w, tx.h, ep.xa, ep.ya, ti);
                    float p00 = a.texE[ti[0]], p01 = a.texE[ti[1]], p10 = a.texE[ti[2]], p11 = a.texE[ti[3]];
                    ep.r = (1.f - ep.ay) * ((1.f - ep.ax) * p00 + ep.ax * p01) +
                           ep.ay * ((1.f - ep.ax) * p10 + ep.ax * p11);
                }
            }
            V3 kaNee{0.f, 0.f, 0.f};
            float krNee = 0.f;
            V3 wo = -dir;
            if (m.hasLe && (bounce == 0u || specular)) {          // PathIntegrator.cpp:128-132
                V3 Le = ld3(m.Le);
                L = L + betaE * Le;
                emit(Le);
            }
            if (!m.isDelta && sc.numLights > 0u) {                 // :134-144, estimateDirect :166-265
                float ul = haltonSample(c_haltonR, a.seed, px, py, s, dim++);
                uint32_t li = (uint32_t)(ul * nL);
                li = li < sc.numLights - 1u ? li : sc.numLights - 1u;
                const DLight& Lt = sc.lights[li];
                V2 u2;
                u2.x = haltonSample(c_haltonR, a.seed, px, py, s, dim++);
                u2.y = haltonSample(c_haltonR, a.seed, px, py, s, dim++);
                LSample ls = lightSample(sc, Lt, p, u2);
                bool shadowOk = false;
                V3 shC{0.f, 0.f, 0.f}, shE{0.f, 0.f, 0.f}, shO{0.f, 0.f, 0.f}, shD{0.f, 0.f, 0.f};
                float shTmax = 0.f, shDr = 0.f;
                V3 shCoef{0.f, 0.f, 0.f}, shDR{0.f, 0.f, 0.f};
                if (ls.pdf > 0.f && !isBlack(ls.Li)) {
                    float bpdf;
                    V3 bf = matEval(sc, m, fr, mIn, wo, ls.wi, bpdf);
                    if (!isBlack(bf)) {
                        V3 tgt = p + ls.wi * ls.dist;
                        V3 sd = tgt - p;
                        float dist = len(sd);
                        sd = sd * (1.f / dist);
                        V3 off = n * 1e-4f;
                        if (dot(sd, n) < 0.f) off = -off;
                        float f = ls.pdf, g = bpdf;
                        float wgt = (f * f) / (f * f + g * g);
                        float cosI = fabsf(dot(ls.wi, n));
                        V3 fe = bf;
                        if (ep.ok) {
                            V3 fd = matEvalDiff(m, fr, ep.base, ep.r, wo, ls.wi, shDr, shDR);
                            if (evalSet) fe = fd;
                        }
                        V3 Ld = fe * ls.Li * cosI * wgt / ls.pdf;
                        shC = betaE * Ld * nL;
                        shE = Ld * nL;
                        shCoef = betaE * (ls.Li * cosI * wgt / ls.pdf) * nL;
                        shO = p + off;
                        shD = sd;
                        shTmax = dist * (1.f - 1e-4f);
                        shadowOk = true;
                    }
                }
                // BSDF-sampling strategy; g++ evaluates get1D (uComponent) before get2D
                float uc = haltonSample(c_haltonR, a.seed, px, py, s, dim++);
                V2 ub;
                ub.x = haltonSample(c_haltonR, a.seed, px, py, s, dim++);
                ub.y = haltonSample(c_haltonR, a.seed, px, py, s, dim++);
                BSample bs = matSample(sc, m, fr, mIn, wo, ub, uc);
                bool misOk = false;
                V3 miX{0.f, 0.f, 0.f}, miC{0.f, 0.f, 0.f}, miO{0.f, 0.f, 0.f}, miCoef{0.f, 0.f, 0.f};
                V3 miDR{0.f, 0.f, 0.f};
                float miDr = 0.f;
                if (bsValid(bs)) {
                    float lp = lightPdf(sc, Lt, p, bs.wi);
                    if (lp > 0.f) {
                        float f = bs.pdf, g = lp;
                        float wgt = bs.delta ? 1.f : (f * f) / (f * f + g * g);
                        V3 off = ng * 1e-4f;
                        if (dot(bs.wi, ng) < 0.f) off = -off;
                        float cosI = (a.flags & AHIP_FLAG_FIX_DOUBLE_COSINE) ? 1.f : fabsf(dot(bs.wi, n));
                        V3 fse = bs.f;
                        if (ep.ok) {
                            float cw = fabsf(toLocal(fr, bs.wi).z);
                            V3 fd = matEvalDiff(m, fr, ep.base, ep.r, wo, bs.wi, miDr, miDR) * cw;
                            if (evalSet) fse = fd;
                            miDr *= cw;
                            miDR = miDR * cw;
                        }
                        V3 xp = fse * cosI * wgt / bs.pdf;
                        miC = betaE * xp * nL;       // k_shade's MIS weight, same association
                        miX = xp * nL;
                        miCoef = betaE * (V3{1.f, 1.f, 1.f} * (cosI * wgt / bs.pdf)) * nL;
                        miO = p + off;
                        misOk = true;
                    }
                }
                // shadow ray (k_shadow): any hit
                if (shadowOk) {
                    TRay sr = makeTRay(shO, shD, 1e-4f, shTmax);
                    float t2, u2h, v2h;
                    if (traverse<true, false>(sc, sr, S, t2, u2h, v2h, tc, a.overflow) == ~0u) {
                        L = L + shC;
                        emit(shE);
                        if (ADJ && ep.ok) {
                            kaNee = kaNee + delta * shCoef * shDr;
                            V3 q = delta * shCoef * shDR;
                            krNee += q.x + q.y + q.z;
                        }
                    }
                }
                // MIS ray (k_mis): emitted radiance where the BSDF sample lands
                if (misOk) {
                    TRay mr = makeTRay(miO, bs.wi, 1e-4f, 1e10f);
                    float t2, u2h, v2h;
                    V3 Li{0.f, 0.f, 0.f};
                    if (sc.hasEmitters) {
                        uint32_t h = traverse<false, false>(sc, mr, S, t2, u2h, v2h, tc, a.overflow);
                        if (h == ~0u) Li = envLe(sc, bs.wi);
                        else {
                            uint32_t mesh = sc.tris[h].meshID;
                            int32_t hm = mesh < sc.numMeshes ? sc.meshMat[mesh] : -1;
                            if (hm >= 0 && sc.mats[hm].hasLe) Li = ld3(sc.mats[hm].Le);
                        }
                    } else {
                        uint32_t h = traverse<true, false>(sc, mr, S, t2, u2h, v2h, tc, a.overflow);
                        if (h == ~0u) Li = envLe(sc, bs.wi);
                    }
                    if (!isBlack(Li)) {
                        L = L + miC * Li;
                        emit(miX * Li);
                        if (ADJ && ep.ok) {
                            V3 c = miCoef * Li;
                            kaNee = kaNee + delta * c * miDr;
                            V3 q = delta * c * miDR;
                            krNee += q.x + q.y + q.z;
                        }
                    }
                }
            }
            // the vertex record (after this vertex's own emission, which belongs to earlier vertices)
            uint32_t me = ~0u;
            if (ADJ && ep.ok && nv < a.maxV) {
                me = nv;
                R.at(F_TRI, me) = __uint_as_float(ep.commit);
                R.at(F_TAP, me) = __uint_as_float(ep.texR ? (ep.xa | (ep.ya << 16)) : ~0u);
                R.at(F_AX, me) = ep.ax;
                R.at(F_AY, me) = ep.ay;
                R.set3(F_TC, me, betaE);
                R.set3(F_ACC, me, V3{0.f, 0.f, 0.f});
                R.at(F_SRHO, me) = 0.f;
                R.set3(F_U, me, V3{0.f, 0.f, 0.f});
                R.set3(F_KA, me, kaNee);
                R.at(F_KR, me) = krNee;
            }
            if (bounce >= a.maxDepth) {
                if (me != ~0u) ++nv;
                break;
            }
            // continuation (PathIntegrator.cpp:146-160)
            float uc = haltonSample(c_haltonR, a.seed, px, py, s, dim++);
            V2 ub;
            ub.x = haltonSample(c_haltonR, a.seed, px, py, s, dim++);
            ub.y = haltonSample(c_haltonR, a.seed, px, py, s, dim++);
            BSample bs = matSample(sc, m, fr, mIn, wo, ub, uc);
            if (!bsValid(bs)) {
                if (me != ~0u) ++nv;
                break;
            }
            specular = bs.delta;
            V3 wS = bs.f / bs.pdf, wE = wS;
            if (ep.ok) {
                float cw = fabsf(toLocal(fr, bs.wi).z);
                float dr;
                V3 gr;
                V3 fe = matEvalDiff(m, fr, ep.base, ep.r, wo, bs.wi, dr, gr);
                if (evalSet) wE = (fe * cw) / bs.pdf;
                if (me != ~0u) {
                    R.at(F_SRHO, me) = dr * cw / bs.pdf;
                    R.set3(F_U, me, delta * (gr * (cw / bs.pdf)));
                }
            }
            if (ADJ)
                for (uint32_t j = 0; j < nv; ++j) R.set3(F_TC, j, R.get3(F_TC, j) * wE);
            if (me != ~0u) ++nv;
            betaS = betaS * wS;
            betaE = betaE * wE;
            V3 off = n * 1e-4f;
            if (dot(bs.wi, n) < 0.f) off = -off;
            rayO = p + off;
            dir = bs.wi;
            if (bounce >= a.minDepth) {
                float qq = 1.f - stdmin(maxc(betaS), 0.95f);
                if (haltonSample(c_haltonR, a.seed, px, py, s, dim++) < qq) break;
                float k = 1.f / (1.f - qq);
                betaS = betaS * k;
                betaE = betaE * k;
                if (ADJ)
                    for (uint32_t j = 0; j < nv; ++j) R.set3(F_TC, j, R.get3(F_TC, j) * k);
            }
        }
        if (!ADJ) {
            a.Lout[(size_t)sl * a.w * a.h + pix] = make_float4(L.x, L.y, L.z, 0.f);
        } else if (finite3(L)) {
            for (uint32_t j = 0; j < nv; ++j) {
                // A derivative that overflowed float (a near-singular GGX lobe seen through a tiny
                // pdf) can only give inf or inf * 0: such a vertex term is dropped, as the forward
                // drops a non-finite radiance sample.
                V3 acc = R.get3(F_ACC, j);
                V3 ka = delta * acc * R.at(F_SRHO, j) + R.get3(F_KA, j);
                uint32_t c = __float_as_uint(R.at(F_TRI, j));
                if (a.gAlb && finite3(ka)) {
                    atomicAdd(a.gAlb + 3u * (size_t)c, (double)ka.x);
                    atomicAdd(a.gAlb + 3u * (size_t)c + 1u, (double)ka.y);
                    atomicAdd(a.gAlb + 3u * (size_t)c + 2u, (double)ka.z);
                }
                uint32_t tap = __float_as_uint(R.at(F_TAP, j));
                V3 uu = R.get3(F_U, j);
                float kr = uu.x * acc.x + uu.y * acc.y + uu.z * acc.z + R.at(F_KR, j);
                if (a.gTex && tap != ~0u && isfinite(kr)) {
                    const DTexture& tx = sc.textures[a.diffTex];
                    uint32_t ti[4];
                    tapIdx(tx.w, tx.h, tap & 0xffffu, tap >> 16, ti);
                    float ax = R.at(F_AX, j), ay = R.at(F_AY, j);
                    const double k = kr;
                    atomicAdd(a.gTex + ti[0], k * ((1.f - ay) * (1.f - ax)));
                    atomicAdd(a.gTex + ti[1], k * ((1.f - ay) * ax));
                    atomicAdd(a.gTex + ti[2], k * (ay * (1.f - ax)));
                    atomicAdd(a.gTex + ti[3], k * (ay * ax));
                }
            }
        }
    }
}

__global__ void k_fill_f64(double* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// Per-pixel sums of the replayed samples in sample order (deterministic), like k_accumulate.
__global__ __launch_bounds__(256) void k_replay_reduce(ReplayArgs a, float4* accum, float* lsq) {
    const size_t n = (size_t)a.w * a.h;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        V3 acc{0.f, 0.f, 0.f};
        float q = 0.f;
        for (uint32_t k = 0; k < a.ns; ++k) {
            float4 l4 = a.Lout[(size_t)k * n + i];
            V3 L{l4.x, l4.y, l4.z};
            if (finite3(L)) {
                acc = acc + L;
                float lum = luminance(L);
                q += lum * lum;
            }
        }
        accum[i] = make_float4(acc.x, acc.y, acc.z, (float)a.ns);
        lsq[i] = q;
    }
}

void launchZeroF64(double* p, size_t n, hipStream_t s) {
    size_t b = (n + 255u) / 256u;
    if (b == 0u) return;
    uint32_t blocks = b > 4096u ? 4096u : (uint32_t)b;
    hipLaunchKernelGGL(k_fill_f64, dim3(blocks), dim3(256), 0, s, p, n);
}

void setHaltonTablesReplay(const HaltonTables& t, hipStream_t s) {
    hipMemcpyToSymbolAsync(HIP_SYMBOL(c_haltonR), &t, sizeof(t), 0, hipMemcpyHostToDevice, s);
}
uint32_t replay_block_size() { return kReplayBS; }
void launchReplay(const DScene& sc, const ReplayArgs& a, bool adjoint, uint32_t blocks, hipStream_t s) {
    if (adjoint) hipLaunchKernelGGL(k_replay<true>, dim3(blocks), dim3(kReplayBS), 0, s, sc, a);
    else hipLaunchKernelGGL(k_replay<false>, dim3(blocks), dim3(kReplayBS), 0, s, sc, a);
}
void launchReplayReduce(const ReplayArgs& a, float4* accum, float* lsq, hipStream_t s) {
    size_t n = (size_t)a.w * a.h;
    size_t b = (n + 255u) / 256u;
    uint32_t blocks = b > 8192u ? 8192u : (b == 0u ? 1u : (uint32_t)b);
    hipLaunchKernelGGL(k_replay_reduce, dim3(blocks), dim3(256), 0, s, a, accum, lsq);
}

}  // namespace ahip
