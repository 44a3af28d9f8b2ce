// ref_driver.cpp — test-infrastructure driver around the reference's own CPU sources.
//
// ORACLE / TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ together
// with the unmodified reference translation units src/accel/BVHBackend.cpp and
// src/integrator/PathIntegrator.cpp, compiled where they lie under the reference checkout
// with the reference's Release flags (cmake/compiler_options.cmake:10-51).  Only tests/,
// the smoke check and bench.py's cpu_baseline leg load it.
//
// What it can drive: PCGRng, HaltonSampler, Camera::makePinhole, BVHBackend (commit/trace),
// PathIntegrator::renderTile with LambertianMaterial / EmissiveMaterial / AreaLight and a
// constant background, the RenderSession non-adaptive tile loop on the reference ThreadPool
// (RenderSession.cpp:511-547), the Kajiya-Kay hair material (MarschnerHairMaterial), and the parts
// of StandardSurfaceMaterial and DomeLight that do not need OpenImageIO:
//   * StandardSurfaceMaterial's virtual sample/evaluate call evalTOV, defined only in the
//     OIIO-only Texture.cpp, and constructing the class would emit its vtable and so need that
//     symbol.  Its private lobe helpers (sampleGGX, sampleDiffuse, evalGGX, evalCombined,
//     StandardSurface.h:511-661) read only member data, so they run here on an object whose
//     constructor is not called: the members are set as the constructor sets them (:152-160) and
//     the opaque branch of sample / evaluate above the helpers (:231-241, 351-405, 446-503) is
//     restated for constant parameters.
//   * DomeLight's constructor calls loadImage, defined in the OIIO-only DomeLight.cpp.  Its
//     distribution build, sampling and lookup (DomeLight.h:117-133, 168-171, 228-364) are header
//     code; they run here on an object whose constructor is not called, with the pixels the loader
//     would have produced set directly.
// CurveBrute.cpp includes spdlog, which this image lacks: it is not built.
//
// Private members are read through an access macro applied only to the reference headers this
// translation unit includes (no reference code is changed or replaced).

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <optional>
#include <queue>
#include <string>
#include <thread>
#include <vector>

#define private public
#include "accel/BVHBackend.h"
#include "shading/StandardSurface.h"    // free GGX helpers; the material's private lobe helpers
#include "shading/MarschnerHair.h"      // the runtime (Kajiya-Kay) hair material
#include "shading/lights/DomeLight.h"   // Distribution1D; the dome's private sampling / lookup
#undef private
#include "integrator/PathIntegrator.h"
#include "integrator/BDPTIntegrator.h"
#include "sampling/HaltonSampler.h"
#include "sampling/PCGRng.h"
#include "shading/Lambertian.h"
#include "shading/lights/AreaLight.h"
#include "shading/lights/DirectionalLight.h"
#include "render/ThreadPool.h"
#include <anacapa/film/Film.h>

#include "../include/anacapa_hip.h"
#include "ref_scene.h"

using namespace anacapa;

namespace {

void writeTile(const TileBuffer& tb, float* out) {
    for (uint32_t i = 0; i < tb.width * tb.height; ++i) {
        out[i * 5 + 0] = tb.pixels[i].r;
        out[i * 5 + 1] = tb.pixels[i].g;
        out[i * 5 + 2] = tb.pixels[i].b;
        out[i * 5 + 3] = tb.pixels[i].weight;
        out[i * 5 + 4] = tb.sumLumSq[i];
    }
}

}  // namespace

extern "C" {

int ref_pcg_u32(uint64_t seed, uint64_t seq, uint32_t n, uint32_t* out) {
    PCGRng rng(seed, seq);
    for (uint32_t i = 0; i < n; ++i) out[i] = rng.nextUint32();
    return 0;
}

int ref_pcg_float(uint64_t seed, uint32_t n, float* out) {
    PCGRng rng(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = rng.nextFloat();
    return 0;
}

// First ndims values of get1D() after startPixelSample(px, py, idx).
int ref_halton_seq(uint64_t seed, uint32_t px, uint32_t py, uint32_t idx, uint32_t ndims, float* out) {
    HaltonSampler s(1, seed);
    s.startPixelSample(px, py, idx);
    for (uint32_t d = 0; d < ndims; ++d) out[d] = s.get1D();
    return 0;
}

// Same stream consumed through get2D() pairs (checks get2D = {get1D(), get1D()} order).
int ref_halton_seq2d(uint64_t seed, uint32_t px, uint32_t py, uint32_t idx, uint32_t npairs, float* out) {
    HaltonSampler s(1, seed);
    s.startPixelSample(px, py, idx);
    for (uint32_t d = 0; d < npairs; ++d) {
        Vec2f v = s.get2D();
        out[2 * d] = v.x;
        out[2 * d + 1] = v.y;
    }
    return 0;
}

int ref_make_pinhole(const float* from, const float* at, const float* up, float vfov, uint32_t w,
                     uint32_t h, ahip_camera* out) {
    Camera c = Camera::makePinhole(v3(from), v3(at), v3(up), vfov, w, h);
    for (int i = 0; i < 3; ++i) {
        out->origin[i] = c.origin[i];
        out->lower_left[i] = c.lowerLeftCorner[i];
        out->horizontal[i] = c.horizontal[i];
        out->vertical[i] = c.vertical[i];
    }
    out->width = c.imageWidth;
    out->height = c.imageHeight;
    return 0;
}

// Camera::generateRay for a pinhole camera: rays[n*6] = origin, direction.
int ref_generate_rays(const ahip_camera* c, uint32_t n, const uint32_t* px, const uint32_t* py,
                      const float* jx, const float* jy, float* rays) {
    Camera cam = toCamera(c);
    for (uint32_t i = 0; i < n; ++i) {
        Ray r = cam.generateRay(px[i], py[i], jx[i], jy[i], 0.5f, 0.5f, 0.f);
        for (int k = 0; k < 3; ++k) {
            rays[i * 6 + k] = r.origin[k];
            rays[i * 6 + 3 + k] = r.direction[k];
        }
    }
    return 0;
}

std::unique_ptr<IMaterial> makeHarnessStandardSurface(const ahip_material& a);
std::unique_ptr<ILight> makeHarnessDome(const ahip_scene_desc* d, const ahip_light& l);

void* ref_scene_create(const ahip_scene_desc* d) {
    return buildRefScene(d, makeHarnessStandardSurface, makeHarnessDome).release();
}

void ref_scene_destroy(void* h) { delete static_cast<RefScene*>(h); }

int ref_bvh_sizes(void* h, uint32_t* nnodes, uint32_t* nprims) {
    auto* s = static_cast<RefScene*>(h);
    *nnodes = (uint32_t)s->accel->m_nodes.size();
    *nprims = (uint32_t)s->accel->m_primIndices.size();
    return 0;
}

// nodes32: 8 words per node = bmin[3], dataA, bmax[3], dataB (bit patterns).
int ref_bvh_dump(void* h, uint32_t* nodes32, uint32_t* prims) {
    auto* s = static_cast<RefScene*>(h);
    const auto& nodes = s->accel->m_nodes;
    std::memcpy(nodes32, nodes.data(), nodes.size() * sizeof(BVHNode));
    std::memcpy(prims, s->accel->m_primIndices.data(), s->accel->m_primIndices.size() * 4);
    return 0;
}

// rays[n*8] = o, d, tMin, tMax. out[n*12] floats: t, meshID(bits), primID(bits), p[3], n[3], ng[3];
// meshID bits = ~0u on a miss.
int ref_trace(void* h, const float* rays, uint32_t n, float* out) {
    auto* s = static_cast<RefScene*>(h);
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        Ray ray(v3(r), v3(r + 3), r[6], r[7]);
        TraceResult tr = s->accel->trace(ray);
        float* o = out + 12 * i;
        uint32_t mesh = tr.hit ? tr.si.meshID : ~0u, prim = tr.hit ? tr.si.primID : ~0u;
        o[0] = tr.hit ? tr.si.t : 0.f;
        std::memcpy(o + 1, &mesh, 4);
        std::memcpy(o + 2, &prim, 4);
        for (int k = 0; k < 3; ++k) {
            o[3 + k] = tr.si.p[k];
            o[6 + k] = tr.si.n[k];
            o[9 + k] = tr.si.ng[k];
        }
    }
    return 0;
}

int ref_occluded(void* h, const float* rays, uint32_t n, uint8_t* out) {
    auto* s = static_cast<RefScene*>(h);
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        out[i] = s->accel->occluded(Ray(v3(r), v3(r + 3), r[6], r[7])) ? 1 : 0;
    }
    return 0;
}

// PathIntegrator::renderTile on one tile; out[w*h*5] = r,g,b (pre-weighted sums), weight, sumLumSq.
int ref_render_tile(void* h, const ahip_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t hh,
                    uint32_t sampleStart, uint32_t sampleCount, uint32_t maxDepth, uint64_t seed,
                    float* out) {
    auto* s = static_cast<RefScene*>(h);
    s->view.camera = toCamera(cam);
    PathIntegrator integ(maxDepth);
    integ.prepare(s->view);
    HaltonSampler base(sampleCount, seed);
    auto sampler = base.clone();
    TileBuffer tb(x0, y0, w, hh);
    TileRequest tile{x0, y0, w, hh, sampleStart, sampleCount};
    integ.renderTile(s->view, tile, cam->width, cam->height, *sampler, tb);
    writeTile(tb, out);
    return 0;
}

// BDPTIntegrator(maxDepth, fireflyClamp)::renderTile on the same scene (RenderSession.cpp:459-461
// constructs it so), out[w*h*5] as ref_render_tile; albedo / normal [w*h*3] the TileBuffer AOV sums.
int ref_render_tile_bdpt(void* h, const ahip_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t hh,
                         uint32_t sampleStart, uint32_t sampleCount, uint32_t maxDepth, uint64_t seed, float clamp,
                         float* out, float* albedo, float* normal) {
    auto* s = static_cast<RefScene*>(h);
    s->view.camera = toCamera(cam);
    BDPTIntegrator integ(maxDepth, clamp);
    integ.prepare(s->view);
    HaltonSampler base(sampleCount, seed);
    auto sampler = base.clone();
    TileBuffer tb(x0, y0, w, hh);
    TileRequest tile{x0, y0, w, hh, sampleStart, sampleCount};
    integ.renderTile(s->view, tile, cam->width, cam->height, *sampler, tb);
    writeTile(tb, out);
    for (uint32_t i = 0; albedo && i < w * hh; ++i) {
        albedo[3 * i + 0] = tb.albedo[i].r;
        albedo[3 * i + 1] = tb.albedo[i].g;
        albedo[3 * i + 2] = tb.albedo[i].b;
        normal[3 * i + 0] = tb.normals[i].r;
        normal[3 * i + 1] = tb.normals[i].g;
        normal[3 * i + 2] = tb.normals[i].b;
    }
    return 0;
}

// The same with the tile's denoising AOVs as TileBuffer holds them after one renderTile call
// (PathIntegrator.cpp:58-66, Film.h:108-118): albedo / normal [w*h*3] = the recorded sums (one
// record per pixel per call).
int ref_render_tile_aovs(void* h, const ahip_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t hh,
                         uint32_t sampleStart, uint32_t sampleCount, uint32_t maxDepth, uint64_t seed, float* out,
                         float* albedo, float* normal) {
    auto* s = static_cast<RefScene*>(h);
    s->view.camera = toCamera(cam);
    PathIntegrator integ(maxDepth);
    integ.prepare(s->view);
    HaltonSampler base(sampleCount, seed);
    auto sampler = base.clone();
    TileBuffer tb(x0, y0, w, hh);
    TileRequest tile{x0, y0, w, hh, sampleStart, sampleCount};
    integ.renderTile(s->view, tile, cam->width, cam->height, *sampler, tb);
    writeTile(tb, out);
    for (uint32_t i = 0; i < w * hh; ++i) {
        albedo[3 * i + 0] = tb.albedo[i].r;
        albedo[3 * i + 1] = tb.albedo[i].g;
        albedo[3 * i + 2] = tb.albedo[i].b;
        normal[3 * i + 0] = tb.normals[i].r;
        normal[3 * i + 1] = tb.normals[i].g;
        normal[3 * i + 2] = tb.normals[i].b;
    }
    return 0;
}

// RenderSession::render non-adaptive CPU branch (RenderSession.cpp:465-547): 64x64 tiles
// (scheduleTiles, RenderSession.cpp:380-397), ThreadPool(nthreads) (0 = hardware
// concurrency), HaltonSampler(spp, seed) cloned per tile, PathIntegrator(maxDepth) ->
// renderTile -> TileBuffer, merged here into out[W*H*5].  *seconds = steady_clock around
// the parallel tile loop only (RenderSession.cpp:479, 665-667).
int ref_render_rows(void* h, const ahip_camera* cam, uint32_t spp, uint32_t maxDepth, uint64_t seed,
                    uint32_t row0, uint32_t row1, uint32_t nthreads, uint32_t tileSize, float* out, double* seconds,
                    uint32_t* threadsUsed);

int ref_render_frame(void* h, const ahip_camera* cam, uint32_t spp, uint32_t maxDepth, uint64_t seed,
                     uint32_t nthreads, uint32_t tileSize, float* out, double* seconds,
                     uint32_t* threadsUsed) {
    return ref_render_rows(h, cam, spp, maxDepth, seed, 0, cam->height, nthreads, tileSize, out, seconds,
                           threadsUsed);
}

// The same tile loop over the rows [row0, row1) of the frame only (a band of the frame for a bounded
// CPU timing sample); out still holds the whole frame.
int ref_render_rows(void* h, const ahip_camera* cam, uint32_t spp, uint32_t maxDepth, uint64_t seed,
                    uint32_t row0, uint32_t row1, uint32_t nthreads, uint32_t tileSize, float* out, double* seconds,
                    uint32_t* threadsUsed) {
    auto* s = static_cast<RefScene*>(h);
    s->view.camera = toCamera(cam);
    const uint32_t W = cam->width, H = cam->height;
    row1 = std::min(row1, H);
    PathIntegrator integ(maxDepth);
    integ.prepare(s->view);
    HaltonSampler base(spp, seed);
    ThreadPool pool(nthreads);
    if (threadsUsed) *threadsUsed = pool.numThreads();
    std::vector<TileRequest> tiles;
    for (uint32_t y = row0; y < row1; y += tileSize)
        for (uint32_t x = 0; x < W; x += tileSize)
            tiles.push_back({x, y, std::min(tileSize, W - x), std::min(tileSize, row1 - y), 0, spp});
    std::memset(out, 0, sizeof(float) * W * H * 5);
    std::mutex mergeMutex;
    auto t0 = std::chrono::steady_clock::now();
    pool.parallelFor((uint32_t)tiles.size(), [&](uint32_t ti) {
        const TileRequest& t = tiles[ti];
        auto sampler = base.clone();
        TileBuffer tb(t.x0, t.y0, t.width, t.height);
        integ.renderTile(s->view, t, W, H, *sampler, tb);
        std::lock_guard<std::mutex> lk(mergeMutex);
        for (uint32_t ty = 0; ty < t.height; ++ty)
            for (uint32_t tx = 0; tx < t.width; ++tx) {
                uint32_t fi = (t.y0 + ty) * W + (t.x0 + tx), li = ty * t.width + tx;
                out[fi * 5 + 0] += tb.pixels[li].r;
                out[fi * 5 + 1] += tb.pixels[li].g;
                out[fi * 5 + 2] += tb.pixels[li].b;
                out[fi * 5 + 3] += tb.pixels[li].weight;
                out[fi * 5 + 4] += tb.sumLumSq[li];
            }
    });
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return 0;
}

// ref_render_frame with BDPTIntegrator(maxDepth, clamp) (RenderSession.cpp:459-461) as the integrator.
int ref_render_frame_bdpt(void* h, const ahip_camera* cam, uint32_t spp, uint32_t maxDepth, uint64_t seed, float clamp,
                          uint32_t nthreads, uint32_t tileSize, float* out, double* seconds, uint32_t* threadsUsed) {
    auto* s = static_cast<RefScene*>(h);
    s->view.camera = toCamera(cam);
    const uint32_t W = cam->width, H = cam->height;
    BDPTIntegrator integ(maxDepth, clamp);
    integ.prepare(s->view);
    HaltonSampler base(spp, seed);
    ThreadPool pool(nthreads);
    if (threadsUsed) *threadsUsed = pool.numThreads();
    std::vector<TileRequest> tiles;
    for (uint32_t y = 0; y < H; y += tileSize)
        for (uint32_t x = 0; x < W; x += tileSize)
            tiles.push_back({x, y, std::min(tileSize, W - x), std::min(tileSize, H - y), 0, spp});
    std::memset(out, 0, sizeof(float) * W * H * 5);
    std::mutex mergeMutex;
    auto t0 = std::chrono::steady_clock::now();
    pool.parallelFor((uint32_t)tiles.size(), [&](uint32_t ti) {
        const TileRequest& t = tiles[ti];
        auto sampler = base.clone();
        TileBuffer tb(t.x0, t.y0, t.width, t.height);
        integ.renderTile(s->view, t, W, H, *sampler, tb);
        std::lock_guard<std::mutex> lk(mergeMutex);
        for (uint32_t ty = 0; ty < t.height; ++ty)
            for (uint32_t tx = 0; tx < t.width; ++tx) {
                uint32_t fi = (t.y0 + ty) * W + (t.x0 + tx), li = ty * t.width + tx;
                out[fi * 5 + 0] += tb.pixels[li].r;
                out[fi * 5 + 1] += tb.pixels[li].g;
                out[fi * 5 + 2] += tb.pixels[li].b;
                out[fi * 5 + 3] += tb.pixels[li].weight;
                out[fi * 5 + 4] += tb.sumLumSq[li];
            }
    });
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return 0;
}

// Distribution1D (DomeLight.h:19-55): build over values, then sample at us[k].
int ref_distribution1d(const float* values, uint32_t count, uint32_t nsamples, const float* us,
                       uint32_t* idx, float* pdf, float* uRemapped, float* cdfOut, float* integral) {
    Distribution1D d;
    d.build(values, count);
    for (uint32_t k = 0; k < nsamples; ++k) idx[k] = d.sample(us[k], pdf[k], uRemapped[k]);
    if (cdfOut) std::memcpy(cdfOut, d.cdf.data(), sizeof(float) * (count + 1));
    if (integral) *integral = d.integral;
    return 0;
}

// Free GGX / Fresnel helpers of StandardSurface.h:20-92, evaluated elementwise.
// which: 0 D_GGX(a, b), 1 G1_Smith(a, b), 2 schlickDielectric(a, b), 3 fresnelDielectric(a, b),
//        4 pdfGGX_reflection(a, b, c), 5 F0_fromIOR(a); 6 sampleGGX_halfvector((a,b), c) -> 3 outs.
int ref_ggx_helper(int which, uint32_t n, const float* a, const float* b, const float* c, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        switch (which) {
            case 0: out[i] = D_GGX(a[i], b[i]); break;
            case 1: out[i] = G1_Smith(a[i], b[i]); break;
            case 2: out[i] = schlickDielectric(a[i], b[i]); break;
            case 3: out[i] = fresnelDielectric(a[i], b[i]); break;
            case 4: out[i] = pdfGGX_reflection(a[i], b[i], c[i]); break;
            case 5: out[i] = F0_fromIOR(a[i]); break;
            case 6: {
                Vec3f w = sampleGGX_halfvector(Vec2f{a[i], b[i]}, c[i]);
                out[3 * i] = w.x; out[3 * i + 1] = w.y; out[3 * i + 2] = w.z;
                break;
            }
            default: return 1;
        }
    }
    return 0;
}

// Lambertian sample / evaluate through the reference material and ShadingContext.
// in[n*9]: normal n, wo, (u.x, u.y, unused); out[n*8]: wi, f(rgb), pdf, valid, evalPdf.
int ref_lambert(const float* albedo, uint32_t n, const float* in, float* out) {
    LambertianMaterial mat(v3(albedo));
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 9 * i;
        SurfaceInteraction si;
        si.n = v3(q);
        si.ng = si.n;
        Vec3f wo = v3(q + 3);
        ShadingContext ctx(si, -wo);
        BSDFSample bs = mat.sample(ctx, wo, Vec2f{q[6], q[7]}, 0.f);
        BSDFEval be = mat.evaluate(ctx, wo, bs.wi);
        float* o = out + 8 * i;
        o[0] = bs.wi.x; o[1] = bs.wi.y; o[2] = bs.wi.z;
        o[3] = bs.f.x; o[4] = bs.f.y; o[5] = bs.pdf;
        o[6] = bs.isValid() ? 1.f : 0.f;
        o[7] = be.pdf;
    }
    return 0;
}

// ---------------------------------------------------------------- hair (Kajiya-Kay at run time)
// Material probe on a curve hit, the layout every probe of this repository shares:
// in[n*19]  = normal n, tangent t, strand colour, wo, wi, u.x, u.y, uComponent, curve (1 here)
// out[n*16] = evaluate(wo, wi): f rgb, pdf, pdfRev | sample(wo, u, uc): wi, f rgb, pdf, valid |
//             effectiveSigmaA rgb
int ref_hair_kk_probe(const ahip_material* mat, uint32_t n, const float* in, float* out) {
    MarschnerHairMaterial::Params p;
    p.eta = mat->hair_eta;
    p.sigma_a = v3(mat->hair_sigma_a);
    p.beta_m = mat->hair_beta_m;
    p.beta_n = mat->hair_beta_n;
    p.alpha = mat->hair_alpha;
    MarschnerHairMaterial m(p);
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

// ---------------------------------------------------------------- StandardSurface (opaque, constant)
namespace {
// Storage whose StandardSurfaceMaterial member is never constructed (see the header comment).
union SSStorage {
    SSStorage() {}
    ~SSStorage() {}
    StandardSurfaceMaterial m;
};

struct SSHolder {
    SSStorage s;
    explicit SSHolder(const ahip_material& a) {
        StandardSurfaceMaterial& m = s.m;
        new (&m.m_p) StandardSurfaceMaterial::Params();
        m.m_p.base_color = Spectrum{a.albedo[0], a.albedo[1], a.albedo[2]};
        m.m_p.base = a.base;
        m.m_p.metalness = a.metalness;
        m.m_p.roughness = a.roughness;
        m.m_p.specular = a.specular;
        m.m_p.specular_color = v3(a.specular_color);
        m.m_p.specular_IOR = a.specular_ior;
        m.m_p.coat = a.coat;
        m.m_p.coat_roughness = a.coat_roughness;
        m.m_p.transmission = a.transmission;
        m.m_p.emission = a.emission_weight;
        m.m_p.emission_color = Spectrum{a.emission[0], a.emission[1], a.emission[2]};
        m.m_p.opacity = a.opacity;
        m.m_p.alphaMask = a.alpha_mask != 0u;
        // StandardSurfaceMaterial(const Params&), StandardSurface.h:152-160
        float r = std::max(1e-4f, m.m_p.roughness.value);
        m.m_alpha = r * r;
        m.m_alpha2 = m.m_alpha * m.m_alpha;
        m.m_coatAlpha = std::max(1e-4f, m.m_p.coat_roughness * m.m_p.coat_roughness);
        m.m_coatAlpha2 = m.m_coatAlpha * m.m_coatAlpha;
        m.m_f0Dielectric = F0_fromIOR(m.m_p.specular_IOR);
        m.m_coatF0 = F0_fromIOR(1.5f);
    }
    ~SSHolder() { s.m.m_p.~Params(); }
    // evalOpacity and transmittanceColor with a constant opacity (StandardSurface.h:197-225)
    float opacity() const { return s.m.m_p.alphaMask ? s.m.m_p.opacity.value : 1.f; }
    Spectrum transmittance() const {
        const StandardSurfaceMaterial& m = s.m;
        if (m.m_p.alphaMask) {
            float t = 1.f - m.m_p.opacity.value;
            if (t <= 0.f) return {};
            return {t, t, t};
        }
        if (m.m_p.transmission < 0.001f) return {};
        if (m.m_p.metalness.value > 0.001f) return {};
        float Fr = fresnelDielectric(1.f, m.m_p.specular_IOR);
        return m.m_p.base_color.value * (m.m_p.transmission * (1.f - Fr));
    }
    SSHolder(const SSHolder&) = delete;
    SSHolder& operator=(const SSHolder&) = delete;

    // Layer weights at wo (StandardSurface.h:355-371 and 478-489, identical in sample and evaluate).
    bool weights(float woz, float metal, float spec, float& wCoat, float& wMetal, float& wSpec,
                 float& wDiff) const {
        const StandardSurfaceMaterial& m = s.m;
        float coatF = schlickDielectric(woz, m.m_coatF0);
        float specF = schlickDielectric(woz, m.m_f0Dielectric);
        wCoat = m.m_p.coat * coatF;
        wMetal = metal * (1.f - wCoat);
        wSpec = spec * (1.f - metal) * specF * (1.f - wCoat);
        wDiff = m.m_p.base * (1.f - metal) * (1.f - spec * specF) * (1.f - wCoat);
        float wSum = wCoat + wMetal + wSpec + wDiff;
        if (wSum <= 0.f) return false;
        float invW = 1.f / wSum;
        wCoat *= invW;
        wMetal *= invW;
        wSpec *= invW;
        wDiff *= invW;
        return true;
    }

    // evaluate with constant parameters: the glass branches (StandardSurface.h:410-474) on the
    // reference's free helpers, else the opaque branch (:446-503)
    BSDFEval evaluate(const ShadingContext& ctx, Vec3f wo, Vec3f wi) const {
        const StandardSurfaceMaterial& m = s.m;
        Vec3f woLocal = ctx.toLocal(wo), wiLocal = ctx.toLocal(wi);
        Spectrum base_color = m.m_p.base_color.value;
        float metal = m.m_p.metalness.value, spec = m.m_p.specular.value, rough = m.m_p.roughness.value;
        if (!sameHemisphere(woLocal, wiLocal)) {
            if (m.m_p.transmission < 0.001f) return {};
            if (rough < 0.001f) return {};
            if (metal > 0.001f) return {};
            float eta = ctx.frontFace ? m.m_p.specular_IOR : (1.f / m.m_p.specular_IOR);
            float alpha2 = rough * rough;
            alpha2 *= alpha2;
            Vec3f wh = safeNormalize(woLocal + wiLocal * eta);
            if (wh.z < 0.f) wh = -wh;
            float cosIH = std::max(0.f, dot(woLocal, wh));
            float cosTH = std::abs(dot(wiLocal, wh));
            float Fr = fresnelDielectric(cosIH, eta);
            float D = D_GGX(wh.z, alpha2);
            float G2 = G2_Smith_Separable(woLocal.z, std::abs(wiLocal.z), alpha2);
            float denom = cosIH + eta * cosTH;
            if (denom < 1e-6f) return {};
            float bsdf_scalar = (1.f - Fr) * D * G2 * cosIH * cosTH / (woLocal.z * std::abs(wiLocal.z) * denom * denom);
            float jacobian = eta * eta * cosTH / (denom * denom);
            BSDFEval e;
            e.f = base_color * bsdf_scalar;
            e.pdf = (1.f - Fr) * D * wh.z * jacobian;
            e.pdfRev = e.pdf;
            return e;
        }
        float alpha = std::max(1e-4f, rough * rough);
        float alpha2 = alpha * alpha;
        if (m.m_p.transmission > 0.001f && metal < 0.001f) {
            if (rough < 0.001f) return {};
            float eta = ctx.frontFace ? m.m_p.specular_IOR : (1.f / m.m_p.specular_IOR);
            Vec3f wh = safeNormalize(woLocal + wiLocal);
            if (wh.z < 0.f) return {};
            float cosIH = std::max(0.f, dot(woLocal, wh));
            float Fr = fresnelDielectric(cosIH, eta);
            float D = D_GGX(wh.z, alpha2);
            float G2 = G2_Smith_Separable(woLocal.z, wiLocal.z, alpha2);
            float cosR = wiLocal.z;
            float pdf = Fr > 0.f ? Fr * pdfGGX_reflection(wh.z, alpha2, cosIH) : 0.f;
            BSDFEval e;
            e.f = base_color * (Fr * D * G2 / (4.f * woLocal.z * cosR));
            e.pdf = pdf;
            e.pdfRev = pdf;
            return e;
        }
        float wCoat, wMetal, wSpec, wDiff;
        if (!weights(woLocal.z, metal, spec, wCoat, wMetal, wSpec, wDiff)) return {};
        float pdfFwd, pdfRev;
        Spectrum f = m.evalCombined(ctx, woLocal, wiLocal, wCoat, wMetal, wSpec, wDiff, base_color, spec, alpha2,
                                    pdfFwd, pdfRev);
        BSDFEval e;
        e.f = f;
        e.pdf = pdfFwd;
        e.pdfRev = pdfRev;
        return e;
    }

    // sample with constant parameters and no normal map: the glass branch (StandardSurface.h:
    // 248-347) on the reference's free helpers, else the opaque branch (:231-241, 351-405)
    BSDFSample sample(const ShadingContext& ctx, Vec3f wo, Vec2f u, float uComponent) const {
        const StandardSurfaceMaterial& m = s.m;
        const ShadingContext& sctx = ctx;
        Vec3f woLocal = sctx.toLocal(wo);
        if (woLocal.z <= 0.f) return {};
        Spectrum base_color = m.m_p.base_color.value;
        float metal = m.m_p.metalness.value, spec = m.m_p.specular.value, rough = m.m_p.roughness.value;
        if (m.m_p.transmission > 0.001f && metal < 0.001f) {
            float eta = ctx.frontFace ? m.m_p.specular_IOR : (1.f / m.m_p.specular_IOR);
            float cosI = woLocal.z;
            if (rough < 0.001f) {
                float Fr = fresnelDielectric(cosI, eta);
                BSDFSample r;
                if (uComponent < Fr) {
                    r.wi = sctx.toWorld(Vec3f{-woLocal.x, -woLocal.y, woLocal.z});
                    r.f = base_color * Fr;
                    r.pdf = Fr;
                    r.pdfRev = Fr;
                    r.flags = BSDFFlag_Specular | BSDFFlag_Reflection;
                } else {
                    float sin2T = std::max(0.f, 1.f - cosI * cosI) / (eta * eta);
                    float cosT = std::sqrt(std::max(0.f, 1.f - sin2T));
                    r.wi = sctx.toWorld(Vec3f{-woLocal.x / eta, -woLocal.y / eta, -cosT});
                    r.f = base_color * (1.f - Fr);
                    r.pdf = 1.f - Fr;
                    r.pdfRev = 1.f - Fr;
                    r.eta = eta;
                    r.flags = BSDFFlag_Specular | BSDFFlag_Transmission;
                }
                return r;
            }
            float alpha = std::max(1e-4f, rough * rough), alpha2 = alpha * alpha;
            Vec3f wh = sampleGGX_halfvector(u, alpha2);
            if (wh.z <= 0.f) return {};
            float cosIH = std::max(0.f, dot(woLocal, wh));
            if (cosIH <= 0.f) return {};
            float Fr = fresnelDielectric(cosIH, eta);
            BSDFSample r;
            if (uComponent < Fr) {
                Vec3f wiLocal = wh * (2.f * cosIH) - woLocal;
                if (wiLocal.z <= 0.f) return {};
                float D = D_GGX(wh.z, alpha2), G2 = G2_Smith_Separable(woLocal.z, wiLocal.z, alpha2);
                Spectrum f0{Fr, Fr, Fr};
                Spectrum bsdf = f0 * (D * G2 / (4.f * woLocal.z));
                r.wi = sctx.toWorld(wiLocal);
                r.f = bsdf * base_color;
                r.pdf = Fr * pdfGGX_reflection(wh.z, alpha2, cosIH);
                r.pdfRev = r.pdf;
                r.flags = BSDFFlag_Glossy | BSDFFlag_Reflection;
                return r;
            }
            float sin2T = std::max(0.f, 1.f - cosIH * cosIH) / (eta * eta);
            if (sin2T >= 1.f) return {};
            float cosT_H = std::sqrt(1.f - sin2T);
            Vec3f wiLocal = woLocal * (-1.f / eta) + wh * (cosIH / eta - cosT_H);
            if (wiLocal.z >= 0.f) return {};
            float cosO = woLocal.z, cosI_t = std::abs(wiLocal.z);
            float D = D_GGX(wh.z, alpha2), G2 = G2_Smith_Separable(cosO, cosI_t, alpha2);
            float denom = cosIH + eta * std::abs(dot(wiLocal, wh));
            if (denom < 1e-6f) return {};
            float absCosT_wh = std::abs(dot(wiLocal, wh));
            float pdf_wh = D * wh.z;
            float jacobian = eta * eta * absCosT_wh / (denom * denom);
            r.wi = sctx.toWorld(wiLocal);
            r.f = base_color * ((1.f - Fr) * D * G2 * cosIH * absCosT_wh / (cosO * denom * denom));
            r.pdf = (1.f - Fr) * pdf_wh * jacobian;
            r.pdfRev = r.pdf;
            r.eta = eta;
            r.flags = BSDFFlag_Glossy | BSDFFlag_Transmission;
            return r;
        }
        float alpha = std::max(1e-4f, rough * rough);
        float alpha2 = alpha * alpha;
        float wCoat, wMetal, wSpec, wDiff;
        if (!weights(woLocal.z, metal, spec, wCoat, wMetal, wSpec, wDiff)) return {};
        BSDFSample result;
        if (uComponent < wCoat) {
            result = m.sampleGGX(sctx, woLocal, u, m.m_coatAlpha, m.m_coatAlpha2,
                                 Spectrum{m.m_coatF0, m.m_coatF0, m.m_coatF0}, false);
        } else if (uComponent < wCoat + wMetal) {
            result = m.sampleGGX(sctx, woLocal, u, alpha, alpha2, base_color * m.m_p.base, false);
        } else if (uComponent < wCoat + wMetal + wSpec) {
            Spectrum f0 = m.m_p.specular_color * (spec * m.m_f0Dielectric);
            result = m.sampleGGX(sctx, woLocal, u, alpha, alpha2, f0, false);
        } else {
            result = m.sampleDiffuse(sctx, woLocal, u, base_color);
        }
        if (!result.isValid()) return {};
        float pdfFwd = 0.f, pdfRev = 0.f;
        Spectrum fCombined = m.evalCombined(ctx, woLocal, sctx.toLocal(result.wi), wCoat, wMetal, wSpec, wDiff,
                                            base_color, spec, alpha2, pdfFwd, pdfRev);
        result.f = fCombined * std::abs(sctx.toLocal(result.wi).z);
        result.pdf = pdfFwd;
        result.pdfRev = pdfRev;
        return result;
    }
};
}  // namespace

// A test material behind the reference's own IMaterial interface for the integrator tests.
// StandardSurfaceMaterial itself cannot be built here (its virtual methods call evalTOV, defined in
// the OpenImageIO-only Texture.cpp), so its methods for constant parameters are restated on SSHolder
// (isDelta :162-169, flags :173-181, Le :183-187, reflectance :189-195, evalOpacity :197-201,
// transmittanceColor :207-225, sample / evaluate as above).  With it the reference's unmodified
// PathIntegrator — Li's alpha test (PathIntegrator.cpp:104-118), estimateDirect's step loop
// (:217-236) and shadowTransmittance (ShadowRay.h:26-61) — runs on glass and alpha-masked scenes.
class HarnessStandardSurface : public IMaterial {
public:
    explicit HarnessStandardSurface(const ahip_material& a) : h_(a) {}
    bool isDelta() const override {
        const auto& p = h_.s.m.m_p;
        float r = p.roughness.value, m = p.metalness.value, s = p.specular.value;
        if (r < 0.001f && p.transmission > 0.001f && m < 0.001f) return true;
        return r < 0.001f && (m > 0.999f || s > 0.f);
    }
    float roughness() const override { return h_.s.m.m_p.roughness.value; }
    float metalness() const override { return h_.s.m.m_p.metalness.value; }
    uint32_t flags() const override {
        const auto& p = h_.s.m.m_p;
        uint32_t f = BSDFFlag_Reflection;
        if (p.transmission > 0.001f) f |= BSDFFlag_Transmission;
        if (p.roughness.value < 0.001f) f |= BSDFFlag_Specular;
        else if (p.roughness.value < 0.3f) f |= BSDFFlag_Glossy;
        else f |= BSDFFlag_Diffuse;
        return f;
    }
    Spectrum Le(const ShadingContext&, Vec3f) const override {
        const auto& p = h_.s.m.m_p;
        if (p.emission > 0.f) return p.emission_color.value * p.emission;
        return {};
    }
    Spectrum reflectance(const ShadingContext&) const override {
        const auto& p = h_.s.m.m_p;
        float metal = p.metalness.value, spec = p.specular.value;
        float diff = (1.f - metal) * (1.f - spec * 0.5f);
        return p.base_color.value * p.base * (diff + metal);
    }
    float evalOpacity(const ShadingContext&) const override { return h_.opacity(); }
    Spectrum transmittanceColor(const ShadingContext&) const override { return h_.transmittance(); }
    BSDFSample sample(const ShadingContext& ctx, Vec3f wo, Vec2f u, float uc) const override {
        BSDFSample r = h_.sample(ctx, wo, u, uc);
        if (r.isValid() && !(h_.s.m.m_p.transmission > 0.001f && h_.s.m.m_p.metalness.value < 0.001f))
            r.flags = flags();                                           // result.flags = flags() (:403)
        return r;
    }
    BSDFEval evaluate(const ShadingContext& ctx, Vec3f wo, Vec3f wi) const override {
        return h_.evaluate(ctx, wo, wi);
    }
    float pdf(const ShadingContext& ctx, Vec3f wo, Vec3f wi) const override { return evaluate(ctx, wo, wi).pdf; }

private:
    SSHolder h_;
};

std::unique_ptr<IMaterial> makeHarnessStandardSurface(const ahip_material& a) {
    if (a.roughness_texture >= 0 || (a.alpha_mask && a.opacity_texture >= 0)) return nullptr;   // constants only
    return std::make_unique<HarnessStandardSurface>(a);
}

// The probe layout above, on a surface hit (ShadingContext from n alone: t, bt by
// buildOrthonormalBasis, ShadingContext.h:45-47); slots 13..15 hold transmittanceColor.
int ref_ss_probe(const ahip_material* mat, uint32_t n, const float* in, float* out) {
    if (mat->type != AHIP_MAT_STANDARD || mat->roughness_texture >= 0 || (mat->alpha_mask && mat->opacity_texture >= 0))
        return 1;
    SSHolder h(*mat);
    const Spectrum tint = h.transmittance();
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 19 * i;
        SurfaceInteraction si;
        si.n = v3(q);
        si.ng = si.n;
        Vec3f wo = v3(q + 9), wi = v3(q + 12);
        ShadingContext ctx(si, -wo);
        BSDFEval e = h.evaluate(ctx, wo, wi);
        BSDFSample s = h.sample(ctx, wo, Vec2f{q[15], q[16]}, q[17]);
        float* o = out + 16 * i;
        o[0] = e.f.x; o[1] = e.f.y; o[2] = e.f.z; o[3] = e.pdf; o[4] = e.pdfRev;
        o[5] = s.wi.x; o[6] = s.wi.y; o[7] = s.wi.z;
        o[8] = s.f.x; o[9] = s.f.y; o[10] = s.f.z; o[11] = s.pdf; o[12] = s.isValid() ? 1.f : 0.f;
        o[13] = tint.x; o[14] = tint.y; o[15] = tint.z;
    }
    return 0;
}

// One GGX lobe through the material's own evalGGX (StandardSurface.h:578-603):
// in[n*10] = woLocal, wiLocal, alpha2, f0 rgb; out[n*5] = f rgb, pdfFwd, pdfRev.
int ref_ss_eval_ggx(uint32_t n, const float* in, float* out) {
    ahip_material a{};
    a.type = AHIP_MAT_STANDARD;
    a.base = 1.f;
    a.roughness = 0.5f;
    a.specular_ior = 1.5f;
    SSHolder h(a);
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 10 * i;
        float pf, pr;
        Spectrum f = h.s.m.evalGGX(v3(q), v3(q + 3), q[6], v3(q + 7), pf, pr);
        out[5 * i] = f.x; out[5 * i + 1] = f.y; out[5 * i + 2] = f.z; out[5 * i + 3] = pf; out[5 * i + 4] = pr;
    }
    return 0;
}

// ---------------------------------------------------------------- DomeLight (pixels set directly)
namespace {
union DomeStorage {
    DomeStorage() {}
    ~DomeStorage() {}
    DomeLight d;
};
struct DomeHolder {
    DomeStorage s;
    DomeHolder(uint32_t w, uint32_t h, const float* rgb, float intensity, const float* rot) {
        DomeLight& d = s.d;
        new (&d.m_pixels) std::vector<float>(rgb, rgb + (size_t)w * h * 3);
        new (&d.m_conditional) std::vector<Distribution1D>();
        new (&d.m_marginal) Distribution1D();
        d.m_width = w;
        d.m_height = h;
        d.m_intensity = intensity;
        d.m_sceneRadius = 10.f;
        d.m_sceneCenter = Vec3f{0.f, 0.f, 0.f};
        d.m_envIntegral = 0.f;
        // DomeLight(...) identity rotation then setRotation (DomeLight.h:95, 105-112)
        d.m_rot[0] = v3(rot);
        d.m_rot[1] = v3(rot + 3);
        d.m_rot[2] = v3(rot + 6);
        d.buildDistribution();
    }
    ~DomeHolder() {
        DomeLight& d = s.d;
        using PixVec = std::vector<float>;
        using CondVec = std::vector<Distribution1D>;
        d.m_pixels.~PixVec();
        d.m_conditional.~CondVec();
        d.m_marginal.~Distribution1D();
    }
    DomeHolder(const DomeHolder&) = delete;
    DomeHolder& operator=(const DomeHolder&) = delete;
};
}  // namespace

// The dome as the renderer's environment light for the integrator runs: an ILight that forwards to
// the reference DomeLight's own methods (all header code, DomeLight.h:117-183), called on the
// holder's object whose constructor (with its OpenImageIO loader) is not run.
namespace {
class HarnessDome : public ILight {
public:
    HarnessDome(uint32_t w, uint32_t h, const float* rgb, float intensity, const float* rot)
        : h_(w, h, rgb, intensity, rot) {}
    LightSample sample(Vec3f p, Vec3f n, Vec2f u) const override { return h_.s.d.DomeLight::sample(p, n, u); }
    float pdf(Vec3f p, Vec3f wi) const override { return h_.s.d.DomeLight::pdf(p, wi); }
    LightLeSample sampleLe(Vec2f a, Vec2f b) const override { return h_.s.d.DomeLight::sampleLe(a, b); }
    Spectrum Le(Vec3f pos, Vec3f n, Vec3f wo) const override { return h_.s.d.DomeLight::Le(pos, n, wo); }
    float power() const override { return h_.s.d.DomeLight::power(); }
    bool isInfinite() const override { return true; }
    bool isDelta() const override { return false; }
    // the scene sphere of sampleLe, as RenderSession / the USD loader set it (DomeLight.h:187-188)
    void setSphere(float radius, Vec3f center) {
        h_.s.d.setSceneRadius(radius);
        h_.s.d.setSceneCenter(center);
    }

private:
    DomeHolder h_;
};
}  // namespace

std::unique_ptr<ILight> makeHarnessDome(const ahip_scene_desc* d, const ahip_light& l) {
    auto dome = std::make_unique<HarnessDome>(d->env_width, d->env_height, d->env_pixels, l.intensity, l.rot);
    if (l.scene_radius > 0.f) dome->setSphere(l.scene_radius, v3(l.scene_center));
    return dome;
}

// A dome over an RGB f32 equirect image (row 0 = zenith) and a world->envmap rotation (3 rows).
void* ref_dome_create(uint32_t w, uint32_t h, const float* rgb, float intensity, const float* rot) {
    return new DomeHolder(w, h, rgb, intensity, rot);
}
void ref_dome_destroy(void* p) { delete static_cast<DomeHolder*>(p); }

// DomeLight::sample (DomeLight.h:117-128): u[n*2] -> out[n*7] = wi, pdf, Li rgb.
int ref_dome_sample(void* p, uint32_t n, const float* u, float* out) {
    const DomeLight& d = static_cast<DomeHolder*>(p)->s.d;
    for (uint32_t i = 0; i < n; ++i) {
        LightSample ls = d.DomeLight::sample(Vec3f{0.f, 0.f, 0.f}, Vec3f{0.f, 1.f, 0.f}, Vec2f{u[2 * i], u[2 * i + 1]});
        float* o = out + 7 * i;
        o[0] = ls.wi.x; o[1] = ls.wi.y; o[2] = ls.wi.z; o[3] = ls.pdf;
        o[4] = ls.Li.x; o[5] = ls.Li.y; o[6] = ls.Li.z;
    }
    return 0;
}

// DomeLight::Le and DomeLight::pdf (DomeLight.h:130-133, 168-171): dirs[n*3] -> out[n*4] = Le rgb, pdf.
int ref_dome_eval(void* p, uint32_t n, const float* dirs, float* out) {
    const DomeLight& d = static_cast<DomeHolder*>(p)->s.d;
    for (uint32_t i = 0; i < n; ++i) {
        Vec3f w = v3(dirs + 3 * i);
        Spectrum le = d.DomeLight::Le(Vec3f{}, Vec3f{}, w);
        out[4 * i] = le.x; out[4 * i + 1] = le.y; out[4 * i + 2] = le.z;
        out[4 * i + 3] = d.DomeLight::pdf(Vec3f{}, w);
    }
    return 0;
}

// The built distribution (DomeLight.h:228-259): marginal func[H] and cdf[H+1], conditional
// func[H*W] and cdf[H*(W+1)], the marginal integral.
int ref_dome_tables(void* p, float* margFunc, float* margCdf, float* condFunc, float* condCdf, float* margInt) {
    const DomeLight& d = static_cast<DomeHolder*>(p)->s.d;
    std::memcpy(margFunc, d.m_marginal.func.data(), sizeof(float) * d.m_height);
    std::memcpy(margCdf, d.m_marginal.cdf.data(), sizeof(float) * (d.m_height + 1));
    for (uint32_t r = 0; r < d.m_height; ++r) {
        std::memcpy(condFunc + (size_t)r * d.m_width, d.m_conditional[r].func.data(), sizeof(float) * d.m_width);
        std::memcpy(condCdf + (size_t)r * (d.m_width + 1), d.m_conditional[r].cdf.data(),
                    sizeof(float) * (d.m_width + 1));
    }
    *margInt = d.m_marginal.integral;
    return 0;
}

}  // extern "C"
