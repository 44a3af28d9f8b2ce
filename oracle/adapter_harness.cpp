// adapter_harness.cpp — runs the reference-side adapter (anacapa_amd/adapter/HipPathIntegrator)
// inside the reference's own render loop.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/libadapter_harness.so from
// the reference's BVHBackend.cpp and PathIntegrator.cpp (compiled where they lie, at its Release
// flags), the adapter's HipPathIntegrator.cpp and this file, linked to the product library
// anacapa_amd/lib/libanacapa_hip.so.  Only tests/ load it.
//
// It does what RenderSession::render does for a non-adaptive render (RenderSession.cpp:441-547):
// build the SceneView, construct the integrator, isValid(), prepare(scene), then 64x64
// TileRequests on the reference ThreadPool, each with a clone of HaltonSampler(spp, seed) and its
// own TileBuffer, merged into the frame — once with the adapter as the IIntegrator, once with the
// reference PathIntegrator, on the same SceneView.  Film (src/film/Film.cpp) needs OpenImageIO
// and OpenImageDenoise, which this image lacks, so the tile merge is done here the way
// Film::mergeTile adds a TileBuffer (pre-weighted sums, weights, luminance squares, AOV sums), and
// the adapter's HipPathIntegratorFilm.cpp (renderFrame / renderAdaptive's merge) is not linked:
// mergeAccumIntoFilm below only reports that it was reached.

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "accel/BVHBackend.h"
#include "integrator/BDPTIntegrator.h"
#include "integrator/PathIntegrator.h"
#include "render/ThreadPool.h"
#include "sampling/HaltonSampler.h"
#include "shading/Lambertian.h"
#include "shading/lights/AreaLight.h"
#include "shading/lights/DirectionalLight.h"
#include <anacapa/film/Film.h>

#include "../anacapa_amd/adapter/HipPathIntegrator.h"
#include "ref_scene.h"

namespace anacapa {
std::atomic<int> g_filmMerges{0};
// HipPathIntegratorFilm.cpp is not part of this harness (Film.cpp is not buildable here).
void mergeAccumIntoFilm(Film&, const float*, uint32_t, uint32_t, uint32_t, uint32_t) { g_filmMerges++; }
}  // namespace anacapa

namespace {

struct Harness {
    std::unique_ptr<RefScene> scene;
    std::unique_ptr<HipPathIntegrator> hip;
};

// RenderSession's tile loop (scheduleTiles + parallelFor, RenderSession.cpp:380-397, 511-547) with
// `integ`, merged into out[W*H*5] (r, g, b sums, weight, luminance-squared sum) and, when given,
// albedo / normal [W*H*3] (TileBuffer AOV sums).
void renderTiles(IIntegrator& integ, const SceneView& view, uint32_t W, uint32_t H, uint32_t spp, uint64_t seed,
                 uint32_t nthreads, uint32_t tileSize, float* out, float* albedo, float* normal) {
    HaltonSampler base(spp, seed);
    ThreadPool pool(nthreads);
    std::vector<TileRequest> tiles;
    for (uint32_t y = 0; y < H; y += tileSize)
        for (uint32_t x = 0; x < W; x += tileSize)
            tiles.push_back({x, y, std::min(tileSize, W - x), std::min(tileSize, H - y), 0, spp});
    std::memset(out, 0, sizeof(float) * W * H * 5);
    if (albedo) std::memset(albedo, 0, sizeof(float) * W * H * 3);
    if (normal) std::memset(normal, 0, sizeof(float) * W * H * 3);
    std::mutex mergeMutex;
    pool.parallelFor((uint32_t)tiles.size(), [&](uint32_t ti) {
        const TileRequest& t = tiles[ti];
        auto sampler = base.clone();
        TileBuffer tb(t.x0, t.y0, t.width, t.height);
        integ.renderTile(view, t, W, H, *sampler, tb);
        std::lock_guard<std::mutex> lk(mergeMutex);
        for (uint32_t ty = 0; ty < t.height; ++ty)
            for (uint32_t tx = 0; tx < t.width; ++tx) {
                uint32_t fi = (t.y0 + ty) * W + (t.x0 + tx), li = ty * t.width + tx;
                out[fi * 5 + 0] += tb.pixels[li].r;
                out[fi * 5 + 1] += tb.pixels[li].g;
                out[fi * 5 + 2] += tb.pixels[li].b;
                out[fi * 5 + 3] += tb.pixels[li].weight;
                out[fi * 5 + 4] += tb.sumLumSq[li];
                if (albedo) {
                    albedo[fi * 3 + 0] += tb.albedo[li].r;
                    albedo[fi * 3 + 1] += tb.albedo[li].g;
                    albedo[fi * 3 + 2] += tb.albedo[li].b;
                }
                if (normal) {
                    normal[fi * 3 + 0] += tb.normals[li].r;
                    normal[fi * 3 + 1] += tb.normals[li].g;
                    normal[fi * 3 + 2] += tb.normals[li].b;
                }
            }
    });
}

}  // namespace

extern "C" {

// The reference SceneView of the scene (Lambert / Emissive / quad area lights) and the adapter on
// HIP device 0, prepared with it.  Returns null when the reference cannot represent the scene.
// integrator / clamp: the adapter's (0 = path tracer, 1 = BDPT with that firefly clamp).
void* ah_create(const ahip_scene_desc* d, const ahip_camera* cam, uint32_t maxDepth, uint32_t integrator, float clamp) {
    auto h = std::make_unique<Harness>();
    h->scene = buildRefScene(d);
    if (!h->scene) return nullptr;
    h->scene->view.camera = toCamera(cam);
    h->hip = std::make_unique<HipPathIntegrator>(0, maxDepth, integrator, clamp);
    h->hip->prepare(h->scene->view);
    return h.release();
}

void ah_destroy(void* p) { delete static_cast<Harness*>(p); }

// isValid(), whether prepare() took the scene, the device name and the last error.
int ah_status(void* p, int* valid, int* prepared, char* name, size_t nameLen, char* err, size_t errLen) {
    auto* h = static_cast<Harness*>(p);
    *valid = h->hip->isValid() ? 1 : 0;
    *prepared = h->hip->prepared() ? 1 : 0;
    std::snprintf(name, nameLen, "%s", h->hip->name().c_str());
    std::snprintf(err, errLen, "%s", h->hip->lastError().c_str());
    return 0;
}

// The frame through RenderSession's tile loop with the adapter (which = 0), the reference
// PathIntegrator(maxDepth) (which = 1) or the reference BDPTIntegrator(maxDepth, clamp) (which = 2)
// as the IIntegrator.
int ah_render(void* p, int which, uint32_t spp, uint32_t maxDepth, uint64_t seed, uint32_t nthreads, uint32_t tileSize,
              float clamp, float* out, float* albedo, float* normal) {
    auto* h = static_cast<Harness*>(p);
    const Camera& cam = *h->scene->view.camera;
    if (which == 0) {
        renderTiles(*h->hip, h->scene->view, cam.imageWidth, cam.imageHeight, spp, seed, nthreads, tileSize, out,
                    albedo, normal);
    } else if (which == 2) {
        BDPTIntegrator integ(maxDepth, clamp);
        integ.prepare(h->scene->view);
        renderTiles(integ, h->scene->view, cam.imageWidth, cam.imageHeight, spp, seed, nthreads, tileSize, out,
                    albedo, normal);
    } else {
        PathIntegrator integ(maxDepth);
        integ.prepare(h->scene->view);
        renderTiles(integ, h->scene->view, cam.imageWidth, cam.imageHeight, spp, seed, nthreads, tileSize, out,
                    albedo, normal);
    }
    return 0;
}

int ah_film_merges() { return anacapa::g_filmMerges.load(); }

}  // extern "C"
