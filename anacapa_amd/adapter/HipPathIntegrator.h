// HipPathIntegrator.h — the reference-side adapter: an IIntegrator that forwards to the
// MI355X backend through include/anacapa_hip.h.  It takes CudaPathIntegrator's place
// (src/gpu/cuda/CudaPathIntegrator.h:21-52) and is compiled inside the reference renderer,
// against its headers; RenderSession selects it the way it selects the CUDA backend
// (src/render/RenderSession.cpp:441-452).
#pragma once

#include <anacapa/integrator/IIntegrator.h>
#include <anacapa/film/Film.h>

#include <memory>
#include <mutex>
#include <string>

struct ahip_context;

namespace anacapa {

// Adds a device accumulation (w*h pixels of r, g, b sums, sample count, luminance-squared sum, as
// ahip_download_accum / ahip_render_tile return it) into `film` at (x0, y0) through a TileBuffer
// and Film::mergeTile (the CudaPathIntegrator merge, CudaPathIntegrator.cu:404-417).  Defined in
// HipPathIntegratorFilm.cpp, the adapter's only use of Film.cpp.
void mergeAccumIntoFilm(Film& film, const float* acc, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h);

class HipPathIntegrator : public IIntegrator {
public:
    // maxDepth: PathIntegrator's (RenderSession passes m_settings.maxDepth, RenderSession.cpp:463).
    // integrator: 0 = the path tracer, 1 = BDPTIntegrator(maxDepth, fireflyClamp) (RenderSession.cpp:
    // 459-461; AHIP_INTEGRATOR_* of anacapa_hip.h); prepare() leaves a BDPT adapter unprepared for
    // scenes the device BDPT does not take, so RenderSession keeps the CPU integrator.
    explicit HipPathIntegrator(int device = 0, uint32_t maxDepth = 8, uint32_t integrator = 0,
                               float fireflyClamp = 10.f);
    ~HipPathIntegrator() override;

    void prepare(const SceneView& scene) override;
    bool renderFrame(const SceneView& scene, uint32_t filmWidth, uint32_t filmHeight, uint32_t sampleStart,
                     uint32_t sampleCount, Film& film) override;
    void renderTile(const SceneView& scene, const TileRequest& tile, uint32_t filmWidth, uint32_t filmHeight,
                    ISampler& sampler, TileBuffer& out) override;

    // RenderSession's adaptive two-pass render (RenderSession.cpp:549-653) in one device call:
    // base pass, per-tile variance allocation of the remaining samples, refinement of the tiles;
    // the result is merged into `film`.  Returns false (RenderSession then renders tiles) on failure.
    bool renderAdaptive(const SceneView& scene, uint32_t filmWidth, uint32_t filmHeight, uint32_t tileSize,
                        uint32_t totalSpp, uint32_t baseSpp, Film& film);

    bool isValid() const;
    bool prepared() const;        // the last prepare() took the scene (else RenderSession uses the CPU)
    void clearAccum();
    std::string name() const;
    std::string lastError() const;

private:
    bool ensureCamera(const SceneView& scene, uint32_t w, uint32_t h);

    ahip_context* m_ctx = nullptr;
    mutable std::mutex m_mutex;      // renderTile is called concurrently (RenderSession.cpp:515-521)
    bool m_prepared = false;
    uint32_t m_maxDepth;
    uint32_t m_integrator;
    float m_fireflyClamp;
    uint32_t m_camW = 0, m_camH = 0;
};

}  // namespace anacapa
