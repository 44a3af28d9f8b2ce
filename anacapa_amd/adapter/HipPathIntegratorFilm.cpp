// HipPathIntegratorFilm.cpp — the adapter's use of Film (src/film/Film.cpp): a device
// accumulation becomes a TileBuffer of per-pixel means and weights, merged with Film::mergeTile
// the way CudaPathIntegrator merges its accumulation (CudaPathIntegrator.cu:404-417).  Kept in its
// own translation unit so everything else in the adapter links without Film.cpp.
#include "HipPathIntegrator.h"

namespace anacapa {

void mergeAccumIntoFilm(Film& film, const float* acc, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h) {
    TileBuffer tb(x0, y0, w, h);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const float* p = &acc[5 * ((size_t)y * w + x)];
            float wt = p[3] > 0.f ? p[3] : 1.f;
            tb.add(x, y, p[0] / wt, p[1] / wt, p[2] / wt, wt);
            tb.addLumSq(x, y, p[4]);
        }
    film.mergeTile(tb);
}

}  // namespace anacapa
