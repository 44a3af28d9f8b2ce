// curve_basis.cpp — hair curves in any of the interchange bases turned into the endpoint-sharing
// cubic Bezier strands the device curve path takes (3n+1 control points per strand, per-CV
// diameters).
//
// Restates the curve handling of the reference's Alembic reader (src/scene/alembic/
// AlembicLoader.cpp:58-154 basis conversion, :159-306 processICurves) without the Alembic SDK:
// the caller supplies what an ICurves sample holds (per-curve vertex counts, positions, widths
// in vertex / curve / constant scope, per-curve colours, the object's world matrix) and gets
// StrandDesc records back:
//   - linear:      each of the N-1 segments becomes a cubic with its inner CVs at 1/3 and 2/3;
//   - Bezier:      taken as is; a curve whose CV count is not 3n+1 (n >= 1) is skipped;
//   - B-spline:    the N-3 uniform cubic B-spline spans become Bezier spans (the standard
//                  B-spline-to-Bezier change of basis), CVs and widths alike;
//   - Catmull-Rom: the N-3 interior spans become Bezier spans with tangents (p[i+2]-p[i])/2;
//   - curves with fewer than 2 CVs are dropped silently, cubic curves with fewer than 4 CVs and
//     results with fewer than 4 CVs are skipped and counted, as the reference does.
// Widths: a per-vertex array is scaled by width_scale, a per-curve array is replicated over the
// curve's CVs (default_width past its end) and scaled, a single value is scaled and shared, and
// no widths (or any other count) gives default_width unscaled (AlembicLoader.cpp:176-209).
// Points go through the world matrix as Mat4f::transformPoint does (column vectors, divide by w;
// include/anacapa/core/Types.h:144-150).
#include "../../include/anacapa_hip.h"

#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

struct P3 {
    float x, y, z;
};
inline P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 operator*(P3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }

// One span's four Bezier points and widths from four input CVs.  Each basis writes the span's
// end points and inner points; spans share end points, so only span 0 stores its first point.
using SpanFn = void (*)(const P3* p, const float* w, P3* b, float* bw);

void bsplineSpan(const P3* p, const float* w, P3* b, float* bw) {
    b[0] = (p[0] + p[1] * 4.f + p[2]) * (1.f / 6.f);
    b[1] = (p[1] * 2.f + p[2]) * (1.f / 3.f);
    b[2] = (p[1] + p[2] * 2.f) * (1.f / 3.f);
    b[3] = (p[1] + p[2] * 4.f + p[3]) * (1.f / 6.f);
    bw[0] = (w[0] + 4.f * w[1] + w[2]) / 6.f;
    bw[1] = (2.f * w[1] + w[2]) / 3.f;
    bw[2] = (w[1] + 2.f * w[2]) / 3.f;
    bw[3] = (w[1] + 4.f * w[2] + w[3]) / 6.f;
}

void catmullRomSpan(const P3* p, const float* w, P3* b, float* bw) {
    b[0] = p[1];
    b[1] = p[1] + (p[2] - p[0]) * (1.f / 6.f);
    b[2] = p[2] - (p[3] - p[1]) * (1.f / 6.f);
    b[3] = p[2];
    bw[0] = w[1];
    bw[1] = w[1] + (w[2] - w[0]) / 6.f;
    bw[2] = w[2] - (w[3] - w[1]) / 6.f;
    bw[3] = w[2];
}

void convertSpans(const std::vector<P3>& p, const std::vector<float>& w, SpanFn span, std::vector<P3>& bp,
                  std::vector<float>& bw) {
    const int n = (int)p.size() - 3;
    if (n <= 0) return;
    bp.resize(3 * (size_t)n + 1);
    bw.resize(3 * (size_t)n + 1);
    for (int i = 0; i < n; ++i) {
        P3 b[4];
        float sw[4];
        span(&p[(size_t)i], &w[(size_t)i], b, sw);
        const int from = i == 0 ? 0 : 1;
        for (int k = from; k < 4; ++k) {
            bp[3 * (size_t)i + k] = b[k];
            bw[3 * (size_t)i + k] = sw[k];
        }
    }
}

void linearSpans(const std::vector<P3>& p, const std::vector<float>& w, std::vector<P3>& bp, std::vector<float>& bw) {
    const int n = (int)p.size() - 1;
    if (n <= 0) return;
    bp.resize(3 * (size_t)n + 1);
    bw.resize(3 * (size_t)n + 1);
    bp[0] = p[0];
    bw[0] = w[0];
    for (int i = 0; i < n; ++i) {
        const P3 a = p[(size_t)i], d = p[(size_t)i + 1] - p[(size_t)i];
        const float wa = w[(size_t)i], dw = w[(size_t)i + 1] - w[(size_t)i];
        bp[3 * (size_t)i + 1] = a + d * (1.f / 3.f);
        bp[3 * (size_t)i + 2] = a + d * (2.f / 3.f);
        bp[3 * (size_t)i + 3] = p[(size_t)i + 1];
        bw[3 * (size_t)i + 1] = wa + dw / 3.f;
        bw[3 * (size_t)i + 2] = wa + dw * 2.f / 3.f;
        bw[3 * (size_t)i + 3] = w[(size_t)i + 1];
    }
}

P3 transformPoint(const float* m, P3 p) {  // row-major m[r*4+c], column vectors
    const float x = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
    const float y = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
    const float z = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
    const float w = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15];
    return {x / w, y / w, z / w};
}

}  // namespace

struct ahip_curve_set {
    std::vector<uint32_t> cvOffset{0}, widthOffset{0}, mesh;
    std::vector<float> cvs, widths, color;
    ahip_strands view;
};

extern "C" {

int ahip_curves_to_bezier(uint32_t basis, uint32_t num_curves, const uint32_t* counts, const float* points,
                          const float* widths, uint32_t num_widths, const float* colors, const float* world,
                          const ahip_curve_options* opts, ahip_curve_set** out, char* err, size_t err_len) {
    auto fail = [&](int st, const char* m) {
        if (err && err_len) std::snprintf(err, err_len, "%s", m);
        return st;
    };
    if (out) *out = nullptr;
    if (!out || !opts || (num_curves && (!counts || !points)))
        return fail(AHIP_E_INVALID_ARG, "ahip_curves_to_bezier: null argument");
    if (basis > AHIP_CURVE_CATMULL_ROM) return fail(AHIP_E_UNSUPPORTED, "ahip_curves_to_bezier: unknown curve basis");
    if (num_widths && !widths) return fail(AHIP_E_INVALID_ARG, "ahip_curves_to_bezier: widths missing");
    try {
        size_t total = 0;
        for (uint32_t c = 0; c < num_curves; ++c) total += counts[c];
        // widths per CV over the whole object (AlembicLoader.cpp:176-209)
        std::vector<float> all;
        if (num_widths && num_widths == total) {
            all.resize(total);
            for (size_t i = 0; i < total; ++i) all[i] = widths[i] * opts->width_scale;
        } else if (num_widths && num_widths == num_curves) {
            all.reserve(total);
            for (uint32_t c = 0; c < num_curves; ++c) {
                const float w = c < num_widths ? widths[c] : opts->default_width;
                for (uint32_t k = 0; k < counts[c]; ++k) all.push_back(w * opts->width_scale);
            }
        } else if (num_widths == 1) {
            all.assign(total, widths[0] * opts->width_scale);
        }
        if (all.empty()) all.assign(total, opts->default_width);

        auto* set = new ahip_curve_set();
        uint32_t skipped = 0;
        size_t off = 0;
        for (uint32_t c = 0; c < num_curves; ++c) {
            const uint32_t nv = counts[c];
            const size_t start = off;
            off += nv;
            if (nv < 2) continue;
            std::vector<P3> raw(nv);
            for (uint32_t i = 0; i < nv; ++i) {
                const float* q = points + 3 * (start + i);
                raw[i] = world ? transformPoint(world, {q[0], q[1], q[2]}) : P3{q[0], q[1], q[2]};
            }
            std::vector<float> rw(all.begin() + (ptrdiff_t)start, all.begin() + (ptrdiff_t)(start + nv));
            std::vector<P3> bp;
            std::vector<float> bw;
            if (basis == AHIP_CURVE_LINEAR) {
                linearSpans(raw, rw, bp, bw);
            } else if (basis == AHIP_CURVE_BEZIER) {
                if (nv < 4 || (nv - 1) % 3 != 0) {
                    ++skipped;
                    continue;
                }
                bp = raw;
                bw = rw;
            } else {
                if (nv < 4) {
                    ++skipped;
                    continue;
                }
                convertSpans(raw, rw, basis == AHIP_CURVE_BSPLINE ? bsplineSpan : catmullRomSpan, bp, bw);
            }
            if (bp.size() < 4) {
                ++skipped;
                continue;
            }
            for (const P3& p : bp) set->cvs.insert(set->cvs.end(), {p.x, p.y, p.z});
            set->widths.insert(set->widths.end(), bw.begin(), bw.end());
            set->cvOffset.push_back((uint32_t)(set->cvs.size() / 3));
            set->widthOffset.push_back((uint32_t)set->widths.size());
            set->mesh.push_back(opts->material_index);
            for (int k = 0; k < 3; ++k) set->color.push_back(colors ? colors[3 * (size_t)c + k] : 1.f);
        }
        ahip_strands& v = set->view;
        std::memset(&v, 0, sizeof v);
        v.num_strands = (uint32_t)set->mesh.size();
        v.num_cvs = (uint32_t)(set->cvs.size() / 3);
        v.num_skipped = skipped;
        v.strand_cv_offset = set->cvOffset.data();
        v.curve_cvs = set->cvs.data();
        v.strand_width_offset = set->widthOffset.data();
        v.strand_widths = set->widths.data();
        v.strand_mesh_index = set->mesh.data();
        v.strand_color = set->color.data();
        *out = set;
        if (err && err_len) err[0] = '\0';
        return AHIP_OK;
    } catch (const std::bad_alloc&) {
        return fail(AHIP_E_OUT_OF_MEMORY, "ahip_curves_to_bezier: out of host memory");
    }
}

const ahip_strands* ahip_curve_set_strands(const ahip_curve_set* set) { return set ? &set->view : nullptr; }

void ahip_curve_set_free(ahip_curve_set* set) { delete set; }

}  // extern "C"
