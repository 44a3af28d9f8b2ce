// oracle_bvh.cpp — the oracle's SAH-12 BVH build (restating src/accel/BVHBackend.cpp:17-228).
// ORACLE / TEST INFRASTRUCTURE ONLY.  Compiled with the reference's Release flags
// (-O3 -mavx2 -mfma, cmake/compiler_options.cmake:10-51) so GCC contracts the surface-area
// and cost arithmetic into fused multiply-adds exactly where it does for the reference, which
// makes near-tie split decisions — and hence the tree — identical.
#include "oracle.h"

namespace orc {

    // BVHBackend.cpp:172-228
int Bvh::sahSplit(const std::vector<PrimInfo>& pi, uint32_t s, uint32_t e, const BBox& cb, int& axisOut) const {
        float bestCost = kInf;
        int bestSplit = -1;
        axisOut = -1;
        BBox parent;
        for (uint32_t i = s; i < e; ++i) parent.expand(pi[i].b);
        V3 d = parent.diag();
        float parentArea = 2.f * (d.x * d.y + d.y * d.z + d.z * d.x);
        if (parentArea < 1e-12f) return -1;
        auto area = [](const BBox& b) {
            V3 d2 = b.diag();
            return 2.f * (d2.x * d2.y + d2.y * d2.z + d2.z * d2.x);
        };
        for (int axis = 0; axis < 3; ++axis) {
            float range = cb.diag()[axis];
            if (range < 1e-7f) continue;
            BBox bb[12];
            uint32_t bc[12] = {0};
            for (uint32_t i = s; i < e; ++i) {
                int b = (int)(12 * ((pi[i].c[axis] - cb.mn[axis]) / range));
                b = std::clamp(b, 0, 11);
                bc[b]++;
                bb[b].expand(pi[i].b);
            }
            BBox lb[11];
            uint32_t lcnt[11];
            BBox acc;
            uint32_t lc = 0;
            for (int i = 0; i < 11; ++i) {
                acc.expand(bb[i]);
                lc += bc[i];
                lb[i] = acc;
                lcnt[i] = lc;
            }
            BBox rb;
            uint32_t rc = 0;
            for (int i = 10; i >= 0; --i) {
                rb.expand(bb[i + 1]);
                rc += bc[i + 1];
                float cost = 1.f + 1.f * (area(lb[i]) * (float)lcnt[i] + area(rb) * (float)rc) / parentArea;
                if (cost < bestCost) { bestCost = cost; bestSplit = i; axisOut = axis; }
            }
        }
        float leafCost = 1.f * (float)(e - s);
        if (bestCost >= leafCost) { axisOut = -1; return -1; }
        return bestSplit;
    }

    // libstdc++'s std::partition for bidirectional ranges (two-pointer swap scheme).
template <class P>
static PrimInfo* partition(PrimInfo* first, PrimInfo* last, P pred) {
        while (true) {
            while (true) {
                if (first == last) return first;
                if (pred(*first)) ++first; else break;
            }
            --last;
            while (true) {
                if (first == last) return first;
                if (!pred(*last)) --last; else break;
            }
            std::swap(*first, *last);
            ++first;
        }
    }

uint32_t Bvh::build(std::vector<PrimInfo>& pi, uint32_t s, uint32_t e) {             // BVHBackend.cpp:121-170
        uint32_t idx = (uint32_t)nodes.size();
        nodes.emplace_back();
        BBox b, cb;
        for (uint32_t i = s; i < e; ++i) { b.expand(pi[i].b); cb.expand(pi[i].c); }
        for (int k = 0; k < 3; ++k) { nodes[idx].bmin[k] = b.mn[k]; nodes[idx].bmax[k] = b.mx[k]; }
        uint32_t count = e - s;
        int axis = -1, split = -1;
        if (count > 4) split = sahSplit(pi, s, e, cb, axis);
        if (axis < 0 || split < 0) {
            uint32_t off = (uint32_t)prims.size();
            for (uint32_t i = s; i < e; ++i) prims.push_back(pi[i].orig);
            nodes[idx].dataA = off;
            nodes[idx].dataB = count | kLeaf;
            return idx;
        }
        float range = cb.diag()[axis];
        uint32_t mid = s;
        if (range > 0.f) {
            float cmin = cb.mn[axis];
            PrimInfo* p = partition(pi.data() + s, pi.data() + e, [&](const PrimInfo& q) {
                int bk = (int)(12 * ((q.c[axis] - cmin) / range));
                return std::clamp(bk, 0, 11) <= split;
            });
            mid = (uint32_t)(p - pi.data());
        }
        if (mid == s || mid == e) mid = s + count / 2;
        build(pi, s, mid);
        uint32_t right = build(pi, mid, e);
        nodes[idx].dataA = right;
        nodes[idx].dataB = (uint32_t)axis;
        return idx;
    }

void Bvh::commit(const ahip_scene_desc* d) {                                           // BVHBackend.cpp:17-111
        nodes.clear(); tris.clear(); prims.clear();
        for (uint32_t m = 0; m < d->num_meshes; ++m) {
            uint32_t vo = d->mesh_vertex_offset[m];
            uint32_t i0 = d->mesh_index_offset[m], i1 = d->mesh_index_offset[m + 1];
            uint32_t nn = d->mesh_normal_count ? d->mesh_normal_count[m] : 0;
            uint32_t nu = d->mesh_uv_count ? d->mesh_uv_count[m] : 0;
            auto P = [&](uint32_t k) {
                const float* q = d->positions + 3 * (vo + k);
                return V3{q[0], q[1], q[2]};
            };
            for (uint32_t ti = 0; ti < (i1 - i0) / 3; ++ti) {
                uint32_t a = d->indices[i0 + 3 * ti], b = d->indices[i0 + 3 * ti + 1], c = d->indices[i0 + 3 * ti + 2];
                Tri t;
                t.v0 = identityPoint(P(a));
                V3 v1 = identityPoint(P(b)), v2 = identityPoint(P(c));
                t.e1 = v1 - t.v0;
                t.e2 = v2 - t.v0;
                t.n = safeNormalize(cross(t.e1, t.e2));
                auto N = [&](uint32_t k) {
                    if (k < nn) {
                        const float* q = d->normals + 3 * (vo + k);
                        return safeNormalize(identityNormal(V3{q[0], q[1], q[2]}));
                    }
                    return t.n;
                };
                auto U = [&](uint32_t k) {
                    V2 r;
                    if (k < nu) { r.x = d->uvs[2 * (vo + k)]; r.y = d->uvs[2 * (vo + k) + 1]; }
                    return r;
                };
                t.sn0 = N(a); t.sn1 = N(b); t.sn2 = N(c);
                t.uv0 = U(a); t.uv1 = U(b); t.uv2 = U(c);
                t.meshID = m;
                t.primID = ti;
                tris.push_back(t);
            }
        }
        if (tris.empty()) return;
        uint32_t n = (uint32_t)tris.size();
        std::vector<PrimInfo> pi(n);
        for (uint32_t i = 0; i < n; ++i) {
            const Tri& t = tris[i];
            BBox b;
            V3 v1 = t.v0 + t.e1, v2 = t.v0 + t.e2;
            b.expand(t.v0); b.expand(v1); b.expand(v2);
            for (int ax = 0; ax < 3; ++ax)
                if (b.diag()[ax] < 1e-7f) { b.mn[ax] -= 1e-7f; b.mx[ax] += 1e-7f; }
            pi[i] = {b, b.centroid(), i};
        }
        // BVHBackend.cpp:106-109: the index list starts as iota(n) and leaves are appended
        // after it, so leaf offsets begin at n.
        prims.resize(n);
        for (uint32_t i = 0; i < n; ++i) prims[i] = i;
        nodes.reserve(2 * n);
        build(pi, 0, n);
    }

// ---------------------------------------------------------------- curve BVH (CurveBrute.cpp:247-396)
namespace {
struct SegWork {
    uint32_t strand, seg;
    float bmin[3], bmax[3], c[3];
};

// StrandDesc::widthAt (CurvePool.h:45-54): none -> 0.001, one -> constant, two -> root/tip ramp,
// more -> linear interpolation between the two diameters around v * (n - 1)
float widthAtN(const float* w, uint32_t n, float v) {
    if (n == 0) return 0.001f;
    if (n == 1) return w[0];
    if (n == 2) return w[0] * (1.f - v) + w[1] * v;
    float idx = v * float(int(n) - 1);
    int lo = std::max(0, int(idx));
    int hi = std::min(int(n) - 1, lo + 1);
    float frac = idx - float(lo);
    return w[lo] * (1.f - frac) + w[hi] * frac;
}

uint32_t buildCurves(std::vector<SegWork>& w, std::vector<CNode>& nodes, std::vector<uint32_t>& order, uint32_t s,
                     uint32_t e) {
    uint32_t idx = (uint32_t)nodes.size();
    nodes.push_back({});
    float bmin[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
    float bmax[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
    for (uint32_t i = s; i < e; ++i)
        for (int k = 0; k < 3; ++k) {
            if (w[i].bmin[k] < bmin[k]) bmin[k] = w[i].bmin[k];
            if (w[i].bmax[k] > bmax[k]) bmax[k] = w[i].bmax[k];
        }
    uint32_t count = e - s;
    if (count <= 4) {
        uint32_t first = (uint32_t)order.size();
        for (uint32_t i = s; i < e; ++i) order.push_back(w[i].strand);   // index into the segment list
        for (int k = 0; k < 3; ++k) { nodes[idx].bmin[k] = bmin[k]; nodes[idx].bmax[k] = bmax[k]; }
        nodes[idx].lp = first;
        nodes[idx].rc = count | 0x80000000u;
        return idx;
    }
    float cmin[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
    float cmax[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
    for (uint32_t i = s; i < e; ++i)
        for (int k = 0; k < 3; ++k) {
            if (w[i].c[k] < cmin[k]) cmin[k] = w[i].c[k];
            if (w[i].c[k] > cmax[k]) cmax[k] = w[i].c[k];
        }
    int axis = 0;
    float ext[3] = {cmax[0] - cmin[0], cmax[1] - cmin[1], cmax[2] - cmin[2]};
    if (ext[1] > ext[0]) axis = 1;
    if (ext[2] > ext[axis]) axis = 2;
    uint32_t mid = (s + e) / 2;
    std::nth_element(w.begin() + s, w.begin() + mid, w.begin() + e,
                     [axis](const SegWork& a, const SegWork& b) { return a.c[axis] < b.c[axis]; });
    uint32_t l = buildCurves(w, nodes, order, s, mid);
    uint32_t r = buildCurves(w, nodes, order, mid, e);
    for (int k = 0; k < 3; ++k) { nodes[idx].bmin[k] = bmin[k]; nodes[idx].bmax[k] = bmax[k]; }
    nodes[idx].lp = l;
    nodes[idx].rc = r;
    return idx;
}
}  // namespace

// exported for the tests: the oracle's StrandDesc::widthAt
extern "C" float orc_width_at(const float* w, uint32_t n, float v) { return widthAtN(w, n, v); }

void Curves::build(const ahip_scene_desc* d) {
    nodes.clear(); segs.clear(); strandMesh.clear(); strandColor.clear();
    if (!d->num_strands) return;
    std::vector<SegWork> work;
    std::vector<Seg> all;
    for (uint32_t s = 0; s < d->num_strands; ++s) {
        uint32_t c0 = d->strand_cv_offset[s], c1 = d->strand_cv_offset[s + 1];
        uint32_t ncv = c1 - c0, N = ncv < 4 ? 0 : (ncv - 1) / 3;
        const float* wv = d->strand_widths ? d->strand_widths + d->strand_width_offset[s] : d->strand_width + 2 * s;
        uint32_t nw = d->strand_widths ? d->strand_width_offset[s + 1] - d->strand_width_offset[s] : 2u;
        strandMesh.push_back(d->strand_mesh_index[s]);
        strandColor.push_back(d->strand_color ? V3{d->strand_color[3 * s], d->strand_color[3 * s + 1],
                                                   d->strand_color[3 * s + 2]}
                                              : V3{1.f, 1.f, 1.f});
        for (uint32_t g = 0; g < N; ++g) {
            Seg sg;
            for (int k = 0; k < 4; ++k) {
                const float* p = d->curve_cvs + 3 * (size_t)(c0 + 3 * g + k);
                sg.P[k] = V3{p[0], p[1], p[2]};
            }
            float v0 = float(g) / float(N), v3 = float(g + 1) / float(N);
            float wa = widthAtN(wv, nw, v0), wb = widthAtN(wv, nw, v3);
            sg.hw0 = wa * 0.5f;
            sg.hw3 = wb * 0.5f;
            sg.tOff = v0;
            sg.tScale = v3 - v0;
            sg.strand = s;
            sg.seg = g;
            float mh = std::max(wa, wb) * 0.5f;
            SegWork sw;
            sw.strand = (uint32_t)all.size();   // index into `all`
            sw.seg = g;
            for (int k = 0; k < 3; ++k) {
                sw.bmin[k] = std::min({sg.P[0][k], sg.P[1][k], sg.P[2][k], sg.P[3][k]}) - mh;
                sw.bmax[k] = std::max({sg.P[0][k], sg.P[1][k], sg.P[2][k], sg.P[3][k]}) + mh;
                sw.c[k] = (sw.bmin[k] + sw.bmax[k]) * 0.5f;
            }
            all.push_back(sg);
            work.push_back(sw);
        }
    }
    if (work.empty()) return;
    std::vector<uint32_t> order;
    buildCurves(work, nodes, order, 0, (uint32_t)work.size());
    for (uint32_t i : order) segs.push_back(all[i]);
}

}  // namespace orc
