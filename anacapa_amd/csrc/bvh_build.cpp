// bvh_build.cpp — host SAH-12 BVH build and flattening into the 64-byte HBM layout.
//
// The tree is the one BVHBackend::commit builds (src/accel/BVHBackend.cpp:17-228): binned SAH
// with 12 buckets per axis, leaves of at most 4 triangles, traversal cost 1, intersection
// cost 1, std::partition on the bucket predicate with a median fallback, depth-first node
// order with the left child at index+1.  Identical trees give identical traversal orders,
// so ties between equally distant triangles resolve exactly as in the reference.  This file
// is compiled by g++ with the reference's Release flags (-O3 -mavx2 -mfma) so that its SAH
// arithmetic contracts into the same fused multiply-adds and every near-tie split decision
// matches.
//
// Flattening: each interior node becomes one 64-B DNode carrying both child boxes; triangles
// are stored in leaf order as 64-B DTri records (intersection data + ids) plus 64-B DTriShade
// records (shading normals and uvs), so the traversal kernel never chases an index list.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#define __host__
#define __device__
#include "device_types.h"
#include "bvh_build.h"

namespace ahip {
namespace {

constexpr int kBuckets = 12;        // BVHBackend.h:66
constexpr uint32_t kMaxLeaf = 4;    // BVHBackend.h:67
constexpr float kTrav = 1.f, kIsect = 1.f;
constexpr float kInf = std::numeric_limits<float>::infinity();

struct Box {
    float lo[3] = {kInf, kInf, kInf};
    float hi[3] = {-kInf, -kInf, -kInf};
    void grow(const float p[3]) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    float extent(int k) const { return hi[k] - lo[k]; }
    // 2 * (dx*dy + dy*dz + dz*dx), BVHBackend.cpp:185, 211-214
    float surface() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Prim {
    Box box;
    float c[3];
    uint32_t id;
};

struct Builder {
    std::vector<RefNode>& nodes;
    std::vector<uint32_t>& order;
    std::vector<Prim>& prims;
    uint32_t maxDepth = 0;

    static int bucketOf(float c, float lo, float range) {
        int b = static_cast<int>(kBuckets * ((c - lo) / range));
        return std::clamp(b, 0, kBuckets - 1);
    }

    // Best split over 3 axes x 11 bucket boundaries; -1 = make a leaf (BVHBackend.cpp:172-228).
    int split(uint32_t s, uint32_t e, const Box& cb, int& axis) const {
        float best = kInf;
        int bestSplit = -1;
        axis = -1;
        Box parent;
        for (uint32_t i = s; i < e; ++i) parent.grow(prims[i].box);
        float parentArea = parent.surface();
        if (parentArea < 1e-12f) return -1;
        for (int ax = 0; ax < 3; ++ax) {
            float range = cb.extent(ax);
            if (range < 1e-7f) continue;
            Box bbox[kBuckets];
            uint32_t bcnt[kBuckets] = {};
            for (uint32_t i = s; i < e; ++i) {
                int b = bucketOf(prims[i].c[ax], cb.lo[ax], range);
                ++bcnt[b];
                bbox[b].grow(prims[i].box);
            }
            Box left[kBuckets - 1];
            uint32_t nleft[kBuckets - 1];
            Box run;
            uint32_t cnt = 0;
            for (int i = 0; i < kBuckets - 1; ++i) {
                run.grow(bbox[i]);
                cnt += bcnt[i];
                left[i] = run;
                nleft[i] = cnt;
            }
            Box right;
            uint32_t nright = 0;
            for (int i = kBuckets - 2; i >= 0; --i) {
                right.grow(bbox[i + 1]);
                nright += bcnt[i + 1];
                float cost = kTrav + kIsect * (left[i].surface() * nleft[i] + right.surface() * nright) / parentArea;
                if (cost < best) { best = cost; bestSplit = i; axis = ax; }
            }
        }
        float leafCost = kIsect * static_cast<float>(e - s);
        if (best >= leafCost) { axis = -1; return -1; }
        return bestSplit;
    }

    uint32_t build(uint32_t s, uint32_t e, uint32_t depth) {
        maxDepth = std::max(maxDepth, depth);
        uint32_t idx = static_cast<uint32_t>(nodes.size());
        nodes.emplace_back();
        Box b, cb;
        for (uint32_t i = s; i < e; ++i) {
            b.grow(prims[i].box);
            cb.grow(prims[i].c);
        }
        for (int k = 0; k < 3; ++k) {
            nodes[idx].bmin[k] = b.lo[k];
            nodes[idx].bmax[k] = b.hi[k];
        }
        uint32_t count = e - s;
        int axis = -1, sp = -1;
        if (count > kMaxLeaf) sp = split(s, e, cb, axis);
        if (axis < 0 || sp < 0) {
            nodes[idx].dataA = static_cast<uint32_t>(order.size());
            nodes[idx].dataB = count | 0x80000000u;
            for (uint32_t i = s; i < e; ++i) order.push_back(prims[i].id);
            return idx;
        }
        float range = cb.extent(axis);
        uint32_t mid = s;
        if (range > 0.f) {
            float lo = cb.lo[axis];
            Prim* p = std::partition(prims.data() + s, prims.data() + e, [&](const Prim& q) {
                return bucketOf(q.c[axis], lo, range) <= sp;
            });
            mid = static_cast<uint32_t>(p - prims.data());
        }
        if (mid == s || mid == e) mid = s + count / 2;
        build(s, mid, depth + 1);
        uint32_t r = build(mid, e, depth + 1);
        nodes[idx].dataA = r;
        nodes[idx].dataB = static_cast<uint32_t>(axis);
        return idx;
    }
};

inline void sub3(const float a[3], const float b[3], float o[3]) {
    o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2];
}
inline void normalizeSafe(float v[3]) {   // Types.h:86-89
    float len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len > 1e-8f) {
        float inv = 1.f / len;
        v[0] *= inv; v[1] *= inv; v[2] *= inv;
    } else {
        v[0] = 0.f; v[1] = 0.f; v[2] = 1.f;
    }
}
// Mat4f::identity().transformPoint / transformNormal as commit applies them to static meshes
// (BVHBackend.cpp:26, 52-61; Types.h:144-167): the same sums of products with 0 and 1.
inline void xformPointIdentity(const float p[3], float o[3]) {
    float x = 1.f * p[0] + 0.f * p[1] + 0.f * p[2] + 0.f;
    float y = 0.f * p[0] + 1.f * p[1] + 0.f * p[2] + 0.f;
    float z = 0.f * p[0] + 0.f * p[1] + 1.f * p[2] + 0.f;
    float w = 0.f * p[0] + 0.f * p[1] + 0.f * p[2] + 1.f;
    o[0] = x / w; o[1] = y / w; o[2] = z / w;
}
inline void xformNormalIdentity(const float n[3], float o[3]) {
    o[0] = 1.f * n[0] + 0.f * n[1] + 0.f * n[2];
    o[1] = 0.f * n[0] + 1.f * n[1] + 0.f * n[2];
    o[2] = 0.f * n[0] + 0.f * n[1] + 1.f * n[2];
}

// Collapse the flattened binary tree into DNode4 records, depth-first with each record before
// its subtrees.  Returns the record index of binary interior node `bi`; `depth` counts records
// on the path.  A group is valid only if its box in the parent equals the union of its two
// children's boxes, which the traversal relies on to reproduce the group's entry distance.
struct Collapse {
    const std::vector<DNode>& bin;
    std::vector<DNode4>& out;
    uint32_t maxDepth = 0;
    bool unionOk = true;

    static void setSlot(DNode4& R, int s, const float* lo, const float* hi, uint32_t ref) {
        for (int k = 0; k < 3; ++k) {
            R.bmin[k][s] = lo[k];
            R.bmax[k][s] = hi[k];
        }
        R.ref[s] = ref;
    }
    static void setEmpty(DNode4& R, int s) {
        for (int k = 0; k < 3; ++k) {
            R.bmin[k][s] = kInf;
            R.bmax[k][s] = -kInf;
        }
        R.ref[s] = kRefEmpty;
    }

    uint32_t run(uint32_t bi, uint32_t depth) {
        maxDepth = std::max(maxDepth, depth);
        const uint32_t idx = static_cast<uint32_t>(out.size());
        out.emplace_back();
        const DNode X = bin[bi];
        uint32_t pending[4];
        for (int g = 0; g < 2; ++g) {
            const uint32_t cref = g == 0 ? X.left : X.right;
            const float* glo = g == 0 ? X.lmin : X.rmin;
            const float* ghi = g == 0 ? X.lmax : X.rmax;
            if (refIsLeaf(cref)) {
                setSlot(out[idx], 2 * g, glo, ghi, cref);
                setEmpty(out[idx], 2 * g + 1);
                pending[2 * g] = pending[2 * g + 1] = kRefEmpty;
                continue;
            }
            const DNode& C = bin[refIndex(cref)];
            for (int k = 0; k < 3; ++k) {
                unionOk = unionOk && glo[k] == std::min(C.lmin[k], C.rmin[k]) &&
                          ghi[k] == std::max(C.lmax[k], C.rmax[k]);
            }
            setSlot(out[idx], 2 * g, C.lmin, C.lmax, C.left);
            setSlot(out[idx], 2 * g + 1, C.rmin, C.rmax, C.right);
            pending[2 * g] = C.left;
            pending[2 * g + 1] = C.right;
        }
        out[idx].pad[0] = out[idx].pad[1] = out[idx].pad[2] = out[idx].pad[3] = 0u;
        for (int s = 0; s < 4; ++s) {
            const uint32_t r = pending[s];
            if (r == kRefEmpty || refIsLeaf(r)) continue;
            const uint32_t child = run(refIndex(r), depth + 1);
            out[idx].ref[s] = child;
        }
        return idx;
    }
};

}  // namespace

int buildBVH(const ahip_scene_desc* d, BVHOutput& out, std::string& err) {
    out = BVHOutput{};
    // 1. commit-order triangles (BVHBackend.cpp:22-77)
    std::vector<DTri> tri;
    std::vector<DTriShade> sh;
    uint64_t total = 0;
    for (uint32_t m = 0; m < d->num_meshes; ++m) {
        uint32_t i0 = d->mesh_index_offset[m], i1 = d->mesh_index_offset[m + 1];
        if (i1 < i0 || (i1 - i0) % 3 != 0 || i1 > d->num_indices) { err = "bad mesh_index_offset"; return 1; }
        total += (i1 - i0) / 3;
    }
    if (total > kMaxIndex) { err = "too many triangles for 28-bit references"; return 1; }
    tri.resize(total);
    sh.resize(total);
    uint32_t t = 0;
    for (uint32_t m = 0; m < d->num_meshes; ++m) {
        uint32_t vo = d->mesh_vertex_offset[m], vn = d->mesh_vertex_offset[m + 1] - vo;
        uint32_t i0 = d->mesh_index_offset[m], i1 = d->mesh_index_offset[m + 1];
        uint32_t nn = d->mesh_normal_count ? d->mesh_normal_count[m] : 0;
        uint32_t nu = d->mesh_uv_count ? d->mesh_uv_count[m] : 0;
        if (nn && !d->normals) nn = 0;
        if (nu && !d->uvs) nu = 0;
        for (uint32_t k = 0; k < (i1 - i0) / 3; ++k, ++t) {
            const uint32_t* ix = d->indices + i0 + 3 * k;
            if (ix[0] >= vn || ix[1] >= vn || ix[2] >= vn) { err = "vertex index out of range"; return 1; }
            float p[3][3];
            for (int j = 0; j < 3; ++j) xformPointIdentity(d->positions + 3 * (size_t)(vo + ix[j]), p[j]);
            DTri& T = tri[t];
            std::memcpy(T.v0, p[0], 12);
            sub3(p[1], p[0], T.e1);
            sub3(p[2], p[0], T.e2);
            T.n[0] = T.e1[1] * T.e2[2] - T.e1[2] * T.e2[1];
            T.n[1] = T.e1[2] * T.e2[0] - T.e1[0] * T.e2[2];
            T.n[2] = T.e1[0] * T.e2[1] - T.e1[1] * T.e2[0];
            normalizeSafe(T.n);
            T.meshID = m;
            T.primID = k;
            T.commitIndex = t;
            T.matIndex = -1;   // resolved against the material table at prepare time
            DTriShade& S = sh[t];
            float* sn[3] = {S.sn0, S.sn1, S.sn2};
            float* uv[3] = {S.uv0, S.uv1, S.uv2};
            for (int j = 0; j < 3; ++j) {
                if (ix[j] < nn) {
                    xformNormalIdentity(d->normals + 3 * (size_t)(vo + ix[j]), sn[j]);
                    normalizeSafe(sn[j]);
                } else {
                    std::memcpy(sn[j], T.n, 12);
                }
                if (ix[j] < nu) {
                    uv[j][0] = d->uvs[2 * (size_t)(vo + ix[j])];
                    uv[j][1] = d->uvs[2 * (size_t)(vo + ix[j]) + 1];
                } else {
                    uv[j][0] = uv[j][1] = 0.f;
                }
            }
            S.pad = 0;
        }
    }
    const uint32_t n = static_cast<uint32_t>(total);
    out.numTris = n;
    if (n == 0) return 0;

    // 2. primitive boxes padded to >= 1e-7 per axis (BVHBackend.cpp:81-104)
    std::vector<Prim> prims(n);
    for (uint32_t i = 0; i < n; ++i) {
        const DTri& T = tri[i];
        float v1[3], v2[3];
        for (int k = 0; k < 3; ++k) { v1[k] = T.v0[k] + T.e1[k]; v2[k] = T.v0[k] + T.e2[k]; }
        Box b;
        b.grow(T.v0); b.grow(v1); b.grow(v2);
        for (int k = 0; k < 3; ++k)
            if (b.hi[k] - b.lo[k] < 1e-7f) { b.lo[k] -= 1e-7f; b.hi[k] += 1e-7f; }
        prims[i].box = b;
        for (int k = 0; k < 3; ++k) prims[i].c[k] = (b.lo[k] + b.hi[k]) * 0.5f;
        prims[i].id = i;
    }

    // 3. build; the reference's index list starts as iota(n) (BVHBackend.cpp:106-107)
    out.refNodes.reserve(2 * (size_t)n);
    out.refOrder.resize(n);
    for (uint32_t i = 0; i < n; ++i) out.refOrder[i] = i;
    Builder B{out.refNodes, out.refOrder, prims};
    B.build(0, n, 1);
    out.maxDepth = B.maxDepth;

    // 4. flatten
    const std::vector<RefNode>& rn = out.refNodes;
    std::vector<uint32_t> interiorIndex(rn.size(), ~0u);
    uint32_t ni = 0;
    for (size_t i = 0; i < rn.size(); ++i)
        if (!(rn[i].dataB & 0x80000000u)) interiorIndex[i] = ni++;
    auto refOf = [&](uint32_t node) -> uint32_t {
        const RefNode& r = rn[node];
        if (r.dataB & 0x80000000u) return makeLeafRef(r.dataA - n, r.dataB & 0x7fffffffu);
        return interiorIndex[node];
    };
    out.nodes.resize(ni);
    for (size_t i = 0; i < rn.size(); ++i) {
        if (interiorIndex[i] == ~0u) continue;
        DNode& D = out.nodes[interiorIndex[i]];
        uint32_t L = static_cast<uint32_t>(i + 1), R = rn[i].dataA;
        std::memcpy(D.lmin, rn[L].bmin, 12);
        std::memcpy(D.lmax, rn[L].bmax, 12);
        std::memcpy(D.rmin, rn[R].bmin, 12);
        std::memcpy(D.rmax, rn[R].bmax, 12);
        D.left = refOf(L);
        D.right = refOf(R);
        D.pad[0] = D.pad[1] = 0;
    }
    out.rootRef = refOf(0);
    std::memcpy(out.rootMin, rn[0].bmin, 12);
    std::memcpy(out.rootMax, rn[0].bmax, 12);
    // triangles in leaf order: leaf entries of the index list follow the iota prefix
    out.tris.resize(n);
    out.shade.resize(n);
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t src = out.refOrder[n + k];
        out.tris[k] = tri[src];
        out.shade[k] = sh[src];
    }
    // 5. two-level records for the traversal kernels
    if (!refIsLeaf(out.rootRef)) {
        out.nodes4.reserve(out.nodes.size() / 2 + 2);
        Collapse c{out.nodes, out.nodes4};
        out.rootRef = c.run(refIndex(out.rootRef), 1);
        out.depth4 = c.maxDepth;
        if (!c.unionOk) {
            err = "BVH node box differs from the union of its children's boxes";
            return AHIP_E_INVALID_ARG;
        }
    }
    return 0;
}

// ---------------------------------------------------------------- curves
// Median split on the longest centroid axis with std::nth_element, leaves of at most 4
// segments, children stored depth-first with explicit left/right indices, segment boxes = hull
// of the 4 control points grown by the larger end half-width (CurveBrute.cpp:247-396).
namespace {

struct SegBox {
    uint32_t rec;               // index into the commit-order segment list
    float lo[3], hi[3], c[3];
};

struct CurveBuilder {
    std::vector<SegBox>& w;
    std::vector<CNode>& nodes;
    std::vector<uint32_t>& leafOrder;
    uint32_t maxDepth = 0;

    uint32_t build(uint32_t s, uint32_t e, uint32_t depth) {
        maxDepth = std::max(maxDepth, depth);
        uint32_t idx = static_cast<uint32_t>(nodes.size());
        nodes.push_back(CNode{});
        float lo[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
        float hi[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
        for (uint32_t i = s; i < e; ++i)
            for (int k = 0; k < 3; ++k) {
                if (w[i].lo[k] < lo[k]) lo[k] = w[i].lo[k];
                if (w[i].hi[k] > hi[k]) hi[k] = w[i].hi[k];
            }
        if (e - s <= 4) {
            CNode& N = nodes[idx];
            std::memcpy(N.bmin, lo, 12);
            std::memcpy(N.bmax, hi, 12);
            N.lp = static_cast<uint32_t>(leafOrder.size());
            N.rc = (e - s) | 0x80000000u;
            for (uint32_t i = s; i < e; ++i) leafOrder.push_back(w[i].rec);
            return idx;
        }
        float cl[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
        float ch[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
        for (uint32_t i = s; i < e; ++i)
            for (int k = 0; k < 3; ++k) {
                if (w[i].c[k] < cl[k]) cl[k] = w[i].c[k];
                if (w[i].c[k] > ch[k]) ch[k] = w[i].c[k];
            }
        float ext[3] = {ch[0] - cl[0], ch[1] - cl[1], ch[2] - cl[2]};
        int axis = 0;
        if (ext[1] > ext[0]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;
        uint32_t mid = (s + e) / 2;
        std::nth_element(w.begin() + s, w.begin() + mid, w.begin() + e,
                         [axis](const SegBox& a, const SegBox& b) { return a.c[axis] < b.c[axis]; });
        uint32_t l = build(s, mid, depth + 1);
        uint32_t r = build(mid, e, depth + 1);
        CNode& N = nodes[idx];
        std::memcpy(N.bmin, lo, 12);
        std::memcpy(N.bmax, hi, 12);
        N.lp = l;
        N.rc = r;
        return idx;
    }
};

// StrandDesc::widthAt (CurvePool.h:45-54) over a strand's diameters w[0, n)
float widthAt(const float* w, uint32_t n, float v) {
    if (n == 0) return 0.001f;
    if (n == 1) return w[0];
    if (n == 2) return w[0] * (1.f - v) + w[1] * v;
    const float x = v * float(int(n) - 1);
    const int lo = std::max(0, int(x));
    const int hi = std::min(int(n) - 1, lo + 1);
    const float f = x - float(lo);
    return w[lo] * (1.f - f) + w[hi] * f;
}

}  // namespace

int buildCurves(const ahip_scene_desc* d, CurveOutput& out, std::string& err) {
    out = CurveOutput{};
    if (!d->num_strands) return 0;
    if (!d->strand_cv_offset || !d->curve_cvs || !(d->strand_width || d->strand_widths) || !d->strand_mesh_index ||
        (d->strand_widths && !d->strand_width_offset)) {
        err = "strand arrays missing";
        return 1;
    }
    std::vector<DSeg> all;
    std::vector<uint32_t> ids;
    std::vector<SegBox> work;
    out.strandInfo.resize(4 * (size_t)d->num_strands);
    for (uint32_t s = 0; s < d->num_strands; ++s) {
        uint32_t c0 = d->strand_cv_offset[s], c1 = d->strand_cv_offset[s + 1];
        if (c1 < c0) {
            err = "strand_cv_offset not increasing";
            return 1;
        }
        uint32_t ncv = c1 - c0, N = ncv < 4 ? 0u : (ncv - 1) / 3;
        const float* sw = d->strand_widths ? d->strand_widths + d->strand_width_offset[s] : d->strand_width + 2 * s;
        uint32_t nw = 2;
        if (d->strand_widths) {
            if (d->strand_width_offset[s + 1] < d->strand_width_offset[s]) {
                err = "strand_width_offset not increasing";
                return 1;
            }
            nw = d->strand_width_offset[s + 1] - d->strand_width_offset[s];
        }
        float* info = &out.strandInfo[4 * (size_t)s];
        for (int k = 0; k < 3; ++k) info[k] = d->strand_color ? d->strand_color[3 * (size_t)s + k] : 1.f;
        uint32_t mesh = d->strand_mesh_index[s];
        std::memcpy(&info[3], &mesh, 4);
        for (uint32_t g = 0; g < N; ++g) {
            DSeg S;
            for (int k = 0; k < 4; ++k) std::memcpy(&S.P[3 * k], d->curve_cvs + 3 * (size_t)(c0 + 3 * g + k), 12);
            float v0 = float(g) / float(N), v3 = float(g + 1) / float(N);
            float wa = widthAt(sw, nw, v0), wb = widthAt(sw, nw, v3);
            S.hw0 = wa * 0.5f;
            S.hw3 = wb * 0.5f;
            S.tOff = v0;
            S.tScale = v3 - v0;
            float mh = std::max(wa, wb) * 0.5f;
            SegBox b;
            b.rec = static_cast<uint32_t>(all.size());
            for (int k = 0; k < 3; ++k) {
                float p0 = S.P[k], p1 = S.P[3 + k], p2 = S.P[6 + k], p3 = S.P[9 + k];
                b.lo[k] = std::min({p0, p1, p2, p3}) - mh;
                b.hi[k] = std::max({p0, p1, p2, p3}) + mh;
                b.c[k] = (b.lo[k] + b.hi[k]) * 0.5f;
            }
            all.push_back(S);
            ids.push_back(s);
            ids.push_back(g);
            work.push_back(b);
        }
    }
    if (work.empty()) return 0;
    if (work.size() >= 0x7fffffffu) {
        err = "too many curve segments";
        return 1;
    }
    std::vector<uint32_t> order;
    CurveBuilder B{work, out.nodes, order};
    B.build(0, static_cast<uint32_t>(work.size()), 1);
    out.maxDepth = B.maxDepth;
    out.segs.reserve(order.size());
    out.segIds.reserve(2 * order.size());
    for (uint32_t r : order) {
        out.segs.push_back(all[r]);
        out.segIds.push_back(ids[2 * r]);
        out.segIds.push_back(ids[2 * r + 1]);
    }
    return 0;
}

}  // namespace ahip
