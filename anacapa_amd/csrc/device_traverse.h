// device_traverse.h — BVH traversal over 64-byte two-box nodes with an LDS short stack.
//
// Semantics are those of BVHBackend::traceImpl / occluded (src/accel/BVHBackend.cpp:233-444):
//  * invDir = 1 / safe(d) with |d| <= 1e-9 mapped to +1e-9, and dirNeg from the sign of d
//    (BVHBackend.cpp:233-247);
//  * slab test rejecting tMax <= tMin (BVHBackend.cpp:249-262);
//  * Moller-Trumbore on the direction rebuilt as 1/invDir, hit accepted for t in
//    [tMin, closest] inclusive so a later leaf triangle at equal t replaces the earlier one
//    (BVHBackend.cpp:264-283, 369-372);
//  * nearer child first, the right child first on equal entry distance, nodes culled when
//    their entry distance exceeds the closest hit (BVHBackend.cpp:349, 376-387).
// Both child boxes live in the parent record, so the entry distance computed when a child
// is pushed is exactly what the reference recomputes when it pops that child; the cull test
// is applied at push and again at pop.  Any-hit queries (shadow rays) return whether a hit
// exists in [tMin, tMax], which is the boolean both BVHBackend::occluded and the closest-hit
// based shadowTransmittance (src/integrator/ShadowRay.h:26-61) produce for opaque scenes.
//
// The intersection arithmetic runs without FP contraction so results are a fixed IEEE
// computation (the CPU oracle computes the same bits).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.h"
#include "device_types.h"

namespace ahip {

struct TRay {
    V3 o;
    V3 inv;     // 1 / safe(d)
    V3 dir;     // 1 / inv, the direction Moller-Trumbore sees (BVHBackend.cpp:267)
    float tMin, tMax;
    uint32_t neg;   // bit k set when d[k] < 0
};

__device__ __forceinline__ TRay makeTRay(V3 o, V3 d, float tMin, float tMax) {
#pragma clang fp contract(off)
    TRay r;
    r.o = o;
    r.tMin = tMin;
    r.tMax = tMax;
    float sx = fabsf(d.x) > 1e-9f ? d.x : 1e-9f;
    float sy = fabsf(d.y) > 1e-9f ? d.y : 1e-9f;
    float sz = fabsf(d.z) > 1e-9f ? d.z : 1e-9f;
    r.inv = V3{1.f / sx, 1.f / sy, 1.f / sz};
    r.dir = V3{1.f / r.inv.x, 1.f / r.inv.y, 1.f / r.inv.z};
    r.neg = (d.x < 0.f ? 1u : 0u) | (d.y < 0.f ? 2u : 0u) | (d.z < 0.f ? 4u : 0u);
    return r;
}

// Slab test of one box; returns hit and the entry distance.
__device__ __forceinline__ bool slab(const float* bmin, const float* bmax, const TRay& r, float& tNear) {
#pragma clang fp contract(off)
    float nx = (r.neg & 1u) ? bmax[0] : bmin[0], fx = (r.neg & 1u) ? bmin[0] : bmax[0];
    float ny = (r.neg & 2u) ? bmax[1] : bmin[1], fy = (r.neg & 2u) ? bmin[1] : bmax[1];
    float nz = (r.neg & 4u) ? bmax[2] : bmin[2], fz = (r.neg & 4u) ? bmin[2] : bmax[2];
    float t0x = (nx - r.o.x) * r.inv.x, t1x = (fx - r.o.x) * r.inv.x;
    float t0y = (ny - r.o.y) * r.inv.y, t1y = (fy - r.o.y) * r.inv.y;
    float t0z = (nz - r.o.z) * r.inv.z, t1z = (fz - r.o.z) * r.inv.z;
    float tmin = fmaxf(fmaxf(fmaxf(r.tMin, t0x), t0y), t0z);
    float tmax = fminf(fminf(fminf(r.tMax, t1x), t1y), t1z);
    tNear = tmin;
    return tmax > tmin;
}

// Contraction-free vector helpers for the intersection arithmetic (the pragma is lexical, so the
// general helpers of device_math.h, which may fuse, are not used here).
__device__ __forceinline__ float dotNC(V3 a, V3 b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ V3 crossNC(V3 a, V3 b) {
#pragma clang fp contract(off)
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ bool hitTri(const DTri& T, const TRay& r, float tMaxCur, float& t, float& u, float& v) {
#pragma clang fp contract(off)
    V3 e1 = ld3(T.e1), e2 = ld3(T.e2), v0 = ld3(T.v0);
    V3 h = crossNC(r.dir, e2);
    float a = dotNC(e1, h);
    if (fabsf(a) < 1e-9f) return false;
    float f = 1.f / a;
    V3 s = V3{r.o.x - v0.x, r.o.y - v0.y, r.o.z - v0.z};
    u = f * dotNC(s, h);
    if (u < 0.f || u > 1.f) return false;
    V3 q = crossNC(s, e1);
    v = f * dotNC(r.dir, q);
    if (v < 0.f || u + v > 1.f) return false;
    t = f * dotNC(e2, q);
    return t >= r.tMin && t <= tMaxCur;
}

// Triangle record fetch as three 16-byte loads (48 of the 64 bytes; ids are read on a hit).
__device__ __forceinline__ DTri loadTriGeom(const DTri* tris, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(tris + i);
    float4 a = p[0], b = p[1], c = p[2];
    DTri T;
    T.v0[0] = a.x; T.v0[1] = a.y; T.v0[2] = a.z;
    T.e1[0] = a.w; T.e1[1] = b.x; T.e1[2] = b.y;
    T.e2[0] = b.z; T.e2[1] = b.w; T.e2[2] = c.x;
    T.n[0] = c.y; T.n[1] = c.z; T.n[2] = c.w;
    return T;
}

struct NodeBoxes {
    float lmin[3], lmax[3], rmin[3], rmax[3];
    uint32_t left, right;
};
__device__ __forceinline__ NodeBoxes loadNode(const DNode* nodes, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    float4 a = p[0], b = p[1], c = p[2];
    uint4 d = reinterpret_cast<const uint4*>(nodes + i)[3];
    NodeBoxes n;
    n.lmin[0] = a.x; n.lmin[1] = a.y; n.lmin[2] = a.z;
    n.lmax[0] = a.w; n.lmax[1] = b.x; n.lmax[2] = b.y;
    n.rmin[0] = b.z; n.rmin[1] = b.w; n.rmin[2] = c.x;
    n.rmax[0] = c.y; n.rmax[1] = c.z; n.rmax[2] = c.w;
    n.left = d.x;
    n.right = d.y;
    return n;
}

// Short stack: K entries per thread in LDS (entry-major so a wave touches consecutive banks),
// deeper entries in a per-thread global overflow region.
struct TStack {
    uint32_t* lref;     // LDS [K * BS]
    float* lt;          // LDS [K * BS]
    uint32_t* gref;     // global [(cap-K) * gstride]
    float* gt;
    uint32_t tid, bs, K, gtid, gstride, cap;
};

__device__ __forceinline__ void spush(const TStack& S, uint32_t& sp, uint32_t ref, float t) {
    if (sp < S.K) {
        S.lref[sp * S.bs + S.tid] = ref;
        S.lt[sp * S.bs + S.tid] = t;
    } else {
        size_t o = (size_t)(sp - S.K) * S.gstride + S.gtid;
        S.gref[o] = ref;
        S.gt[o] = t;
    }
    ++sp;
}
__device__ __forceinline__ void spop(const TStack& S, uint32_t& sp, uint32_t& ref, float& t) {
    --sp;
    if (sp < S.K) {
        ref = S.lref[sp * S.bs + S.tid];
        t = S.lt[sp * S.bs + S.tid];
    } else {
        size_t o = (size_t)(sp - S.K) * S.gstride + S.gtid;
        ref = S.gref[o];
        t = S.gt[o];
    }
}

struct TraceCounters {
    uint32_t nodes, tris;
};

// Closest hit (ANY = false) or any hit (ANY = true).  Returns the leaf-order triangle index
// of the hit, or ~0u.  *overflow is set when the stack would exceed its capacity.
template <bool ANY, bool COUNT>
__device__ uint32_t traverse(const DScene& sc, const TRay& r, const TStack& S, float& tHit, float& uHit, float& vHit,
                             TraceCounters& cnt, uint32_t* overflow) {
    float closest = r.tMax;
    uint32_t hit = ~0u;
    float hu = 0.f, hv = 0.f;
    if (!sc.hasBVH) return ~0u;
    float tn;
    if (COUNT) ++cnt.nodes;
    if (!slab(sc.rootMin, sc.rootMax, r, tn)) return ~0u;
    uint32_t cur = sc.rootRef;
    uint32_t sp = 0;
    while (true) {
        if (refIsLeaf(cur)) {
            uint32_t first = refIndex(cur), n = refCount(cur);
            for (uint32_t k = 0; k < n; ++k) {
                DTri T = loadTriGeom(sc.tris, first + k);
                float t, u, v;
                if (COUNT) ++cnt.tris;
                if (hitTri(T, r, closest, t, u, v)) {
                    closest = t;
                    hit = first + k;
                    hu = u;
                    hv = v;
                    if (ANY) {
                        tHit = t;
                        return hit;
                    }
                }
            }
        } else {
            NodeBoxes nd = loadNode(sc.nodes, refIndex(cur));
            if (COUNT) ++cnt.nodes;
            float tL, tR;
            bool hL = slab(nd.lmin, nd.lmax, r, tL) && !(tL > closest);
            bool hR = slab(nd.rmin, nd.rmax, r, tR) && !(tR > closest);
            if (hL && hR) {
                if (sp + 1u >= S.cap) {
                    *overflow = 1u;
                    break;
                }
                if (tL < tR) {
                    spush(S, sp, nd.right, tR);
                    cur = nd.left;
                } else {
                    spush(S, sp, nd.left, tL);
                    cur = nd.right;
                }
                continue;
            }
            if (hL) { cur = nd.left; continue; }
            if (hR) { cur = nd.right; continue; }
        }
        bool found = false;
        while (sp > 0u) {
            float t;
            spop(S, sp, cur, t);
            if (!(t > closest)) { found = true; break; }
        }
        if (!found) break;
    }
    tHit = closest;
    uHit = hu;
    vHit = hv;
    return hit;
}

}  // namespace ahip
