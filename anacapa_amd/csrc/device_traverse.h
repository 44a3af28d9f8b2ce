// device_traverse.h — BVH traversal over 128-byte two-level nodes with an LDS short stack.
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
// The kernels walk DNode4 records (two binary levels per record) in exactly the binary
// traversal's order and with its culls (see traverse below); the entry distance stored with a
// stacked slot is the one the reference recomputes when it pops that node, and the cull test is
// applied at push and again at pop.  Any-hit queries (shadow rays) return whether a hit
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
    V3 d;       // original direction (the curve tests use it, CurveBrute.cpp:41-47, 478-498)
    V3 inv;     // 1 / safe(d)
    V3 dir;     // 1 / inv, the direction Moller-Trumbore sees (BVHBackend.cpp:267)
    float tMin, tMax;
    uint32_t neg;   // bit k set when d[k] < 0
};

__device__ __forceinline__ TRay makeTRay(V3 o, V3 d, float tMin, float tMax) {
#pragma clang fp contract(off)
    TRay r;
    r.o = o;
    r.d = d;
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

// Slab test of one slot of a DNode4 given its near and far planes (already chosen by the
// direction sign); also returns the per-axis near distances the group entry is built from.
// The arithmetic is that of slab() above, so a slot's entry distance and verdict are the bits
// the binary tree's test of the same box gives.
struct SlotSlab {
    float t0x, t0y, t0z, tmin;
    bool hit;
};
__device__ __forceinline__ SlotSlab slotSlab(float nx, float ny, float nz, float fx, float fy, float fz,
                                             const TRay& r) {
#pragma clang fp contract(off)
    SlotSlab o;
    o.t0x = (nx - r.o.x) * r.inv.x;
    o.t0y = (ny - r.o.y) * r.inv.y;
    o.t0z = (nz - r.o.z) * r.inv.z;
    float t1x = (fx - r.o.x) * r.inv.x, t1y = (fy - r.o.y) * r.inv.y, t1z = (fz - r.o.z) * r.inv.z;
    o.tmin = fmaxf(fmaxf(fmaxf(r.tMin, o.t0x), o.t0y), o.t0z);
    float tmax = fminf(fminf(fminf(r.tMax, t1x), t1y), t1z);
    o.hit = tmax > o.tmin;
    return o;
}

// Entry distance of the union of two slot boxes.  The near plane of the union on each axis is
// the nearer of the two slots' planes, and (p - o) * inv rounds monotonically in p, so the
// union's per-axis distance is the smaller of the slots' distances, bit for bit what the slab
// test of the union box itself computes.
__device__ __forceinline__ float groupEntry(const SlotSlab& a, const SlotSlab& b, float tMin) {
    return fmaxf(fmaxf(fmaxf(tMin, fminf(a.t0x, b.t0x)), fminf(a.t0y, b.t0y)), fminf(a.t0z, b.t0z));
}

// Short stack: K entries per thread in LDS (entry-major so a wave touches consecutive banks),
// each entry one 8-byte (reference, entry distance) pair; deeper entries in a per-thread global
// overflow region.  The LDS side is typed in the LDS address space so the compiler keeps the
// LDS and global accesses apart (ds_read / global_load) instead of merging them into one
// generic-address load.
typedef __attribute__((address_space(3))) uint64_t lds_u64;
struct TStack {
    lds_u64* l;         // LDS [K * BS]: reference | entry distance bits << 32
    uint32_t* gref;     // global [(cap-K) * gstride]
    float* gt;
    uint32_t tid, bs, K, gtid, gstride, cap;
    __device__ TStack(uint64_t* lds, uint32_t* gr, float* g, uint32_t tid_, uint32_t bs_, uint32_t K_, uint32_t gtid_,
                      uint32_t gs, uint32_t cap_)
        : l((lds_u64*)lds), gref(gr), gt(g), tid(tid_), bs(bs_), K(K_), gtid(gtid_), gstride(gs), cap(cap_) {}
};

__device__ __forceinline__ void spush(const TStack& S, uint32_t& sp, uint32_t ref, float t) {
    if (sp < S.K) {
        S.l[sp * S.bs + S.tid] = (uint64_t)ref | ((uint64_t)__float_as_uint(t) << 32);
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
        uint64_t e = S.l[sp * S.bs + S.tid];
        ref = (uint32_t)e;
        t = __uint_as_float((uint32_t)(e >> 32));
    } else {
        size_t o = (size_t)(sp - S.K) * S.gstride + S.gtid;
        ref = S.gref[o];
        t = S.gt[o];
    }
}

struct TraceCounters {
    uint32_t nodes, tris;
};

// Per-ray traversal state, so a kernel can advance many rays one step at a time.
struct TState {
    uint32_t cur;        // node or leaf reference being visited
    uint32_t sp;         // stack depth
    uint32_t triIdx;     // next triangle of the current leaf
    uint32_t triLeft;    // triangles of the current leaf still to test
    uint32_t hit;        // leaf-order triangle index of the closest hit so far, or ~0u
    float closest, hu, hv;
};

// Root test; false when the ray misses the scene box (or there is no BVH) and the query is over.
__device__ __forceinline__ bool traverseBegin(const DScene& sc, const TRay& r, TState& st) {
    st.cur = sc.rootRef;
    st.sp = 0u;
    st.triIdx = 0u;
    st.triLeft = 0u;
    st.hit = ~0u;
    st.closest = r.tMax;
    st.hu = 0.f;
    st.hv = 0.f;
    if (!sc.hasBVH) return false;
    float tn;
    return slab(sc.rootMin, sc.rootMax, r, tn);   // root box from the kernel argument
}

// One step: a node visit, or ONE triangle of the current leaf, then a pop when the step leaves
// nothing to enter.  Returns true when the query is over (stack exhausted, any-hit found, or
// stack overflow, which sets *overflow).  The first three 16-byte loads are common to a node and
// a triangle, so lanes at a leaf and lanes at a node issue their loads together and the wave
// waits once per step.
template <bool ANY, bool COUNT>
__device__ __forceinline__ bool traverseStep(const DScene& sc, const TRay& r, const TStack& S, TState& st,
                                             TraceCounters& cnt, uint32_t* overflow) {
    if (st.triLeft == 0u && refIsLeaf(st.cur)) {
        st.triIdx = refIndex(st.cur);
        st.triLeft = refCount(st.cur);
    }
    const bool isTri = st.triLeft != 0u;
    const float4* p = isTri ? reinterpret_cast<const float4*>(sc.tris + st.triIdx)
                            : reinterpret_cast<const float4*>(sc.nodes4 + refIndex(st.cur));
    const float4 mnx = p[0], mny = p[1], mnz = p[2];
    float4 mxx, mxy, mxz;
    uint4 rf;
    if (!isTri) {
        mxx = p[3];
        mxy = p[4];
        mxz = p[5];
        rf = reinterpret_cast<const uint4*>(p)[6];
    }
    if (isTri) {
        DTri T;
        T.v0[0] = mnx.x; T.v0[1] = mnx.y; T.v0[2] = mnx.z;
        T.e1[0] = mnx.w; T.e1[1] = mny.x; T.e1[2] = mny.y;
        T.e2[0] = mny.z; T.e2[1] = mny.w; T.e2[2] = mnz.x;
        float t, u, v;
        if (COUNT) ++cnt.tris;
        if (hitTri(T, r, st.closest, t, u, v)) {
            st.closest = t;
            st.hit = st.triIdx;
            st.hu = u;
            st.hv = v;
            if (ANY) return true;
        }
        ++st.triIdx;
        if (--st.triLeft != 0u) return false;
    } else {
        if (COUNT) ++cnt.nodes;
        const bool nx = (r.neg & 1u) != 0u, ny = (r.neg & 2u) != 0u, nz = (r.neg & 4u) != 0u;
        const float4 ax = nx ? mxx : mnx, bx = nx ? mnx : mxx;
        const float4 ay = ny ? mxy : mny, by = ny ? mny : mxy;
        const float4 az = nz ? mxz : mnz, bz = nz ? mnz : mxz;
        const SlotSlab s0 = slotSlab(ax.x, ay.x, az.x, bx.x, by.x, bz.x, r);
        const SlotSlab s1 = slotSlab(ax.y, ay.y, az.y, bx.y, by.y, bz.y, r);
        const SlotSlab s2 = slotSlab(ax.z, ay.z, az.z, bx.z, by.z, bz.z, r);
        const SlotSlab s3 = slotSlab(ax.w, ay.w, az.w, bx.w, by.w, bz.w, r);
        uint32_t q0, q1, q2, q3;
        float u0, u1, u2, u3;
        bool v0, v1, v2, v3;
        if constexpr (ANY) {
            // An any-hit query ends at its first hit and culls nothing before it (a slot that
            // passes the slab test enters before tMax = closest), and the answer does not
            // depend on the order, so the live slots are taken in record order.
            v0 = s0.hit & (rf.x != kRefEmpty);
            v1 = s1.hit & (rf.y != kRefEmpty);
            v2 = s2.hit & (rf.z != kRefEmpty);
            v3 = s3.hit & (rf.w != kRefEmpty);
            q0 = rf.x; q1 = rf.y; q2 = rf.z; q3 = rf.w;
            u0 = s0.tmin; u1 = s1.tmin; u2 = s2.tmin; u3 = s3.tmin;
        } else {
            const bool l0 = s0.hit & !(s0.tmin > st.closest) & (rf.x != kRefEmpty);
            const bool l1 = s1.hit & !(s1.tmin > st.closest) & (rf.y != kRefEmpty);
            const bool l2 = s2.hit & !(s2.tmin > st.closest) & (rf.z != kRefEmpty);
            const bool l3 = s3.hit & !(s3.tmin > st.closest) & (rf.w != kRefEmpty);
            // order within each group, then of the groups (left first only when strictly
            // nearer); the group entries are computed unconditionally (no branch)
            const float gA = groupEntry(s0, s1, r.tMin), gB = groupEntry(s2, s3, r.tMin);
            const bool aSwap = l0 & l1 & !(s0.tmin < s1.tmin);
            const bool bSwap = l2 & l3 & !(s2.tmin < s3.tmin);
            const bool gSwap = (l0 | l1) & (l2 | l3) & !(gA < gB);
            const uint32_t ra0 = aSwap ? rf.y : rf.x, ra1 = aSwap ? rf.x : rf.y;
            const float ta0 = aSwap ? s1.tmin : s0.tmin, ta1 = aSwap ? s0.tmin : s1.tmin;
            const bool la0 = aSwap ? l1 : l0, la1 = aSwap ? l0 : l1;
            const uint32_t rb0 = bSwap ? rf.w : rf.z, rb1 = bSwap ? rf.z : rf.w;
            const float tb0 = bSwap ? s3.tmin : s2.tmin, tb1 = bSwap ? s2.tmin : s3.tmin;
            const bool lb0 = bSwap ? l3 : l2, lb1 = bSwap ? l2 : l3;
            // visiting order q0, q1, q2, q3
            q0 = gSwap ? rb0 : ra0; q1 = gSwap ? rb1 : ra1; q2 = gSwap ? ra0 : rb0; q3 = gSwap ? ra1 : rb1;
            u0 = gSwap ? tb0 : ta0; u1 = gSwap ? tb1 : ta1; u2 = gSwap ? ta0 : tb0; u3 = gSwap ? ta1 : tb1;
            v0 = gSwap ? lb0 : la0; v1 = gSwap ? lb1 : la1; v2 = gSwap ? la0 : lb0; v3 = gSwap ? la1 : lb1;
        }
        const uint32_t live = (uint32_t)v0 + (uint32_t)v1 + (uint32_t)v2 + (uint32_t)v3;
        if (live != 0u) {
            if (st.sp + live - 1u >= S.cap) {
                *overflow = 1u;
                return true;
            }
            // push the live slots after the first, farthest in the order first
            if (v3 && (v0 || v1 || v2)) spush(S, st.sp, q3, u3);
            if (v2 && (v0 || v1)) spush(S, st.sp, q2, u2);
            if (v1 && v0) spush(S, st.sp, q1, u1);
            st.cur = v0 ? q0 : v1 ? q1 : v2 ? q2 : q3;
            return false;
        }
    }
    if constexpr (ANY) {
        // nothing is culled before the first hit: the top entry is always taken
        if (st.sp == 0u) return true;
        float t;
        spop(S, st.sp, st.cur, t);
        return false;
    } else {
        while (st.sp > 0u) {
            float t;
            spop(S, st.sp, st.cur, t);
            if (!(t > st.closest)) return false;
        }
        return true;
    }
}

// Closest hit (ANY = false) or any hit (ANY = true).  Returns the leaf-order triangle index
// of the hit, or ~0u.  *overflow is set when the stack would exceed its capacity.
//
// Each DNode4 visit stands for binary node X and its children A (slots 0,1) and B (slots 2,3).
// The binary traversal enters the nearer of A and B (B on a tie), then the nearer of that
// group's children (the second on a tie), and pops the rest later, culling each against the
// closest hit at that moment.  Here the slots are put in that same order (group entries from
// groupEntry, slot entries from slotSlab), the first live slot is entered and the other live
// slots are pushed in reverse.  Culling is monotone (the closest distance only shrinks and a
// child's entry is never below its group's), so testing each slot against the closest hit when
// it is entered or popped skips exactly the subtrees the binary traversal skips, leaves are
// reached in the same order, and ties between equally distant triangles resolve identically.
template <bool ANY, bool COUNT>
__device__ uint32_t traverse(const DScene& sc, const TRay& r, const TStack& S, float& tHit, float& uHit, float& vHit,
                             TraceCounters& cnt, uint32_t* overflow) {
    TState st;
    if (!traverseBegin(sc, r, st)) return ~0u;
    while (!traverseStep<ANY, COUNT>(sc, r, S, st, cnt, overflow)) {
    }
    tHit = st.closest;
    uHit = st.hu;
    vHit = st.hv;
    return st.hit;
}

// ---------------------------------------------------------------- curves (CurveBrute.cpp)
// Ray-space ribbon test and the curve BVH, with the reference's semantics: CVs projected into a
// Frisvad frame around the ray direction; recursive de Casteljau midpoint splits to depth 6 with
// hull pruning against [tMin, best]; a leaf accepts the closest chord point within the
// interpolated half width, strictly nearer than the best hit; the left half is searched before
// the right half, so the right half is pruned with the left half's result.  The recursion runs
// on a private stack of pending right halves (at most one per level).
struct CurveHit {
    float tSeg;
    V3 tang;
    uint32_t rec;
};

// Frisvad basis (Types.h:108-114) without contraction: the ray-space frame of the curve test.
__device__ __forceinline__ void onbNC(V3 n, V3& t, V3& bt) {
#pragma clang fp contract(off)
    float sign = copysignf(1.f, n.z);
    float a = -1.f / (sign + n.z);
    float b = n.x * n.y * a;
    t = V3{1.f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    bt = V3{b, sign + n.y * n.y * a, -n.y};
}

struct SubFrame {
    V3 P[4];
    float hw0, hw3, tOff, tScale;
    int depth;
};

__device__ __forceinline__ V3 projNC(V3 p, V3 o, V3 x, V3 y, V3 z) {
#pragma clang fp contract(off)
    V3 q{p.x - o.x, p.y - o.y, p.z - o.z};
    return V3{dotNC(q, x), dotNC(q, y), dotNC(q, z)};
}
__device__ __forceinline__ V3 midNC(V3 a, V3 b) {
#pragma clang fp contract(off)
    return V3{(a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f};
}

// The root frame of a segment record: its control points, end half widths and parameter range,
// `depth` de Casteljau levels above the chord test (the reference's MAX_DEPTH = 6,
// CurveBrute.cpp:194; 1..6, which the 7-entry stack of pending halves holds).
__device__ __forceinline__ SubFrame rootFrame(const DSeg& S, int depth) {
    SubFrame f;
    f.P[0] = V3{S.P[0], S.P[1], S.P[2]};
    f.P[1] = V3{S.P[3], S.P[4], S.P[5]};
    f.P[2] = V3{S.P[6], S.P[7], S.P[8]};
    f.P[3] = V3{S.P[9], S.P[10], S.P[11]};
    f.hw0 = S.hw0;
    f.hw3 = S.hw3;
    f.tOff = S.tOff;
    f.tScale = S.tScale;
    f.depth = depth;
    return f;
}

// One frame of the reference's recursion (CurveBrute.cpp:85-176), as an explicit depth-first
// walk: the frame `f` is tested in ray space against [tMin, best]; a frame that passes and is not
// a leaf is split, its left half becomes `f` and its right half waits on `st` (at most one per
// level, in the order the recursion visits them); otherwise the next pending half is taken.
// Returns true when no frame is left.  `acc` is set when the leaf chord of `f` was accepted,
// strictly nearer than `best`, with its distance `tr`, strand parameter `tSeg` and unit tangent.
__device__ __forceinline__ bool segmentStep(SubFrame& f, SubFrame* st, int& sp, V3 o, V3 d, float tMin, V3 fx,
                                            V3 fy, float best, bool& acc, float& tr, float& tSeg, V3& tang) {
#pragma clang fp contract(off)
    acc = false;
    {
        V3 c0 = projNC(f.P[0], o, fx, fy, d), c1 = projNC(f.P[1], o, fx, fy, d);
        V3 c2 = projNC(f.P[2], o, fx, fy, d), c3 = projNC(f.P[3], o, fx, fy, d);
        float mh = fmaxf(f.hw0, f.hw3);
        float zMin = fminf(fminf(fminf(c0.z, c1.z), c2.z), c3.z), zMax = fmaxf(fmaxf(fmaxf(c0.z, c1.z), c2.z), c3.z);
        if (zMin - mh > best) goto next;
        if (zMax + mh < tMin) goto next;
        float xMin = fminf(fminf(fminf(c0.x, c1.x), c2.x), c3.x), xMax = fmaxf(fmaxf(fmaxf(c0.x, c1.x), c2.x), c3.x);
        float yMin = fminf(fminf(fminf(c0.y, c1.y), c2.y), c3.y), yMax = fmaxf(fmaxf(fmaxf(c0.y, c1.y), c2.y), c3.y);
        float cx = (0.f < xMin) ? xMin : (0.f > xMax ? xMax : 0.f);
        float cy = (0.f < yMin) ? yMin : (0.f > yMax ? yMax : 0.f);
        if (cx * cx + cy * cy > mh * mh) goto next;
        if (f.depth == 0) {
            float ax = c0.x, ay = c0.y, dx = c3.x - ax, dy = c3.y - ay;
            float l2 = dx * dx + dy * dy;
            float tl = l2 < 1e-12f ? 0.f : stdmax(0.f, stdmin(1.f, (-ax * dx - ay * dy) / l2));
            float px = ax + dx * tl, py = ay + dy * tl;
            float d2 = px * px + py * py;
            float hw = f.hw0 * (1.f - tl) + f.hw3 * tl;
            if (d2 > hw * hw) goto next;
            float t = c0.z + (c3.z - c0.z) * tl;
            if (t < tMin || t >= best) goto next;
            // bezierTangent (CurveBrute.cpp:59-64)
            float mt = 1.f - tl;
            float a0 = 3.f * mt * mt, a1 = 6.f * mt * tl, a2 = 3.f * tl * tl;
            V3 d10{f.P[1].x - f.P[0].x, f.P[1].y - f.P[0].y, f.P[1].z - f.P[0].z};
            V3 d21{f.P[2].x - f.P[1].x, f.P[2].y - f.P[1].y, f.P[2].z - f.P[1].z};
            V3 d32{f.P[3].x - f.P[2].x, f.P[3].y - f.P[2].y, f.P[3].z - f.P[2].z};
            V3 tg{d10.x * a0 + d21.x * a1 + d32.x * a2, d10.y * a0 + d21.y * a1 + d32.y * a2,
                  d10.z * a0 + d21.z * a1 + d32.z * a2};
            if (tg.x * tg.x + tg.y * tg.y + tg.z * tg.z < 1e-12f)
                tg = V3{f.P[3].x - f.P[0].x, f.P[3].y - f.P[0].y, f.P[3].z - f.P[0].z};
            float l = sqrtf(tg.x * tg.x + tg.y * tg.y + tg.z * tg.z);
            if (l > 1e-8f) {
                float il = 1.f / l;
                tg = V3{tg.x * il, tg.y * il, tg.z * il};
            } else {
                tg = V3{0.f, 0.f, 1.f};
            }
            acc = true;
            tr = t;
            tSeg = f.tOff + tl * f.tScale;
            tang = tg;
            goto next;
        }
        V3 q01 = midNC(f.P[0], f.P[1]), q12 = midNC(f.P[1], f.P[2]), q23 = midNC(f.P[2], f.P[3]);
        V3 r01 = midNC(q01, q12), r12 = midNC(q12, q23), m = midNC(r01, r12);
        float hm = (f.hw0 + f.hw3) * 0.5f, ts = f.tScale * 0.5f;
        SubFrame R;
        R.P[0] = m; R.P[1] = r12; R.P[2] = q23; R.P[3] = f.P[3];
        R.hw0 = hm; R.hw3 = f.hw3; R.tOff = f.tOff + ts; R.tScale = ts; R.depth = f.depth - 1;
        st[sp++] = R;
        f.P[1] = q01; f.P[2] = r01; f.P[3] = m;
        f.hw3 = hm; f.tScale = ts; f.depth = f.depth - 1;
        return false;
    }
next:
    if (sp == 0) return true;
    f = st[--sp];
    return false;
}

// Returns true when a strictly nearer hit on this segment was found (best, out updated).
__device__ bool segmentHit(const DSeg& S, const TRay& r, V3 fx, V3 fy, float& best, float& tSeg, V3& tang,
                           int depth) {
    SubFrame st[7];
    int sp = 0;
    SubFrame f = rootFrame(S, depth);
    bool hit = false;
    while (true) {
        bool acc;
        float t, ts;
        V3 tg;
        const bool done = segmentStep(f, st, sp, r.o, r.d, r.tMin, fx, fy, best, acc, t, ts, tg);
        if (acc) {
            best = t;
            tSeg = ts;
            tang = tg;
            hit = true;
        }
        if (done) break;
    }
    return hit;
}

// CurveBrute::aabbHit (CurveBrute.cpp:478-498)
// invD = 1 / r.d per axis, computed once per ray (the same quotient the box test divides out).
__device__ __forceinline__ bool curveBox(const CNode& n, const TRay& r, const V3& invD, float maxT) {
#pragma clang fp contract(off)
    float tmin = r.tMin, tmax = maxT;
    const float o[3] = {r.o.x, r.o.y, r.o.z}, dd[3] = {r.d.x, r.d.y, r.d.z}, id[3] = {invD.x, invD.y, invD.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (fabsf(dd[i]) < 1e-9f) {
            if (o[i] < n.bmin[i] || o[i] > n.bmax[i]) return false;
            continue;
        }
        float inv = id[i];
        float t0 = (n.bmin[i] - o[i]) * inv, t1 = (n.bmax[i] - o[i]) * inv;
        if (inv < 0.f) { float s = t0; t0 = t1; t1 = s; }
        tmin = stdmax(tmin, t0);
        tmax = stdmin(tmax, t1);
        if (tmax < tmin) return false;
    }
    return true;
}

// curveBox that also gives the entry distance (the larger of tMin and the near slab planes), by
// which a walk orders two children.
__device__ __forceinline__ bool curveBoxT(const float* bmin, const float* bmax, V3 o, V3 d, const V3& invD, float tMin,
                                          float maxT, float& entry) {
#pragma clang fp contract(off)
    float tmin = tMin, tmax = maxT;
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z}, id[3] = {invD.x, invD.y, invD.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (fabsf(dd[i]) < 1e-9f) {
            if (oo[i] < bmin[i] || oo[i] > bmax[i]) return false;
            continue;
        }
        float inv = id[i];
        float t0 = (bmin[i] - oo[i]) * inv, t1 = (bmax[i] - oo[i]) * inv;
        if (inv < 0.f) { float s = t0; t0 = t1; t1 = s; }
        tmin = stdmax(tmin, t0);
        tmax = stdmin(tmax, t1);
        if (tmax < tmin) return false;
    }
    entry = tmin;
    return true;
}

// A child-pair record as four 16-byte loads issued together (the box tests then run on registers).
__device__ __forceinline__ CWide loadCWide(const CWide* w, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(w + i);
    const float4 a = p[0], b = p[1], c = p[2], d = p[3];
    CWide r;
    r.a[0] = a.x; r.a[1] = a.y; r.a[2] = a.z; r.a[3] = a.w; r.a[4] = b.x; r.a[5] = b.y;
    r.b[0] = b.z; r.b[1] = b.w; r.b[2] = c.x; r.b[3] = c.y; r.b[4] = c.z; r.b[5] = c.w;
    r.refA = __float_as_uint(d.x);
    r.refB = __float_as_uint(d.y);
    r.pad0 = r.pad1 = 0u;
    return r;
}

__device__ __forceinline__ CNode loadCNode(const CNode* nodes, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    float4 a = p[0], b = p[1];
    CNode n;
    n.bmin[0] = a.x; n.bmin[1] = a.y; n.bmin[2] = a.z;
    n.bmax[0] = a.w; n.bmax[1] = b.x; n.bmax[2] = b.y;
    n.lp = __float_as_uint(b.z);
    n.rc = __float_as_uint(b.w);
    return n;
}

// Curve part of CurveBrute::trace (CurveBrute.cpp:398-440) after the triangle BVH: `best` is the
// triangle hit distance (or tMax).  ANY stops at the first accepted segment.
template <bool ANY, bool COUNT>
__device__ bool traverseCurves(const DScene& sc, const TRay& r, const TStack& S, float& best, CurveHit& ch,
                               TraceCounters& cnt, uint32_t* overflow) {
    if (!sc.hasCurves) return false;
    V3 fx, fy;
    onbNC(r.d, fx, fy);
    const V3 invD{1.f / r.d.x, 1.f / r.d.y, 1.f / r.d.z};   // used only where |d| >= 1e-9
    bool hit = false;
    uint32_t sp = 0;
    spush(S, sp, 0u, 0.f);
    while (sp > 0u) {
        uint32_t ni;
        float unused;
        spop(S, sp, ni, unused);
        CNode n = loadCNode(sc.cnodes, ni);
        if (COUNT) ++cnt.nodes;
        if (!curveBox(n, r, invD, best)) continue;
        if (n.rc & 0x80000000u) {
            uint32_t cntSeg = n.rc & 0x7fffffffu;
            for (uint32_t k = 0; k < cntSeg; ++k) {
                const DSeg& Sg = sc.segs[n.lp + k];
                if (COUNT) ++cnt.tris;
                float tSeg;
                V3 tg;
                if (segmentHit(Sg, r, fx, fy, best, tSeg, tg, (int)sc.curveDepth)) {
                    hit = true;
                    ch.tSeg = tSeg;
                    ch.tang = tg;
                    ch.rec = n.lp + k;
                    if (ANY) return true;
                }
            }
        } else {
            if (sp + 2u >= S.cap) {
                *overflow = 1u;
                return hit;
            }
            spush(S, sp, n.rc, 0.f);
            spush(S, sp, n.lp, 0.f);
        }
    }
    return hit;
}

}  // namespace ahip
