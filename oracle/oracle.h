// oracle.h — shared declarations of the CPU oracle (test infrastructure only; see oracle.cpp).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../include/anacapa_hip.h"

namespace orc {

// ---------------------------------------------------------------- math (core/Types.h)
struct V2 { float x = 0.f, y = 0.f; };
struct V3 {
    float x = 0.f, y = 0.f, z = 0.f;
    V3() = default;
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float& operator[](int i) { return (&x)[i]; }
    float operator[](int i) const { return (&x)[i]; }
};
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
static inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
static inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }   // Types.h:70-72
static inline V3 cross(V3 a, V3 b) {                                                 // Types.h:74-78
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline float lengthSq(V3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
static inline float length(V3 v) { return std::sqrt(lengthSq(v)); }
static inline V3 normalize(V3 v) { return v * (1.f / length(v)); }                 // Types.h:80-84
static inline V3 safeNormalize(V3 v, V3 fb = {0, 0, 1}) {                           // Types.h:86-89
    float len = length(v);
    return len > 1e-8f ? v * (1.f / len) : fb;
}
static inline float absDot(V3 a, V3 b) { return std::abs(dot(a, b)); }
static inline bool isBlack(V3 s) { return s.x == 0.f && s.y == 0.f && s.z == 0.f; }
static inline float fmaxs(float a, float b) { return (a < b) ? b : a; }               // std::max
static inline float fmins(float a, float b) { return (b < a) ? b : a; }               // std::min
static inline float maxComponent(V3 v) {                                              // Types.h:64
    float m = v.x;
    if (m < v.y) m = v.y;
    if (m < v.z) m = v.z;
    return m;
}
static inline bool isFinite(V3 v) {
    return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z);
}
static inline float luminance(V3 s) {                                                 // Types.h:245-248
    return 0.2722287f * s.x + 0.6740818f * s.y + 0.0536895f * s.z;
}
static inline void onb(V3 n, V3& t, V3& bt) {                                         // Types.h:108-114
    float sign = std::copysign(1.f, n.z);
    float a = -1.f / (sign + n.z);
    float b = n.x * n.y * a;
    t = V3{1.f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    bt = V3{b, sign + n.y * n.y * a, -n.y};
}
// Identity Mat4f::transformPoint / transformNormal (Types.h:144-167), as BVHBackend::commit
// applies them to static meshes (BVHBackend.cpp:26, 52-61): the same sums of products.
static inline V3 identityPoint(V3 p) {
    float x = 1.f * p.x + 0.f * p.y + 0.f * p.z + 0.f;
    float y = 0.f * p.x + 1.f * p.y + 0.f * p.z + 0.f;
    float z = 0.f * p.x + 0.f * p.y + 1.f * p.z + 0.f;
    float w = 0.f * p.x + 0.f * p.y + 0.f * p.z + 1.f;
    return {x / w, y / w, z / w};
}
static inline V3 identityNormal(V3 n) {
    return {1.f * n.x + 0.f * n.y + 0.f * n.z, 0.f * n.x + 1.f * n.y + 0.f * n.z,
            0.f * n.x + 0.f * n.y + 1.f * n.z};
}

constexpr float kInf = std::numeric_limits<float>::infinity();
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float k2Pi = 6.28318530717958647692f;

struct Ray {                                                                          // Types.h:253-266
    V3 o, d;
    float tMin = 1e-4f, tMax = kInf;
};
static inline Ray spawnRay(V3 origin, V3 normal, V3 dir) {                            // Types.h:269-275
    V3 off = normal * 1e-4f;
    if (dot(dir, normal) < 0.f) off = -off;
    Ray r;
    r.o = origin + off;
    r.d = dir;
    return r;
}
static inline Ray spawnRayTo(V3 origin, V3 normal, V3 target) {                       // Types.h:277-284
    V3 dir = target - origin;
    float dist = length(dir);
    dir = dir * (1.f / dist);
    Ray r = spawnRay(origin, normal, dir);
    r.tMax = dist * (1.f - 1e-4f);
    return r;
}

// ---------------------------------------------------------------- PCG32 (sampling/PCGRng.h:21-41)
struct Pcg {
    uint64_t state = 0x853c49e6748fea9bULL, inc = 0xda3e39cb94b95bdbULL;
    Pcg() = default;
    explicit Pcg(uint64_t seed, uint64_t seq = 1) {
        state = 0;
        inc = (seq << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    float nextFloat() { return (float)next() * 2.3283064365386963e-10f; }
};

// ---------------------------------------------------------------- Halton (sampling/HaltonSampler.h)
static const uint32_t kPrimes[128] = {                                                // HaltonSampler.h:68-79
    2,   3,   5,   7,   11,  13,  17,  19,  23,  29,  31,  37,  41,  43,  47,  53,  59,  61,  67,
    71,  73,  79,  83,  89,  97,  101, 103, 107, 109, 113, 127, 131, 137, 139, 149, 151, 157, 163,
    167, 173, 179, 181, 191, 193, 197, 199, 211, 223, 227, 229, 233, 239, 241, 251, 257, 263, 269,
    271, 277, 281, 283, 293, 307, 311, 313, 317, 331, 337, 347, 349, 353, 359, 367, 373, 379, 383,
    389, 397, 401, 409, 419, 421, 431, 433, 439, 443, 449, 457, 461, 463, 467, 479, 487, 491, 499,
    503, 509, 521, 523, 541, 547, 557, 563, 569, 571, 577, 587, 593, 599, 601, 607, 613, 617, 619,
    631, 641, 643, 647, 653, 659, 661, 673, 677, 683, 691, 701, 709, 719};

// HaltonSampler.h:82-93.  The reference Release build (-O3 -mfma) contracts the
// accumulation `result += float(d) * invBN` into one fused multiply-add.
static float radicalInverse(uint32_t b, uint64_t n) {
    float invB = 1.f / (float)b;
    float invBN = 1.f, result = 0.f;
    while (n > 0) {
        uint64_t d = n % b;
        n /= b;
        invBN *= invB;
        result = std::fma((float)d, invBN, result);
    }
    return result;
}

struct Halton {                                                                       // HaltonSampler.h:20-113
    uint64_t seed = 0;
    uint32_t px = 0, py = 0, idx = 0, dim = 0;
    Pcg rng;
    void start(uint32_t x, uint32_t y, uint32_t i) {                                  // :33-41
        px = x; py = y; idx = i; dim = 0;
        rng = Pcg(seed ^ ((uint64_t)x << 32 | y));
    }
    float scrambled(uint32_t d, uint64_t n) const {                                   // :96-104
        Pcg r(seed ^ ((uint64_t)px * 2654435761ULL) ^ ((uint64_t)py * 805459861ULL) ^
              ((uint64_t)d * 999999893ULL));
        uint64_t scramble = r.next();
        return radicalInverse(kPrimes[d], n ^ scramble);
    }
    float get1D() {                                                                   // :43-50
        uint32_t d = dim++;
        if (d >= 128) return rng.nextFloat();
        return scrambled(d, idx);
    }
    V2 get2D() {                                                                      // :52-54
        V2 v;
        v.x = get1D();
        v.y = get1D();
        return v;
    }
};

// ---------------------------------------------------------------- camera (IIntegrator.h:27-143)
struct Camera {
    V3 origin, llc, hor, ver;
    uint32_t w = 1, h = 1;
};
static Camera makePinhole(V3 from, V3 at, V3 up, float vfov, uint32_t W, uint32_t H) {  // :49-74
    float theta = vfov * 3.14159265f / 180.f;
    float hh = std::tan(theta * 0.5f);
    float aspect = (float)W / (float)H;
    V3 w = normalize(from - at);
    V3 u = normalize(cross(up, w));
    V3 v = cross(w, u);
    float vpH = 2.f * hh, vpW = aspect * vpH;
    Camera c;
    c.origin = from;
    c.hor = u * vpW;
    c.ver = v * vpH;
    c.llc = from - u * (vpW * 0.5f) - v * (vpH * 0.5f) - w;
    c.w = W;
    c.h = H;
    return c;
}
static Ray generateRay(const Camera& c, uint32_t px, uint32_t py, float jx, float jy) {  // :90-108
    float s = ((float)px + jx) / (float)c.w;
    float t = ((float)(c.h - 1 - py) + jy) / (float)c.h;
    V3 target = c.llc + c.hor * s + c.ver * t;
    Ray r;
    r.o = c.origin;
    r.d = normalize(target - c.origin);
    return r;
}

// ---------------------------------------------------------------- BVH (accel/BVHBackend.{h,cpp})
struct BBox {
    V3 mn{kInf, kInf, kInf}, mx{-kInf, -kInf, -kInf};
    void expand(V3 p) {
        mn = {fmins(mn.x, p.x), fmins(mn.y, p.y), fmins(mn.z, p.z)};
        mx = {fmaxs(mx.x, p.x), fmaxs(mx.y, p.y), fmaxs(mx.z, p.z)};
    }
    void expand(const BBox& o) {
        mn = {fmins(mn.x, o.mn.x), fmins(mn.y, o.mn.y), fmins(mn.z, o.mn.z)};
        mx = {fmaxs(mx.x, o.mx.x), fmaxs(mx.y, o.mx.y), fmaxs(mx.z, o.mx.z)};
    }
    V3 centroid() const { return (mn + mx) * 0.5f; }
    V3 diag() const { return mx - mn; }
};
struct Node {                                                                         // BVHBackend.h:25-39
    float bmin[3];
    uint32_t dataA;
    float bmax[3];
    uint32_t dataB;
};
constexpr uint32_t kLeaf = 0x80000000u;
struct Tri {                                                                          // BVHBackend.h:49-59
    V3 v0, e1, e2, n, sn0, sn1, sn2;
    V2 uv0, uv1, uv2;
    uint32_t meshID, primID;
};
struct PrimInfo { BBox b; V3 c; uint32_t orig; };

struct Bvh {
    std::vector<Node> nodes;
    std::vector<Tri> tris;
    std::vector<uint32_t> prims;

    // Build (BVHBackend.cpp:17-228), in oracle_bvh.cpp: that translation unit is compiled
    // with the reference's Release flags so its SAH arithmetic contracts the same way.
    int sahSplit(const std::vector<PrimInfo>& pi, uint32_t s, uint32_t e, const BBox& cb, int& axisOut) const;
    uint32_t build(std::vector<PrimInfo>& pi, uint32_t s, uint32_t e);
    void commit(const ahip_scene_desc* d);

    struct R4 { V3 o, inv; int neg[3]; float tMin, tMax; };
    static R4 makeRay4(const Ray& r) {                                                // BVHBackend.cpp:233-247
        R4 q;
        q.o = r.o; q.tMin = r.tMin; q.tMax = r.tMax;
        auto safe = [](float v) { return std::abs(v) > 1e-9f ? v : 1e-9f; };
        q.inv = {1.f / safe(r.d.x), 1.f / safe(r.d.y), 1.f / safe(r.d.z)};
        q.neg[0] = r.d.x < 0.f; q.neg[1] = r.d.y < 0.f; q.neg[2] = r.d.z < 0.f;
        return q;
    }
    static bool hitBox(const Node& nd, const R4& r, float& tNear) {                   // :249-262
        const float* bb[2] = {nd.bmin, nd.bmax};
        float t0m = r.tMin, t1m = r.tMax;
        for (int i = 0; i < 3; ++i) {
            float t0 = (bb[r.neg[i]][i] - r.o[i]) * r.inv[i];
            float t1 = (bb[1 - r.neg[i]][i] - r.o[i]) * r.inv[i];
            t0m = fmaxs(t0m, t0);
            t1m = fmins(t1m, t1);
            if (t1m <= t0m) return false;
        }
        tNear = t0m;
        return true;
    }
    static bool hitTri(const Tri& tr, const R4& r, float& t, float& u, float& v) {    // :264-283
        V3 dir = {1.f / r.inv.x, 1.f / r.inv.y, 1.f / r.inv.z};
        V3 h = cross(dir, tr.e2);
        float a = dot(tr.e1, h);
        if (std::abs(a) < 1e-9f) return false;
        float f = 1.f / a;
        V3 s = r.o - tr.v0;
        u = f * dot(s, h);
        if (u < 0.f || u > 1.f) return false;
        V3 q = cross(s, tr.e1);
        v = f * dot(dir, q);
        if (v < 0.f || u + v > 1.f) return false;
        t = f * dot(tr.e2, q);
        return t >= r.tMin && t <= r.tMax;
    }

    // Closest hit (BVHBackend.cpp:335-404); returns triangle index or ~0u.
    uint32_t trace(const Ray& ray, float& tHit, float& uHit, float& vHit, uint64_t* nodesVisited = nullptr,
                   uint64_t* trisTested = nullptr) const {
        if (nodes.empty()) return ~0u;
        R4 r = makeRay4(ray);
        uint32_t st[64];
        int top = 0;
        st[top++] = 0;
        float closest = r.tMax, tn;
        uint32_t hit = ~0u;
        float hu = 0.f, hv = 0.f;
        while (top > 0) {
            uint32_t ni = st[--top];
            const Node& nd = nodes[ni];
            if (nodesVisited) ++*nodesVisited;
            if (!hitBox(nd, r, tn) || tn > closest) continue;
            if (nd.dataB & kLeaf) {
                uint32_t cnt = nd.dataB & ~kLeaf;
                for (uint32_t i = 0; i < cnt; ++i) {
                    uint32_t idx = prims[nd.dataA + i];
                    R4 rc = r;
                    rc.tMax = closest;
                    float t, u, v;
                    if (trisTested) ++*trisTested;
                    if (hitTri(tris[idx], rc, t, u, v)) { closest = t; hit = idx; hu = u; hv = v; }
                }
            } else {
                uint32_t L = ni + 1, Rr = nd.dataA;
                float tL = kInf, tR = kInf;
                bool hL = hitBox(nodes[L], r, tL), hR = hitBox(nodes[Rr], r, tR);
                if (hL && hR) {
                    if (tL < tR) { st[top++] = Rr; st[top++] = L; }
                    else { st[top++] = L; st[top++] = Rr; }
                } else if (hL) st[top++] = L;
                else if (hR) st[top++] = Rr;
            }
        }
        tHit = closest; uHit = hu; vHit = hv;
        return hit;
    }

    // BVHBackend::occluded (BVHBackend.cpp:409-444)
    bool occluded(const Ray& ray) const {
        if (nodes.empty()) return false;
        R4 r = makeRay4(ray);
        uint32_t st[64];
        int top = 0;
        st[top++] = 0;
        float tn, t, u, v;
        while (top > 0) {
            uint32_t ni = st[--top];
            const Node& nd = nodes[ni];
            if (!hitBox(nd, r, tn)) continue;
            if (nd.dataB & kLeaf) {
                for (uint32_t i = 0; i < (nd.dataB & ~kLeaf); ++i)
                    if (hitTri(tris[prims[nd.dataA + i]], r, t, u, v)) return true;
            } else {
                st[top++] = ni + 1;
                st[top++] = nd.dataA;
            }
        }
        return false;
    }
};

// ---------------------------------------------------------------- curves (accel/CurveBrute.{h,cpp})
struct CNode {                                                                        // CurveBrute.h:27-34
    float bmin[3], bmax[3];
    uint32_t lp, rc;   // interior: left, right child; leaf: first segment, count | 0x80000000
};
struct Seg {
    V3 P[4];
    float hw0, hw3;        // widthAt(v0) * 0.5, widthAt(v3) * 0.5 (CurveBrute.cpp:199-202)
    float tOff, tScale;    // v0, v3 - v0
    uint32_t strand, seg;
};
struct Curves {
    std::vector<CNode> nodes;
    std::vector<Seg> segs;           // leaf order (CurveBrute m_segRefs)
    std::vector<uint32_t> strandMesh;
    std::vector<V3> strandColor;
    void build(const ahip_scene_desc* d);   // oracle_bvh.cpp (CurveBrute.cpp:247-396)
};

}  // namespace orc
