// device_types.h — HBM layouts shared by the host code and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace ahip {

// A stack / child reference packs a leaf flag, a leaf primitive count (1..8) and an index:
//   bit 31      leaf flag
//   bits 28..30 count-1 (leaves hold at most 4 triangles, BVHBackend.h:67)
//   bits 0..27  node index (interior) or first triangle index (leaf)
constexpr uint32_t kRefLeaf = 0x80000000u;
constexpr uint32_t kRefIndexMask = 0x0fffffffu;
constexpr uint32_t kMaxIndex = kRefIndexMask;
__host__ __device__ inline uint32_t makeLeafRef(uint32_t first, uint32_t count) {
    return kRefLeaf | ((count - 1u) << 28) | first;
}
__host__ __device__ inline bool refIsLeaf(uint32_t r) { return (r & kRefLeaf) != 0u; }
__host__ __device__ inline uint32_t refIndex(uint32_t r) { return r & kRefIndexMask; }
__host__ __device__ inline uint32_t refCount(uint32_t r) { return ((r >> 28) & 7u) + 1u; }

// One 64-byte record per interior node of the reference binary BVH (BVHBackend.h:25-39):
// the boxes of BOTH children, so a visit costs one aligned 64-B fetch and tests two boxes.
struct alignas(64) DNode {
    float lmin[3], lmax[3];   // left child box   (bytes  0..23)
    float rmin[3], rmax[3];   // right child box  (bytes 24..47)
    uint32_t left, right;     // child references (bytes 48..55)
    uint32_t pad[2];
};
static_assert(sizeof(DNode) == 64, "DNode must be 64 bytes");

// Two levels of the binary BVH in one 128-byte record, the node the traversal kernels read.
// For a binary interior node X with children A (left) and B (right), slots 0,1 hold A's two
// children and slots 2,3 hold B's; a child that is a leaf takes its group's first slot and
// leaves the second empty (box +inf..-inf, ref kRefEmpty).  Interior slots refer to the record
// built for that grandchild.  Bounds are stored per axis so one 16-byte load fetches one bound
// of all four slots.  A group's own box is the union of its slots' boxes (each node's box is
// grown over its own primitives), so the traversal recomputes the group's slab entry distance
// from the slots' per-axis values and visits the slots in exactly the binary tree's order.
struct alignas(128) DNode4 {
    float bmin[3][4];         // [axis][slot]  (bytes   0..47)
    float bmax[3][4];         //               (bytes  48..95)
    uint32_t ref[4];          //               (bytes  96..111)
    uint32_t pad[4];
};
static_assert(sizeof(DNode4) == 128, "DNode4 must be 128 bytes");
constexpr uint32_t kRefEmpty = 0xffffffffu;

// Leaf-ordered triangle record used by intersection (BVHTriangle v0/e1/e2/n, BVHBackend.h:49-59).
struct alignas(64) DTri {
    float v0[3];
    float e1[3];
    float e2[3];
    float n[3];               // geometric normal (normalised cross(e1, e2))
    uint32_t meshID, primID;  // SurfaceInteraction ids
    uint32_t commitIndex;     // index in commit order (reference m_tris)
    int32_t matIndex;         // scene.materials[meshID] resolved at prepare time, -1 = none
};
static_assert(sizeof(DTri) == 64, "DTri must be 64 bytes");

// Shading attributes, fetched only for the closest hit.
struct alignas(64) DTriShade {
    float sn0[3], sn1[3], sn2[3];
    float uv0[2], uv1[2], uv2[2];
    uint32_t pad;
};
static_assert(sizeof(DTriShade) == 64, "DTriShade must be 64 bytes");

// Curve BVH node in the reference's own 32-byte layout (CurveNode, CurveBrute.h:27-34): the
// curve traversal tests a child's box when it is popped, against the current best distance.
struct CNode {
    float bmin[3], bmax[3];
    uint32_t lp;   // interior: left child; leaf: first segment record
    uint32_t rc;   // interior: right child; leaf: count | 0x80000000
};
static_assert(sizeof(CNode) == 32, "CNode must be 32 bytes");

// 64-byte cubic Bezier segment record in leaf order: 4 control points, the half widths at the
// ends and the strand-parameter range [tOff, tOff + tScale] (CurveBrute.cpp:182-202).
struct alignas(64) DSeg {
    float P[12];
    float hw0, hw3, tOff, tScale;
};
static_assert(sizeof(DSeg) == 64, "DSeg must be 64 bytes");

constexpr uint32_t kCurveHit = 0x80000000u;   // hit code flag: low bits = segment record

// Both children of an interior curve-tree node in one 64-byte record, for a walk that orders the
// children before it descends: each child's box (bmin, bmax) and reference, the child's node index
// if it is interior, else kCWLeaf | (record count - 1) << 29 | first record.
struct alignas(64) CWide {
    float a[6], b[6];
    uint32_t refA, refB, pad0, pad1;
};
static_assert(sizeof(CWide) == 64, "CWide must be 64 bytes");
constexpr uint32_t kCWLeaf = 0x80000000u;
constexpr uint32_t kMaxCurveDepth = 6;        // ribbon subdivision levels (CurveBrute.cpp:194 fixes 6)

struct DMat {
    uint32_t type;            // AHIP_MAT_*
    uint32_t isDelta;         // StandardSurface::isDelta (StandardSurface.h:162-169)
    uint32_t specularFlag;    // flags() & Specular (roughness < 0.001)
    int32_t roughTex;         // -1 or texture slot
    float albedo[3];          // Lambert albedo / StandardSurface base_color
    float Le[3];              // emitted radiance (already * emission weight)
    float base, metal, spec, rough;
    float specColor[3];
    float coat, coatA2, f0d, coatF0;
    uint32_t hasLe;
    float hairSigma[3];       // MarschnerHairMaterial default sigma_a
    float hairEta;            // Marschner R/TT/TRT: index of the cortex,
    float hairV[3];           //   longitudinal variance of R, TT, TRT,
    float hairS;              //   azimuthal logistic scale,
    float hairAlpha[3];       //   cuticle tilts (radians) (MarschnerHair.h:200-216)
    // Transparency (StandardSurface.h:162-169, 197-225, 248-347, 410-474), flags set on the host:
    uint32_t glass;           // transmission > 0.001 && metalness < 0.001: the dielectric branch of sample
                              // and the Fresnel-weighted reflection of evaluate in one hemisphere
    uint32_t trans;           // !(transmission < 0.001) && !(metalness > 0.001): evaluate's transmission
                              // lobe across the surface and a non-black transmittanceColor
    float eta;                // specular_IOR
    float trScale;            // transmission * (1 - fresnelDielectric(1, IOR)): the shadow tint per unit base colour
    uint32_t alphaMask;       // Params::alphaMask: evalOpacity = opacity, shadow tint 1 - opacity
    float opacity;            // constant opacity (opacityTex < 0)
    int32_t opacityTex;       // -1 or texture slot
};

struct DLight {
    uint32_t type;            // AHIP_LIGHT_*
    float center[3], uA[3], vA[3], Le[3], normal[3];
    float area, uLen, vLen;
    float dir[3];             // directional: unit dirToLight
    float cosCone;            // directional: cos of the angular radius (1 = hard)
    float sceneC[3];          // dome, directional: centre and radius of the scene sphere where
    float sceneR;             //   BDPT's light subpaths start (sampleLe)
};

struct DTexture {
    const float* px;
    uint32_t w, h;
};

struct DDome {
    const float* px;          // W*H texels of 4 floats (rgb + pad), 16-byte aligned
    const float* margFunc;    // [H]
    const float* margCdf;     // [H+1]
    const float* condFunc;    // [H*W]
    const float* condCdf;     // [H*(W+1)]
    const float* condInt;     // [H]
    const uint32_t* margGuide;  // [margM+1]: first marginal CDF index >= k / margM
    const uint32_t* condGuide;  // [H*(condM+1)]: the same per row of the conditional CDF
    uint32_t margM, condM;      // powers of two
    float margInt;
    uint32_t W, H;
    float intensity;
    float rot[9];
};

// Halton bases and their reciprocals (1.f / (float)prime, the reference's invB), in constant memory.
// Per-dimension base, its float reciprocal and the constants of exact division by the base
// (Granlund & Montgomery, "Division by invariant integers using multiplication", PLDI 1994): for
// 94 of the 128 bases a 32-bit multiplier exists and n / b == mulhi(magic, n) >> sh (shifts =
// sh | 1 << 16, Theorem 4.2); for the others n / b == (t + ((n - t) >> sh1)) >> sh2 with
// t = mulhi(magic, n) (figure 4.1; shifts = sh1 | sh2 << 8), for every 32-bit n.
struct HaltonTables {
    uint32_t prime[128];
    float invPrime[128];
    uint32_t magic[128];
    uint32_t shifts[128];
};

struct DCamera {
    float origin[3], llc[3], hor[3], ver[3];
    uint32_t w, h;
};

// Everything a kernel needs to see of the scene (passed by value as a kernel argument).
struct DScene {
    const DNode4* nodes4;     // two-level records, root at index 0 when rootRef is interior
    const DTri* tris;
    const DTriShade* shade;
    const DMat* mats;
    const int32_t* meshMat;
    const DLight* lights;
    const DTexture* textures;
    uint32_t numMeshes, numLights, numNodes, numTris;
    float rootMin[3], rootMax[3];
    uint32_t rootRef;         // reference of the root (a leaf ref when the whole BVH is one leaf)
    uint32_t hasBVH;
    uint32_t hasEnv;          // dome present
    uint32_t hasEmitters;     // any material with non-zero Le
    float envRadiance[3];
    DDome dome;
    // curves
    const CNode* cnodes;
    const CWide* cwide;       // [nodes] child pairs of the interior nodes (null: not built)
    const DSeg* segs;
    const uint32_t* segIds;   // [2 * segments]: strand, segment index within the strand
    const float* strandInfo;  // [4 * strands]: colour rgb, meshID (bits)
    uint32_t hasCurves;
    const float* triAlbedo;   // [3 * tris] per-triangle base colour in commit order, or null
    uint32_t hasHair;         // any hair material
    uint32_t curveDepth;      // de Casteljau depth of the ribbon test (the reference's MAX_DEPTH is 6)
    uint32_t hasTransparent;  // any glass or alpha-masked material: shadow and BSDF-sampled NEE rays step
                              // through transmissive surfaces (ShadowRay.h:26-61, PathIntegrator.cpp:217-236)
    uint32_t hasAlpha;        // any alpha-masked material: the bounce loop's pass-through test runs
    uint32_t hasDirectional;  // any directional (delta) light
};

}  // namespace ahip
