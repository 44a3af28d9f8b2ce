// kernels.h — launch interface between the host context (abi.cpp) and kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/anacapa_hip.h"
#include "device_sampling.h"
#include "device_types.h"

namespace ahip {

// One camera view (a whole film or a tile of it) with its own accumulation target.
struct ViewDesc {
    DCamera cam;
    uint32_t x0, y0, w, h;       // region of the film rendered by this view
    uint32_t tileW, tileH;       // pixel tile that one 64-lane wavefront covers (tileW * tileH = 64)
    uint32_t tilesX;             // tiles across the region
    uint32_t slots;              // paths per sample index: tilesX * tilesY * 64
    uint32_t stride;             // pixels per row of the accumulation arrays (w, or the film width
                                 // when the region is a tile of a frame buffer)
    float4* accum;               // region [h rows of stride]: r, g, b sums and weight
    float* accumLumSq;           // region [w*h]
    float* accumAov;             // region [w*h*6]: first-hit reflectance and normal sums, or null
};

// A pass packs whole sample indices of one or more views into one path range.
constexpr uint32_t kMaxPassSegments = 128;
struct PassSeg {
    uint32_t view, sampleStart, samples;
    uint32_t pathBase;           // first path id of this segment
    uint32_t slotBase;           // first accumulate slot of this segment
    uint32_t pad[3];
};
struct PassTable {
    uint32_t numSegs, totalPaths, totalSlots, pad;
    PassSeg seg[kMaxPassSegments];
};

struct RenderConsts {
    uint64_t seed;
    uint32_t maxDepth, minDepth, flags, pad;
};

// Device queues of the bounce loop (all structure-of-arrays, 16-byte elements).
struct WaveQueues {
    float4* ro[2];    // continuation ray origin | path id
    float4* rd[2];    // direction | sampler dimension (bit 31: specular bounce)
    float4* beta[2];  // throughput
    float4* hit;      // t, u, v | leaf-order triangle (~0 = miss; bit 31 = curve segment, u = strand v)
    float4* hitAux;   // curve hits: fiber tangent
    float4* so;       // shadow ray origin | path id
    float4* sd;       // direction | tMax
    float4* sc;       // contribution if unoccluded
    float4* mo;       // MIS ray origin | path id
    float4* md;       // direction | tMax
    float4* mc;       // weight; times emitted radiance
    float4* L;        // per path radiance
    float4* aovA;     // per path first-hit reflectance (null: no view of the pass records AOVs)
    float4* aovN;     // per path first-hit shading normal
    uint2* pix;       // per path: px | py << 16, sample index
    uint32_t* counters;  // [(maxDepth + 2) * 4]
    uint32_t* ovfRef;    // traversal stack overflow (entries past the LDS part)
    float* ovfT;
    const ViewDesc* views;
    const PassTable* tables;
    uint32_t* passIdx;   // index of the pass being run; advanced by the pass's last kernel
    uint32_t stackCap;   // total stack entries available per thread
    uint32_t cap;        // path capacity
    // curve scenes (null otherwise): the pooled curve queries
    uint32_t* cState;    // [cap] 1 when the ray's curve walk is to run (after the triangle part)
    uint32_t* cwRef;     // curve walk stack entries past the LDS part, [(cwCap - kPoolStackK) * threads]
    float* cwT;
    uint32_t cwCap;      // curve walk stack entries per lane
    uint32_t poolBlocks; // grid of k_curve_pool (one wave per block)
};

// The curve queries' ray interval starts here (CurveBrute.cpp:398-440 tests from the ray's tMin,
// which every queued ray sets to 1e-4).
constexpr float kCurveTMin = 1e-4f;
// k_curve_pool: records waiting per wave (a power of two); the walkers step while fewer than
// kPoolLow wait, so the pool never holds more than kPoolLow + 64 * 7 (a leaf holds at most 7)
constexpr uint32_t kPoolCap = 512;
constexpr uint32_t kPoolLow = 64;
constexpr uint32_t kPoolSlotShift = 26;          // pool entry: slot << 26 | record (records < 2^26)
constexpr uint32_t kPoolIterCap = 1u << 24;      // steps of one wave before it gives up (flagged)
constexpr uint32_t kPoolStackK = 2;              // walk stack entries per lane in LDS
static_assert(kPoolLow + 64u * 7u <= kPoolCap, "the curve pool must hold a full walk step");

// Detached path-replay estimator (adjoint.hip): one thread per (pixel, sample) over the film.
constexpr uint32_t kRecFields = 18;   // floats per differentiable vertex record
struct ReplayArgs {
    DCamera cam;
    uint32_t w, h, s0, ns;
    uint32_t tilesX, slots;      // 8x8 pixel tiles: slots = tilesX * tilesY * 64 paths per sample
    uint64_t seed;
    uint32_t maxDepth, minDepth, flags;
    uint32_t diffTex;            // texture whose texels are the roughness parameters (~0u: none)
    const float* albE;           // evaluation albedo [3 * tris] commit order (null: sampling values)
    const float* texE;           // evaluation texels of diffTex (null: sampling values)
    float4* Lout;                // primal: radiance per (sample, pixel) [ns * w * h]
    const float* dLdI;           // adjoint: [h * w * 3]
    double* gAlb;                // adjoint: [3 * tris]; double so that sums of ~1e3 path terms of
    double* gTex;                // either sign do not lose the result to float32 rounding
    float* rec;                  // adjoint: vertex records [kRecFields * maxV * nthreads]
    uint32_t maxV, nthreads;
    uint32_t* ovfRef;            // traversal stack overflow, stride nthreads
    float* ovfT;
    uint32_t stackCap;
    uint32_t* overflow;          // stack overflow counter
};

void setHaltonTablesReplay(const HaltonTables& t, hipStream_t s);
uint32_t replay_block_size();
void launchReplay(const DScene& sc, const ReplayArgs& a, bool adjoint, uint32_t blocks, hipStream_t s);
void launchReplayReduce(const ReplayArgs& a, float4* accum, float* lsq, hipStream_t s);
void launchZeroF64(double* p, size_t n, hipStream_t s);

// The bidirectional path tracer (bdpt.hip): a chunk of `slots` (pixel, sample) slots of a tile,
// pixel-major (slot = pixel * ns + sample), pixels pix0.. of the tile in row-major order.
struct BdptArgs {
    DCamera cam;
    uint32_t x0, y0, w, h;       // tile on the film
    uint32_t s0, ns;             // sample range of every pixel
    uint32_t pix0, slots;        // first pixel of the chunk, slots of the chunk (a multiple of ns)
    uint64_t seed;
    uint32_t maxDepth, maxV;     // BDPTIntegrator maxDepth; vertices per subpath (maxDepth + 2)
    float clamp;                 // firefly clamp of the s, t >= 2 strategies (0 = off)
    uint32_t cap;                // slot capacity of the stores below
    float4* verts;               // [2 subpaths][maxV][5 fields][cap]
    uint32_t* counts;            // [cap] camera vertices | light vertices << 16
    float4* aovA;                // [cap] first-hit reflectance | 1 when the camera ray hit a surface
    float4* aovN;                // [cap] first-hit normal
    float4* sampleL;             // [cap] the sample's radiance
    float4* terms;               // [R][cap] weighted strategy contributions | visibility (0 / 1)
    uint32_t* nTerms;            // [cap] terms of the slot
    float4* stageO;              // [R][cap] a slot's shadow rays before they are queued
    float4* stageD;
    float4* connO;               // connection queue: shadow ray origin | term index
    float4* connD;               // direction | tMax
    uint32_t* connCount;         // connection queue length
    const float* aliasProb;      // LightSampler alias table (LightSampler.h:28-69)
    const uint32_t* alias;
    const float* lightW;         // light selection pdfs
    uint32_t* ovfRef;            // traversal stack overflow (stride: grid threads)
    float* ovfT;
    uint32_t stackCap;
    uint32_t* overflow;
    float4* outAcc;              // output sums r, g, b (pre-weighted), weight; [rows * outStride]
    float* outLsq;               // luminance-squared sums
    float* outAov;               // null or [rows * outStride * 6]: mean albedo, mean normal per call
    uint32_t outStride, outX0, outY0;
};
void setHaltonTablesBdpt(const HaltonTables& t, hipStream_t s);
// blocks: the slot kernels' grid; traceBlocks: the connection queue's (grid-stride over its length)
void launchBdpt(const DScene& sc, const BdptArgs& a, uint32_t blocks, uint32_t traceBlocks, hipStream_t s);

void setHaltonTables(const HaltonTables& t, hipStream_t s);
uint32_t lds_stack_entries();
uint32_t trace_block_size();
// Blocks of k_curve_pool that stay resident on one CU together (its LDS use bounds them).
uint32_t curve_pool_blocks_per_cu();
void launchFillU32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
void launchMaterialProbe(const DMat& m, bool marschner, uint32_t n, const float* in, float* out, hipStream_t s);
void launchFillF32(float* p, size_t n, float v, hipStream_t s);
void launchHalton(uint64_t seed, const uint32_t* px, const uint32_t* py, const uint32_t* idx, const uint32_t* dim,
                  uint32_t n, float* out, hipStream_t s);
void launchPassEnd(const WaveQueues& q, uint32_t maxDepth, unsigned long long* totals, hipStream_t s);
void launchGenerate(const RenderConsts& rc, const WaveQueues& q, uint32_t blocks, hipStream_t s);
void launchExtend(const DScene& sc, const WaveQueues& q, uint32_t bounce, uint32_t cur, uint32_t blocks,
                  hipStream_t s);
void launchShade(const DScene& sc, const RenderConsts& rc, const WaveQueues& q, uint32_t bounce, uint32_t cur,
                 uint32_t blocks, hipStream_t s);
void launchShadow(const DScene& sc, const WaveQueues& q, uint32_t bounce, uint32_t blocks, hipStream_t s);
void launchMis(const DScene& sc, const WaveQueues& q, uint32_t bounce, uint32_t blocks, hipStream_t s);
void launchAccumulate(const WaveQueues& q, uint32_t blocks, hipStream_t s);
void launchResolve(const float4* acc, const float* lsq, const float* aov, size_t n, float evScale, uint8_t* srgb,
                   float* rgb, float* var, float* alb, float* nrm, hipStream_t s);
void launchTileVariance(const float4* acc, const float* lsq, uint32_t W, uint32_t H, uint32_t tile, float* out,
                        hipStream_t s);
void launchTraceBatch(const DScene& sc, const float* rays, uint32_t n, ahip_hit* hits, uint8_t* occ, bool any,
                      uint32_t* ovfRef, float* ovfT, uint32_t cap, uint32_t blocks, uint32_t* overflow,
                      uint32_t* counts, hipStream_t s);
void launchCount(const DScene& sc, const WaveQueues& q, uint32_t bounce, uint32_t cur, int kind, uint32_t blocks,
                 unsigned long long* sums, hipStream_t s);

}  // namespace ahip
