// abi.cpp — the extern "C" boundary (include/anacapa_hip.h): context, scene upload and the
// host side of the wavefront bounce loop.
//
// The context plays CudaPathIntegrator's role (src/gpu/cuda/CudaPathIntegrator.cu:165-489):
// it owns one HIP stream and every device allocation, copies the scene into HBM in prepare,
// keeps a persistent accumulation buffer that successive render_frame calls add into, and
// reports failures as status codes plus a message instead of aborting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/anacapa_hip.h"
#include "bvh_build.h"
#include "device_types.h"
#include "host_scene.h"
#include "kernels.h"

using namespace ahip;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        release();
        if (n == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        else p = nullptr;
        return e;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    int kind = 0;     // 0 trace, 1 shade
    int bounce = -1;  // the bounce a traversal / shade launch belongs to (-1: generate, accumulate)
};

// ------------------------------------------------------------ no-allocation traversal scratch
// ahip_trace_rays_noalloc runs on this fixed device array instead of allocating: a 256-byte header
// (word 0: the stack-overflow flag), then the traversal stack entries past the LDS part as
// [AHIP_NOALLOC_MAX_STACK][kNoAllocThreads] child references followed by as many entry distances
// (the layout TStack indexes with the grid's thread count as stride).  Every GPU holds its own copy
// until the process ends.  In host code `scratch` names a host-side stand-in, not device memory:
// host code only ever uses the per-device address hipGetSymbolAddress returns for it.
constexpr uint32_t kNoAllocThreads = 262144;   // 2048 blocks of 128 threads, grid-stride over the rays
constexpr size_t kNoAllocHeader = 256;
constexpr size_t kNoAllocStackWords = (size_t)AHIP_NOALLOC_MAX_STACK * kNoAllocThreads;
constexpr size_t kNoAllocBytes = kNoAllocHeader + 2u * kNoAllocStackWords * sizeof(uint32_t);
static_assert(kNoAllocBytes == (64u << 20) + 256u, "no-allocation scratch: 64 MiB of stack + header");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratchMu;
std::vector<void*> g_scratchAddr;   // per device ordinal, resolved once

// The device address of `scratch` on the current device (ordinal dev), resolved on first use.
hipError_t scratchAddress(int dev, void** out) {
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if ((int)g_scratchAddr.size() <= dev) g_scratchAddr.resize(dev + 1, nullptr);
    if (!g_scratchAddr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratchAddr[dev] = p;
    }
    *out = g_scratchAddr[dev];
    return hipSuccess;
}

}  // namespace

struct ahip_context {
    std::mutex mu;
    int device = -1;
    bool valid = false;
    bool failNext = false;
    std::string name;
    std::string err;
    hipStream_t stream = nullptr;
    uint32_t numCU = 0;

    // scene
    bool prepared = false;
    BVHOutput bvh;
    DevBuf dNodes, dTris, dShade, dMats, dMeshMat, dLights, dTexDesc, dDomePx, dDomeMF, dDomeMC, dDomeCF, dDomeCC, dDomeMG, dDomeCG,
        dDomeCI;
    std::vector<DevBuf*> texBufs;
    CurveOutput curves;
    DevBuf dCNodes, dCWide, dSegs, dSegIds, dStrandInfo;
    DevBuf dTriAlb;                      // sampling-parameter per-triangle albedo (commit order)
    uint32_t numCommit = 0;              // triangles in commit order
    std::vector<uint32_t> texW, texH;
    DScene dsc{};

    // detached replay / adjoint
    DevBuf dAlbE, dTexE, dLout, dRepAcc, dRepLsq, dDLdI, dGAlb, dGTex, dRec, dRepCtr;
    bool albESet = false, texESet = false;
    uint32_t diffTex = 0;

    // render settings
    ahip_camera cam{};
    bool hasCam = false;
    ahip_render_params params{8, 2, 0, 0, 0, AHIP_INTEGRATOR_PATH, 10.f};
    uint32_t maxPaths = 1u << 25;

    // persistent accumulation
    DevBuf dAccum, dAccumLsq, dAccumAov;   // persistent frame sums (AOVs when aovOn)
    DevBuf dResolve[5];                    // ahip_resolve outputs, kept for the next call (preview loops)
    uint32_t accW = 0, accH = 0;

    // wavefront queues
    DevBuf qRo[2], qRd[2], qBeta[2], qHit, qHitAux, qSo, qSd, qSc, qMo, qMd, qMc, qL, qCtr, qOvfRef, qOvfT, qTotals;
    DevBuf qCState, qCwRef, qCwT;      // curve scenes: the pooled curve queries
    uint32_t qCurveCap = 0, qCurveStack = 0;
    uint32_t poolBlocks = 0;           // k_curve_pool grid (one wave per block)
    // BDPT (bdpt.hip): why the prepared scene is refused (empty: accepted), the light sampler's
    // alias table and the vertex stores of a chunk of slots
    std::string bdptRefusal = "no scene";
    DevBuf dAliasProb, dAlias, dLightW, dBdptVerts, dBdptCounts, dBdptAovA, dBdptAovN, dBdptL, dBdptCtr;
    DevBuf dBdptTerms, dBdptNTerms, dBdptConnO, dBdptConnD, dBdptStageO, dBdptStageD;
    size_t bdptCap = 0, bdptMaxV = 0;
    DevBuf qPix, qPassIdx, dViews, dTables, dBatchAcc, dBatchLsq;
    uint32_t tileW = 8;          // pixel tile per wavefront: tileW x (64 / tileW)
    bool frameGraph = false;     // render_frame: per-launch timing events (the roofline figure needs them)
    bool batchGraph = true;      // render_batch: one captured pass replayed per pass
    uint32_t qCap = 0, qCtrSlots = 0;
    uint32_t traceBlocks = 0, shadeBlocks = 0, stackCap = 0;

    // statistics
    ahip_stats stats{};
    std::vector<unsigned long long> bounceRays;   // per bounce: continuation, shadow, MIS rays
    std::vector<double> bounceMs;                 // per bounce: timed traversal + shade launches
    std::vector<EventPair> events;
    size_t eventsUsed = 0;
    hipEvent_t renderBeg = nullptr, renderEnd = nullptr;
    hipGraphExec_t passGraph = nullptr;          // the last captured pass (render_batch)
    std::vector<unsigned char> passGraphKey;     // the kernel arguments it was captured with
    void dropPassGraph() {
        if (passGraph) hipGraphExecDestroy(passGraph);
        passGraph = nullptr;
        passGraphKey.clear();
    }
    void* scratch = nullptr;     // this device's no-allocation scratch (ahip_trace_rays_noalloc)
    std::string scratchErr;      // why it could not be resolved at create
    bool lastRenderValid = false;
    bool aovOn = false;          // ahip_set_aovs
    bool accumHasSamples = false; // the frame accumulation holds samples since the last clear
    bool aovPartial = false;     // AOV sums started after samples were already accumulated
    DevBuf qAovA, qAovN;         // per-path first-hit AOVs (allocated when a render records them)
    uint32_t qAovCap = 0;
    uint32_t lastW = 0, lastH = 0, lastS0 = 0, lastNs = 0;

    ~ahip_context() {
        dropPassGraph();
        for (auto* b : texBufs) delete b;
        for (auto& e : events) {
            if (e.a) hipEventDestroy(e.a);
            if (e.b) hipEventDestroy(e.b);
        }
        if (renderBeg) hipEventDestroy(renderBeg);
        if (renderEnd) hipEventDestroy(renderEnd);
        if (stream) hipStreamDestroy(stream);
    }
};

namespace {

int fail(ahip_context* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(ctx, _e == hipErrorOutOfMemory ? AHIP_E_OUT_OF_MEMORY : AHIP_E_HIP, \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                 \
    } while (0)

template <class T>
int upload(ahip_context* ctx, DevBuf& b, const T* src, size_t n) {
    HIPCHK(b.alloc(std::max<size_t>(n, 1) * sizeof(T)));
    if (n) HIPCHK(hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return AHIP_OK;
}

int checkInjected(ahip_context* ctx) {
    if (ctx->failNext) {
        ctx->failNext = false;
        return fail(ctx, AHIP_E_INJECTED, "injected launch failure (ahip_fail_next_launch)");
    }
    return AHIP_OK;
}

EventPair& nextEvent(ahip_context* ctx, int kind, int bounce) {
    if (ctx->eventsUsed == ctx->events.size()) {
        EventPair e;
        hipEventCreate(&e.a);
        hipEventCreate(&e.b);
        ctx->events.push_back(e);
    }
    EventPair& e = ctx->events[ctx->eventsUsed++];
    e.kind = kind;
    e.bounce = bounce;
    return e;
}

int ensureQueues(ahip_context* ctx, uint32_t cap) {
    uint32_t slots = (ctx->params.max_depth + 2u) * 4u;
    if (!ctx->curves.nodes.empty()) {
        // a walk that pushes a node's children after testing them holds at most one pending child
        // per level below the root, plus the two just pushed
        const uint32_t pc = std::max(cap, ctx->qCap), sc = ctx->curves.maxDepth + 2u;
        if (ctx->qCurveCap != pc || ctx->qCurveStack != sc || !ctx->qCState.p) {
            HIPCHK(ctx->qCState.alloc((size_t)pc * sizeof(uint32_t)));
            const size_t deep = (size_t)(sc > kPoolStackK ? sc - kPoolStackK : 1u) * ctx->poolBlocks * 64u;
            HIPCHK(ctx->qCwRef.alloc(deep * sizeof(uint32_t)));
            HIPCHK(ctx->qCwT.alloc(deep * sizeof(float)));
            ctx->qCurveCap = pc;
            ctx->qCurveStack = sc;
        }
    }
    if (cap <= ctx->qCap && slots <= ctx->qCtrSlots && ctx->qOvfRef.p) return AHIP_OK;
    cap = std::max(cap, ctx->qCap);
    size_t f4 = (size_t)cap * sizeof(float4);
    for (int k = 0; k < 2; ++k) {
        HIPCHK(ctx->qRo[k].alloc(f4));
        HIPCHK(ctx->qRd[k].alloc(f4));
        HIPCHK(ctx->qBeta[k].alloc(f4));
    }
    HIPCHK(ctx->qHit.alloc(f4));
    HIPCHK(ctx->qHitAux.alloc(f4));
    HIPCHK(ctx->qSo.alloc(f4));
    HIPCHK(ctx->qSd.alloc(f4));
    HIPCHK(ctx->qSc.alloc(f4));
    HIPCHK(ctx->qMo.alloc(f4));
    HIPCHK(ctx->qMd.alloc(f4));
    HIPCHK(ctx->qMc.alloc(f4));
    HIPCHK(ctx->qL.alloc(f4));
    HIPCHK(ctx->qPix.alloc((size_t)cap * sizeof(uint2)));
    HIPCHK(ctx->qPassIdx.alloc(sizeof(uint32_t)));
    HIPCHK(ctx->qCtr.alloc(slots * sizeof(uint32_t)));
    HIPCHK(ctx->qTotals.alloc((8 + (size_t)slots) * sizeof(unsigned long long)));
    ctx->qCap = cap;
    ctx->qCtrSlots = slots;
    return AHIP_OK;
}

int ensureStack(ahip_context* ctx) {
    uint32_t K = lds_stack_entries();
    // a DNode4 visit pushes at most 3 slots and every stacked slot belongs to a record on the
    // current root-to-node path; the curve traversal pushes at most one entry per level
    uint32_t capEntries = std::max(3u * ctx->bvh.depth4, ctx->curves.maxDepth) + 2u;
    ctx->stackCap = capEntries;
    size_t extra = capEntries > K ? capEntries - K : 1u;
    size_t threads = (size_t)ctx->traceBlocks * 128u;
    HIPCHK(ctx->qOvfRef.alloc(extra * threads * sizeof(uint32_t)));
    HIPCHK(ctx->qOvfT.alloc(extra * threads * sizeof(float)));
    return AHIP_OK;
}

WaveQueues queues(ahip_context* ctx) {
    WaveQueues q{};
    for (int k = 0; k < 2; ++k) {
        q.ro[k] = ctx->qRo[k].as<float4>();
        q.rd[k] = ctx->qRd[k].as<float4>();
        q.beta[k] = ctx->qBeta[k].as<float4>();
    }
    q.hit = ctx->qHit.as<float4>();
    q.hitAux = ctx->qHitAux.as<float4>();
    q.so = ctx->qSo.as<float4>();
    q.sd = ctx->qSd.as<float4>();
    q.sc = ctx->qSc.as<float4>();
    q.mo = ctx->qMo.as<float4>();
    q.md = ctx->qMd.as<float4>();
    q.mc = ctx->qMc.as<float4>();
    q.L = ctx->qL.as<float4>();
    q.pix = ctx->qPix.as<uint2>();
    q.counters = ctx->qCtr.as<uint32_t>();
    q.ovfRef = ctx->qOvfRef.as<uint32_t>();
    q.ovfT = ctx->qOvfT.as<float>();
    q.stackCap = ctx->stackCap;
    q.cap = ctx->qCap;
    // the pool entry holds a record in kPoolSlotShift bits; larger curve sets use the whole-ray kernels
    if (!ctx->curves.nodes.empty() && ctx->qCurveCap == ctx->qCap && ctx->dsc.cwide &&
        ctx->curves.segs.size() < (size_t(1) << kPoolSlotShift)) {
        q.cState = ctx->qCState.as<uint32_t>();
        q.cwRef = ctx->qCwRef.as<uint32_t>();
        q.cwT = ctx->qCwT.as<float>();
        q.cwCap = ctx->qCurveStack;
        q.poolBlocks = ctx->poolBlocks;
    }
    return q;
}

DCamera toDCam(const ahip_camera& c) {
    DCamera d;
    for (int i = 0; i < 3; ++i) {
        d.origin[i] = c.origin[i];
        d.llc[i] = c.lower_left[i];
        d.hor[i] = c.horizontal[i];
        d.ver[i] = c.vertical[i];
    }
    d.w = c.width;
    d.h = c.height;
    return d;
}

// One render request: a camera, a region of its film, a sample range and where to add the sums.
struct ViewJob {
    ahip_camera cam;
    uint32_t x0, y0, w, h;
    uint32_t s0, ns;
    float4* acc;
    float* lsq;
    float* aov = nullptr;        // [w*h*6] AOV sums, or null
    uint32_t stride = 0;         // pixels per row of acc / lsq / aov; 0 = w
};

// Pack whole sample indices of the jobs, in order, into passes of at most `cap` paths (a job whose
// single sample needs more paths than cap raises cap).  Every pass is a list of segments.
int planPasses(ahip_context* ctx, const std::vector<ViewJob>& jobs, std::vector<ViewDesc>& views,
               std::vector<PassTable>& tables, uint32_t& cap, uint32_t& maxSlots) {
    const uint32_t tw = ctx->tileW, th = 64u / ctx->tileW;
    views.clear();
    tables.clear();
    uint64_t need = 0, widest = 64;
    for (const ViewJob& j : jobs) {
        ViewDesc v{};
        v.cam = toDCam(j.cam);
        v.x0 = j.x0; v.y0 = j.y0; v.w = j.w; v.h = j.h;
        v.tileW = tw;
        v.tileH = th;
        v.tilesX = (j.w + tw - 1u) / tw;
        uint64_t slots = (uint64_t)v.tilesX * ((j.h + th - 1u) / th) * 64u;
        if (slots > 0x7fffffffull || j.x0 + j.w > 65536u || j.y0 + j.h > 65536u)
            return fail(ctx, AHIP_E_INVALID_ARG, "view too large");
        v.slots = (uint32_t)slots;
        v.stride = j.stride ? j.stride : j.w;
        v.accum = j.acc;
        v.accumLumSq = j.lsq;
        v.accumAov = j.aov;
        widest = std::max<uint64_t>(widest, slots);
        need += slots * j.ns;
        views.push_back(v);
    }
    // no more paths per pass than the jobs need (queue memory scales with the capacity), never
    // fewer than one sample of the widest view
    uint64_t cap64 = std::max<uint64_t>(widest, std::min<uint64_t>(std::max(64u, ctx->maxPaths), need));
    if (cap64 > 0x7fffffffull) return fail(ctx, AHIP_E_INVALID_ARG, "pass capacity too large");
    cap = (uint32_t)cap64;
    maxSlots = 0;
    PassTable cur{};
    auto close = [&]() {
        if (cur.numSegs) {
            maxSlots = std::max(maxSlots, cur.totalSlots);
            tables.push_back(cur);
        }
        cur = PassTable{};
    };
    for (uint32_t vi = 0; vi < (uint32_t)jobs.size(); ++vi) {
        const uint32_t slots = views[vi].slots;
        uint32_t start = jobs[vi].s0, left = jobs[vi].ns;
        while (left > 0u) {
            uint32_t fit = (cap - cur.totalPaths) / slots;
            if (fit == 0u || cur.numSegs == kMaxPassSegments) {
                close();
                continue;
            }
            uint32_t k = std::min(fit, left);
            PassSeg& sg = cur.seg[cur.numSegs++];
            sg = PassSeg{vi, start, k, cur.totalPaths, cur.totalSlots, {0, 0, 0}};
            cur.totalPaths += k * slots;
            cur.totalSlots += slots;
            start += k;
            left -= k;
        }
    }
    close();
    return AHIP_OK;
}

// Launches of one pass.  The kernels find their pass through the device-side pass index, so the
// same sequence (or a graph captured from it) serves every pass.
void enqueuePass(ahip_context* ctx, const WaveQueues& q, const RenderConsts& rc, uint32_t cap, uint32_t maxSlots,
                 bool events, unsigned long long* countSums) {
    hipStream_t s = ctx->stream;
    const uint32_t maxDepth = rc.maxDepth;
    auto timed = [&](int kind, int bounce, auto&& fn) {
        if (!events) {
            fn();
            return;
        }
        EventPair& e = nextEvent(ctx, kind, bounce);
        hipEventRecord(e.a, s);
        fn();
        hipEventRecord(e.b, s);
    };
    uint32_t gb = std::max(1u, std::min((cap + 255u) / 256u, ctx->shadeBlocks));
    timed(1, -1, [&] { launchGenerate(rc, q, gb, s); });
    uint32_t cur = 0;
    for (uint32_t b = 0; b <= maxDepth; ++b) {
        if (countSums) launchCount(ctx->dsc, q, b, cur, 0, ctx->traceBlocks, countSums, s);
        timed(0, (int)b, [&] { launchExtend(ctx->dsc, q, b, cur, ctx->traceBlocks, s); });
        timed(1, (int)b, [&] { launchShade(ctx->dsc, rc, q, b, cur, ctx->shadeBlocks, s); });
        if (countSums) {
            launchCount(ctx->dsc, q, b, cur, 1, ctx->traceBlocks, countSums, s);
            launchCount(ctx->dsc, q, b, cur, 2, ctx->traceBlocks, countSums, s);
        }
        timed(0, (int)b, [&] { launchShadow(ctx->dsc, q, b, ctx->traceBlocks, s); });
        timed(0, (int)b, [&] { launchMis(ctx->dsc, q, b, ctx->traceBlocks, s); });
        cur ^= 1u;
    }
    uint32_t ab = std::max(1u, std::min((maxSlots + 255u) / 256u, 8192u));
    timed(1, -1, [&] { launchAccumulate(q, ab, s); });
    launchPassEnd(q, maxDepth, ctx->qTotals.as<unsigned long long>(), s);
}

// Render the jobs.  graph = true captures one pass into a HIP graph and replays it per pass
// (no host work between passes); otherwise every launch is issued with timing events.
int renderJobs(ahip_context* ctx, const std::vector<ViewJob>& jobs, bool graph, unsigned long long* countSums) {
    std::vector<ViewDesc> views;
    std::vector<PassTable> tables;
    uint32_t cap = 0, maxSlots = 0;
    int st = planPasses(ctx, jobs, views, tables, cap, maxSlots);
    if (st) return st;
    if (tables.empty()) return AHIP_OK;
    if ((st = ensureQueues(ctx, cap))) return st;
    HIPCHK(ctx->dViews.alloc(views.size() * sizeof(ViewDesc)));
    HIPCHK(ctx->dTables.alloc(tables.size() * sizeof(PassTable)));
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(ctx->dViews.p, views.data(), views.size() * sizeof(ViewDesc), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(ctx->dTables.p, tables.data(), tables.size() * sizeof(PassTable), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));   // the host vectors go out of scope after this call
    WaveQueues q = queues(ctx);
    bool aov = false;
    for (const ViewJob& j : jobs) aov = aov || j.aov != nullptr;
    if (aov) {
        if (ctx->qAovCap < cap || !ctx->qAovA.p) {
            HIPCHK(ctx->qAovA.alloc((size_t)cap * sizeof(float4)));
            HIPCHK(ctx->qAovN.alloc((size_t)cap * sizeof(float4)));
            ctx->qAovCap = cap;
        }
        q.aovA = ctx->qAovA.as<float4>();
        q.aovN = ctx->qAovN.as<float4>();
    }
    q.views = ctx->dViews.as<ViewDesc>();
    q.tables = ctx->dTables.as<PassTable>();
    q.passIdx = ctx->qPassIdx.as<uint32_t>();
    RenderConsts rc{ctx->params.seed, ctx->params.max_depth, ctx->params.min_depth, ctx->params.flags, 0};
    launchFillU32(q.counters, ctx->qCtrSlots, 0u, s);
    launchFillU32(q.passIdx, 1, 0u, s);
    if (graph && !countSums) {
        // One pass captured into a graph and replayed per pass.  The graph bakes in the kernel
        // arguments (queue and scene pointers, render constants, pass capacity), so an executable
        // graph whose arguments are byte-identical is reused across calls instead of re-captured.
        std::vector<unsigned char> key;
        auto put = [&](const void* p, size_t n) {
            const unsigned char* b = static_cast<const unsigned char*>(p);
            key.insert(key.end(), b, b + n);
        };
        put(&q, sizeof(q));
        put(&rc, sizeof(rc));
        put(&ctx->dsc, sizeof(ctx->dsc));
        const uint32_t dims[4] = {cap, maxSlots, ctx->traceBlocks, ctx->shadeBlocks};
        put(dims, sizeof(dims));
        if (!ctx->passGraph || key != ctx->passGraphKey) {
            ctx->dropPassGraph();
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            enqueuePass(ctx, q, rc, cap, maxSlots, false, nullptr);
            hipError_t e = hipStreamEndCapture(s, &g);
            if (e != hipSuccess) return fail(ctx, AHIP_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            hipGraphExec_t ge = nullptr;
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) return fail(ctx, AHIP_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            ctx->passGraph = ge;
            ctx->passGraphKey = std::move(key);
            ctx->stats.graph_captures += 1;
        }
        hipError_t e = hipSuccess;
        for (size_t p = 0; p < tables.size() && e == hipSuccess; ++p) e = hipGraphLaunch(ctx->passGraph, s);
        hipError_t e2 = hipStreamSynchronize(s);
        if (e != hipSuccess || e2 != hipSuccess) {
            ctx->dropPassGraph();
            return fail(ctx, AHIP_E_HIP, std::string("graph launch: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        }
        ctx->stats.graph_passes += tables.size();
    } else {
        for (size_t p = 0; p < tables.size(); ++p) enqueuePass(ctx, q, rc, cap, maxSlots, true, countSums);
    }
    for (const ViewJob& j : jobs) ctx->stats.samples += (uint64_t)j.w * j.h * j.ns;
    ctx->stats.passes += tables.size();
    HIPCHK(hipGetLastError());
    return AHIP_OK;
}

// Collect per-launch event timings and ray totals after the stream has been synchronised.
int finishStats(ahip_context* ctx) {
    for (size_t i = 0; i < ctx->eventsUsed; ++i) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->events[i].a, ctx->events[i].b));
        if (ctx->events[i].kind == 0) {
            ctx->stats.trace_ms += ms;
            ctx->stats.trace_launches += 1;
        } else {
            ctx->stats.shade_ms += ms;
            ctx->stats.shade_launches += 1;
        }
        const int b = ctx->events[i].bounce;
        if (b >= 0) {
            if (ctx->bounceMs.size() <= (size_t)b) ctx->bounceMs.resize(b + 1, 0.0);
            ctx->bounceMs[b] += ms;
        }
    }
    ctx->eventsUsed = 0;
    const uint32_t nb = ctx->params.max_depth + 1u;
    std::vector<unsigned long long> all(8 + 4 * (size_t)nb);
    HIPCHK(hipMemcpy(all.data(), ctx->qTotals.p, all.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    if (ctx->bounceRays.size() < 3 * (size_t)nb) ctx->bounceRays.resize(3 * (size_t)nb, 0ull);
    for (uint32_t b = 0; b < nb; ++b)
        for (uint32_t k = 0; k < 3; ++k) ctx->bounceRays[3 * b + k] += all[8 + 4 * b + k];
    const unsigned long long* tot = all.data();
    // tot[0] continuation rays, tot[1] shadow rays, tot[2] MIS rays, tot[3] stack overflows
    ctx->stats.rays_closest += tot[0] + (ctx->dsc.hasEmitters ? tot[2] : 0ull);
    ctx->stats.rays_shadow += tot[1] + (ctx->dsc.hasEmitters ? 0ull : tot[2]);
    if (tot[3]) return fail(ctx, AHIP_E_HIP, "traversal stack capacity exceeded");
    return AHIP_OK;
}

}  // namespace

extern "C" {

int ahip_abi_version(void) { return AHIP_ABI_VERSION; }
const char* ahip_backend_type(void) { return "HIP"; }

int ahip_create(int device_ordinal, ahip_context** out) {
    if (!out) return AHIP_E_INVALID_ARG;
    auto* ctx = new ahip_context();
    *out = ctx;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(ctx, AHIP_E_NO_DEVICE, "no HIP device available");
    if (device_ordinal < 0 || device_ordinal >= n) return fail(ctx, AHIP_E_INVALID_ARG, "device ordinal out of range");
    ctx->device = device_ordinal;
    HIPCHK(hipSetDevice(device_ordinal));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device_ordinal));
    ctx->numCU = (uint32_t)prop.multiProcessorCount;
    // the marketing name can be empty (it comes from a lookup table the runtime may not find)
    ctx->name = std::string(prop.name[0] ? prop.name : "AMD GPU") + " (" + prop.gcnArchName + ", " +
                std::to_string(ctx->numCU) + " CUs)";
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&ctx->renderBeg));
    HIPCHK(hipEventCreate(&ctx->renderEnd));
    // Grid-stride grids oversubscribed well past the resident blocks (8 waves per SIMD = 16 blocks
    // of 128 threads per CU): rays differ widely in cost, and blocks queued behind the resident
    // ones keep the CUs evenly loaded to the end of a launch (measured on the 1080p frame:
    // traversal 18.6 ms at 16 blocks per CU, 16.1 ms at 128; shade 5.9 ms at 8, 5.45 ms at 32).
    ctx->traceBlocks = std::max(1u, ctx->numCU) * 128u;
    // the pooled curve queries: exactly the one-wave blocks that stay resident together (each
    // loops over its share of the rays, so a block queued behind them would run alone at the end;
    // measured on the hair frame at 8 spp: 773 ms at the 6 resident blocks per CU, 1181 ms at 7)
    ctx->poolBlocks = std::max(1u, ctx->numCU) * curve_pool_blocks_per_cu();
    ctx->shadeBlocks = std::max(1u, ctx->numCU) * 32u;
    HaltonTables T;
    fillHaltonTables(T);
    setHaltonTables(T, ctx->stream);
    setHaltonTablesReplay(T, ctx->stream);
    setHaltonTablesBdpt(T, ctx->stream);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // the no-allocation scratch of this device (the device is current); a failure here only
    // disables ahip_trace_rays_noalloc, which reports it
    hipError_t se = scratchAddress(device_ordinal, &ctx->scratch);
    if (se != hipSuccess) {
        ctx->scratch = nullptr;
        ctx->scratchErr = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
    }
    ctx->valid = true;
    return AHIP_OK;
}

void ahip_destroy(ahip_context* ctx) {
    if (!ctx) return;
    if (ctx->valid) {
        hipSetDevice(ctx->device);
        hipStreamSynchronize(ctx->stream);
    }
    delete ctx;
}

int ahip_is_valid(const ahip_context* ctx) { return ctx && ctx->valid ? 1 : 0; }
const char* ahip_device_name(const ahip_context* ctx) { return ctx ? ctx->name.c_str() : ""; }
const char* ahip_last_error(const ahip_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

#define ENTER()                                                                   \
    if (!ctx) return AHIP_E_INVALID_ARG;                                          \
    std::lock_guard<std::mutex> _lk(ctx->mu);                                     \
    if (!ctx->valid) return fail(ctx, AHIP_E_NO_DEVICE, "context is not valid"); \
    HIPCHK(hipSetDevice(ctx->device))

int ahip_prepare(ahip_context* ctx, const ahip_scene_desc* d) {
    ENTER();
    if (!d) return fail(ctx, AHIP_E_INVALID_ARG, "null scene");
    ctx->prepared = false;
    HostScene hs;
    std::string err;
    int st = buildHostScene(d, hs, err);
    if (st) return fail(ctx, st, err);
    if (buildBVH(d, ctx->bvh, err)) return fail(ctx, AHIP_E_INVALID_ARG, err);
    if (buildCurves(d, ctx->curves, err)) return fail(ctx, AHIP_E_INVALID_ARG, err);
    DScene& S = ctx->dsc;
    S = DScene{};
    if ((st = upload(ctx, ctx->dCNodes, ctx->curves.nodes.data(), ctx->curves.nodes.size()))) return st;
    if ((st = upload(ctx, ctx->dSegs, ctx->curves.segs.data(), ctx->curves.segs.size()))) return st;
    if ((st = upload(ctx, ctx->dSegIds, ctx->curves.segIds.data(), ctx->curves.segIds.size()))) return st;
    if ((st = upload(ctx, ctx->dStrandInfo, ctx->curves.strandInfo.data(), ctx->curves.strandInfo.size()))) return st;
    S.cnodes = ctx->dCNodes.as<CNode>();
    // child pairs for the pooled walk; not built (the whole-ray kernels run) when a leaf holds
    // more than 4 records or starts past 2^29 (the reference's builder makes leaves of at most 4)
    S.cwide = nullptr;
    if (!ctx->curves.nodes.empty()) {
        const std::vector<CNode>& cn = ctx->curves.nodes;
        std::vector<CWide> wide(cn.size());
        bool ok = true;
        auto ref = [&](uint32_t c) -> uint32_t {
            const CNode& n = cn[c];
            if (!(n.rc & 0x80000000u)) return c;
            const uint32_t cnt = n.rc & 0x7fffffffu;
            if (cnt == 0u || cnt > 4u || n.lp >= (1u << 29)) ok = false;
            return kCWLeaf | ((cnt - 1u) & 3u) << 29 | (n.lp & ((1u << 29) - 1u));
        };
        for (size_t i = 0; i < cn.size() && ok; ++i) {
            if (cn[i].rc & 0x80000000u) continue;
            CWide& w = wide[i];
            const CNode &a = cn[cn[i].lp], &b = cn[cn[i].rc];
            for (int k = 0; k < 3; ++k) {
                w.a[k] = a.bmin[k];
                w.a[3 + k] = a.bmax[k];
                w.b[k] = b.bmin[k];
                w.b[3 + k] = b.bmax[k];
            }
            w.refA = ref(cn[i].lp);
            w.refB = ref(cn[i].rc);
            w.pad0 = w.pad1 = 0u;
        }
        if (ok && (cn[0].rc & 0x80000000u)) ref(0u);   // a one-leaf tree: its root is pushed as a leaf
        if (ok) {
            if ((st = upload(ctx, ctx->dCWide, wide.data(), wide.size()))) return st;
            S.cwide = ctx->dCWide.as<CWide>();
        }
    }
    S.segs = ctx->dSegs.as<DSeg>();
    S.segIds = ctx->dSegIds.as<uint32_t>();
    S.strandInfo = ctx->dStrandInfo.as<float>();
    S.hasCurves = ctx->curves.nodes.empty() ? 0u : 1u;
    S.hasHair = hs.hasHair ? 1u : 0u;
    S.curveDepth = ctx->params.curve_depth ? ctx->params.curve_depth : kMaxCurveDepth;
    if ((st = upload(ctx, ctx->dNodes, ctx->bvh.nodes4.data(), ctx->bvh.nodes4.size()))) return st;
    for (DTri& t : ctx->bvh.tris)   // the material of each triangle, so shading skips the mesh table
        t.matIndex = t.meshID < hs.meshMat.size() ? hs.meshMat[t.meshID] : -1;
    if ((st = upload(ctx, ctx->dTris, ctx->bvh.tris.data(), ctx->bvh.tris.size()))) return st;
    if ((st = upload(ctx, ctx->dShade, ctx->bvh.shade.data(), ctx->bvh.shade.size()))) return st;
    if ((st = upload(ctx, ctx->dMats, hs.mats.data(), hs.mats.size()))) return st;
    if ((st = upload(ctx, ctx->dMeshMat, hs.meshMat.data(), hs.meshMat.size()))) return st;
    if ((st = upload(ctx, ctx->dLights, hs.lights.data(), hs.lights.size()))) return st;
    for (auto* b : ctx->texBufs) delete b;
    ctx->texBufs.clear();
    std::vector<DTexture> tex;
    for (size_t i = 0; i < hs.texPixels.size(); ++i) {
        auto* b = new DevBuf();
        ctx->texBufs.push_back(b);
        if ((st = upload(ctx, *b, hs.texPixels[i].data(), hs.texPixels[i].size()))) return st;
        tex.push_back(DTexture{b->as<float>(), hs.texW[i], hs.texH[i]});
    }
    if ((st = upload(ctx, ctx->dTexDesc, tex.data(), tex.size()))) return st;
    ctx->texW = hs.texW;
    ctx->texH = hs.texH;
    ctx->numCommit = d->num_indices / 3u;
    ctx->dTriAlb.release();
    if (d->tri_albedo && ctx->numCommit) {
        if ((st = upload(ctx, ctx->dTriAlb, d->tri_albedo, 3 * (size_t)ctx->numCommit))) return st;
        S.triAlbedo = ctx->dTriAlb.as<float>();
    }
    ctx->albESet = ctx->texESet = false;
    ctx->diffTex = 0;
    S.nodes4 = ctx->dNodes.as<DNode4>();
    S.tris = ctx->dTris.as<DTri>();
    S.shade = ctx->dShade.as<DTriShade>();
    S.mats = ctx->dMats.as<DMat>();
    S.meshMat = ctx->dMeshMat.as<int32_t>();
    S.lights = ctx->dLights.as<DLight>();
    S.textures = ctx->dTexDesc.as<DTexture>();
    S.numMeshes = d->num_meshes;
    S.numLights = (uint32_t)hs.lights.size();
    S.numNodes = (uint32_t)ctx->bvh.nodes4.size();
    S.numTris = ctx->bvh.numTris;
    S.hasBVH = ctx->bvh.numTris > 0 ? 1u : 0u;
    S.rootRef = ctx->bvh.rootRef;
    std::memcpy(S.rootMin, ctx->bvh.rootMin, 12);
    std::memcpy(S.rootMax, ctx->bvh.rootMax, 12);
    S.hasEmitters = hs.hasEmitters ? 1u : 0u;
    S.hasTransparent = hs.hasTransparent ? 1u : 0u;
    S.hasAlpha = hs.hasAlpha ? 1u : 0u;
    S.hasDirectional = 0u;
    for (const DLight& L : hs.lights) S.hasDirectional |= L.type == AHIP_LIGHT_DIRECTIONAL ? 1u : 0u;
    std::memcpy(S.envRadiance, d->env_radiance, 12);
    if (hs.hasDome) {
        const HostDome& D = hs.dome;
        if ((st = upload(ctx, ctx->dDomePx, D.px4.data(), D.px4.size()))) return st;
        if ((st = upload(ctx, ctx->dDomeMF, D.margFunc.data(), D.margFunc.size()))) return st;
        if ((st = upload(ctx, ctx->dDomeMC, D.margCdf.data(), D.margCdf.size()))) return st;
        if ((st = upload(ctx, ctx->dDomeCF, D.condFunc.data(), D.condFunc.size()))) return st;
        if ((st = upload(ctx, ctx->dDomeCC, D.condCdf.data(), D.condCdf.size()))) return st;
        if ((st = upload(ctx, ctx->dDomeCI, D.condInt.data(), D.condInt.size()))) return st;
        if ((st = upload(ctx, ctx->dDomeMG, D.margGuide.data(), D.margGuide.size()))) return st;
        if ((st = upload(ctx, ctx->dDomeCG, D.condGuide.data(), D.condGuide.size()))) return st;
        S.hasEnv = 1u;
        S.dome.px = ctx->dDomePx.as<float>();
        S.dome.margFunc = ctx->dDomeMF.as<float>();
        S.dome.margCdf = ctx->dDomeMC.as<float>();
        S.dome.condFunc = ctx->dDomeCF.as<float>();
        S.dome.condCdf = ctx->dDomeCC.as<float>();
        S.dome.condInt = ctx->dDomeCI.as<float>();
        S.dome.margGuide = ctx->dDomeMG.as<uint32_t>();
        S.dome.condGuide = ctx->dDomeCG.as<uint32_t>();
        S.dome.margM = D.margM;
        S.dome.condM = D.condM;
        S.dome.margInt = D.margInt;
        S.dome.W = D.W;
        S.dome.H = D.H;
        S.dome.intensity = D.intensity;
        std::memcpy(S.dome.rot, D.rot, sizeof(D.rot));
    }
    // BDPT takes triangles with Lambert / Emissive / StandardSurface materials without a roughness
    // texture (the vertices keep no uv, and connections evaluate the BSDF at stored vertices; an
    // opacity texture is read only at fresh hits: the camera subpath's alpha test and the
    // connections' transmittance) lit by quad area lights, the dome and directional lights
    ctx->bdptRefusal.clear();
    if (!ctx->curves.nodes.empty()) ctx->bdptRefusal = "curves";
    else if (d->env_radiance[0] != 0.f || d->env_radiance[1] != 0.f || d->env_radiance[2] != 0.f)
        ctx->bdptRefusal = "a constant environment radiance (no light)";
    if (ctx->bdptRefusal.empty() && d->tri_albedo) ctx->bdptRefusal = "per-triangle albedo";
    for (const DMat& m : hs.mats)
        if (ctx->bdptRefusal.empty() && m.type != AHIP_MAT_LAMBERT && m.type != AHIP_MAT_EMISSIVE &&
            !(m.type == AHIP_MAT_STANDARD && m.roughTex < 0))
            ctx->bdptRefusal = "a material other than Lambert / Emissive / StandardSurface without a roughness texture";
    for (const DLight& L : hs.lights)
        if (ctx->bdptRefusal.empty() && L.type != AHIP_LIGHT_AREA && !(L.sceneR > 0.f))
            ctx->bdptRefusal = "a dome or directional light without a scene_radius";
    if (ctx->bdptRefusal.empty() && !hs.lights.empty()) {
        // LightSampler::build (LightSampler.h:28-69): Vose's alias table over the lights' power:
        // AreaLight::power (AreaLight.h:114-117), DomeLight::power (DomeLight.h:176-179, its
        // integral from :255-257), DirectionalLight::power (DirectionalLight.h:136-139, 41)
        constexpr float kPiF = 3.14159265358979323846f, k2PiF = 6.28318530717958647692f;
        const uint32_t n = (uint32_t)hs.lights.size();
        std::vector<float> w(n), prob(n);
        std::vector<uint32_t> alias(n, 0u), small, large;
        float total = 0.f;
        for (uint32_t i = 0; i < n; ++i) {
            const DLight& L = hs.lights[i];
            const float lum = 0.2722287f * L.Le[0] + 0.6740818f * L.Le[1] + 0.0536895f * L.Le[2];
            float pw;
            if (L.type == AHIP_LIGHT_DOME) {
                const float envIntegral = hs.dome.margInt * k2PiF * kPiF / (float)(hs.dome.W * hs.dome.H);
                pw = envIntegral * 4.f * kPiF * hs.dome.intensity;
            } else if (L.type == AHIP_LIGHT_DIRECTIONAL) {
                const float diskArea = kPiF * L.sceneR * L.sceneR;
                pw = lum * diskArea * kPiF;
            } else {
                pw = lum * L.area * kPiF;
            }
            w[i] = std::max(0.f, pw);
            total += w[i];
        }
        if (total <= 0.f) {
            for (float& x : w) x = 1.f;
            total = (float)n;
        }
        const float inv = 1.f / total;
        for (float& x : w) x *= inv;
        for (uint32_t i = 0; i < n; ++i) {
            prob[i] = w[i] * (float)n;
            (prob[i] < 1.f ? small : large).push_back(i);
        }
        while (!small.empty() && !large.empty()) {
            const uint32_t sm = small.back(), lg = large.back();
            small.pop_back();
            large.pop_back();
            alias[sm] = lg;
            prob[lg] = prob[lg] + prob[sm] - 1.f;
            (prob[lg] < 1.f ? small : large).push_back(lg);
        }
        for (uint32_t i : small) prob[i] = 1.f;
        for (uint32_t i : large) prob[i] = 1.f;
        if ((st = upload(ctx, ctx->dAliasProb, prob.data(), n))) return st;
        if ((st = upload(ctx, ctx->dAlias, alias.data(), n))) return st;
        if ((st = upload(ctx, ctx->dLightW, w.data(), n))) return st;
    }
    if ((st = ensureStack(ctx))) return st;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->prepared = true;
    return AHIP_OK;
}

int ahip_bvh_info(ahip_context* ctx, uint32_t* num_nodes, uint32_t* num_tris, uint32_t* max_depth) {
    ENTER();
    if (!ctx->prepared) return fail(ctx, AHIP_E_NOT_PREPARED, "prepare first");
    if (num_nodes) *num_nodes = (uint32_t)ctx->bvh.nodes.size();
    if (num_tris) *num_tris = ctx->bvh.numTris;
    if (max_depth) *max_depth = ctx->bvh.maxDepth;
    return AHIP_OK;
}

int ahip_bvh_dump(ahip_context* ctx, float* nodes32, uint32_t* prim_indices) {
    ENTER();
    if (!ctx->prepared) return fail(ctx, AHIP_E_NOT_PREPARED, "prepare first");
    if (nodes32) std::memcpy(nodes32, ctx->bvh.refNodes.data(), ctx->bvh.refNodes.size() * sizeof(RefNode));
    if (prim_indices) std::memcpy(prim_indices, ctx->bvh.refOrder.data(), ctx->bvh.refOrder.size() * 4);
    return AHIP_OK;
}

int ahip_set_camera(ahip_context* ctx, const ahip_camera* cam) {
    ENTER();
    if (!cam || cam->width == 0 || cam->height == 0) return fail(ctx, AHIP_E_INVALID_ARG, "bad camera");
    ctx->cam = *cam;
    ctx->hasCam = true;
    return AHIP_OK;
}

int ahip_set_params(ahip_context* ctx, const ahip_render_params* p) {
    ENTER();
    if (!p || p->max_depth > 1000u) return fail(ctx, AHIP_E_INVALID_ARG, "bad render params");
    if (p->curve_depth > kMaxCurveDepth) return fail(ctx, AHIP_E_INVALID_ARG, "curve_depth must be 0..6");
    if (p->integrator > AHIP_INTEGRATOR_BDPT) return fail(ctx, AHIP_E_INVALID_ARG, "unknown integrator");
    if (!(p->firefly_clamp >= 0.f) || !std::isfinite(p->firefly_clamp))
        return fail(ctx, AHIP_E_INVALID_ARG, "firefly_clamp must be finite and >= 0");
    if (p->integrator == AHIP_INTEGRATOR_BDPT && ctx->prepared && !ctx->bdptRefusal.empty())
        return fail(ctx, AHIP_E_UNSUPPORTED, "BDPT takes triangles with Lambert / Emissive / StandardSurface materials "
                                             "and area / dome / directional lights; the scene has " + ctx->bdptRefusal);
    ctx->params = *p;
    ctx->dsc.curveDepth = p->curve_depth ? p->curve_depth : kMaxCurveDepth;
    return AHIP_OK;
}

int ahip_set_launch(ahip_context* ctx, uint32_t max_paths, uint32_t tile) {
    ENTER();
    if (max_paths) ctx->maxPaths = std::max(64u, max_paths);
    if (tile) {
        if (tile != 4u && tile != 8u && tile != 16u) return fail(ctx, AHIP_E_INVALID_ARG, "tile width must be 4, 8 or 16");
        ctx->tileW = tile;
    }
    return AHIP_OK;
}

int ahip_clear_accum(ahip_context* ctx) {
    ENTER();
    if (ctx->dAccum.p) {
        launchFillF32(ctx->dAccum.as<float>(), (size_t)ctx->accW * ctx->accH * 4, 0.f, ctx->stream);
        launchFillF32(ctx->dAccumLsq.as<float>(), (size_t)ctx->accW * ctx->accH, 0.f, ctx->stream);
        if (ctx->dAccumAov.p) launchFillF32(ctx->dAccumAov.as<float>(), (size_t)ctx->accW * ctx->accH * 6, 0.f, ctx->stream);
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    ctx->accumHasSamples = false;
    ctx->aovPartial = false;
    return AHIP_OK;
}

static int renderCall(ahip_context* ctx, const std::vector<ViewJob>& jobs, bool graph, unsigned long long* counts) {
    int st = checkInjected(ctx);
    if (st) return st;
    st = ensureQueues(ctx, 64);
    if (st) return st;
    launchFillU32(reinterpret_cast<uint32_t*>(ctx->qTotals.p), (uint32_t)(ctx->qTotals.bytes / 4), 0u, ctx->stream);
    HIPCHK(hipEventRecord(ctx->renderBeg, ctx->stream));
    st = renderJobs(ctx, jobs, graph, counts);
    if (st) return st;
    HIPCHK(hipEventRecord(ctx->renderEnd, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->renderBeg, ctx->renderEnd));
    ctx->stats.render_ms = ms;
    return finishStats(ctx);
}

// Make sure the persistent film-sized accumulation (and, when on, the AOV sums) exists.
static int ensureFrameAccum(ahip_context* ctx, uint32_t width, uint32_t height);

// BDPTIntegrator::renderTile over the tile (x0, y0, w, h) x [s0, s0 + ns): chunks of whole pixels,
// each (pixel, sample) slot on one lane (bdpt.hip), added into outAcc / outLsq / outAov at
// (outX0 + x, outY0 + y) with row stride outStride.
static int bdptRender(ahip_context* ctx, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t s0, uint32_t ns,
                      float4* outAcc, float* outLsq, float* outAov, uint32_t outStride, uint32_t outX0,
                      uint32_t outY0) {
    if (!ctx->bdptRefusal.empty())
        return fail(ctx, AHIP_E_UNSUPPORTED, "BDPT takes triangles with Lambert / Emissive / StandardSurface materials "
                                             "and area / dome / directional lights; the scene has " + ctx->bdptRefusal);
    if (ns == 0u) return AHIP_OK;
    int st = checkInjected(ctx);
    if (st) return st;
    st = ensureQueues(ctx, 64);
    if (st) return st;
    // about 512K slots per chunk; per slot 2 x (maxDepth + 2) vertices of 80 B, up to
    // (maxDepth + 2)(maxDepth + 3) strategy terms of 16 B and as many staged and queued shadow rays
    // of 32 B each (10.4 KB at depth 8: 5.5 GB per chunk)
    const size_t maxV = (size_t)ctx->params.max_depth + 2u, R = maxV * (maxV + 1u);
    const size_t pixPerChunk = std::max<size_t>(1u, (size_t)(1u << 19) / ns);
    const size_t cap = pixPerChunk * ns;
    if (cap * R > (size_t)1 << 31) return fail(ctx, AHIP_E_INVALID_ARG, "BDPT chunk too large");
    if (ctx->bdptCap < cap || ctx->bdptMaxV < maxV || !ctx->dBdptVerts.p) {
        HIPCHK(ctx->dBdptVerts.alloc(cap * 2u * maxV * 5u * sizeof(float4)));
        HIPCHK(ctx->dBdptCounts.alloc(cap * sizeof(uint32_t)));
        HIPCHK(ctx->dBdptAovA.alloc(cap * sizeof(float4)));
        HIPCHK(ctx->dBdptAovN.alloc(cap * sizeof(float4)));
        HIPCHK(ctx->dBdptL.alloc(cap * sizeof(float4)));
        HIPCHK(ctx->dBdptTerms.alloc(cap * R * sizeof(float4)));
        HIPCHK(ctx->dBdptNTerms.alloc(cap * sizeof(uint32_t)));
        HIPCHK(ctx->dBdptConnO.alloc(cap * R * sizeof(float4)));
        HIPCHK(ctx->dBdptConnD.alloc(cap * R * sizeof(float4)));
        HIPCHK(ctx->dBdptStageO.alloc(cap * R * sizeof(float4)));
        HIPCHK(ctx->dBdptStageD.alloc(cap * R * sizeof(float4)));
        HIPCHK(ctx->dBdptCtr.alloc(2 * sizeof(uint32_t)));
        ctx->bdptCap = cap;
        ctx->bdptMaxV = maxV;
    }
    BdptArgs a{};
    a.cam = toDCam(ctx->cam);
    a.x0 = x0; a.y0 = y0; a.w = w; a.h = h; a.s0 = s0; a.ns = ns;
    a.seed = ctx->params.seed;
    a.maxDepth = ctx->params.max_depth;
    a.maxV = (uint32_t)maxV;
    a.clamp = ctx->params.firefly_clamp;
    a.cap = (uint32_t)ctx->bdptCap;
    a.verts = ctx->dBdptVerts.as<float4>();
    a.counts = ctx->dBdptCounts.as<uint32_t>();
    a.aovA = ctx->dBdptAovA.as<float4>();
    a.aovN = ctx->dBdptAovN.as<float4>();
    a.sampleL = ctx->dBdptL.as<float4>();
    a.terms = ctx->dBdptTerms.as<float4>();
    a.nTerms = ctx->dBdptNTerms.as<uint32_t>();
    a.connO = ctx->dBdptConnO.as<float4>();
    a.connD = ctx->dBdptConnD.as<float4>();
    a.stageO = ctx->dBdptStageO.as<float4>();
    a.stageD = ctx->dBdptStageD.as<float4>();
    a.connCount = ctx->dBdptCtr.as<uint32_t>() + 1;
    a.aliasProb = ctx->dAliasProb.as<float>();
    a.alias = ctx->dAlias.as<uint32_t>();
    a.lightW = ctx->dLightW.as<float>();
    a.ovfRef = ctx->qOvfRef.as<uint32_t>();
    a.ovfT = ctx->qOvfT.as<float>();
    a.stackCap = ctx->stackCap;
    a.overflow = ctx->dBdptCtr.as<uint32_t>();
    a.outAcc = outAcc;
    a.outLsq = outLsq;
    a.outAov = outAov;
    a.outStride = outStride;
    a.outX0 = outX0;
    a.outY0 = outY0;
    launchFillU32(a.overflow, 1, 0u, ctx->stream);
    HIPCHK(hipEventRecord(ctx->renderBeg, ctx->stream));
    const size_t npix = (size_t)w * h;
    for (size_t p0 = 0; p0 < npix; p0 += pixPerChunk) {
        a.pix0 = (uint32_t)p0;
        a.slots = (uint32_t)(std::min(pixPerChunk, npix - p0) * ns);
        // every slot on its own lane; never more threads than the traversal overflow stacks hold
        const uint32_t blocks = std::min<uint32_t>(ctx->traceBlocks, (a.slots + 127u) / 128u);
        launchBdpt(ctx->dsc, a, blocks, ctx->traceBlocks, ctx->stream);
    }
    HIPCHK(hipEventRecord(ctx->renderEnd, ctx->stream));
    uint32_t ovf = 0;
    HIPCHK(hipMemcpyAsync(&ovf, a.overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ovf) return fail(ctx, AHIP_E_HIP, "traversal stack capacity exceeded");
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->renderBeg, ctx->renderEnd));
    ctx->stats.render_ms = ms;
    ctx->stats.samples += (uint64_t)npix * ns;
    return AHIP_OK;
}

static int renderFrameImpl(ahip_context* ctx, uint32_t width, uint32_t height, uint32_t sample_start,
                           uint32_t sample_count) {
    if (!ctx->prepared) return fail(ctx, AHIP_E_NOT_PREPARED, "prepare first");
    if (!ctx->hasCam) return fail(ctx, AHIP_E_INVALID_ARG, "set_camera first");
    if (width != ctx->cam.width || height != ctx->cam.height)
        return fail(ctx, AHIP_E_INVALID_ARG, "film size differs from the camera's");
    int st = ensureFrameAccum(ctx, width, height);
    if (st) return st;
    if (ctx->params.integrator == AHIP_INTEGRATOR_BDPT) {
        if (ctx->aovOn) return fail(ctx, AHIP_E_UNSUPPORTED, "AOVs of BDPT frames (render tiles with AOVs instead)");
        st = bdptRender(ctx, 0, 0, width, height, sample_start, sample_count, ctx->dAccum.as<float4>(),
                        ctx->dAccumLsq.as<float>(), nullptr, width, 0, 0);
    } else {
        std::vector<ViewJob> jobs{ViewJob{ctx->cam, 0, 0, width, height, sample_start, sample_count,
                                          ctx->dAccum.as<float4>(), ctx->dAccumLsq.as<float>(),
                                          ctx->aovOn ? ctx->dAccumAov.as<float>() : nullptr}};
        st = renderCall(ctx, jobs, ctx->frameGraph, nullptr);
    }
    if (st) return st;
    ctx->accumHasSamples = ctx->accumHasSamples || sample_count > 0;
    ctx->lastRenderValid = true;
    ctx->lastW = width; ctx->lastH = height; ctx->lastS0 = sample_start; ctx->lastNs = sample_count;
    return AHIP_OK;
}

int ahip_render_frame(ahip_context* ctx, uint32_t width, uint32_t height, uint32_t sample_start,
                      uint32_t sample_count) {
    ENTER();
    return renderFrameImpl(ctx, width, height, sample_start, sample_count);
}

static int ensureFrameAccum(ahip_context* ctx, uint32_t width, uint32_t height) {
    int st = ensureQueues(ctx, 64);
    if (st) return st;
    if (ctx->accW != width || ctx->accH != height || !ctx->dAccum.p) {
        HIPCHK(ctx->dAccum.alloc((size_t)width * height * sizeof(float4)));
        HIPCHK(ctx->dAccumLsq.alloc((size_t)width * height * sizeof(float)));
        ctx->dAccumAov.release();
        ctx->accW = width;
        ctx->accH = height;
        launchFillF32(ctx->dAccum.as<float>(), (size_t)width * height * 4, 0.f, ctx->stream);
        launchFillF32(ctx->dAccumLsq.as<float>(), (size_t)width * height, 0.f, ctx->stream);
        ctx->accumHasSamples = false;
        ctx->aovPartial = false;
    }
    if (ctx->aovOn && !ctx->dAccumAov.p) {
        // AOV sums start at zero when turned on; if the radiance sums already hold earlier samples,
        // the AOV sums cover fewer samples than the weight until the next clear
        ctx->aovPartial = ctx->accumHasSamples;
        HIPCHK(ctx->dAccumAov.alloc((size_t)width * height * 6 * sizeof(float)));
        launchFillF32(ctx->dAccumAov.as<float>(), (size_t)width * height * 6, 0.f, ctx->stream);
    }
    return AHIP_OK;
}

// RenderSession's adaptive two-pass render (RenderSession.cpp:549-653) on the device: the base
// pass over the whole frame, then the CPU tile path's allocation of the remaining samples by tile
// variance, rendered as one list of tile jobs into the same accumulation.
int ahip_render_adaptive(ahip_context* ctx, uint32_t width, uint32_t height, uint32_t tile_size, uint32_t total_spp,
                         uint32_t base_spp, uint32_t* extra_per_tile, float* tile_variance) {
    ENTER();
    if (ctx->params.integrator != AHIP_INTEGRATOR_PATH)
        return fail(ctx, AHIP_E_UNSUPPORTED, "ahip_render_adaptive runs the path tracer only");
    if (tile_size == 0 || total_spp == 0) return fail(ctx, AHIP_E_INVALID_ARG, "tile size and sample count must be > 0");
    uint32_t base = base_spp ? base_spp : std::min(32u, std::max(16u, total_spp / 16u));
    base = std::min(base, total_spp);
    const uint32_t extra = total_spp - base;
    int st = renderFrameImpl(ctx, width, height, 0, base);
    if (st) return st;
    const uint32_t tilesX = (width + tile_size - 1u) / tile_size, tilesY = (height + tile_size - 1u) / tile_size;
    const uint32_t nt = tilesX * tilesY;
    std::vector<float> var(nt, 0.f);
    std::vector<uint32_t> per(nt, 0u);
    if (extra > 0) {
        DevBuf dVar;
        HIPCHK(dVar.alloc(nt * sizeof(float)));
        launchTileVariance(ctx->dAccum.as<float4>(), ctx->dAccumLsq.as<float>(), width, height, tile_size,
                           dVar.as<float>(), ctx->stream);
        HIPCHK(hipMemcpyAsync(var.data(), dVar.p, nt * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        // half the extra samples everywhere, the other half by log(1 + variance) share
        std::vector<float> logW(nt);
        float totalW = 0.f;
        for (uint32_t i = 0; i < nt; ++i) logW[i] = std::log(1.f + var[i]);
        for (uint32_t i = 0; i < nt; ++i) totalW = totalW + logW[i];
        const uint32_t uniformHalf = extra / 2u, bonusHalf = extra - uniformHalf;
        std::vector<ViewJob> jobs;
        for (uint32_t i = 0; i < nt; ++i) {
            uint32_t bonus = 0;
            if (totalW > 1e-6f)
                bonus = (uint32_t)std::round(logW[i] / totalW * (float)(bonusHalf * nt));
            per[i] = std::min(uniformHalf + bonus, extra);
            if (per[i] == 0) continue;
            const uint32_t x0 = (i % tilesX) * tile_size, y0 = (i / tilesX) * tile_size;
            const size_t o = (size_t)y0 * width + x0;
            ViewJob j{ctx->cam, x0, y0, std::min(tile_size, width - x0), std::min(tile_size, height - y0), base, per[i],
                      ctx->dAccum.as<float4>() + o, ctx->dAccumLsq.as<float>() + o,
                      ctx->aovOn ? ctx->dAccumAov.as<float>() + 6 * o : nullptr};
            j.stride = width;
            jobs.push_back(j);
        }
        if (!jobs.empty()) {
            st = renderCall(ctx, jobs, ctx->frameGraph, nullptr);
            if (st) return st;
        }
    }
    if (extra_per_tile) std::copy(per.begin(), per.end(), extra_per_tile);
    if (tile_variance) std::copy(var.begin(), var.end(), tile_variance);
    return AHIP_OK;
}

int ahip_resolve(ahip_context* ctx, float exposure, uint8_t* srgb8, float* rgb, float* variance, float* albedo,
                 float* normal) {
    ENTER();
    if (!ctx->dAccum.p) return fail(ctx, AHIP_E_NOT_PREPARED, "nothing rendered");
    if ((albedo || normal) && !ctx->dAccumAov.p)
        return fail(ctx, AHIP_E_INVALID_ARG, "AOVs were not accumulated (ahip_set_aovs before rendering)");
    if ((albedo || normal) && ctx->aovPartial)
        return fail(ctx, AHIP_E_INVALID_ARG, "AOVs were turned on after samples were accumulated (clear first)");
    const size_t n = (size_t)ctx->accW * ctx->accH;
    DevBuf &dS = ctx->dResolve[0], &dR = ctx->dResolve[1], &dV = ctx->dResolve[2], &dA = ctx->dResolve[3],
           &dN = ctx->dResolve[4];
    if (srgb8) HIPCHK(dS.alloc(n * 3));
    if (rgb) HIPCHK(dR.alloc(n * 3 * sizeof(float)));
    if (variance) HIPCHK(dV.alloc(n * sizeof(float)));
    if (albedo) HIPCHK(dA.alloc(n * 3 * sizeof(float)));
    if (normal) HIPCHK(dN.alloc(n * 3 * sizeof(float)));
    const float evScale = std::pow(2.f, exposure);
    launchResolve(ctx->dAccum.as<float4>(), ctx->dAccumLsq.as<float>(), ctx->dAccumAov.as<float>(), n, evScale,
                  srgb8 ? dS.as<uint8_t>() : nullptr, rgb ? dR.as<float>() : nullptr,
                  variance ? dV.as<float>() : nullptr, albedo ? dA.as<float>() : nullptr,
                  normal ? dN.as<float>() : nullptr, ctx->stream);
    HIPCHK(hipGetLastError());
    auto get = [&](const DevBuf& d, void* host, size_t bytes) -> hipError_t {
        return host ? hipMemcpyAsync(host, d.p, bytes, hipMemcpyDeviceToHost, ctx->stream) : hipSuccess;
    };
    HIPCHK(get(dS, srgb8, n * 3));
    HIPCHK(get(dR, rgb, n * 3 * sizeof(float)));
    HIPCHK(get(dV, variance, n * sizeof(float)));
    HIPCHK(get(dA, albedo, n * 3 * sizeof(float)));
    HIPCHK(get(dN, normal, n * 3 * sizeof(float)));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return AHIP_OK;
}

static int downloadRegion(ahip_context* ctx, const DevBuf& acc, const DevBuf& lsq, size_t n, float* out) {
    std::vector<float> a(n * 4), l(n);
    HIPCHK(hipMemcpy(a.data(), acc.p, n * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(l.data(), lsq.p, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
        out[5 * i + 0] = a[4 * i + 0];
        out[5 * i + 1] = a[4 * i + 1];
        out[5 * i + 2] = a[4 * i + 2];
        out[5 * i + 3] = a[4 * i + 3];
        out[5 * i + 4] = l[i];
    }
    return AHIP_OK;
}

int ahip_download_accum(ahip_context* ctx, float* out) {
    ENTER();
    if (!out) return fail(ctx, AHIP_E_INVALID_ARG, "null output");
    if (!ctx->dAccum.p) return fail(ctx, AHIP_E_NOT_PREPARED, "nothing rendered");
    return downloadRegion(ctx, ctx->dAccum, ctx->dAccumLsq, (size_t)ctx->accW * ctx->accH, out);
}

static int renderTileImpl(ahip_context* ctx, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t sample_start,
                          uint32_t sample_count, float* out, float* albedo, float* normal) {
    if (!ctx->prepared) return fail(ctx, AHIP_E_NOT_PREPARED, "prepare first");
    if (!ctx->hasCam) return fail(ctx, AHIP_E_INVALID_ARG, "set_camera first");
    if (!out || w == 0 || h == 0) return fail(ctx, AHIP_E_INVALID_ARG, "bad tile");
    const bool aov = albedo != nullptr || normal != nullptr;
    if (aov && (!albedo || !normal)) return fail(ctx, AHIP_E_INVALID_ARG, "albedo and normal both needed");
    w = std::min(w, ctx->cam.width > x0 ? ctx->cam.width - x0 : 0u);
    h = std::min(h, ctx->cam.height > y0 ? ctx->cam.height - y0 : 0u);
    if (w == 0 || h == 0) return fail(ctx, AHIP_E_INVALID_ARG, "tile outside the film");
    int st = ensureQueues(ctx, 64);
    if (st) return st;
    const size_t n = (size_t)w * h;
    DevBuf acc, lsq, aovBuf;
    HIPCHK(acc.alloc(n * sizeof(float4)));
    HIPCHK(lsq.alloc(n * sizeof(float)));
    launchFillF32(acc.as<float>(), n * 4, 0.f, ctx->stream);
    launchFillF32(lsq.as<float>(), n, 0.f, ctx->stream);
    if (aov) {
        HIPCHK(aovBuf.alloc(n * 6 * sizeof(float)));
        launchFillF32(aovBuf.as<float>(), n * 6, 0.f, ctx->stream);
    }
    const bool bdpt = ctx->params.integrator == AHIP_INTEGRATOR_BDPT;
    if (bdpt) {
        st = bdptRender(ctx, x0, y0, w, h, sample_start, sample_count, acc.as<float4>(), lsq.as<float>(),
                        aov ? aovBuf.as<float>() : nullptr, w, 0, 0);
    } else {
        std::vector<ViewJob> jobs{ViewJob{ctx->cam, x0, y0, w, h, sample_start, sample_count, acc.as<float4>(),
                                          lsq.as<float>(), aov ? aovBuf.as<float>() : nullptr}};
        st = renderCall(ctx, jobs, false, nullptr);
    }
    if (st) return st;
    if ((st = downloadRegion(ctx, acc, lsq, n, out))) return st;
    if (!aov) return AHIP_OK;
    if (bdpt) {   // BDPTIntegrator::renderTile records each pixel's means once (BDPTIntegrator.cpp:567-580)
        std::vector<float> sums(n * 6);
        HIPCHK(hipMemcpy(sums.data(), aovBuf.p, n * 6 * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) {
                albedo[3 * i + c] = sums[6 * i + c];
                normal[3 * i + c] = sums[6 * i + 3 + c];
            }
        return AHIP_OK;
    }
    // PathIntegrator::renderTile (PathIntegrator.cpp:58-66): mean over the tile's samples; the mean
    // normal scaled to unit length when longer than 1e-6
    std::vector<float> sums(n * 6);
    HIPCHK(hipMemcpy(sums.data(), aovBuf.p, n * 6 * sizeof(float), hipMemcpyDeviceToHost));
    const float invN = sample_count ? 1.f / (float)sample_count : 0.f;
    for (size_t i = 0; i < n; ++i) {
        const float* a = sums.data() + 6 * i;
        for (int c = 0; c < 3; ++c) albedo[3 * i + c] = a[c] * invN;
        float nx = a[3] * invN, ny = a[4] * invN, nz = a[5] * invN;
        float len = std::sqrt(nx * nx + ny * ny + nz * nz);   // Vec3f::length (Types.h)
        if (len > 1e-6f) {
            float il = 1.f / len;
            nx *= il; ny *= il; nz *= il;
        }
        normal[3 * i + 0] = nx;
        normal[3 * i + 1] = ny;
        normal[3 * i + 2] = nz;
    }
    return AHIP_OK;
}

int ahip_render_tile(ahip_context* ctx, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t sample_start,
                     uint32_t sample_count, float* out) {
    ENTER();
    return renderTileImpl(ctx, x0, y0, w, h, sample_start, sample_count, out, nullptr, nullptr);
}

int ahip_render_tile_aovs(ahip_context* ctx, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                          uint32_t sample_start, uint32_t sample_count, float* out, float* albedo, float* normal) {
    ENTER();
    if (!albedo || !normal) return fail(ctx, AHIP_E_INVALID_ARG, "null AOV output");
    return renderTileImpl(ctx, x0, y0, w, h, sample_start, sample_count, out, albedo, normal);
}

int ahip_set_aovs(ahip_context* ctx, int enable) {
    ENTER();
    ctx->aovOn = enable != 0;
    if (!ctx->aovOn) ctx->dAccumAov.release();
    return AHIP_OK;
}

int ahip_download_aovs(ahip_context* ctx, float* out) {
    ENTER();
    if (!out) return fail(ctx, AHIP_E_INVALID_ARG, "null output");
    if (!ctx->dAccumAov.p) return fail(ctx, AHIP_E_NOT_PREPARED, "no AOVs recorded (ahip_set_aovs, then render)");
    HIPCHK(hipMemcpy(out, ctx->dAccumAov.p, (size_t)ctx->accW * ctx->accH * 6 * sizeof(float), hipMemcpyDeviceToHost));
    return AHIP_OK;
}

int ahip_render_batch(ahip_context* ctx, const ahip_view* views, uint32_t num_views) {
    ENTER();
    if (ctx->params.integrator != AHIP_INTEGRATOR_PATH)
        return fail(ctx, AHIP_E_UNSUPPORTED, "ahip_render_batch runs the path tracer only");
    if (!ctx->prepared) return fail(ctx, AHIP_E_NOT_PREPARED, "prepare first");
    if (!views && num_views) return fail(ctx, AHIP_E_INVALID_ARG, "null views");
    // One accumulation buffer holds every view's film; views sharing a max depth share passes.
    std::vector<size_t> off(num_views + 1, 0);
    for (uint32_t v = 0; v < num_views; ++v) {
        const ahip_camera& c = views[v].camera;
        if (!c.width || !c.height) return fail(ctx, AHIP_E_INVALID_ARG, "bad view");
        off[v + 1] = off[v] + (size_t)c.width * c.height;
    }
    if (!num_views) return AHIP_OK;
    HIPCHK(ctx->dBatchAcc.alloc(off[num_views] * sizeof(float4)));
    HIPCHK(ctx->dBatchLsq.alloc(off[num_views] * sizeof(float)));
    launchFillF32(ctx->dBatchAcc.as<float>(), off[num_views] * 4, 0.f, ctx->stream);
    launchFillF32(ctx->dBatchLsq.as<float>(), off[num_views], 0.f, ctx->stream);
    ahip_render_params savedP = ctx->params;
    std::vector<uint32_t> depths;
    for (uint32_t v = 0; v < num_views; ++v) {
        uint32_t d = views[v].max_depth ? views[v].max_depth : savedP.max_depth;
        if (std::find(depths.begin(), depths.end(), d) == depths.end()) depths.push_back(d);
    }
    int st = AHIP_OK;
    double ms = 0.0;
    for (uint32_t d : depths) {
        std::vector<ViewJob> jobs;
        for (uint32_t v = 0; v < num_views; ++v) {
            const ahip_view& V = views[v];
            if ((V.max_depth ? V.max_depth : savedP.max_depth) != d) continue;
            jobs.push_back(ViewJob{V.camera, 0, 0, V.camera.width, V.camera.height, V.sample_start, V.sample_count,
                                   ctx->dBatchAcc.as<float4>() + off[v], ctx->dBatchLsq.as<float>() + off[v]});
        }
        ctx->params.max_depth = d;
        st = renderCall(ctx, jobs, ctx->batchGraph, nullptr);
        ms += ctx->stats.render_ms;
        if (st) break;
    }
    ctx->params = savedP;
    if (st) return st;
    ctx->stats.render_ms = ms;
    std::vector<float> a(off[num_views] * 4), l(off[num_views]);
    HIPCHK(hipMemcpy(a.data(), ctx->dBatchAcc.p, a.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(l.data(), ctx->dBatchLsq.p, l.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t v = 0; v < num_views; ++v) {
        if (!views[v].accum) continue;
        for (size_t i = off[v]; i < off[v + 1]; ++i) {
            float* o = views[v].accum + 5 * (i - off[v]);
            o[0] = a[4 * i]; o[1] = a[4 * i + 1]; o[2] = a[4 * i + 2]; o[3] = a[4 * i + 3]; o[4] = l[i];
        }
    }
    return AHIP_OK;
}

int ahip_trace_rays(ahip_context* ctx, const float* rays, uint32_t n, ahip_hit* hits) {
    ENTER();
    if (!ctx->prepared) return fail(ctx, AHIP_E_NOT_PREPARED, "prepare first");
    if (n == 0) return AHIP_OK;
    if (!rays || !hits) return fail(ctx, AHIP_E_INVALID_ARG, "null arrays");
    int st = checkInjected(ctx);
    if (st) return st;
    DevBuf dr, dh, dflag;
    HIPCHK(dr.alloc((size_t)n * 32));
    HIPCHK(dh.alloc((size_t)n * sizeof(ahip_hit)));
    HIPCHK(dflag.alloc(4));
    HIPCHK(hipMemcpyAsync(dr.p, rays, (size_t)n * 32, hipMemcpyHostToDevice, ctx->stream));
    launchFillU32(dflag.as<uint32_t>(), 1, 0u, ctx->stream);
    launchTraceBatch(ctx->dsc, dr.as<float>(), n, dh.as<ahip_hit>(), nullptr, false, ctx->qOvfRef.as<uint32_t>(),
                     ctx->qOvfT.as<float>(), ctx->stackCap, ctx->traceBlocks, dflag.as<uint32_t>(), nullptr,
                     ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hits, dh.p, (size_t)n * sizeof(ahip_hit), hipMemcpyDeviceToHost, ctx->stream));
    uint32_t flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, dflag.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (flag) return fail(ctx, AHIP_E_HIP, "traversal stack capacity exceeded");
    return AHIP_OK;
}

// ahip_trace_rays on caller-owned device arrays, with the traversal's overflow stack and flag in
// this device's fixed scratch array: no allocation.  The same kernel and tree as ahip_trace_rays,
// so the hits are identical (the stack's grid stride differs, which only moves where entries live).
int ahip_trace_rays_noalloc(ahip_context* ctx, const float* rays, uint32_t n, ahip_hit* hits) {
    ENTER();
    if (!ctx->prepared) return fail(ctx, AHIP_E_NOT_PREPARED, "prepare first");
    if (n == 0) return AHIP_OK;
    if (!rays || !hits) return fail(ctx, AHIP_E_INVALID_ARG, "null arrays");
    if (n > AHIP_NOALLOC_MAX_RAYS)
        return fail(ctx, AHIP_E_INVALID_ARG, "ahip_trace_rays_noalloc: " + std::to_string(n) +
                                                 " rays, more than AHIP_NOALLOC_MAX_RAYS");
    if (!ctx->scratch) return fail(ctx, AHIP_E_HIP, "ahip_trace_rays_noalloc: " + ctx->scratchErr);
    const uint32_t K = lds_stack_entries();
    const uint32_t extra = ctx->stackCap > K ? ctx->stackCap - K : 0u;
    if (extra > AHIP_NOALLOC_MAX_STACK)
        return fail(ctx, AHIP_E_UNSUPPORTED,
                    "ahip_trace_rays_noalloc: the scene's BVH needs " + std::to_string(extra) +
                        " traversal stack entries beyond the LDS ones; the fixed scratch holds " +
                        std::to_string(AHIP_NOALLOC_MAX_STACK));
    const uint32_t blocks = kNoAllocThreads / trace_block_size();
    if (blocks * trace_block_size() != kNoAllocThreads)
        return fail(ctx, AHIP_E_HIP, "ahip_trace_rays_noalloc: block size does not divide the scratch grid");
    int st = checkInjected(ctx);
    if (st) return st;
    unsigned char* base = static_cast<unsigned char*>(ctx->scratch);
    uint32_t* flag = reinterpret_cast<uint32_t*>(base);
    uint32_t* ovfRef = reinterpret_cast<uint32_t*>(base + kNoAllocHeader);
    float* ovfT = reinterpret_cast<float*>(base + kNoAllocHeader + kNoAllocStackWords * sizeof(uint32_t));
    // the array keeps the previous call's contents: the flag must start at zero (a fill kernel, so
    // the call stays valid inside a captured graph); stack entries are written before they are read
    launchFillU32(flag, 1, 0u, ctx->stream);
    launchTraceBatch(ctx->dsc, rays, n, hits, nullptr, false, ovfRef, ovfT, ctx->stackCap, blocks, flag, nullptr,
                     ctx->stream);
    HIPCHK(hipGetLastError());
    uint32_t f = 0;
    HIPCHK(hipMemcpyAsync(&f, flag, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (f) return fail(ctx, AHIP_E_HIP, "traversal stack capacity exceeded");
    return AHIP_OK;
}

int ahip_occluded_rays(ahip_context* ctx, const float* rays, uint32_t n, uint8_t* occluded) {
    ENTER();
    if (!ctx->prepared) return fail(ctx, AHIP_E_NOT_PREPARED, "prepare first");
    if (n == 0) return AHIP_OK;
    if (!rays || !occluded) return fail(ctx, AHIP_E_INVALID_ARG, "null arrays");
    int st = checkInjected(ctx);
    if (st) return st;
    DevBuf dr, dout, dflag;
    HIPCHK(dr.alloc((size_t)n * 32));
    HIPCHK(dout.alloc(n));
    HIPCHK(dflag.alloc(4));
    HIPCHK(hipMemcpyAsync(dr.p, rays, (size_t)n * 32, hipMemcpyHostToDevice, ctx->stream));
    launchFillU32(dflag.as<uint32_t>(), 1, 0u, ctx->stream);
    launchTraceBatch(ctx->dsc, dr.as<float>(), n, nullptr, dout.as<uint8_t>(), true, ctx->qOvfRef.as<uint32_t>(),
                     ctx->qOvfT.as<float>(), ctx->stackCap, ctx->traceBlocks, dflag.as<uint32_t>(), nullptr,
                     ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(occluded, dout.p, n, hipMemcpyDeviceToHost, ctx->stream));
    uint32_t flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, dflag.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (flag) return fail(ctx, AHIP_E_HIP, "traversal stack capacity exceeded");
    return AHIP_OK;
}

int ahip_halton_eval(ahip_context* ctx, uint64_t seed, const uint32_t* px, const uint32_t* py, const uint32_t* idx,
                     const uint32_t* dim, uint32_t n, float* out) {
    ENTER();
    if (n == 0) return AHIP_OK;
    if (!px || !py || !idx || !dim || !out) return fail(ctx, AHIP_E_INVALID_ARG, "null arrays");
    for (uint32_t i = 0; i < n; ++i)
        if (dim[i] > 4096u) return fail(ctx, AHIP_E_INVALID_ARG, "dimension beyond 4096");
    DevBuf b;
    HIPCHK(b.alloc((size_t)n * 20));
    uint32_t* d = b.as<uint32_t>();
    HIPCHK(hipMemcpyAsync(d, px, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(d + n, py, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(d + 2 * (size_t)n, idx, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(d + 3 * (size_t)n, dim, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    launchHalton(seed, d, d + n, d + 2 * (size_t)n, d + 3 * (size_t)n, n, reinterpret_cast<float*>(d + 4 * (size_t)n),
                 ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, d + 4 * (size_t)n, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return AHIP_OK;
}

// IMaterial::evaluate / sample of one material on probe rows, on the device (the per-lobe parity
// checks of the shading code against the reference's material classes).
int ahip_material_probe(ahip_context* ctx, const ahip_material* mat, uint32_t flags, uint32_t n, const float* in,
                        float* out) {
    ENTER();
    if (!mat) return fail(ctx, AHIP_E_INVALID_ARG, "null material");
    if (n == 0) return AHIP_OK;
    if (!in || !out) return fail(ctx, AHIP_E_INVALID_ARG, "null arrays");
    if (n > (1u << 24)) return fail(ctx, AHIP_E_INVALID_ARG, "more than 2^24 probe rows");
    DMat D;
    std::string err;
    int st = convertMaterial(*mat, 0u, D, err);
    if (st) return fail(ctx, st, err);
    DevBuf b;
    const size_t inB = (size_t)n * 19u * sizeof(float), outB = (size_t)n * 16u * sizeof(float);
    HIPCHK(b.alloc(inB + outB));
    float* din = b.as<float>();
    float* dout = din + (size_t)n * 19u;
    HIPCHK(hipMemcpyAsync(din, in, inB, hipMemcpyHostToDevice, ctx->stream));
    launchMaterialProbe(D, (flags & AHIP_FLAG_HAIR_MARSCHNER) != 0u, n, din, dout, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dout, outB, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return AHIP_OK;
}

// ----------------------------------------------------------------- detached replay / adjoint
int ahip_set_eval_params(ahip_context* ctx, const float* tri_albedo, uint32_t texture, const float* texels) {
    ENTER();
    if (!ctx->prepared) return fail(ctx, AHIP_E_NOT_PREPARED, "prepare first");
    if (texels && texture >= ctx->texW.size()) return fail(ctx, AHIP_E_INVALID_ARG, "texture index out of range");
    ctx->diffTex = texture;
    int st;
    ctx->albESet = tri_albedo != nullptr;
    if (tri_albedo && (st = upload(ctx, ctx->dAlbE, tri_albedo, 3 * (size_t)ctx->numCommit))) return st;
    ctx->texESet = texels != nullptr;
    if (texels && (st = upload(ctx, ctx->dTexE, texels, (size_t)ctx->texW[texture] * ctx->texH[texture]))) return st;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return AHIP_OK;
}

namespace {

// Shared set-up of ahip_replay / ahip_adjoint: the film of the current camera, samples
// [s0, s0 + ns), one thread per (pixel, sample) in 8x8 pixel tiles.
int replayArgs(ahip_context* ctx, uint32_t s0, uint32_t ns, ReplayArgs& a, uint32_t& blocks) {
    if (!ctx->prepared) return fail(ctx, AHIP_E_NOT_PREPARED, "prepare first");
    if (!ctx->hasCam) return fail(ctx, AHIP_E_INVALID_ARG, "set_camera first");
    if (ctx->dsc.hasCurves) return fail(ctx, AHIP_E_UNSUPPORTED, "replay / adjoint support triangle scenes only");
    if (ctx->dsc.hasTransparent)
        return fail(ctx, AHIP_E_UNSUPPORTED, "replay / adjoint support opaque scenes only (no glass or alpha masks)");
    if (ctx->dsc.hasDirectional)
        return fail(ctx, AHIP_E_UNSUPPORTED, "replay / adjoint support area and dome lights only");
    if (ns == 0) return fail(ctx, AHIP_E_INVALID_ARG, "sample_count must be positive");
    a = ReplayArgs{};
    a.cam = toDCam(ctx->cam);
    a.w = ctx->cam.width;
    a.h = ctx->cam.height;
    a.s0 = s0;
    a.ns = ns;
    a.tilesX = (a.w + 7u) / 8u;
    uint64_t slots = (uint64_t)a.tilesX * ((a.h + 7u) / 8u) * 64u;
    if (slots * ns > 0xffffffffull - 1u) return fail(ctx, AHIP_E_INVALID_ARG, "too many samples for one call");
    a.slots = (uint32_t)slots;
    a.seed = ctx->params.seed;
    a.maxDepth = ctx->params.max_depth;
    a.minDepth = ctx->params.min_depth;
    a.flags = ctx->params.flags;
    a.diffTex = ctx->diffTex < ctx->texW.size() ? ctx->diffTex : ~0u;
    if (a.diffTex != ~0u && (ctx->texW[a.diffTex] > 65535u || ctx->texH[a.diffTex] > 65535u))
        return fail(ctx, AHIP_E_UNSUPPORTED, "differentiable texture wider than 65535 texels");
    a.albE = ctx->albESet ? ctx->dAlbE.as<float>() : nullptr;
    a.texE = ctx->texESet ? ctx->dTexE.as<float>() : nullptr;
    a.maxV = ctx->params.max_depth + 1u;
    // blocks: 48 of 2 waves per CU, oversubscribed like the wavefront kernels (the kernel is held
    // to 168 VGPRs, 3 waves per SIMD, so 6 are resident; measured: the adjoint takes 8.9 ms at 6
    // blocks per CU, 7.5 ms at 48), never more threads than the traversal overflow stacks were
    // sized for
    const uint32_t bs = replay_block_size();
    blocks = std::max(1u, ctx->numCU) * 48u;
    blocks = std::min(blocks, (ctx->traceBlocks * 128u) / bs);
    a.nthreads = blocks * bs;
    a.ovfRef = ctx->qOvfRef.as<uint32_t>();
    a.ovfT = ctx->qOvfT.as<float>();
    a.stackCap = ctx->stackCap;
    HIPCHK(ctx->dRepCtr.alloc(sizeof(uint32_t)));
    a.overflow = ctx->dRepCtr.as<uint32_t>();
    launchFillU32(a.overflow, 1, 0u, ctx->stream);
    return AHIP_OK;
}

int replayOverflowCheck(ahip_context* ctx) {
    uint32_t ovf = 0;
    HIPCHK(hipMemcpy(&ovf, ctx->dRepCtr.p, sizeof(ovf), hipMemcpyDeviceToHost));
    if (ovf) return fail(ctx, AHIP_E_HIP, "traversal stack capacity exceeded");
    return AHIP_OK;
}

}  // namespace

int ahip_replay(ahip_context* ctx, uint32_t sample_start, uint32_t sample_count, float* out) {
    ENTER();
    if (ctx->params.integrator != AHIP_INTEGRATOR_PATH)
        return fail(ctx, AHIP_E_UNSUPPORTED, "ahip_replay runs the path tracer only");
    if (!out) return fail(ctx, AHIP_E_INVALID_ARG, "null output");
    ReplayArgs a;
    uint32_t blocks = 0;
    int st = replayArgs(ctx, sample_start, sample_count, a, blocks);
    if (st) return st;
    if ((st = checkInjected(ctx))) return st;
    const size_t n = (size_t)a.w * a.h;
    HIPCHK(ctx->dLout.alloc(n * a.ns * sizeof(float4)));
    HIPCHK(ctx->dRepAcc.alloc(n * sizeof(float4)));
    HIPCHK(ctx->dRepLsq.alloc(n * sizeof(float)));
    a.Lout = ctx->dLout.as<float4>();
    HIPCHK(hipEventRecord(ctx->renderBeg, ctx->stream));
    launchReplay(ctx->dsc, a, false, blocks, ctx->stream);
    launchReplayReduce(a, ctx->dRepAcc.as<float4>(), ctx->dRepLsq.as<float>(), ctx->stream);
    HIPCHK(hipEventRecord(ctx->renderEnd, ctx->stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->renderBeg, ctx->renderEnd));
    ctx->stats.render_ms = ms;
    if ((st = replayOverflowCheck(ctx))) return st;
    return downloadRegion(ctx, ctx->dRepAcc, ctx->dRepLsq, n, out);
}

int ahip_adjoint(ahip_context* ctx, uint32_t width, uint32_t height, uint32_t sample_start, uint32_t sample_count,
                 const float* dL_dimage, float* d_albedo, float* d_roughness_tex) {
    ENTER();
    if (ctx->params.integrator != AHIP_INTEGRATOR_PATH)
        return fail(ctx, AHIP_E_UNSUPPORTED, "ahip_adjoint runs the path tracer only");
    if (!dL_dimage) return fail(ctx, AHIP_E_INVALID_ARG, "null dL_dimage");
    ReplayArgs a;
    uint32_t blocks = 0;
    int st = replayArgs(ctx, sample_start, sample_count, a, blocks);
    if (st) return st;
    if (width != a.w || height != a.h) return fail(ctx, AHIP_E_INVALID_ARG, "width/height must match the camera film");
    if ((st = checkInjected(ctx))) return st;
    const size_t n = (size_t)a.w * a.h;
    const size_t nAlb = 3 * (size_t)ctx->numCommit;
    const size_t nTex = a.diffTex != ~0u ? (size_t)ctx->texW[a.diffTex] * ctx->texH[a.diffTex] : 0;
    if ((st = upload(ctx, ctx->dDLdI, dL_dimage, 3 * n))) return st;
    HIPCHK(ctx->dGAlb.alloc(std::max<size_t>(nAlb, 1) * sizeof(double)));
    HIPCHK(ctx->dGTex.alloc(std::max<size_t>(nTex, 1) * sizeof(double)));
    HIPCHK(ctx->dRec.alloc((size_t)kRecFields * a.maxV * a.nthreads * sizeof(float)));
    launchZeroF64(ctx->dGAlb.as<double>(), nAlb, ctx->stream);
    launchZeroF64(ctx->dGTex.as<double>(), nTex, ctx->stream);
    a.dLdI = ctx->dDLdI.as<float>();
    a.gAlb = d_albedo && nAlb ? ctx->dGAlb.as<double>() : nullptr;
    a.gTex = d_roughness_tex && nTex ? ctx->dGTex.as<double>() : nullptr;
    a.rec = ctx->dRec.as<float>();
    HIPCHK(hipEventRecord(ctx->renderBeg, ctx->stream));
    launchReplay(ctx->dsc, a, true, blocks, ctx->stream);
    HIPCHK(hipEventRecord(ctx->renderEnd, ctx->stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->renderBeg, ctx->renderEnd));
    ctx->stats.render_ms = ms;
    if ((st = replayOverflowCheck(ctx))) return st;
    std::vector<double> g;
    if (d_albedo && nAlb) {
        g.resize(nAlb);
        HIPCHK(hipMemcpy(g.data(), ctx->dGAlb.p, nAlb * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nAlb; ++i) d_albedo[i] = (float)g[i];
    }
    if (d_roughness_tex && nTex) {
        g.resize(nTex);
        HIPCHK(hipMemcpy(g.data(), ctx->dGTex.p, nTex * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nTex; ++i) d_roughness_tex[i] = (float)g[i];
    }
    return AHIP_OK;
}

int ahip_get_stats(ahip_context* ctx, ahip_stats* out) {
    if (!ctx || !out) return AHIP_E_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    *out = ctx->stats;
    return AHIP_OK;
}

int ahip_reset_stats(ahip_context* ctx) {
    if (!ctx) return AHIP_E_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->stats = ahip_stats{};
    ctx->bounceRays.clear();
    ctx->bounceMs.clear();
    return AHIP_OK;
}

int ahip_get_bounce_stats(ahip_context* ctx, uint32_t capacity, uint64_t* rays, double* ms, uint32_t* num_bounces) {
    if (!ctx || !num_bounces) return AHIP_E_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    const uint32_t nb = (uint32_t)std::max(ctx->bounceRays.size() / 3, ctx->bounceMs.size());
    *num_bounces = nb;
    for (uint32_t b = 0; b < std::min(nb, capacity); ++b) {
        for (uint32_t k = 0; k < 3; ++k)
            if (rays) rays[3 * b + k] = 3 * (size_t)b + k < ctx->bounceRays.size() ? ctx->bounceRays[3 * b + k] : 0ull;
        if (ms) ms[b] = b < ctx->bounceMs.size() ? ctx->bounceMs[b] : 0.0;
    }
    return AHIP_OK;
}

int ahip_count_traffic(ahip_context* ctx, uint32_t width, uint32_t height, uint32_t sample_start,
                       uint32_t sample_count) {
    ENTER();
    if (ctx->params.integrator != AHIP_INTEGRATOR_PATH)
        return fail(ctx, AHIP_E_UNSUPPORTED, "ahip_count_traffic runs the path tracer only");
    if (!ctx->prepared) return fail(ctx, AHIP_E_NOT_PREPARED, "prepare first");
    if (ctx->dsc.hasTransparent)
        return fail(ctx, AHIP_E_UNSUPPORTED, "ahip_count_traffic counts opaque scenes only (no glass or alpha masks)");
    if (!ctx->hasCam || width != ctx->cam.width || height != ctx->cam.height)
        return fail(ctx, AHIP_E_INVALID_ARG, "film size differs from the camera's");
    int st = ensureQueues(ctx, 64);
    if (st) return st;
    DevBuf acc, lsq, sums;
    HIPCHK(acc.alloc((size_t)width * height * sizeof(float4)));
    HIPCHK(lsq.alloc((size_t)width * height * sizeof(float)));
    HIPCHK(sums.alloc(2 * sizeof(unsigned long long)));
    launchFillF32(acc.as<float>(), (size_t)width * height * 4, 0.f, ctx->stream);
    launchFillF32(lsq.as<float>(), (size_t)width * height, 0.f, ctx->stream);
    launchFillU32(sums.as<uint32_t>(), 4, 0u, ctx->stream);
    ahip_stats saved = ctx->stats;
    std::vector<ViewJob> jobs{
        ViewJob{ctx->cam, 0, 0, width, height, sample_start, sample_count, acc.as<float4>(), lsq.as<float>()}};
    st = renderCall(ctx, jobs, false, sums.as<unsigned long long>());
    unsigned long long s[2] = {0, 0};
    if (!st) {
        hipError_t e = hipMemcpy(s, sums.p, sizeof(s), hipMemcpyDeviceToHost);
        if (e != hipSuccess) st = fail(ctx, AHIP_E_HIP, hipGetErrorString(e));
    }
    ahip_stats counted = ctx->stats;
    ctx->stats = saved;
    ctx->stats.nodes_visited = s[0];
    ctx->stats.prims_tested = s[1];
    ctx->stats.rays_closest = counted.rays_closest - saved.rays_closest;   // the counted frame only
    ctx->stats.rays_shadow = counted.rays_shadow - saved.rays_shadow;
    return st;
}

int ahip_fail_next_launch(ahip_context* ctx) {
    if (!ctx) return AHIP_E_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->failNext = true;
    return AHIP_OK;
}

int ahip_camera_pinhole(const float from[3], const float at[3], const float up[3], float vfov_degrees, uint32_t width,
                        uint32_t height, ahip_camera* out) {
    if (!from || !at || !up || !out || width == 0 || height == 0) return AHIP_E_INVALID_ARG;
    return makePinholeCamera(from, at, up, vfov_degrees, width, height, out);
}

}  // extern "C"
