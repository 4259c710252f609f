// Host-side launchers of the light-transport kernels (implemented in kernels.hip).
#pragma once
#include <algorithm>
#include <vector>
#include <hip/hip_runtime.h>
#include "dev_types.h"
#include "radiance_cache.h"

namespace vlra {

// Depths rendered as separate per-depth kernels before the paths move to the tail kernel
// (0: the tail kernel runs every depth, primary rays included).  On the 1080p spectral
// icosphere grid (2.07 M, 1.14 M, 0.24 M, 0.11 M, 48 k ... rays at depths 0, 1, 2, 3, 4 ...)
// a frame takes 3.88 ms with every depth in the tail kernel, 4.35 ms with depth 0 as its own
// kernels and 4.88 ms with depths 0 and 1: the tail kernel, which refills its lanes per path,
// carries the whole frame best.  The denoiser guides need depth 0 as its own kernels.
constexpr uint32_t kDefaultWavefrontDepth = 0;

struct FrameParams {
    uint32_t width, height;        // buffer size (stride = width)
    uint32_t imageW, imageH;       // rendered sub-image (shrinkCoeff)
    uint32_t frameIndex;           // numAccumFrames (1-based)
    uint32_t maxPathLength;
    uint32_t traversalStats;       // count nodes / triangles / instances per traversal
    uint32_t wavefrontDepth;       // depths run as separate kernels; the rest in the tail kernel
    uint32_t tailBlocks;           // persistent grid of the tail kernel (blocks of 128)
    // radiance cache
    uint32_t cacheMode;            // 0 off; 1 record training vertices only; 2 also end paths in the cache
    uint32_t frameStamp;           // grows every frame: marks the cache records written this frame
    float cacheThreshold;          // c of the spread test a(x_1..x_n) > c a_0
    uint32_t cacheFullRadiance;    // the cache query replaces the last vertex's next-event estimate too
    uint32_t trainStride, trainOffset;   // pixel i traces a training path when i % stride == offset
    // the frame's paths are pixels pixelBegin .. pixelBegin + pixelCount - 1 of the image
    // (row-major over imageW); a frame may be rendered as several such parts
    uint32_t pixelBegin, pixelCount;
    uint32_t sortShade;            // shade each per-depth queue in order of shading class
    uint32_t guideReset;           // the denoiser guides restart with this frame
};

// Shading classes of the material sort: 0 = miss, 1 + material type for a hit.
constexpr uint32_t kSortClasses = 16;

constexpr uint32_t kCacheTrainVerts = 3;
// frames of at least this many pixels are rendered as two halves on two streams
constexpr uint32_t kSplitMinPixels = 1u << 18;   // vertices 1..3 of a training path become samples

// Radiance-cache records written by the path tracer (indexed by path, or by training slot =
// training path * kCacheTrainVerts + vertex - 1); a record is live when its stamp equals the
// frame's stamp.
struct CacheBuffers {
    float4* queryRaw;              // 4 float4 per path: the raw input record
    float4* queryThroughput;       // path throughput at the query vertex
    uint32_t* queryStamp;
    float2* spread;                // (a_0, a) per path
    float4* trainRaw;              // 4 float4 per slot
    float4* trainThroughput;       // throughput on arrival at the vertex
    float4* trainContrib;          // path contribution after the vertex's next-event estimate
    uint32_t* trainStamp;          // the vertex was reached this frame
    uint32_t* trainAfterStamp;     // the path went on to a next vertex (trainContrib is set)
};

constexpr uint32_t kDefaultTailBlocks = 2048;   // 8 blocks of 128 per CU

// Frame-schedule autotuner.  The image does not depend on the schedule, so candidates are
// timed on real frames: first the wavefront depth (tail grid fixed), then the tail grid at
// the best depth, then the material sort off and on (not at depth 0, which has no per-depth
// queues to sort); each candidate gets
// kFramesPerCandidate frames and keeps its fastest.
struct ScheduleTuner {
    static constexpr uint32_t kFramesPerCandidate = 3;
    struct Cand {
        uint32_t wavefrontDepth, tailBlocks, sortShade;
        float bestMs;
    };
    std::vector<Cand> cands;
    uint32_t phase = 0;           // 0: depths, 1: tail grids, 2: material sort off / on, 3: done
    uint32_t index = 0, frames = 0;
    uint64_t key = 0;
    bool done() const { return phase >= 3; }
    void start(uint64_t k) {
        key = k;
        phase = 0;
        index = frames = 0;
        cands.clear();
        for (uint32_t d : {0u, 1u, 2u, 3u, 4u}) cands.push_back({d, kDefaultTailBlocks, 0u, 1e30f});
    }
    const Cand &current() const { return cands[index]; }
    Cand best() const {
        Cand b = cands[0];
        for (const Cand &c : cands)
            if (c.bestMs < b.bestMs) b = c;
        return b;
    }
    void report(float ms) {
        Cand &c = cands[index];
        c.bestMs = std::min(c.bestMs, ms);
        if (++frames < kFramesPerCandidate) return;
        frames = 0;
        if (++index < cands.size()) return;
        Cand b = best();
        index = 0;
        if (phase == 0) {
            phase = 1;
            cands.clear();
            for (uint32_t t : {512u, 1024u, 2048u, 4096u}) cands.push_back({b.wavefrontDepth, t, 0u, 1e30f});
        } else if (phase == 1 && b.wavefrontDepth == 0) {
            // no per-depth queues to sort: the material sort has nothing to change
            phase = 3;
            cands.assign(1, b);
        } else if (phase == 1) {
            phase = 2;
            cands.clear();
            for (uint32_t so : {0u, 1u}) cands.push_back({b.wavefrontDepth, b.tailBlocks, so, 1e30f});
        } else {
            phase = 3;
            cands.assign(1, b);
        }
    }
};

struct FrameBuffers {
    PathState ps;
    RayQueue rays[2];
    HitBuffer hits;
    ShadowQueue shadow;
    uint32_t* counters;            // kNumCounters
    // material sort of the per-depth queues: class per queue entry, the entries in class
    // order, and per depth kSortClasses class counts then kSortClasses cursors
    uint32_t* sortKey;
    uint32_t* sortPerm;
    uint32_t* sortCounts;          // kMaxDepthCounters * 2 * kSortClasses
    float* accum;                  // W*H*channels*2 (sum, comp)
    float4* output;                // W*H
    CacheBuffers cache;
    // denoiser guides (null when off): the first hit's albedo binned like the colour (3 or 16
    // planes of W*H floats) and its shading normal (3 planes), summed over frames
    float* guideAlbedo;
    float* guideNormal;
};

// timing hook: called before/after each launch (may be null)
struct LaunchTimer {
    virtual void begin(int stage) = 0;
    virtual void end(int stage) = 0;
    virtual ~LaunchTimer() {}
};
enum Stage {
    STAGE_GENERATE = 0, STAGE_INTERSECT, STAGE_SHADE, STAGE_SHADOW, STAGE_ACCUMULATE, STAGE_CONVERT, STAGE_TAIL,
    STAGE_LIGHT_PATHS, STAGE_EYE_PATHS, STAGE_CACHE_QUERY, STAGE_CACHE_TRAIN, NUM_STAGES
};

// LVC-BPT frame state (lvc_bpt.cu:144-971).  Light paths write their vertices to fixed slots
// (slotsPerPath per path); a prefix sum of the per-path counts compacts them into `cache` in
// light-path order, so the cache does not depend on thread scheduling.
constexpr uint32_t kLightVertexFloat4s = 5;     // 80 bytes per cached vertex
constexpr uint32_t kMaxLightPaths = 1024u * 1024u;   // two-level scan of 1024-wide blocks
struct BptParams {
    float wlOffset;              // common wavelengths of the frame (context.cpp:1464-1467)
    uint32_t wlSel;
    float wlProb;                // their probability density
    uint32_t numLightPaths;
    uint32_t slotsPerPath;       // maxPathLength + 1
    float4* slots;               // numLightPaths * slotsPerPath vertices
    uint32_t* slotCounts;        // vertices written by each light path
    uint32_t* slotOffsets;       // exclusive prefix sum inside each 1024-path block
    uint32_t* blockSums;         // per-block totals, then their exclusive prefix sum
    float4* cache;               // compacted vertices
    uint32_t* numVertices;
    uint64_t* lightRng;          // PCG32 state per light path
    float* splat;                // lens-connection splats: channel planes of width * height
};

// Radiance-cache work of a frame: the network, this step's optimiser settings and the batch
// buffers the training samples are compacted into.
struct CacheLaunch {
    RcDevice net;
    RcTrainParams train;
    uint32_t numSlots;             // training slots this frame (training paths * kCacheTrainVerts)
    float4* batchRaw;              // kMaxBatch records of 4 float4
    float4* batchTarget;
    uint32_t* batchCount;
    uint32_t* slotFlags;           // per slot: a live sample
    uint32_t* slotOffsets;         // exclusive scan inside blocks of 1024 slots
    uint32_t* slotBlockSums;       // per block total, then its exclusive scan (<= 64 blocks)
    float4* slotTargets;
};

// Generate, trace, shade and accumulate the pixels of fp's part (and convert the whole image
// to RGBA when `convert`).
// `afterWavefront` (optional) is recorded once the per-depth kernels are enqueued, before the
// tail kernel.
hipError_t launchFrame(const DevScene& scene, const FrameParams& fp, const FrameBuffers& fb, bool spectral,
                       hipStream_t stream, LaunchTimer* timer, const CacheLaunch* cache = nullptr,
                       bool convert = true, hipEvent_t afterWavefront = nullptr);
hipError_t launchConvert(const DevScene& scene, const FrameParams& fp, const FrameBuffers& fb, bool spectral,
                         hipStream_t stream, LaunchTimer* timer);

// A batch of render jobs over one set of path buffers (paths of job j are the pixels
// jobOffsets[j] .. jobOffsets[j + 1] - 1).  fp.imageW is the total pixel count (imageH = 1),
// fp.maxPathLength the largest of the jobs', fp.wavefrontDepth / tailBlocks the schedule.
struct DevJob {
    DevScene scene;                // with the job's camera
    float* accum;                  // Kahan planes of width * height floats
    float4* output;
    uint32_t pixelOffset, width, height, spp, maxPathLength;
    uint32_t frameIndex;           // samples started (advanced on the device at each step)
    uint32_t active;               // the job renders in the current step
    uint32_t pad;
};
struct BatchLaunch {
    DevScene anyScene;             // shared tables (colour matching) for the kernels' fixed arguments
    FrameParams fp;
    FrameBuffers fb;
    DevJob* jobs;
    const uint32_t* jobOffsets;
    uint32_t* activeOffsets;       // numJobs + 1, written on the device at each step
    uint32_t numJobs;
    bool nodePrograms;             // some job's scene has node programs: the kernels that run them
};
hipError_t launchBatchStep(const BatchLaunch& b, bool spectral, hipStream_t stream);
hipError_t launchBatchConvert(const BatchLaunch& b, bool spectral, hipStream_t stream);

// One LVC-BPT frame: light paths, cache compaction, eye paths, accumulation, conversion.
hipError_t launchBptFrame(const DevScene& scene, const FrameParams& fp, const FrameBuffers& fb, const BptParams& bp,
                          bool spectral, hipStream_t stream, LaunchTimer* timer);
hipError_t launchLightTraceFrame(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                 const BptParams &bp, bool spectral, hipStream_t s, LaunchTimer *timer);

hipError_t launchTraceRays(const DevScene& scene, const void* rays, uint32_t numRays, void* hits, bool anyHit,
                           hipStream_t stream);

// Node program (base, count) at device-resident surface points (20 floats each), 4 floats out
// per point: the spectrum (Spectrum plugs) or the value in the plug's own type.
hipError_t launchEvalNodePlug(const DevScene& scene, uint32_t base, uint32_t count, uint32_t plugType, bool spectral,
                              const float* points, uint32_t n, float* out, hipStream_t stream);
hipError_t launchBsdfQueries(const DevScene& scene, const DevMaterial& material, bool spectral, const void* queries,
                             uint32_t n, void* results, hipStream_t stream);

} // namespace vlra
