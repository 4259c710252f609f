// Host-side launchers of the light-transport kernels (implemented in kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "dev_types.h"

namespace vlra {

// Depths rendered as separate per-depth kernels before the paths move to the tail kernel.
// On the 1080p icosphere grid (2.07 M, 1.14 M, 0.24 M, 0.11 M, 48 k ... rays at depths 0, 1, 2,
// 3, 4 ...) a frame takes 14.7 ms with every depth as its own kernels, 9.8 ms with 3 or 4.
constexpr uint32_t kDefaultWavefrontDepth = 3;

struct FrameParams {
    uint32_t width, height;        // buffer size (stride = width)
    uint32_t imageW, imageH;       // rendered sub-image (shrinkCoeff)
    uint32_t frameIndex;           // numAccumFrames (1-based)
    uint32_t maxPathLength;
    uint32_t traversalStats;       // count nodes / triangles / instances per traversal
    uint32_t wavefrontDepth;       // depths run as separate kernels; the rest in the tail kernel
};

struct FrameBuffers {
    PathState ps;
    RayQueue rays[2];
    HitBuffer hits;
    ShadowQueue shadow;
    uint32_t* counters;            // kNumCounters
    float* accum;                  // W*H*channels*2 (sum, comp)
    float4* output;                // W*H
};

// timing hook: called before/after each launch (may be null)
struct LaunchTimer {
    virtual void begin(int stage) = 0;
    virtual void end(int stage) = 0;
    virtual ~LaunchTimer() {}
};
enum Stage { STAGE_GENERATE = 0, STAGE_INTERSECT, STAGE_SHADE, STAGE_SHADOW, STAGE_ACCUMULATE, STAGE_CONVERT, STAGE_TAIL, NUM_STAGES };

hipError_t launchFrame(const DevScene& scene, const FrameParams& fp, const FrameBuffers& fb, bool spectral,
                       hipStream_t stream, LaunchTimer* timer);

hipError_t launchTraceRays(const DevScene& scene, const void* rays, uint32_t numRays, void* hits, bool anyHit,
                           hipStream_t stream);

hipError_t launchBsdfQueries(const DevScene& scene, const DevMaterial& material, bool spectral, const void* queries,
                             uint32_t n, void* results, hipStream_t stream);

} // namespace vlra
