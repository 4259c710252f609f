/* Extensions of the VLR C API on MI355X.
 *
 * None of these has a counterpart in the reference's vlr.h.  They are of three kinds:
 *   1. runtime knobs for values the reference fixes at compile time, with defaults that
 *      reproduce the reference (MaxPathLength = 25, path_tracing.cu:78; RGB rendering,
 *      common_internal.h:23; 500,000 light paths, context.h:303);
 *   2. progressive-render state access (per-pixel PCG32 states and the Kahan accumulation
 *      buffer, context.cpp:1167-1190), which makes resume-after-checkpoint bitwise equal to
 *      an uninterrupted render;
 *   3. direct entry points into the hot kernels for parity tests and benchmarks: the BVH
 *      traversal (replaces optixTrace, path_tracing.cu:89-93 and
 *      light_transport_common.h:169-172), the BVH builders (replace the OptiX GAS / IAS
 *      builds, scene.cpp:1149-1202, 1319-1329) and the flattened BSDF code (replaces the
 *      BSDF direct callables, materials.cu:282-328).
 */
#pragma once

#include <stdint.h>
#include <stddef.h>

#if defined(__cplusplus)
extern "C" {
#endif

/* ---- knobs ---- */
VLR_API VLRResult vlrContextSetMaxPathLength(VLRContext context, uint32_t maxPathLength);
VLR_API VLRResult vlrContextGetMaxPathLength(VLRContextConst context, uint32_t* maxPathLength);
/* Spectral (4 hero wavelengths, Meng upsampling) vs RGB transport.  Takes effect at the
 * next render with firstFrame = true. */
VLR_API VLRResult vlrContextSetSpectralRendering(VLRContext context, bool enable);
VLR_API VLRResult vlrContextGetSpectralRendering(VLRContextConst context, bool* enabled);

/* ---- statistics ---- */
typedef struct VLRRenderStats {
    uint64_t numFrames;          /* frames rendered since the last reset */
    uint64_t paths;              /* camera paths started */
    uint64_t closestRays;        /* extension (closest-hit) rays traced */
    uint64_t shadowRays;         /* shadow (any-hit) rays traced */
    uint64_t droppedNonFinite;   /* samples dropped for a non-finite contribution */
    /* device time accumulated with HIP events when kernel timing is enabled, in ms */
    double msGenerate, msIntersect, msShade, msShadow, msAccumulate, msConvert;
    uint64_t launchesIntersect, launchesShadow, launchesShade;
    /* traversal work, counted only while traversal statistics are enabled: inner BVH nodes
     * fetched, triangles tested and instances entered, for closest-hit and shadow rays */
    uint64_t nodesClosest, trianglesClosest, instancesClosest;
    uint64_t nodesShadow, trianglesShadow, instancesShadow;
    double msTail;               /* the persistent tail kernel (see vlrContextSetWavefrontDepth) */
    uint64_t launchesTail;
    /* rays traced by the per-depth kernels; closestRays - closestRaysWavefront went through the tail */
    uint64_t closestRaysWavefront, shadowRaysWavefront;
    /* VLRRenderer_BPT: light-path pass (tracing and cache compaction) and eye-path pass */
    double msLightPaths, msEyePaths;
    /* radiance cache: queries (inference) and training (sample compaction + one Adam step) */
    double msCacheQuery, msCacheTrain;
} VLRRenderStats;
/* Counting the rays reads back one small counter buffer per render call. */
VLR_API VLRResult vlrContextGetRenderStats(VLRContextConst context, VLRRenderStats* stats);
VLR_API VLRResult vlrContextResetRenderStats(VLRContext context);
/* Brackets every kernel launch of vlrContextRender with HIP events on the stream it runs on.
 * The events are read back when the statistics are read (or once 16384 are pending), so timed
 * frames are not synchronised.  Off by default. */
VLR_API VLRResult vlrContextSetKernelTiming(VLRContext context, bool enable);

/* ---- denoiser guides (path_tracing.cu:95-108, lvc_bpt.cu:528-540, aux_buffer_generator.cu:22-93,
 * post_process.cu:70-90) ----
 * With the guides on, every renderer sums each sample's first-hit albedo (the BSDF's base colour
 * times the D65 white point over the wavelength pdf; zero for a miss) and shading normal per
 * pixel, restarting with the accumulation: VLRRenderer_PathTracing and VLRRenderer_BPT at their
 * eye paths' first hits, VLRRenderer_LightTracing with one primary ray per pixel before each
 * frame's light paths, drawn from the pixel's generator as the reference's aux-buffer pass
 * draws it (with the guides off that pass is skipped, and the pixel generators stay as they
 * are; the reference runs it on every light-tracing frame).  Reading converts them as the
 * reference's copyBuffers does: linear Rec.709 albedo averaged over the samples, and the
 * normal sum rotated into the camera's frame with x negated and normalised (zero stays zero);
 * width*height RGBA float4 each, alpha 1 (either pointer may be NULL).  The guides are what
 * the reference hands its denoiser; the denoiser itself (OptiX) is not part of this library,
 * and vlrContextRender's `denoise` flag is accepted and ignored.  Off by default. */
VLR_API VLRResult vlrContextSetDenoiserAuxBuffers(VLRContext context, bool enable);
VLR_API VLRResult vlrContextReadDenoiserAuxBuffers(VLRContextConst context, float* albedoRGBA, float* normalRGBA);
/* Renders with the traversal kernels' counting variant (slower; off by default). */
VLR_API VLRResult vlrContextSetTraversalStatistics(VLRContext context, bool enable);
/* Frame schedule: depths below `depth` run as separate intersect / shade / shadow kernels over
 * device queues; from `depth` on every remaining path runs to its end inside one persistent
 * kernel; 0 (the default) runs every depth, primary rays included, in that kernel.  With the
 * denoiser guides on, depth 0 always runs as its own kernels.  The image does not depend on it
 * (paths draw and add in the same order). */
VLR_API VLRResult vlrContextSetWavefrontDepth(VLRContext context, uint32_t depth);
VLR_API VLRResult vlrContextGetWavefrontDepth(VLRContextConst context, uint32_t* depth);
/* Schedule autotuning (on by default; setting the wavefront depth turns it off): the first
 * frames after a change of scene, size or mode time candidate schedules (wavefront depth,
 * then the tail kernel's grid) and keep the fastest.  Tuning frames synchronise the stream. */
VLR_API VLRResult vlrContextSetScheduleAutotuning(VLRContext context, bool enable);
/* Frames of at least 2^18 pixels rendered as two pixel halves on two streams, each half
 * running its own frames, so that the deep paths of one half overlap the per-depth kernels of
 * the other (off by default; never with the radiance cache).  The image does not depend on it. */
VLR_API VLRResult vlrContextSetSplitFrames(VLRContext context, bool enable);
/* Material sort of the per-depth shade queues: each depth's hits are shaded in order of
 * material type (misses first), so a wave runs one material's code at a time.  The image does
 * not change.  Setting it fixes the schedule (turns autotuning off); the autotuner otherwise
 * times sorting on and off.  Get returns what the last frame used. */
VLR_API VLRResult vlrContextSetShadeSorting(VLRContext context, bool enable);
VLR_API VLRResult vlrContextGetShadeSorting(VLRContextConst context, bool* enabled);
/* The schedule of the last rendered frame and whether tuning has finished. */
VLR_API VLRResult vlrContextGetSchedule(VLRContextConst context, uint32_t* wavefrontDepth, uint32_t* tailBlocks,
                                        bool* tuned);
/* Queue sizes of the last rendered frame: closest[d] extension rays and shadow[d] shadow rays
 * traced at depth d, for d < maxDepth (entries past the frame's path length are 0). */
VLR_API VLRResult vlrContextGetQueueCounts(VLRContextConst context, uint32_t* closest, uint32_t* shadow,
                                           uint32_t maxDepth);

/* Light paths per frame of VLRRenderer_BPT (OptiX::numLightPaths = 500,000, context.h:303),
 * at most 1,048,576.  Changing it re-seeds the light-path RNG states. */
VLR_API VLRResult vlrContextSetNumLightPaths(VLRContext context, uint32_t numLightPaths);
VLR_API VLRResult vlrContextGetNumLightPaths(VLRContextConst context, uint32_t* numLightPaths);

/* ---- batches of render jobs ----
 * Renders many jobs (scene, camera, image size, samples, max path length) together: each step
 * renders one sample per pixel of every job that has samples left, over one set of path
 * buffers, with the step's launches captured once into a HIP graph and replayed.  Each job's
 * pixels are seeded like a single render, so its image is bit-identical to rendering it alone
 * with vlrContextRender (samples from firstFrame, VLRRenderer_PathTracing, no radiance cache).
 * The spectral/RGB mode is the context's.  With autotuning on (the default), the first steps
 * of a batch of 8 or more samples time candidate wavefront depths and keep the fastest.  The
 * call returns once the batch has finished; completionMs (host array of numJobs, may be
 * NULL) receives each job's completion time from the start of the batch. */
typedef struct VLRBatchJob {
    VLRSceneConst scene;
    VLRCameraConst camera;
    uint32_t width, height;
    uint32_t numSamples;         /* >= 1 */
    uint32_t maxPathLength;      /* 1 .. 63 */
} VLRBatchJob;
VLR_API VLRResult vlrContextRenderBatch(VLRContext context, hipStream_t stream, const VLRBatchJob* jobs,
                                        uint32_t numJobs, float* completionMs);
/* RGBA32F image of job `job` of the last batch (width * height float4, row-major). */
VLR_API VLRResult vlrContextReadBatchOutput(VLRContextConst context, uint32_t job, float* rgba);
/* Replay captured graphs (default) or launch every kernel of every step directly. */
VLR_API VLRResult vlrContextSetBatchGraphs(VLRContext context, bool enable);
/* The wavefront depth and tail grid the last batch ran with. */
VLR_API VLRResult vlrContextGetBatchSchedule(VLRContextConst context, uint32_t* wavefrontDepth, uint32_t* tailBlocks);

/* ---- progressive state (host pointers, width*height entries, row-major) ---- */
VLR_API VLRResult vlrContextReadRNGStates(VLRContextConst context, uint64_t* states);
VLR_API VLRResult vlrContextWriteRNGStates(VLRContext context, const uint64_t* states);
/* PCG32 states of the BPT light paths (numLightPaths entries; linearRngBuffer,
 * context.cpp:618-626 of the reference) */
VLR_API VLRResult vlrContextReadLightRNGStates(VLRContextConst context, uint64_t* states);
VLR_API VLRResult vlrContextWriteLightRNGStates(VLRContext context, const uint64_t* states);
/* Kahan accumulation pairs: (sum, compensation) per channel, channels = 3 (RGB) or 16
 * (spectral strata) as the context's requested mode (vlrContextSetSpectral) gives them;
 * numFloats must equal width*height*channels*2, or the call returns InvalidArgument.  Reading
 * fails with InvalidArgument while the context holds the accumulation of the other mode (the
 * mode was changed and no first frame has been rendered since), so a checkpoint of one mode
 * is never read as the other. */
VLR_API VLRResult vlrContextGetAccumulationChannels(VLRContextConst context, uint32_t* channels);
VLR_API VLRResult vlrContextReadAccumulation(VLRContextConst context, float* data, size_t numFloats);
VLR_API VLRResult vlrContextWriteAccumulation(VLRContext context, const float* data, size_t numFloats,
                                              uint32_t numAccumFrames);
/* The per-frame wavelength generator of VLRRenderer_BPT (a default-seeded std::mt19937 with
 * two draws per frame, context.cpp:1464-1467 of the reference): the number of frames that
 * drew from it since the context was created.  Setting it restarts the generator and replays
 * that many frames' draws, so a resumed BPT render draws the wavelengths of the original. */
VLR_API VLRResult vlrContextGetBptWavelengthFrames(VLRContextConst context, uint64_t* frames);
VLR_API VLRResult vlrContextSetBptWavelengthFrames(VLRContext context, uint64_t frames);

/* ---- radiance cache ----
 * A 5 x 64 ReLU MLP (BF16 weights and activations, FP32 accumulation) trained online on the
 * path tracer's own vertices and queried to end paths once their spread passes a threshold
 * (Mueller et al. 2021, "Real-time Neural Radiance Caching for Path Tracing").  It has no
 * counterpart in the reference; the path tracer without it is the reference's.  Each frame
 * a set of training paths (one pixel in `stride`, a different offset every frame) is traced
 * to the end; vertices 1..3 of each become samples whose target is the radiance the path
 * gathered after that vertex divided by the throughput on arrival; the other paths end in
 * the cache at the first vertex after a bounce whose spread exceeds threshold * a_0.  The
 * first `warmupSteps` frames after a reset only train. */
typedef struct VLRRadianceCacheSettings {
    float learningRate;          /* Adam step size, 1e-2 */
    float beta1, beta2, epsilon; /* 0.9, 0.99, 1e-15 */
    uint32_t trainingPaths;      /* training paths per frame, 1 .. 21845 (default 21845) */
    float terminationThreshold;  /* spread threshold c, 0.01 */
    uint32_t warmupSteps;        /* training-only frames after a reset, 16 */
    uint32_t seed;               /* weight initialisation (Glorot uniform from PCG32) */
    uint32_t fullRadiance;       /* 1 (default): the cache supplies all the radiance the path's
                                  * last vertex reflects, its next-event estimate included, and
                                  * samples are what a training path gathers after a vertex's
                                  * emission; 0: the last vertex keeps its next-event estimate and
                                  * the cache supplies what a continued path would add after it */
} VLRRadianceCacheSettings;
VLR_API VLRResult vlrContextSetRadianceCache(VLRContext context, bool enable);
VLR_API VLRResult vlrContextGetRadianceCache(VLRContextConst context, bool* enabled);
VLR_API VLRResult vlrContextSetRadianceCacheSettings(VLRContext context, const VLRRadianceCacheSettings* settings);
VLR_API VLRResult vlrContextGetRadianceCacheSettings(VLRContextConst context, VLRRadianceCacheSettings* settings);
/* Re-initialises the weights from the seed and clears the optimiser state. */
VLR_API VLRResult vlrContextResetRadianceCache(VLRContext context);
/* FP32 master parameters (host arrays of vlrRadianceCacheNumParameters floats): per layer
 * l = 0..5 the row-major weights [rows_l][cols_l] (rows 64, the output layer padded to 16 rows
 * of which 4 are used; cols 96 for l = 0, else 64), then the biases (5 x 64, then 16).
 * Writing rounds them to BF16 for the network and clears the optimiser state. */
VLR_API VLRResult vlrRadianceCacheNumParameters(uint32_t* numParameters);
VLR_API VLRResult vlrContextReadRadianceCacheWeights(VLRContextConst context, float* parameters);
VLR_API VLRResult vlrContextWriteRadianceCacheWeights(VLRContext context, const float* parameters);
/* Gradient of the loss of the last training step (same layout), its mean loss (relative L2,
 * (y - t)^2 / (y^2 + 0.01) averaged over samples and outputs), and the number of steps taken. */
VLR_API VLRResult vlrContextReadRadianceCacheGradient(VLRContextConst context, float* gradient);
VLR_API VLRResult vlrContextGetRadianceCacheLoss(VLRContextConst context, float* loss, uint32_t* steps);
/* Direct entry points (device-resident arrays): records are 16 floats each (position in the
 * scene's bounding cube [0,1]^3, outgoing direction theta/pi and phi/(2 pi), shading normal
 * (n+1)/2, roughness, diffuse RGB, specular RGB, hero-wavelength offset), outputs and targets
 * 4 floats each.  Training takes one Adam step on the first numOutputs outputs. */
VLR_API VLRResult vlrContextRadianceCacheInfer(VLRContext context, hipStream_t stream, const float* records,
                                               uint32_t numRecords, float* outputs);
VLR_API VLRResult vlrContextRadianceCacheTrain(VLRContext context, hipStream_t stream, const float* records,
                                               const float* targets, uint32_t numRecords, uint32_t numOutputs);

/* ---- traversal entry point ---- */
typedef struct VLRRay { float org[3]; float tmin; float dir[3]; float tmax; } VLRRay;
typedef struct VLRHit { int32_t instIndex; int32_t geomInstIndex; int32_t primIndex; float t, b1, b2; } VLRHit;
/* Sets up `scene` on the device if needed, then traces numRays device-resident rays.
 * anyHit = false: closest hit (instIndex = -1 on a miss).  anyHit = true: occlusion test
 * in [tmin, tmax); hits[i].instIndex >= 0 when blocked.  Indices: instances in
 * depth-first child order of parent nodes that hold surface nodes, geometry instances
 * (material groups) in order of first appearance. */
VLR_API VLRResult vlrContextTraceRays(VLRContext context, hipStream_t stream, VLRScene scene, const VLRRay* rays,
                                      uint32_t numRays, VLRHit* hits, bool anyHit);

/* ---- acceleration structures ----
 * Replace the OptiX GAS / IAS builds (Scene::setup, scene.cpp:1149-1202 and 1319-1329,
 * PreferFastTrace at scene.cpp:1049): one BLAS per distinct geometry group, a TLAS over the
 * instances, 64-byte quantised 4-wide nodes.  The device builder (default) and the host
 * builder produce the same trees up to node numbering. */
#define VLR_ACCEL_BUILDER_DEVICE 0u
#define VLR_ACCEL_BUILDER_HOST 1u
VLR_API VLRResult vlrContextSetAccelBuilder(VLRContext context, uint32_t builder);
VLR_API VLRResult vlrContextGetAccelBuilder(VLRContextConst context, uint32_t* builder);
typedef struct VLRAccelStats {
    uint32_t numBlas, numTlasInstances;
    uint64_t numTriangles;
    uint32_t numBlasNodes, numTlasNodes;
    float msBlas, msTlas;        /* wall time of the last scene setup's builds */
    uint32_t blasReused;         /* 1: no geometry changed since the previous setup, BLASes kept */
    uint32_t numNodeProgramSteps; /* shader-node program steps of the set-up scene; 0: the
                                     kernels without node evaluation run */
    uint32_t tlasDepth, blasDepth; /* levels of 4-wide nodes of the TLAS and of the deepest BLAS;
                                      a scene whose 3 (tlasDepth + blasDepth) + 1 exceeds the
                                      traversal stack (80 entries) is refused at set-up */
    uint32_t tlasRefit;          /* 1: the last set-up refitted the TLAS instead of rebuilding it */
} VLRAccelStats;
/* With refit on (the default), a set-up after which only instance transforms changed (the same
 * instances over the same geometry) keeps the TLAS's tree and recomputes its boxes; it rebuilds
 * instead once the refitted tree's cost (the summed areas of its nodes over the root's) exceeds
 * twice the cost it had when built.  Off: the TLAS is rebuilt on every set-up, as the reference
 * rebuilds its IAS (scene.cpp:1325-1327). */
VLR_API VLRResult vlrContextSetTlasRefit(VLRContext context, bool enable);
VLR_API VLRResult vlrContextGetAccelStats(VLRContextConst context, VLRAccelStats* stats);
/* Copies one part of the acceleration structure of `scene`, as set up for rendering, to host
 * memory: which = 0 TLAS nodes, 1 BLAS nodes (64 B each), 2 BLAS triangle records (48 B: v0,
 * geometry instance, v1, primitive, v2, pad), 3 per-instance pairs (first BLAS node, first
 * triangle record; ~0 for instances without triangles).  dst = NULL queries the size. */
VLR_API VLRResult vlrContextReadAccel(VLRContext context, VLRScene scene, uint32_t which, void* dst, size_t capacity,
                                      size_t* bytes);
/* Builds a 4-wide BVH over numPrims boxes (6 floats each: lo xyz, hi xyz) with at most
 * maxLeafPrims (1..8) primitives per leaf.  Writes *numNodes nodes (at most
 * vlrBvh4NodeCapacity(numPrims)) and the primitive order (leaf codes index it).  With
 * leafRemap (maxLeafPrims 1) a leaf holds ~(leafRemap[primitive] << 3) instead.
 * vlrBuildBvh4Host works on host memory; vlrContextBuildBvh4 on device memory, on `stream`,
 * returning when the build has finished. */
VLR_API VLRResult vlrBvh4NodeCapacity(uint32_t numPrims, uint32_t* capacity);
VLR_API VLRResult vlrBuildBvh4Host(const float* boxes, uint32_t numPrims, uint32_t maxLeafPrims,
                                   const uint32_t* leafRemap, void* nodes, uint32_t nodeCapacity, uint32_t* order,
                                   uint32_t* numNodes);
VLR_API VLRResult vlrContextBuildBvh4(VLRContext context, hipStream_t stream, const float* boxes, uint32_t numPrims,
                                      uint32_t maxLeafPrims, const uint32_t* leafRemap, void* nodes,
                                      uint32_t nodeCapacity, uint32_t* order, uint32_t* numNodes);

/* vlrBuildBvh4NoAlloc: no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the build of vlrContextBuildBvh4 (same arguments and results, no context), on the current
 * GPU, calling no hipMalloc / hipFree.  Its scratch is one static region per GPU (about 35 MB,
 * held until the process ends) that every call on that GPU shares, so calls that overlap on
 * one GPU race.  numPrims above VLR_BVH4_NO_ALLOC_MAX_PRIMS is refused with InvalidArgument.
 * Returns when the build has finished on `stream`. */
#define VLR_BVH4_NO_ALLOC_MAX_PRIMS 131072u
VLR_API VLRResult vlrBuildBvh4NoAlloc(hipStream_t stream, const float* boxes, uint32_t numPrims,
                                      uint32_t maxLeafPrims, const uint32_t* leafRemap, void* nodes,
                                      uint32_t nodeCapacity, uint32_t* order, uint32_t* numNodes);

/* ---- shader-node entry point ----
 * Evaluates one shader-node plug (the node programs of shader_nodes.cu:20-483 with calcNode's
 * conversions, kernel_common.h:658-737) at numPoints device-resident surface points, each 20
 * floats: position, geometric normal, shading frame x, y, z, texture coordinates, wavelength
 * offset (spectral: the 4 wavelengths 360 + 470 (i + offset) / 4 nm), 2 unused.  The instance
 * transform is the identity.  values (device): 4 floats per point, the spectrum at the 4
 * wavelengths (RGB and 0 in the RGB build) for a Spectrum plug, else the value in the plug's
 * own type. */
VLR_API VLRResult vlrContextEvaluateShaderNodePlug(VLRContext context, hipStream_t stream, VLRShaderNodePlug plug,
                                                   bool spectral, const float* points, uint32_t numPoints,
                                                   float* values);

/* ---- BSDF entry point ---- */
typedef struct VLRBSDFQuery {
    float dirLocal[3];
    float geomNormalLocal[3];
    float wlOffset, wlSelect;
    float uComponent, uDir[2];
    float evalDir[3];
    int32_t singleIsSelected;
    int32_t transportMode;       /* 0 radiance, 1 importance (adjoint) */
} VLRBSDFQuery;
typedef struct VLRBSDFResult {
    float sampledDir[3];
    float sampledPDF;
    uint32_t sampledType;
    float sampledValue[4];
    float evalValue[4];
    float evalPDF;
    float revValue[4];           /* reverse (bidirectional) value and pdf of evaluate / evaluatePDF */
    float revPDF;
    float sampledRevValue[4];    /* reverse value and pdf of the sample */
    float sampledRevPDF;
} VLRBSDFResult;
/* Evaluates the material's BSDF (immediate parameters) on device-resident queries:
 * sample / evaluate / evaluatePDF in the shading frame, radiance or importance transport,
 * with the reverse quantities bidirectional methods use. */
VLR_API VLRResult vlrContextEvaluateBSDFs(VLRContext context, hipStream_t stream, VLRSurfaceMaterialConst material,
                                          bool spectral, const VLRBSDFQuery* queries, uint32_t numQueries,
                                          VLRBSDFResult* results);

#if defined(__cplusplus)
}
#endif
