// Context: output/accumulation/RNG buffers and render orchestration.
//
// Follows libVLR/context.cpp: bindOutputBuffer (:1130-1236) allocates the output, the Kahan
// accumulation buffer and one PCG32 state per pixel seeded by mt19937_64(591842031321323413)
// in row-major order (:1182-1187); render (:1267-1591) sets the scene up, resets the
// accumulation on firstFrame and renders one sample per pixel while
// numAccumFrames < limitNumAccumFrames (:1416-1443), then converts to RGBA (:1582-1589).
#include "runtime.h"

#include <cstring>
#include <random>
#include <stdexcept>

namespace vlra {

Context::Context(hipCtx_t hctx) : Object(ObjType::Context, nullptr) {
    ctx = this;
    (void)hctx;
    // no device work here: the scene graph can be built on a host without a GPU; device
    // buffers are allocated by bindOutputBuffer / the first render or trace call
    if (!loadSpectralTables()) throw std::runtime_error("spectral tables not found next to the library");
}

void Context::ensureDevice() {
    if (device >= 0) return;
    VLRA_HIP(hipGetDevice(&device));
    counters.reserve(kNumCounters * sizeof(uint32_t));
    VLRA_HIP(hipMemset(counters.ptr, 0, kNumCounters * sizeof(uint32_t)));
}

Context::~Context() {
    if (device >= 0) (void)hipDeviceSynchronize();
    if (outputArray) (void)hipFreeArray(outputArray);
    for (hipEvent_t e : tuneEvents)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : partEvents)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : wavefrontEvents)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t ps : partStreams)
        if (ps) (void)hipStreamDestroy(ps);
    for (Object *o : std::set<Object *>(objects)) delete o;
}

void Context::resizeAccum() {
    VLRA_HIP(hipDeviceSynchronize());   // frames of split halves may still run on their streams
    accumChannels = spectralActive ? 16 : 3;
    size_t n = size_t(width) * height;
    accum.reserve(n * accumChannels * 2 * sizeof(float));
    VLRA_HIP(hipMemset(accum.ptr, 0, n * accumChannels * 2 * sizeof(float)));
}

void Context::bindOutputBuffer(uint32_t w, uint32_t h) {
    if (w == 0 || h == 0) throw ApiError{3, "output size must be positive"};
    ensureDevice();
    VLRA_HIP(hipDeviceSynchronize());
    if (outputArray && (w != width || h != height)) {
        VLRA_HIP(hipFreeArray(outputArray));
        outputArray = nullptr;
    }
    width = w;
    height = h;
    const size_t n = size_t(w) * h;
    output.reserve(n * sizeof(float4));
    VLRA_HIP(hipMemset(output.ptr, 0, n * sizeof(float4)));
    rng.reserve(n * sizeof(uint64_t));
    alpha.reserve(n * sizeof(float4));
    contrib.reserve(n * sizeof(float4));
    meta.reserve(n * sizeof(float4));
    for (int k = 0; k < 2; ++k) {
        rayOrg[k].reserve(n * sizeof(float4));
        rayDir[k].reserve(n * sizeof(float4));
    }
    hitA.reserve(n * sizeof(int4));
    hitB.reserve(n * sizeof(float2));
    shOrg.reserve(n * sizeof(float4));
    shDir.reserve(n * sizeof(float4));
    shVal.reserve(n * sizeof(float4));
    resizeAccum();
    std::vector<uint64_t> seeds(n);
    std::mt19937_64 seedRng(591842031321323413ULL);
    for (size_t i = 0; i < n; ++i) seeds[i] = seedRng();
    VLRA_HIP(hipMemcpy(rng.ptr, seeds.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice));
    numAccumFrames = 0;
}

void Context::ensureScene(const Scene &s) {
    ensureDevice();
    if (gpu.valid && gpu.scene == &s && gpu.sceneVersion == version && gpu.spectral == spectralActive) return;
    VLRA_HIP(hipDeviceSynchronize());
    buildSceneGPU(*this, s, spectralActive, &gpu);
    gpu.scene = &s;
    gpu.sceneVersion = version;
}

namespace {
// Records a pair of events around each launch, from a pool the context keeps, and folds the
// elapsed times into the per-stage totals once the frame has been enqueued.
} // namespace

Context::EventTimer::~EventTimer() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}
hipEvent_t Context::EventTimer::next() {
    if (used == pool.size()) {
        hipEvent_t e;
        VLRA_HIP(hipEventCreate(&e));
        pool.push_back(e);
    }
    return pool[used++];
}
void Context::EventTimer::begin(int) { VLRA_HIP(hipEventRecord(next(), stream)); }
void Context::EventTimer::end(int stage) {
    VLRA_HIP(hipEventRecord(next(), stream));
    stages.push_back(stage);
}
void Context::EventTimer::collect(double *ms, uint64_t *launches) {
    if (used == 0) return;
    VLRA_HIP(hipEventSynchronize(pool[used - 1]));
    for (size_t k = 0; k < stages.size(); ++k) {
        float t = 0;
        VLRA_HIP(hipEventElapsedTime(&t, pool[2 * k], pool[2 * k + 1]));
        ms[stages[k]] += t;
        launches[stages[k]] += 1;
    }
    stages.clear();
    used = 0;
}

Context::EventTimer *Context::timer(int k, hipStream_t s) {
    if (!kernelTiming) return nullptr;
    EventTimer &t = timers[k];
    // events of one launch are recorded on one stream; launches of different streams do not mix
    if (t.used && t.stream != s) t.collect(stageMs, launches);
    if (t.used >= EventTimer::kMaxPending) t.collect(stageMs, launches);
    t.stream = s;
    return &t;
}

void Context::collectTimers() {
    for (EventTimer &t : timers) t.collect(stageMs, launches);
}

void Context::render(hipStream_t stream, const Camera *camera, uint32_t shrinkCoeff, bool firstFrame, uint32_t limit,
                     uint32_t *numAccum) {
    if (!scene) throw ApiError{3, "no scene is set"};
    if (!camera) throw ApiError{3, "camera is null"};
    if (!camera->perspective) throw ApiError{4, "only the perspective camera is on the render path"};
    if (width == 0 || height == 0) throw ApiError{3, "bind an output buffer first"};
    if (renderer != VLRRenderer_PathTracing && renderer != VLRRenderer_BPT && renderer != VLRRenderer_LightTracing)
        throw ApiError{3, "renderer not available on this build: VLRRenderer_PathTracing, VLRRenderer_LightTracing "
                          "and VLRRenderer_BPT"};
    const bool bpt = renderer == VLRRenderer_BPT;
    const bool lightTracing = renderer == VLRRenderer_LightTracing;
    if (shrinkCoeff == 0) shrinkCoeff = 1;
    if (firstFrame) {
        if (spectralActive != spectralRequested) {
            spectralActive = spectralRequested;
            resizeAccum();
        }
        numAccumFrames = 0;
        // the camera is read on the first frame only (context.cpp:1416-1421)
        frameCamera = camera->toDevice();
        resumePending = false;
    } else if (resumePending) {
        // the first frame after a written accumulation (a resumed render) reads the camera too
        frameCamera = camera->toDevice();
        resumePending = false;
    }
    ensureScene(*scene);
    gpu.dev.camera = frameCamera;
    if (numAccumFrames + 1 <= limit) {
        ++numAccumFrames;
        FrameParams fp;
        fp.width = width;
        fp.height = height;
        fp.imageW = width / shrinkCoeff;
        fp.imageH = height / shrinkCoeff;
        fp.frameIndex = numAccumFrames;
        fp.maxPathLength = std::min<uint32_t>(maxPathLength, kMaxDepthCounters - 1);
        fp.traversalStats = traversalStats ? 1 : 0;
        fp.wavefrontDepth = wavefrontDepth;
        fp.tailBlocks = tailBlocks;
        fp.sortShade = shadeSort ? 1 : 0;
        fp.guideReset = 0;
        fp.cacheMode = 0;
        fp.frameStamp = 0;
        fp.cacheThreshold = 0.0f;
        fp.cacheFullRadiance = 0;
        fp.trainStride = 1;
        fp.trainOffset = 0;
        fp.pixelBegin = 0;
        fp.pixelCount = fp.imageW * fp.imageH;
        if (bpt || lightTracing) {
            renderBptFrame(stream, fp, firstFrame, lightTracing);
            updateOutputArray(stream);
            if (numAccum) *numAccum = numAccumFrames;
            return;
        }
        // autotune the schedule per (scene version, size, mode, path length)
        const uint64_t tuneKey = (version * 1000003ull) ^ (uint64_t(width) << 40) ^ (uint64_t(height) << 20) ^
                                 (uint64_t(fp.imageW) << 8) ^ (spectralActive ? 1u : 0u) ^
                                 (uint64_t(fp.maxPathLength) << 2) ^ (cacheEnabled ? 0x2ull : 0ull) ^
                                 (splitFrames ? 0x1000000000000000ull : 0ull);
        bool tuning = false;
        if (autotune && !traversalStats) {
            if (tuner.key != tuneKey || tuner.cands.empty()) tuner.start(tuneKey);
            if (!tuner.done()) {
                tuning = true;
                fp.wavefrontDepth = tuner.current().wavefrontDepth;
                fp.tailBlocks = tuner.current().tailBlocks;
                fp.sortShade = tuner.current().sortShade;
            } else {
                fp.wavefrontDepth = tuner.best().wavefrontDepth;
                fp.tailBlocks = tuner.best().tailBlocks;
                fp.sortShade = tuner.best().sortShade;
            }
        }
        if (tuning) {
            if (!tuneEvents[0]) {
                VLRA_HIP(hipEventCreate(&tuneEvents[0]));
                VLRA_HIP(hipEventCreate(&tuneEvents[1]));
            }
            VLRA_HIP(hipEventRecord(tuneEvents[0], stream));
        }
        if (splitFrames && !cacheEnabled && fp.pixelCount >= kSplitMinPixels) fp.sortShade = 0;   // halves share one sort buffer
        if (fp.sortShade) {
            const size_t n = size_t(width) * height;
            sortKey.reserve(n * sizeof(uint32_t));
            sortPerm.reserve(n * sizeof(uint32_t));
            sortCounts.reserve(2 * kSortClasses * kMaxDepthCounters * sizeof(uint32_t));
        }
        FrameBuffers fb = frameBuffers();
        prepareGuides(fp, fb);
        CacheLaunch cl;
        const CacheLaunch *clp = nullptr;
        if (cacheEnabled) {
            ensureCacheNet();
            ensureCacheFrameBuffers();
            const uint32_t n = fp.imageW * fp.imageH;
            const uint32_t stride = std::max(1u, (n + cacheSettings.trainingPaths - 1) / cacheSettings.trainingPaths);
            fp.frameStamp = ++frameStamp;
            fp.cacheMode = cacheStep >= cacheSettings.warmupSteps ? 2 : 1;
            fp.cacheThreshold = cacheSettings.threshold;
            fp.cacheFullRadiance = cacheSettings.fullRadiance ? 1 : 0;
            fp.trainStride = stride;
            fp.trainOffset = fp.frameStamp % stride;
            fb.cache.queryRaw = static_cast<float4 *>(cqRaw.ptr);
            fb.cache.queryThroughput = static_cast<float4 *>(cqThroughput.ptr);
            fb.cache.queryStamp = static_cast<uint32_t *>(cqStamp.ptr);
            fb.cache.spread = static_cast<float2 *>(cSpread.ptr);
            fb.cache.trainRaw = static_cast<float4 *>(ctRaw.ptr);
            fb.cache.trainThroughput = static_cast<float4 *>(ctThroughput.ptr);
            fb.cache.trainContrib = static_cast<float4 *>(ctContrib.ptr);
            fb.cache.trainStamp = static_cast<uint32_t *>(ctStamp.ptr);
            fb.cache.trainAfterStamp = static_cast<uint32_t *>(ctAfterStamp.ptr);
            cl.net = rcDevice();
            cl.train.learningRate = cacheSettings.learningRate;
            cl.train.beta1 = cacheSettings.beta1;
            cl.train.beta2 = cacheSettings.beta2;
            cl.train.epsilon = cacheSettings.epsilon;
            cl.train.step = ++cacheStep;
            cl.train.numOut = spectralActive ? 4 : 3;
            cl.numSlots = ((n + stride - 1) / stride) * kCacheTrainVerts;
            cl.batchRaw = static_cast<float4 *>(cbRaw.ptr);
            cl.batchTarget = static_cast<float4 *>(cbTarget.ptr);
            cl.batchCount = static_cast<uint32_t *>(cbCount.ptr);
            cl.slotFlags = static_cast<uint32_t *>(csFlags.ptr);
            cl.slotOffsets = static_cast<uint32_t *>(csOffsets.ptr);
            cl.slotBlockSums = static_cast<uint32_t *>(csBlockSums.ptr);
            cl.slotTargets = static_cast<float4 *>(csTargets.ptr);
            clp = &cl;
        }
        const bool split = splitFrames && !cacheEnabled && fp.pixelCount >= kSplitMinPixels;
        lastFrameSplit = split;
        if (!split) joinHalves(stream);
        if (split) {
            renderSplitFrame(stream, fp, fb, tuning);
        } else {
            VLRA_HIP(launchFrame(gpu.dev, fp, fb, spectralActive, stream, timer(0, stream), clp));
        }
        if (tuning) {
            VLRA_HIP(hipEventRecord(tuneEvents[1], stream));
            VLRA_HIP(hipEventSynchronize(tuneEvents[1]));
            float ms = 0.0f;
            VLRA_HIP(hipEventElapsedTime(&ms, tuneEvents[0], tuneEvents[1]));
            tuner.report(ms);
        }
        lastFrame = fp;
        ++framesRendered;
        pathsStarted += uint64_t(fp.imageW) * fp.imageH;
        updateOutputArray(stream);
    }
    if (numAccum) *numAccum = numAccumFrames;
}

// The output as a 2D array of float4 (the reference's cudau::Array, context.cpp:1243), made on
// the first request; from then on every rendered frame is copied into it on the render stream.
// A GPU without image support (CDNA parts such as gfx950 report none) has no arrays: the
// request fails with InternalError and the output stays readable as the linear buffer.
hipArray_t Context::outputArrayHandle() {
    ensureDevice();
    if (!width) throw ApiError{3, "bind an output buffer first"};
    int images = 0;
    VLRA_HIP(hipDeviceGetAttribute(&images, hipDeviceAttributeImageSupport, device));
    if (!images)
        throw ApiError{5, "this GPU has no image support (hipDeviceAttributeImageSupport is 0), so the output has "
                          "no hipArray form; read it with vlrContextReadOutputBuffer"};
    if (!outputArray) {
        hipChannelFormatDesc desc = hipCreateChannelDesc<float4>();
        VLRA_HIP(hipMallocArray(&outputArray, &desc, width, height, hipArraySurfaceLoadStore));
        VLRA_HIP(hipDeviceSynchronize());
        VLRA_HIP(hipMemcpy2DToArray(outputArray, 0, 0, output.ptr, size_t(width) * sizeof(float4),
                                    size_t(width) * sizeof(float4), height, hipMemcpyDeviceToDevice));
    }
    return outputArray;
}

void Context::updateOutputArray(hipStream_t stream) {
    if (!outputArray) return;
    joinHalves(stream);
    VLRA_HIP(hipMemcpy2DToArrayAsync(outputArray, 0, 0, output.ptr, size_t(width) * sizeof(float4),
                                     size_t(width) * sizeof(float4), height, hipMemcpyDeviceToDevice, stream));
}

// The frame as two pixel halves, each generated, traced, shaded, accumulated and converted on
// its own stream with its own queues (the halves of the queue buffers) and counters.  A half
// depends only on its own previous frame, so the streams are not joined after a frame: the
// next frame of one half starts while the other half's deepest paths of this frame still
// run.  Every call that reads or reuses the frame buffers joins first (joinHalves, or a
// device synchronisation).  Every path computes exactly what it computes in a one-part frame.
void Context::renderSplitFrame(hipStream_t stream, const FrameParams &fp, const FrameBuffers &fb, bool join) {
    if (!partStreams[0]) {
        for (auto &ps : partStreams) VLRA_HIP(hipStreamCreateWithFlags(&ps, hipStreamNonBlocking));
        for (auto &e : partEvents) VLRA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (auto &e : wavefrontEvents) VLRA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    if (!counters2.ptr) {
        counters2.reserve(kNumCounters * sizeof(uint32_t));
        VLRA_HIP(hipMemset(counters2.ptr, 0, kNumCounters * sizeof(uint32_t)));
    }
    VLRA_HIP(hipEventRecord(partEvents[2], stream));
    const uint32_t half = fp.pixelCount / 2;
    for (int k = 0; k < 2; ++k) {
        VLRA_HIP(hipStreamWaitEvent(partStreams[k], partEvents[2], 0));
        // the halves take turns with their per-depth kernels: each half's tail kernel (a few
        // long path chains) overlaps the other half's per-depth kernels
        if (wavefrontRecorded[1 - k]) VLRA_HIP(hipStreamWaitEvent(partStreams[k], wavefrontEvents[1 - k], 0));
        FrameParams p = fp;
        p.pixelBegin = k == 0 ? 0 : half;
        p.pixelCount = k == 0 ? half : fp.pixelCount - half;
        FrameBuffers b = fb;
        const uint32_t o = p.pixelBegin;
        for (int q = 0; q < 2; ++q) {
            b.rays[q].org += o;
            b.rays[q].dir += o;
        }
        b.hits.a += o;
        b.hits.b += o;
        b.shadow.org += o;
        b.shadow.dir += o;
        b.shadow.value += o;
        if (k == 1) b.counters = static_cast<uint32_t *>(counters2.ptr);
        VLRA_HIP(launchFrame(gpu.dev, p, b, spectralActive, partStreams[k], timer(k, partStreams[k]), nullptr,
                             true, wavefrontEvents[k]));
        wavefrontRecorded[k] = true;
        VLRA_HIP(hipEventRecord(partEvents[k], partStreams[k]));
    }
    halvesPending = true;
    if (join) joinHalves(stream);
}

void Context::joinHalves(hipStream_t stream) {
    if (!halvesPending) return;
    for (int k = 0; k < 2; ++k) VLRA_HIP(hipStreamWaitEvent(stream, partEvents[k], 0));
    halvesPending = false;
}

FrameBuffers Context::frameBuffers() {
    FrameBuffers fb;
    fb.ps.rng = static_cast<uint64_t *>(rng.ptr);
    fb.ps.alpha = static_cast<float4 *>(alpha.ptr);
    fb.ps.contrib = static_cast<float4 *>(contrib.ptr);
    fb.ps.meta = static_cast<float4 *>(meta.ptr);
    for (int k = 0; k < 2; ++k) {
        fb.rays[k].org = static_cast<float4 *>(rayOrg[k].ptr);
        fb.rays[k].dir = static_cast<float4 *>(rayDir[k].ptr);
    }
    fb.hits.a = static_cast<int4 *>(hitA.ptr);
    fb.hits.b = static_cast<float2 *>(hitB.ptr);
    fb.shadow.org = static_cast<float4 *>(shOrg.ptr);
    fb.shadow.dir = static_cast<float4 *>(shDir.ptr);
    fb.shadow.value = static_cast<float4 *>(shVal.ptr);
    fb.counters = static_cast<uint32_t *>(counters.ptr);
    fb.sortKey = static_cast<uint32_t *>(sortKey.ptr);
    fb.sortPerm = static_cast<uint32_t *>(sortPerm.ptr);
    fb.sortCounts = static_cast<uint32_t *>(sortCounts.ptr);
    fb.accum = static_cast<float *>(accum.ptr);
    fb.output = static_cast<float4 *>(output.ptr);
    fb.cache = CacheBuffers{};
    fb.guideAlbedo = fb.guideNormal = nullptr;
    return fb;
}

// Denoiser guides of the frame, when switched on: their buffers, the restart rule and the
// frame buffers' pointers.
void Context::prepareGuides(FrameParams &fp, FrameBuffers &fb) {
    if (!guidesEnabled) return;
    const size_t n = size_t(width) * height, bytes = n * (accumChannels + 3) * sizeof(float);
    if (guideBytes != bytes) {
        guideAlbedo.reserve(n * accumChannels * sizeof(float));
        guideNormal.reserve(n * 3 * sizeof(float));
        guideBytes = bytes;
        guideFrames = 0;
    }
    // the guides restart with the accumulation, or when they were just switched on
    fp.guideReset = (numAccumFrames == 1 || guideFrames == 0) ? 1 : 0;
    guideFrames = fp.guideReset ? 1 : guideFrames + 1;
    fb.guideAlbedo = static_cast<float *>(guideAlbedo.ptr);
    fb.guideNormal = static_cast<float *>(guideNormal.ptr);
}

// ---- LVC-BPT (context.cpp:1444-1481 of the reference) ----
void Context::setNumLightPaths(uint32_t n) {
    if (n == 0 || n > kMaxLightPaths) throw ApiError{3, "number of light paths must be in [1, 1048576]"};
    numLightPaths = n;
}

void Context::ensureBpt(uint32_t slotsPerPath) {
    ensureDevice();
    if (lightRngPaths != numLightPaths) {
        // the light-path RNG states are seeded once, as the reference does at context creation
        std::vector<uint64_t> seeds(numLightPaths);
        std::mt19937_64 seedRng(459182033132123413ULL);
        for (auto &v : seeds) v = seedRng();
        lightRng.upload(seeds.data(), seeds.size() * sizeof(uint64_t));
        lightRngPaths = numLightPaths;
    }
    const size_t vertices = size_t(numLightPaths) * slotsPerPath;
    lvSlots.reserve(vertices * kLightVertexFloat4s * sizeof(float4));
    lvCache.reserve(vertices * kLightVertexFloat4s * sizeof(float4));
    lvCounts.reserve(size_t(numLightPaths) * sizeof(uint32_t));
    lvOffsets.reserve(size_t(numLightPaths) * sizeof(uint32_t));
    lvBlockSums.reserve(1024 * sizeof(uint32_t));
    lvNum.reserve(sizeof(uint32_t));
    const size_t splatBytes = size_t(width) * height * accumChannels * sizeof(float);
    void *before = splat.ptr;
    splat.reserve(splatBytes);
    if (splat.ptr != before || splatLayout != splatBytes) {
        VLRA_HIP(hipMemset(splat.ptr, 0, splatBytes));
        splatLayout = splatBytes;
    }
}

void Context::renderBptFrame(hipStream_t stream, const FrameParams &frame, bool firstFrame, bool lightTracing) {
    joinHalves(stream);
    FrameParams fp = frame;
    const uint32_t slotsPerPath = fp.maxPathLength + 1;
    ensureBpt(slotsPerPath);
    if (firstFrame)
        VLRA_HIP(hipMemsetAsync(splat.ptr, 0, size_t(width) * height * accumChannels * sizeof(float), stream));
    // common wavelengths of a BPT frame: two draws, offset first (WavelengthSamples::
    // createWithEqualOffsets, spectrum_types.h); light tracing draws per light path instead
    float wlOffset = 0.0f, wlSelect = 0.0f;
    if (!lightTracing) bptWavelengths(&wlOffset, &wlSelect);
    const uint32_t N = spectralActive ? 4 : 3;
    BptParams bp;
    bp.wlOffset = wlOffset;
    bp.wlSel = std::min<uint32_t>(static_cast<uint16_t>(N * wlSelect), N - 1);
    bp.wlProb = spectralActive ? 4 / (830.0f - 360.0f) : 1.0f;
    bp.numLightPaths = numLightPaths;
    bp.slotsPerPath = slotsPerPath;
    bp.slots = static_cast<float4 *>(lvSlots.ptr);
    bp.slotCounts = static_cast<uint32_t *>(lvCounts.ptr);
    bp.slotOffsets = static_cast<uint32_t *>(lvOffsets.ptr);
    bp.blockSums = static_cast<uint32_t *>(lvBlockSums.ptr);
    bp.cache = static_cast<float4 *>(lvCache.ptr);
    bp.numVertices = static_cast<uint32_t *>(lvNum.ptr);
    bp.lightRng = static_cast<uint64_t *>(lightRng.ptr);
    bp.splat = static_cast<float *>(splat.ptr);
    FrameBuffers fb = frameBuffers();
    prepareGuides(fp, fb);
    auto launch = lightTracing ? launchLightTraceFrame : launchBptFrame;
    VLRA_HIP(launch(gpu.dev, fp, fb, bp, spectralActive, stream, timer(0, stream)));
    lastFrame = fp;
    ++framesRendered;
    pathsStarted += lightTracing ? uint64_t(numLightPaths) : uint64_t(fp.imageW) * fp.imageH;
}

void Context::readLightRngStates(uint64_t *dst) {
    ensureBpt(std::min<uint32_t>(maxPathLength, kMaxDepthCounters - 1) + 1);
    VLRA_HIP(hipDeviceSynchronize());
    VLRA_HIP(hipMemcpy(dst, lightRng.ptr, size_t(numLightPaths) * 8, hipMemcpyDeviceToHost));
}

void Context::writeLightRngStates(const uint64_t *src) {
    ensureBpt(std::min<uint32_t>(maxPathLength, kMaxDepthCounters - 1) + 1);
    VLRA_HIP(hipDeviceSynchronize());
    VLRA_HIP(hipMemcpy(lightRng.ptr, src, size_t(numLightPaths) * 8, hipMemcpyHostToDevice));
}

// ---- radiance cache ----
namespace {
uint16_t bf16Host(float x) {   // round to nearest even
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
// PCG32 (O'Neill 2014) for the weight initialisation
struct HostPcg {
    uint64_t state;
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + 1ULL;
        uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    float uniform() { return (next() >> 8) * (1.0f / 16777216.0f); }
};
}  // namespace

RcDevice Context::rcDevice() {
    RcDevice d;
    d.w = static_cast<uint16_t *>(rcW.ptr);
    d.wt = static_cast<uint16_t *>(rcWT.ptr);
    d.bias = static_cast<float *>(rcBias.ptr);
    d.master = static_cast<float *>(rcMaster.ptr);
    d.adamM = static_cast<float *>(rcM.ptr);
    d.adamV = static_cast<float *>(rcV.ptr);
    d.grad = static_cast<float *>(rcGrad.ptr);
    d.partial = static_cast<float *>(rcPartial.ptr);
    d.act = static_cast<uint16_t *>(rcAct.ptr);
    d.delta = static_cast<uint16_t *>(rcDelta.ptr);
    d.lossPartial = static_cast<float *>(rcLossPartial.ptr);
    d.loss = static_cast<float *>(rcLoss.ptr);
    d.batchStride = rc::kMaxBatch;
    return d;
}

void Context::writeCacheWeights(const float *master) {
    ensureDevice();
    std::vector<uint16_t> w(rc::kNumWeights), wt(rc::kNumWeights);
    for (uint32_t l = 0; l < rc::kLayers; ++l) {
        const uint32_t rows = rc::layerRows(l), cols = rc::layerCols(l), off = rc::weightOffset(l);
        for (uint32_t r = 0; r < rows; ++r)
            for (uint32_t c = 0; c < cols; ++c) {
                uint16_t h = bf16Host(master[off + r * cols + c]);
                w[off + r * cols + c] = h;
                wt[off + c * rows + r] = h;
            }
    }
    rcMaster.upload(master, rc::kNumParams * sizeof(float));
    rcW.upload(w.data(), w.size() * 2);
    rcWT.upload(wt.data(), wt.size() * 2);
    rcBias.upload(master + rc::kNumWeights, rc::kNumBiases * sizeof(float));
    std::vector<float> zeros(rc::kNumParams, 0.0f);
    rcM.upload(zeros.data(), zeros.size() * sizeof(float));
    rcV.upload(zeros.data(), zeros.size() * sizeof(float));
    rcGrad.upload(zeros.data(), zeros.size() * sizeof(float));
    cacheStep = 0;
}

void Context::readCacheWeights(float *master) {
    ensureCacheNet();
    VLRA_HIP(hipDeviceSynchronize());
    VLRA_HIP(hipMemcpy(master, rcMaster.ptr, rc::kNumParams * sizeof(float), hipMemcpyDeviceToHost));
}

// Glorot-uniform weights from PCG32(seed), zero biases; padding rows of the output layer stay 0.
void Context::resetCache() {
    ensureDevice();
    std::vector<float> master(rc::kNumParams, 0.0f);
    HostPcg rng{0x853c49e6748fea9bULL ^ (static_cast<uint64_t>(cacheSettings.seed) * 0x9E3779B97F4A7C15ULL)};
    for (uint32_t l = 0; l < rc::kLayers; ++l) {
        const uint32_t rows = rc::layerRows(l), cols = rc::layerCols(l), off = rc::weightOffset(l);
        const uint32_t fanOut = l == rc::kLayers - 1 ? rc::kOut : rows;
        const float a = std::sqrt(6.0f / static_cast<float>(cols + fanOut));
        for (uint32_t r = 0; r < rows; ++r)
            for (uint32_t c = 0; c < cols; ++c) {
                float u = rng.uniform();
                master[off + r * cols + c] = r < fanOut ? a * (2.0f * u - 1.0f) : 0.0f;
            }
    }
    rcPartial.reserve(size_t(rc::kMaxSlices) * rc::kNumParams * sizeof(float));
    rcAct.reserve(size_t(rc::kActRows) * rc::kMaxBatch * 2);
    rcDelta.reserve(size_t(rc::kDeltaRows) * rc::kMaxBatch * 2);
    rcLossPartial.reserve(rc::kMaxBatch / 128 * sizeof(float));
    rcLoss.reserve(sizeof(float));
    VLRA_HIP(hipMemset(rcLoss.ptr, 0, sizeof(float)));
    cbRaw.reserve(size_t(rc::kMaxBatch) * 4 * sizeof(float4));
    cbTarget.reserve(size_t(rc::kMaxBatch) * sizeof(float4));
    cbCount.reserve(sizeof(uint32_t));
    csFlags.reserve(size_t(rc::kMaxBatch) * sizeof(uint32_t));
    csOffsets.reserve(size_t(rc::kMaxBatch) * sizeof(uint32_t));
    csBlockSums.reserve(64 * sizeof(uint32_t));
    csTargets.reserve(size_t(rc::kMaxBatch) * sizeof(float4));
    writeCacheWeights(master.data());
    cacheInitialized = true;
}

void Context::ensureCacheNet() {
    if (!cacheInitialized) resetCache();
}

void Context::ensureCacheFrameBuffers() {
    const size_t n = size_t(width) * height;
    const size_t bytesRaw = n * 4 * sizeof(float4);
    if (cqRaw.bytes < bytesRaw) {
        cqRaw.reserve(bytesRaw);
        cqThroughput.reserve(n * sizeof(float4));
        cqStamp.reserve(n * sizeof(uint32_t));
        VLRA_HIP(hipMemset(cqStamp.ptr, 0, n * sizeof(uint32_t)));
        cSpread.reserve(n * sizeof(float2));
    }
    if (!ctRaw.ptr) {
        ctRaw.reserve(size_t(rc::kMaxBatch) * 4 * sizeof(float4));
        ctThroughput.reserve(size_t(rc::kMaxBatch) * sizeof(float4));
        ctContrib.reserve(size_t(rc::kMaxBatch) * sizeof(float4));
        ctStamp.reserve(size_t(rc::kMaxBatch) * sizeof(uint32_t));
        VLRA_HIP(hipMemset(ctStamp.ptr, 0, size_t(rc::kMaxBatch) * sizeof(uint32_t)));
        ctAfterStamp.reserve(size_t(rc::kMaxBatch) * sizeof(uint32_t));
        VLRA_HIP(hipMemset(ctAfterStamp.ptr, 0, size_t(rc::kMaxBatch) * sizeof(uint32_t)));
    }
}

void Context::cacheInfer(hipStream_t stream, const float *records, uint32_t n, float *out) {
    ensureCacheNet();
    VLRA_HIP(rcLaunchInfer(rcDevice(), records, nullptr, n, out, stream));
}

void Context::cacheTrain(hipStream_t stream, const float *records, const float *targets, uint32_t n,
                         uint32_t numOut) {
    if (n == 0 || n > rc::kMaxBatch) throw ApiError{3, "training batch must hold 1 to 65536 samples"};
    if (numOut < 1 || numOut > 4) throw ApiError{3, "outputs in the loss must be 1 to 4"};
    ensureCacheNet();
    RcTrainParams p;
    p.learningRate = cacheSettings.learningRate;
    p.beta1 = cacheSettings.beta1;
    p.beta2 = cacheSettings.beta2;
    p.epsilon = cacheSettings.epsilon;
    p.step = ++cacheStep;
    p.numOut = numOut;
    VLRA_HIP(rcLaunchTrain(rcDevice(), p, records, targets, nullptr, n, stream));
}

// ---- batches ----
BatchState::~BatchState() {
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {
struct GraphExec {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GraphExec() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};
}  // namespace

// Renders every job to its sample count.  Each step renders one sample per pixel of every job
// that has samples left, over one set of path buffers; a step's launch sequence is captured
// once into a graph and replayed.  Pixels are seeded per job like a single render
// (mt19937_64(591842031321323413), row-major), so each job's image is bit-identical to
// rendering it alone.
void Context::renderBatch(hipStream_t userStream, const std::vector<BatchJobDesc> &jobs, float *completionMs) {
    ensureDevice();
    if (jobs.empty()) return;
    for (const BatchJobDesc &j : jobs) {
        if (!j.scene || !j.camera) throw ApiError{3, "batch job without scene or camera"};
        if (!j.camera->perspective) throw ApiError{4, "only the perspective camera is on the render path"};
        if (j.width == 0 || j.height == 0 || j.spp == 0) throw ApiError{3, "batch job with an empty image or no samples"};
        if (j.maxPathLength == 0 || j.maxPathLength >= kMaxDepthCounters)
            throw ApiError{3, "batch job max path length must be 1 to 63"};
    }
    VLRA_HIP(hipDeviceSynchronize());
    BatchState &b = batch;
    const bool spectral = spectralRequested;
    if (b.sceneVersion != version || b.spectral != spectral) {
        b.scenes.clear();
        b.sceneVersion = version;
        b.spectral = spectral;
    }
    for (const BatchJobDesc &j : jobs) {
        auto &g = b.scenes[j.scene];
        if (!g) {
            g = std::make_unique<SceneGPU>();
            buildSceneGPU(*this, *j.scene, spectral, g.get());
            g->scene = j.scene;
            g->sceneVersion = version;
        }
    }
    // path buffers over all jobs' pixels
    const uint32_t C = spectral ? 16 : 3;
    std::vector<uint32_t> offsets(jobs.size());
    size_t total = 0;
    uint32_t maxLen = 0, maxSpp = 0;
    for (size_t k = 0; k < jobs.size(); ++k) {
        offsets[k] = static_cast<uint32_t>(total);
        total += size_t(jobs[k].width) * jobs[k].height;
        maxLen = std::max(maxLen, jobs[k].maxPathLength);
        maxSpp = std::max(maxSpp, jobs[k].spp);
    }
    if (total >= (1ull << 31)) throw ApiError{3, "batch too large"};
    const size_t n = total;
    b.rng.reserve(n * sizeof(uint64_t));
    b.alpha.reserve(n * sizeof(float4));
    b.contrib.reserve(n * sizeof(float4));
    b.meta.reserve(n * sizeof(float4));
    for (int k = 0; k < 2; ++k) {
        b.rayOrg[k].reserve(n * sizeof(float4));
        b.rayDir[k].reserve(n * sizeof(float4));
    }
    b.hitA.reserve(n * sizeof(int4));
    b.hitB.reserve(n * sizeof(float2));
    b.shOrg.reserve(n * sizeof(float4));
    b.shDir.reserve(n * sizeof(float4));
    b.shVal.reserve(n * sizeof(float4));
    b.accum.resize(jobs.size());
    b.output.resize(jobs.size());
    b.widths.resize(jobs.size());
    b.heights.resize(jobs.size());
    std::vector<uint64_t> seeds(n);
    std::vector<DevJob> table(jobs.size());
    for (size_t k = 0; k < jobs.size(); ++k) {
        const BatchJobDesc &j = jobs[k];
        const size_t px = size_t(j.width) * j.height;
        if (!b.accum[k]) b.accum[k] = std::make_unique<DeviceBuffer>();
        if (!b.output[k]) b.output[k] = std::make_unique<DeviceBuffer>();
        b.accum[k]->reserve(px * C * 2 * sizeof(float));
        b.output[k]->reserve(px * sizeof(float4));
        b.widths[k] = j.width;
        b.heights[k] = j.height;
        std::mt19937_64 seedRng(591842031321323413ULL);
        for (size_t i = 0; i < px; ++i) seeds[offsets[k] + i] = seedRng();
        DevJob &d = table[k];
        std::memset(&d, 0, sizeof(d));
        d.scene = b.scenes[j.scene]->dev;
        d.scene.camera = j.camera->toDevice();
        d.accum = static_cast<float *>(b.accum[k]->ptr);
        d.output = static_cast<float4 *>(b.output[k]->ptr);
        d.pixelOffset = offsets[k];
        d.width = j.width;
        d.height = j.height;
        d.spp = j.spp;
        d.maxPathLength = j.maxPathLength;
    }
    VLRA_HIP(hipMemcpy(b.rng.ptr, seeds.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice));
    b.jobs.upload(table.data(), table.size() * sizeof(DevJob));
    b.jobOffsets.upload(offsets.data(), offsets.size() * sizeof(uint32_t));
    b.activeOffsets.reserve((jobs.size() + 1) * sizeof(uint32_t));
    if (!b.stream) VLRA_HIP(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));

    BatchLaunch L;
    L.anyScene = table[0].scene;
    std::memset(&L.fp, 0, sizeof(L.fp));
    L.fp.width = L.fp.imageW = static_cast<uint32_t>(n);
    L.fp.height = L.fp.imageH = 1;
    L.fp.maxPathLength = maxLen;
    L.fp.trainStride = 1;
    L.fb = FrameBuffers{};
    L.fb.ps.rng = static_cast<uint64_t *>(b.rng.ptr);
    L.fb.ps.alpha = static_cast<float4 *>(b.alpha.ptr);
    L.fb.ps.contrib = static_cast<float4 *>(b.contrib.ptr);
    L.fb.ps.meta = static_cast<float4 *>(b.meta.ptr);
    for (int k = 0; k < 2; ++k) {
        L.fb.rays[k].org = static_cast<float4 *>(b.rayOrg[k].ptr);
        L.fb.rays[k].dir = static_cast<float4 *>(b.rayDir[k].ptr);
    }
    L.fb.hits.a = static_cast<int4 *>(b.hitA.ptr);
    L.fb.hits.b = static_cast<float2 *>(b.hitB.ptr);
    L.fb.shadow.org = static_cast<float4 *>(b.shOrg.ptr);
    L.fb.shadow.dir = static_cast<float4 *>(b.shDir.ptr);
    L.fb.shadow.value = static_cast<float4 *>(b.shVal.ptr);
    L.fb.counters = static_cast<uint32_t *>(counters.ptr);
    L.fb.sortKey = L.fb.sortPerm = L.fb.sortCounts = nullptr;   // no material sort in batches
    L.jobs = static_cast<DevJob *>(b.jobs.ptr);
    L.jobOffsets = static_cast<const uint32_t *>(b.jobOffsets.ptr);
    L.activeOffsets = static_cast<uint32_t *>(b.activeOffsets.ptr);
    L.numJobs = static_cast<uint32_t>(jobs.size());
    L.nodePrograms = false;
    for (const auto &kv : b.scenes) {
        L.nodePrograms = L.nodePrograms || kv.second->dev.hasNodePrograms;
        // the batch kernels draw no random numbers during traversal
        if (kv.second->dev.hasAlpha) throw ApiError{4, "alpha nodes are not supported in render batches"};
    }

    // schedule candidates (wavefront depth; the tail grid fixed), each timed on one step when
    // autotuning is on and there are enough steps; otherwise the context's schedule
    std::vector<uint32_t> cands;
    if (autotune && maxSpp >= 8) cands = {0, 1, 2, 3, 4, 6};
    else cands = {wavefrontDepth};
    std::map<uint32_t, std::unique_ptr<GraphExec>> graphs;
    auto stepLaunch = [&](uint32_t wd, hipStream_t s) {
        BatchLaunch Lw = L;
        Lw.fp.wavefrontDepth = wd;
        Lw.fp.tailBlocks = kDefaultTailBlocks;
        if (!batchGraphs) {
            VLRA_HIP(launchBatchStep(Lw, spectral, s));
            return;
        }
        auto &ge = graphs[wd];
        if (!ge) {
            ge = std::make_unique<GraphExec>();
            VLRA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            hipError_t le = launchBatchStep(Lw, spectral, s);
            hipError_t ce = hipStreamEndCapture(s, &ge->graph);
            VLRA_HIP(le);
            VLRA_HIP(ce);
            VLRA_HIP(hipGraphInstantiate(&ge->exec, ge->graph, nullptr, nullptr, 0));
        }
        VLRA_HIP(hipGraphLaunch(ge->exec, s));
    };
    std::vector<hipEvent_t> ev(maxSpp + 1);
    for (auto &e : ev) VLRA_HIP(hipEventCreate(&e));
    hipEvent_t ready;
    VLRA_HIP(hipEventCreate(&ready));
    VLRA_HIP(hipEventRecord(ready, userStream));
    VLRA_HIP(hipStreamWaitEvent(b.stream, ready, 0));
    VLRA_HIP(hipEventRecord(ev[0], b.stream));
    uint32_t chosen = cands[0];
    float best = 1e30f;
    for (uint32_t step = 0; step < maxSpp; ++step) {
        const bool tuning = cands.size() > 1 && step < cands.size();
        const uint32_t wd = tuning ? cands[step] : chosen;
        stepLaunch(wd, b.stream);
        VLRA_HIP(hipEventRecord(ev[step + 1], b.stream));
        if (tuning) {
            VLRA_HIP(hipEventSynchronize(ev[step + 1]));
            float ms = 0.0f;
            VLRA_HIP(hipEventElapsedTime(&ms, ev[step], ev[step + 1]));
            if (ms < best) {
                best = ms;
                chosen = wd;
            }
        }
    }
    VLRA_HIP(launchBatchConvert(L, spectral, b.stream));
    hipEvent_t done;
    VLRA_HIP(hipEventCreate(&done));
    VLRA_HIP(hipEventRecord(done, b.stream));
    VLRA_HIP(hipStreamWaitEvent(userStream, done, 0));
    b.wavefrontDepth = chosen;
    b.tailBlocks = kDefaultTailBlocks;
    VLRA_HIP(hipEventSynchronize(done));
    if (completionMs) {
        for (size_t k = 0; k < jobs.size(); ++k) {
            float ms = 0.0f;
            VLRA_HIP(hipEventElapsedTime(&ms, ev[0], ev[jobs[k].spp]));
            completionMs[k] = ms;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(ready);
    (void)hipEventDestroy(done);
    framesRendered += maxSpp;
    for (const BatchJobDesc &j : jobs) pathsStarted += uint64_t(j.width) * j.height * j.spp;
}

void Context::readBatchOutput(uint32_t job, float *rgba) {
    if (job >= batch.output.size() || !batch.output[job]) throw ApiError{3, "no such batch job"};
    VLRA_HIP(hipDeviceSynchronize());
    VLRA_HIP(hipMemcpy(rgba, batch.output[job]->ptr, size_t(batch.widths[job]) * batch.heights[job] * sizeof(float4),
                       hipMemcpyDeviceToHost));
}

void Context::readOutput(float *dst) {
    VLRA_HIP(hipDeviceSynchronize());
    VLRA_HIP(hipMemcpy(dst, output.ptr, size_t(width) * height * sizeof(float4), hipMemcpyDeviceToHost));
}

void Context::traceRays(hipStream_t stream, const Scene &s, const void *rays, uint32_t n, void *hits, bool anyHit) {
    joinHalves(stream);
    ensureScene(s);
    VLRA_HIP(launchTraceRays(gpu.dev, rays, n, hits, anyHit, stream));
}

void Context::setAccelBuilder(uint32_t builder) {
    if (builder != kAccelBuilderDevice && builder != kAccelBuilderHost) throw ApiError{3, "unknown builder"};
    if (builder != accelBuilder) {
        accelBuilder = builder;
        touch();
    }
}

void Context::readAccel(const Scene &s, uint32_t which, void *dst, size_t capacity, size_t *bytes) {
    if (which > 3) throw ApiError{3, "unknown acceleration-structure part"};
    ensureScene(s);
    VLRA_HIP(hipDeviceSynchronize());
    size_t n = 0;
    const void *src = nullptr;
    std::vector<uint32_t> pairs;
    if (which == 0) {
        n = size_t(gpu.numTlasNodes) * sizeof(BVH4QNode);
        src = gpu.tlasNodes.ptr;
    } else if (which == 1) {
        n = size_t(gpu.numBlasNodes) * sizeof(BVH4QNode);
        src = gpu.blasNodes.ptr;
    } else if (which == 2) {
        n = size_t(gpu.numTriangles) * sizeof(TriRecord);
        src = gpu.blasTris.ptr;
    } else {
        std::vector<DevInstance> insts(gpu.numInstances);
        if (!insts.empty())
            VLRA_HIP(hipMemcpy(insts.data(), gpu.instances.ptr, insts.size() * sizeof(DevInstance), hipMemcpyDeviceToHost));
        for (const DevInstance &I : insts) {
            pairs.push_back(I.hasGeometry ? I.blasNodeBase : ~0u);
            pairs.push_back(I.hasGeometry ? I.blasTriBase : ~0u);
        }
        n = pairs.size() * sizeof(uint32_t);
    }
    *bytes = n;
    if (!dst) return;
    if (capacity < n) throw ApiError{3, "destination too small"};
    if (which == 3) std::memcpy(dst, pairs.data(), n);
    else if (n) VLRA_HIP(hipMemcpy(dst, src, n, hipMemcpyDeviceToHost));
}

void Context::evaluateNodePlug(hipStream_t stream, const PlugRef &plug, bool spectral, const float *points, uint32_t n,
                               float *values) {
    ensureDevice();
    if (!loadSpectralTables()) throw std::runtime_error("spectral tables not found next to the library");
    ShaderTable T;
    T.spectral = spectral;
    const uint32_t out = plug.plugType == VLRShaderNodePlugType_Spectrum ? 1u : 8u;
    const auto prog = T.program({{plug, out}});
    if (T.nodes.empty()) T.nodes.emplace_back();
    if (T.steps.empty()) T.steps.emplace_back();
    if (T.textures.empty()) T.textures.emplace_back();
    if (T.texels.empty()) T.texels.push_back(make_float4(0, 0, 0, 0));
    if (T.pool.empty()) T.pool.push_back(0.0f);
    DevInstance ident;
    std::memset(&ident, 0, sizeof(ident));
    for (int r = 0; r < 3; ++r) ident.o2w[r * 4 + r] = ident.w2o[r * 4 + r] = 1.0f;
    DeviceBuffer nodesB, stepsB, texB, texelsB, poolB, instB, gridB, pointsB;
    nodesB.upload(T.nodes.data(), T.nodes.size() * sizeof(DevShaderNode));
    stepsB.upload(T.steps.data(), T.steps.size() * sizeof(DevNodeStep));
    texB.upload(T.textures.data(), T.textures.size() * sizeof(DevTexture));
    texelsB.upload(T.texels.data(), T.texels.size() * sizeof(float4));
    poolB.upload(T.pool.data(), T.pool.size() * sizeof(float));
    instB.upload(&ident, sizeof(ident));
    gridB.upload(mengGrid().data(), mengGrid().size() * sizeof(MengCellDev));
    pointsB.upload(mengPoints().data(), mengPoints().size() * sizeof(float));
    DevScene sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.nodes = static_cast<const DevShaderNode *>(nodesB.ptr);
    sc.nodeSteps = static_cast<const DevNodeStep *>(stepsB.ptr);
    sc.textures = static_cast<const DevTexture *>(texB.ptr);
    sc.texels = static_cast<const float4 *>(texelsB.ptr);
    sc.nodePool = static_cast<const float *>(poolB.ptr);
    sc.instances = static_cast<const DevInstance *>(instB.ptr);
    sc.mengGrid = static_cast<const MengCellDev *>(gridB.ptr);
    sc.mengPoints = static_cast<const float *>(pointsB.ptr);
    sc.numInstances = 1;
    sc.hasNodePrograms = 1;
    VLRA_HIP(launchEvalNodePlug(sc, prog.first, prog.second, plug.plugType, spectral, points, n, values, stream));
    VLRA_HIP(hipStreamSynchronize(stream));
}

void Context::evaluateBSDFs(hipStream_t stream, const SurfaceMaterial &m, bool spectral, const void *q, uint32_t n,
                            void *r) {
    if (!m.supportedOnRenderPath()) throw ApiError{4, "material type not on the render path"};
    ensureDevice();
    bool usesTex;
    const Image2D *img;
    DevMaterial dm = m.toDevice(spectral, &usesTex, &img, nullptr);
    // the BSDF kernels only read the Meng tables from the scene record
    if (!gpu.mengPoints.ptr) {
        gpu.mengGrid.upload(mengGrid().data(), mengGrid().size() * sizeof(MengCellDev));
        gpu.mengPoints.upload(mengPoints().data(), mengPoints().size() * sizeof(float));
    }
    DevScene sc = gpu.dev;
    sc.mengGrid = static_cast<const MengCellDev *>(gpu.mengGrid.ptr);
    sc.mengPoints = static_cast<const float *>(gpu.mengPoints.ptr);
    VLRA_HIP(launchBsdfQueries(sc, dm, spectral, q, n, r, stream));
}

void Context::statistics(uint64_t *closest, uint64_t *shadow, uint64_t *dropped, uint64_t trav[6],
                         uint64_t wavefront[2]) {
    *closest = *shadow = *dropped = 0;
    for (int k = 0; k < 6; ++k) trav[k] = 0;
    wavefront[0] = wavefront[1] = 0;
    if (device < 0) return;
    for (int part = 0; part < 2; ++part) {
        const DeviceBuffer &cb = part == 0 ? counters : counters2;
        if (!cb.ptr) continue;
        uint32_t host[kNumCounters];
        VLRA_HIP(hipDeviceSynchronize());
        VLRA_HIP(hipMemcpy(host, cb.ptr, sizeof(host), hipMemcpyDeviceToHost));
        uint64_t v[2];
        std::memcpy(v, host + kStatClosest, 8);
        *closest += v[0];
        std::memcpy(v, host + kStatShadow, 8);
        *shadow += v[0];
        *dropped += host[kStatDropped];
        uint64_t t[6];
        std::memcpy(t, host + kStatTravClosest, 24);
        std::memcpy(t + 3, host + kStatTravShadow, 24);
        for (int k = 0; k < 6; ++k) trav[k] += t[k];
        std::memcpy(v, host + kStatWavefront, 16);
        wavefront[0] += v[0];
        wavefront[1] += v[1];
    }
}

void Context::queueCounts(uint32_t *closest, uint32_t *shadow, uint32_t maxDepth) {
    uint32_t host[kNumCounters] = {}, host2[kNumCounters] = {};
    if (device >= 0) {
        VLRA_HIP(hipDeviceSynchronize());
        VLRA_HIP(hipMemcpy(host, counters.ptr, sizeof(host), hipMemcpyDeviceToHost));
        // the second half's queues of the last frame, when it was split
        if (counters2.ptr && lastFrameSplit)
            VLRA_HIP(hipMemcpy(host2, counters2.ptr, sizeof(host2), hipMemcpyDeviceToHost));
    }
    for (uint32_t d = 0; d < maxDepth; ++d) {
        const bool inFrame = d < kMaxDepthCounters && d < lastFrame.maxPathLength;
        closest[d] = inFrame ? host[2 * d] + host2[2 * d] : 0;
        shadow[d] = inFrame ? host[2 * d + 1] + host2[2 * d + 1] : 0;
    }
}

void Context::bptWavelengths(float *offset, float *select) {
    std::uniform_real_distribution<float> u01;
    *offset = u01(bptWlRng);
    *select = u01(bptWlRng);
    ++bptWlFrames;
}

void Context::setBptWavelengthFrames(uint64_t frames) {
    bptWlRng = std::mt19937();
    bptWlFrames = 0;
    float o, s;
    for (uint64_t f = 0; f < frames; ++f) bptWavelengths(&o, &s);
}

void Context::readGuides(float *albedoRGBA, float *normalRGBA) {
    const size_t n = size_t(width) * height;
    const uint32_t C = accumChannels;
    std::vector<float> alb(n * C, 0.0f), nrm(n * 3, 0.0f);
    VLRA_HIP(hipDeviceSynchronize());
    if (guideFrames > 0) {
        VLRA_HIP(hipMemcpy(alb.data(), guideAlbedo.ptr, alb.size() * sizeof(float), hipMemcpyDeviceToHost));
        VLRA_HIP(hipMemcpy(nrm.data(), guideNormal.ptr, nrm.size() * sizeof(float), hipMemcpyDeviceToHost));
    }
    float cx[16], cy[16], cz[16], integral = 1.0f;
    if (C == 16) cmfBins(cx, cy, cz, &integral);
    const float rec = guideFrames > 0 ? 1.0f / guideFrames : 0.0f;
    const float *R = frameCamera.rot;   // column-major; its transpose is the inverse orientation
    for (size_t p = 0; p < n; ++p) {
        float X = 0.0f, Y = 0.0f, Z = 0.0f;
        if (C == 16) {
            for (uint32_t k = 0; k < 16; ++k) {
                const float v = alb[k * n + p];
                X += cx[k] * v;
                Y += cy[k] * v;
                Z += cz[k] * v;
            }
            X /= integral;
            Y /= integral;
            Z /= integral;
        } else {
            const float r = alb[p], g = alb[n + p], b = alb[2 * n + p];
            X = 0.4124564f * r + 0.3575761f * g + 0.1804375f * b;
            Y = 0.2126729f * r + 0.7151522f * g + 0.0721750f * b;
            Z = 0.0193339f * r + 0.1191920f * g + 0.9503041f * b;
        }
        X *= rec;
        Y *= rec;
        Z *= rec;
        if (albedoRGBA) {
            albedoRGBA[4 * p + 0] = 3.2404542f * X + -1.5371385f * Y + -0.4985314f * Z;
            albedoRGBA[4 * p + 1] = -0.9692660f * X + 1.8760108f * Y + 0.0415560f * Z;
            albedoRGBA[4 * p + 2] = 0.0556434f * X + -0.2040259f * Y + 1.0572252f * Z;
            albedoRGBA[4 * p + 3] = 1.0f;
        }
        if (normalRGBA) {
            const float nx = nrm[p], ny = nrm[n + p], nz = nrm[2 * n + p];
            float c[3];
            for (int i = 0; i < 3; ++i) c[i] = R[3 * i + 0] * nx + R[3 * i + 1] * ny + R[3 * i + 2] * nz;
            c[0] *= -1;
            if (c[0] != 0 || c[1] != 0 || c[2] != 0) {
                const float l = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
                for (float &v : c) v /= l;
            }
            normalRGBA[4 * p + 0] = c[0];
            normalRGBA[4 * p + 1] = c[1];
            normalRGBA[4 * p + 2] = c[2];
            normalRGBA[4 * p + 3] = 1.0f;
        }
    }
}

void Context::readAccumulation(float *perPixel) {
    const size_t n = size_t(width) * height, planes = size_t(accumChannels) * 2;
    std::vector<float> dev(n * planes);
    VLRA_HIP(hipDeviceSynchronize());
    VLRA_HIP(hipMemcpy(dev.data(), accum.ptr, dev.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < planes; ++q)
        for (size_t p = 0; p < n; ++p) perPixel[p * planes + q] = dev[q * n + p];
}

void Context::writeAccumulation(const float *perPixel) {
    const size_t n = size_t(width) * height, planes = size_t(accumChannels) * 2;
    std::vector<float> dev(n * planes);
    for (size_t q = 0; q < planes; ++q)
        for (size_t p = 0; p < n; ++p) dev[q * n + p] = perPixel[p * planes + q];
    VLRA_HIP(hipDeviceSynchronize());
    VLRA_HIP(hipMemcpy(accum.ptr, dev.data(), dev.size() * sizeof(float), hipMemcpyHostToDevice));
}

void Context::resetStats() {
    collectTimers();
    if (device >= 0) {
        VLRA_HIP(hipDeviceSynchronize());
        VLRA_HIP(hipMemset(counters.ptr, 0, kNumCounters * sizeof(uint32_t)));
        if (counters2.ptr) VLRA_HIP(hipMemset(counters2.ptr, 0, kNumCounters * sizeof(uint32_t)));
    }
    for (int i = 0; i < NUM_STAGES; ++i) {
        stageMs[i] = 0;
        launches[i] = 0;
    }
    framesRendered = 0;
    pathsStarted = 0;
}

} // namespace vlra
