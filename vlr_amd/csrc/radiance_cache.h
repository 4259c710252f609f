// Radiance cache: a 5 x 64 ReLU MLP (BF16 weights and activations, FP32 accumulation on MFMA)
// trained online on the path tracer's own vertices and queried to end paths early.  There is
// no counterpart in the reference; the query sits after Russian roulette in the path tracer's
// hit program (path_tracing.cu:185-189).
//
// Network: 96 inputs -> 64 -> 64 -> 64 -> 64 -> 64 -> 4 outputs.  Inputs are a 16-float raw
// record encoded on the fly (12-frequency sin/cos of the normalised position, the outgoing
// direction in spherical coordinates, the shading normal, roughness, diffuse and specular
// albedo, and the hero-wavelength offset), zero-padded to 96.  Outputs are radiance at the
// path's 4 hero wavelengths (spectral) or RGB.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace vlra {

namespace rc {
constexpr uint32_t kRawFloats = 16;     // raw record: see encodeRow in radiance_cache.hip
constexpr uint32_t kIn = 96;            // encoded input width (85 used)
constexpr uint32_t kHidden = 64;
constexpr uint32_t kOut = 4;
constexpr uint32_t kOutPad = 16;        // the output layer is stored as 16 rows (rows 4..15 zero)
constexpr uint32_t kLayers = 6;         // 5 hidden layers + output layer
constexpr uint32_t kFreqs = 12;

// per layer: rows (outputs, padded), columns (inputs)
__host__ __device__ constexpr uint32_t layerRows(uint32_t l) { return l == kLayers - 1 ? kOutPad : kHidden; }
__host__ __device__ constexpr uint32_t layerCols(uint32_t l) { return l == 0 ? kIn : kHidden; }
__host__ __device__ constexpr uint32_t weightOffset(uint32_t l) {
    return l == 0 ? 0u : kHidden * kIn + (l - 1) * kHidden * kHidden;
}
constexpr uint32_t kNumWeights = kHidden * kIn + 4 * kHidden * kHidden + kOutPad * kHidden;   // 23,552
__host__ __device__ constexpr uint32_t biasOffset(uint32_t l) { return l * kHidden; }
constexpr uint32_t kNumBiases = 5 * kHidden + kOutPad;                                         // 336
constexpr uint32_t kNumParams = kNumWeights + kNumBiases;                                      // 23,888
// parameters that are not padding: 23,108
constexpr uint32_t kNumRealParams = kHidden * kIn + 4 * kHidden * kHidden + kOut * kHidden + 5 * kHidden + kOut;

// activations kept for the backward pass, [row][sample] (rows: x_0 96, x_1..x_5 64 each)
constexpr uint32_t kActRows = kIn + 5 * kHidden;      // 416
__host__ __device__ constexpr uint32_t actRow(uint32_t l) { return l == 0 ? 0u : kIn + (l - 1) * kHidden; }
// gradients w.r.t. pre-activations, [row][sample] (delta_0..delta_4 64 rows, delta_5 16 rows)
constexpr uint32_t kDeltaRows = 5 * kHidden + kOutPad;   // 336
__host__ __device__ constexpr uint32_t deltaRow(uint32_t l) { return l * kHidden; }

constexpr uint32_t kGradSliceSamples = 2048;   // samples per weight-gradient partial
constexpr uint32_t kMaxBatch = 65536;
constexpr uint32_t kMaxSlices = kMaxBatch / kGradSliceSamples;
}  // namespace rc

// Device state of the cache (pointers into one allocation per buffer).
struct RcDevice {
    uint16_t* w;          // bf16 weights, row-major [rows][cols] per layer
    uint16_t* wt;         // bf16 transposed weights [cols][rows] per layer
    float* bias;          // fp32 biases (kNumBiases)
    float* master;        // fp32 master parameters: weights then biases (kNumParams)
    float* adamM;         // first moments (kNumParams)
    float* adamV;         // second moments (kNumParams)
    float* grad;          // reduced gradient of the last step (kNumParams)
    float* partial;       // gradient partials [slice][kNumParams] (weights, then biases)
    uint16_t* act;        // [kActRows][batchStride] bf16
    uint16_t* delta;      // [kDeltaRows][batchStride] bf16
    float* lossPartial;   // per training block
    float* loss;          // mean loss of the last step
    uint32_t batchStride; // samples per row of act / delta: the largest batch they hold
};

struct RcTrainParams {
    float learningRate;   // Adam (Kingma and Ba 2015)
    float beta1, beta2, epsilon;
    uint32_t step;        // 1-based step index of this update
    uint32_t numOut;      // channels in the loss (3 RGB, 4 spectral)
};

// Launchers (radiance_cache.hip).  `raw` holds n records of kRawFloats floats; `count` is a
// device pointer to the number of valid records when `countDev` is set (n is then the capacity).
hipError_t rcLaunchInfer(const RcDevice& d, const float* raw, const uint32_t* countDev, uint32_t n, float* out,
                         hipStream_t s);
hipError_t rcLaunchTrain(const RcDevice& d, const RcTrainParams& p, const float* raw, const float* targets,
                         const uint32_t* countDev, uint32_t n, hipStream_t s);
// Path p's query record is live when stamp[p] == frameStamp; its path contribution gets
// throughput * max(prediction, 0).
hipError_t rcLaunchQuery(const RcDevice& d, const float* raw, const float4* throughput, const uint32_t* stamp,
                         uint32_t frameStamp, uint32_t numPaths, float4* contrib, uint32_t numChannels,
                         hipStream_t s);

}  // namespace vlra
