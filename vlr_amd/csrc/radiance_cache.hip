// Radiance-cache MLP kernels for gfx950: inference, query (path termination), and one training
// step (forward + backward, weight gradients over sample slices, deterministic reduction, Adam).
//
// All matrix products run on v_mfma_f32_16x16x32_bf16 (16x16 output tile, K = 32, BF16
// operands, FP32 accumulation).  Operand layout per lane l (g = l / 16, c = l % 16):
//   A (16 x 32):  row c, k = 8g .. 8g+7          (8 consecutive BF16 = one 16-byte load)
//   B (32 x 16):  column c, k = 8g .. 8g+7
//   D (16 x 16):  column c, rows 4g .. 4g+3
// Each wave owns 64 samples.  Activations live in LDS as [sample][feature] BF16 rows, so a B
// operand (8 consecutive features of one sample) is one 16-byte LDS read; a layer's output is
// written back in place once the wave has read that tile's inputs.  Weight tiles (A operands)
// are read from global memory (23 k parameters, L2-resident) once per layer and reused over
// the wave's 4 sample tiles.  The training step keeps every activation and every gradient
// w.r.t. a pre-activation in [row][sample] layout in HBM, so the weight-gradient products
// (a reduction over samples) read 8 consecutive samples per lane with one 16-byte load.
#include "radiance_cache.h"

namespace vlra {
using namespace rc;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));

constexpr float kPiF = 3.14159265358979323846f;
constexpr uint32_t kRow = 104;        // BF16 per LDS activation row (96 + 8: rows start 16-byte aligned)
constexpr uint32_t kDeltaRowLds = 72; // BF16 per LDS gradient row (64 + 8)

__device__ __forceinline__ uint16_t toBf16(float x) {
    __bf16 b = static_cast<__bf16>(x);   // round to nearest even
    return __builtin_bit_cast(uint16_t, b);
}
__device__ __forceinline__ float fromBf16(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t *p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(p));
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u16x8{0, 0, 0, 0, 0, 0, 0, 0}); }
__device__ __forceinline__ f32x4 mfma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Raw record -> 96 encoded inputs, written as BF16 into an LDS row.
//   raw[0..2]   position normalised to the scene box, [0, 1]
//   raw[3..4]   outgoing direction: theta / pi, phi / (2 pi)
//   raw[5..7]   shading normal, (n + 1) / 2
//   raw[8]      roughness;  raw[9..11] diffuse albedo;  raw[12..14] specular albedo
//   raw[15]     hero-wavelength offset (0 for RGB)
// inputs 24a + 2f, 24a + 2f + 1 = sin, cos(2^f pi p_a) for f < 12 (double-angle recurrence
// from one sincos per axis); inputs 72..84 = raw[3..15]; 85..95 = 0.
__device__ __forceinline__ void encodeRow(const float *raw, bool valid, uint16_t *row) {
    float r[kRawFloats];
#pragma unroll
    for (uint32_t k = 0; k < kRawFloats; ++k) r[k] = valid ? raw[k] : 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float s, c;
        sincosf(kPiF * r[a], &s, &c);
        if (!valid) s = c = 0.0f;
        uint16_t v[24];
#pragma unroll
        for (int f = 0; f < 12; ++f) {
            v[2 * f] = toBf16(s);
            v[2 * f + 1] = toBf16(c);
            float s2 = 2.0f * s * c;
            float c2 = 1.0f - 2.0f * s * s;
            s = s2;
            c = c2;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
            *reinterpret_cast<u16x8 *>(row + 24 * a + 8 * k) =
                u16x8{v[8 * k], v[8 * k + 1], v[8 * k + 2], v[8 * k + 3], v[8 * k + 4], v[8 * k + 5], v[8 * k + 6],
                      v[8 * k + 7]};
    }
    *reinterpret_cast<u16x8 *>(row + 72) = u16x8{toBf16(r[3]), toBf16(r[4]), toBf16(r[5]), toBf16(r[6]),
                                                 toBf16(r[7]), toBf16(r[8]), toBf16(r[9]), toBf16(r[10])};
    *reinterpret_cast<u16x8 *>(row + 80) =
        u16x8{toBf16(r[11]), toBf16(r[12]), toBf16(r[13]), toBf16(r[14]), toBf16(r[15]), 0, 0, 0};
    *reinterpret_cast<u16x8 *>(row + 88) = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
}

// Forward pass of one wave over its 64 samples (rows of `act`).  With STORE, x_1..x_5 are also
// written to d.act in [row][sample] layout at samples sampleBase + 0..63.  Lanes with g = 0
// return the 4 outputs of sample 16q + c in y[q].
template <bool STORE>
__device__ __forceinline__ void forwardWave(const RcDevice &d, uint16_t *act, uint32_t lane, uint32_t sampleBase,
                                            f32x4 y[4]) {
    const uint32_t c = lane & 15, g = lane >> 4;
#pragma unroll 1
    for (uint32_t l = 0; l < kLayers - 1; ++l) {
        const uint32_t K = layerCols(l);
        const uint32_t ks = K / 32;
        const uint16_t *W = d.w + weightOffset(l);
        bf16x8 A[4][3];
        float bias[4][4];
#pragma unroll
        for (uint32_t t = 0; t < 4; ++t) {
#pragma unroll
            for (uint32_t s = 0; s < 3; ++s)
                A[t][s] = s < ks ? ld8(W + (16 * t + c) * K + 32 * s + 8 * g) : zero8();
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r) bias[t][r] = d.bias[biasOffset(l) + 16 * t + 4 * g + r];
        }
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            uint16_t *row = act + (16 * q + c) * kRow;
            bf16x8 B[3];
#pragma unroll
            for (uint32_t s = 0; s < 3; ++s) B[s] = s < ks ? ld8(row + 32 * s + 8 * g) : zero8();
            f32x4 acc[4];
#pragma unroll
            for (uint32_t t = 0; t < 4; ++t) {
                acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (uint32_t s = 0; s < 3; ++s)
                    if (s < ks) acc[t] = mfma(A[t][s], B[s], acc[t]);
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (uint32_t t = 0; t < 4; ++t) {
                uint16_t h[4];
#pragma unroll
                for (uint32_t r = 0; r < 4; ++r) h[r] = toBf16(fmaxf(acc[t][r] + bias[t][r], 0.0f));
                *reinterpret_cast<u16x4 *>(row + 16 * t + 4 * g) = u16x4{h[0], h[1], h[2], h[3]};
                if constexpr (STORE) {
                    const uint32_t sample = sampleBase + 16 * q + c;
#pragma unroll
                    for (uint32_t r = 0; r < 4; ++r)
                        d.act[(actRow(l + 1) + 16 * t + 4 * g + r) * d.batchStride + sample] = h[r];
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    // output layer: 16 padded rows, rows 0..3 live in lanes g = 0
    const uint16_t *W = d.w + weightOffset(kLayers - 1);
    bf16x8 A0 = ld8(W + c * kHidden + 8 * g), A1 = ld8(W + c * kHidden + 32 + 8 * g);
    float bias[4];
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r) bias[r] = d.bias[biasOffset(kLayers - 1) + 4 * g + r];
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
        const uint16_t *row = act + (16 * q + c) * kRow;
        bf16x8 B0 = ld8(row + 8 * g), B1 = ld8(row + 32 + 8 * g);
        f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        acc = mfma(A0, B0, acc);
        acc = mfma(A1, B1, acc);
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) acc[r] += bias[r];
        y[q] = acc;
    }
}

constexpr uint32_t kFwdBlock = 256;   // 4 waves x 64 samples
constexpr uint32_t kTrainBlock = 128; // 2 waves x 64 samples

__global__ void __launch_bounds__(kFwdBlock) k_rcInfer(RcDevice d, const float *raw, const uint32_t *countDev,
                                                        uint32_t n, float *out) {
    __shared__ __attribute__((aligned(16))) uint16_t lds[4 * 64 * kRow];
    const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x & 63;
    const uint32_t count = countDev ? min(*countDev, n) : n;
    const uint32_t base = blockIdx.x * kFwdBlock + wave * 64;
    if (base >= count) return;   // wave-uniform
    uint16_t *act = lds + wave * 64 * kRow;
    const uint32_t s = base + lane;
    encodeRow(raw + static_cast<size_t>(min(s, count - 1)) * kRawFloats, s < count, act + lane * kRow);
    __builtin_amdgcn_wave_barrier();
    f32x4 y[4];
    forwardWave<false>(d, act, lane, 0, y);
    const uint32_t c = lane & 15, g = lane >> 4;
    if (g == 0) {
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t sq = base + 16 * q + c;
            if (sq < count) *reinterpret_cast<float4 *>(out + static_cast<size_t>(sq) * 4) = make_float4(y[q][0], y[q][1], y[q][2], y[q][3]);
        }
    }
}

__global__ void __launch_bounds__(kFwdBlock) k_rcQuery(RcDevice d, const float *raw, const float4 *throughput,
                                                        const uint32_t *stamp, uint32_t frameStamp,
                                                        uint32_t numPaths, float4 *contrib, uint32_t numChannels) {
    __shared__ __attribute__((aligned(16))) uint16_t lds[4 * 64 * kRow];
    const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x & 63;
    const uint32_t base = blockIdx.x * kFwdBlock + wave * 64;
    const uint32_t s = base + lane;
    const bool live = s < numPaths && stamp[s] == frameStamp;
    if (__ballot(live) == 0) return;   // wave-uniform
    uint16_t *act = lds + wave * 64 * kRow;
    encodeRow(raw + static_cast<size_t>(live ? s : base) * kRawFloats, live, act + lane * kRow);
    __builtin_amdgcn_wave_barrier();
    f32x4 y[4];
    forwardWave<false>(d, act, lane, 0, y);
    const uint32_t c = lane & 15, g = lane >> 4;
    if (g == 0) {
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t sq = base + 16 * q + c;
            if (sq >= numPaths || stamp[sq] != frameStamp) continue;
            float4 T = throughput[sq];
            float4 acc = contrib[sq];
            acc.x += T.x * fmaxf(y[q][0], 0.0f);
            acc.y += T.y * fmaxf(y[q][1], 0.0f);
            acc.z += T.z * fmaxf(y[q][2], 0.0f);
            if (numChannels == 4) acc.w += T.w * fmaxf(y[q][3], 0.0f);
            contrib[sq] = acc;
        }
    }
}

// Forward + backward of one training step for 128 samples per block.  Writes x_0..x_5 and
// delta_0..delta_5 ([row][sample], BF16) for every sample slot below the padded batch, zeros
// past `count`, and the block's summed loss.
__global__ void __launch_bounds__(kTrainBlock) k_rcTrain(RcDevice d, RcTrainParams p, const float *raw,
                                                          const float *targets, const uint32_t *countDev,
                                                          uint32_t n) {
    __shared__ __attribute__((aligned(16))) uint16_t actLds[2 * 64 * kRow];
    __shared__ __attribute__((aligned(16))) uint16_t deltaLds[2 * 64 * kDeltaRowLds];
    __shared__ float lossLds[2];
    const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x & 63;
    const uint32_t c = lane & 15, g = lane >> 4;
    const uint32_t count = countDev ? min(*countDev, n) : n;
    const uint32_t base = blockIdx.x * kTrainBlock + wave * 64;
    uint16_t *act = actLds + wave * 64 * kRow;
    uint16_t *dl = deltaLds + wave * 64 * kDeltaRowLds;
    const uint32_t s = base + lane;
    const bool valid = s < count;
    encodeRow(raw + static_cast<size_t>(valid ? s : 0) * kRawFloats, valid, act + lane * kRow);
    // x_0 to HBM
#pragma unroll 8
    for (uint32_t k = 0; k < kIn; ++k) d.act[(actRow(0) + k) * d.batchStride + s] = act[lane * kRow + k];
    __builtin_amdgcn_wave_barrier();
    f32x4 y[4];
    forwardWave<true>(d, act, lane, base, y);

    // loss and its gradient w.r.t. the outputs: relative L2, (y - t)^2 / (sg(y)^2 + 0.01)
    float lossSum = 0.0f;
    const float norm = 1.0f / (static_cast<float>(max(count, 1u)) * static_cast<float>(p.numOut));
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t sq = base + 16 * q + c;
        uint16_t h[4] = {0, 0, 0, 0};
        if (g == 0 && sq < count) {
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                if (k >= p.numOut) continue;
                float t = targets[static_cast<size_t>(sq) * 4 + k];
                float yy = y[q][k];
                float den = yy * yy + 0.01f;
                float diff = yy - t;
                lossSum += diff * diff / den;
                h[k] = toBf16(2.0f * diff / den * norm);
            }
        }
        // delta_5 (16 padded rows) to HBM; lanes g = 0 hold rows 0..3
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) d.delta[(deltaRow(5) + 4 * g + r) * d.batchStride + sq] = h[r];
        // LDS row of sample 16q + c: columns 0..31 = delta_5 rows 0..31 (rows >= 4 are 0)
        if (g == 0) {
            uint16_t *row = dl + (16 * q + c) * kDeltaRowLds;
            *reinterpret_cast<u16x8 *>(row) = u16x8{h[0], h[1], h[2], h[3], 0, 0, 0, 0};
            *reinterpret_cast<u16x8 *>(row + 8) = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
            *reinterpret_cast<u16x8 *>(row + 16) = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
            *reinterpret_cast<u16x8 *>(row + 24) = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
        }
    }
    __builtin_amdgcn_wave_barrier();

    // delta_{l-1} = (W_l^T delta_l) * [x_l > 0] for l = 5 .. 1
#pragma unroll 1
    for (int l = kLayers - 1; l >= 1; --l) {
        const uint32_t rows = layerRows(l);        // K of this product (16 for the output layer)
        const uint32_t ks = l == kLayers - 1 ? 1 : 2;
        const uint16_t *WT = d.wt + weightOffset(l);   // [cols][rows]
        bf16x8 A[4][2];
#pragma unroll
        for (uint32_t t = 0; t < 4; ++t)
#pragma unroll
            for (uint32_t s2 = 0; s2 < 2; ++s2) {
                const uint32_t k0 = 32 * s2 + 8 * g;
                A[t][s2] = (s2 < ks && k0 < rows) ? ld8(WT + (16 * t + c) * rows + k0) : zero8();
            }
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            uint16_t *row = dl + (16 * q + c) * kDeltaRowLds;
            bf16x8 B[2];
#pragma unroll
            for (uint32_t s2 = 0; s2 < 2; ++s2) B[s2] = s2 < ks ? ld8(row + 32 * s2 + 8 * g) : zero8();
            f32x4 acc[4];
#pragma unroll
            for (uint32_t t = 0; t < 4; ++t) {
                acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (uint32_t s2 = 0; s2 < 2; ++s2)
                    if (s2 < ks) acc[t] = mfma(A[t][s2], B[s2], acc[t]);
            }
            __builtin_amdgcn_wave_barrier();
            const uint32_t sample = base + 16 * q + c;
#pragma unroll
            for (uint32_t t = 0; t < 4; ++t) {
                uint16_t h[4];
#pragma unroll
                for (uint32_t r = 0; r < 4; ++r) {
                    const uint32_t nrn = 16 * t + 4 * g + r;
                    const float x = fromBf16(d.act[(actRow(l) + nrn) * d.batchStride + sample]);
                    h[r] = toBf16(x > 0.0f ? acc[t][r] : 0.0f);
                    d.delta[(deltaRow(l - 1) + nrn) * d.batchStride + sample] = h[r];
                }
                *reinterpret_cast<u16x4 *>(row + 16 * t + 4 * g) = u16x4{h[0], h[1], h[2], h[3]};
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    // block loss
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) lossSum += __shfl_xor(lossSum, off);
    if (lane == 0) lossLds[wave] = lossSum;
    __syncthreads();
    if (threadIdx.x == 0) d.lossPartial[blockIdx.x] = lossLds[0] + lossLds[1];
}

// Weight-gradient tiles: tile index -> (layer, row tile, column tile)
__device__ __forceinline__ void tileOf(uint32_t tile, uint32_t *l, uint32_t *ti, uint32_t *tj) {
    if (tile < 24) {
        *l = 0;
        *ti = tile / 6;
        *tj = tile % 6;
    } else if (tile < 88) {
        *l = 1 + (tile - 24) / 16;
        *ti = ((tile - 24) % 16) / 4;
        *tj = (tile - 24) % 4;
    } else {
        *l = 5;
        *ti = 0;
        *tj = tile - 88;
    }
}
constexpr uint32_t kNumTiles = 92;

// One wave per (tile, slice of kGradSliceSamples samples): dW tile = sum_s delta_l[i][s] x_l[j][s].
// The waves of column tile 0 also sum delta_l over the slice: the bias gradient of rows i.
__global__ void __launch_bounds__(256) k_rcWeightGrad(RcDevice d, uint32_t nPad, uint32_t numSlices) {
    const uint32_t waveId = blockIdx.x * 4 + threadIdx.x / 64;
    if (waveId >= kNumTiles * numSlices) return;   // wave-uniform
    const uint32_t lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
    const uint32_t tile = waveId % kNumTiles, slice = waveId / kNumTiles;
    uint32_t l, ti, tj;
    tileOf(tile, &l, &ti, &tj);
    const uint32_t s0 = slice * kGradSliceSamples;
    const uint32_t s1 = min(s0 + kGradSliceSamples, nPad);
    const uint16_t *Arow = d.delta + static_cast<size_t>(deltaRow(l) + 16 * ti + c) * d.batchStride;
    const uint16_t *Brow = d.act + static_cast<size_t>(actRow(l) + 16 * tj + c) * d.batchStride;
    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float rowSum = 0.0f;
    for (uint32_t s = s0; s < s1; s += 32) {
        const u16x8 a = *reinterpret_cast<const u16x8 *>(Arow + s + 8 * g);
        acc = mfma(__builtin_bit_cast(bf16x8, a), ld8(Brow + s + 8 * g), acc);
        if (tj == 0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) rowSum += fromBf16(a[k]);
        }
    }
    const uint32_t cols = layerCols(l);
    float *dst = d.partial + static_cast<size_t>(slice) * kNumParams;
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r) dst[weightOffset(l) + (16 * ti + 4 * g + r) * cols + 16 * tj + c] = acc[r];
    if (tj == 0) {
        rowSum += __shfl_xor(rowSum, 16);
        rowSum += __shfl_xor(rowSum, 32);
        if (g == 0) dst[kNumWeights + biasOffset(l) + 16 * ti + c] = rowSum;
    }
}

// Sum of the gradient partials over slices (fixed order) and the Adam update of every parameter;
// writes the FP32 master, the BF16 weights in both layouts and the FP32 biases.
__global__ void __launch_bounds__(256) k_rcAdam(RcDevice d, RcTrainParams p, uint32_t numSlices,
                                                 uint32_t numLossBlocks, const uint32_t *countDev, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x == 0) {
        // the loss: the blocks' partial sums added in a fixed order (per thread a strided run,
        // then a tree over the 256 threads), so the value repeats from run to run
        __shared__ float part[256];
        float sum = 0.0f;
        for (uint32_t b = threadIdx.x; b < numLossBlocks; b += 256) sum += d.lossPartial[b];
        part[threadIdx.x] = sum;
        __syncthreads();
        for (uint32_t w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const uint32_t count = countDev ? min(*countDev, n) : n;
            *d.loss = part[0] / (static_cast<float>(max(count, 1u)) * static_cast<float>(p.numOut));
        }
    }
    if (i >= kNumParams) return;
    float gsum = 0.0f;
    for (uint32_t s = 0; s < numSlices; ++s) gsum += d.partial[static_cast<size_t>(s) * kNumParams + i];
    d.grad[i] = gsum;
    float m = p.beta1 * d.adamM[i] + (1.0f - p.beta1) * gsum;
    float v = p.beta2 * d.adamV[i] + (1.0f - p.beta2) * gsum * gsum;
    d.adamM[i] = m;
    d.adamV[i] = v;
    const float mHat = m / (1.0f - powf(p.beta1, static_cast<float>(p.step)));
    const float vHat = v / (1.0f - powf(p.beta2, static_cast<float>(p.step)));
    const float w = d.master[i] - p.learningRate * mHat / (sqrtf(vHat) + p.epsilon);
    d.master[i] = w;
    if (i < kNumWeights) {
        uint32_t l = 0;
        while (l + 1 < kLayers && i >= weightOffset(l + 1)) ++l;
        const uint32_t local = i - weightOffset(l), cols = layerCols(l), rows = layerRows(l);
        const uint32_t r = local / cols, cc = local % cols;
        const uint16_t h = toBf16(w);
        d.w[i] = h;
        d.wt[weightOffset(l) + cc * rows + r] = h;
    } else {
        d.bias[i - kNumWeights] = w;
    }
}

inline uint32_t blocksFor(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

}  // namespace

hipError_t rcLaunchInfer(const RcDevice &d, const float *raw, const uint32_t *countDev, uint32_t n, float *out,
                         hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rcInfer, dim3(blocksFor(n, kFwdBlock)), dim3(kFwdBlock), 0, s, d, raw, countDev, n, out);
    return hipGetLastError();
}

hipError_t rcLaunchQuery(const RcDevice &d, const float *raw, const float4 *throughput, const uint32_t *stamp,
                         uint32_t frameStamp, uint32_t numPaths, float4 *contrib, uint32_t numChannels,
                         hipStream_t s) {
    if (numPaths == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rcQuery, dim3(blocksFor(numPaths, kFwdBlock)), dim3(kFwdBlock), 0, s, d, raw, throughput,
                       stamp, frameStamp, numPaths, contrib, numChannels);
    return hipGetLastError();
}

hipError_t rcLaunchTrain(const RcDevice &d, const RcTrainParams &p, const float *raw, const float *targets,
                         const uint32_t *countDev, uint32_t n, hipStream_t s) {
    if (n == 0 || n > d.batchStride || d.batchStride % kGradSliceSamples != 0) return hipErrorInvalidValue;
    const uint32_t blocks = blocksFor(n, kTrainBlock);
    const uint32_t nPad = blocks * kTrainBlock;
    const uint32_t slices = blocksFor(nPad, kGradSliceSamples);
    hipLaunchKernelGGL(k_rcTrain, dim3(blocks), dim3(kTrainBlock), 0, s, d, p, raw, targets, countDev, n);
    hipLaunchKernelGGL(k_rcWeightGrad, dim3(blocksFor(kNumTiles * slices, 4)), dim3(256), 0, s, d, nPad, slices);
    hipLaunchKernelGGL(k_rcAdam, dim3(blocksFor(kNumParams, 256)), dim3(256), 0, s, d, p, slices, blocks, countDev,
                       n);
    return hipGetLastError();
}

}  // namespace vlra
