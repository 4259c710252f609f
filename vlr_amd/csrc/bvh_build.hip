// Device build of the acceleration structures (the algorithm, shared with the host build, is
// described in bvh.h).
//
//   wide phase    nodes of more than kBvhSmallPrims primitives, one level per round: blocks of
//                 kBinChunk primitives bin centroids in LDS and merge the bins with atomic
//                 min / max on order-preserving integer keys (exact, so independent of the
//                 order the blocks run in); one wave per node evaluates the 45 SAH boundaries;
//                 a stable partition places every primitive by its rank (per-block counts, a
//                 per-node scan of them, ballot ranks inside the block).  The host reads back
//                 one small record per node per round and numbers the children in node order.
//   subtrees      one workgroup per node of at most kBvhSmallPrims primitives: boxes, ids and
//                 the local order live in LDS; the four waves take the nodes of one depth in
//                 turn and split each alone (wave-synchronous binning, sweep and partition).  A
//                 node that splits at local position m is given id base + m - 1: each
//                 boundary between two primitives is the split of exactly one node, so ids do
//                 not depend on the order the waves run in, and each node writes its id into
//                 its parent's slot.
//   collapse      breadth-first, one thread per 4-wide node: greedy opening of the largest
//                 inner child, quantisation (quantizeBvh4Node), a scan numbers the next level.
#include "bvh.h"
#include "runtime.h"

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

namespace vlra {
namespace {

constexpr uint32_t kBinChunk = 4096;   // primitives per block in the wide phase
constexpr int kBinWords = 3 * kBvhBins * 8;   // per node: [axis][bin] -> lo xyz, hi xyz, count, pad
constexpr int kLdsBinWords = 3 * kBvhBins * 7;
constexpr uint32_t kMaxFront = kBvhSmallPrims / 2;   // nodes of one depth inside a subtree
constexpr uint32_t kNoParent = 0xFFFFFFFFu;

__device__ inline uint32_t fkey(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float keyf(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k); }
float hostKeyf(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__device__ inline void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline float waveMin(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float waveMax(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline unsigned long long waveMinU64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o);
        v = w < v ? w : v;
    }
    return v;
}

struct PrimBox {
    float lo[3], hi[3];
};
__device__ inline PrimBox loadBox(const float4 *bx, uint32_t p) {
    const float4 a = bx[2 * p], b = bx[2 * p + 1];
    return {{a.x, a.y, a.z}, {b.x, b.y, b.z}};
}
__device__ inline float centroid(const PrimBox &b, int a) { return 0.5f * (b.lo[a] + b.hi[a]); }

// Candidate key of the SAH sweep: the smallest key wins; equal costs go to the lower axis, then
// to the higher boundary (the order of the host sweep).  ~0 = no candidate.
__device__ inline unsigned long long sahKey(float cost, int axis, int k) {
    return (static_cast<unsigned long long>(__float_as_uint(cost)) << 32) |
           static_cast<unsigned long long>(axis * 16 + (kBvhBins - 1 - k));
}

// Evaluates boundary k (1..kBvhBins-1) on `axis` from bins laid out as [axis][bin][stride]
// with lo keys at 0..2, hi keys at 3..5 and the count at 6.
template <int STRIDE>
__device__ inline unsigned long long sahCandidate(const uint32_t *bins, int axis, int k, uint32_t *leftCount) {
    float llo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, lhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float rlo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, rhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    uint32_t lc = 0, rc = 0;
    for (int j = 0; j < kBvhBins; ++j) {
        const uint32_t *w = bins + (axis * kBvhBins + j) * STRIDE;
        const uint32_t c = w[6];
        if (!c) continue;
        float *lo = j < k ? llo : rlo, *hi = j < k ? lhi : rhi;
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], keyf(w[a]));
            hi[a] = fmaxf(hi[a], keyf(w[3 + a]));
        }
        if (j < k) lc += c;
        else rc += c;
    }
    *leftCount = lc;
    if (!lc || !rc) return ~0ull;
    const float cost = bvhArea(llo, lhi) * static_cast<float>(lc) + bvhArea(rlo, rhi) * static_cast<float>(rc);
    if (!(cost < FLT_MAX)) return ~0ull;
    return sahKey(cost, axis, k);
}

// ---- primitive boxes ----
__global__ void k_boxesFromFloats(const float *in, uint32_t n, float4 *bx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *b = in + 6 * size_t(i);
    bx[2 * i] = make_float4(bvhNoNegZero(b[0]), bvhNoNegZero(b[1]), bvhNoNegZero(b[2]), 0.0f);
    bx[2 * i + 1] = make_float4(bvhNoNegZero(b[3]), bvhNoNegZero(b[4]), bvhNoNegZero(b[5]), 0.0f);
}

__global__ void k_boxesFromTris(const TriRecord *tris, uint32_t n, float4 *bx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const TriRecord &t = tris[i];
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = bvhNoNegZero(fminf(fminf(t.v0[a], t.v1[a]), t.v2[a]));
        hi[a] = bvhNoNegZero(fmaxf(fmaxf(t.v0[a], t.v1[a]), t.v2[a]));
    }
    bx[2 * i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    bx[2 * i + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
}

__global__ void k_iota(uint32_t *ord, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ord[i] = i;
}

// Per segment: keys[12 * seg + 0..5] centroid lo / hi, + 6..11 box lo / hi.
__global__ void k_segInit(uint32_t *keys, uint32_t numSegs, uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *err = 0;
    if (i >= 12 * numSegs) return;
    const uint32_t f = i % 12;
    keys[i] = (f < 3 || (f >= 6 && f < 9)) ? 0xFFFFFFFFu : 0u;
}

// ---- wide phase ----
struct WTask {                 // a node of more than kBvhSmallPrims primitives
    uint32_t begin, count, depth, node;
    uint32_t firstBlock, numBlocks, pad[2];
    float clo[3], chi[3], pad2[2];
};
static_assert(sizeof(WTask) == 64, "task layout");
struct WBlock {
    uint32_t task, begin, end, pad;
};
struct WOut {
    int32_t axis, split;       // axis < 0: split at the middle of the range
    uint32_t leftCount, pad;
    uint32_t cent[2][6];       // children's centroid bounds (keys: lo xyz, hi xyz)
    uint32_t box[2][6];        // children's boxes
};
static_assert(sizeof(WOut) == 112, "output layout");

__global__ void __launch_bounds__(256) k_segBounds(const float4 *bx, const WBlock *blocks, uint32_t *keys) {
    const WBlock blk = blocks[blockIdx.x];
    float v[12];
    for (int a = 0; a < 3; ++a) {
        v[a] = v[6 + a] = FLT_MAX;
        v[3 + a] = v[9 + a] = -FLT_MAX;
    }
    for (uint32_t i = blk.begin + threadIdx.x; i < blk.end; i += blockDim.x) {
        const PrimBox b = loadBox(bx, i);
        for (int a = 0; a < 3; ++a) {
            const float c = centroid(b, a);
            v[a] = fminf(v[a], c);
            v[3 + a] = fmaxf(v[3 + a], c);
            v[6 + a] = fminf(v[6 + a], b.lo[a]);
            v[9 + a] = fmaxf(v[9 + a], b.hi[a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        v[a] = waveMin(v[a]);
        v[3 + a] = waveMax(v[3 + a]);
        v[6 + a] = waveMin(v[6 + a]);
        v[9 + a] = waveMax(v[9 + a]);
    }
    if ((threadIdx.x & 63) == 0) {
        uint32_t *k = keys + 12 * size_t(blk.task);
        for (int a = 0; a < 3; ++a) {
            atomicMin(&k[a], fkey(v[a]));
            atomicMax(&k[3 + a], fkey(v[3 + a]));
            atomicMin(&k[6 + a], fkey(v[6 + a]));
            atomicMax(&k[9 + a], fkey(v[9 + a]));
        }
    }
}

__global__ void k_levelClear(uint32_t *bins, WOut *outs, uint32_t numTasks) {
    const uint32_t t = blockIdx.x;
    if (t >= numTasks) return;
    for (int i = threadIdx.x; i < kBinWords; i += blockDim.x) {
        const int f = i & 7;
        bins[size_t(t) * kBinWords + i] = f < 3 ? 0xFFFFFFFFu : 0u;
    }
    if (threadIdx.x < 24) {
        const int s = threadIdx.x / 12, f = threadIdx.x % 12;
        uint32_t *w = f < 6 ? &outs[t].cent[s][f] : &outs[t].box[s][f - 6];
        *w = (f % 6) < 3 ? 0xFFFFFFFFu : 0u;
    }
}

__global__ void __launch_bounds__(256) k_bin(const float4 *bx, const uint32_t *ord, const WTask *tasks,
                                            const WBlock *blocks, uint32_t *bins) {
    __shared__ uint32_t sb[kLdsBinWords];
    for (int i = threadIdx.x; i < kLdsBinWords; i += blockDim.x) sb[i] = (i % 7) < 3 ? 0xFFFFFFFFu : 0u;
    __syncthreads();
    const WBlock blk = blocks[blockIdx.x];
    const WTask &T = tasks[blk.task];
    float clo[3], ext[3];
    for (int a = 0; a < 3; ++a) {
        clo[a] = T.clo[a];
        ext[a] = T.chi[a] - T.clo[a];
    }
    for (uint32_t i = blk.begin + threadIdx.x; i < blk.end; i += blockDim.x) {
        const PrimBox b = loadBox(bx, ord[i]);
        for (int a = 0; a < 3; ++a) {
            if (!(ext[a] > 0.0f)) continue;
            const int k = bvhBin(centroid(b, a), clo[a], ext[a]);
            uint32_t *w = sb + (a * kBvhBins + k) * 7;
            atomicMin(&w[0], fkey(b.lo[0]));
            atomicMin(&w[1], fkey(b.lo[1]));
            atomicMin(&w[2], fkey(b.lo[2]));
            atomicMax(&w[3], fkey(b.hi[0]));
            atomicMax(&w[4], fkey(b.hi[1]));
            atomicMax(&w[5], fkey(b.hi[2]));
            atomicAdd(&w[6], 1u);
        }
    }
    __syncthreads();
    uint32_t *g = bins + size_t(blk.task) * kBinWords;
    for (int i = threadIdx.x; i < kLdsBinWords; i += blockDim.x) {
        const int bin = i / 7, f = i % 7;
        if (sb[bin * 7 + 6] == 0) continue;
        uint32_t *w = g + bin * 8 + f;
        if (f < 3) atomicMin(w, sb[i]);
        else if (f < 6) atomicMax(w, sb[i]);
        else atomicAdd(w, sb[i]);
    }
}

__global__ void __launch_bounds__(64) k_split(const WTask *tasks, const uint32_t *bins, WOut *outs) {
    const uint32_t t = blockIdx.x;
    const int lane = threadIdx.x;
    const WTask &T = tasks[t];
    const uint32_t *b = bins + size_t(t) * kBinWords;
    unsigned long long key = ~0ull;
    if (T.depth < kBvhSahMaxDepth && lane < 3 * (kBvhBins - 1)) {
        const int a = lane / (kBvhBins - 1), k = lane % (kBvhBins - 1) + 1;
        uint32_t lc;
        if (T.chi[a] - T.clo[a] > 0.0f) key = sahCandidate<8>(b, a, k, &lc);
    }
    key = waveMinU64(key);
    if (lane == 0) {
        WOut &O = outs[t];
        if (key != ~0ull) {
            const int low = static_cast<int>(key & 0xFF);
            const int a = low >> 4, k = kBvhBins - 1 - (low & 15);
            uint32_t lc = 0;
            for (int j = 0; j < k; ++j) lc += b[(a * kBvhBins + j) * 8 + 6];
            O.axis = a;
            O.split = k;
            O.leftCount = lc;
        } else {
            O.axis = -1;
            O.split = 0;
            O.leftCount = T.count / 2;
        }
    }
}

__device__ inline bool goesLeft(const PrimBox &b, const WTask &T, const WOut &O, uint32_t pos) {
    if (O.axis < 0) return pos - T.begin < O.leftCount;
    const int a = O.axis;
    return bvhBin(centroid(b, a), T.clo[a], T.chi[a] - T.clo[a]) < O.split;
}

__global__ void __launch_bounds__(256) k_partCount(const float4 *bx, const uint32_t *ord, const WTask *tasks,
                                                  const WBlock *blocks, const WOut *outs, uint32_t *blockLeft) {
    const WBlock blk = blocks[blockIdx.x];
    const WTask &T = tasks[blk.task];
    const WOut &O = outs[blk.task];
    uint32_t cnt = 0;
    for (uint32_t base = blk.begin; base < blk.end; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        bool f = false;
        if (i < blk.end) f = goesLeft(loadBox(bx, ord[i]), T, O, i);
        cnt += __syncthreads_count(f);
    }
    if (threadIdx.x == 0) blockLeft[blockIdx.x] = cnt;
}

// Exclusive offsets of the blocks' left counts inside each node.
__global__ void k_partScan(const WTask *tasks, uint32_t numTasks, const uint32_t *blockLeft, uint32_t *blockOff) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= numTasks) return;
    uint32_t run = 0;
    for (uint32_t j = 0; j < tasks[t].numBlocks; ++j) {
        const uint32_t b = tasks[t].firstBlock + j;
        blockOff[b] = run;
        run += blockLeft[b];
    }
}

__global__ void __launch_bounds__(256) k_partWrite(const float4 *bx, const uint32_t *ord, uint32_t *tmp,
                                                  const WTask *tasks, const WBlock *blocks, WOut *outs,
                                                  const uint32_t *blockOff, uint32_t *err) {
    __shared__ uint32_t sl[4], sr[4];
    const WBlock blk = blocks[blockIdx.x];
    const WTask &T = tasks[blk.task];
    WOut &O = outs[blk.task];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t leftBefore = blockOff[blockIdx.x];
    uint32_t rightBefore = (blk.begin - T.begin) - leftBefore;
    float st[2][12];
    for (int s = 0; s < 2; ++s)
        for (int a = 0; a < 3; ++a) {
            st[s][a] = st[s][6 + a] = FLT_MAX;
            st[s][3 + a] = st[s][9 + a] = -FLT_MAX;
        }
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t base = blk.begin; base < blk.end; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < blk.end;
        uint32_t p = 0;
        PrimBox b = {};
        bool f = false;
        if (valid) {
            p = ord[i];
            b = loadBox(bx, p);
            f = goesLeft(b, T, O, i);
        }
        const unsigned long long lm = __ballot(valid && f), rm = __ballot(valid && !f);
        if (lane == 0) {
            sl[wave] = __popcll(lm);
            sr[wave] = __popcll(rm);
        }
        __syncthreads();
        uint32_t lo = 0, ro = 0, lt = 0, rt = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wave) {
                lo += sl[w];
                ro += sr[w];
            }
            lt += sl[w];
            rt += sr[w];
        }
        if (valid) {
            const uint32_t dst = f ? T.begin + leftBefore + lo + __popcll(lm & below)
                                   : T.begin + O.leftCount + rightBefore + ro + __popcll(rm & below);
            if (dst >= T.begin && dst < T.begin + T.count) tmp[dst] = p;
            else atomicOr(err, 1u);
            const int s = f ? 0 : 1;
            for (int a = 0; a < 3; ++a) {
                const float c = centroid(b, a);
                st[s][a] = fminf(st[s][a], c);
                st[s][3 + a] = fmaxf(st[s][3 + a], c);
                st[s][6 + a] = fminf(st[s][6 + a], b.lo[a]);
                st[s][9 + a] = fmaxf(st[s][9 + a], b.hi[a]);
            }
        }
        leftBefore += lt;
        rightBefore += rt;
        __syncthreads();
    }
    for (int s = 0; s < 2; ++s)
        for (int j = 0; j < 12; ++j) st[s][j] = (j % 6) < 3 ? waveMin(st[s][j]) : waveMax(st[s][j]);
    if (lane == 0) {
        for (int s = 0; s < 2; ++s)
            for (int j = 0; j < 6; ++j) {
                uint32_t *cw = &O.cent[s][j], *bw = &O.box[s][j];
                if (j < 3) {
                    atomicMin(cw, fkey(st[s][j]));
                    atomicMin(bw, fkey(st[s][6 + j]));
                } else {
                    atomicMax(cw, fkey(st[s][j]));
                    atomicMax(bw, fkey(st[s][6 + j]));
                }
            }
    }
}

__global__ void k_copyBack(const WBlock *blocks, const uint32_t *tmp, uint32_t *ord) {
    const WBlock blk = blocks[blockIdx.x];
    for (uint32_t i = blk.begin + threadIdx.x; i < blk.end; i += blockDim.x) ord[i] = tmp[i];
}

// ---- subtrees ----
struct STask {
    uint32_t begin, count, node, depth, resBase, pad[3];
};

struct SubtreeLds {
    float lo[3][kBvhSmallPrims], hi[3][kBvhSmallPrims];
    uint32_t gid[kBvhSmallPrims];
    uint16_t perm[kBvhSmallPrims], perm2[kBvhSmallPrims];
    uint32_t bins[4][kLdsBinWords];
    uint32_t front[2][kMaxFront][3];   // (b | e << 16, parent, slot | depth << 8)
    uint32_t frontCount[2];
};

__device__ inline PrimBox ldsBox(const SubtreeLds &L, uint32_t l) {
    return {{L.lo[0][l], L.lo[1][l], L.lo[2][l]}, {L.hi[0][l], L.hi[1][l], L.hi[2][l]}};
}

// Splits one node of a subtree; called by all lanes of one wave with the same entry.
__device__ void splitSubtreeNode(SubtreeLds &L, const STask &S, const uint32_t *entry, int next, uint32_t maxLeaf,
                                 BVHNode *nodes, uint32_t *err) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t b = entry[0] & 0xFFFF, e = entry[0] >> 16, parent = entry[1];
    const uint32_t slot = entry[2] & 0xFF, depth = entry[2] >> 8;
    const uint32_t n = e - b;
    // centroid bounds
    float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = b + lane; i < e; i += 64) {
        const PrimBox bb = ldsBox(L, L.perm[i]);
        for (int a = 0; a < 3; ++a) {
            clo[a] = fminf(clo[a], centroid(bb, a));
            chi[a] = fmaxf(chi[a], centroid(bb, a));
        }
    }
    for (int a = 0; a < 3; ++a) {
        clo[a] = waveMin(clo[a]);
        chi[a] = waveMax(chi[a]);
    }
    // 0: SAH boundary, 1: middle of the range, 2: median of the widest centroid axis
    int mode = 1, axis = 0, split = 0;
    uint32_t leftCount = n / 2;
    if (depth < kBvhSahMaxDepth) {
        uint32_t *bins = L.bins[wave];
        for (int i = lane; i < kLdsBinWords; i += 64) bins[i] = (i % 7) < 3 ? 0xFFFFFFFFu : 0u;
        waveSync();
        for (uint32_t i = b + lane; i < e; i += 64) {
            const PrimBox bb = ldsBox(L, L.perm[i]);
            for (int a = 0; a < 3; ++a) {
                const float ext = chi[a] - clo[a];
                if (!(ext > 0.0f)) continue;
                uint32_t *w = bins + (a * kBvhBins + bvhBin(centroid(bb, a), clo[a], ext)) * 7;
                atomicMin(&w[0], fkey(bb.lo[0]));
                atomicMin(&w[1], fkey(bb.lo[1]));
                atomicMin(&w[2], fkey(bb.lo[2]));
                atomicMax(&w[3], fkey(bb.hi[0]));
                atomicMax(&w[4], fkey(bb.hi[1]));
                atomicMax(&w[5], fkey(bb.hi[2]));
                atomicAdd(&w[6], 1u);
            }
        }
        waveSync();
        unsigned long long key = ~0ull;
        if (lane < 3 * (kBvhBins - 1)) {
            const int a = lane / (kBvhBins - 1), k = lane % (kBvhBins - 1) + 1;
            uint32_t lc;
            if (chi[a] - clo[a] > 0.0f) key = sahCandidate<7>(bins, a, k, &lc);
        }
        key = waveMinU64(key);
        if (key != ~0ull) {
            const int low = static_cast<int>(key & 0xFF);
            axis = low >> 4;
            split = kBvhBins - 1 - (low & 15);
            leftCount = 0;
            for (int j = 0; j < split; ++j) leftCount += bins[(axis * kBvhBins + j) * 7 + 6];
            mode = 0;
        }
        waveSync();
    } else if (n <= kBvhSmallPrims) {
        mode = 2;
        for (int a = 1; a < 3; ++a)
            if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
    }
    // reorder [b, e)
    if (mode == 0) {
        const float ext = chi[axis] - clo[axis];
        const unsigned long long below = (1ull << lane) - 1ull;
        uint32_t lrun = 0, rrun = 0;
        for (uint32_t base = b; base < e; base += 64) {
            const uint32_t i = base + lane;
            const bool valid = i < e;
            uint32_t l = 0;
            bool f = false;
            if (valid) {
                l = L.perm[i];
                f = bvhBin(centroid(ldsBox(L, l), axis), clo[axis], ext) < split;
            }
            const unsigned long long lm = __ballot(valid && f), rm = __ballot(valid && !f);
            if (valid) {
                const uint32_t dst = f ? b + lrun + __popcll(lm & below) : b + leftCount + rrun + __popcll(rm & below);
                if (dst < e) L.perm2[dst] = static_cast<uint16_t>(l);
                else atomicOr(err, 2u);
            }
            lrun += __popcll(lm);
            rrun += __popcll(rm);
        }
        if (lrun != leftCount && lane == 0) atomicOr(err, 2u);
    } else if (mode == 2) {
        for (uint32_t i = b + lane; i < e; i += 64) {
            const float ki = centroid(ldsBox(L, L.perm[i]), axis);
            uint32_t rank = 0;
            for (uint32_t j = b; j < e; ++j) {
                const float kj = centroid(ldsBox(L, L.perm[j]), axis);
                rank += (kj < ki || (kj == ki && j < i)) ? 1u : 0u;
            }
            L.perm2[b + rank] = L.perm[i];
        }
    }
    if (mode != 1) {
        waveSync();
        for (uint32_t i = b + lane; i < e; i += 64) L.perm[i] = L.perm2[i];
        waveSync();
    }
    const uint32_t mid = b + leftCount;
    // children's boxes
    float cb[2][6];
    for (int s = 0; s < 2; ++s)
        for (int a = 0; a < 3; ++a) {
            cb[s][a] = FLT_MAX;
            cb[s][3 + a] = -FLT_MAX;
        }
    for (uint32_t i = b + lane; i < e; i += 64) {
        const PrimBox bb = ldsBox(L, L.perm[i]);
        const int s = i < mid ? 0 : 1;
        for (int a = 0; a < 3; ++a) {
            cb[s][a] = fminf(cb[s][a], bb.lo[a]);
            cb[s][3 + a] = fmaxf(cb[s][3 + a], bb.hi[a]);
        }
    }
    for (int s = 0; s < 2; ++s)
        for (int a = 0; a < 3; ++a) {
            cb[s][a] = waveMin(cb[s][a]);
            cb[s][3 + a] = waveMax(cb[s][3 + a]);
        }
    if (mid <= b || mid >= e) {   // never for a correct split; the node is dropped
        if (lane == 0) atomicOr(err, 4u);
        return;
    }
    const uint32_t self = parent == kNoParent ? S.node : S.resBase + mid - 1;
    if (lane == 0) {
        const uint32_t rb[2] = {b, mid}, re[2] = {mid, e};
        int32_t codes[2];
        uint32_t inner = 0;
        for (int s = 0; s < 2; ++s) {
            const uint32_t cnt = re[s] - rb[s];
            codes[s] = cnt <= maxLeaf ? static_cast<int32_t>(~(((S.begin + rb[s]) << kLeafShift) | (cnt - 1))) : 0;
            if (cnt > maxLeaf) ++inner;
        }
        BVHNode N;
        for (int a = 0; a < 3; ++a) {
            N.lo0[a] = cb[0][a];
            N.hi0[a] = cb[0][3 + a];
            N.lo1[a] = cb[1][a];
            N.hi1[a] = cb[1][3 + a];
        }
        N.child[0] = codes[0];
        N.child[1] = codes[1];
        N.pad[0] = N.pad[1] = 0;
        nodes[self] = N;
        if (parent != kNoParent) nodes[parent].child[slot] = static_cast<int32_t>(self);
        if (inner) {
            const uint32_t at = atomicAdd(&L.frontCount[next], inner);
            uint32_t k = at;
            for (int s = 0; s < 2; ++s) {
                if (re[s] - rb[s] <= maxLeaf) continue;
                if (k < kMaxFront) {
                    L.front[next][k][0] = rb[s] | (re[s] << 16);
                    L.front[next][k][1] = self;
                    L.front[next][k][2] = static_cast<uint32_t>(s) | ((depth + 1) << 8);
                } else {
                    atomicOr(err, 8u);
                }
                ++k;
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_subtrees(const float4 *bx, uint32_t *ord, const STask *tasks,
                                                 uint32_t maxLeaf, BVHNode *nodes, uint32_t *err) {
    __shared__ SubtreeLds L;
    const STask S = tasks[blockIdx.x];
    const int wave = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < S.count; i += blockDim.x) {
        const uint32_t p = ord[S.begin + i];
        const PrimBox b = loadBox(bx, p);
        L.gid[i] = p;
        L.perm[i] = static_cast<uint16_t>(i);
        for (int a = 0; a < 3; ++a) {
            L.lo[a][i] = b.lo[a];
            L.hi[a][i] = b.hi[a];
        }
    }
    if (threadIdx.x == 0) {
        L.front[0][0][0] = 0u | (S.count << 16);
        L.front[0][0][1] = kNoParent;
        L.front[0][0][2] = S.depth << 8;
        L.frontCount[0] = 1;
        L.frontCount[1] = 0;
    }
    __syncthreads();
    int cur = 0;
    for (;;) {
        const uint32_t nf = min(L.frontCount[cur], kMaxFront);
        if (nf == 0) break;
        for (uint32_t j = wave; j < nf; j += 4) {
            uint32_t entry[3] = {L.front[cur][j][0], L.front[cur][j][1], L.front[cur][j][2]};
            splitSubtreeNode(L, S, entry, cur ^ 1, maxLeaf, nodes, err);
        }
        __syncthreads();
        if (threadIdx.x == 0) L.frontCount[cur] = 0;
        cur ^= 1;
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < S.count; i += blockDim.x) ord[S.begin + i] = L.gid[L.perm[i]];
}

// ---- collapse to 4-wide nodes ----
struct CSlot {
    float lo[3], hi[3];
    int32_t code;
    uint32_t pad;
};

__device__ inline CSlot slotOf(const BVHNode &N, int s) {
    CSlot r;
    const float *lo = s == 0 ? N.lo0 : N.lo1, *hi = s == 0 ? N.hi0 : N.hi1;
    for (int a = 0; a < 3; ++a) {
        r.lo[a] = lo[a];
        r.hi[a] = hi[a];
    }
    r.code = N.child[s];
    r.pad = 0;
    return r;
}

__global__ void k_collapseSlots(const BVHNode *nodes, const uint32_t *frontier, uint32_t F, CSlot *slotsOut,
                                uint32_t *innerCount) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    const BVHNode N = nodes[frontier[i]];
    CSlot slots[4];
    int n = 0;
    for (int s = 0; s < 2; ++s) {
        const CSlot c = slotOf(N, s);
        if (c.code != 0) slots[n++] = c;
    }
    while (n < 4) {
        int best = -1;
        float bestArea = -1.0f;
        for (int k = 0; k < n; ++k)
            if (slots[k].code > 0 && bvhArea(slots[k].lo, slots[k].hi) > bestArea) {
                best = k;
                bestArea = bvhArea(slots[k].lo, slots[k].hi);
            }
        if (best < 0) break;
        const BVHNode M = nodes[slots[best].code];
        const CSlot a = slotOf(M, 0), c = slotOf(M, 1);
        CSlot tmp[4];
        int m = 0;
        for (int k = 0; k < n; ++k) {
            if (k == best) {
                if (a.code != 0) tmp[m++] = a;
                if (c.code != 0) tmp[m++] = c;
            } else {
                tmp[m++] = slots[k];
            }
        }
        for (int k = 0; k < m; ++k) slots[k] = tmp[k];
        n = m;
    }
    uint32_t inner = 0;
    for (int k = 0; k < 4; ++k) {
        CSlot s = k < n ? slots[k] : CSlot{{0, 0, 0}, {0, 0, 0}, 0, 0};
        if (s.code > 0) ++inner;
        slotsOut[4 * size_t(i) + k] = s;
    }
    innerCount[i] = inner;
}

// Segment s's root is node s: inner child codes count from it, leaf codes from the segment's
// first primitive.
__global__ void k_collapseWrite(const CSlot *slots, const uint32_t *innerOff, const uint32_t *frontSeg, uint32_t F,
                                uint32_t base, BVH4QNode *out, uint32_t *nextFrontier, uint32_t *nextSeg,
                                const uint32_t *segFirst, const uint32_t *leafRemap, const uint32_t *ord) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    const uint32_t seg = frontSeg[i];
    BVH4Node f;
    uint32_t r = innerOff[i];
    for (int k = 0; k < 4; ++k) {
        const CSlot s = slots[4 * size_t(i) + k];
        int32_t code = s.code;
        if (code > 0) {
            nextFrontier[r] = static_cast<uint32_t>(code);
            nextSeg[r] = seg;
            code = static_cast<int32_t>(base + F + r - seg);
            ++r;
        } else if (code < 0) {
            const uint32_t u = static_cast<uint32_t>(~code);
            const uint32_t first = u >> kLeafShift;
            code = leafRemap ? static_cast<int32_t>(~(leafRemap[ord[first]] << kLeafShift))
                             : static_cast<int32_t>(~(((first - segFirst[seg]) << kLeafShift) | (u & ((1u << kLeafShift) - 1))));
        }
        f.loX[k] = s.lo[0];
        f.loY[k] = s.lo[1];
        f.loZ[k] = s.lo[2];
        f.hiX[k] = s.hi[0];
        f.hiY[k] = s.hi[1];
        f.hiZ[k] = s.hi[2];
        f.child[k] = code;
        f.pad[k] = 0;
    }
    out[base + i] = quantizeBvh4Node(f);
}

// ---- exclusive scan of uint32 (1024 values per block) ----
__device__ inline uint32_t blockExclusiveScan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t sw[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sw[wave] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int w = 0; w < 4; ++w) {
        if (w < wave) off += sw[w];
        tot += sw[w];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

__global__ void __launch_bounds__(256) k_scanLocal(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *sums) {
    const uint32_t i0 = blockIdx.x * 1024 + threadIdx.x * 4;
    uint32_t v[4], s = 0;
    for (int j = 0; j < 4; ++j) {
        v[j] = i0 + j < n ? in[i0 + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t run = blockExclusiveScan(s, &total);
    for (int j = 0; j < 4; ++j) {
        if (i0 + j < n) out[i0 + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(256) k_scanSums(uint32_t *sums, uint32_t nb, uint32_t *total) {
    uint32_t carry = 0;
    for (uint32_t c = 0; c < nb; c += 1024) {
        const uint32_t i0 = c + threadIdx.x * 4;
        uint32_t v[4], s = 0;
        for (int j = 0; j < 4; ++j) {
            v[j] = i0 + j < nb ? sums[i0 + j] : 0u;
            s += v[j];
        }
        uint32_t t;
        uint32_t run = carry + blockExclusiveScan(s, &t);
        for (int j = 0; j < 4; ++j) {
            if (i0 + j < nb) sums[i0 + j] = run;
            run += v[j];
        }
        carry += t;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scanAdd(uint32_t *out, uint32_t n, const uint32_t *sums) {
    const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
    const uint32_t add = sums[blockIdx.x];
    for (int j = 0; j < 4; ++j)
        if (i + 256 * j < n) out[i + 256 * j] += add;
}

__global__ void k_gatherTris(const TriRecord *src, const uint32_t *order, uint32_t n, TriRecord *dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[order[i]];
}

uint32_t blocksFor(size_t n, uint32_t per) { return static_cast<uint32_t>((n + per - 1) / per); }

} // namespace

// A scratch buffer of the builder: device memory it owns, or (arena mode) a range of one
// region of device memory it was given, carved at 256-byte alignment.
struct ScratchBuf {
    DeviceBuffer own;
    void *ptr = nullptr;
};

struct GpuBvhBuilder::Impl {
    ScratchBuf boxes, ord, tmp, tasks, blocks, bins, outs, blockLeft, blockOff, misc, errBuf, segFirstDev, bvh2, small,
        frontA, frontB, segA, segB, slots, inner, innerOff, scanSums, scanTotal;
    // arena mode: every buffer comes from [arena, arena + arenaBytes); nothing is allocated
    uint8_t *arena = nullptr;
    size_t arenaBytes = 0, arenaTop = 0;
    // arena mode: the primitive order is written straight into this device array
    uint32_t *orderOut = nullptr;

    void take(ScratchBuf &b, size_t bytes) {
        if (!arena) {
            b.own.reserve(bytes);
            b.ptr = b.own.ptr;
            return;
        }
        const size_t at = (arenaTop + 255) & ~size_t(255);
        if (at + bytes > arenaBytes) throw std::runtime_error("BVH build: scratch region exhausted");
        b.ptr = arena + at;
        arenaTop = at + bytes;
    }
    void put(ScratchBuf &b, const void *src, size_t bytes, hipStream_t s) {
        take(b, bytes);
        if (bytes) VLRA_HIP(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, s));
    }

    // scanSums / scanTotal must hold blocksFor(n, 1024) + 1 words
    void scan(const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t s, uint32_t *hostTotal) {
        const uint32_t nb = std::max(1u, blocksFor(n, 1024));
        auto *sums = static_cast<uint32_t *>(scanSums.ptr);
        auto *tot = static_cast<uint32_t *>(scanTotal.ptr);
        k_scanLocal<<<nb, 256, 0, s>>>(in, out, n, sums);
        k_scanSums<<<1, 256, 0, s>>>(sums, nb, tot);
        k_scanAdd<<<nb, 256, 0, s>>>(out, n, sums);
        VLRA_HIP(hipMemcpyAsync(hostTotal, tot, 4, hipMemcpyDeviceToHost, s));
        VLRA_HIP(hipStreamSynchronize(s));
    }

    // Builds one tree per segment [segFirst[k], segFirst[k] + segCount[k]) of the primitives
    // (consecutive, covering [0, n)); segment k's root is node k.
    uint32_t run(const std::vector<uint32_t> &segFirst, const std::vector<uint32_t> &segCount, uint32_t maxLeaf,
                 const uint32_t *leafRemap, BVH4QNode *out, uint32_t cap, hipStream_t s, BvhBuildTimes *info) {
        const uint32_t numSegs = static_cast<uint32_t>(segFirst.size());
        uint32_t n = 0;
        for (uint32_t k = 0; k < numSegs; ++k) {
            if (segFirst[k] != n || segCount[k] == 0) throw std::runtime_error("BVH build: bad segments");
            n += segCount[k];
        }
        const float4 *bx = static_cast<const float4 *>(boxes.ptr);
        if (orderOut)
            ord.ptr = orderOut;
        else
            take(ord, size_t(n) * 4);
        take(tmp, size_t(n) * 4);
        take(misc, size_t(numSegs) * 12 * 4);
        take(errBuf, 4);
        put(segFirstDev, segFirst.data(), size_t(numSegs) * 4, s);
        // buffers of one wide level, and after the wide phase those of the collapse, start here
        const size_t levelMark = arenaTop;
        auto *ordp = static_cast<uint32_t *>(ord.ptr);
        auto *keys = static_cast<uint32_t *>(misc.ptr);
        auto *err = static_cast<uint32_t *>(errBuf.ptr);
        k_iota<<<blocksFor(n, 256), 256, 0, s>>>(ordp, n);
        k_segInit<<<blocksFor(size_t(numSegs) * 12, 256), 256, 0, s>>>(keys, numSegs, err);
        std::vector<WBlock> blk;
        for (uint32_t k = 0; k < numSegs; ++k)
            for (uint32_t j = 0; j < blocksFor(segCount[k], kBinChunk); ++j)
                blk.push_back({k, segFirst[k] + j * kBinChunk, std::min(segFirst[k] + segCount[k], segFirst[k] + (j + 1) * kBinChunk), 0});
        put(blocks, blk.data(), blk.size() * sizeof(WBlock), s);
        k_segBounds<<<static_cast<uint32_t>(blk.size()), 256, 0, s>>>(bx, static_cast<const WBlock *>(blocks.ptr), keys);
        std::vector<uint32_t> rootKeys(size_t(numSegs) * 12);
        VLRA_HIP(hipMemcpyAsync(rootKeys.data(), keys, rootKeys.size() * 4, hipMemcpyDeviceToHost, s));
        VLRA_HIP(hipStreamSynchronize(s));

        std::vector<BVHNode> wideNodes(numSegs);
        std::memset(wideNodes.data(), 0, wideNodes.size() * sizeof(BVHNode));
        std::vector<STask> smallTasks;
        std::vector<WTask> level;
        auto leafCode = [](uint32_t first, uint32_t count) {
            return static_cast<int32_t>(~((first << kLeafShift) | (count - 1)));
        };
        for (uint32_t k = 0; k < numSegs; ++k) {
            const uint32_t *rk = rootKeys.data() + 12 * size_t(k);
            if (segCount[k] <= maxLeaf) {
                BVHNode &R = wideNodes[k];
                for (int a = 0; a < 3; ++a) {
                    R.lo0[a] = hostKeyf(rk[6 + a]);
                    R.hi0[a] = hostKeyf(rk[9 + a]);
                }
                R.child[0] = leafCode(segFirst[k], segCount[k]);
            } else if (segCount[k] <= kBvhSmallPrims) {
                smallTasks.push_back({segFirst[k], segCount[k], k, 1, 0, {0, 0, 0}});
            } else {
                WTask t = {};
                t.begin = segFirst[k];
                t.count = segCount[k];
                t.depth = 1;
                t.node = k;
                for (int a = 0; a < 3; ++a) {
                    t.clo[a] = hostKeyf(rk[a]);
                    t.chi[a] = hostKeyf(rk[3 + a]);
                }
                level.push_back(t);
            }
        }
        // wide phase
        std::vector<WOut> outsHost;
        uint32_t levels = 0;
        while (!level.empty()) {
            ++levels;
            blk.clear();
            for (uint32_t t = 0; t < level.size(); ++t) {
                WTask &T = level[t];
                T.firstBlock = static_cast<uint32_t>(blk.size());
                T.numBlocks = blocksFor(T.count, kBinChunk);
                for (uint32_t j = 0; j < T.numBlocks; ++j)
                    blk.push_back({t, T.begin + j * kBinChunk, std::min(T.begin + T.count, T.begin + (j + 1) * kBinChunk), 0});
            }
            const uint32_t nt = static_cast<uint32_t>(level.size()), nb = static_cast<uint32_t>(blk.size());
            arenaTop = levelMark;
            take(tasks, nt * sizeof(WTask));
            take(blocks, nb * sizeof(WBlock));
            take(bins, size_t(nt) * kBinWords * 4);
            take(outs, nt * sizeof(WOut));
            take(blockLeft, nb * 4);
            take(blockOff, nb * 4);
            VLRA_HIP(hipMemcpyAsync(tasks.ptr, level.data(), nt * sizeof(WTask), hipMemcpyHostToDevice, s));
            VLRA_HIP(hipMemcpyAsync(blocks.ptr, blk.data(), nb * sizeof(WBlock), hipMemcpyHostToDevice, s));
            auto *tp = static_cast<const WTask *>(tasks.ptr);
            auto *bp = static_cast<const WBlock *>(blocks.ptr);
            auto *op = static_cast<WOut *>(outs.ptr);
            auto *binp = static_cast<uint32_t *>(bins.ptr);
            auto *bl = static_cast<uint32_t *>(blockLeft.ptr);
            auto *bo = static_cast<uint32_t *>(blockOff.ptr);
            k_levelClear<<<nt, 256, 0, s>>>(binp, op, nt);
            k_bin<<<nb, 256, 0, s>>>(bx, ordp, tp, bp, binp);
            k_split<<<nt, 64, 0, s>>>(tp, binp, op);
            k_partCount<<<nb, 256, 0, s>>>(bx, ordp, tp, bp, op, bl);
            k_partScan<<<blocksFor(nt, 256), 256, 0, s>>>(tp, nt, bl, bo);
            k_partWrite<<<nb, 256, 0, s>>>(bx, ordp, static_cast<uint32_t *>(tmp.ptr), tp, bp, op, bo, err);
            k_copyBack<<<nb, 256, 0, s>>>(bp, static_cast<const uint32_t *>(tmp.ptr), ordp);
            outsHost.resize(nt);
            VLRA_HIP(hipMemcpyAsync(outsHost.data(), op, nt * sizeof(WOut), hipMemcpyDeviceToHost, s));
            VLRA_HIP(hipStreamSynchronize(s));
            std::vector<WTask> nextLevel;
            for (uint32_t t = 0; t < nt; ++t) {
                const WTask &T = level[t];
                const WOut &O = outsHost[t];
                const uint32_t cnt[2] = {O.leftCount, T.count - O.leftCount};
                if (cnt[0] == 0 || cnt[1] == 0) throw std::runtime_error("BVH build: empty child");
                const uint32_t first[2] = {T.begin, T.begin + O.leftCount};
                BVHNode N;
                std::memset(&N, 0, sizeof(N));
                for (int a = 0; a < 3; ++a) {
                    N.lo0[a] = hostKeyf(O.box[0][a]);
                    N.hi0[a] = hostKeyf(O.box[0][3 + a]);
                    N.lo1[a] = hostKeyf(O.box[1][a]);
                    N.hi1[a] = hostKeyf(O.box[1][3 + a]);
                }
                for (int c = 0; c < 2; ++c) {
                    if (cnt[c] <= maxLeaf) {
                        N.child[c] = leafCode(first[c], cnt[c]);
                        continue;
                    }
                    const uint32_t id = static_cast<uint32_t>(wideNodes.size());
                    wideNodes.emplace_back();
                    std::memset(&wideNodes.back(), 0, sizeof(BVHNode));
                    N.child[c] = static_cast<int32_t>(id);
                    if (cnt[c] <= kBvhSmallPrims) {
                        smallTasks.push_back({first[c], cnt[c], id, T.depth + 1, 0, {0, 0, 0}});
                    } else {
                        WTask u = {};
                        u.begin = first[c];
                        u.count = cnt[c];
                        u.depth = T.depth + 1;
                        u.node = id;
                        for (int a = 0; a < 3; ++a) {
                            u.clo[a] = hostKeyf(O.cent[c][a]);
                            u.chi[a] = hostKeyf(O.cent[c][3 + a]);
                        }
                        nextLevel.push_back(u);
                    }
                }
                wideNodes[T.node] = N;
            }
            level.swap(nextLevel);
        }
        // subtrees: node ids base + (split position - 1) inside each
        uint32_t total = static_cast<uint32_t>(wideNodes.size());
        for (STask &t : smallTasks) {
            t.resBase = total;
            total += t.count - 1;
        }
        arenaTop = levelMark;
        take(bvh2, size_t(total) * sizeof(BVHNode));
        auto *b2 = static_cast<BVHNode *>(bvh2.ptr);
        VLRA_HIP(hipMemcpyAsync(b2, wideNodes.data(), wideNodes.size() * sizeof(BVHNode), hipMemcpyHostToDevice, s));
        if (!smallTasks.empty()) {
            put(small, smallTasks.data(), smallTasks.size() * sizeof(STask), s);
            k_subtrees<<<static_cast<uint32_t>(smallTasks.size()), 256, 0, s>>>(bx, ordp, static_cast<const STask *>(small.ptr),
                                                                                 maxLeaf, b2, err);
        }
        // collapse: level 0 holds the segments' roots
        take(frontA, size_t(total) * 4);
        take(frontB, size_t(total) * 4);
        take(segA, size_t(total) * 4);
        take(segB, size_t(total) * 4);
        take(slots, size_t(total) * 4 * sizeof(CSlot));
        take(inner, size_t(total) * 4);
        take(innerOff, size_t(total) * 4);
        take(scanSums, (size_t(std::max(1u, blocksFor(total, 1024))) + 1) * 4);
        take(scanTotal, 4);
        std::vector<uint32_t> roots(numSegs);
        for (uint32_t k = 0; k < numSegs; ++k) roots[k] = k;
        VLRA_HIP(hipMemcpyAsync(frontA.ptr, roots.data(), size_t(numSegs) * 4, hipMemcpyHostToDevice, s));
        VLRA_HIP(hipMemcpyAsync(segA.ptr, roots.data(), size_t(numSegs) * 4, hipMemcpyHostToDevice, s));
        uint32_t F = numSegs, base = 0, clevels = 0;
        if (F > cap) throw std::runtime_error("BVH build: node capacity exceeded");
        auto *fa = static_cast<uint32_t *>(frontA.ptr), *fb = static_cast<uint32_t *>(frontB.ptr);
        auto *sa = static_cast<uint32_t *>(segA.ptr), *sb = static_cast<uint32_t *>(segB.ptr);
        while (F > 0) {
            ++clevels;
            k_collapseSlots<<<blocksFor(F, 256), 256, 0, s>>>(b2, fa, F, static_cast<CSlot *>(slots.ptr),
                                                              static_cast<uint32_t *>(inner.ptr));
            uint32_t next = 0;
            scan(static_cast<const uint32_t *>(inner.ptr), static_cast<uint32_t *>(innerOff.ptr), F, s, &next);
            if (size_t(base) + F + next > cap) throw std::runtime_error("BVH build: node capacity exceeded");
            k_collapseWrite<<<blocksFor(F, 256), 256, 0, s>>>(static_cast<const CSlot *>(slots.ptr),
                                                              static_cast<const uint32_t *>(innerOff.ptr), sa, F, base,
                                                              out, fb, sb, static_cast<const uint32_t *>(segFirstDev.ptr),
                                                              leafRemap, ordp);
            base += F;
            F = next;
            std::swap(fa, fb);
            std::swap(sa, sb);
        }
        uint32_t errHost = 0;
        VLRA_HIP(hipMemcpyAsync(&errHost, err, 4, hipMemcpyDeviceToHost, s));
        VLRA_HIP(hipStreamSynchronize(s));
        if (errHost) throw std::runtime_error("BVH build: inconsistent partition (code " + std::to_string(errHost) + ")");
        if (info) {
            info->largeLevels = levels;
            info->smallTasks = static_cast<uint32_t>(smallTasks.size());
            info->collapseLevels = clevels;
        }
        return base;
    }
};

GpuBvhBuilder::GpuBvhBuilder() : impl_(new Impl) {}
GpuBvhBuilder::~GpuBvhBuilder() { delete impl_; }

const uint32_t *GpuBvhBuilder::order() const { return static_cast<const uint32_t *>(impl_->ord.ptr); }

static void checkArgs(uint32_t n, uint32_t maxLeaf, const uint32_t *leafRemap) {
    if (n >= kBvhMaxPrims) throw std::runtime_error("BVH build: too many primitives");
    if (maxLeaf < 1 || maxLeaf > kMaxLeafPrims) throw std::runtime_error("BVH build: leaf size out of range");
    if (leafRemap && maxLeaf != 1) throw std::runtime_error("BVH build: leaf remapping needs one primitive per leaf");
}

uint32_t GpuBvhBuilder::build(const float *boxes, uint32_t n, uint32_t maxLeaf, const uint32_t *leafRemap,
                              BVH4QNode *nodes, hipStream_t s, BvhBuildTimes *info) {
    checkArgs(n, maxLeaf, leafRemap);
    if (n == 0) return 0;
    impl_->take(impl_->boxes, size_t(n) * 32);
    k_boxesFromFloats<<<blocksFor(n, 256), 256, 0, s>>>(boxes, n, static_cast<float4 *>(impl_->boxes.ptr));
    return impl_->run({0u}, {n}, maxLeaf, leafRemap, nodes, nodeCapacity(n), s, info);
}

uint32_t GpuBvhBuilder::buildTriangles(const TriRecord *tris, const std::vector<uint32_t> &segFirst,
                                       const std::vector<uint32_t> &segCount, uint32_t maxLeaf, BVH4QNode *nodes,
                                       hipStream_t s, BvhBuildTimes *info) {
    uint64_t n = 0, cap = 0;
    for (uint32_t c : segCount) {
        n += c;
        cap += nodeCapacity(c);
    }
    if (n >= kBvhMaxPrims) throw std::runtime_error("BVH build: too many primitives");
    checkArgs(static_cast<uint32_t>(n), maxLeaf, nullptr);
    if (segFirst.empty()) return 0;
    impl_->take(impl_->boxes, size_t(n) * 32);
    k_boxesFromTris<<<blocksFor(n, 256), 256, 0, s>>>(tris, static_cast<uint32_t>(n),
                                                      static_cast<float4 *>(impl_->boxes.ptr));
    return impl_->run(segFirst, segCount, maxLeaf, nullptr, nodes, static_cast<uint32_t>(cap), s, info);
}

void gatherTriangles(const TriRecord *src, const uint32_t *order, uint32_t n, TriRecord *dst, hipStream_t s) {
    if (n) k_gatherTris<<<blocksFor(n, 256), 256, 0, s>>>(src, order, n, dst);
}

// ---- builds that allocate nothing ----
// Scratch of a single-tree build over n <= kNoAllocMaxPrims primitives, taken in the order
// run() takes it: the boxes, the partition buffer, the segment keys, error word and segment
// table; then the larger of one wide level's buffers and the collapse's.  Every buffer may
// cost 255 bytes of alignment.
constexpr size_t alignUp(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t noAllocScratchBytes(size_t n) {
    // wide levels: tasks hold more than kBvhSmallPrims primitives each, blocks cover kBinChunk
    const size_t tasks = n / (kBvhSmallPrims + 1), blks = n / kBinChunk + 1 + tasks;
    const size_t level = alignUp(tasks * sizeof(WTask)) + alignUp(blks * sizeof(WBlock)) +
                         alignUp(tasks * kBinWords * 4) + alignUp(tasks * sizeof(WOut)) + 2 * alignUp(blks * 4);
    // collapse: at most n - 1 binary nodes, subtree tasks of two or more primitives each
    const size_t total = n > 1 ? n - 1 : 1;
    const size_t collapse = alignUp(total * sizeof(BVHNode)) + alignUp((n / 2) * sizeof(STask)) +
                            6 * alignUp(total * 4) + alignUp(total * 4 * sizeof(CSlot)) +
                            alignUp(((total + 1023) / 1024 + 1) * 4) + alignUp(4);
    const size_t fixed = alignUp(n * 32) + alignUp(n * 4) + 3 * alignUp(12 * 4);
    return fixed + (level > collapse ? level : collapse) + 256;
}
constexpr size_t kNoAllocScratchBytes = noAllocScratchBytes(kNoAllocMaxPrims);
static_assert(kNoAllocScratchBytes < (size_t(40) << 20), "the no-alloc scratch stays modest");

// One region per GPU for the builds that allocate nothing, held until the process ends.  The
// name is a host-side stand-in: host code uses the device address that hipGetSymbolAddress
// returns for the current GPU, never the name itself.
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex gScratchMutex;
uint8_t *gScratchPtr[kNoAllocMaxDevices] = {};

// The current GPU's copy of `scratch`; looked up once per GPU, with that GPU current.
uint8_t *scratchOfCurrentDevice() {
    int dev = -1;
    VLRA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) throw std::runtime_error("BVH build: device index out of range");
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchPtr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("BVH build: no address for the scratch region: ") + hipGetErrorString(e));
        gScratchPtr[dev] = static_cast<uint8_t *>(p);
    }
    return gScratchPtr[dev];
}
}  // namespace

uint32_t buildBvh4NoAlloc(const float *boxes, uint32_t n, uint32_t maxLeaf, const uint32_t *leafRemap,
                          BVH4QNode *nodes, uint32_t *order, hipStream_t s) {
    checkArgs(n, maxLeaf, leafRemap);
    if (n > kNoAllocMaxPrims) throw std::runtime_error("BVH build: more primitives than the no-alloc maximum");
    if (n == 0) return 0;
    uint8_t *region = scratchOfCurrentDevice();
    GpuBvhBuilder::Impl impl;
    impl.arena = region;
    impl.arenaBytes = kNoAllocScratchBytes;
    impl.orderOut = order;
    impl.take(impl.boxes, size_t(n) * 32);
    k_boxesFromFloats<<<blocksFor(n, 256), 256, 0, s>>>(boxes, n, static_cast<float4 *>(impl.boxes.ptr));
    return impl.run({0u}, {n}, maxLeaf, leafRemap, nodes, GpuBvhBuilder::nodeCapacity(n), s, nullptr);
}

} // namespace vlra
