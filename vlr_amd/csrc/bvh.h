// Acceleration-structure builders (replace the OptiX GAS / IAS builds of
// libVLR/scene.cpp:1149-1202, 1319-1329, configured PreferFastTrace at scene.cpp:1049).
//
// One algorithm, two implementations that give the same tree bit for bit:
//   bvh_host.cpp   sequential recursion on the host (kept as the measured CPU baseline and as
//                  an independent implementation the device build is tested against)
//   bvh_build.hip  the device build: wide binning passes for nodes of more than
//                  kBvhSmallPrims primitives, one workgroup per smaller subtree, then the
//                  collapse to 4-wide nodes and the 8-bit quantisation on the GPU
//
// The binary tree: binned SAH over centroids with kBvhBins bins per axis, the cost
// area(L) * |L| + area(R) * |R| over the 3 * (kBvhBins - 1) bin boundaries, first minimum in
// the order (axis 0..2, boundary from the highest down), a stable partition.  Nodes with at
// most maxLeaf primitives are leaves.  Where no boundary separates the centroids the node is
// split at the middle of its range; beyond depth kBvhSahMaxDepth a node of at most
// kBvhSmallPrims primitives is split at the median of a stable sort of its centroids along the
// widest centroid axis, and a larger one at the middle of its range (guards that bound the
// depth on degenerate input).  Primitive boxes are taken with -0 replaced by +0 so that the
// minima and maxima do not depend on the order they are taken in.
//
// The 4-wide tree: from a binary node's two children the inner child with the largest box area
// (the first of equal ones) is replaced by its own two children until four slots are filled or
// only leaves remain; slot order carries over.  Node 0 is the root, a child code > 0 is a node,
// < 0 a leaf ~((first << kLeafShift) | (count - 1)) over the primitive order, 0 an empty slot.
// Node numbering differs between the two builders (depth-first on the host, breadth-first on
// the device); the trees are equal up to that renumbering.
#pragma once

#include <hip/hip_runtime.h>

#include "dev_types.h"

#include <cmath>
#include <cstdint>
#include <vector>

namespace vlra {

constexpr int kBvhBins = 16;
constexpr uint32_t kBvhSahMaxDepth = 40;     // root split is depth 1
constexpr uint32_t kBvhSmallPrims = 1024;    // subtrees of at most this many primitives: one workgroup
constexpr uint32_t kBvhMaxPrims = 1u << 28;  // leaf code limit (first << kLeafShift)

// Traversal node stack per ray: kLdsStack entries in LDS, kScratchStack more in scratch.  A
// visit of an inner 4-wide node pushes at most 3 entries and descends one level; entering an
// instance pushes one marker.  A two-level walk therefore never holds more than
// travStackBound(TLAS depth, deepest BLAS depth) entries, and scene set-up refuses trees whose
// bound exceeds the capacity, so no entry is ever lost.
constexpr int kLdsStack = 16;
constexpr int kScratchStack = 64;
constexpr uint32_t kTravStackCapacity = kLdsStack + kScratchStack;
__host__ __device__ inline uint32_t travStackBound(uint32_t tlasDepth, uint32_t blasDepth) {
    return 3 * (tlasDepth + blasDepth) + 1;
}

__host__ __device__ inline float bvhNoNegZero(float v) { return v + 0.0f; }

// Box area as the SAH uses it (0 for an empty box).
__host__ __device__ inline float bvhArea(const float lo[3], const float hi[3]) {
    if (lo[0] > hi[0]) return 0.0f;
    const float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
    return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
}

// Bin of a centroid coordinate c in [clo, clo + ext], ext > 0.
__host__ __device__ inline int bvhBin(float c, float clo, float ext) {
    int k = static_cast<int>(static_cast<float>(kBvhBins) * (c - clo) / ext);
    return k < 0 ? 0 : (k > kBvhBins - 1 ? kBvhBins - 1 : k);
}

// One quantised node: per axis the grid step is the power of two 2^k with 255 * 2^k >= the
// node extent; each child plane is rounded outwards and checked with the decoding arithmetic
// the traversal uses (origin + float(q) * step: an exact product and one rounded sum).
__host__ __device__ inline BVH4QNode quantizeBvh4Node(const BVH4Node &f) {
    BVH4QNode q;
    for (int a = 0; a < 3; ++a) {
        q.origin[a] = 0.0f;
        q.qlo[a] = 0;
        q.qhi[a] = 0;
    }
    q.scaleExp = 0;
    q.pad[0] = q.pad[1] = 0;
    const float *lo[3] = {f.loX, f.loY, f.loZ}, *hi[3] = {f.hiX, f.hiY, f.hiZ};
    for (int k = 0; k < 4; ++k) q.child[k] = f.child[k];
    for (int a = 0; a < 3; ++a) {
        float blo = INFINITY, bhi = -INFINITY;
        for (int k = 0; k < 4; ++k)
            if (f.child[k] != 0) {
                blo = fminf(blo, lo[a][k]);
                bhi = fmaxf(bhi, hi[a][k]);
            }
        if (!(blo <= bhi)) blo = bhi = 0.0f;   // no child
        q.origin[a] = blo;
        int e;
        frexpf(fmaxf((bhi - blo) / 255.0f, 1e-30f), &e);   // (ext / 255) < 2^e
        e = e < -126 ? -126 : e;
        for (;; ++e) {
            const float step = ldexpf(1.0f, e);
            uint32_t wl = 0, wh = 0;
            bool ok = true;
            for (int k = 0; k < 4; ++k) {
                if (f.child[k] == 0) continue;
                const float ql = floorf((lo[a][k] - blo) / step), qh = ceilf((hi[a][k] - blo) / step);
                int il = static_cast<int>(fminf(fmaxf(ql, 0.0f), 255.0f));
                int ih = static_cast<int>(fminf(fmaxf(qh, 0.0f), 255.0f));
                while (il > 0 && blo + static_cast<float>(il) * step > lo[a][k]) --il;
                while (ih < 255 && blo + static_cast<float>(ih) * step < hi[a][k]) ++ih;
                if (blo + static_cast<float>(il) * step > lo[a][k] || blo + static_cast<float>(ih) * step < hi[a][k]) {
                    ok = false;
                    break;
                }
                wl |= static_cast<uint32_t>(il) << (8 * k);
                wh |= static_cast<uint32_t>(ih) << (8 * k);
            }
            if (!ok) continue;
            q.qlo[a] = wl;
            q.qhi[a] = wh;
            q.scaleExp |= static_cast<uint32_t>(e + 127) << (8 * a);
            break;
        }
    }
    return q;
}

// ---- host builder (bvh_host.cpp) ----
// boxes: 6 floats per primitive (lo xyz, hi xyz).  leafRemap (optional, maxLeaf 1): leaf codes
// carry leafRemap[primitive] in place of the position in the order.
struct HostBvh4 {
    std::vector<BVH4QNode> nodes;
    std::vector<uint32_t> order;   // position -> primitive
    uint32_t depth = 0;            // levels of 4-wide nodes (the root alone: 1)
};
void hostBuildBvh4(const float *boxes, uint32_t numPrims, uint32_t maxLeaf, const uint32_t *leafRemap,
                   HostBvh4 *out);

// ---- device builder (bvh_build.hip) ----
struct DeviceBuffer;
struct BvhBuildTimes {
    uint32_t largeLevels = 0, smallTasks = 0, collapseLevels = 0;
};
class GpuBvhBuilder {
public:
    GpuBvhBuilder();
    ~GpuBvhBuilder();
    GpuBvhBuilder(const GpuBvhBuilder &) = delete;
    GpuBvhBuilder &operator=(const GpuBvhBuilder &) = delete;

    // Upper bound of the 4-wide nodes a build over n primitives writes.
    static uint32_t nodeCapacity(uint32_t numPrims) { return numPrims > 1 ? numPrims - 1 : 1; }

    // boxes: device, 6 floats per primitive.  Writes the nodes (device, nodeCapacity entries)
    // and returns their count; the primitive order stays in order() until the next build.
    // leafRemap: device or null.  Runs on `stream` and returns after it has finished.
    uint32_t build(const float *boxes, uint32_t numPrims, uint32_t maxLeaf, const uint32_t *leafRemap,
                   BVH4QNode *nodes, hipStream_t stream, BvhBuildTimes *info = nullptr);
    // One tree per segment of triangle records (device): segment k covers records
    // [segFirst[k], segFirst[k] + segCount[k]) (consecutive from 0, none empty).  Segment k's
    // root is node k, its inner child codes count from node k and its leaf codes from
    // segFirst[k], so a segment is traversed with base node k and base record segFirst[k].
    // The nodes array holds sum(nodeCapacity(segCount[k])) entries; order() maps positions to
    // records over all segments.
    uint32_t buildTriangles(const TriRecord *tris, const std::vector<uint32_t> &segFirst,
                            const std::vector<uint32_t> &segCount, uint32_t maxLeaf, BVH4QNode *nodes,
                            hipStream_t stream, BvhBuildTimes *info = nullptr);
    const uint32_t *order() const;

    struct Impl;

private:
    Impl *impl_;
};

// Largest build of buildBvh4NoAlloc.
constexpr uint32_t kNoAllocMaxPrims = 1u << 17;
// A single-tree build over n <= kNoAllocMaxPrims device boxes (6 floats each) that allocates no
// memory: its scratch is one static region per GPU, shared by every call on that GPU (calls that
// overlap on one GPU race), and the primitive order goes straight to `order` (device, n
// entries).  Returns the node count after the build has finished on `stream`.
uint32_t buildBvh4NoAlloc(const float *boxes, uint32_t numPrims, uint32_t maxLeaf, const uint32_t *leafRemap,
                          BVH4QNode *nodes, uint32_t *order, hipStream_t stream);

// dst[i] = src[order[i]] on the device.
void gatherTriangles(const TriRecord *src, const uint32_t *order, uint32_t n, TriRecord *dst, hipStream_t stream);

} // namespace vlra
