// Wavefront path tracer for MI355X (gfx950).
//
// One vlrContextRender call = one sample per pixel, run as a sequence of kernels joined by
// device queues (the reference runs the same integrator as one OptiX megakernel,
// libVLR/GPU_kernels/path_tracing.cu:30-337):
//   k_frameBegin  clear the per-depth queue counters
//   k_generate    camera ray per pixel (path_tracing.cu:30-70, cameras.cu:43-76, 145-157), in
//                 8 x 8 pixel tiles per wave
//   per depth d:  k_intersect  closest hit over the two-level BVH
//                 k_shade      hit/miss programs (path_tracing.cu:133-337): implicit emission
//                              with MIS, Russian roulette, next-event estimation (builds a
//                              shadow ray, light_transport_common.h:147-192), BSDF sampling;
//                              appends the next ray and the shadow ray with wave-aggregated
//                              atomics
//                 k_shadow     any-hit traversal, adds the NEE contribution if unoccluded
//   k_accumulate  Kahan accumulation (path_tracing.cu:119-127)
//   k_convert     XYZ -> Rec.709 output (post_process.cu:93-163)
// Per-path random numbers are drawn in the reference's order, so a path consumes the same
// PCG32 stream as the reference's megakernel.
//
// Traversal: 4-wide nodes with 8-bit quantised child boxes (64 B), inner nodes and leaves in
// separate loops, per-thread short stack in LDS with an overflow area in scratch, watertight
// ray/triangle test (Woop, Benthin and Wald, JCGT 2(1) 2013), compiled without floating-point
// contraction so that closest hits reproduce the oracle's brute force bit for bit.
#include "bvh.h"
#include "dev_shading.h"
#include "shader_nodes.h"
#include "launch.h"
#include "radiance_cache.h"

#pragma clang fp contract(off)

// The build compiles this file once per unit (VLRA_UNIT): 1 RGB path tracing, 2 the same
// spectral, 3 RGB LVC-BPT and light tracing, 4 the same spectral, 5 dispatchers, conversion
// and the direct query kernels, 6 RGB render batches, 7 the same spectral, 8 RGB path
// tracing of scenes with node programs, 9 the same spectral; 0 (or no definition) emits
// everything.  The units
// compile in parallel.  Each unit's kernels live in a namespace of their own, so no kernel is
// defined by two units.
#ifndef VLRA_UNIT
#define VLRA_UNIT 0
#endif
#if VLRA_UNIT == 1
#define VLRA_UNIT_NS unit_pt_rgb
#elif VLRA_UNIT == 2
#define VLRA_UNIT_NS unit_pt_spectral
#elif VLRA_UNIT == 3
#define VLRA_UNIT_NS unit_bpt_rgb
#elif VLRA_UNIT == 4
#define VLRA_UNIT_NS unit_bpt_spectral
#elif VLRA_UNIT == 5
#define VLRA_UNIT_NS unit_misc
#elif VLRA_UNIT == 6
#define VLRA_UNIT_NS unit_batch_rgb
#elif VLRA_UNIT == 7
#define VLRA_UNIT_NS unit_batch_spectral
#elif VLRA_UNIT == 8
#define VLRA_UNIT_NS unit_pt_nodes_rgb
#elif VLRA_UNIT == 9
#define VLRA_UNIT_NS unit_pt_nodes_spectral
#else
#define VLRA_UNIT_NS unit_all
#endif

namespace vlra {
namespace VLRA_UNIT_NS {

constexpr int kTraceBlock = 128;
// per block: kLdsStack entries per lane of node codes, then as many entry distances
constexpr int kTraceLds = 2 * kLdsStack * kTraceBlock;
// A CU's 160 KB of LDS holds 9 blocks of kTraceLds (4.5 waves per SIMD).  The per-depth
// closest-hit kernel keeps 12 entries per lane (12 KB per block), so that its registers (5 waves
// per SIMD) set the limit; the per-depth any-hit kernel needs no entry distances (8 KB per block
// for 16 entries; 7 waves per SIMD).  The tail kernel, limited to 4 waves per SIMD by its
// registers, keeps the full kTraceLds.
constexpr int kIntersectStack = 12;
constexpr int kIntersectLds = 2 * kIntersectStack * kTraceBlock;
constexpr int kShadowLds = kLdsStack * kTraceBlock;
#ifndef VLRA_TAIL_WAVES_PER_EU
#define VLRA_TAIL_WAVES_PER_EU 4
#endif
constexpr int kTailWavesPerEU = VLRA_TAIL_WAVES_PER_EU;   // occupancy asked of the tail kernel
// Occupancy asked of the per-path LVC-BPT and light-tracing kernels.  Left to themselves they
// take 130-230 VGPRs and run 2 waves per SIMD; with some registers spilled, more waves hide
// more of their traversals' fetch latency (C3 box at 1280x720: light paths 1.55 -> 1.39 ms,
// eye paths 2.40 -> 1.89 ms, light tracing 2.60 -> 1.79 ms per frame).
constexpr int kBptLightWavesPerEU = 4;
constexpr int kBptEyeWavesPerEU = 3;
constexpr int kLightTraceWavesPerEU = 4;
constexpr int32_t kSentinel = INT32_MIN;

// ---------------------------------------------------------------------------------------------
// ray set-up and primitive tests
struct RayPre {
    V3 org, dir, inv;
    int kx, ky, kz;
    float Sx, Sy, Sz;
    __device__ __forceinline__ void init(const V3 &o, const V3 &d) {
        org = o;
        dir = d;
        kz = 0;
        float adx = fabsf(d.x), ady = fabsf(d.y), adz = fabsf(d.z);
        float best = adx;
        if (ady > best) { kz = 1; best = ady; }
        if (adz > best) { kz = 2; }
        kx = kz == 2 ? 0 : kz + 1;
        ky = kx == 2 ? 0 : kx + 1;
        float dkz = d[kz];
        if (dkz < 0.0f) { int t = kx; kx = ky; ky = t; }
        Sx = d[kx] / dkz;
        Sy = d[ky] / dkz;
        Sz = 1.0f / dkz;
        inv = V3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    }
};

__device__ __forceinline__ bool intersectTri(const RayPre &r, const V3 &p0, const V3 &p1, const V3 &p2, float tmin,
                                             float tmax, float *tOut, float *b1Out, float *b2Out) {
    const V3 A = p0 - r.org, B = p1 - r.org, C = p2 - r.org;
    const float Akz = A[r.kz], Bkz = B[r.kz], Ckz = C[r.kz];
    const float Ax = A[r.kx] - r.Sx * Akz;
    const float Ay = A[r.ky] - r.Sy * Akz;
    const float Bx = B[r.kx] - r.Sx * Bkz;
    const float By = B[r.ky] - r.Sy * Bkz;
    const float Cx = C[r.kx] - r.Sx * Ckz;
    const float Cy = C[r.ky] - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = static_cast<float>(static_cast<double>(Cx) * static_cast<double>(By) - static_cast<double>(Cy) * static_cast<double>(Bx));
        V = static_cast<float>(static_cast<double>(Ax) * static_cast<double>(Cy) - static_cast<double>(Ay) * static_cast<double>(Cx));
        W = static_cast<float>(static_cast<double>(Bx) * static_cast<double>(Ay) - static_cast<double>(By) * static_cast<double>(Ax));
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T = U * (r.Sz * Akz) + V * (r.Sz * Bkz) + W * (r.Sz * Ckz);
    const float rcpDet = 1.0f / det;
    const float t = T * rcpDet;
    if (!(t >= tmin && t < tmax)) return false;
    *tOut = t;
    *b1Out = V * rcpDet;
    *b2Out = W * rcpDet;
    return true;
}

// Conservative slab test: the exit distance is widened by a few ulps so rounding never culls
// a box that holds a hit the triangle test would report.
__device__ __forceinline__ bool slab(const V3 &lo, const V3 &hi, const RayPre &r, float tmin, float tmax,
                                     float *tEnter) {
    float tx0 = (lo.x - r.org.x) * r.inv.x, tx1 = (hi.x - r.org.x) * r.inv.x;
    float ty0 = (lo.y - r.org.y) * r.inv.y, ty1 = (hi.y - r.org.y) * r.inv.y;
    float tz0 = (lo.z - r.org.z) * r.inv.z, tz1 = (hi.z - r.org.z) * r.inv.z;
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    tf = fminf(tf * (1.0f + 8.0f * __FLT_EPSILON__), tmax * (1.0f + 8.0f * __FLT_EPSILON__));
    *tEnter = tn;
    return tn <= tf;
}

struct TraceHit {
    int32_t inst, geomInst, prim;
    float t, b1, b2;
};

// Work counters of one traversal (inner nodes fetched, triangles tested, instances entered),
// kept only by the statistics instantiation.
struct TravCount {
    uint32_t nodes = 0, tris = 0, insts = 0;
};

// Traversal state: one ray's walk through the two-level BVH.  The node stack
// lives in LDS (kLdsStack entries per lane) with an overflow area the caller keeps in scratch.
constexpr int32_t kTravDone = INT32_MIN + 1;
struct TravState {
    RayPre world, ray;
    float tmin, tmax;
    int32_t cur, curInst;
    int sp;
    uint32_t blasNode, blasTri;   // bases of the instance being traversed
    bool inBlas, found, anyHit;
    TraceHit hit;
    Pcg32 *rng;                   // the path's generator for the stochastic alpha test
    float vis;                    // fractional visibility of an any-hit ray through alpha surfaces
};

// alpha of a triangle group's alpha node at a candidate hit (renderer_common.h:532-548)
__device__ __forceinline__ float alphaAt(const DevScene &sc, int32_t inst, uint32_t geomInst, uint32_t prim, float b1,
                                         float b2);

__device__ __forceinline__ void travBegin(TravState &st, const DevScene &sc, const V3 &org, const V3 &dir, float tmin,
                                          float tmax, bool anyHit) {
    st.world.init(org, dir);
    st.ray = st.world;
    st.tmin = tmin;
    st.tmax = tmax;
    st.cur = sc.hasTlas ? 0 : kTravDone;
    st.curInst = -1;
    st.sp = 0;
    st.blasNode = st.blasTri = 0;
    st.inBlas = false;
    st.found = false;
    st.anyHit = anyHit;
    st.hit.inst = -1;
    st.hit.geomInst = st.hit.prim = -1;
    st.hit.t = st.hit.b1 = st.hit.b2 = 0.0f;
    st.rng = nullptr;
    st.vis = 1.0f;
}

// Runs the traversal to its end (then st.found and st.hit hold the closest hit, or any hit for
// an any-hit ray).  The loop is a while-while loop: a lane walks inner nodes until it reaches a
// leaf, an instance's end marker or the end of the walk, and only then handles that one item
// (enter the instance, test the leaf's triangles, or leave the instance) before walking inner
// nodes again.  Each lane visits the same nodes in the same order as a one-item-per-iteration
// loop; the wave just runs its lanes' inner-node visits together and their leaves together.
// With ALPHA, triangles of groups with an alpha node pass the stochastic alpha test
// (path_tracing.cu:14-25) for closest-hit rays and scale the fractional visibility
// (light_transport_common.h:205-218) for any-hit rays.
// ANY: an any-hit ray, which needs no visiting order: its children are stacked as they come,
// without entry distances.
// SD: stack entries per lane in LDS (lds holds SD * kTraceBlock node codes, then as many entry
// distances unless ANY); the other kTravStackCapacity - SD entries are in ovf / ovfT.
template <bool ANY, bool STATS, bool ALPHA = false, int SD = kLdsStack>
__device__ __forceinline__ void traverse(TravState &st, const DevScene &sc, int32_t *lds, int lane, int32_t *ovf,
                                         float *ovfT, TravCount *tc) {
    constexpr int kOvf = static_cast<int>(kTravStackCapacity) - SD;
    float *ldsT = ANY ? nullptr : reinterpret_cast<float *>(lds + SD * kTraceBlock);
    // each stack entry keeps the distance at which the ray enters its box; once a closer hit
    // has shrunk tmax past it (with the slab test's widening), the entry is dropped unvisited.
    // Scene set-up bounds the depth by kTravStackCapacity (travStackBound), so the index never
    // reaches past the scratch area; the clamp only keeps the access inside the array.
    auto push = [&](int32_t v, float tEnter) {
        if (st.sp < SD) {
            lds[st.sp * kTraceBlock + lane] = v;
            if constexpr (!ANY) ldsT[st.sp * kTraceBlock + lane] = tEnter;
        } else {
            ovf[min(st.sp - SD, kOvf - 1)] = v;
            if constexpr (!ANY) ovfT[min(st.sp - SD, kOvf - 1)] = tEnter;
        }
        ++st.sp;
    };
    // pop, dropping entries that lie beyond the closest hit so far (an any-hit ray stops at its
    // first hit, so every entry it pops is still open)
    auto pop = [&]() -> int32_t {
        if constexpr (ANY) {
            if (st.sp == 0) return kTravDone;
            --st.sp;
            return st.sp < SD ? lds[st.sp * kTraceBlock + lane] : ovf[min(st.sp - SD, kOvf - 1)];
        } else {
            const float tCull = st.tmax * (1.0f + 8.0f * __FLT_EPSILON__);
            while (st.sp > 0) {
                --st.sp;
                float te;
                int32_t v;
                if (st.sp < SD) {
                    v = lds[st.sp * kTraceBlock + lane];
                    te = ldsT[st.sp * kTraceBlock + lane];
                } else {
                    v = ovf[min(st.sp - SD, kOvf - 1)];
                    te = ovfT[min(st.sp - SD, kOvf - 1)];
                }
                if (te <= tCull) return v;
            }
            return kTravDone;
        }
    };
    while (st.cur != kTravDone) {
        // inner nodes
        while (st.cur >= 0) {
            const int32_t cur = st.cur;
            if constexpr (STATS) ++tc->nodes;
            const BVH4QNode *nodes = st.inBlas ? sc.blasNodes + st.blasNode : sc.tlasNodes;
            const uint4 *np = reinterpret_cast<const uint4 *>(nodes + cur);
            const uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3];
            // decode: origin + q * step per axis (step = 2^(e - 127), exact product)
            const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
            const float sx = __uint_as_float((w0.w & 0xFFu) << 23), sy = __uint_as_float(((w0.w >> 8) & 0xFFu) << 23),
                        sz = __uint_as_float(((w0.w >> 16) & 0xFFu) << 23);
            const uint32_t qlx = w1.x, qly = w1.y, qlz = w1.z, qhx = w1.w, qhy = w2.x, qhz = w2.y;
            auto lo = [&](int k) {
                return V3(ox + static_cast<float>((qlx >> (8 * k)) & 0xFFu) * sx,
                          oy + static_cast<float>((qly >> (8 * k)) & 0xFFu) * sy,
                          oz + static_cast<float>((qlz >> (8 * k)) & 0xFFu) * sz);
            };
            auto hi = [&](int k) {
                return V3(ox + static_cast<float>((qhx >> (8 * k)) & 0xFFu) * sx,
                          oy + static_cast<float>((qhy >> (8 * k)) & 0xFFu) * sy,
                          oz + static_cast<float>((qhz >> (8 * k)) & 0xFFu) * sz);
            };
            // child code 0 (the root's index) marks an empty slot
            constexpr float kMiss = __builtin_huge_valf();
            float t0 = kMiss, t1 = kMiss, t2 = kMiss, t3 = kMiss;
            int32_t c0 = static_cast<int32_t>(w2.z), c1 = static_cast<int32_t>(w2.w), c2 = static_cast<int32_t>(w3.x),
                    c3 = static_cast<int32_t>(w3.y);
            float tt;
            if (c0 != 0 && slab(lo(0), hi(0), st.ray, st.tmin, st.tmax, &tt)) t0 = tt;
            if (c1 != 0 && slab(lo(1), hi(1), st.ray, st.tmin, st.tmax, &tt)) t1 = tt;
            if (c2 != 0 && slab(lo(2), hi(2), st.ray, st.tmin, st.tmax, &tt)) t2 = tt;
            if (c3 != 0 && slab(lo(3), hi(3), st.ray, st.tmin, st.tmax, &tt)) t3 = tt;
            const int hits = (t0 != kMiss) + (t1 != kMiss) + (t2 != kMiss) + (t3 != kMiss);
            if (ANY && hits > 0) {
                // any order will do: stack the hit children but the first, visit that one next
                int32_t next = 0;   // 0 is never a child code
                auto take = [&](int32_t c, float t) {
                    if (t == kMiss) return;
                    if (next == 0) next = c;
                    else push(c, t);
                };
                take(c0, t0);
                take(c1, t1);
                take(c2, t2);
                take(c3, t3);
                st.cur = next;
            } else if (hits > 0) {
                // sort the four (distance, child) pairs, nearest first; misses sort last
                auto cswap = [](float &ta, int32_t &ca, float &tb, int32_t &cb) {
                    if (tb < ta) {
                        float t = ta; ta = tb; tb = t;
                        int32_t c = ca; ca = cb; cb = c;
                    }
                };
                cswap(t0, c0, t1, c1);
                cswap(t2, c2, t3, c3);
                cswap(t0, c0, t2, c2);
                cswap(t1, c1, t3, c3);
                cswap(t1, c1, t2, c2);
                if (hits > 3) push(c3, t3);
                if (hits > 2) push(c2, t2);
                if (hits > 1) push(c1, t1);
                st.cur = c0;
            } else {
                st.cur = pop();
            }
        }
        if (st.cur == kTravDone) break;
        const int32_t cur = st.cur;
        if (cur != kSentinel) {
            const uint32_t code = ~static_cast<uint32_t>(cur);
            const uint32_t first = code >> kLeafShift;
            const uint32_t count = (code & (kMaxLeafPrims - 1)) + 1;
            if (!st.inBlas) {
                // enter the instance: object-space ray, BLAS root, return marker on the stack
                if constexpr (STATS) ++tc->insts;
                const DevInstance &I = sc.instances[first];
                V3 o = xfmPoint(I.w2o, st.world.org);
                V3 d = xfmVector(I.w2o, st.world.dir);
                st.ray.init(o, d);
                st.curInst = static_cast<int32_t>(first);
                st.blasNode = I.blasNodeBase;
                st.blasTri = I.blasTriBase;
                st.inBlas = true;
                push(kSentinel, -__builtin_huge_valf());
                st.cur = 0;
                continue;
            }
            if constexpr (STATS) tc->tris += count;
            const TriRecord *tris = sc.blasTris + st.blasTri;
            for (uint32_t k = 0; k < count; ++k) {
                const float4 *tp = reinterpret_cast<const float4 *>(tris + first + k);
                const float4 q0 = tp[0], q1 = tp[1], q2 = tp[2];
                float t, b1, b2;
                if (intersectTri(st.ray, xyz(q0), xyz(q1), xyz(q2), st.tmin, st.tmax, &t, &b1, &b2)) {
                    if constexpr (ALPHA) {
                        const uint32_t gi = __float_as_uint(q0.w);
                        if (sc.hasAlpha && sc.geomInsts[gi].alphaCount) {
                            const float alpha = alphaAt(sc, st.curInst, gi, __float_as_uint(q1.w), b1, b2);
                            if (ANY) {
                                st.vis *= 1 - alpha;
                                if (st.vis != 0.0f) continue;
                            } else if (st.rng && st.rng->uniform() >= alpha) {
                                continue;
                            }
                        }
                    }
                    st.tmax = t;
                    st.found = true;
                    st.hit.inst = st.curInst;
                    st.hit.geomInst = static_cast<int32_t>(__float_as_uint(q0.w));
                    st.hit.prim = static_cast<int32_t>(__float_as_uint(q1.w));
                    st.hit.t = t;
                    st.hit.b1 = b1;
                    st.hit.b2 = b2;
                    if (ANY) break;
                }
            }
            if (ANY && st.found) {
                st.cur = kTravDone;
                break;
            }
        } else {
            // leave the instance
            st.inBlas = false;
            st.ray = st.world;
        }
        st.cur = pop();
    }
}

// rng: the path's generator (closest-hit rays through alpha surfaces); vis: the fractional
// visibility of an any-hit ray (0 when occluded)
template <bool ANY, bool STATS = false, bool ALPHA = false, int SD = kLdsStack>
__device__ __forceinline__ bool traceRay(const DevScene &sc, const V3 &org, const V3 &dir, float tmin, float tmax,
                                         TraceHit *hit, int32_t *lds, int lane, TravCount *tc = nullptr,
                                         Pcg32 *rng = nullptr, float *vis = nullptr) {
    TravState st;
    travBegin(st, sc, org, dir, tmin, tmax, ANY);
    st.rng = rng;
    int32_t ovf[kTravStackCapacity - SD];
    float ovfT[ANY ? 1 : kTravStackCapacity - SD];
    traverse<ANY, STATS, ALPHA, SD>(st, sc, lds, lane, ovf, ovfT, tc);
    if (st.found) *hit = st.hit;
    if (vis) *vis = st.found ? 0.0f : st.vis;
    return st.found;
}

// wave-aggregated queue append: one atomic per wave, lanes get consecutive slots
__device__ __forceinline__ uint32_t appendSlot(uint32_t *counter, bool active) {
    uint64_t mask = __ballot(active);
    if (!active) return 0;
    uint32_t lanes = __popcll(mask);
    uint32_t lane = __lane_id();
    uint32_t rank = __popcll(mask & ((1ull << lane) - 1ull));
    uint32_t leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, lanes);
    base = __shfl(base, static_cast<int>(leader));
    return base + rank;
}

// ---------------------------------------------------------------------------------------------
// distributions (shared.h:182-298)
__device__ __forceinline__ uint32_t nextPow2(uint32_t x) { return x <= 1 ? 1 : 1u << (32 - __clz(x - 1)); }

__device__ __forceinline__ uint32_t discreteSearch(const float *cdf, uint32_t n, float u) {
    int idx = 0;
    for (int d = static_cast<int>(nextPow2(n) >> 1); d >= 1; d >>= 1) {
        if (idx + d >= static_cast<int>(n)) continue;
        if (cdf[idx + d] <= u) idx += d;
    }
    return static_cast<uint32_t>(idx);
}
// Discrete1D: PMF at base, CDF (n + 1 entries) after it
__device__ __forceinline__ uint32_t sampleDiscrete(const float *pool, uint32_t base, uint32_t n, float u, float *prob,
                                                   float *remapped) {
    const float *pmf = pool + base, *cdf = pool + base + n;
    uint32_t idx = discreteSearch(cdf, n, u);
    *prob = pmf[idx];
    if (remapped) *remapped = (u - cdf[idx]) / (cdf[idx + 1] - cdf[idx]);
    return idx;
}
// RegularConstantContinuous1D: PDF at base, CDF after it
__device__ __forceinline__ float sampleContinuous(const float *pool, uint32_t base, uint32_t n, float u, float *pd) {
    const float *pdf = pool + base, *cdf = pool + base + n;
    uint32_t idx = discreteSearch(cdf, n, u);
    *pd = pdf[idx];
    float t = (u - cdf[idx]) / (cdf[idx + 1] - cdf[idx]);
    return (idx + t) / n;
}
__device__ __forceinline__ float continuousPDF(const float *pool, uint32_t base, uint32_t n, float smp) {
    int32_t idx = min(static_cast<int32_t>(n) - 1, static_cast<int32_t>(smp * n));
    return pool[base + idx];
}
__device__ __forceinline__ void envMapSample(const DevScene &sc, float u0, float u1, float *d0, float *d1, float *pd) {
    float topPDF;
    *d1 = sampleContinuous(sc.distPool, sc.envTopBase, sc.envMapH, u1, &topPDF);
    uint32_t row = min(static_cast<uint32_t>(sc.envMapH * *d1), sc.envMapH - 1);
    *d0 = sampleContinuous(sc.distPool, sc.envRowBase + row * (2 * sc.envMapW + 1), sc.envMapW, u0, pd);
    *pd *= topPDF;
}
__device__ __forceinline__ float envMapPDF(const DevScene &sc, float d0, float d1) {
    uint32_t row = min(static_cast<uint32_t>(sc.envMapH * d1), sc.envMapH - 1);
    return continuousPDF(sc.distPool, sc.envTopBase, sc.envMapH, d1) *
           continuousPDF(sc.distPool, sc.envRowBase + row * (2 * sc.envMapW + 1), sc.envMapW, d0);
}

// bilinear environment fetch, repeat wrapping in both axes
__device__ __forceinline__ float4 envFetch(const DevScene &sc, float u, float v) {
    float x = u * sc.envW - 0.5f, y = v * sc.envH - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y);
    float wx = x - fx0, wy = y - fy0;
    int32_t x0 = static_cast<int32_t>(fx0), y0 = static_cast<int32_t>(fy0);
    int32_t W = static_cast<int32_t>(sc.envW), H = static_cast<int32_t>(sc.envH);
    int32_t xa = ((x0 % W) + W) % W, xb = (((x0 + 1) % W) + W) % W;
    int32_t ya = ((y0 % H) + H) % H, yb = (((y0 + 1) % H) + H) % H;
    float4 t00 = sc.envTex[ya * W + xa], t10 = sc.envTex[ya * W + xb];
    float4 t01 = sc.envTex[yb * W + xa], t11 = sc.envTex[yb * W + xb];
    float4 r;
    r.x = (1 - wy) * ((1 - wx) * t00.x + wx * t10.x) + wy * ((1 - wx) * t01.x + wx * t11.x);
    r.y = (1 - wy) * ((1 - wx) * t00.y + wx * t10.y) + wy * ((1 - wx) * t01.y + wx * t11.y);
    r.z = (1 - wy) * ((1 - wx) * t00.z + wx * t10.z) + wy * ((1 - wx) * t01.z + wx * t11.z);
    r.w = 1.0f;
    return r;
}

template <int N>
__device__ __forceinline__ Spec<N> envEmittance(const DevScene &sc, const DevMaterial &m, float texU, float texV,
                                                const Wls<N> &wls) {
    Spec<N> e;
    if (!m.useEnvTexture || sc.envTex == nullptr) {
        e = evalSpectrum<N>(sc, m.spec[0], wls);
    } else {
        float4 t = envFetch(sc, texU, texV);
        if constexpr (N == 3) {
            e[0] = t.x; e[1] = t.y; e[2] = t.z;
        } else {
            uint32_t adj, st;
            mengAdjacents(sc, t.x, t.y, &adj, &st);
            e = mengEvaluate(sc, adj, st, t.z / 0.009355121400914532f, wls);
        }
    }
    return e * m.scalars[0];
}

// triangle.cu:6-53 + renderer_common.h:609-626: decode in object space, then to world
template <bool NODES = true>
__device__ __forceinline__ void decodeHit(const DevScene &sc, int32_t instIdx, int32_t giIdx, int32_t prim, float b1,
                                          float b2, SurfPt *sp, float *hypAreaPDF) {
    const DevInstance &I = sc.instances[instIdx];
    const DevGeomInst &G = sc.geomInsts[giIdx];
    const uint32_t *ix = sc.indices + G.indexBase + 3 * prim;
    const DevVertex &v0 = sc.vertices[G.vertexBase + ix[0]];
    const DevVertex &v1 = sc.vertices[G.vertexBase + ix[1]];
    const DevVertex &v2 = sc.vertices[G.vertexBase + ix[2]];
    V3 p0 = ld3(v0.p), p1 = ld3(v1.p), p2 = ld3(v2.p);
    V3 gn = cross(p1 - p0, p2 - p0);
    float area = gn.length() / 2;
    gn = gn / (2 * area);
    float probLightPrim = G.distIntegral > 0.0f ? area / G.distIntegral : 0.0f;
    *hypAreaPDF = probLightPrim / area;
    float b0 = 1 - b1 - b2;
    V3 position = b0 * p0 + b1 * p1 + b2 * p2;
    V3 sn = normalize(b0 * ld3(v0.n) + b1 * ld3(v1.n) + b2 * ld3(v2.n));
    V3 tc0 = b0 * ld3(v0.t) + b1 * ld3(v1.t) + b2 * ld3(v2.t);
    if (!sn.allFinite() || !tc0.allFinite()) {
        V3 bt;
        sn = gn;
        makeCoordinateSystem(sn, &tc0, &bt);
    }
    sp->texU = b0 * v0.u + b1 * v1.u + b2 * v2.u;
    sp->texV = b0 * v0.v + b1 * v1.v + b2 * v2.v;
    float dotNT = dot(sn, tc0);
    tc0 = tc0 - dotNT * sn;
    sp->position = xfmPoint(I.o2w, position);
    sp->frame = Frame(normalize(xfmVector(I.o2w, tc0)), normalize(xfmNormal(I.w2o, sn)));
    sp->gn = normalize(xfmNormal(I.w2o, gn));
    sp->isPoint = false;
    sp->atInfinity = false;
    sp->instIndex = static_cast<uint32_t>(instIdx);
    sp->geomInstIndex = static_cast<uint32_t>(giIdx);
    sp->primIndex = static_cast<uint32_t>(prim);
    sp->u = b0;
    sp->v = b1;
    if constexpr (NODES) applyGroupNodes(sc, G, sp);   // normal and tangent nodes (renderer_common.h:620-625)
}

__device__ __forceinline__ float alphaAt(const DevScene &sc, int32_t inst, uint32_t geomInst, uint32_t prim, float b1,
                                         float b2) {
    SurfPt sp;
    float hypAreaPDF;
    decodeHit<false>(sc, inst, static_cast<int32_t>(geomInst), static_cast<int32_t>(prim), b1, b2, &sp, &hypAreaPDF);
    const DevGeomInst &G = sc.geomInsts[geomInst];
    NodeOutputs<3> o;
    Wls<3> none;
    runNodeProgram<3, false>(sc, G.alphaBase, G.alphaCount, sp, none, VC_F1, VC_F1, &o);
    return (o.has & 8u) ? o.vec[0].x : 1.0f;   // calcNode's default 1
}

// environment surface point from a direction (path_tracing.cu:282-310)
__device__ __forceinline__ void envPointFromDir(const DevScene &sc, const V3 &direction, SurfPt *sp, float *hypAreaPDF) {
    float theta = acosf(fminf(fmaxf(direction.y, -1.0f), 1.0f));
    float phi = fmodf(atan2f(-direction.x, direction.z) + 2 * kPiF, 2 * kPiF);
    float sinPhi, cosPhi;
    sincosf(phi, &sinPhi, &cosPhi);
    V3 tc0 = normalize(V3(-cosPhi, 0.0f, -sinPhi));
    sp->frame.x = tc0;
    sp->frame.z = -direction;
    sp->frame.y = cross(sp->frame.z, sp->frame.x);
    sp->position = direction;
    sp->gn = -direction;
    sp->isPoint = false;
    sp->atInfinity = true;
    sp->instIndex = sp->geomInstIndex = sp->primIndex = 0;
    sp->u = phi;
    sp->v = theta;
    phi += sc.envRotation;
    phi = phi - floorf(phi / (2 * kPiF)) * 2 * kPiF;
    sp->texU = phi / (2 * kPiF);
    sp->texV = theta / kPiF;
    float uvPDF = envMapPDF(sc, phi / (2 * kPiF), theta / kPiF);
    *hypAreaPDF = uvPDF / (2 * kPiF * kPiF * sinf(theta));
}

struct LightSample {
    SurfPt sp;
    float areaPDF;
    uint32_t posType;
    uint32_t material;
};

// selectSurfaceLight + sample{TriangleMesh, Point, InfiniteSphere}
// (light_transport_common.h:179-192, triangle.cu:113-177, point.cu:38-80, infinite_sphere.cu:60-114)
__device__ __forceinline__ void sampleLight(const DevScene &sc, float uLight, float uP0, float uP1, LightSample *ls,
                                            float *lightProb) {
    float instProb, uInst;
    uint32_t li = sampleDiscrete(sc.distPool, sc.lightInstBase, sc.numLightInsts, uLight, &instProb, &uInst);
    if (li == 0) {
        *lightProb = instProb;
        float u, v, uvPDF;
        envMapSample(sc, uP0, uP1, &u, &v, &uvPDF);
        float phi = 2 * kPiF * u;
        float theta = kPiF * v;
        float posPhi = phi - sc.envRotation;
        posPhi = posPhi - floorf(posPhi / (2 * kPiF)) * 2 * kPiF;
        V3 dir = fromPolarYUp(posPhi, theta);
        float sinPhi, cosPhi;
        sincosf(posPhi, &sinPhi, &cosPhi);
        V3 tc0 = normalize(V3(-cosPhi, 0.0f, -sinPhi));
        V3 gn = -dir;
        ls->sp.position = dir;
        ls->sp.frame.x = tc0;
        ls->sp.frame.z = gn;
        ls->sp.frame.y = cross(gn, tc0);
        ls->sp.gn = gn;
        ls->sp.isPoint = false;
        ls->sp.atInfinity = true;
        ls->sp.texU = phi / (2 * kPiF);
        ls->sp.texV = theta / kPiF;
        ls->sp.instIndex = ls->sp.geomInstIndex = ls->sp.primIndex = 0;
        ls->sp.u = posPhi;
        ls->sp.v = theta;
        ls->areaPDF = uvPDF / (2 * kPiF * kPiF * sinf(theta));
        ls->posType = dt::Reflection | dt::LowFreq;
        ls->material = sc.envMaterial;
        return;
    }
    const DevInstance &I = sc.instances[li - 1];
    float geomProb, uPrim;
    uint32_t lgi = sampleDiscrete(sc.distPool, I.lightDistBase, I.numGeomInsts, uInst, &geomProb, &uPrim);
    uint32_t gi = sc.instGeomInsts[I.geomInstListBase + lgi];
    const DevGeomInst &G = sc.geomInsts[gi];
    *lightProb = instProb * geomProb;
    float primProb;
    uint32_t prim = sampleDiscrete(sc.distPool, G.distBase, G.numPrims, uPrim, &primProb, nullptr);
    ls->material = G.material;
    ls->sp.instIndex = li - 1;
    ls->sp.geomInstIndex = gi;
    ls->sp.primIndex = prim;
    ls->sp.u = ls->sp.v = 0.0f;
    if (G.isPoints) {
        const DevVertex &v = sc.vertices[G.vertexBase + sc.indices[G.indexBase + prim]];
        ls->areaPDF = primProb / (4 * kPiF);
        ls->posType = dt::Reflection | dt::Delta0D;
        ls->sp.position = xfmPoint(I.o2w, ld3(v.p));
        V3 sn = normalize(xfmNormal(I.w2o, ld3(v.n)));
        V3 tc0 = normalize(xfmVector(I.o2w, ld3(v.t)));
        ls->sp.frame = Frame(tc0, sn);
        ls->sp.gn = sn;
        ls->sp.isPoint = true;
        ls->sp.atInfinity = false;
        ls->sp.texU = ls->sp.texV = 0.0f;
        return;
    }
    const uint32_t *ix = sc.indices + G.indexBase + 3 * prim;
    const DevVertex &v0 = sc.vertices[G.vertexBase + ix[0]];
    const DevVertex &v1 = sc.vertices[G.vertexBase + ix[1]];
    const DevVertex &v2 = sc.vertices[G.vertexBase + ix[2]];
    V3 e1 = xfmVector(I.o2w, ld3(v1.p) - ld3(v0.p));
    V3 e2 = xfmVector(I.o2w, ld3(v2.p) - ld3(v0.p));
    V3 gn = cross(e1, e2);
    float area = gn.length() / 2;
    gn = gn / (2 * area);
    ls->areaPDF = primProb / area;
    ls->posType = dt::Reflection | dt::LowFreq;
    float b0, b1;
    uniformSampleTriangle(uP0, uP1, &b0, &b1);
    float b2 = 1.0f - b0 - b1;
    V3 position = b0 * ld3(v0.p) + b1 * ld3(v1.p) + b2 * ld3(v2.p);
    V3 sn = b0 * ld3(v0.n) + b1 * ld3(v1.n) + b2 * ld3(v2.n);
    V3 tc0 = b0 * ld3(v0.t) + b1 * ld3(v1.t) + b2 * ld3(v2.t);
    ls->sp.texU = b0 * v0.u + b1 * v1.u + b2 * v2.u;
    ls->sp.texV = b0 * v0.v + b1 * v1.v + b2 * v2.v;
    position = xfmPoint(I.o2w, position);
    sn = normalize(xfmNormal(I.w2o, sn));
    tc0 = xfmVector(I.o2w, tc0);
    float dotNT = dot(sn, tc0);
    tc0 = normalize(tc0 - dotNT * sn);
    ls->sp.position = position;
    ls->sp.frame = Frame(tc0, sn);
    ls->sp.gn = gn;
    ls->sp.isPoint = false;
    ls->sp.atInfinity = false;
    ls->sp.u = b0;
    ls->sp.v = b1;
}

// EDF of an emitter material at a surface point (materials.cu:2040-2096, 2194-2250, 2830-2896):
// returns emittance M (evaluateEmittance) and the directional factor for dirLocal.
template <int N, bool NODES = true>
__device__ __forceinline__ bool evalEDF(const DevScene &sc, const DevMaterial &m, const SurfPt &sp, const Wls<N> &wls,
                                        const V3 &dirLocal, Spec<N> *M, float *fe) {
    switch (m.type) {
    case MAT_DIFFUSE_EMITTER: {
        NodeOutputs<N> o;
        if constexpr (NODES) materialNodes<N>(sc, m, &sp, wls, &o);
        else o.has = 0;
        *M = materialSpectrum<N>(sc, m, 0, o, wls) * m.scalars[0];
        *fe = dirLocal.z > 0.0f ? 1.0f / kPiF : 0.0f;
        return true;
    }
    case MAT_POINT_EMITTER: {
        NodeOutputs<N> o;
        if constexpr (NODES) materialNodes<N>(sc, m, &sp, wls, &o);
        else o.has = 0;
        *M = materialSpectrum<N>(sc, m, 0, o, wls) * m.scalars[0];
        *fe = 1.0f / (4 * kPiF);
        return true;
    }
    case MAT_ENVIRONMENT_EMITTER:
        *M = kPiF * envEmittance<N>(sc, m, sp.texU, sp.texV, wls);
        *fe = dirLocal.z > 0.0f ? 1.0f / kPiF : 0.0f;
        return true;
    default:
        return false;
    }
}

// ---------------------------------------------------------------------------------------------
// path-state packing
__device__ __forceinline__ uint32_t packFlags(uint32_t sel, bool single, uint32_t prevType, uint32_t pathLen) {
    return (sel & kFlagSelMask) | (single ? kFlagSingle : 0u) | ((prevType & 0x7Fu) << kFlagPrevTypeShift) |
           (pathLen << kFlagPathLenShift);
}

template <int N>
__device__ __forceinline__ Spec<N> loadSpec(const float4 &v) {
    Spec<N> s;
    s[0] = v.x; s[1] = v.y; s[2] = v.z;
    if constexpr (N == 4) s[3] = v.w;
    return s;
}
template <int N>
__device__ __forceinline__ float4 storeSpec(const Spec<N> &s) {
    return make_float4(s[0], s[1], s[2], N == 4 ? s[N - 1] : 0.0f);
}

struct KArgs {
    DevScene sc;
    FrameParams fp;
    FrameBuffers fb;
    uint32_t depth;
    // batches: path i belongs to the job j with jobOffsets[j] <= i < jobOffsets[j + 1]
    DevJob *jobs;
    const uint32_t *jobOffsets;
    // per step: where each job's pixels start among the pixels of the jobs taking part (exclusive
    // prefix over the jobs in order, inactive jobs counting zero), their total at [numJobs]
    uint32_t *activeOffsets;
    uint32_t numJobs;
};

__device__ __forceinline__ uint32_t jobOf(const KArgs &a, uint32_t path) {
    uint32_t lo = 0, hi = a.numJobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.jobOffsets[mid] <= path) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One block: a job takes part in this step while it has samples left; the exclusive prefix of
// the taking-part jobs' pixel counts places their pixels one after another in the step's queue.
constexpr uint32_t kStepBeginThreads = 256;
__global__ void __launch_bounds__(kStepBeginThreads) k_batchStepBegin(DevJob *jobs, uint32_t numJobs,
                                                                       uint32_t *activeOffsets) {
    __shared__ uint32_t waveTotals[kStepBeginThreads / 64];
    const uint32_t t = threadIdx.x, lane = __lane_id(), wave = t / 64;
    uint32_t carry = 0;
    for (uint32_t c = 0; c < numJobs; c += kStepBeginThreads) {
        const uint32_t j = c + t;
        uint32_t cnt = 0;
        if (j < numJobs) {
            DevJob &J = jobs[j];
            if (J.frameIndex < J.spp) {
                J.frameIndex += 1;
                J.active = 1;
                cnt = J.width * J.height;
            } else {
                J.active = 0;
            }
        }
        uint32_t incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t up = __shfl_up(incl, off);
            if (static_cast<int>(lane) >= off) incl += up;
        }
        if (lane == 63) waveTotals[wave] = incl;
        __syncthreads();
        uint32_t before = carry, total = carry;
        for (uint32_t w = 0; w < kStepBeginThreads / 64; ++w) {
            if (w < wave) before += waveTotals[w];
            total += waveTotals[w];
        }
        if (j < numJobs) activeOffsets[j] = before + incl - cnt;
        carry = total;
        __syncthreads();
    }
    if (t == 0) activeOffsets[numJobs] = carry;
}

__global__ void k_frameBegin(uint32_t *counters, uint32_t numPaths, uint32_t *sortCounts) {
    uint32_t i = threadIdx.x;
    for (uint32_t k = i; k < 2 * kMaxDepthCounters; k += blockDim.x) counters[k] = 0;
    if (sortCounts)
        for (uint32_t k = i; k < 2 * kSortClasses * kMaxDepthCounters; k += blockDim.x) sortCounts[k] = 0;
    if (i == 0) {
        counters[0] = numPaths;
        counters[kTailWork] = 0;
    }
}

// ---- generate: path_tracing.cu:30-70 ----
// Starts the path of pixel (px, py) of a W x H image: path state and primary ray.
template <int N>
__device__ __forceinline__ void generatePath(const FrameBuffers &fb, const DevCamera &cam, uint32_t W, uint32_t H,
                                             uint32_t px, uint32_t py, uint32_t path, float4 *rayOrg,
                                             float4 *rayDirOut) {
    Pcg32 rng{fb.ps.rng[path]};
    float pX = px + rng.uniform();
    float pY = py + rng.uniform();
    float wlOffset = rng.uniform();
    float wlSelect = rng.uniform();
    uint32_t sel = min(static_cast<uint32_t>(static_cast<uint16_t>(N * wlSelect)), static_cast<uint32_t>(N - 1));
    float selectWLPDF = N == 4 ? N / (830.0f - 360.0f) : 1.0f;
    float ul0 = rng.uniform(), ul1 = rng.uniform();
    float lx, ly;
    concentricSampleDisk(ul0, ul1, &lx, &ly);
    V3 c0(cam.rot[0], cam.rot[1], cam.rot[2]), c1(cam.rot[3], cam.rot[4], cam.rot[5]), c2(cam.rot[6], cam.rot[7], cam.rot[8]);
    V3 camN = normalize(c2);
    Frame cf;
    cf.z = camN;
    cf.x = normalize(c0);
    cf.y = cross(cf.z, cf.x);
    V3 orgLocal(cam.lensRadius * lx, cam.lensRadius * ly, 0.0f);
    V3 lensPos = (c0 * orgLocal.x + c1 * orgLocal.y + c2 * orgLocal.z) + ld3(cam.position);
    float lensAreaPDF = cam.lensRadius > 0.0f ? 1.0f / (kPiF * cam.lensRadius * cam.lensRadius) : 1.0f;
    float uD0 = pX / W, uD1 = pY / H;
    V3 pFocus(cam.opWidth * (0.5f - uD0), cam.opHeight * (0.5f - uD1), cam.objPlaneDistance);
    V3 dirLocal = normalize(pFocus - V3(cam.lensRadius * lx, cam.lensRadius * ly, 0.0f));
    float dirPDF = sq(cam.imgPlaneDistance) / (cube(dirLocal.z) * cam.imgPlaneArea);
    dirPDF *= W * H;
    V3 rayDir = cf.fromLocal(dirLocal);
    Spec<N> alpha(cam.sensitivity);
    alpha = alpha * (absDot(rayDir, camN) / (lensAreaPDF * dirPDF * selectWLPDF));
    float initImportance = alpha.importance(sel);
    fb.ps.rng[path] = rng.s;
    fb.ps.alpha[path] = storeSpec<N>(alpha);
    fb.ps.contrib[path] = make_float4(0, 0, 0, 0);
    uint32_t flags = packFlags(sel, false, dt::Reflection | dt::LowFreq, 0);
    fb.ps.meta[path] = make_float4(wlOffset, dirPDF, initImportance, __uint_as_float(flags));
    *rayOrg = make_float4(lensPos.x, lensPos.y, lensPos.z, __uint_as_float(path));
    *rayDirOut = make_float4(rayDir.x, rayDir.y, rayDir.z, 3.402823466e+38f);
}

// Queue entry i -> pixel of a whole W x H frame: bands of 8 rows, each cut into 8 x 8 tiles
// (the band's last tile as wide as the image leaves), so the 64 primary rays of a wave cover
// one tile and leave the camera close together.  A bijection of [0, W*H).
__device__ __forceinline__ void tilePixel(uint32_t i, uint32_t W, uint32_t H, uint32_t *px, uint32_t *py) {
    const uint32_t band = i / (8 * W);
    const uint32_t bandRows = min(8u, H - 8 * band);
    const uint32_t j = i - band * 8 * W;
    const uint32_t tile = j / (8 * bandRows);
    const uint32_t tileW = min(8u, W - 8 * tile);
    const uint32_t l = j - tile * 8 * bandRows;
    *px = 8 * tile + l % tileW;
    *py = 8 * band + l / tileW;
}

template <int N, bool MULTI = false>
__global__ void __launch_bounds__(256) k_generate(KArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (MULTI) {
        // queue entry i: the jobs taking part in this step one after another, each in 8 x 8 tiles
        const uint32_t total = a.activeOffsets[a.numJobs];
        if (i == 0) a.fb.counters[0] = total;
        if (i >= total) return;
        uint32_t lo = 0, hi = a.numJobs;   // the last job whose pixels start at or before i
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.activeOffsets[mid] <= i) lo = mid;
            else hi = mid;
        }
        const DevJob &J = a.jobs[lo];
        uint32_t px, py;
        tilePixel(i - a.activeOffsets[lo], J.width, J.height, &px, &py);
        float4 o, d;
        generatePath<N>(a.fb, J.scene.camera, J.width, J.height, px, py, J.pixelOffset + py * J.width + px, &o, &d);
        a.fb.rays[0].org[i] = o;
        a.fb.rays[0].dir[i] = d;
    } else {
        if (i >= a.fp.pixelCount) return;
        uint32_t px, py;
        if (a.fp.pixelCount == a.fp.imageW * a.fp.imageH) {
            tilePixel(i, a.fp.imageW, a.fp.imageH, &px, &py);
        } else {   // a part of a frame: its pixels in row-major order
            const uint32_t pix = a.fp.pixelBegin + i;
            px = pix % a.fp.imageW;
            py = pix / a.fp.imageW;
        }
        float4 o, d;
        generatePath<N>(a.fb, a.sc.camera, a.fp.imageW, a.fp.imageH, px, py, py * a.fp.width + px, &o, &d);
        a.fb.rays[0].org[i] = o;
        a.fb.rays[0].dir[i] = d;
    }
}

// wave-wide sum (all 64 lanes must be active)
__device__ __forceinline__ uint32_t waveSum(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
// adds one wave's traversal counters to three 64-bit statistics slots
__device__ __forceinline__ void addTravStats(uint32_t *counters, uint32_t slot, const TravCount &tc) {
    uint32_t n = waveSum(tc.nodes), t = waveSum(tc.tris), in = waveSum(tc.insts);
    if (__lane_id() == 0) {
        unsigned long long *c64 = reinterpret_cast<unsigned long long *>(counters + slot);
        atomicAdd(c64 + 0, static_cast<unsigned long long>(n));
        atomicAdd(c64 + 1, static_cast<unsigned long long>(t));
        atomicAdd(c64 + 2, static_cast<unsigned long long>(in));
    }
}

// ---- intersect: closest hit ----
template <bool STATS, bool MULTI = false, bool ALPHA = false>
__global__ void __launch_bounds__(kTraceBlock) k_intersect(KArgs a) {
    __shared__ int32_t lds[kIntersectLds];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const RayQueue q = a.fb.rays[a.depth & 1];
    const uint32_t count = a.fb.counters[2 * a.depth];
    TravCount tc;
    uint32_t cls = kSortClasses;   // shading class for the material sort; kSortClasses = none
    if (i < count) {
        float4 o = q.org[i], d = q.dir[i];
        TraceHit h;
        h.inst = -1;
        bool found;
        if constexpr (MULTI)
            found = traceRay<false, STATS, false, kIntersectStack>(a.jobs[jobOf(a, __float_as_uint(o.w))].scene,
                                                                   xyz(o), xyz(d), 0.0f, d.w, &h, lds, threadIdx.x, &tc);
        else if constexpr (ALPHA) {
            // the alpha test draws from the path's generator before its shading does
            const uint32_t path = __float_as_uint(o.w);
            Pcg32 rng{a.fb.ps.rng[path]};
            found = traceRay<false, STATS, true, kIntersectStack>(a.sc, xyz(o), xyz(d), 0.0f, d.w, &h, lds, threadIdx.x,
                                                                  &tc, &rng);
            a.fb.ps.rng[path] = rng.s;
        } else
            found = traceRay<false, STATS, false, kIntersectStack>(a.sc, xyz(o), xyz(d), 0.0f, d.w, &h, lds,
                                                                   threadIdx.x, &tc);
        if (found) {
            a.fb.hits.a[i] = make_int4(h.inst, h.geomInst, h.prim, __float_as_int(h.t));
            a.fb.hits.b[i] = make_float2(h.b1, h.b2);
        } else {
            a.fb.hits.a[i] = make_int4(-1, -1, -1, 0);
        }
        if (!MULTI && a.fp.sortShade) {
            cls = found ? min(1u + a.sc.materials[a.sc.geomInsts[h.geomInst].material].type, kSortClasses - 1) : 0u;
            a.fb.sortKey[i] = cls;
        }
    }
    if (!MULTI && a.fp.sortShade) {
        // class histogram of this depth's queue, one atomic per class present in the wave
        uint32_t *counts = a.fb.sortCounts + 2 * kSortClasses * a.depth;
        unsigned long long left = __ballot(cls < kSortClasses);
        const int lane = threadIdx.x & 63;
        while (left) {
            const int leader = __ffsll(static_cast<long long>(left)) - 1;
            const uint32_t c = __shfl(cls, leader);
            const unsigned long long m = __ballot(cls == c);
            if (lane == leader) atomicAdd(&counts[c], static_cast<uint32_t>(__popcll(m)));
            left &= ~m;
        }
    }
    if constexpr (STATS) addTravStats(a.fb.counters, kStatTravClosest, tc);
}

// Material sort: the queue entries of one depth in order of shading class (within a class,
// in order of arrival), so that the waves of k_shade run one material's code at a time.
__global__ void __launch_bounds__(256) k_sortScatter(KArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = a.fb.counters[2 * a.depth];
    uint32_t *counts = a.fb.sortCounts + 2 * kSortClasses * a.depth;
    const uint32_t cls = i < count ? a.fb.sortKey[i] : kSortClasses;
    uint32_t base = 0;
    if (cls < kSortClasses)
        for (uint32_t c = 0; c < cls; ++c) base += counts[c];
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long left = __ballot(cls < kSortClasses);
    uint32_t slot = 0;
    while (left) {
        const int leader = __ffsll(static_cast<long long>(left)) - 1;
        const uint32_t c = __shfl(cls, leader);
        const unsigned long long m = __ballot(cls == c);
        uint32_t first = 0;
        if (lane == leader) first = atomicAdd(&counts[kSortClasses + c], static_cast<uint32_t>(__popcll(m)));
        first = __shfl(first, leader);
        if (cls == c) slot = first + static_cast<uint32_t>(__popcll(m & below));
        left &= ~m;
    }
    if (cls < kSortClasses && base + slot < count) a.fb.sortPerm[base + slot] = i;
}

// ---- shade: path_tracing.cu:133-259 (hit) and :266-337 (miss) ----
// Per-path registers of the integrator between two bounces.
template <int N>
struct PathRegs {
    Spec<N> alpha, contrib;
    Wls<N> wls;
    float prevDirPDF, initImportance;
    uint32_t prevType;
    Pcg32 rng;
    bool rngUsed;
    float a0, a;    // radiance cache: path spread at the first vertex and accumulated since
};

template <int N>
__device__ __forceinline__ void loadPath(const FrameBuffers &fb, uint32_t path, PathRegs<N> *P) {
    P->alpha = loadSpec<N>(fb.ps.alpha[path]);
    P->contrib = loadSpec<N>(fb.ps.contrib[path]);
    float4 meta = fb.ps.meta[path];
    uint32_t flags = __float_as_uint(meta.w);
    P->wls.offset = meta.x;
    P->wls.sel = flags & kFlagSelMask;
    P->wls.single = (flags & kFlagSingle) != 0;
    P->prevDirPDF = meta.y;
    P->initImportance = meta.z;
    P->prevType = (flags >> kFlagPrevTypeShift) & 0x7Fu;
    P->rng.s = fb.ps.rng[path];
    P->rngUsed = false;
    P->a0 = P->a = 0.0f;
    if (fb.cache.spread) {
        float2 sp = fb.cache.spread[path];
        P->a0 = sp.x;
        P->a = sp.y;
    }
}

// Radiance-cache view of a frame for the shading code
struct CacheCtx {
    uint32_t mode, stamp, stride, offset, width, imageW;
    float threshold;
    // 1: the cache supplies all the radiance a path's last vertex reflects (its next-event
    // estimate included); 0: the vertex keeps its next-event estimate and the cache supplies
    // what a continued path would add after it
    uint32_t full;
    CacheBuffers b;
};
__device__ __forceinline__ CacheCtx cacheCtx(const FrameParams &fp, const FrameBuffers &fb) {
    CacheCtx c;
    c.mode = fp.cacheMode;
    c.full = fp.cacheFullRadiance;
    c.stamp = fp.frameStamp;
    c.stride = fp.trainStride;
    c.offset = fp.trainOffset;
    c.width = fp.width;
    c.imageW = fp.imageW;
    c.threshold = fp.cacheThreshold;
    c.b = fb.cache;
    return c;
}

// Raw radiance-cache input record of a surface point (4 float4; the layout the cache's
// encoder reads): position in the scene's bounding cube, outgoing direction (theta / pi,
// phi / 2 pi), shading normal (n + 1) / 2, roughness, diffuse and specular albedo, hero offset.
__device__ __forceinline__ void writeCacheRecord(const DevScene &sc, const SurfPt &sp, uint32_t material,
                                                 const V3 &rayDir, float wlOffset, float4 *dst) {
    const float inv = 0.5f / sc.worldRadius;
    const float px = fminf(fmaxf((sp.position.x - sc.center[0]) * inv + 0.5f, 0.0f), 1.0f);
    const float py = fminf(fmaxf((sp.position.y - sc.center[1]) * inv + 0.5f, 0.0f), 1.0f);
    const float pz = fminf(fmaxf((sp.position.z - sc.center[2]) * inv + 0.5f, 0.0f), 1.0f);
    const V3 wo = -rayDir;
    const float theta = acosf(fminf(fmaxf(wo.y, -1.0f), 1.0f)) / kPiF;
    const float phi = (atan2f(wo.z, wo.x) + kPiF) / (2 * kPiF);
    const V3 n = sp.frame.z;
    const float4 m0 = sc.cacheMaterials[2 * material], m1 = sc.cacheMaterials[2 * material + 1];
    dst[0] = make_float4(px, py, pz, theta);
    dst[1] = make_float4(phi, 0.5f * (n.x + 1.0f), 0.5f * (n.y + 1.0f), 0.5f * (n.z + 1.0f));
    dst[2] = make_float4(m0.w, m0.x, m0.y, m0.z);
    dst[3] = make_float4(m1.x, m1.y, m1.z, wlOffset);
}

// One bounce of path_tracing.cu:133-337 for a ray that hit (ha.x >= 0) or missed: implicit
// emission with MIS, Russian roulette, next-event estimation (returns the shadow ray and the
// contribution it carries when unoccluded) and BSDF sampling (returns the next ray).
template <int N, bool NODES = true>
__device__ __forceinline__ void shadeStep(const DevScene &sc, uint32_t pathLength, uint32_t maxPathLength,
                                          uint32_t path, const V3 &rayOrg, const V3 &rayDir, int4 ha, float2 hb,
                                          PathRegs<N> &P, const CacheCtx &cc, bool *emitShadow, float4 *shOrg,
                                          float4 *shDir, float4 *shVal, bool *emitNext, float4 *nextOrg,
                                          float4 *nextDir) {
    *emitShadow = false;
    *emitNext = false;
    const bool maxLengthTerminate = pathLength >= maxPathLength;
    if (cc.mode != 0 && !cc.full && pathLength >= 2 && pathLength - 1 <= kCacheTrainVerts) {
        // training path: the contribution before this vertex's emission closes the previous
        // vertex's sample (everything it gathered up to and including its next-event estimate)
        const uint32_t pix = (path / cc.width) * cc.imageW + path % cc.width;
        if (pix % cc.stride == cc.offset) {
            const uint32_t slot = (pix / cc.stride) * kCacheTrainVerts + pathLength - 2;
            cc.b.trainContrib[slot] = storeSpec<N>(P.contrib);
            cc.b.trainAfterStamp[slot] = cc.stamp;
        }
    }
    if (ha.x < 0) {
        // miss: environment light with MIS
        if (sc.hasEnv && sc.envImportance > 0.0f) {
            SurfPt sp;
            float hypAreaPDF;
            envPointFromDir(sc, rayDir, &sp, &hypAreaPDF);
            const DevMaterial &m = sc.materials[sc.envMaterial];
            V3 dirOutLocal = sp.frame.toLocal(-rayDir);
            Spec<N> M;
            float fe;
            evalEDF<N, NODES>(sc, m, sp, P.wls, dirOutLocal, &M, &fe);
            if (!M.isZero()) {
                float mis = 1.0f;
                if (!dt::isDelta(P.prevType) && pathLength > 1) {
                    float instProb = 1.0f / sc.lightInstIntegral;
                    float bsdfPDF = P.prevDirPDF;
                    float lightPDF = instProb * hypAreaPDF * 1.0f / fabsf(dirOutLocal.z);
                    mis = (bsdfPDF * bsdfPDF) / (lightPDF * lightPDF + bsdfPDF * bsdfPDF);
                }
                P.contrib += P.alpha * (M * fe) * mis;
            }
        }
        return;
    }
    SurfPt sp;
    float hypAreaPDF;
    decodeHit<NODES>(sc, ha.x, ha.y, ha.z, hb.x, hb.y, &sp, &hypAreaPDF);
    const DevGeomInst &G = sc.geomInsts[ha.y];
    const DevMaterial &mat = sc.materials[G.material];
    V3 dirOutLocal = sp.frame.toLocal(-rayDir);
    if (mat.isEmitter) {
        Spec<N> M;
        float fe;
        evalEDF<N, NODES>(sc, mat, sp, P.wls, dirOutLocal, &M, &fe);
        if (!M.isZero()) {
            float mis = 1.0f;
            if (!dt::isDelta(P.prevType) && pathLength > 1) {
                const DevInstance &I = sc.instances[ha.x];
                float instProb = I.lightDistIntegral / sc.lightInstIntegral;
                float geomInstProb = G.importance / I.lightDistIntegral;
                float bsdfPDF = P.prevDirPDF;
                float dist2 = (sp.position - rayOrg).sqLength();
                float lightPDF = instProb * geomInstProb * hypAreaPDF * dist2 / fabsf(dirOutLocal.z);
                mis = (bsdfPDF * bsdfPDF) / (lightPDF * lightPDF + bsdfPDF * bsdfPDF);
            }
            P.contrib += P.alpha * (M * fe) * mis;
        }
    }
    bool cacheEnd = false;
    if (cc.mode != 0) {
        // path spread (Mueller et al. 2021, "Real-time Neural Radiance Caching for Path
        // Tracing"): a_0 from the primary vertex, then one term per non-delta bounce
        const float dist2 = (sp.position - rayOrg).sqLength();
        const float cosT = fabsf(dirOutLocal.z);
        if (pathLength == 1) {
            P.a0 = sqrtf(dist2 / (4 * kPiF * cosT));
            P.a = 0.0f;
        } else if (!dt::isDelta(P.prevType)) {
            P.a += sqrtf(dist2 / (P.prevDirPDF * cosT));
        }
        const uint32_t pix = (path / cc.width) * cc.imageW + path % cc.width;
        if (pix % cc.stride == cc.offset) {
            // training path: vertices 1..kCacheTrainVerts become samples once the path ends
            if (pathLength <= kCacheTrainVerts) {
                const uint32_t slot = (pix / cc.stride) * kCacheTrainVerts + pathLength - 1;
                writeCacheRecord(sc, sp, G.material, rayDir, N == 4 ? P.wls.offset : 0.0f, cc.b.trainRaw + 4 * slot);
                cc.b.trainThroughput[slot] = storeSpec<N>(P.alpha);
                cc.b.trainStamp[slot] = cc.stamp;
                if (cc.full) {
                    // the sample is everything the path gathers after this vertex's emission
                    cc.b.trainContrib[slot] = storeSpec<N>(P.contrib);
                    cc.b.trainAfterStamp[slot] = cc.stamp;
                }
            }
        } else {
            cacheEnd = cc.mode == 2 && pathLength >= 2 && !maxLengthTerminate && !P.wls.single &&
                       P.a > cc.threshold * P.a0;
        }
    }
    if (maxLengthTerminate) return;
    P.rngUsed = true;
    if (!cacheEnd) {
        float continueProb = fminf(P.alpha.importance(P.wls.sel) / P.initImportance, 1.0f);
        if (P.rng.uniform() >= continueProb) return;
        P.alpha = P.alpha / continueProb;
    }
    Bsdf<N> bsdf = setupBsdf<N, NODES>(sc, mat, P.wls, &sp);
    if (cacheEnd && cc.full) {
        if (bsdf.hasNonDelta()) {
            // the cache supplies all the radiance this vertex reflects; no next-event estimate
            writeCacheRecord(sc, sp, G.material, rayDir, N == 4 ? P.wls.offset : 0.0f, cc.b.queryRaw + 4 * path);
            cc.b.queryThroughput[path] = storeSpec<N>(P.alpha);
            cc.b.queryStamp[path] = cc.stamp;
            return;
        }
        // a purely specular vertex does not end in the cache: the path goes on
        cacheEnd = false;
        float continueProb = fminf(P.alpha.importance(P.wls.sel) / P.initImportance, 1.0f);
        if (P.rng.uniform() >= continueProb) return;
        P.alpha = P.alpha / continueProb;
    }
    V3 gnL = sp.frame.toLocal(sp.gn);
    if (bsdf.hasNonDelta()) {
        float uLight = P.rng.uniform();
        float uP0 = P.rng.uniform(), uP1 = P.rng.uniform();
        // a scene without emitters has nothing to sample (the draws are still made)
        if (sc.lightInstIntegral > 0.0f) {
            LightSample ls;
            float lightProb;
            sampleLight(sc, uLight, uP0, uP1, &ls, &lightProb);
            const DevMaterial &lm = sc.materials[ls.material];
            // shadow ray direction (SurfacePoint::calcDirectionFrom)
            V3 sdir;
            float dist2;
            if (ls.sp.atInfinity) {
                dist2 = 1.0f;
                sdir = normalize(ls.sp.position);
            } else {
                V3 dv = ls.sp.position - sp.position;
                dist2 = dv.sqLength();
                sdir = dv / sqrtf(dist2);
            }
            V3 sdl = ls.sp.frame.toLocal(-sdir);
            Spec<N> M;
            float fe;
            evalEDF<N, NODES>(sc, lm, ls.sp, P.wls, sdl, &M, &fe);
            if (!M.isZero()) {
                float rec = 1.0f / dist2;
                V3 sdsn = sp.frame.toLocal(sdir);
                float lightPDF = lightProb * ls.areaPDF;
                Spec<N> fs = bsdf.evaluate(dirOutLocal, gnL, sdsn);
                float cosLight = ls.sp.isPoint ? 1.0f : absDot(-sdir, ls.sp.gn);
                float bsdfPDF = bsdf.evaluatePDF(dirOutLocal, sdsn, P.wls.sel) * cosLight * rec;
                float mis = 1.0f;
                if (!dt::isDelta(ls.posType) && !isinf(lightPDF))
                    mis = (lightPDF * lightPDF) / (lightPDF * lightPDF + bsdfPDF * bsdfPDF);
                float Gf = 1.0f * absDot(sdsn, gnL) * cosLight * rec;
                float scalarCoeff = Gf * mis / lightPDF;
                Spec<N> value = P.alpha * (M * fe) * fs * scalarCoeff;
                // testVisibility (light_transport_common.h:147-177)
                bool front = dot(sp.gn, sdir) > 0;
                V3 so = offsetRayOrigin(sp.position, front ? sp.gn : -sp.gn);
                front = -dot(ls.sp.gn, sdir) > 0;
                V3 lo = offsetRayOrigin(ls.sp.position, front ? ls.sp.gn : -ls.sp.gn);
                float tmax = 3.402823466e+38f;
                if (!ls.sp.atInfinity) tmax = (lo - so).length() * 0.9999f;
                *emitShadow = true;
                *shOrg = make_float4(so.x, so.y, so.z, __uint_as_float(path));
                *shDir = make_float4(sdir.x, sdir.y, sdir.z, tmax);
                *shVal = storeSpec<N>(value);
            }
        }
    }
    if (cacheEnd) {
        // the cache supplies the indirect radiance leaving this vertex (what a continued path
        // would add after the next-event estimate above); the path ends here
        writeCacheRecord(sc, sp, G.material, rayDir, N == 4 ? P.wls.offset : 0.0f, cc.b.queryRaw + 4 * path);
        cc.b.queryThroughput[path] = storeSpec<N>(P.alpha);
        cc.b.queryStamp[path] = cc.stamp;
        return;
    }
    float uC = P.rng.uniform(), uD0 = P.rng.uniform(), uD1 = P.rng.uniform();
    V3 dirL;
    float dirPDF;
    uint32_t sType;
    Spec<N> fs = bsdf.sample(dirOutLocal, gnL, P.wls.sel, uC, uD0, uD1, &dirL, &dirPDF, &sType);
    if (fs.isZero() || dirPDF == 0.0f) return;
    if ((sType & dt::Dispersive) && !P.wls.single) {
        dirPDF /= N;
        P.wls.single = true;
    }
    float cosFactor = dot(dirL, gnL);
    P.alpha = P.alpha * (fs * (fabsf(cosFactor) / dirPDF));
    V3 dirIn = sp.frame.fromLocal(dirL);
    V3 no = offsetRayOrigin(sp.position, cosFactor > 0.0f ? sp.gn : -sp.gn);
    *emitNext = true;
    *nextOrg = make_float4(no.x, no.y, no.z, __uint_as_float(path));
    *nextDir = make_float4(dirIn.x, dirIn.y, dirIn.z, 3.402823466e+38f);
    P.prevDirPDF = dirPDF;
    P.prevType = sType;
}

template <int N>
__device__ __forceinline__ void storePath(const FrameBuffers &fb, uint32_t path, uint32_t pathLength,
                                          const PathRegs<N> &P, bool continues) {
    fb.ps.contrib[path] = storeSpec<N>(P.contrib);
    if (P.rngUsed) fb.ps.rng[path] = P.rng.s;
    if (continues) {
        fb.ps.alpha[path] = storeSpec<N>(P.alpha);
        fb.ps.meta[path] = make_float4(P.wls.offset, P.prevDirPDF, P.initImportance,
                                       __uint_as_float(packFlags(P.wls.sel, P.wls.single, P.prevType, pathLength)));
        if (fb.cache.spread) fb.cache.spread[path] = make_float2(P.a0, P.a);
    }
}

// Denoiser guides of a path's first hit (path_tracing.cu:95-108, 154-157, 272-275;
// lvc_bpt.cu:528-540; aux_buffer_generator.cu:57-67): the BSDF's base colour times the D65
// white point over the wavelength pdf, added to the pixel's strata like a colour sample
// (without compensation), and the shading normal; a miss adds zero.  Path index = pixel index
// (row stride fp.width).
template <int N, bool NODES>
__device__ __forceinline__ void addGuides(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                          uint32_t path, const int4 &ha, const float2 &hb, const Wls<N> &wls,
                                          float selectWLPDF) {
    Spec<N> albedo(0.0f);
    V3 n(0.0f);
    if (ha.x >= 0) {
        SurfPt sp;
        float hypAreaPDF;
        decodeHit<NODES>(sc, ha.x, ha.y, ha.z, hb.x, hb.y, &sp, &hypAreaPDF);
        const DevMaterial &mat = sc.materials[sc.geomInsts[ha.y].material];
        albedo = setupBsdf<N, NODES>(sc, mat, wls, &sp).baseColor() * evalSpectrum<N>(sc, sc.whitePoint, wls) /
                 selectWLPDF;
        n = sp.frame.z;
    }
    const size_t plane = static_cast<size_t>(fp.width) * fp.height;
    const bool reset = fp.guideReset != 0;
    float *alb = fb.guideAlbedo + path;
    if constexpr (N == 4) {
        const float recBinWidth = 16 / (830.0f - 360.0f);
        float addend[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) addend[k] = 0.0f;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            float lam = 360.0f + (830.0f - 360.0f) * (w + wls.offset) / 4;
            uint32_t bin = min(static_cast<uint32_t>((lam - 360.0f) / (830.0f - 360.0f) * 16), 15u);
#pragma unroll
            for (int k = 0; k < 16; ++k)
                if (k == static_cast<int>(bin)) addend[k] += albedo[w] * recBinWidth;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) alb[k * plane] = (reset ? 0.0f : alb[k * plane]) + addend[k];
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) alb[k * plane] = (reset ? 0.0f : alb[k * plane]) + albedo[k];
    }
    float *nrm = fb.guideNormal + path;
    nrm[0] = (reset ? 0.0f : nrm[0]) + n.x;
    nrm[plane] = (reset ? 0.0f : nrm[plane]) + n.y;
    nrm[2 * plane] = (reset ? 0.0f : nrm[2 * plane]) + n.z;
}

template <int N, bool NODES>
__device__ __forceinline__ void firstHitGuides(const KArgs &a, uint32_t path, const int4 &ha, const float2 &hb,
                                               const Wls<N> &wls) {
    addGuides<N, NODES>(a.sc, a.fp, a.fb, path, ha, hb, wls, N == 4 ? 4 / (830.0f - 360.0f) : 1.0f);
}

// The light tracer's guides (aux_buffer_generator.cu:22-93, launched before every light-tracing
// frame, context.cpp:1449-1451): per pixel a primary ray drawn from the pixel's generator as the
// path tracer draws one (pixel jitter, wavelengths, lens point), traced to its first hit (alpha
// tests draw from the same generator), whose guides are added.
template <int N, bool NODES>
__global__ void __launch_bounds__(kTraceBlock) k_ltGuides(KArgs a) {
    __shared__ int32_t lds[kTraceLds];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.fp.imageW * a.fp.imageH) return;
    const uint32_t px = i % a.fp.imageW, py = i / a.fp.imageW;
    const uint32_t path = py * a.fp.width + px;
    float4 o, d;
    generatePath<N>(a.fb, a.sc.camera, a.fp.imageW, a.fp.imageH, px, py, path, &o, &d);
    Pcg32 rng{a.fb.ps.rng[path]};
    TraceHit h;
    h.inst = -1;
    h.b1 = h.b2 = 0.0f;
    const bool found = traceRay<false, false, NODES>(a.sc, xyz(o), xyz(d), 0.0f, d.w, &h, lds, threadIdx.x, nullptr,
                                                     &rng);
    a.fb.ps.rng[path] = rng.s;
    const float4 meta = a.fb.ps.meta[path];
    Wls<N> wls;
    wls.offset = meta.x;
    wls.sel = __float_as_uint(meta.w) & kFlagSelMask;
    wls.single = false;
    const int4 ha = found ? make_int4(h.inst, h.geomInst, h.prim, 0) : make_int4(-1, -1, -1, 0);
    firstHitGuides<N, NODES>(a, path, ha, make_float2(h.b1, h.b2), wls);
}

template <int N, bool MULTI = false, bool NODES = true, bool GUIDES = false>
__global__ void __launch_bounds__(256) k_shade(KArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t depth = a.depth;
    const uint32_t count = a.fb.counters[2 * depth];
    bool emitNext = false, emitShadow = false;
    float4 nextOrg, nextDir, shOrg, shDir, shVal;
    if (i < count) {
        const RayQueue q = a.fb.rays[depth & 1];
        const uint32_t j = (!MULTI && a.fp.sortShade) ? a.fb.sortPerm[i] : i;   // queue entry
        float4 o4 = q.org[j], d4 = q.dir[j];
        const uint32_t path = __float_as_uint(o4.w);
        int4 ha = a.fb.hits.a[j];
        float2 hb = ha.x >= 0 ? a.fb.hits.b[j] : make_float2(0.0f, 0.0f);
        PathRegs<N> P;
        loadPath<N>(a.fb, path, &P);
        if constexpr (GUIDES) {
            if (depth == 0) firstHitGuides<N, NODES>(a, path, ha, hb, P.wls);
        }
        if constexpr (MULTI) {
            const DevJob &J = a.jobs[jobOf(a, path)];
            shadeStep<N, NODES>(J.scene, depth + 1, J.maxPathLength, path, xyz(o4), xyz(d4), ha, hb, P, cacheCtx(a.fp, a.fb),
                         &emitShadow, &shOrg, &shDir, &shVal, &emitNext, &nextOrg, &nextDir);
        } else {
            shadeStep<N, NODES>(a.sc, depth + 1, a.fp.maxPathLength, path, xyz(o4), xyz(d4), ha, hb, P, cacheCtx(a.fp, a.fb),
                         &emitShadow, &shOrg, &shDir, &shVal, &emitNext, &nextOrg, &nextDir);
        }
        storePath<N>(a.fb, path, depth + 1, P, emitNext);
    }
    // queue appends (every lane of the wave takes part in the ballots)
    uint32_t slot = appendSlot(&a.fb.counters[2 * (depth + 1)], emitNext);
    if (emitNext) {
        const RayQueue nq = a.fb.rays[(depth + 1) & 1];
        nq.org[slot] = nextOrg;
        nq.dir[slot] = nextDir;
    }
    uint32_t sslot = appendSlot(&a.fb.counters[2 * depth + 1], emitShadow);
    if (emitShadow) {
        a.fb.shadow.org[sslot] = shOrg;
        a.fb.shadow.dir[sslot] = shDir;
        a.fb.shadow.value[sslot] = shVal;
    }
}

// ---- shadow: any-hit, adds the NEE contribution when unoccluded ----
template <bool STATS, bool MULTI = false, bool ALPHA = false>
__global__ void __launch_bounds__(kTraceBlock) k_shadow(KArgs a) {
    __shared__ int32_t lds[kShadowLds];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = a.fb.counters[2 * a.depth + 1];
    TravCount tc;
    if (i < count) {
        float4 o = a.fb.shadow.org[i], d = a.fb.shadow.dir[i];
        TraceHit h;
        bool occluded;
        float vis = 1.0f;
        if constexpr (MULTI)
            occluded = traceRay<true, STATS>(a.jobs[jobOf(a, __float_as_uint(o.w))].scene, xyz(o), xyz(d), 0.0f,
                                             d.w, &h, lds, threadIdx.x, &tc);
        else
            occluded = traceRay<true, STATS, ALPHA>(a.sc, xyz(o), xyz(d), 0.0f, d.w, &h, lds, threadIdx.x, &tc,
                                                    nullptr, &vis);
        if (!occluded) {
            uint32_t path = __float_as_uint(o.w);
            float4 v = a.fb.shadow.value[i];
            if constexpr (ALPHA) v = make_float4(v.x * vis, v.y * vis, v.z * vis, v.w * vis);
            float4 c = a.fb.ps.contrib[path];
            a.fb.ps.contrib[path] = make_float4(c.x + v.x, c.y + v.y, c.z + v.z, c.w + v.w);
        }
    }
    if constexpr (STATS) addTravStats(a.fb.counters, kStatTravShadow, tc);
}

// ---- tail: from depth a.depth on, each path runs intersect -> shade -> shadow in a loop ----
// Persistent waves; a lane whose path has ended takes the next queue entry at the top of the
// loop (one atomic per wave for all lanes that need work), so lanes stay busy while the queue
// lasts instead of waiting for the longest path of a fixed chunk.  Every path consumes its
// random numbers and adds its contributions in the same order as the per-depth kernels, so a
// frame's result does not depend on the depth at which paths switch to this kernel.  Rays per
// depth are counted in LDS and added to the frame counters once per block.
template <int N, bool STATS, bool MULTI = false, bool NODES = true>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(kTailWavesPerEU)))
k_tail(KArgs a) {
    __shared__ int32_t lds[kTraceLds];
    __shared__ uint32_t depthCounts[2 * kMaxDepthCounters];
    const DevScene &sc = a.sc;
    const uint32_t depth0 = a.depth;
    const uint32_t count = a.fb.counters[2 * depth0];
    const RayQueue q = a.fb.rays[depth0 & 1];
    const CacheCtx cc = cacheCtx(a.fp, a.fb);
    for (uint32_t k = threadIdx.x; k < 2 * kMaxDepthCounters; k += blockDim.x) depthCounts[k] = 0;
    __syncthreads();
    TravCount tc, tcShadow;
    const uint32_t lane = __lane_id();
    bool active = false, drained = false;
    uint32_t path = 0, depth = depth0, job = 0;
    PathRegs<N> P;
    V3 org(0.0f), dir(0.0f, 0.0f, 1.0f);
    float tmax = 0.0f;
    while (true) {
        if (!drained) {
            const uint64_t need = __ballot(!active);
            if (need) {
                const uint32_t n = __popcll(need);
                const uint32_t leader = __ffsll(static_cast<unsigned long long>(need)) - 1;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(&a.fb.counters[kTailWork], n);
                base = __shfl(base, static_cast<int>(leader));
                if (base + n >= count) drained = true;
                if (!active) {
                    const uint32_t i = base + __popcll(need & ((1ull << lane) - 1ull));
                    if (i < count) {
                        const float4 o4 = q.org[i], d4 = q.dir[i];
                        path = __float_as_uint(o4.w);
                        if constexpr (MULTI) job = jobOf(a, path);
                        loadPath<N>(a.fb, path, &P);
                        org = xyz(o4);
                        dir = xyz(d4);
                        tmax = d4.w;
                        depth = depth0;
                        active = true;
                    }
                }
            }
        }
        if (!__ballot(active)) break;
        if (!active) continue;
        const DevScene &psc = MULTI ? a.jobs[job].scene : sc;
        const uint32_t maxLen = MULTI ? a.jobs[job].maxPathLength : a.fp.maxPathLength;
        TraceHit h;
        h.inst = -1;
        h.b1 = h.b2 = 0.0f;
        constexpr bool kAlpha = NODES && !MULTI;
        if constexpr (kAlpha) P.rngUsed = P.rngUsed || sc.hasAlpha;
        bool found = traceRay<false, STATS, kAlpha>(psc, org, dir, 0.0f, tmax, &h, lds, threadIdx.x, &tc, &P.rng);
        int4 ha = found ? make_int4(h.inst, h.geomInst, h.prim, 0) : make_int4(-1, -1, -1, 0);
        bool emitShadow, emitNext;
        float4 shOrg, shDir, shVal, nextOrg, nextDir;
        shadeStep<N, NODES>(psc, depth + 1, maxLen, path, org, dir, ha, make_float2(h.b1, h.b2), P, cc, &emitShadow,
                     &shOrg, &shDir, &shVal, &emitNext, &nextOrg, &nextDir);
        if (emitShadow) {
            atomicAdd(&depthCounts[2 * depth + 1], 1u);
            TraceHit sh;
            float vis = 1.0f;
            bool occluded = traceRay<true, STATS, kAlpha>(psc, xyz(shOrg), xyz(shDir), 0.0f, shDir.w, &sh, lds,
                                                          threadIdx.x, &tcShadow, nullptr, &vis);
            if (!occluded) P.contrib += loadSpec<N>(shVal) * vis;
        }
        if (!emitNext) {
            storePath<N>(a.fb, path, depth + 1, P, false);
            active = false;
            continue;
        }
        ++depth;
        atomicAdd(&depthCounts[2 * depth], 1u);
        org = xyz(nextOrg);
        dir = xyz(nextDir);
        tmax = nextDir.w;
    }
    if constexpr (STATS) {
        addTravStats(a.fb.counters, kStatTravClosest, tc);
        addTravStats(a.fb.counters, kStatTravShadow, tcShadow);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 2 * kMaxDepthCounters; k += blockDim.x)
        if (depthCounts[k]) atomicAdd(&a.fb.counters[k], depthCounts[k]);
}

// ---- radiance cache: training samples from the finished training paths ----
// target of slot k (vertex v of training path t) = (C_final - C_v) / T_v per channel: the
// radiance the path gathered after the next-event estimate at v (the indirect light leaving
// v, Russian roulette included), divided by the throughput on arrival at v.
template <int N>
__device__ __forceinline__ bool cacheTarget(const KArgs &a, uint32_t k, float4 *target) {
    const FrameParams &fp = a.fp;
    const CacheBuffers &cb = a.fb.cache;
    if (cb.trainStamp[k] != fp.frameStamp) return false;
    const uint32_t pix = (k / kCacheTrainVerts) * fp.trainStride + fp.trainOffset;
    if (pix >= fp.imageW * fp.imageH) return false;
    const uint32_t path = (pix / fp.imageW) * fp.width + pix % fp.imageW;
    const float4 cf = a.fb.ps.contrib[path], T = cb.trainThroughput[k];
    // a path that ended at this vertex gathered nothing after its next-event estimate
    const float4 c = cb.trainAfterStamp[k] == fp.frameStamp ? cb.trainContrib[k] : cf;
    float t[4] = {(cf.x - c.x) / T.x, (cf.y - c.y) / T.y, (cf.z - c.z) / T.z, N == 4 ? (cf.w - c.w) / T.w : 0.0f};
    const float tt[4] = {T.x, T.y, T.z, T.w};
    for (int ch = 0; ch < N; ++ch)
        if (!(tt[ch] > 0.0f) || !isfinite(t[ch])) return false;
    *target = make_float4(t[0], t[1], t[2], t[3]);
    return true;
}

// Compaction of the live slots in slot order (deterministic): flags and targets per slot, an
// exclusive scan in blocks of 1024 slots plus one scan of the block totals, then a gather.
template <int N>
__global__ void __launch_bounds__(256) k_rcSlotFlags(KArgs a, uint32_t numSlots, uint32_t *flags,
                                                      float4 *targets) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= numSlots) return;
    float4 tg = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    flags[k] = cacheTarget<N>(a, k, &tg) ? 1u : 0u;
    targets[k] = tg;
}

constexpr uint32_t kSlotScanBlock = 1024;
__global__ void __launch_bounds__(kSlotScanBlock) k_rcScanSlots(const uint32_t *flags, uint32_t numSlots,
                                                                uint32_t *offsets, uint32_t *blockSums) {
    __shared__ uint32_t waveTotals[kSlotScanBlock / 64];
    const uint32_t k = blockIdx.x * kSlotScanBlock + threadIdx.x;
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
    const uint32_t v = k < numSlots ? flags[k] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= static_cast<uint32_t>(off)) x += y;
    }
    if (lane == 63) waveTotals[wave] = x;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t t = threadIdx.x < kSlotScanBlock / 64 ? waveTotals[threadIdx.x] : 0u, st = t;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t y = __shfl_up(st, off);
            if (lane >= static_cast<uint32_t>(off)) st += y;
        }
        if (threadIdx.x < kSlotScanBlock / 64) waveTotals[threadIdx.x] = st - t;
        if (threadIdx.x == kSlotScanBlock / 64 - 1) blockSums[blockIdx.x] = st;
    }
    __syncthreads();
    if (k < numSlots) offsets[k] = waveTotals[wave] + x - v;
}

// one wave: exclusive scan of at most 64 block totals, and the total
__global__ void __launch_bounds__(64) k_rcScanSlotBlocks(uint32_t *blockSums, uint32_t numBlocks, uint32_t *total) {
    const uint32_t lane = threadIdx.x;
    const uint32_t v = lane < numBlocks ? blockSums[lane] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= static_cast<uint32_t>(off)) x += y;
    }
    if (lane < numBlocks) blockSums[lane] = x - v;
    if (lane == 63) *total = x;
}

__global__ void __launch_bounds__(256) k_rcGatherSlots(const float4 *trainRaw, const uint32_t *flags,
                                                        const uint32_t *offsets, const uint32_t *blockOffsets,
                                                        const float4 *targets, uint32_t numSlots, float4 *batchRaw,
                                                        float4 *batchTarget) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= numSlots || !flags[k]) return;
    const uint32_t out = blockOffsets[k / kSlotScanBlock] + offsets[k];
    const float4 *src = trainRaw + 4 * k;
    float4 *dst = batchRaw + 4 * out;
    dst[0] = src[0];
    dst[1] = src[1];
    dst[2] = src[2];
    dst[3] = src[3];
    batchTarget[out] = targets[k];
}

// ---- accumulate: Kahan sum per channel (common_internal.h:350-368, spectrum_types.h:674-684) ----
// XYZ (through the 16 CMF bins, spectral) or linear Rec.709 (RGB) of a pixel's sums, averaged
// over the frames, as Rec.709 scaled by the image area (post_process.cu:93-163)
template <int N>
__device__ __forceinline__ float4 convertSums(const DevScene &sc, const float *sums, uint32_t frames, uint32_t W,
                                              uint32_t H) {
    float X, Y, Z;
    if constexpr (N == 4) {
        X = Y = Z = 0.0f;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            X += sc.cmfX[k] * sums[k];
            Y += sc.cmfY[k] * sums[k];
            Z += sc.cmfZ[k] * sums[k];
        }
        X /= sc.integralCMF;
        Y /= sc.integralCMF;
        Z /= sc.integralCMF;
    } else {
        const float r = sums[0], g = sums[1], b = sums[2];
        X = 0.4124564f * r + 0.3575761f * g + 0.1804375f * b;
        Y = 0.2126729f * r + 0.7151522f * g + 0.0721750f * b;
        Z = 0.0193339f * r + 0.1191920f * g + 0.9503041f * b;
    }
    float rec = 1.0f / frames;
    X *= rec; Y *= rec; Z *= rec;
    float R = 3.2404542f * X + -1.5371385f * Y + -0.4985314f * Z;
    float G = -0.9692660f * X + 1.8760108f * Y + 0.0415560f * Z;
    float B = 0.0556434f * X + -0.2040259f * Y + 1.0572252f * Z;
    float res = static_cast<float>(W) * static_cast<float>(H);
    return make_float4(R * res, G * res, B * res, 1.0f);
}

// CONVERT (single renders): also writes the pixel's output from the sums just stored, which
// k_convert would read back.
template <int N, bool MULTI = false, bool CONVERT = false>
__global__ void __launch_bounds__(256) k_accumulate(KArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t numPaths = a.fp.imageW * a.fp.imageH;
    if (i >= numPaths) return;
    uint32_t p;
    size_t plane;
    float *acc;
    bool reset;
    if constexpr (MULTI) {
        const DevJob &J = a.jobs[jobOf(a, i)];
        if (!J.active) return;
        p = i;
        plane = static_cast<size_t>(J.width) * J.height;
        acc = J.accum + (i - J.pixelOffset);
        reset = J.frameIndex == 1;
    } else {
        if (i >= a.fp.pixelCount) return;
        const uint32_t pix = a.fp.pixelBegin + i;
        const uint32_t px = pix % a.fp.imageW, py = pix / a.fp.imageW;
        p = py * a.fp.width + px;
        // planes of width*height floats: sum of channel k at plane 2k, compensation at 2k + 1
        plane = static_cast<size_t>(a.fp.width) * a.fp.height;
        acc = a.fb.accum + p;
        reset = a.fp.frameIndex == 1;
    }
    constexpr int C = N == 4 ? 16 : 3;
    Spec<N> c = loadSpec<N>(a.fb.ps.contrib[p]);
    bool finite = c.allFinite();
    if (!finite) {
        atomicAdd(&a.fb.counters[kStatDropped], 1u);
        if (reset) {
            for (int k = 0; k < 2 * C; ++k) acc[k * plane] = 0.0f;
        }
        if constexpr (CONVERT && !MULTI) {
            float sums[C];
#pragma unroll
            for (int k = 0; k < C; ++k) sums[k] = reset ? 0.0f : acc[(2 * k) * plane];
            a.fb.output[p] = convertSums<N>(a.sc, sums, a.fp.frameIndex, a.fp.imageW, a.fp.imageH);
        }
        return;
    }
    float addend[C];
    if constexpr (N == 4) {
        float wlOffset = a.fb.ps.meta[p].x;
        const float recBinWidth = 16 / (830.0f - 360.0f);
#pragma unroll
        for (int k = 0; k < 16; ++k) addend[k] = 0.0f;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            float lam = 360.0f + (830.0f - 360.0f) * (w + wlOffset) / 4;
            uint32_t bin = min(static_cast<uint32_t>((lam - 360.0f) / (830.0f - 360.0f) * 16), 15u);
#pragma unroll
            for (int k = 0; k < 16; ++k)
                if (k == static_cast<int>(bin)) addend[k] += c[w] * recBinWidth;
        }
    } else {
        addend[0] = c[0]; addend[1] = c[1]; addend[2] = c[2];
    }
    float sums[C];
#pragma unroll
    for (int k = 0; k < C; ++k) {
        float sum = reset ? 0.0f : acc[(2 * k) * plane];
        float comp = reset ? 0.0f : acc[(2 * k + 1) * plane];
        float cIn = addend[k] - comp;
        float tmp = sum + cIn;
        comp = (tmp - sum) - cIn;
        acc[(2 * k) * plane] = tmp;
        acc[(2 * k + 1) * plane] = comp;
        sums[k] = tmp;
    }
    if constexpr (CONVERT && !MULTI)
        a.fb.output[p] = convertSums<N>(a.sc, sums, a.fp.frameIndex, a.fp.imageW, a.fp.imageH);
}

// ---- convert: post_process.cu:93-163 ----
template <int N, bool MULTI = false>
__global__ void __launch_bounds__(256) k_convert(KArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t numPaths = a.fp.imageW * a.fp.imageH;
    if (i >= numPaths) return;
    size_t plane;
    const float *acc;
    float4 *out;
    uint32_t frames, W, H;
    const DevScene *scp;
    if constexpr (MULTI) {
        const DevJob &J = a.jobs[jobOf(a, i)];
        plane = static_cast<size_t>(J.width) * J.height;
        acc = J.accum + (i - J.pixelOffset);
        out = J.output + (i - J.pixelOffset);
        frames = J.frameIndex;
        W = J.width;
        H = J.height;
        scp = &J.scene;
    } else {
        if (i >= a.fp.pixelCount) return;
        const uint32_t pix = a.fp.pixelBegin + i;
        const uint32_t px = pix % a.fp.imageW, py = pix / a.fp.imageW;
        const uint32_t p = py * a.fp.width + px;
        plane = static_cast<size_t>(a.fp.width) * a.fp.height;
        acc = a.fb.accum + p;
        out = a.fb.output + p;
        frames = a.fp.frameIndex;
        W = a.fp.imageW;
        H = a.fp.imageH;
        scp = nullptr;
    }
    const DevScene &sc = MULTI ? *scp : a.sc;
    if (frames == 0) return;
    constexpr int C = N == 4 ? 16 : 3;
    float sums[C];
#pragma unroll
    for (int k = 0; k < C; ++k) sums[k] = acc[(2 * k) * plane];
    *out = convertSums<N>(sc, sums, frames, W, H);
}

__global__ void k_frameEnd(uint32_t *counters, uint32_t maxPathLength, uint32_t wavefrontDepth) {
    // fold this frame's queue sizes into the 64-bit ray statistics
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t closest = 0, shadow = 0, wfClosest = 0, wfShadow = 0;
    for (uint32_t d = 0; d < maxPathLength && d < kMaxDepthCounters - 1; ++d) {
        closest += counters[2 * d];
        shadow += counters[2 * d + 1];
        if (d < wavefrontDepth) {
            wfClosest += counters[2 * d];
            wfShadow += counters[2 * d + 1];
        }
    }
    uint64_t *c64 = reinterpret_cast<uint64_t *>(counters + kStatClosest);
    c64[0] += closest;
    c64[1] += shadow;
    uint64_t *w64 = reinterpret_cast<uint64_t *>(counters + kStatWavefront);
    w64[0] += wfClosest;
    w64[1] += wfShadow;
}

// =============================================================================================
// LVC-BPT (lvc_bpt.cu:144-971).  Per frame: light paths (one thread each) write their vertices
// into fixed per-path slots; a prefix sum over the per-path counts compacts them into a cache
// ordered by light path and depth; eye paths (one thread per pixel) connect each vertex,
// including the lens vertex, to one uniformly chosen cache vertex, lens connections being
// splatted to the pixel they project to; the eye contributions and the splats are then added
// to the Kahan accumulation.
constexpr int kLVF4 = kLightVertexFloat4s;

struct BptArgs {
    DevScene sc;
    FrameParams fp;
    FrameBuffers fb;
    BptParams bp;
};

// a: inst, geomInst, prim, flags | b: u, v, probDensity, prevProbDensity
// c: secondPrevPartialDenomMisWeight, secondPrevProbRatioToFirst, backwardConversionFactor, dirIn.x
// d: dirIn.y, dirIn.z, 0, 0 | e: flux
constexpr uint32_t kLVDelta = 1u << 2, kLVPrevDelta = 1u << 3, kLVWlSel = 1u << 4;
struct LVert {
    uint32_t kind, inst, geomInst, prim, pathLength;
    bool delta, prevDelta, wlSelected;
    float u, v, probDensity, prevProbDensity, spdmw, sprtf, bcf;
    V3 dirIn;
    float4 flux;
};

template <int N>
__device__ __forceinline__ void storeLV(float4 *dst, const SurfPt &sp, const Spec<N> &flux, const V3 &dirIn,
                                        float bcf, float pd, float ppd, float spdmw, float sprtf, bool delta,
                                        bool prevDelta, bool wlSelected, uint32_t pathLength) {
    uint32_t kind = sp.atInfinity ? 2u : (sp.isPoint ? 1u : 0u);
    uint32_t flags = kind | (delta ? kLVDelta : 0u) | (prevDelta ? kLVPrevDelta : 0u) | (wlSelected ? kLVWlSel : 0u) |
                     (pathLength << 8);
    dst[0] = make_float4(__uint_as_float(sp.instIndex), __uint_as_float(sp.geomInstIndex),
                         __uint_as_float(sp.primIndex), __uint_as_float(flags));
    dst[1] = make_float4(sp.u, sp.v, pd, ppd);
    dst[2] = make_float4(spdmw, sprtf, bcf, dirIn.x);
    dst[3] = make_float4(dirIn.y, dirIn.z, 0.0f, 0.0f);
    dst[4] = storeSpec<N>(flux);
}

__device__ __forceinline__ LVert loadLV(const float4 *src) {
    float4 a = src[0], b = src[1], c = src[2], d = src[3];
    LVert r;
    uint32_t flags = __float_as_uint(a.w);
    r.kind = flags & 3u;
    r.delta = (flags & kLVDelta) != 0;
    r.prevDelta = (flags & kLVPrevDelta) != 0;
    r.wlSelected = (flags & kLVWlSel) != 0;
    r.pathLength = flags >> 8;
    r.inst = __float_as_uint(a.x);
    r.geomInst = __float_as_uint(a.y);
    r.prim = __float_as_uint(a.z);
    r.u = b.x; r.v = b.y; r.probDensity = b.z; r.prevProbDensity = b.w;
    r.spdmw = c.x; r.sprtf = c.y; r.bcf = c.z;
    r.dirIn = V3(c.w, d.x, d.y);
    r.flux = src[4];
    return r;
}

// decodeHitPointForTriangle (triangle.cu:55-109), ForPoint (point.cu:6-35), ForInfiniteSphere
// (infinite_sphere.cu:17-56); returns the material index
template <bool NODES = true>
__device__ __forceinline__ uint32_t decodeLightVertex(const DevScene &sc, const LVert &lv, SurfPt *sp) {
    sp->instIndex = lv.inst;
    sp->geomInstIndex = lv.geomInst;
    sp->primIndex = lv.prim;
    if (lv.kind == 2) {
        float posPhi = lv.u, theta = lv.v;
        float phi = posPhi + sc.envRotation;
        phi = phi - floorf(phi / (2 * kPiF)) * 2 * kPiF;
        V3 direction = fromPolarYUp(posPhi, theta);
        float sinPhi, cosPhi;
        sincosf(posPhi, &sinPhi, &cosPhi);
        V3 tc0 = normalize(V3(-cosPhi, 0.0f, -sinPhi));
        V3 gn = -direction;
        sp->frame.x = tc0;
        sp->frame.z = gn;
        sp->frame.y = cross(sp->frame.z, sp->frame.x);
        sp->position = direction;
        sp->gn = gn;
        sp->isPoint = false;
        sp->atInfinity = true;
        sp->u = posPhi;
        sp->v = theta;
        sp->texU = phi / (2 * kPiF);
        sp->texV = theta / kPiF;
        return sc.envMaterial;
    }
    const DevInstance &I = sc.instances[lv.inst];
    const DevGeomInst &G = sc.geomInsts[lv.geomInst];
    sp->atInfinity = false;
    if (lv.kind == 1) {
        const DevVertex &vx = sc.vertices[G.vertexBase + sc.indices[G.indexBase + lv.prim]];
        sp->position = xfmPoint(I.o2w, ld3(vx.p));
        V3 sn = normalize(xfmNormal(I.w2o, ld3(vx.n)));
        V3 tc0 = normalize(xfmVector(I.o2w, ld3(vx.t)));
        sp->frame = Frame(tc0, sn);
        sp->isPoint = true;
        sp->gn = sn;
        sp->u = lv.u;
        sp->v = lv.v;
        sp->texU = sp->texV = 0.0f;
        return G.material;
    }
    const uint32_t *ix = sc.indices + G.indexBase + 3 * lv.prim;
    const DevVertex &v0 = sc.vertices[G.vertexBase + ix[0]];
    const DevVertex &v1 = sc.vertices[G.vertexBase + ix[1]];
    const DevVertex &v2 = sc.vertices[G.vertexBase + ix[2]];
    V3 e1 = xfmVector(I.o2w, ld3(v1.p) - ld3(v0.p));
    V3 e2 = xfmVector(I.o2w, ld3(v2.p) - ld3(v0.p));
    V3 gn = cross(e1, e2);
    float area = gn.length() / 2;
    gn = gn / (2 * area);
    float b0 = lv.u, b1 = lv.v, b2 = 1.0f - lv.u - lv.v;
    V3 position = b0 * ld3(v0.p) + b1 * ld3(v1.p) + b2 * ld3(v2.p);
    V3 sn = b0 * ld3(v0.n) + b1 * ld3(v1.n) + b2 * ld3(v2.n);
    V3 tc0 = b0 * ld3(v0.t) + b1 * ld3(v1.t) + b2 * ld3(v2.t);
    sp->texU = b0 * v0.u + b1 * v1.u + b2 * v2.u;
    sp->texV = b0 * v0.v + b1 * v1.v + b2 * v2.v;
    position = xfmPoint(I.o2w, position);
    sn = normalize(xfmNormal(I.w2o, sn));
    tc0 = xfmVector(I.o2w, tc0);
    if (!sn.allFinite() || !tc0.allFinite()) {
        V3 bt;
        sn = gn;
        makeCoordinateSystem(sn, &tc0, &bt);
    }
    float dotNT = dot(sn, tc0);
    tc0 = normalize(tc0 - dotNT * sn);
    sp->position = position;
    sp->frame = Frame(tc0, sn);
    sp->isPoint = false;
    sp->gn = gn;
    sp->u = b0;
    sp->v = b1;
    if constexpr (NODES) applyGroupNodes(sc, G, sp);
    return G.material;
}

// computeRRProbability (lvc_bpt.cu:137-141)
template <int N>
__device__ __forceinline__ float rrProbability(const Spec<N> &fs, const V3 &dirLocal, float dirDensity, const V3 &gnLocal,
                                               uint32_t sel) {
    Spec<N> localThroughput = fs * absDot(dirLocal, gnLocal) / dirDensity;
    return fminf(localThroughput.importance(sel), 1.0f);
}

// EDF directional sampling and pdf (materials.cu:2052-2096, 2206-2229, 2849-2880)
__device__ __forceinline__ float edfPDF(uint32_t type, const V3 &d, float worldDiscArea) {
    switch (type) {
    case MAT_DIFFUSE_EMITTER: return d.z <= 0.0f ? 0.0f : d.z / kPiF;
    case MAT_POINT_EMITTER: return 1.0f / (4 * kPiF);
    case MAT_ENVIRONMENT_EMITTER: return d.z <= 0.0f ? 0.0f : 1.0f / worldDiscArea;
    default: return 0.0f;
    }
}

// The emitter's directional distribution used as the BSDF of a light-source vertex
template <int N>
__device__ __forceinline__ Bsdf<N> edfAsBsdf(const DevScene &sc, uint32_t type) {
    Bsdf<N> b;
    b.kind = type == MAT_DIFFUSE_EMITTER ? BSDF_EDF_DIFFUSE
           : type == MAT_POINT_EMITTER ? BSDF_EDF_POINT
           : type == MAT_ENVIRONMENT_EMITTER ? BSDF_EDF_ENV : BSDF_NULL;
    b.a0 = sc.worldDiscArea;
    b.a1 = sc.worldRadius;
    b.a2 = 0.0f;
    b.dispersive = false;
    return b;
}

// PerspectiveCameraIDF (cameras.cu:124-211)
struct LensIdf {
    float xOnLens, yOnLens, opWidth, opHeight, imgPlaneDistance, objPlaneDistance, imgPlaneArea;
    __device__ __forceinline__ bool project(const V3 &d, float *u0, float *u1) const {
        if (d.z <= 0.0f) return false;
        float objPlaneZ = objPlaneDistance / d.z;
        V3 pFocus = d * objPlaneZ + V3(xOnLens, yOnLens, 0.0f);
        *u0 = -pFocus.x / opWidth + 0.5f;
        *u1 = -pFocus.y / opHeight + 0.5f;
        return !(*u0 < 0.0f || *u0 >= 1.0f || *u1 < 0.0f || *u1 >= 1.0f);
    }
    __device__ __forceinline__ float pdf(const V3 &d) const {
        float a, b;
        if (!project(d, &a, &b)) return 0.0f;
        return sq(imgPlaneDistance) / (cube(d.z) * imgPlaneArea);
    }
};

// testVisibility (light_transport_common.h:147-177) returning the connection direction and,
// with ALPHA, the fractional visibility through alpha surfaces
template <bool ALPHA = false>
__device__ __forceinline__ bool connectVisible(const DevScene &sc, const SurfPt &shading, const SurfPt &light, V3 *dir,
                                               float *dist2, int32_t *lds, int lane, float *vis = nullptr) {
    if (light.atInfinity) {
        *dist2 = 1.0f;
        *dir = normalize(light.position);
    } else {
        V3 dv = light.position - shading.position;
        *dist2 = dv.sqLength();
        *dir = dv / sqrtf(*dist2);
    }
    bool front = dot(shading.gn, *dir) > 0;
    V3 so = offsetRayOrigin(shading.position, front ? shading.gn : -shading.gn);
    front = -dot(light.gn, *dir) > 0;
    V3 lo = offsetRayOrigin(light.position, front ? light.gn : -light.gn);
    float tmax = 3.402823466e+38f;
    if (!light.atInfinity) tmax = (lo - so).length() * 0.9999f;
    TraceHit h;
    return !traceRay<true, false, ALPHA>(sc, so, *dir, 0.0f, tmax, &h, lds, lane, nullptr, nullptr, vis);
}

template <int N>
__device__ __forceinline__ Wls<N> commonWls(const BptParams &bp, bool single) {
    Wls<N> w;
    w.offset = bp.wlOffset;
    w.sel = bp.wlSel;
    w.single = single;
    return w;
}

// ---- light subpaths (lvc_bpt.cu:144-334) ----
template <int N, bool NODES = true>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(kBptLightWavesPerEU)))
k_bptLight(BptArgs a) {
    __shared__ int32_t lds[kTraceLds];
    const DevScene &sc = a.sc;
    const BptParams &bp = a.bp;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t closestRays = 0;
    if (i < bp.numLightPaths) {
        float4 *slots = bp.slots + static_cast<size_t>(i) * bp.slotsPerPath * kLVF4;
        uint32_t written = 0;
        auto store = [&](const SurfPt &sp, const Spec<N> &flux, const V3 &dirIn, float bcf, float pd, float ppd,
                         float spdmw, float sprtf, bool delta, bool prevDelta, bool wlSel, uint32_t pl) {
            if (written < bp.slotsPerPath)
                storeLV<N>(slots + written * kLVF4, sp, flux, dirIn, bcf, pd, ppd, spdmw, sprtf, delta, prevDelta,
                           wlSel, pl);
            ++written;
        };
        Pcg32 rng{bp.lightRng[i]};
        const Wls<N> wlsC = commonWls<N>(bp, false);
        if (sc.lightInstIntegral > 0.0f) {
            float uLight = rng.uniform();
            float uP0 = rng.uniform(), uP1 = rng.uniform();
            LightSample ls;
            float lightProb;
            sampleLight(sc, uLight, uP0, uP1, &ls, &lightProb);
            const DevMaterial &lm = sc.materials[ls.material];
            Spec<N> M;
            float fe;
            evalEDF<N, NODES>(sc, lm, ls.sp, wlsC, V3(0, 0, 1), &M, &fe);
            float probDensity0 = bp.numLightPaths * lightProb * ls.areaPDF;
            Spec<N> alpha = M / probDensity0;
            store(ls.sp, alpha, V3(0, 0, 1), 0.0f, probDensity0, 1.0f, 0.0f, 1.0f, dt::isDelta(ls.posType), false,
                  false, 0);
            float uC = rng.uniform(), uD0 = rng.uniform(), uD1 = rng.uniform();
            (void)uC;
            // EDF::sample (materials.cu:2052-2063, 2206-2213, 2849-2860)
            V3 dirLocal;
            float dirPDF;
            float fe1;
            if (lm.type == MAT_POINT_EMITTER) {
                dirLocal = uniformSampleSphere(uD0, uD1);
                dirPDF = 1.0f / (4 * kPiF);
                fe1 = 1.0f / (4 * kPiF);
            } else if (lm.type == MAT_ENVIRONMENT_EMITTER) {
                float dx, dy;
                concentricSampleDisk(uD0, uD1, &dx, &dy);
                dirLocal = V3(sc.worldRadius * dx, sc.worldRadius * dy, 1);
                dirPDF = 1.0f / sc.worldDiscArea;
                fe1 = 1 / kPiF;
            } else {
                dirLocal = cosineSampleHemisphere(uD0, uD1);
                dirPDF = dirLocal.z / kPiF;
                fe1 = 1.0f / kPiF;
            }
            const uint32_t sampledType0 = dt::Reflection | dt::LowFreq;
            V3 rayOrg = offsetRayOrigin(ls.sp.position, ls.sp.gn);
            if (ls.sp.atInfinity) {
                rayOrg = ld3(sc.center) + 1.1f * sc.worldRadius * ls.sp.position + ls.sp.frame.x * dirLocal.x +
                         ls.sp.frame.y * dirLocal.y;
                dirLocal.x = 0;
                dirLocal.y = 0;
            }
            V3 rayDir = ls.sp.frame.fromLocal(dirLocal);
            float cosTerm = ls.sp.cosTerm(rayDir);
            alpha *= Spec<N>(fe1) * (cosTerm / dirPDF);
            float roDirPDF = dirPDF, roCosTerm = cosTerm, roPrevRevAreaPDF = 1.0f;
            bool roSecondPrevDelta = false, roPrevDelta = dt::isDelta(ls.posType);
            uint32_t roSampledType = sampledType0;
            float probDensity = probDensity0, prevProbDensity = 1.0f;
            float spdmw = 0.0f, sprtf = 1.0f;
            bool single = false;
            bool originIsPoint = ls.sp.isPoint, originIsInfinity = ls.sp.atInfinity;
            uint32_t pathLength = 0;
            while (true) {
                bool terminate = true;
                ++pathLength;
                ++closestRays;
                TraceHit h;
                h.inst = -1;
                h.b1 = h.b2 = 0.0f;
                bool found = traceRay<false, false, NODES>(sc, rayOrg, rayDir, 0.0f, 3.402823466e+38f, &h, lds,
                                                           threadIdx.x, nullptr, &rng);
                V3 nextOrg, nextDir;
                float woDirPDF = 0, woCosTerm = 0, woRevAreaPDF = 0;
                uint32_t woSampledType = 0;
                if (found) {
                    const Wls<N> wls = commonWls<N>(bp, single);
                    SurfPt sp;
                    float hypAreaPDF;
                    decodeHit<NODES>(sc, h.inst, h.geomInst, h.prim, h.b1, h.b2, &sp, &hypAreaPDF);
                    const DevGeomInst &G = sc.geomInsts[h.geomInst];
                    Bsdf<N> bsdf = setupBsdf<N, NODES>(sc, sc.materials[G.material], wls, &sp);
                    V3 dirInLocal = sp.frame.toLocal(-rayDir);
                    V3 gnLocal = sp.frame.toLocal(sp.gn);
                    bool secondLastValid = (!roPrevDelta && !roSecondPrevDelta) && pathLength > 2;
                    float probRatio = roPrevRevAreaPDF / prevProbDensity;
                    spdmw = sq(probRatio) * (spdmw + (secondLastValid ? 1 : 0));
                    sprtf *= probRatio;
                    prevProbDensity = probDensity;
                    float lastDist2 = (sp.position - rayOrg).sqLength();
                    if (originIsInfinity) lastDist2 = 1;
                    probDensity = roDirPDF * absDot(dirInLocal, gnLocal) / lastDist2;
                    store(sp, alpha, dirInLocal, roCosTerm / lastDist2, probDensity, prevProbDensity, spdmw, sprtf,
                          dt::isDelta(roSampledType), roPrevDelta, single, pathLength);
                    float sC = rng.uniform(), s0 = rng.uniform(), s1 = rng.uniform();
                    V3 dl;
                    float pdf;
                    uint32_t stype;
                    BsdfRev<N> rev;
                    rev.value = Spec<N>(0.0f);
                    Spec<N> fs = bsdf.sample(dirInLocal, gnLocal, wls.sel, sC, s0, s1, &dl, &pdf, &stype, true, &rev);
                    if (!(fs.isZero() || pdf == 0.0f)) {
                        if ((stype & dt::Dispersive) && !single) single = true;
                        float cosT = dot(dl, gnLocal);
                        Spec<N> localThroughput = fs * (fabsf(cosT) / pdf);
                        alpha *= localThroughput;
                        float continueProb = fminf(localThroughput.importance(bp.wlSel), 1.0f);
                        if (rng.uniform() < continueProb) {
                            alpha = alpha / continueProb;
                            terminate = false;
                            nextDir = sp.frame.fromLocal(dl);
                            nextOrg = offsetRayOrigin(sp.position, cosT > 0.0f ? sp.gn : -sp.gn);
                            woDirPDF = pdf * continueProb;
                            woCosTerm = fabsf(cosT);
                            woRevAreaPDF = rev.pdf * roCosTerm / lastDist2;
                            woRevAreaPDF *= rrProbability<N>(rev.value, dirInLocal, rev.pdf, gnLocal, bp.wlSel);
                            if (originIsPoint) woRevAreaPDF = 0;
                            originIsPoint = false;
                            originIsInfinity = false;
                            woSampledType = stype;
                        }
                    }
                }
                if (pathLength >= a.fp.maxPathLength) terminate = true;
                if (terminate) break;
                rayOrg = nextOrg;
                rayDir = nextDir;
                roDirPDF = woDirPDF;
                roCosTerm = woCosTerm;
                roPrevRevAreaPDF = woRevAreaPDF;
                roSecondPrevDelta = roPrevDelta;
                roPrevDelta = dt::isDelta(roSampledType);
                roSampledType = woSampledType;
            }
        }
        bp.slotCounts[i] = min(written, bp.slotsPerPath);
        bp.lightRng[i] = rng.s;
    }
    uint32_t c = waveSum(closestRays);
    if (__lane_id() == 0 && c) atomicAdd(reinterpret_cast<unsigned long long *>(a.fb.counters + kStatClosest), c);
}

// ---- cache compaction: exclusive prefix sum of the per-path counts, then a gather ----
constexpr uint32_t kScanBlock = 1024;
__global__ void __launch_bounds__(kScanBlock) k_bptScanBlocks(BptParams bp) {
    __shared__ uint32_t waveTotals[kScanBlock / 64];
    const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    uint32_t v = i < bp.numLightPaths ? bp.slotCounts[i] : 0u;
    // inclusive scan inside the wave
    uint32_t x = v;
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= static_cast<uint32_t>(off)) x += y;
    }
    const uint32_t wave = threadIdx.x / 64;
    if (lane == 63) waveTotals[wave] = x;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t t = threadIdx.x < kScanBlock / 64 ? waveTotals[threadIdx.x] : 0u;
        uint32_t s = t;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t y = __shfl_up(s, off);
            if (lane >= static_cast<uint32_t>(off)) s += y;
        }
        if (threadIdx.x < kScanBlock / 64) waveTotals[threadIdx.x] = s - t;   // exclusive wave offsets
        if (threadIdx.x == kScanBlock / 64 - 1) bp.blockSums[blockIdx.x] = s;
    }
    __syncthreads();
    if (i < bp.numLightPaths) bp.slotOffsets[i] = waveTotals[wave] + x - v;
}

// one block: exclusive scan of the block totals (at most 1024 blocks), total vertex count
__global__ void __launch_bounds__(kScanBlock) k_bptScanTop(BptParams bp, uint32_t numBlocks) {
    __shared__ uint32_t waveTotals[kScanBlock / 64];
    const uint32_t lane = __lane_id();
    uint32_t v = threadIdx.x < numBlocks ? bp.blockSums[threadIdx.x] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= static_cast<uint32_t>(off)) x += y;
    }
    const uint32_t wave = threadIdx.x / 64;
    if (lane == 63) waveTotals[wave] = x;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t t = threadIdx.x < kScanBlock / 64 ? waveTotals[threadIdx.x] : 0u;
        uint32_t s = t;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t y = __shfl_up(s, off);
            if (lane >= static_cast<uint32_t>(off)) s += y;
        }
        if (threadIdx.x < kScanBlock / 64) waveTotals[threadIdx.x] = s - t;
        if (threadIdx.x == kScanBlock / 64 - 1) *bp.numVertices = s;
    }
    __syncthreads();
    if (threadIdx.x < numBlocks) bp.blockSums[threadIdx.x] = waveTotals[wave] + x - v;
}

// One thread per float4 of a light path's slots: thread (path i, element k) copies element k if
// the path wrote that many vertices, so consecutive threads read and write consecutive float4s.
__global__ void __launch_bounds__(256) k_bptCompact(BptParams bp) {
    const uint32_t perPath = bp.slotsPerPath * kLVF4;
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t i = g / perPath;
    const uint32_t k = static_cast<uint32_t>(g - i * perPath);
    if (i >= bp.numLightPaths) return;
    if (k >= bp.slotCounts[i] * kLVF4) return;
    const uint32_t base = bp.blockSums[i / kScanBlock] + bp.slotOffsets[i];
    bp.cache[static_cast<size_t>(base) * kLVF4 + k] = bp.slots[i * perPath + k];
}

// ---- eye subpaths (lvc_bpt.cu:338-971) ----
template <int N, bool NODES = true, bool GUIDES = false>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(kBptEyeWavesPerEU)))
k_bptEye(BptArgs a) {
    __shared__ int32_t lds[kTraceLds];
    const DevScene &sc = a.sc;
    const BptParams &bp = a.bp;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t numPaths = a.fp.imageW * a.fp.imageH;
    uint32_t closestRays = 0, shadowRays = 0;
    if (i < numPaths) {
        constexpr int C = N == 4 ? 16 : 3;
        const uint32_t px = i % a.fp.imageW, py = i / a.fp.imageW;
        const uint32_t path = py * a.fp.width + px;
        const uint32_t W = a.fp.imageW, H = a.fp.imageH;
        const uint32_t numLightVertices = *bp.numVertices;
        const float vertexProb = 1.0f / numLightVertices;
        const size_t plane = static_cast<size_t>(a.fp.width) * a.fp.height;
        Pcg32 rng{a.fb.ps.rng[path]};
        const Wls<N> wlsC = commonWls<N>(bp, false);
        float pX = px + rng.uniform();
        float pY = py + rng.uniform();
        const DevCamera &cam = sc.camera;
        float lx, ly;
        float ul0 = rng.uniform(), ul1 = rng.uniform();
        concentricSampleDisk(ul0, ul1, &lx, &ly);
        V3 c0(cam.rot[0], cam.rot[1], cam.rot[2]), c1(cam.rot[3], cam.rot[4], cam.rot[5]),
            c2(cam.rot[6], cam.rot[7], cam.rot[8]);
        SurfPt spE0;
        spE0.frame.z = normalize(c2);
        spE0.frame.x = normalize(c0);
        spE0.frame.y = cross(spE0.frame.z, spE0.frame.x);
        V3 orgLocal(cam.lensRadius * lx, cam.lensRadius * ly, 0.0f);
        spE0.position = (c0 * orgLocal.x + c1 * orgLocal.y + c2 * orgLocal.z) + ld3(cam.position);
        spE0.gn = spE0.frame.z;
        spE0.isPoint = false;
        spE0.atInfinity = false;
        float lensAreaPDF = cam.lensRadius > 0.0f ? 1.0f / (kPiF * cam.lensRadius * cam.lensRadius) : 1.0f;
        const uint32_t lensPosType = cam.lensRadius > 0.0f ? dt::LowFreq : dt::Delta0D;
        LensIdf idf{cam.lensRadius * lx, cam.lensRadius * ly, cam.opWidth, cam.opHeight, cam.imgPlaneDistance,
                    cam.objPlaneDistance, cam.imgPlaneArea};
        Spec<N> alpha = Spec<N>(cam.sensitivity) / (lensAreaPDF * bp.wlProb * W * H);

        // lens vertex connection (t = 1), splatted
        if (numLightVertices > 0) {
            uint32_t idx = min(static_cast<uint32_t>(numLightVertices * rng.uniform()), numLightVertices - 1);
            LVert vtx = loadLV(bp.cache + static_cast<size_t>(idx) * kLVF4);
            SurfPt spL;
            uint32_t matL = decodeLightVertex<NODES>(sc, vtx, &spL);
            V3 conRayDir;
            float sqDist;
            ++shadowRays;
            float vis = 1.0f;
            if (connectVisible<NODES>(sc, spE0, spL, &conRayDir, &sqDist, lds, threadIdx.x, &vis)) {
                float recSq = 1.0f / sqDist;
                const DevMaterial &mL = sc.materials[matL];
                Bsdf<N> bL = vtx.pathLength == 0 ? edfAsBsdf<N>(sc, mL.type) : setupBsdf<N, NODES>(sc, mL, wlsC, &spL);
                if (bL.hasNonDelta()) {
                    V3 gnLocalL = spL.frame.toLocal(spL.gn);
                    V3 conLocalL = spL.frame.toLocal(-conRayDir);
                    V3 conLocalE = spE0.frame.toLocal(conRayDir);
                    float cosL = spL.cosTerm(conRayDir);
                    float cosE = spE0.cosTerm(conRayDir);
                    float G = cosL * cosE * recSq;
                    float scalarConTerm = G * vis / vertexProb;
                    if (vtx.wlSelected) scalarConTerm *= N;
                    Spec<N> backwardFsL;
                    Spec<N> forwardFsL = bL.evaluate(vtx.dirIn, gnLocalL, conLocalL, wlsC.sel, true, &backwardFsL);
                    float backwardDirDensityL;
                    bL.evaluatePDF(vtx.dirIn, conLocalL, wlsC.sel, &backwardDirDensityL);
                    backwardDirDensityL *= rrProbability<N>(backwardFsL, vtx.dirIn, backwardDirDensityL, gnLocalL, bp.wlSel);
                    float backwardAreaDensityL = backwardDirDensityL * vtx.bcf;
                    if (vtx.prevDelta) backwardAreaDensityL = 0;
                    float u0 = 0.0f, u1 = 0.0f;
                    bool onScreen = idf.project(conLocalE, &u0, &u1);
                    Spec<N> forwardFsE(onScreen ? 1.0f : 0.0f);
                    float forwardDirDensityE = idf.pdf(conLocalE);
                    float forwardAreaDensityE = forwardDirDensityE * cosL * recSq;
                    if (spL.isPoint) forwardAreaDensityE = 0;
                    float partialL;
                    {
                        bool lastValid = (!vtx.delta && !vtx.prevDelta) && vtx.pathLength > 1;
                        float r = backwardAreaDensityL / vtx.prevProbDensity;
                        partialL = sq(r) * (vtx.spdmw + (lastValid ? 1 : 0));
                        float probRatioToFirst = vtx.sprtf;
                        if (vtx.pathLength > 0) probRatioToFirst *= r;
                        bool curValid = !vtx.delta && vtx.pathLength > 0;
                        float r2 = forwardAreaDensityE / vtx.probDensity;
                        partialL = sq(r2) * (partialL + (curValid ? 1 : 0));
                        probRatioToFirst *= r2;
                        partialL += sq(probRatioToFirst / vertexProb);
                    }
                    Spec<N> conTerm = forwardFsL * scalarConTerm * forwardFsE;
                    Spec<N> unweighted = loadSpec<N>(vtx.flux) * conTerm * alpha;
                    Spec<N> contribution = (1.0f / (1.0f + partialL)) * unweighted;
                    if (contribution.allFinite() && onScreen) {
                        uint32_t ipx = static_cast<uint32_t>(u0 * W), ipy = static_cast<uint32_t>(u1 * H);
                        if (ipx < W && ipy < H) {
                            float *dst = bp.splat + static_cast<size_t>(ipy) * a.fp.width + ipx;
                            if constexpr (N == 4) {
                                const float recBinWidth = 16 / (830.0f - 360.0f);
#pragma unroll
                                for (int w = 0; w < 4; ++w) {
                                    float lam = 360.0f + (830.0f - 360.0f) * (w + wlsC.offset) / 4;
                                    uint32_t bin = min(static_cast<uint32_t>((lam - 360.0f) / (830.0f - 360.0f) * 16), 15u);
                                    atomicAdd(dst + bin * plane, contribution[w] * recBinWidth);
                                }
                            } else {
#pragma unroll
                                for (int c = 0; c < 3; ++c) atomicAdd(dst + c * plane, contribution[c]);
                            }
                        }
                    }
                }
            }
        }

        // first eye ray (cameras.cu:145-157)
        float uD0 = pX / W, uD1 = pY / H;
        V3 pFocus(cam.opWidth * (0.5f - uD0), cam.opHeight * (0.5f - uD1), cam.objPlaneDistance);
        V3 dirLocal = normalize(pFocus - V3(idf.xOnLens, idf.yOnLens, 0.0f));
        float dirPDF = sq(cam.imgPlaneDistance) / (cube(dirLocal.z) * cam.imgPlaneArea);
        V3 rayOrg = spE0.position;
        V3 rayDir = spE0.frame.fromLocal(dirLocal);
        float cosTerm0 = spE0.cosTerm(rayDir);
        alpha *= Spec<N>(1.0f) * (cosTerm0 / dirPDF);
        float roDirPDF = dirPDF, roCosTerm = cosTerm0, roPrevRevAreaPDF = 1.0f;
        bool roSecondPrevDelta = false, roPrevDelta = dt::isDelta(lensPosType);
        uint32_t roSampledType = dt::Reflection | dt::LowFreq;
        float probDensity = lensAreaPDF, prevProbDensity = 1.0f, spdmw = 0.0f;
        bool single = false;
        uint32_t pathLength = 0;
        Spec<N> contribution(0.0f);
        while (true) {
            bool terminate = true;
            ++pathLength;
            ++closestRays;
            TraceHit h;
            h.inst = -1;
            h.b1 = h.b2 = 0.0f;
            bool found = traceRay<false, false, NODES>(sc, rayOrg, rayDir, 0.0f, 3.402823466e+38f, &h, lds,
                                                       threadIdx.x, nullptr, &rng);
            const Wls<N> wls = commonWls<N>(bp, single);
            if (GUIDES && pathLength == 1)
                addGuides<N, NODES>(sc, a.fp, a.fb, path,
                                    found ? make_int4(h.inst, h.geomInst, h.prim, 0) : make_int4(-1, -1, -1, 0),
                                    make_float2(h.b1, h.b2), wls, bp.wlProb);
            V3 nextOrg, nextDir;
            float woDirPDF = 0, woCosTerm = 0, woRevAreaPDF = 0;
            uint32_t woSampledType = 0;
            if (!found) {
                // miss program (lvc_bpt.cu:841-971)
                if (sc.hasEnv && sc.envImportance > 0.0f) {
                    SurfPt spE;
                    float hypAreaPDF;
                    envPointFromDir(sc, rayDir, &spE, &hypAreaPDF);
                    V3 dirOutLocalE = spE.frame.toLocal(-rayDir);
                    bool secondLastValid = (!roPrevDelta && !roSecondPrevDelta) && pathLength > 2;
                    spdmw = sq(roPrevRevAreaPDF / prevProbDensity) * (spdmw + (secondLastValid ? 1 : 0));
                    prevProbDensity = probDensity;
                    const float lastDist2 = 1.0f;
                    probDensity = roDirPDF / lastDist2;
                    Spec<N> M;
                    float fe;
                    evalEDF<N, NODES>(sc, sc.materials[sc.envMaterial], spE, wls, dirOutLocalE, &M, &fe);
                    if (!M.isZero()) {
                        Spec<N> unweighted = alpha * (M * fe);
                        if (single) unweighted = unweighted * static_cast<float>(N);
                        float instProb = 1.0f / sc.lightInstIntegral;
                        float forwardAreaDensityL = bp.numLightPaths * instProb * 1.0f * hypAreaPDF;
                        float backwardDirDensityE = edfPDF(MAT_ENVIRONMENT_EMITTER, dirOutLocalE, sc.worldDiscArea);
                        float backwardAreaDensityE = backwardDirDensityE * roCosTerm / lastDist2;
                        float partialE;
                        {
                            bool lastValid = (!dt::isDelta(roSampledType) && !roPrevDelta) && pathLength > 1;
                            float r = backwardAreaDensityE / prevProbDensity;
                            partialE = sq(r) * (spdmw + (lastValid ? 1 : 0));
                            bool curValid = !dt::isDelta(roSampledType);
                            float r2 = forwardAreaDensityL / probDensity;
                            partialE = sq(r2) * (partialE + (curValid ? 1 : 0));
                            partialE *= sq(vertexProb);
                        }
                        Spec<N> c = (1.0f / (1.0f + partialE)) * unweighted;
                        if (c.allFinite()) contribution += c;
                    }
                }
                break;
            }
            // closest-hit program for eye paths (lvc_bpt.cu:570-837)
            SurfPt spE;
            float hypAreaPDF;
            decodeHit<NODES>(sc, h.inst, h.geomInst, h.prim, h.b1, h.b2, &spE, &hypAreaPDF);
            const DevGeomInst &G = sc.geomInsts[h.geomInst];
            const DevMaterial &matE = sc.materials[G.material];
            Bsdf<N> bsdfE = setupBsdf<N, NODES>(sc, matE, wls, &spE);
            V3 dirOutLocalE = spE.frame.toLocal(-rayDir);
            V3 gnLocalE = spE.frame.toLocal(spE.gn);
            bool secondLastValidE = (!roPrevDelta && !roSecondPrevDelta) && pathLength > 2;
            spdmw = sq(roPrevRevAreaPDF / prevProbDensity) * (spdmw + (secondLastValidE ? 1 : 0));
            prevProbDensity = probDensity;
            float lastDist2 = (rayOrg - spE.position).sqLength();
            probDensity = roDirPDF * absDot(dirOutLocalE, gnLocalE) / lastDist2;

            if (matE.isEmitter) {
                Spec<N> M;
                float fe;
                evalEDF<N, NODES>(sc, matE, spE, wls, dirOutLocalE, &M, &fe);
                if (!M.isZero()) {
                    Spec<N> unweighted = alpha * (M * fe);
                    if (single) unweighted = unweighted * static_cast<float>(N);
                    const DevInstance &I = sc.instances[h.inst];
                    float instProb = I.lightDistIntegral / sc.lightInstIntegral;
                    float geomInstProb = G.importance / I.lightDistIntegral;
                    float forwardAreaDensityL = bp.numLightPaths * instProb * geomInstProb * hypAreaPDF;
                    float backwardDirDensityE = edfPDF(matE.type, dirOutLocalE, sc.worldDiscArea);
                    float backwardAreaDensityE = backwardDirDensityE * roCosTerm / lastDist2;
                    float partialE;
                    {
                        bool lastValid = (!dt::isDelta(roSampledType) && !roPrevDelta) && pathLength > 1;
                        float r = backwardAreaDensityE / prevProbDensity;
                        partialE = sq(r) * (spdmw + (lastValid ? 1 : 0));
                        bool curValid = !dt::isDelta(roSampledType);
                        float r2 = forwardAreaDensityL / probDensity;
                        partialE = sq(r2) * (partialE + (curValid ? 1 : 0));
                        partialE *= sq(vertexProb);
                    }
                    Spec<N> c = (1.0f / (1.0f + partialE)) * unweighted;
                    if (c.allFinite()) contribution += c;
                }
            }

            if (bsdfE.hasNonDelta() && numLightVertices > 0) {
                uint32_t idx = min(static_cast<uint32_t>(numLightVertices * rng.uniform()), numLightVertices - 1);
                LVert vtx = loadLV(bp.cache + static_cast<size_t>(idx) * kLVF4);
                SurfPt spL;
                uint32_t matL = decodeLightVertex<NODES>(sc, vtx, &spL);
                V3 conRayDir;
                float sqDist;
                ++shadowRays;
                float vis = 1.0f;
                if (connectVisible<NODES>(sc, spE, spL, &conRayDir, &sqDist, lds, threadIdx.x, &vis)) {
                    float recSq = 1.0f / sqDist;
                    const DevMaterial &mL = sc.materials[matL];
                    Bsdf<N> bL = vtx.pathLength == 0 ? edfAsBsdf<N>(sc, mL.type) : setupBsdf<N, NODES>(sc, mL, wls, &spL);
                    if (bL.hasNonDelta()) {
                        V3 gnLocalL = spL.frame.toLocal(spL.gn);
                        V3 conLocalL = spL.frame.toLocal(-conRayDir);
                        V3 conLocalE = spE.frame.toLocal(conRayDir);
                        float cosL = spL.cosTerm(conRayDir);
                        float cosE = spE.cosTerm(conRayDir);
                        float Gc = cosL * cosE * recSq;
                        float scalarConTerm = Gc * vis / vertexProb;
                        if (vtx.wlSelected || single) scalarConTerm *= N;
                        Spec<N> backwardFsL;
                        Spec<N> forwardFsL = bL.evaluate(vtx.dirIn, gnLocalL, conLocalL, wls.sel, true, &backwardFsL);
                        float backwardDirDensityL;
                        float forwardDirDensityL = bL.evaluatePDF(vtx.dirIn, conLocalL, wls.sel, &backwardDirDensityL);
                        if (vtx.pathLength > 0)
                            forwardDirDensityL *= rrProbability<N>(forwardFsL, conLocalL, forwardDirDensityL, gnLocalL, bp.wlSel);
                        backwardDirDensityL *= rrProbability<N>(backwardFsL, vtx.dirIn, backwardDirDensityL, gnLocalL, bp.wlSel);
                        float forwardAreaDensityL = forwardDirDensityL * cosE * recSq;
                        float backwardAreaDensityL = backwardDirDensityL * vtx.bcf;
                        if (vtx.prevDelta) backwardAreaDensityL = 0;
                        Spec<N> backwardFsE;
                        Spec<N> forwardFsE = bsdfE.evaluate(dirOutLocalE, gnLocalE, conLocalE, wls.sel, false, &backwardFsE);
                        float backwardDirDensityE;
                        float forwardDirDensityE = bsdfE.evaluatePDF(dirOutLocalE, conLocalE, wls.sel, &backwardDirDensityE);
                        forwardDirDensityE *= rrProbability<N>(forwardFsE, conLocalE, forwardDirDensityE, gnLocalE, bp.wlSel);
                        backwardDirDensityE *= rrProbability<N>(backwardFsE, dirOutLocalE, backwardDirDensityE, gnLocalE, bp.wlSel);
                        float forwardAreaDensityE = forwardDirDensityE * cosL * recSq;
                        if (spL.isPoint) forwardAreaDensityE = 0;
                        float backwardAreaDensityE = backwardDirDensityE * roCosTerm / lastDist2;
                        float partialL;
                        {
                            bool lastValid = (!vtx.delta && !vtx.prevDelta) && vtx.pathLength > 1;
                            float r = backwardAreaDensityL / vtx.prevProbDensity;
                            partialL = sq(r) * (vtx.spdmw + (lastValid ? 1 : 0));
                            float probRatioToFirst = vtx.sprtf;
                            if (vtx.pathLength > 0) probRatioToFirst *= r;
                            bool curValid = !vtx.delta && vtx.pathLength > 0;
                            float r2 = forwardAreaDensityE / vtx.probDensity;
                            partialL = sq(r2) * (partialL + (curValid ? 1 : 0));
                            probRatioToFirst *= r2;
                            partialL += sq(probRatioToFirst / vertexProb);
                        }
                        float partialE;
                        {
                            bool lastValid = !dt::isDelta(roSampledType) && !roPrevDelta;
                            float r = backwardAreaDensityE / prevProbDensity;
                            partialE = sq(r) * (spdmw + (lastValid ? 1 : 0));
                            bool curValid = !dt::isDelta(roSampledType);
                            float r2 = forwardAreaDensityL / probDensity;
                            partialE = sq(r2) * (partialE + (curValid ? 1 : 0));
                        }
                        Spec<N> conTerm = forwardFsL * scalarConTerm * forwardFsE;
                        Spec<N> unweighted = loadSpec<N>(vtx.flux) * conTerm * alpha;
                        Spec<N> c = (1.0f / (1.0f + partialL + partialE)) * unweighted;
                        if (c.allFinite()) contribution += c;
                    }
                }
            }

            float sC = rng.uniform(), s0 = rng.uniform(), s1 = rng.uniform();
            V3 dl;
            float pdf;
            uint32_t stype;
            BsdfRev<N> rev;
            rev.value = Spec<N>(0.0f);
            Spec<N> fs = bsdfE.sample(dirOutLocalE, gnLocalE, wls.sel, sC, s0, s1, &dl, &pdf, &stype, false, &rev);
            if (!(fs.isZero() || pdf == 0.0f)) {
                if ((stype & dt::Dispersive) && !single) single = true;
                float cosT = dot(dl, gnLocalE);
                Spec<N> localThroughput = fs * (fabsf(cosT) / pdf);
                alpha *= localThroughput;
                float continueProb = fminf(localThroughput.importance(bp.wlSel), 1.0f);
                if (rng.uniform() < continueProb) {
                    alpha = alpha / continueProb;
                    terminate = false;
                    nextDir = spE.frame.fromLocal(dl);
                    nextOrg = offsetRayOrigin(spE.position, cosT > 0.0f ? spE.gn : -spE.gn);
                    woDirPDF = pdf * continueProb;
                    woCosTerm = fabsf(cosT);
                    woRevAreaPDF = rev.pdf * roCosTerm / lastDist2;
                    woRevAreaPDF *= rrProbability<N>(rev.value, dirOutLocalE, rev.pdf, gnLocalE, bp.wlSel);
                    woSampledType = stype;
                }
            }
            if (pathLength >= a.fp.maxPathLength) terminate = true;
            if (terminate) break;
            rayOrg = nextOrg;
            rayDir = nextDir;
            roDirPDF = woDirPDF;
            roCosTerm = woCosTerm;
            roPrevRevAreaPDF = woRevAreaPDF;
            roSecondPrevDelta = roPrevDelta;
            roPrevDelta = dt::isDelta(roSampledType);
            roSampledType = woSampledType;
        }
        a.fb.ps.rng[path] = rng.s;
        a.fb.ps.contrib[path] = storeSpec<N>(contribution);
    }
    uint32_t c = waveSum(closestRays), sh = waveSum(shadowRays);
    if (__lane_id() == 0) {
        if (c) atomicAdd(reinterpret_cast<unsigned long long *>(a.fb.counters + kStatClosest), c);
        if (sh) atomicAdd(reinterpret_cast<unsigned long long *>(a.fb.counters + kStatShadow), sh);
    }
}

// ---- light tracing (light_tracing.cu:49-180 ray generation, 182-263 closest hit) ----
// One thread per light path with its own wavelengths.  The light vertex and every later
// vertex with a non-delta BSDF are connected to a sampled lens point; the connection is
// splatted to the pixel it projects to.  The path ends at a miss, a zero BSDF sample, Russian
// roulette on the throughput's importance relative to its initial value, or maxPathLength.
template <int N>
__device__ __forceinline__ void splatContribution(const BptArgs &a, const Wls<N> &wls, const Spec<N> &c, float u0,
                                                  float u1) {
    // atomicAddToBuffer (light_tracing.cu:32-46): values off the image or not finite are dropped
    const uint32_t W = a.fp.imageW, H = a.fp.imageH;
    const uint32_t ipx = static_cast<uint32_t>(u0 * W), ipy = static_cast<uint32_t>(u1 * H);
    if (ipx >= W || ipy >= H || !c.allFinite()) return;
    const size_t plane = static_cast<size_t>(a.fp.width) * a.fp.height;
    float *dst = a.bp.splat + static_cast<size_t>(ipy) * a.fp.width + ipx;
    if constexpr (N == 4) {
        const float recBinWidth = 16 / (830.0f - 360.0f);
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const float lam = wls.lambda(w);
            const uint32_t bin = min(static_cast<uint32_t>((lam - 360.0f) / (830.0f - 360.0f) * 16), 15u);
            atomicAdd(dst + bin * plane, c[w] * recBinWidth);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) atomicAdd(dst + k * plane, c[k]);
    }
}

// a sampled point on the perspective camera's lens (cameras.cu:43-76) and its importance function
__device__ __forceinline__ void sampleLens(const DevCamera &cam, float ul0, float ul1, SurfPt *spE, LensIdf *idf,
                                           float *lensAreaPDF) {
    float lx, ly;
    concentricSampleDisk(ul0, ul1, &lx, &ly);
    V3 c0(cam.rot[0], cam.rot[1], cam.rot[2]), c1(cam.rot[3], cam.rot[4], cam.rot[5]), c2(cam.rot[6], cam.rot[7], cam.rot[8]);
    spE->frame.z = normalize(c2);
    spE->frame.x = normalize(c0);
    spE->frame.y = cross(spE->frame.z, spE->frame.x);
    V3 orgLocal(cam.lensRadius * lx, cam.lensRadius * ly, 0.0f);
    spE->position = (c0 * orgLocal.x + c1 * orgLocal.y + c2 * orgLocal.z) + ld3(cam.position);
    spE->gn = spE->frame.z;
    spE->isPoint = false;
    spE->atInfinity = false;
    *lensAreaPDF = cam.lensRadius > 0.0f ? 1.0f / (kPiF * cam.lensRadius * cam.lensRadius) : 1.0f;
    *idf = LensIdf{cam.lensRadius * lx, cam.lensRadius * ly, cam.opWidth, cam.opHeight, cam.imgPlaneDistance,
                   cam.objPlaneDistance, cam.imgPlaneArea};
}

template <int N, bool NODES = true>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(kLightTraceWavesPerEU)))
k_lightTrace(BptArgs a) {
    __shared__ int32_t lds[kTraceLds];
    const DevScene &sc = a.sc;
    const BptParams &bp = a.bp;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t closestRays = 0, shadowRays = 0;
    if (i < bp.numLightPaths) {
        Pcg32 rng{bp.lightRng[i]};
        // the path's own wavelengths, offset drawn first (createWithEqualOffsets)
        Wls<N> wls;
        wls.offset = rng.uniform();
        const float wlSelect = rng.uniform();
        wls.sel = min(static_cast<uint32_t>(static_cast<uint16_t>(N * wlSelect)), static_cast<uint32_t>(N - 1));
        wls.single = false;
        const float selectWLPDF = N == 4 ? N / (830.0f - 360.0f) : 1.0f;
        const DevCamera &cam = sc.camera;
        if (sc.lightInstIntegral > 0.0f) {
            const float uLight = rng.uniform();
            const float uP0 = rng.uniform(), uP1 = rng.uniform();
            LightSample ls;
            float lightProb;
            sampleLight(sc, uLight, uP0, uP1, &ls, &lightProb);
            const DevMaterial &lm = sc.materials[ls.material];
            Spec<N> Le0;
            float fe;
            evalEDF<N, NODES>(sc, lm, ls.sp, wls, V3(0, 0, 1), &Le0, &fe);
            Spec<N> alpha = Le0 / (bp.numLightPaths * lightProb * ls.areaPDF * selectWLPDF);
            {
                // lens connection of the light vertex
                const float ul0 = rng.uniform(), ul1 = rng.uniform();
                SurfPt spE;
                LensIdf idf;
                float lensPDF;
                sampleLens(cam, ul0, ul1, &spE, &idf, &lensPDF);
                V3 dir;
                float dist2;
                if (cam.sensitivity != 0.0f) {
                    ++shadowRays;
                    float vis = 1.0f;
                    if (connectVisible<NODES>(sc, spE, ls.sp, &dir, &dist2, lds, threadIdx.x, &vis)) {
                        const V3 dirLens = spE.frame.toLocal(dir);
                        const V3 dirSn = ls.sp.frame.toLocal(-dir);
                        float u0 = 0.0f, u1 = 0.0f;
                        const bool onScreen = idf.project(dirLens, &u0, &u1);
                        Spec<N> M;
                        float fe1;
                        evalEDF<N, NODES>(sc, lm, ls.sp, wls, dirSn, &M, &fe1);
                        const float G = vis * ls.sp.cosTerm(-dir) * spE.cosTerm(dir) / dist2;
                        const float scalarCoeff = G / lensPDF;
                        const Spec<N> contribution = alpha * Spec<N>(cam.sensitivity) * Spec<N>(fe1) * scalarCoeff;
                        if (onScreen) splatContribution<N>(a, wls, contribution, u0, u1);
                    }
                }
            }
            // EDF::sample (materials.cu:2052-2063, 2206-2213, 2849-2860)
            const float uC = rng.uniform(), uD0 = rng.uniform(), uD1 = rng.uniform();
            (void)uC;
            V3 dirLocal;
            float dirPDF, fe1;
            if (lm.type == MAT_POINT_EMITTER) {
                dirLocal = uniformSampleSphere(uD0, uD1);
                dirPDF = 1.0f / (4 * kPiF);
                fe1 = 1.0f / (4 * kPiF);
            } else if (lm.type == MAT_ENVIRONMENT_EMITTER) {
                float dx, dy;
                concentricSampleDisk(uD0, uD1, &dx, &dy);
                dirLocal = V3(sc.worldRadius * dx, sc.worldRadius * dy, 1);
                dirPDF = 1.0f / sc.worldDiscArea;
                fe1 = 1 / kPiF;
            } else {
                dirLocal = cosineSampleHemisphere(uD0, uD1);
                dirPDF = dirLocal.z / kPiF;
                fe1 = 1.0f / kPiF;
            }
            V3 rayOrg = offsetRayOrigin(ls.sp.position, ls.sp.gn);
            if (ls.sp.atInfinity) {
                rayOrg = ld3(sc.center) + 1.1f * sc.worldRadius * ls.sp.position + ls.sp.frame.x * dirLocal.x +
                         ls.sp.frame.y * dirLocal.y;
                dirLocal.x = 0;
                dirLocal.y = 0;
            }
            V3 rayDir = ls.sp.frame.fromLocal(dirLocal);
            alpha *= Spec<N>(fe1) * (ls.sp.cosTerm(rayDir) / dirPDF);
            const float initImportance = alpha.importance(wls.sel);
            // two generators from here on, as light_tracing.cu:138-176: the hit program draws
            // from a copy (rng, the state stored at the end), Russian roulette from the
            // ray-generation program's own (rrRng)
            Pcg32 rrRng = rng;
            uint32_t pathLength = 0;
            while (true) {
                ++pathLength;
                if (pathLength >= a.fp.maxPathLength) break;
                ++closestRays;
                TraceHit h;
                h.inst = -1;
                h.b1 = h.b2 = 0.0f;
                if (!traceRay<false, false, NODES>(sc, rayOrg, rayDir, 0.0f, 3.402823466e+38f, &h, lds, threadIdx.x,
                                                   nullptr, &rng))
                    break;
                SurfPt sp;
                float hypAreaPDF;
                decodeHit<NODES>(sc, h.inst, h.geomInst, h.prim, h.b1, h.b2, &sp, &hypAreaPDF);
                const DevGeomInst &G = sc.geomInsts[h.geomInst];
                const Bsdf<N> bsdf = setupBsdf<N, NODES>(sc, sc.materials[G.material], wls, &sp);
                const V3 dirInLocal = sp.frame.toLocal(-rayDir);
                const V3 gnLocal = sp.frame.toLocal(sp.gn);
                if (bsdf.hasNonDelta()) {
                    const float ul0 = rng.uniform(), ul1 = rng.uniform();
                    SurfPt spE;
                    LensIdf idf;
                    float lensPDF;
                    sampleLens(cam, ul0, ul1, &spE, &idf, &lensPDF);
                    V3 dir;
                    float dist2;
                    if (cam.sensitivity != 0.0f) {
                        ++shadowRays;
                        float vis = 1.0f;
                        if (connectVisible<NODES>(sc, spE, sp, &dir, &dist2, lds, threadIdx.x, &vis)) {
                            const V3 dirLens = spE.frame.toLocal(dir);
                            const V3 dirSn = sp.frame.toLocal(-dir);
                            float u0 = 0.0f, u1 = 0.0f;
                            const bool onScreen = idf.project(dirLens, &u0, &u1);
                            const Spec<N> fs = bsdf.evaluate(dirInLocal, gnLocal, dirSn, wls.sel, true);
                            const float Gc = vis * absDot(dirSn, gnLocal) * spE.cosTerm(dir) / dist2;
                            const float scalarCoeff = Gc / lensPDF;
                            const Spec<N> contribution = alpha * Spec<N>(cam.sensitivity) * fs * scalarCoeff;
                            if (onScreen) splatContribution<N>(a, wls, contribution, u0, u1);
                        }
                    }
                }
                const float sC = rng.uniform(), s0 = rng.uniform(), s1 = rng.uniform();
                V3 dl;
                float pdf;
                uint32_t stype;
                const Spec<N> fs = bsdf.sample(dirInLocal, gnLocal, wls.sel, sC, s0, s1, &dl, &pdf, &stype, true);
                if (fs.isZero() || pdf == 0.0f) break;
                if ((stype & dt::Dispersive) && !wls.single) {
                    pdf /= N;
                    wls.single = true;
                }
                const float cosFactor = dot(dl, gnLocal);
                alpha *= fs * (fabsf(cosFactor) / pdf);
                rayOrg = offsetRayOrigin(sp.position, cosFactor > 0.0f ? sp.gn : -sp.gn);
                rayDir = sp.frame.fromLocal(dl);
                // Russian roulette (light_tracing.cu:165-169), from the ray-generation generator
                const float continueProb = fminf(alpha.importance(wls.sel) / initImportance, 1.0f);
                if (rrRng.uniform() >= continueProb) break;
                alpha = alpha / continueProb;
            }
        }
        bp.lightRng[i] = rng.s;
    }
    const uint32_t c = waveSum(closestRays), sh = waveSum(shadowRays);
    if (__lane_id() == 0) {
        if (c) atomicAdd(reinterpret_cast<unsigned long long *>(a.fb.counters + kStatClosest), c);
        if (sh) atomicAdd(reinterpret_cast<unsigned long long *>(a.fb.counters + kStatShadow), sh);
    }
}

// Eye contribution, then the frame's splats, into the Kahan accumulation (lvc_bpt.cu:560-565,
// post_process.cu:18-31); the splat planes are cleared for the next frame.
template <int N, bool EYE = true>
__global__ void __launch_bounds__(256) k_bptAccumulate(BptArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t numPaths = a.fp.imageW * a.fp.imageH;
    if (i >= numPaths) return;
    const uint32_t px = i % a.fp.imageW, py = i / a.fp.imageW;
    const uint32_t p = py * a.fp.width + px;
    constexpr int C = N == 4 ? 16 : 3;
    const size_t plane = static_cast<size_t>(a.fp.width) * a.fp.height;
    float *acc = a.fb.accum + p;
    const bool reset = a.fp.frameIndex == 1;
    // light tracing (EYE = false) has no eye contribution: only the splats are added
    Spec<N> c = EYE ? loadSpec<N>(a.fb.ps.contrib[p]) : Spec<N>(0.0f);
    bool finite = c.allFinite();
    if (!finite) atomicAdd(&a.fb.counters[kStatDropped], 1u);
    float addend[C];
    if constexpr (N == 4) {
        const float recBinWidth = 16 / (830.0f - 360.0f);
#pragma unroll
        for (int k = 0; k < 16; ++k) addend[k] = 0.0f;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            float lam = 360.0f + (830.0f - 360.0f) * (w + a.bp.wlOffset) / 4;
            uint32_t bin = min(static_cast<uint32_t>((lam - 360.0f) / (830.0f - 360.0f) * 16), 15u);
#pragma unroll
            for (int k = 0; k < 16; ++k)
                if (k == static_cast<int>(bin)) addend[k] += c[w] * recBinWidth;
        }
    } else {
        addend[0] = c[0]; addend[1] = c[1]; addend[2] = c[2];
    }
#pragma unroll
    for (int k = 0; k < C; ++k) {
        float sum = reset ? 0.0f : acc[(2 * k) * plane];
        float comp = reset ? 0.0f : acc[(2 * k + 1) * plane];
        if (EYE && finite) {
            float cIn = addend[k] - comp;
            float tmp = sum + cIn;
            comp = (tmp - sum) - cIn;
            sum = tmp;
        }
        float sv = a.bp.splat[k * plane + p];
        float cIn = sv - comp;
        float tmp = sum + cIn;
        comp = (tmp - sum) - cIn;
        acc[(2 * k) * plane] = tmp;
        acc[(2 * k + 1) * plane] = comp;
        a.bp.splat[k * plane + p] = 0.0f;
    }
}

// ---------------------------------------------------------------------------------------------
// entry kernels for parity tests
struct RayIn { float org[3]; float tmin; float dir[3]; float tmax; };
struct HitOut { int32_t inst, geomInst, prim; float t, b1, b2; };

template <bool ANY>
__global__ void __launch_bounds__(kTraceBlock) k_traceRays(DevScene sc, const RayIn *rays, uint32_t n, HitOut *hits) {
    __shared__ int32_t lds[kTraceLds];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    RayIn r = rays[i];
    TraceHit h;
    h.inst = -1; h.geomInst = -1; h.prim = -1; h.t = 0; h.b1 = 0; h.b2 = 0;
    bool f = traceRay<ANY>(sc, V3(r.org[0], r.org[1], r.org[2]), V3(r.dir[0], r.dir[1], r.dir[2]), r.tmin, r.tmax,
                           &h, lds, threadIdx.x);
    HitOut o;
    o.inst = f ? h.inst : -1;
    o.geomInst = f ? h.geomInst : -1;
    o.prim = f ? h.prim : -1;
    o.t = f ? h.t : 0.0f;
    o.b1 = f ? h.b1 : 0.0f;
    o.b2 = f ? h.b2 : 0.0f;
    hits[i] = o;
}

struct BsdfQueryIn {
    float dirLocal[3];
    float gnLocal[3];
    float wlOffset, wlSelect;
    float uComponent, uDir[2];
    float evalDir[3];
    int32_t single;
    int32_t transportMode;
};
struct BsdfResultOut {
    float sampledDir[3];
    float sampledPDF;
    uint32_t sampledType;
    float sampledValue[4];
    float evalValue[4];
    float evalPDF;
    float revValue[4];
    float revPDF;
    float sampledRevValue[4];
    float sampledRevPDF;
};

template <int N>
__global__ void k_bsdfQueries(DevScene sc, DevMaterial m, const BsdfQueryIn *q, uint32_t n, BsdfResultOut *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    BsdfQueryIn Q = q[i];
    Wls<N> wls;
    wls.offset = Q.wlOffset;
    wls.sel = min(static_cast<uint32_t>(static_cast<uint16_t>(N * Q.wlSelect)), static_cast<uint32_t>(N - 1));
    wls.single = Q.single != 0;
    const bool importance = Q.transportMode != 0;
    Bsdf<N> b = setupBsdf<N>(sc, m, wls, static_cast<const SurfPt *>(nullptr));
    V3 qd(Q.dirLocal[0], Q.dirLocal[1], Q.dirLocal[2]);
    V3 gn(Q.gnLocal[0], Q.gnLocal[1], Q.gnLocal[2]);
    V3 dir(0, 0, 0);
    float pdf;
    uint32_t type;
    BsdfRev<N> srev;
    srev.value = Spec<N>(0.0f);
    Spec<N> f = b.sample(qd, gn, wls.sel, Q.uComponent, Q.uDir[0], Q.uDir[1], &dir, &pdf, &type, importance, &srev);
    BsdfResultOut o;
    o.sampledDir[0] = dir.x; o.sampledDir[1] = dir.y; o.sampledDir[2] = dir.z;
    o.sampledPDF = pdf;
    o.sampledType = type;
    V3 ed(Q.evalDir[0], Q.evalDir[1], Q.evalDir[2]);
    Spec<N> rv;
    Spec<N> e = b.evaluate(qd, gn, ed, wls.sel, importance, &rv);
    float rp = 0.0f;
    o.evalPDF = b.evaluatePDF(qd, ed, wls.sel, &rp);
    o.revPDF = rp;
    o.sampledRevPDF = pdf > 0.0f ? srev.pdf : 0.0f;
    for (int c = 0; c < 4; ++c) {
        o.sampledValue[c] = c < N ? f[c] : 0.0f;
        o.evalValue[c] = c < N ? e[c] : 0.0f;
        o.revValue[c] = c < N ? rv[c] : 0.0f;
        o.sampledRevValue[c] = (c < N && pdf > 0.0f) ? srev.value[c] : 0.0f;
    }
    out[i] = o;
}

// ---------------------------------------------------------------------------------------------
// launchers
static inline uint32_t blocksFor(uint32_t n, uint32_t bs) { return (n + bs - 1) / bs; }

template <int N, bool NODES>
static hipError_t launchFrameT(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb, hipStream_t s,
                               LaunchTimer *timer, const CacheLaunch *cache, bool convert,
                               hipEvent_t afterWavefront) {
    KArgs a;
    a.sc = sc;
    a.fp = fp;
    a.fb = fb;
    a.depth = 0;
    a.jobs = nullptr;
    a.jobOffsets = nullptr;
    a.activeOffsets = nullptr;
    a.numJobs = 0;
    const uint32_t n = fp.pixelCount;
    auto t0 = [&](int st) { if (timer) timer->begin(st); };
    auto t1 = [&](int st) { if (timer) timer->end(st); };
    hipLaunchKernelGGL(k_frameBegin, dim3(1), dim3(256), 0, s, fb.counters, n, fp.sortShade ? fb.sortCounts : nullptr);
    t0(STAGE_GENERATE);
    hipLaunchKernelGGL(k_generate<N>, dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
    t1(STAGE_GENERATE);
    const uint32_t numDepths = std::min<uint32_t>(fp.maxPathLength, kMaxDepthCounters - 1);
    // the guides are written by the depth-0 shade kernel
    const bool guides = fb.guideAlbedo != nullptr;
    const uint32_t wavefrontDepths = std::min(std::max(fp.wavefrontDepth, guides ? 1u : 0u), numDepths);
    for (uint32_t d = 0; d < wavefrontDepths; ++d) {
        a.depth = d;
        t0(STAGE_INTERSECT);
        // scenes with alpha nodes run the alpha variants, which keep no traversal statistics
        if (sc.hasAlpha)
            hipLaunchKernelGGL((k_intersect<false, false, true>), dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
        else if (fp.traversalStats)
            hipLaunchKernelGGL(k_intersect<true>, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
        else
            hipLaunchKernelGGL(k_intersect<false>, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
        t1(STAGE_INTERSECT);
        t0(STAGE_SHADE);
        if (fp.sortShade) hipLaunchKernelGGL(k_sortScatter, dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
        if (guides && d == 0)
            hipLaunchKernelGGL((k_shade<N, false, NODES, true>), dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_shade<N, false, NODES>), dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
        t1(STAGE_SHADE);
        t0(STAGE_SHADOW);
        if (sc.hasAlpha)
            hipLaunchKernelGGL((k_shadow<false, false, true>), dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
        else if (fp.traversalStats)
            hipLaunchKernelGGL(k_shadow<true>, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
        else
            hipLaunchKernelGGL(k_shadow<false>, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
        t1(STAGE_SHADOW);
    }
    if (afterWavefront) {
        hipError_t e = hipEventRecord(afterWavefront, s);
        if (e != hipSuccess) return e;
    }
    if (wavefrontDepths < numDepths) {
        a.depth = wavefrontDepths;
        const uint32_t blocks = std::max(1u, std::min(blocksFor(n, kTraceBlock), fp.tailBlocks));
        t0(STAGE_TAIL);
        if (fp.traversalStats)
            hipLaunchKernelGGL((k_tail<N, true, false, NODES>), dim3(blocks), dim3(kTraceBlock), 0, s, a);
        else
            hipLaunchKernelGGL((k_tail<N, false, false, NODES>), dim3(blocks), dim3(kTraceBlock), 0, s, a);
        t1(STAGE_TAIL);
    }
    if (cache && fp.cacheMode == 2) {
        t0(STAGE_CACHE_QUERY);
        hipError_t e = rcLaunchQuery(cache->net, reinterpret_cast<const float *>(fb.cache.queryRaw),
                                     fb.cache.queryThroughput, fb.cache.queryStamp, fp.frameStamp,
                                     fp.width * fp.height, fb.ps.contrib, N, s);
        if (e != hipSuccess) return e;
        t1(STAGE_CACHE_QUERY);
    }
    // the accumulation also writes the converted output when the frame asks for it
    t0(STAGE_ACCUMULATE);
    if (convert)
        hipLaunchKernelGGL((k_accumulate<N, false, true>), dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_accumulate<N>, dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
    t1(STAGE_ACCUMULATE);
    hipLaunchKernelGGL(k_frameEnd, dim3(1), dim3(64), 0, s, fb.counters, fp.maxPathLength, wavefrontDepths);
    if (cache && fp.cacheMode != 0) {
        t0(STAGE_CACHE_TRAIN);
        const uint32_t ns = cache->numSlots, scanBlocks = blocksFor(ns, kSlotScanBlock);
        if (scanBlocks > 64) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_rcSlotFlags<N>, dim3(blocksFor(ns, 256)), dim3(256), 0, s, a, ns, cache->slotFlags,
                           cache->slotTargets);
        hipLaunchKernelGGL(k_rcScanSlots, dim3(scanBlocks), dim3(kSlotScanBlock), 0, s, cache->slotFlags, ns,
                           cache->slotOffsets, cache->slotBlockSums);
        hipLaunchKernelGGL(k_rcScanSlotBlocks, dim3(1), dim3(64), 0, s, cache->slotBlockSums, scanBlocks,
                           cache->batchCount);
        hipLaunchKernelGGL(k_rcGatherSlots, dim3(blocksFor(ns, 256)), dim3(256), 0, s, fb.cache.trainRaw,
                           cache->slotFlags, cache->slotOffsets, cache->slotBlockSums, cache->slotTargets, ns,
                           cache->batchRaw, cache->batchTarget);
        hipError_t e = rcLaunchTrain(cache->net, cache->train, reinterpret_cast<const float *>(cache->batchRaw),
                                     reinterpret_cast<const float *>(cache->batchTarget), cache->batchCount,
                                     cache->numSlots, s);
        if (e != hipSuccess) return e;
        t1(STAGE_CACHE_TRAIN);
    }
    return hipGetLastError();
}

// One step of a batch: every job with samples left renders one more sample per pixel.  The
// launch sequence reads nothing from the host that changes between steps (the jobs' frame
// counters live on the device), so it can be captured once into a graph and replayed.
template <int N>
static hipError_t launchBatchStepT(const BatchLaunch &b, hipStream_t s) {
    KArgs a;
    a.sc = b.anyScene;
    a.fp = b.fp;
    a.fb = b.fb;
    a.depth = 0;
    a.jobs = b.jobs;
    a.jobOffsets = b.jobOffsets;
    a.activeOffsets = b.activeOffsets;
    a.numJobs = b.numJobs;
    const uint32_t n = b.fp.imageW * b.fp.imageH;
    hipLaunchKernelGGL(k_batchStepBegin, dim3(1), dim3(kStepBeginThreads), 0, s, b.jobs, b.numJobs, b.activeOffsets);
    hipLaunchKernelGGL(k_frameBegin, dim3(1), dim3(256), 0, s, b.fb.counters, 0u, static_cast<uint32_t *>(nullptr));
    hipLaunchKernelGGL((k_generate<N, true>), dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
    const uint32_t numDepths = std::min<uint32_t>(b.fp.maxPathLength, kMaxDepthCounters - 1);
    const uint32_t wavefrontDepths = std::min(b.fp.wavefrontDepth, numDepths);
    for (uint32_t d = 0; d < wavefrontDepths; ++d) {
        a.depth = d;
        hipLaunchKernelGGL((k_intersect<false, true>), dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
        if (b.nodePrograms)
            hipLaunchKernelGGL((k_shade<N, true, true>), dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_shade<N, true, false>), dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
        hipLaunchKernelGGL((k_shadow<false, true>), dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
    }
    if (wavefrontDepths < numDepths) {
        a.depth = wavefrontDepths;
        const uint32_t blocks = std::max(1u, std::min(blocksFor(n, kTraceBlock), b.fp.tailBlocks));
        if (b.nodePrograms)
            hipLaunchKernelGGL((k_tail<N, false, true, true>), dim3(blocks), dim3(kTraceBlock), 0, s, a);
        else
            hipLaunchKernelGGL((k_tail<N, false, true, false>), dim3(blocks), dim3(kTraceBlock), 0, s, a);
    }
    hipLaunchKernelGGL((k_accumulate<N, true>), dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_frameEnd, dim3(1), dim3(64), 0, s, b.fb.counters, b.fp.maxPathLength, wavefrontDepths);
    return hipGetLastError();
}

template <int N>
static hipError_t launchBptFrameT(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                  const BptParams &bp, hipStream_t s, LaunchTimer *timer) {
    BptArgs a;
    a.sc = sc;
    a.fp = fp;
    a.fb = fb;
    a.bp = bp;
    KArgs k;
    k.sc = sc;
    k.fp = fp;
    k.fb = fb;
    k.depth = 0;
    k.jobs = nullptr;
    k.jobOffsets = nullptr;
    k.activeOffsets = nullptr;
    k.numJobs = 0;
    const uint32_t n = fp.imageW * fp.imageH;
    const uint32_t scanBlocks = blocksFor(bp.numLightPaths, kScanBlock);
    if (bp.numLightPaths == 0 || scanBlocks > kScanBlock) return hipErrorInvalidValue;
    auto t0 = [&](int st) { if (timer) timer->begin(st); };
    auto t1 = [&](int st) { if (timer) timer->end(st); };
    t0(STAGE_LIGHT_PATHS);
    if (sc.hasNodePrograms)
        hipLaunchKernelGGL((k_bptLight<N, true>), dim3(blocksFor(bp.numLightPaths, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_bptLight<N, false>), dim3(blocksFor(bp.numLightPaths, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
    hipLaunchKernelGGL(k_bptScanBlocks, dim3(scanBlocks), dim3(kScanBlock), 0, s, bp);
    hipLaunchKernelGGL(k_bptScanTop, dim3(1), dim3(kScanBlock), 0, s, bp, scanBlocks);
    const uint64_t compactThreads = static_cast<uint64_t>(bp.numLightPaths) * bp.slotsPerPath * kLVF4;
    hipLaunchKernelGGL(k_bptCompact, dim3(static_cast<uint32_t>((compactThreads + 255) / 256)), dim3(256), 0, s, bp);
    t1(STAGE_LIGHT_PATHS);
    t0(STAGE_EYE_PATHS);
    const bool guides = fb.guideAlbedo != nullptr;
    if (sc.hasNodePrograms && guides)
        hipLaunchKernelGGL((k_bptEye<N, true, true>), dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
    else if (sc.hasNodePrograms)
        hipLaunchKernelGGL((k_bptEye<N, true>), dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
    else if (guides)
        hipLaunchKernelGGL((k_bptEye<N, false, true>), dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_bptEye<N, false>), dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
    t1(STAGE_EYE_PATHS);
    t0(STAGE_ACCUMULATE);
    hipLaunchKernelGGL(k_bptAccumulate<N>, dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
    t1(STAGE_ACCUMULATE);
    t0(STAGE_CONVERT);
    hipLaunchKernelGGL(k_convert<N>, dim3(blocksFor(n, 256)), dim3(256), 0, s, k);
    t1(STAGE_CONVERT);
    return hipGetLastError();
}

// light tracing (context.cpp:1444-1463): the light paths splat, then the splats are added to
// the accumulation (reset on the first frame) and converted
template <int N>
static hipError_t launchLightTraceFrameT(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                         const BptParams &bp, hipStream_t s, LaunchTimer *timer) {
    BptArgs a;
    a.sc = sc;
    a.fp = fp;
    a.fb = fb;
    a.bp = bp;
    KArgs k;
    k.sc = sc;
    k.fp = fp;
    k.fb = fb;
    k.depth = 0;
    k.jobs = nullptr;
    k.jobOffsets = nullptr;
    k.activeOffsets = nullptr;
    k.numJobs = 0;
    const uint32_t n = fp.imageW * fp.imageH;
    if (bp.numLightPaths == 0) return hipErrorInvalidValue;
    auto t0 = [&](int st) { if (timer) timer->begin(st); };
    auto t1 = [&](int st) { if (timer) timer->end(st); };
    if (fb.guideAlbedo) {
        // the guides' primary rays come first, as the reference launches its aux-buffer pass
        t0(STAGE_EYE_PATHS);
        if (sc.hasNodePrograms)
            hipLaunchKernelGGL((k_ltGuides<N, true>), dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, k);
        else
            hipLaunchKernelGGL((k_ltGuides<N, false>), dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, k);
        t1(STAGE_EYE_PATHS);
    }
    t0(STAGE_LIGHT_PATHS);
    if (sc.hasNodePrograms)
        hipLaunchKernelGGL((k_lightTrace<N, true>), dim3(blocksFor(bp.numLightPaths, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_lightTrace<N, false>), dim3(blocksFor(bp.numLightPaths, kTraceBlock)), dim3(kTraceBlock), 0, s, a);
    t1(STAGE_LIGHT_PATHS);
    t0(STAGE_ACCUMULATE);
    hipLaunchKernelGGL((k_bptAccumulate<N, false>), dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
    t1(STAGE_ACCUMULATE);
    t0(STAGE_CONVERT);
    hipLaunchKernelGGL(k_convert<N>, dim3(blocksFor(n, 256)), dim3(256), 0, s, k);
    t1(STAGE_CONVERT);
    return hipGetLastError();
}

// ---- direct evaluation of a node program at given surface points ----
// point: position 3, geometric normal 3, shading frame x 3, y 3, z 3, texture coordinates 2,
// wavelength offset 1, pad 2 (20 floats); instance 0 of the scene record is the identity
template <int N>
__global__ void __launch_bounds__(128) k_evalNodePlug(DevScene sc, uint32_t base, uint32_t count, uint32_t cls,
                                                      const float *points, uint32_t n, float4 *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *p = points + 20 * size_t(i);
    SurfPt sp;
    sp.position = V3(p[0], p[1], p[2]);
    sp.gn = V3(p[3], p[4], p[5]);
    sp.frame.x = V3(p[6], p[7], p[8]);
    sp.frame.y = V3(p[9], p[10], p[11]);
    sp.frame.z = V3(p[12], p[13], p[14]);
    sp.texU = p[15];
    sp.texV = p[16];
    sp.isPoint = sp.atInfinity = false;
    sp.instIndex = sp.geomInstIndex = sp.primIndex = 0;
    sp.u = sp.v = 0.0f;
    Wls<N> wls;
    wls.offset = p[17];
    wls.sel = 0;
    wls.single = false;
    NodeOutputs<N> o;
    runNodeProgram<N, true>(sc, base, count, sp, wls, cls, cls, &o);
    float4 r = make_float4(0, 0, 0, 0);
    if (cls == VC_SPEC) {
        if (o.has & 1u) {
            r.x = o.spec[0][0];
            r.y = o.spec[0][1];
            r.z = o.spec[0][2];
            if constexpr (N == 4) r.w = o.spec[0][3];
        }
    } else if (o.has & 8u) {
        r = o.vec[0];
    }
    out[i] = r;
}

}  // namespace VLRA_UNIT_NS

// ---- entry points ----
// launchers of one channel count each (units 1-4), called by the dispatchers (unit 5)
hipError_t launchFrameRgb(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb, hipStream_t s,
                          LaunchTimer *timer, const CacheLaunch *cache, bool convert, hipEvent_t afterWavefront);
hipError_t launchFrameSpectral(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb, hipStream_t s,
                               LaunchTimer *timer, const CacheLaunch *cache, bool convert,
                               hipEvent_t afterWavefront);
hipError_t launchFrameNodesRgb(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb, hipStream_t s,
                               LaunchTimer *timer, const CacheLaunch *cache, bool convert,
                               hipEvent_t afterWavefront);
hipError_t launchFrameNodesSpectral(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                    hipStream_t s, LaunchTimer *timer, const CacheLaunch *cache, bool convert,
                                    hipEvent_t afterWavefront);
hipError_t launchBatchStepRgb(const BatchLaunch &b, hipStream_t s);
hipError_t launchBatchStepSpectral(const BatchLaunch &b, hipStream_t s);
hipError_t launchBptFrameRgb(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb, const BptParams &bp,
                             hipStream_t s, LaunchTimer *timer);
hipError_t launchBptFrameSpectral(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                  const BptParams &bp, hipStream_t s, LaunchTimer *timer);
hipError_t launchLightTraceFrameRgb(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                    const BptParams &bp, hipStream_t s, LaunchTimer *timer);
hipError_t launchLightTraceFrameSpectral(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                         const BptParams &bp, hipStream_t s, LaunchTimer *timer);

#if VLRA_UNIT == 0 || VLRA_UNIT == 1
hipError_t launchFrameRgb(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb, hipStream_t s,
                          LaunchTimer *timer, const CacheLaunch *cache, bool convert, hipEvent_t afterWavefront) {
    return VLRA_UNIT_NS::launchFrameT<3, false>(sc, fp, fb, s, timer, cache, convert, afterWavefront);
}
#endif
#if VLRA_UNIT == 0 || VLRA_UNIT == 8
hipError_t launchFrameNodesRgb(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb, hipStream_t s,
                               LaunchTimer *timer, const CacheLaunch *cache, bool convert,
                               hipEvent_t afterWavefront) {
    return VLRA_UNIT_NS::launchFrameT<3, true>(sc, fp, fb, s, timer, cache, convert, afterWavefront);
}
#endif
#if VLRA_UNIT == 0 || VLRA_UNIT == 6
hipError_t launchBatchStepRgb(const BatchLaunch &b, hipStream_t s) { return VLRA_UNIT_NS::launchBatchStepT<3>(b, s); }
#endif
#if VLRA_UNIT == 0 || VLRA_UNIT == 2
hipError_t launchFrameSpectral(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb, hipStream_t s,
                               LaunchTimer *timer, const CacheLaunch *cache, bool convert,
                               hipEvent_t afterWavefront) {
    return VLRA_UNIT_NS::launchFrameT<4, false>(sc, fp, fb, s, timer, cache, convert, afterWavefront);
}
#endif
#if VLRA_UNIT == 0 || VLRA_UNIT == 9
hipError_t launchFrameNodesSpectral(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                    hipStream_t s, LaunchTimer *timer, const CacheLaunch *cache, bool convert,
                                    hipEvent_t afterWavefront) {
    return VLRA_UNIT_NS::launchFrameT<4, true>(sc, fp, fb, s, timer, cache, convert, afterWavefront);
}
#endif
#if VLRA_UNIT == 0 || VLRA_UNIT == 7
hipError_t launchBatchStepSpectral(const BatchLaunch &b, hipStream_t s) {
    return VLRA_UNIT_NS::launchBatchStepT<4>(b, s);
}
#endif
#if VLRA_UNIT == 0 || VLRA_UNIT == 3
hipError_t launchBptFrameRgb(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb, const BptParams &bp,
                             hipStream_t s, LaunchTimer *timer) {
    return VLRA_UNIT_NS::launchBptFrameT<3>(sc, fp, fb, bp, s, timer);
}
hipError_t launchLightTraceFrameRgb(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                    const BptParams &bp, hipStream_t s, LaunchTimer *timer) {
    return VLRA_UNIT_NS::launchLightTraceFrameT<3>(sc, fp, fb, bp, s, timer);
}
#endif
#if VLRA_UNIT == 0 || VLRA_UNIT == 4
hipError_t launchBptFrameSpectral(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                  const BptParams &bp, hipStream_t s, LaunchTimer *timer) {
    return VLRA_UNIT_NS::launchBptFrameT<4>(sc, fp, fb, bp, s, timer);
}
hipError_t launchLightTraceFrameSpectral(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                         const BptParams &bp, hipStream_t s, LaunchTimer *timer) {
    return VLRA_UNIT_NS::launchLightTraceFrameT<4>(sc, fp, fb, bp, s, timer);
}
#endif

#if VLRA_UNIT == 0 || VLRA_UNIT == 5
using namespace VLRA_UNIT_NS;

hipError_t launchFrame(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb, bool spectral,
                       hipStream_t s, LaunchTimer *timer, const CacheLaunch *cache, bool convert,
                       hipEvent_t afterWavefront) {
    if (sc.hasNodePrograms)
        return spectral ? launchFrameNodesSpectral(sc, fp, fb, s, timer, cache, convert, afterWavefront)
                        : launchFrameNodesRgb(sc, fp, fb, s, timer, cache, convert, afterWavefront);
    return spectral ? launchFrameSpectral(sc, fp, fb, s, timer, cache, convert, afterWavefront)
                    : launchFrameRgb(sc, fp, fb, s, timer, cache, convert, afterWavefront);
}

hipError_t launchConvert(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb, bool spectral,
                         hipStream_t s, LaunchTimer *timer) {
    KArgs a;
    a.sc = sc;
    a.fp = fp;
    a.fb = fb;
    a.depth = 0;
    a.jobs = nullptr;
    a.jobOffsets = nullptr;
    a.activeOffsets = nullptr;
    a.numJobs = 0;
    const uint32_t n = fp.pixelCount;
    if (timer) timer->begin(STAGE_CONVERT);
    if (spectral)
        hipLaunchKernelGGL(k_convert<4>, dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_convert<3>, dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
    if (timer) timer->end(STAGE_CONVERT);
    return hipGetLastError();
}

hipError_t launchBatchStep(const BatchLaunch &b, bool spectral, hipStream_t s) {
    if (b.numJobs == 0) return hipSuccess;
    return spectral ? launchBatchStepSpectral(b, s) : launchBatchStepRgb(b, s);
}

hipError_t launchBatchConvert(const BatchLaunch &b, bool spectral, hipStream_t s) {
    if (b.numJobs == 0) return hipSuccess;
    KArgs a;
    a.sc = b.anyScene;
    a.fp = b.fp;
    a.fb = b.fb;
    a.depth = 0;
    a.jobs = b.jobs;
    a.jobOffsets = b.jobOffsets;
    a.activeOffsets = b.activeOffsets;
    a.numJobs = b.numJobs;
    const uint32_t n = b.fp.imageW * b.fp.imageH;
    if (spectral)
        hipLaunchKernelGGL((k_convert<4, true>), dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_convert<3, true>), dim3(blocksFor(n, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launchLightTraceFrame(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb,
                                 const BptParams &bp, bool spectral, hipStream_t s, LaunchTimer *timer) {
    return spectral ? launchLightTraceFrameSpectral(sc, fp, fb, bp, s, timer)
                    : launchLightTraceFrameRgb(sc, fp, fb, bp, s, timer);
}

hipError_t launchBptFrame(const DevScene &sc, const FrameParams &fp, const FrameBuffers &fb, const BptParams &bp,
                          bool spectral, hipStream_t s, LaunchTimer *timer) {
    return spectral ? launchBptFrameSpectral(sc, fp, fb, bp, s, timer) : launchBptFrameRgb(sc, fp, fb, bp, s, timer);
}

hipError_t launchTraceRays(const DevScene &sc, const void *rays, uint32_t n, void *hits, bool anyHit, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (anyHit)
        hipLaunchKernelGGL(k_traceRays<true>, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, sc,
                           static_cast<const RayIn *>(rays), n, static_cast<HitOut *>(hits));
    else
        hipLaunchKernelGGL(k_traceRays<false>, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, s, sc,
                           static_cast<const RayIn *>(rays), n, static_cast<HitOut *>(hits));
    return hipGetLastError();
}

hipError_t launchEvalNodePlug(const DevScene &sc, uint32_t base, uint32_t count, uint32_t plugType, bool spectral,
                              const float *points, uint32_t n, float *out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint32_t cls = valueClass(plugType);
    if (spectral)
        hipLaunchKernelGGL(k_evalNodePlug<4>, dim3(blocksFor(n, 128)), dim3(128), 0, s, sc, base, count, cls, points, n,
                           reinterpret_cast<float4 *>(out));
    else
        hipLaunchKernelGGL(k_evalNodePlug<3>, dim3(blocksFor(n, 128)), dim3(128), 0, s, sc, base, count, cls, points, n,
                           reinterpret_cast<float4 *>(out));
    return hipGetLastError();
}

hipError_t launchBsdfQueries(const DevScene &sc, const DevMaterial &m, bool spectral, const void *q, uint32_t n,
                             void *r, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (spectral)
        hipLaunchKernelGGL(k_bsdfQueries<4>, dim3(blocksFor(n, 128)), dim3(128), 0, s, sc, m,
                           static_cast<const BsdfQueryIn *>(q), n, static_cast<BsdfResultOut *>(r));
    else
        hipLaunchKernelGGL(k_bsdfQueries<3>, dim3(blocksFor(n, 128)), dim3(128), 0, s, sc, m,
                           static_cast<const BsdfQueryIn *>(q), n, static_cast<BsdfResultOut *>(r));
    return hipGetLastError();
}

#endif

} // namespace vlra
