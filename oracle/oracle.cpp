// Oracle (test infrastructure only): CPU restatement of the reference's unidirectional path
// tracer and the scene set-up it depends on.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this library.
//
// Sources followed (libVLR/ unless stated):
//   per-pixel PCG32 seeding          context.cpp:1182-1187
//   path tracer RG / CH / MS         GPU_kernels/path_tracing.cu:30-128, 133-259, 266-337
//   light selection, visibility      shared/light_transport_common.h:147-192
//   hit decode, light sampling       GPU_kernels/triangle.cu:6-53, 113-177; point.cu:38-80;
//                                    infinite_sphere.cu:60-114
//   surface point                    shared/renderer_common.h:609-626
//   distributions                    shared/shared.h:182-298, context.cpp:1797-1941
//   geometry/instance set-up         scene.cpp:184-297, 1204-1402
//   scene bounds                     GPU_kernels/setup_scene.cu:7-92
//   accumulate / convert             GPU_kernels/post_process.cu:93-163
//   env texture node, uvs format     GPU_kernels/shader_nodes.cu:457-483, image.cpp:542-565, 1009-1019
//   env importance map               shader_nodes.cpp:1992-2010, image.cpp:792-860
// Intersection: OptiX's traversal is closed; the oracle defines closest hit by the watertight
// test of Woop, Benthin and Wald (JCGT 2(1), 2013) on the raw vertex positions, minimum t,
// ties resolved by first-found.  The product uses the same definition.
#include "oracle.h"
#include "ov_bsdf.h"

#include <vector>
#include <random>
#include <memory>
#include <atomic>
#include <cfloat>
#include <omp.h>

namespace ov {

static SpectralTables g_tables;
SpectralTables &tables() { return g_tables; }

// ------------------------------------------------------------------------------------------
// IEEE half <-> float (round to nearest even), the storage format of RGBA16F/uvsA16F images.
static inline float halfToFloat(uint16_t h) {
    uint32_t sign = uint32_t(h & 0x8000) << 16;
    uint32_t exp = (h >> 10) & 0x1F;
    uint32_t man = h & 0x3FF;
    uint32_t bits;
    if (exp == 0) {
        if (man == 0) {
            bits = sign;
        } else {
            int e = -1;
            do { ++e; man <<= 1; } while ((man & 0x400) == 0);
            man &= 0x3FF;
            bits = sign | uint32_t(127 - 15 - e) << 23 | (man << 13);
        }
    } else if (exp == 31) {
        bits = sign | 0x7F800000u | (man << 13);
    } else {
        bits = sign | ((exp - 15 + 127) << 23) | (man << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}
static inline uint16_t floatToHalf(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000;
    uint32_t absx = x & 0x7FFFFFFF;
    if (absx >= 0x7F800000)
        return uint16_t(sign | 0x7C00 | (absx > 0x7F800000 ? 0x200 : 0));
    if (absx >= 0x477FF000) // rounds to >= 65520 -> inf
        return uint16_t(sign | 0x7C00);
    if (absx < 0x38800000) { // subnormal half
        if (absx < 0x33000000)
            return uint16_t(sign);
        uint32_t e = absx >> 23;
        uint32_t m = (absx & 0x7FFFFF) | 0x800000;
        uint32_t shift = 126 - e; // 14 - (e - 127) + 13 - 1 ... gives mantissa in half units
        uint32_t hm = m >> (shift);
        uint32_t rem = m & ((1u << shift) - 1);
        uint32_t halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (hm & 1)))
            ++hm;
        return uint16_t(sign | hm);
    }
    uint32_t e = (absx >> 23) - 127 + 15;
    uint32_t m = absx & 0x7FFFFF;
    uint32_t hm = m >> 13;
    uint32_t rem = m & 0x1FFF;
    uint32_t h = (e << 10) | hm;
    if (rem > 0x1000 || (rem == 0x1000 && (hm & 1)))
        ++h;
    return uint16_t(sign | h);
}

// ------------------------------------------------------------------------------------------
// Distributions (shared.h:182-298, context.cpp:1797-1941)
struct Discrete1D {
    std::vector<float> PMF, CDF;
    float integral = 0;
    uint32_t n = 0;
    void init(const float *values, size_t num) {
        n = static_cast<uint32_t>(num);
        PMF.assign(values, values + num);
        CDF.resize(num + 1);
        Kahan<float> sum;
        sum.reset();
        for (uint32_t i = 0; i < n; ++i) {
            CDF[i] = sum.result;
            sum.add(PMF[i]);
        }
        integral = sum.result;
        for (uint32_t i = 0; i < n; ++i) {
            PMF[i] /= integral;
            CDF[i] /= integral;
        }
        CDF[n] = 1.0f;
    }
    static uint32_t nextPow2(uint32_t x) { uint32_t p = 1; while (p < x) p <<= 1; return p; }
    uint32_t search(float u) const {
        int idx = 0;
        for (int d = nextPow2(n) >> 1; d >= 1; d >>= 1) {
            if (idx + d >= static_cast<int>(n))
                continue;
            if (CDF[idx + d] <= u)
                idx += d;
        }
        return idx;
    }
    uint32_t sample(float u, float *prob) const {
        uint32_t idx = search(u);
        *prob = PMF[idx];
        return idx;
    }
    uint32_t sample(float u, float *prob, float *remapped) const {
        uint32_t idx = search(u);
        *prob = PMF[idx];
        *remapped = (u - CDF[idx]) / (CDF[idx + 1] - CDF[idx]);
        return idx;
    }
};

struct RegConst1D {
    std::vector<float> PDF, CDF;
    float integral = 0;
    uint32_t n = 0;
    void init(const float *values, size_t num) {
        n = static_cast<uint32_t>(num);
        PDF.assign(values, values + num);
        CDF.resize(num + 1);
        Kahan<float> sum;
        sum.reset();
        for (uint32_t i = 0; i < n; ++i) {
            CDF[i] = sum.result;
            sum.add(PDF[i] / n);
        }
        integral = sum.result;
        for (uint32_t i = 0; i < n; ++i) {
            PDF[i] /= integral;
            CDF[i] /= integral;
        }
        CDF[n] = 1.0f;
    }
    float sample(float u, float *pd) const {
        int idx = 0;
        for (int d = Discrete1D::nextPow2(n) >> 1; d >= 1; d >>= 1) {
            if (idx + d >= static_cast<int>(n))
                continue;
            if (CDF[idx + d] <= u)
                idx += d;
        }
        *pd = PDF[idx];
        float t = (u - CDF[idx]) / (CDF[idx + 1] - CDF[idx]);
        return (idx + t) / n;
    }
    float evaluatePDF(float smp) const {
        int32_t idx = std::min<int32_t>(n - 1, static_cast<int32_t>(smp * n));
        return PDF[idx];
    }
};

struct RegConst2D {
    std::vector<RegConst1D> rows;
    RegConst1D top;
    void init(const float *values, size_t numD1, size_t numD2) {
        rows.resize(numD2);
        std::vector<float> integrals(numD2);
        for (size_t i = 0; i < numD2; ++i) {
            rows[i].init(values + i * numD1, numD1);
            integrals[i] = rows[i].integral;
        }
        top.init(integrals.data(), numD2);
    }
    void sample(float u0, float u1, float *d0, float *d1, float *pd) const {
        float topPDF;
        *d1 = top.sample(u1, &topPDF);
        uint32_t idx = std::min(uint32_t(top.n * *d1), top.n - 1);
        *d0 = rows[idx].sample(u0, pd);
        *pd *= topPDF;
    }
    float evaluatePDF(float d0, float d1) const {
        uint32_t idx = std::min(uint32_t(top.n * d1), top.n - 1);
        return top.evaluatePDF(d1) * rows[idx].evaluatePDF(d0);
    }
};

// ------------------------------------------------------------------------------------------
struct PCG32 {
    uint64_t state;
    // random_distributions.h:6-27
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + 1;
        uint32_t xorshifted = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = old >> 59u;
        return (xorshifted >> rot) | (xorshifted << ((-static_cast<int32_t>(rot)) & 31));
    }
    float uniform() {
        uint32_t bits = (next() >> 9) | 0x3f800000;
        float f;
        std::memcpy(&f, &bits, 4);
        return f - 1.0f;
    }
};

// ------------------------------------------------------------------------------------------
// Watertight ray/triangle test (Woop, Benthin, Wald 2013)
struct RayPre {
    V3 org, dir;
    int kx, ky, kz;
    float Sx, Sy, Sz;
    V3 invDir;
    void init(const V3 &o, const V3 &d) {
        org = o;
        dir = d;
        kz = 0;
        if (std::fabs(d.y) > std::fabs(d[kz])) kz = 1;
        if (std::fabs(d.z) > std::fabs(d[kz])) kz = 2;
        kx = kz == 2 ? 0 : kz + 1;
        ky = kx == 2 ? 0 : kx + 1;
        if (d[kz] < 0.0f)
            std::swap(kx, ky);
        Sx = d[kx] / d[kz];
        Sy = d[ky] / d[kz];
        Sz = 1.0f / d[kz];
        invDir = V3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    }
};

static inline bool intersectTriangle(const RayPre &r, const V3 &p0, const V3 &p1, const V3 &p2, float tmin,
                                     float tmax, float *tOut, float *b1Out, float *b2Out) {
    const V3 A = p0 - r.org, B = p1 - r.org, C = p2 - r.org;
    const float Ax = A[r.kx] - r.Sx * A[r.kz];
    const float Ay = A[r.ky] - r.Sy * A[r.kz];
    const float Bx = B[r.kx] - r.Sx * B[r.kz];
    const float By = B[r.ky] - r.Sy * B[r.kz];
    const float Cx = C[r.kx] - r.Sx * C[r.kz];
    const float Cy = C[r.ky] - r.Sy * C[r.kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = static_cast<float>(static_cast<double>(Cx) * By - static_cast<double>(Cy) * Bx);
        V = static_cast<float>(static_cast<double>(Ax) * Cy - static_cast<double>(Ay) * Cx);
        W = static_cast<float>(static_cast<double>(Bx) * Ay - static_cast<double>(By) * Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return false;
    const float det = U + V + W;
    if (det == 0.0f)
        return false;
    const float Az = r.Sz * A[r.kz];
    const float Bz = r.Sz * B[r.kz];
    const float Cz = r.Sz * C[r.kz];
    const float T = U * Az + V * Bz + W * Cz;
    const float rcpDet = 1.0f / det;
    const float t = T * rcpDet;
    if (!(t >= tmin && t < tmax))
        return false;
    *tOut = t;
    *b1Out = V * rcpDet;
    *b2Out = W * rcpDet;
    return true;
}

struct AABB {
    V3 lo = V3(FLT_MAX), hi = V3(-FLT_MAX);
    void unify(const V3 &p) {
        lo = V3(std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z));
        hi = V3(std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z));
    }
    void unify(const AABB &b) { unify(b.lo); unify(b.hi); }
    V3 centroid() const { return 0.5f * (lo + hi); }
    float area() const {
        V3 d = hi - lo;
        if (d.x < 0) return 0;
        return 2 * (d.x * d.y + d.y * d.z + d.z * d.x);
    }
};

// Conservative slab test: the far distances and the current closest distance are widened by
// 8 ulp-sized relative steps, so float rounding (including the re-rounding of t in another
// instance's object space) never culls a box that holds a closer hit.
static inline bool hitBox(const AABB &b, const RayPre &r, float tmin, float tmax) {
    constexpr float kWiden = 1.0f + 8.0f * FLT_EPSILON;
    float t0 = tmin, t1 = tmax * kWiden;
    for (int a = 0; a < 3; ++a) {
        float tn = (b.lo[a] - r.org[a]) * r.invDir[a];
        float tf = (b.hi[a] - r.org[a]) * r.invDir[a];
        if (tn > tf) std::swap(tn, tf);
        tf *= kWiden;
        if (tn == tn) t0 = tn > t0 ? tn : t0;  // NaN (0*inf) leaves the interval as is
        if (tf == tf) t1 = tf < t1 ? tf : t1;
        if (t0 > t1)
            return false;
    }
    return true;
}

// Binned-SAH BVH2 used by the CPU baseline; correctness is pinned against brute force.
struct BVH {
    struct Node { AABB box; int32_t left = -1, right = -1; uint32_t first = 0, count = 0; };
    std::vector<Node> nodes;
    std::vector<uint32_t> order;
    void build(const std::vector<AABB> &boxes) {
        order.resize(boxes.size());
        for (size_t i = 0; i < boxes.size(); ++i) order[i] = static_cast<uint32_t>(i);
        nodes.clear();
        nodes.reserve(2 * boxes.size() + 1);
        if (boxes.empty()) {
            nodes.push_back(Node());
            return;
        }
        std::vector<V3> cent(boxes.size());
        for (size_t i = 0; i < boxes.size(); ++i) cent[i] = boxes[i].centroid();
        buildRec(boxes, cent, 0, static_cast<uint32_t>(boxes.size()));
    }
    int32_t buildRec(const std::vector<AABB> &boxes, const std::vector<V3> &cent, uint32_t begin, uint32_t end) {
        int32_t idx = static_cast<int32_t>(nodes.size());
        nodes.push_back(Node());
        AABB box, cbox;
        for (uint32_t i = begin; i < end; ++i) {
            box.unify(boxes[order[i]]);
            cbox.unify(cent[order[i]]);
        }
        nodes[idx].box = box;
        uint32_t n = end - begin;
        if (n <= 4) {
            nodes[idx].first = begin;
            nodes[idx].count = n;
            return idx;
        }
        constexpr int NB = 16;
        float bestCost = FLT_MAX;
        int bestAxis = -1, bestSplit = -1;
        for (int a = 0; a < 3; ++a) {
            float ext = cbox.hi[a] - cbox.lo[a];
            if (ext <= 0) continue;
            AABB bb[NB];
            uint32_t bc[NB] = {};
            for (uint32_t i = begin; i < end; ++i) {
                int b = std::min(NB - 1, static_cast<int>(NB * (cent[order[i]][a] - cbox.lo[a]) / ext));
                bb[b].unify(boxes[order[i]]);
                ++bc[b];
            }
            for (int s = 1; s < NB; ++s) {
                AABB l, r;
                uint32_t nl = 0, nr = 0;
                for (int b = 0; b < s; ++b) { if (bc[b]) l.unify(bb[b]); nl += bc[b]; }
                for (int b = s; b < NB; ++b) { if (bc[b]) r.unify(bb[b]); nr += bc[b]; }
                if (!nl || !nr) continue;
                float c = l.area() * nl + r.area() * nr;
                if (c < bestCost) { bestCost = c; bestAxis = a; bestSplit = s; }
            }
        }
        uint32_t mid;
        if (bestAxis < 0) {
            mid = begin + n / 2;
        } else {
            float ext = cbox.hi[bestAxis] - cbox.lo[bestAxis];
            auto it = std::partition(order.begin() + begin, order.begin() + end, [&](uint32_t p) {
                int b = std::min(NB - 1, static_cast<int>(NB * (cent[p][bestAxis] - cbox.lo[bestAxis]) / ext));
                return b < bestSplit;
            });
            mid = static_cast<uint32_t>(it - order.begin());
            if (mid == begin || mid == end) mid = begin + n / 2;
        }
        int32_t l = buildRec(boxes, cent, begin, mid);
        int32_t r = buildRec(boxes, cent, mid, end);
        nodes[idx].left = l;
        nodes[idx].right = r;
        return idx;
    }
};

// ------------------------------------------------------------------------------------------
// Scene after set-up
struct Material {
    int type = 0;
    Triplet t[3];
    float scalars[4] = {0, 0, 0, 0};
    bool useEnvTex = false;
    bool emitting() const {
        return type == ORACLE_MAT_DIFFUSE_EMITTER || type == ORACLE_MAT_POINT_EMITTER ||
               type == ORACLE_MAT_ENVIRONMENT_EMITTER;
    }
};

struct GeomInst {
    std::vector<OracleVertex> verts;
    std::vector<uint32_t> idx;
    bool isPoints = false;
    int material = 0;
    float importance = 0;
    float alpha = -1.0f;  // constant alpha node of the group; < 0: none
    Discrete1D primDist;  // only for emitters (scene.cpp:209-219, 378-393)
    AABB aabb;
    uint32_t numPrims() const { return isPoints ? uint32_t(idx.size()) : uint32_t(idx.size() / 3); }
    V3 pos(uint32_t v) const { return V3(verts[v].position[0], verts[v].position[1], verts[v].position[2]); }
    V3 nrm(uint32_t v) const { return V3(verts[v].normal[0], verts[v].normal[1], verts[v].normal[2]); }
    V3 tc0(uint32_t v) const { return V3(verts[v].tc0Direction[0], verts[v].tc0Direction[1], verts[v].tc0Direction[2]); }
};

struct Instance {
    Xfm xfm;
    std::vector<uint32_t> geomInsts;
    Discrete1D lightDist;
    float importance = 0;
    AABB childAabb, worldAabb;
    int blas = -1;
};

struct BLAS {
    std::vector<uint32_t> geomInsts;   // triangle-mesh geomInsts of the instance
    std::vector<uint64_t> prims;       // (local geomInst << 32) | prim
    BVH bvh;
};

struct Scene {
    bool spectral = false;
    std::vector<Material> materials;
    std::vector<GeomInst> geomInsts;
    std::vector<Instance> instances;
    std::vector<BLAS> blases;
    BVH tlas;
    // Environment: the light-instance list is [env, instances...] (scene.cpp:1380-1392).
    bool hasEnv = false;
    int envMaterial = -1;
    float envImportance = 0;
    float envRotation = 0;
    uint32_t envW = 0, envH = 0;
    std::vector<float> envTex;       // 4 floats per texel after half quantisation (RGB or u,v,s)
    RegConst2D envImportanceMap;
    Discrete1D lightInstDist;        // index 0 = env, i+1 = instances[i]
    V3 center;
    float worldRadius = 0, worldDiscArea = 0;
    PerspectiveCamera camera;
};

static ColorSpace toCS(int c) { return static_cast<ColorSpace>(c); }

static Scene *buildScene(const OracleScene *d, bool spectral) {
    auto *s = new Scene;
    s->spectral = spectral;
    s->materials.resize(d->numMaterials);
    for (uint32_t i = 0; i < d->numMaterials; ++i) {
        const OracleMaterial &m = d->materials[i];
        Material &o = s->materials[i];
        o.type = m.type;
        for (int k = 0; k < 4; ++k) o.scalars[k] = m.scalars[k];
        o.useEnvTex = m.useEnvTexture != 0;
        SpectrumType types[3] = {SpectrumType::Reflectance, SpectrumType::Reflectance, SpectrumType::Reflectance};
        switch (m.type) {
        case ORACLE_MAT_SPECULAR_REFLECTION:
        case ORACLE_MAT_SPECULAR_SCATTERING:
        case ORACLE_MAT_MICROFACET_SCATTERING:
            types[1] = types[2] = SpectrumType::IndexOfRefraction;
            break;
        case ORACLE_MAT_MICROFACET_REFLECTION:
            types[0] = types[1] = SpectrumType::IndexOfRefraction;
            break;
        case ORACLE_MAT_DIFFUSE_EMITTER:
        case ORACLE_MAT_POINT_EMITTER:
        case ORACLE_MAT_ENVIRONMENT_EMITTER:
            types[0] = SpectrumType::LightSource;
            break;
        default:
            break;
        }
        for (int k = 0; k < 3; ++k)
            o.t[k] = Triplet::create(types[k], toCS(m.colorSpace[k]), m.spectra[k][0], m.spectra[k][1], m.spectra[k][2]);
    }

    s->geomInsts.resize(d->numGeomInsts);
    for (uint32_t i = 0; i < d->numGeomInsts; ++i) {
        const OracleGeomInst &g = d->geomInsts[i];
        GeomInst &o = s->geomInsts[i];
        o.verts.assign(g.vertices, g.vertices + g.numVertices);
        o.idx.assign(g.indices, g.indices + g.numIndices);
        o.isPoints = g.isPoints != 0;
        o.material = g.material;
        o.alpha = g.isPoints ? -1.0f : g.alpha;
        bool emit = s->materials[g.material].emitting();
        o.importance = emit ? 1.0f : 0.0f;
        if (o.isPoints) {
            std::vector<float> imp(o.idx.size(), 1.0f);
            if (emit && !imp.empty()) o.primDist.init(imp.data(), imp.size());
        } else {
            uint32_t nt = o.numPrims();
            std::vector<float> areas(nt);
            for (uint32_t t = 0; t < nt; ++t) {
                V3 p0 = o.pos(o.idx[3 * t]), p1 = o.pos(o.idx[3 * t + 1]), p2 = o.pos(o.idx[3 * t + 2]);
                areas[t] = 0.5f * cross(p1 - p0, p2 - p0).length();
                o.aabb.unify(p0);
                o.aabb.unify(p1);
                o.aabb.unify(p2);
            }
            if (emit && nt) o.primDist.init(areas.data(), areas.size());
        }
    }

    s->instances.resize(d->numInstances);
    std::vector<float> lightImp;
    s->hasEnv = d->envMaterial >= 0;
    s->envMaterial = d->envMaterial;
    s->envImportance = s->hasEnv && s->materials[d->envMaterial].emitting() ? 1.0f : 0.0f;
    lightImp.push_back(s->envImportance > 0 ? 1.0f : 0.0f);
    AABB sceneBox;
    for (uint32_t i = 0; i < d->numInstances; ++i) {
        const OracleInstance &in = d->instances[i];
        Instance &o = s->instances[i];
        o.xfm = Xfm::fromColumnMajor4x4(in.transform);
        o.geomInsts.assign(in.geomInsts, in.geomInsts + in.numGeomInsts);
        std::vector<float> imp;
        float sumImp = 0;
        for (uint32_t g : o.geomInsts) {
            imp.push_back(s->geomInsts[g].importance);
            sumImp += s->geomInsts[g].importance;
            if (!s->geomInsts[g].isPoints) o.childAabb.unify(s->geomInsts[g].aabb);
        }
        if (!imp.empty()) o.lightDist.init(imp.data(), imp.size());
        o.importance = sumImp > 0.0f ? 1.0f : 0.0f;
        lightImp.push_back(o.importance);
        if (o.childAabb.lo.x <= o.childAabb.hi.x) {
            for (int c = 0; c < 8; ++c) {
                V3 p((c & 1) ? o.childAabb.hi.x : o.childAabb.lo.x, (c & 2) ? o.childAabb.hi.y : o.childAabb.lo.y,
                     (c & 4) ? o.childAabb.hi.z : o.childAabb.lo.z);
                o.worldAabb.unify(o.xfm.point(p));
            }
            sceneBox.unify(o.worldAabb);
        }
        // BLAS over the instance's triangle geometry
        BLAS b;
        std::vector<AABB> boxes;
        for (uint32_t li = 0; li < o.geomInsts.size(); ++li) {
            const GeomInst &g = s->geomInsts[o.geomInsts[li]];
            if (g.isPoints) continue;
            for (uint32_t t = 0; t < g.numPrims(); ++t) {
                AABB bb;
                bb.unify(g.pos(g.idx[3 * t]));
                bb.unify(g.pos(g.idx[3 * t + 1]));
                bb.unify(g.pos(g.idx[3 * t + 2]));
                boxes.push_back(bb);
                b.prims.push_back((uint64_t(li) << 32) | t);
            }
        }
        b.bvh.build(boxes);
        o.blas = static_cast<int>(s->blases.size());
        s->blases.push_back(std::move(b));
    }
    s->lightInstDist.init(lightImp.data(), lightImp.size());
    {
        std::vector<AABB> boxes;
        for (auto &in : s->instances) boxes.push_back(in.worldAabb);
        s->tlas.build(boxes);
    }
    // setup_scene.cu:82-92
    s->center = sceneBox.centroid();
    s->worldRadius = (sceneBox.hi - s->center).length();
    s->worldDiscArea = kPi * pow2(s->worldRadius);

    // Environment texture: RGBA16F input.  RGB build keeps the half RGB; the spectral build
    // converts every texel to (u, v, s) halves (image.cpp:542-565) as "Light Source" Rec709(D65).
    s->envRotation = d->envRotation;
    if (s->hasEnv && d->envTextureHalf && d->envWidth && d->envHeight) {
        s->envW = d->envWidth;
        s->envH = d->envHeight;
        size_t np = size_t(s->envW) * s->envH;
        s->envTex.resize(np * 4);
        for (size_t p = 0; p < np; ++p) {
            float RGB[3] = {halfToFloat(d->envTextureHalf[4 * p]), halfToFloat(d->envTextureHalf[4 * p + 1]),
                            halfToFloat(d->envTextureHalf[4 * p + 2])};
            if (!spectral) {
                for (int c = 0; c < 3; ++c) s->envTex[4 * p + c] = RGB[c];
                s->envTex[4 * p + 3] = 1.0f;
            } else {
                float XYZ[3];
                transformTristimulus(mat_Rec709_D65_to_XYZ, RGB, XYZ);
                float b = XYZ[0] + XYZ[1] + XYZ[2];
                float xy[2] = {b > 0.0f ? XYZ[0] / b : (1.0f / 3.0f), b > 0.0f ? XYZ[1] / b : (1.0f / 3.0f)};
                float uv[2];
                Meng::xy_to_uv(xy, uv);
                s->envTex[4 * p + 0] = halfToFloat(floatToHalf(uv[0]));
                s->envTex[4 * p + 1] = halfToFloat(floatToHalf(uv[1]));
                s->envTex[4 * p + 2] = halfToFloat(floatToHalf(b));
                s->envTex[4 * p + 3] = 1.0f;
            }
        }
        // importance map: box-shrink by 4, luminance, sin(theta) weight (shader_nodes.cpp:1992-2010)
        uint32_t mw = std::max<uint32_t>(1, s->envW / 4), mh = std::max<uint32_t>(1, s->envH / 4);
        uint32_t fx = s->envW / mw, fy = s->envH / mh;
        std::vector<float> lum(size_t(mw) * mh);
        for (uint32_t y = 0; y < mh; ++y) {
            float theta = kPi * (y + 0.5f) / mh;
            for (uint32_t x = 0; x < mw; ++x) {
                float acc[4] = {0, 0, 0, 0};
                for (uint32_t yy = 0; yy < fy; ++yy)
                    for (uint32_t xx = 0; xx < fx; ++xx)
                        for (int c = 0; c < 4; ++c)
                            acc[c] += s->envTex[4 * ((size_t(y * fy + yy)) * s->envW + x * fx + xx) + c];
                for (int c = 0; c < 4; ++c) acc[c] /= float(fx * fy);
                float Y;
                if (!spectral) {
                    Y = mat_Rec709_D65_to_XYZ[1] * acc[0] + mat_Rec709_D65_to_XYZ[4] * acc[1] + mat_Rec709_D65_to_XYZ[7] * acc[2];
                } else {
                    float xy1 = 0.022853819546830627f * acc[0] + 0.05077639329371236f * acc[1] - 0.006895157122499944f;
                    Y = xy1 * acc[2];
                }
                lum[size_t(y) * mw + x] = Y * std::sin(theta);
            }
        }
        s->envImportanceMap.init(lum.data(), mw, mh);
    } else if (s->hasEnv) {
        float one = 1.0f;
        s->envImportanceMap.init(&one, 1, 1);
    }

    const OracleCamera &c = d->camera;
    PerspectiveCamera &pc = s->camera;
    pc.position = V3(c.position[0], c.position[1], c.position[2]);
    pc.orientation = Quat{c.orientation[0], c.orientation[1], c.orientation[2], c.orientation[3]};
    pc.sensitivity = c.sensitivity;
    pc.aspect = c.aspect;
    pc.fovY = c.fovY;
    pc.lensRadius = c.lensRadius;
    pc.objPlaneDistance = c.opDistance;
    pc.setImagePlaneArea();
    return s;
}

// ------------------------------------------------------------------------------------------
// Tracing
struct HitRec {
    int32_t inst = -1, geomInst = -1, prim = -1;
    float t = 0, b1 = 0, b2 = 0;
};

// Stochastic alpha test of closest-hit rays (path_tracing.cu:14-25) with the path's generator,
// fractional visibility of any-hit rays (light_transport_common.h:205-218); set per trace by
// the caller.  Candidates are met in this BVH's traversal order, not the device's, so only the
// statistics of alpha scenes compare.
static thread_local PCG32 *tlsAlphaRng = nullptr;
static thread_local float tlsVisibility = 1.0f;

// false: the candidate is ignored
static inline bool alphaAccept(const GeomInst &g, bool anyHit) {
    if (g.alpha < 0.0f) return true;
    if (anyHit) {
        tlsVisibility *= 1 - g.alpha;
        return tlsVisibility == 0.0f;
    }
    return !(tlsAlphaRng && tlsAlphaRng->uniform() >= g.alpha);
}

static void traceInstanceBrute(const Scene &s, uint32_t ii, const RayPre &rw, float tmin, float &tmax, HitRec &h,
                               bool anyHit, bool *found) {
    const Instance &in = s.instances[ii];
    RayPre r;
    r.init(in.xfm.invPoint(rw.org), in.xfm.invVector(rw.dir));
    for (uint32_t gi : in.geomInsts) {
        const GeomInst &g = s.geomInsts[gi];
        if (g.isPoints) continue;
        for (uint32_t t = 0; t < g.numPrims(); ++t) {
            float tt, b1, b2;
            if (intersectTriangle(r, g.pos(g.idx[3 * t]), g.pos(g.idx[3 * t + 1]), g.pos(g.idx[3 * t + 2]), tmin, tmax,
                                  &tt, &b1, &b2)) {
                if (!alphaAccept(g, anyHit)) continue;
                tmax = tt;
                h.inst = static_cast<int32_t>(ii);
                h.geomInst = static_cast<int32_t>(gi);
                h.prim = static_cast<int32_t>(t);
                h.t = tt;
                h.b1 = b1;
                h.b2 = b2;
                *found = true;
                if (anyHit) return;
            }
        }
    }
}

static void traceInstanceBVH(const Scene &s, uint32_t ii, const RayPre &rw, float tmin, float &tmax, HitRec &h,
                             bool anyHit, bool *found) {
    const Instance &in = s.instances[ii];
    const BLAS &b = s.blases[in.blas];
    if (b.prims.empty()) return;
    RayPre r;
    r.init(in.xfm.invPoint(rw.org), in.xfm.invVector(rw.dir));
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BVH::Node &n = b.bvh.nodes[stack[--sp]];
        if (!hitBox(n.box, r, tmin, tmax)) continue;
        if (n.left < 0) {
            for (uint32_t k = 0; k < n.count; ++k) {
                uint64_t pr = b.prims[b.bvh.order[n.first + k]];
                uint32_t li = uint32_t(pr >> 32), t = uint32_t(pr & 0xFFFFFFFF);
                uint32_t gi = in.geomInsts[li];
                const GeomInst &g = s.geomInsts[gi];
                float tt, b1, b2;
                if (intersectTriangle(r, g.pos(g.idx[3 * t]), g.pos(g.idx[3 * t + 1]), g.pos(g.idx[3 * t + 2]), tmin,
                                      tmax, &tt, &b1, &b2)) {
                    if (!alphaAccept(g, anyHit)) continue;
                    tmax = tt;
                    h.inst = static_cast<int32_t>(ii);
                    h.geomInst = static_cast<int32_t>(gi);
                    h.prim = static_cast<int32_t>(t);
                    h.t = tt;
                    h.b1 = b1;
                    h.b2 = b2;
                    *found = true;
                    if (anyHit) return;
                }
            }
        } else {
            stack[sp++] = n.right;
            stack[sp++] = n.left;
        }
    }
}

static bool traceRay(const Scene &s, const V3 &o, const V3 &d, float tmin, float tmax, HitRec *h, bool anyHit,
                     bool useBVH) {
    RayPre rw;
    rw.init(o, d);
    bool found = false;
    HitRec hr;
    if (!useBVH) {
        for (uint32_t i = 0; i < s.instances.size(); ++i) {
            traceInstanceBrute(s, i, rw, tmin, tmax, hr, anyHit, &found);
            if (found && anyHit) break;
        }
    } else if (!s.instances.empty()) {
        int32_t stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const BVH::Node &n = s.tlas.nodes[stack[--sp]];
            if (!hitBox(n.box, rw, tmin, tmax)) continue;
            if (n.left < 0) {
                for (uint32_t k = 0; k < n.count; ++k) {
                    traceInstanceBVH(s, s.tlas.order[n.first + k], rw, tmin, tmax, hr, anyHit, &found);
                    if (found && anyHit) break;
                }
                if (found && anyHit) break;
            } else {
                stack[sp++] = n.right;
                stack[sp++] = n.left;
            }
        }
    }
    if (h) *h = hr;
    return found;
}

// ------------------------------------------------------------------------------------------
struct SurfPt {
    uint32_t instIndex = 0, geomInstIndex = 0, primIndex = 0;
    float u = 0, v = 0;
    V3 position, geometricNormal;
    Frame frame;
    float texU = 0, texV = 0;
    bool isPoint = false, atInfinity = false;
    float calcSquaredDistance(const V3 &sp) const { return atInfinity ? 1.0f : sqDistance(position, sp); }
    V3 calcDirectionFrom(const V3 &sp, float *dist2) const {
        if (atInfinity) {
            *dist2 = 1.0f;
            return normalize(position);
        }
        V3 ret = position - sp;
        *dist2 = ret.sqLength();
        return ret / std::sqrt(*dist2);
    }
    float calcCosTerm(const V3 &w) const { return isPoint ? 1.0f : absDot(w, geometricNormal); }
};

// triangle.cu:6-53 followed by renderer_common.h:609-626 (object space decode, then world)
static void calcSurfacePoint(const Scene &s, const HitRec &h, SurfPt *sp, float *hypAreaPDF) {
    const Instance &in = s.instances[h.inst];
    const GeomInst &g = s.geomInsts[h.geomInst];
    uint32_t i0 = g.idx[3 * h.prim], i1 = g.idx[3 * h.prim + 1], i2 = g.idx[3 * h.prim + 2];
    V3 p0 = g.pos(i0), p1 = g.pos(i1), p2 = g.pos(i2);
    V3 e1 = p1 - p0, e2 = p2 - p0;
    V3 gn = cross(e1, e2);
    float area = gn.length() / 2;
    gn /= 2 * area;
    float probLightPrim = g.primDist.n ? area / g.primDist.integral : 0.0f;
    *hypAreaPDF = probLightPrim / area;
    float b0 = 1 - h.b1 - h.b2, b1 = h.b1, b2 = h.b2;
    V3 position = b0 * p0 + b1 * p1 + b2 * p2;
    V3 sn = normalize(b0 * g.nrm(i0) + b1 * g.nrm(i1) + b2 * g.nrm(i2));
    V3 tc0 = b0 * g.tc0(i0) + b1 * g.tc0(i1) + b2 * g.tc0(i2);
    if (!sn.allFinite() || !tc0.allFinite()) {
        V3 bt;
        sn = gn;
        sn.makeCoordinateSystem(&tc0, &bt);
    }
    float tu = b0 * g.verts[i0].texCoord[0] + b1 * g.verts[i1].texCoord[0] + b2 * g.verts[i2].texCoord[0];
    float tv = b0 * g.verts[i0].texCoord[1] + b1 * g.verts[i1].texCoord[1] + b2 * g.verts[i2].texCoord[1];
    float dotNT = dot(sn, tc0);
    tc0 = tc0 - dotNT * sn;
    sp->instIndex = h.inst;
    sp->geomInstIndex = h.geomInst;
    sp->primIndex = h.prim;
    sp->position = in.xfm.point(position);
    sp->frame = Frame(normalize(in.xfm.vector(tc0)), normalize(in.xfm.normal(sn)));
    sp->geometricNormal = normalize(in.xfm.normal(gn));
    sp->isPoint = false;
    sp->atInfinity = false;
    sp->u = b0;
    sp->v = b1;
    sp->texU = tu;
    sp->texV = tv;
}

// Bilinear fetch with repeat wrapping in both axes (shader_nodes.cpp:1910-1919), texel centres
// at half-integers, float weights.
static void envFetch(const Scene &s, float u, float v, float out[4]) {
    float x = u * s.envW - 0.5f, y = v * s.envH - 0.5f;
    float fx0 = std::floor(x), fy0 = std::floor(y);
    float wx = x - fx0, wy = y - fy0;
    int32_t x0 = static_cast<int32_t>(fx0), y0 = static_cast<int32_t>(fy0);
    auto wrap = [](int32_t a, int32_t n) { int32_t m = a % n; return m < 0 ? m + n : m; };
    int32_t xs[2] = {wrap(x0, s.envW), wrap(x0 + 1, s.envW)};
    int32_t ys[2] = {wrap(y0, s.envH), wrap(y0 + 1, s.envH)};
    for (int c = 0; c < 4; ++c) {
        float t00 = s.envTex[4 * (size_t(ys[0]) * s.envW + xs[0]) + c];
        float t10 = s.envTex[4 * (size_t(ys[0]) * s.envW + xs[1]) + c];
        float t01 = s.envTex[4 * (size_t(ys[1]) * s.envW + xs[0]) + c];
        float t11 = s.envTex[4 * (size_t(ys[1]) * s.envW + xs[1]) + c];
        out[c] = (1 - wy) * ((1 - wx) * t00 + wx * t10) + wy * ((1 - wx) * t01 + wx * t11);
    }
}

template <int N>
static SS<N> evalTriplet(const Triplet &t, const WLS<N> &wls) { return t.template eval<N>(wls); }

template <int N>
static SS<N> envEmittance(const Scene &s, const Material &m, const SurfPt &sp, const WLS<N> &wls) {
    if (!m.useEnvTex || s.envTex.empty())
        return evalTriplet<N>(m.t[0], wls);
    float tex[4];
    envFetch(s, sp.texU, sp.texV, tex);
    if constexpr (N == 3) {
        SS<3> r;
        r[0] = tex[0]; r[1] = tex[1]; r[2] = tex[2];
        return r;
    } else {
        Meng mg = Meng::fromUVS(tex[0], tex[1], tex[2] / Meng::EER);
        return mg.evaluate(wls);
    }
}

template <int N>
static BSDF<N> setupBSDF(const Material &m, const WLS<N> &wls) {
    BSDF<N> b;
    switch (m.type) {
    case ORACLE_MAT_MATTE:
        b.kind = BSDFKind::Matte;
        b.c0 = evalTriplet<N>(m.t[0], wls);
        break;
    case ORACLE_MAT_SPECULAR_REFLECTION:
        b.kind = BSDFKind::SpecularBRDF;
        b.c0 = evalTriplet<N>(m.t[0], wls);
        b.c1 = evalTriplet<N>(m.t[1], wls);
        b.c2 = evalTriplet<N>(m.t[2], wls);
        break;
    case ORACLE_MAT_SPECULAR_SCATTERING:
        b.kind = BSDFKind::SpecularBSDF;
        b.c0 = evalTriplet<N>(m.t[0], wls);
        b.c1 = evalTriplet<N>(m.t[1], wls);
        b.c2 = evalTriplet<N>(m.t[2], wls);
        b.dispersive = !wls.single;
        break;
    case ORACLE_MAT_MICROFACET_REFLECTION: {
        b.kind = BSDFKind::MicrofacetBRDF;
        b.c0 = evalTriplet<N>(m.t[0], wls);
        b.c1 = evalTriplet<N>(m.t[1], wls);
        float alpha = pow2(m.scalars[0]);
        float aspect = std::sqrt(1.0f - 0.9f * m.scalars[1]);
        b.a0 = std::fmax(0.001f, alpha / aspect);
        b.a1 = std::fmax(0.001f, alpha * aspect);
        b.a2 = 2 * kPi * m.scalars[2];
        break;
    }
    case ORACLE_MAT_MICROFACET_SCATTERING: {
        // MicrofacetScatteringSurfaceMaterial_setupBSDF (materials.cu:1516-1536)
        b.kind = BSDFKind::MicrofacetBSDF;
        b.c0 = evalTriplet<N>(m.t[0], wls);
        b.c1 = evalTriplet<N>(m.t[1], wls);
        b.c2 = evalTriplet<N>(m.t[2], wls);
        float alpha = pow2(m.scalars[0]);
        float aspect = std::sqrt(1.0f - 0.9f * m.scalars[1]);
        b.a0 = std::fmax(0.001f, alpha / aspect);
        b.a1 = std::fmax(0.001f, alpha * aspect);
        b.a2 = 2 * kPi * m.scalars[2];
        break;
    }
    case ORACLE_MAT_UE4: {
        b.kind = BSDFKind::DiffuseAndSpecular;
        SS<N> base = evalTriplet<N>(m.t[0], wls);
        float roughness = std::fmax(0.01f, m.scalars[1]);
        float metallic = m.scalars[2];
        const float specular = 0.5f;
        b.c0 = base * (1 - metallic);
        b.c1 = lerpS(0.08f * specular * SS<N>(1.0f), base, metallic);
        b.a0 = roughness;
        break;
    }
    default:
        b.kind = BSDFKind::Null;
        break;
    }
    return b;
}

template <int N>
static EDF<N> setupEDF(const Scene &s, const Material &m, const SurfPt &sp, const WLS<N> &wls) {
    EDF<N> e;
    switch (m.type) {
    case ORACLE_MAT_DIFFUSE_EMITTER:
        e.kind = EDFKind::Diffuse;
        e.emittance = evalTriplet<N>(m.t[0], wls) * m.scalars[0];
        break;
    case ORACLE_MAT_POINT_EMITTER:
        e.kind = EDFKind::Point;
        e.emittance = evalTriplet<N>(m.t[0], wls) * m.scalars[0];
        break;
    case ORACLE_MAT_ENVIRONMENT_EMITTER:
        e.kind = EDFKind::Environment;
        e.emittance = envEmittance<N>(s, m, sp, wls) * m.scalars[0];
        e.worldRadius = s.worldRadius;
        e.worldDiscArea = s.worldDiscArea;
        break;
    default:
        break;
    }
    return e;
}

// Light position sample (triangle.cu:113-177, point.cu:38-80, infinite_sphere.cu:60-114)
struct LightSample {
    SurfPt sp;
    float areaPDF = 0;
    uint32_t posType = 0;
    int material = 0;
};

static void sampleLight(const Scene &s, float uLight, float uP0, float uP1, LightSample *ls, float *lightProb,
                        bool *valid) {
    float uInst, instProb;
    uint32_t li = s.lightInstDist.sample(uLight, &instProb, &uInst);
    *valid = true;
    if (li == 0) {
        // environment: single geomInst with importance 1
        float geomInstProb = 1.0f;
        *lightProb = instProb * geomInstProb;
        float u, v, uvPDF;
        s.envImportanceMap.sample(uP0, uP1, &u, &v, &uvPDF);
        float phi = 2 * kPi * u;
        float theta = kPi * v;
        float posPhi = phi - s.envRotation;
        posPhi = posPhi - std::floor(posPhi / (2 * kPi)) * 2 * kPi;
        V3 dir = V3::fromPolarYUp(posPhi, theta);
        float sinPhi = std::sin(posPhi), cosPhi = std::cos(posPhi);
        V3 tc0 = normalize(V3(-cosPhi, 0.0f, -sinPhi));
        V3 gn = -dir;
        SurfPt &sp = ls->sp;
        sp.position = dir;
        sp.frame.x = tc0;
        sp.frame.z = gn;
        sp.frame.y = cross(gn, tc0);
        sp.isPoint = false;
        sp.atInfinity = true;
        sp.geometricNormal = gn;
        sp.u = posPhi;
        sp.v = theta;
        sp.texU = phi / (2 * kPi);
        sp.texV = theta / kPi;
        ls->areaPDF = uvPDF / (2 * kPi * kPi * std::sin(theta));
        ls->posType = DT::Emission | DT::LowFreq;
        ls->material = s.envMaterial;
        return;
    }
    const Instance &in = s.instances[li - 1];
    float uPrim, geomInstProb;
    uint32_t lgi = in.lightDist.sample(uInst, &geomInstProb, &uPrim);
    uint32_t gi = in.geomInsts[lgi];
    const GeomInst &g = s.geomInsts[gi];
    *lightProb = instProb * geomInstProb;
    float primProb;
    uint32_t prim = g.primDist.sample(uPrim, &primProb);
    SurfPt &sp = ls->sp;
    sp.instIndex = li - 1;
    sp.geomInstIndex = gi;
    sp.primIndex = prim;
    ls->material = g.material;
    if (g.isPoints) {
        const float area = 4 * kPi;
        ls->areaPDF = primProb / area;
        ls->posType = DT::Emission | DT::Delta0D;
        uint32_t pi = g.idx[prim];
        sp.position = in.xfm.point(g.pos(pi));
        V3 sn = normalize(in.xfm.normal(g.nrm(pi)));
        V3 tc0 = normalize(in.xfm.vector(g.tc0(pi)));
        sp.frame = Frame(tc0, sn);
        sp.isPoint = true;
        sp.atInfinity = false;
        sp.geometricNormal = sn;
        sp.u = sp.v = 0;
        sp.texU = sp.texV = 0;
        return;
    }
    uint32_t i0 = g.idx[3 * prim], i1 = g.idx[3 * prim + 1], i2 = g.idx[3 * prim + 2];
    V3 e1 = in.xfm.vector(g.pos(i1) - g.pos(i0));
    V3 e2 = in.xfm.vector(g.pos(i2) - g.pos(i0));
    V3 gn = cross(e1, e2);
    float area = gn.length() / 2;
    gn /= 2 * area;
    ls->areaPDF = primProb / area;
    ls->posType = DT::Emission | DT::LowFreq;
    float b0, b1, b2;
    uniformSampleTriangle(uP0, uP1, &b0, &b1);
    b2 = 1.0f - b0 - b1;
    V3 position = b0 * g.pos(i0) + b1 * g.pos(i1) + b2 * g.pos(i2);
    V3 sn = b0 * g.nrm(i0) + b1 * g.nrm(i1) + b2 * g.nrm(i2);
    V3 tc0 = b0 * g.tc0(i0) + b1 * g.tc0(i1) + b2 * g.tc0(i2);
    float tu = b0 * g.verts[i0].texCoord[0] + b1 * g.verts[i1].texCoord[0] + b2 * g.verts[i2].texCoord[0];
    float tv = b0 * g.verts[i0].texCoord[1] + b1 * g.verts[i1].texCoord[1] + b2 * g.verts[i2].texCoord[1];
    position = in.xfm.point(position);
    sn = normalize(in.xfm.normal(sn));
    tc0 = in.xfm.vector(tc0);
    float dotNT = dot(sn, tc0);
    tc0 = normalize(tc0 - dotNT * sn);
    sp.position = position;
    sp.frame = Frame(tc0, sn);
    sp.isPoint = false;
    sp.atInfinity = false;
    sp.geometricNormal = gn;
    sp.u = b0;
    sp.v = b1;
    sp.texU = tu;
    sp.texV = tv;
}

// light_transport_common.h:147-177
static bool testVisibility(const Scene &s, const SurfPt &shading, const SurfPt &light, V3 *shadowDir, float *dist2,
                           uint64_t *shadowRays, float *fractionalVisibility = nullptr) {
    *shadowDir = light.calcDirectionFrom(shading.position, dist2);
    bool front = dot(shading.geometricNormal, *shadowDir) > 0;
    V3 sp = offsetRayOrigin(shading.position, front ? shading.geometricNormal : -shading.geometricNormal);
    front = -dot(light.geometricNormal, *shadowDir) > 0;
    V3 lp = offsetRayOrigin(light.position, front ? light.geometricNormal : -light.geometricNormal);
    float tmax = FLT_MAX;
    if (!light.atInfinity)
        tmax = distance(lp, sp) * 0.9999f;
    ++*shadowRays;
    tlsVisibility = 1.0f;
    const bool visible = !traceRay(s, sp, *shadowDir, 0.0f, tmax, nullptr, true, true);
    if (fractionalVisibility) *fractionalVisibility = visible ? tlsVisibility : 0.0f;
    return visible;
}

// ------------------------------------------------------------------------------------------
// The path tracer (path_tracing.cu:30-337), one pixel sample.
// What the path's first hit gives the denoiser's guide buffers (path_tracing.cu:95-108,
// 154-157, 272-275): the BSDF's base colour times the D65 white point over the wavelength
// pdf, and the shading normal; zero for a miss.
template <int N>
struct FirstHit {
    SS<N> albedo;
    V3 normal;
};

template <int N>
static SS<N> tracePath(const Scene &s, PCG32 &rng, uint32_t px, uint32_t py, uint32_t W, uint32_t H,
                       uint32_t maxPathLength, WLS<N> *wlsOut, OracleRenderStats &st,
                       FirstHit<N> *firstHit = nullptr) {
    float pX = px + rng.uniform();
    float pY = py + rng.uniform();
    float selectWLPDF;
    float wlOff = rng.uniform(), wlSel = rng.uniform();
    WLS<N> wls = WLS<N>::createWithEqualOffsets(wlOff, wlSel, &selectWLPDF);

    // cameras.cu:43-76
    const PerspectiveCamera &cam = s.camera;
    M3 rot = cam.orientation.toMatrix3x3();
    float lx, ly;
    float ul0 = rng.uniform(), ul1 = rng.uniform();
    concentricSampleDisk(ul0, ul1, &lx, &ly);
    V3 orgLocal(cam.lensRadius * lx, cam.lensRadius * ly, 0.0f);
    V3 camN = normalize(rot * V3(0, 0, 1));
    Frame camFrame;
    camFrame.z = camN;
    camFrame.x = normalize(rot * V3(1, 0, 0));
    camFrame.y = cross(camFrame.z, camFrame.x);
    V3 lensPos = rot * orgLocal + cam.position;
    float lensAreaPDF = cam.lensRadius > 0.0f ? 1.0f / (kPi * cam.lensRadius * cam.lensRadius) : 1.0f;
    // PerspectiveCameraIDF (cameras.cu:124-157, 213-222)
    float xOnLens = cam.lensRadius * lx, yOnLens = cam.lensRadius * ly;
    SS<N> We0(cam.sensitivity);
    float uD0 = pX / W, uD1 = pY / H;
    V3 pFocus(cam.opWidth * (0.5f - uD0), cam.opHeight * (0.5f - uD1), cam.objPlaneDistance);
    V3 dirLocal = normalize(pFocus - V3(xOnLens, yOnLens, 0.0f));
    float dirPDF = pow2(cam.imgPlaneDistance) / (pow3(dirLocal.z) * cam.imgPlaneArea);
    uint32_t sampledType = DT::Acquisition | DT::LowFreq;
    SS<N> We1(1.0f);
    dirPDF *= W * H;

    V3 rayOrg = lensPos;
    V3 rayDir = camFrame.fromLocal(dirLocal);
    SS<N> alpha = (We0 * We1) * (absDot(rayDir, camN) / (lensAreaPDF * dirPDF * selectWLPDF));
    float initImportance = alpha.importance(wls.sel);
    float prevDirPDF = dirPDF;
    uint32_t prevSampledType = sampledType;
    uint32_t pathLength = 0;
    bool maxLengthTerminate = false;
    SS<N> contribution(0.0f);

    while (true) {
        ++pathLength;
        if (pathLength >= maxPathLength)
            maxLengthTerminate = true;
        HitRec hit;
        ++st.closestRays;
        tlsAlphaRng = &rng;
        bool found = traceRay(s, rayOrg, rayDir, 0.0f, FLT_MAX, &hit, false, true);
        tlsAlphaRng = nullptr;
        if (firstHit && pathLength == 1 && !found) {
            firstHit->albedo = SS<N>(0.0f);
            firstHit->normal = V3(0.0f, 0.0f, 0.0f);
        }
        if (!found) {
            // miss program (path_tracing.cu:266-337)
            if (!s.hasEnv || s.envImportance == 0)
                break;
            V3 direction = rayDir;
            float phi, theta;
            direction.toPolarYUp(&theta, &phi);
            float sinPhi = std::sin(phi), cosPhi = std::cos(phi);
            V3 tc0 = normalize(V3(-cosPhi, 0.0f, -sinPhi));
            SurfPt sp;
            sp.frame.x = tc0;
            sp.frame.z = -direction;
            sp.frame.y = cross(sp.frame.z, sp.frame.x);
            sp.position = direction;
            sp.isPoint = false;
            sp.atInfinity = true;
            sp.geometricNormal = -direction;
            sp.u = phi;
            sp.v = theta;
            phi += s.envRotation;
            phi = phi - std::floor(phi / (2 * kPi)) * 2 * kPi;
            sp.texU = phi / (2 * kPi);
            sp.texV = theta / kPi;
            float uvPDF = s.envImportanceMap.evaluatePDF(phi / (2 * kPi), theta / kPi);
            float hypAreaPDF = uvPDF / (2 * kPi * kPi * std::sin(theta));
            const Material &m = s.materials[s.envMaterial];
            EDF<N> edf = setupEDF<N>(s, m, sp, wls);
            V3 dirOutLocal = sp.frame.toLocal(-direction);
            SS<N> spEmittance = edf.evaluateEmittance();
            if (spEmittance.hasNonZero()) {
                SS<N> Le = spEmittance * edf.evaluate(dirOutLocal);
                float MISWeight = 1.0f;
                if (!DT::isDelta(prevSampledType) && pathLength > 1) {
                    float instProb = 1.0f / s.lightInstDist.integral;  // env inst importance 1
                    float geomInstProb = 1.0f;
                    float bsdfPDF = prevDirPDF;
                    float dist2 = sp.calcSquaredDistance(rayOrg);
                    float lightPDF = instProb * geomInstProb * hypAreaPDF * dist2 / std::fabs(dirOutLocal.z);
                    MISWeight = (bsdfPDF * bsdfPDF) / (lightPDF * lightPDF + bsdfPDF * bsdfPDF);
                }
                contribution += alpha * Le * MISWeight;
            }
            break;
        }

        // closest-hit program (path_tracing.cu:133-259)
        SurfPt sp;
        float hypAreaPDF;
        calcSurfacePoint(s, hit, &sp, &hypAreaPDF);
        const GeomInst &g = s.geomInsts[hit.geomInst];
        const Material &mat = s.materials[g.material];
        BSDF<N> bsdf = setupBSDF<N>(mat, wls);
        EDF<N> edf = setupEDF<N>(s, mat, sp, wls);
        V3 dirOutLocal = sp.frame.toLocal(-rayDir);
        if (firstHit && pathLength == 1) {
            const Triplet white = Triplet::create(SpectrumType::LightSource, ColorSpace::Rec709_D65, 1, 1, 1);
            firstHit->albedo = bsdf.getBaseColor() * white.eval<N>(wls) / selectWLPDF;
            firstHit->normal = sp.frame.z;
        }

        SS<N> spEmittance = edf.evaluateEmittance();
        if (spEmittance.hasNonZero()) {
            SS<N> Le = spEmittance * edf.evaluate(dirOutLocal);
            float MISWeight = 1.0f;
            if (!DT::isDelta(prevSampledType) && pathLength > 1) {
                const Instance &in = s.instances[hit.inst];
                float instProb = in.lightDist.integral / s.lightInstDist.integral;
                float geomInstProb = g.importance / in.lightDist.integral;
                float bsdfPDF = prevDirPDF;
                float dist2 = sp.calcSquaredDistance(rayOrg);
                float lightPDF = instProb * geomInstProb * hypAreaPDF * dist2 / std::fabs(dirOutLocal.z);
                MISWeight = (bsdfPDF * bsdfPDF) / (lightPDF * lightPDF + bsdfPDF * bsdfPDF);
            }
            contribution += alpha * Le * MISWeight;
        }
        if (maxLengthTerminate)
            break;

        // Russian roulette
        float continueProb = std::fmin(alpha.importance(wls.sel) / initImportance, 1.0f);
        if (rng.uniform() >= continueProb)
            break;
        alpha /= continueProb;

        V3 gnLocal = sp.frame.toLocal(sp.geometricNormal);
        BSDFQuery<N> q{dirOutLocal, gnLocal, DT::All, TransportMode::Radiance, wls.sel};

        // next event estimation
        if (bsdf.hasNonDelta()) {
            float uLight = rng.uniform();
            float uP0 = rng.uniform(), uP1 = rng.uniform();
            // a scene without emitters has nothing to sample (the draws are still made)
            if (s.lightInstDist.integral > 0.0f) {
            LightSample ls;
            float lightProb;
            bool valid;
            sampleLight(s, uLight, uP0, uP1, &ls, &lightProb, &valid);
            const Material &lm = s.materials[ls.material];
            EDF<N> ledf = setupEDF<N>(s, lm, ls.sp, wls);
            SS<N> M = ledf.evaluateEmittance();
            V3 shadowDir;
            float dist2;
            float fracVis = 1.0f;
            if (M.hasNonZero() && testVisibility(s, sp, ls.sp, &shadowDir, &dist2, &st.shadowRays, &fracVis)) {
                float rec = 1.0f / dist2;
                V3 sdl = ls.sp.frame.toLocal(-shadowDir);
                V3 sdsn = sp.frame.toLocal(shadowDir);
                SS<N> Le = M * ledf.evaluate(sdl);
                float lightPDF = lightProb * ls.areaPDF;
                SS<N> fs = bsdf.evaluate(q, sdsn);
                float cosLight = ls.sp.calcCosTerm(-shadowDir);
                float bsdfPDF = bsdf.evaluatePDF(q, sdsn) * cosLight * rec;
                float MISWeight = 1.0f;
                if (!DT::isDelta(ls.posType) && !std::isinf(lightPDF))
                    MISWeight = (lightPDF * lightPDF) / (lightPDF * lightPDF + bsdfPDF * bsdfPDF);
                float G = fracVis * absDot(sdsn, gnLocal) * cosLight * rec;
                float scalarCoeff = G * MISWeight / lightPDF;
                contribution += alpha * Le * fs * scalarCoeff;
            }
            }
        }

        float uC = rng.uniform(), uD0b = rng.uniform(), uD1b = rng.uniform();
        BSDFResult<N> res;
        SS<N> fs = bsdf.sample(q, uC, uD0b, uD1b, &res);
        if (fs == SS<N>(0.0f) || res.dirPDF == 0.0f)
            break;
        if (DT::isDispersive(res.sampledType) && !wls.single) {
            res.dirPDF /= N;
            wls.single = true;
        }
        float cosFactor = dot(res.dirLocal, gnLocal);
        alpha *= fs * (std::fabs(cosFactor) / res.dirPDF);
        V3 dirIn = sp.frame.fromLocal(res.dirLocal);
        rayOrg = offsetRayOrigin(sp.position, cosFactor > 0.0f ? sp.geometricNormal : -sp.geometricNormal);
        rayDir = dirIn;
        prevDirPDF = res.dirPDF;
        prevSampledType = res.sampledType;
    }
    *wlsOut = wls;
    return contribution;
}

// ==========================================================================================
// LVC-BPT (lvc_bpt.cu:144-971): light subpaths are traced first and every vertex is stored in
// a light-vertex cache; each eye vertex, including the lens vertex, is connected to one cache
// vertex chosen uniformly, with the partial MIS bookkeeping of the reference (Davidovic et
// al. 2014).  Lens connections are splatted to the pixel they project to.
template <int N>
struct LightVertex {
    int32_t kind = 0;                 // 0 triangle, 1 point, 2 environment
    uint32_t inst = 0, geomInst = 0, prim = 0;
    float u = 0, v = 0;
    float probDensity = 0, prevProbDensity = 0;
    float secondPrevPartialDenomMisWeight = 0, secondPrevProbRatioToFirst = 0;
    float backwardConversionFactor = 0;
    SS<N> flux;
    V3 dirInLocal;
    bool deltaSampled = false, prevDeltaSampled = false, wlSelected = false;
    uint32_t pathLength = 0;
};

template <int N>
static LightVertex<N> makeLightVertex(const SurfPt &sp, const SS<N> &flux, const V3 &dirInLocal, float bcf,
                                      float probDensity, float prevProbDensity, float spdmw, float sprtf,
                                      bool delta, bool prevDelta, bool wlSelected, uint32_t pathLength) {
    LightVertex<N> lv;
    lv.kind = sp.atInfinity ? 2 : (sp.isPoint ? 1 : 0);
    lv.inst = sp.instIndex;
    lv.geomInst = sp.geomInstIndex;
    lv.prim = sp.primIndex;
    lv.u = sp.u;
    lv.v = sp.v;
    lv.probDensity = probDensity;
    lv.prevProbDensity = prevProbDensity;
    lv.secondPrevPartialDenomMisWeight = spdmw;
    lv.secondPrevProbRatioToFirst = sprtf;
    lv.backwardConversionFactor = bcf;
    lv.flux = flux;
    lv.dirInLocal = dirInLocal;
    lv.deltaSampled = delta;
    lv.prevDeltaSampled = prevDelta;
    lv.wlSelected = wlSelected;
    lv.pathLength = pathLength;
    return lv;
}

// decodeHitPointForTriangle (triangle.cu:55-109), decodeHitPointForPoint (point.cu:6-35),
// decodeHitPointForInfiniteSphere (infinite_sphere.cu:17-56)
template <int N>
static int decodeLightVertex(const Scene &s, const LightVertex<N> &lv, SurfPt *sp) {
    if (lv.kind == 2) {
        float posPhi = lv.u, theta = lv.v;
        float phi = posPhi + s.envRotation;
        phi = phi - std::floor(phi / (2 * kPi)) * 2 * kPi;
        V3 direction = V3::fromPolarYUp(posPhi, theta);
        float sinPhi = std::sin(posPhi), cosPhi = std::cos(posPhi);
        V3 tc0 = normalize(V3(-cosPhi, 0.0f, -sinPhi));
        V3 gn = -direction;
        sp->frame.x = tc0;
        sp->frame.z = gn;
        sp->frame.y = cross(sp->frame.z, sp->frame.x);
        sp->position = direction;
        sp->geometricNormal = gn;
        sp->isPoint = false;
        sp->atInfinity = true;
        sp->u = posPhi;
        sp->v = theta;
        sp->texU = phi / (2 * kPi);
        sp->texV = theta / kPi;
        return s.envMaterial;
    }
    const Instance &in = s.instances[lv.inst];
    const GeomInst &g = s.geomInsts[lv.geomInst];
    sp->instIndex = lv.inst;
    sp->geomInstIndex = lv.geomInst;
    sp->primIndex = lv.prim;
    sp->atInfinity = false;
    if (lv.kind == 1) {
        uint32_t pi = g.idx[lv.prim];
        sp->position = in.xfm.point(g.pos(pi));
        V3 sn = normalize(in.xfm.normal(g.nrm(pi)));
        V3 tc0 = normalize(in.xfm.vector(g.tc0(pi)));
        sp->frame = Frame(tc0, sn);
        sp->isPoint = true;
        sp->geometricNormal = sn;
        sp->u = lv.u;
        sp->v = lv.v;
        sp->texU = sp->texV = 0;
        return g.material;
    }
    uint32_t i0 = g.idx[3 * lv.prim], i1 = g.idx[3 * lv.prim + 1], i2 = g.idx[3 * lv.prim + 2];
    V3 e1 = in.xfm.vector(g.pos(i1) - g.pos(i0));
    V3 e2 = in.xfm.vector(g.pos(i2) - g.pos(i0));
    V3 gn = cross(e1, e2);
    float area = gn.length() / 2;
    gn /= 2 * area;
    float b0 = lv.u, b1 = lv.v, b2 = 1.0f - lv.u - lv.v;
    V3 position = b0 * g.pos(i0) + b1 * g.pos(i1) + b2 * g.pos(i2);
    V3 sn = b0 * g.nrm(i0) + b1 * g.nrm(i1) + b2 * g.nrm(i2);
    V3 tc0 = b0 * g.tc0(i0) + b1 * g.tc0(i1) + b2 * g.tc0(i2);
    sp->texU = b0 * g.verts[i0].texCoord[0] + b1 * g.verts[i1].texCoord[0] + b2 * g.verts[i2].texCoord[0];
    sp->texV = b0 * g.verts[i0].texCoord[1] + b1 * g.verts[i1].texCoord[1] + b2 * g.verts[i2].texCoord[1];
    position = in.xfm.point(position);
    sn = normalize(in.xfm.normal(sn));
    tc0 = in.xfm.vector(tc0);
    if (!sn.allFinite() || !tc0.allFinite()) {
        V3 bt;
        sn = gn;
        sn.makeCoordinateSystem(&tc0, &bt);
    }
    float dotNT = dot(sn, tc0);
    tc0 = normalize(tc0 - dotNT * sn);
    sp->position = position;
    sp->frame = Frame(tc0, sn);
    sp->isPoint = false;
    sp->geometricNormal = gn;
    sp->u = b0;
    sp->v = b1;
    return g.material;
}

// computeRRProbability (lvc_bpt.cu:137-141)
template <int N>
static float rrProbability(const SS<N> &fs, const V3 &dirLocal, float dirDensity, const V3 &gnLocal, uint32_t sel) {
    SS<N> localThroughput = fs * absDot(dirLocal, gnLocal) / dirDensity;
    return std::fmin(localThroughput.importance(sel), 1.0f);
}

// BSDF of a cached light vertex: the material's BSDF, or the emitter's EDF used as a BSDF at
// the light source itself (the fromEDF constructor of renderer_common.h:192-216), with the
// importance-transport shading-normal correction of renderer_common.h:224-264.
template <int N>
struct VertexBSDF {
    bool fromEDF = false;
    BSDF<N> bsdf;
    EDF<N> edf;
    bool hasNonDelta() const { return fromEDF ? edf.matches(DT::WholeSphere | DT::NonDelta) : bsdf.hasNonDelta(); }
    SS<N> evaluate(const BSDFQuery<N> &q, const V3 &dir, SS<N> *rev) const {
        if (!fromEDF) return bsdf.evaluate(q, dir, rev);
        float sn = std::fabs(q.dirLocal.z / dot(q.dirLocal, q.geometricNormalLocal));
        *rev = SS<N>(0.0f);
        if (!std::isfinite(sn)) return SS<N>(0.0f);
        return edf.evaluate(dir) * sn;
    }
    float evaluatePDF(const BSDFQuery<N> &q, const V3 &dir, float *rev) const {
        if (!fromEDF) return bsdf.evaluatePDF(q, dir, rev);
        *rev = 0.0f;
        if (!edf.matches(q.dirTypeFilter)) return 0.0f;
        return edf.evaluatePDF(dir);
    }
};

// PerspectiveCameraIDF (cameras.cu:124-211)
struct LensIDF {
    float sensitivity, xOnLens, yOnLens, opWidth, opHeight, imgPlaneDistance, objPlaneDistance, imgPlaneArea;
    bool project(const V3 &d, float *u0, float *u1) const {
        if (d.z <= 0.0f) return false;
        float objPlaneZ = objPlaneDistance / d.z;
        V3 pFocus = d * objPlaneZ + V3(xOnLens, yOnLens, 0.0f);
        *u0 = -pFocus.x / opWidth + 0.5f;
        *u1 = -pFocus.y / opHeight + 0.5f;
        return !(*u0 < 0.0f || *u0 >= 1.0f || *u1 < 0.0f || *u1 >= 1.0f);
    }
    float directionalImportance(const V3 &d) const {
        float a, b;
        return project(d, &a, &b) ? 1.0f : 0.0f;
    }
    float pdf(const V3 &d) const {
        float a, b;
        if (!project(d, &a, &b)) return 0.0f;
        return pow2(imgPlaneDistance) / (pow3(d.z) * imgPlaneArea);
    }
};

struct BPTSettings {
    uint32_t numLightPaths;
    uint32_t maxPathLength;
};

template <int N>
static void traceLightPath(const Scene &s, const WLS<N> &wlsCommon, const BPTSettings &bs, PCG32 &rng,
                           std::vector<LightVertex<N>> *cache, OracleRenderStats &st) {
    if (s.lightInstDist.integral <= 0.0f) return;
    float uLight = rng.uniform();
    float uP0 = rng.uniform(), uP1 = rng.uniform();
    LightSample ls;
    float lightProb;
    bool valid;
    sampleLight(s, uLight, uP0, uP1, &ls, &lightProb, &valid);
    const Material &lmat = s.materials[ls.material];
    EDF<N> edf = setupEDF<N>(s, lmat, ls.sp, wlsCommon);
    float probDensity0 = bs.numLightPaths * lightProb * ls.areaPDF;
    SS<N> Le0 = edf.evaluateEmittance();
    SS<N> alpha = Le0 / probDensity0;
    cache->push_back(makeLightVertex<N>(ls.sp, alpha, V3(0, 0, 1), 0.0f, probDensity0, 1.0f, 0.0f, 1.0f,
                                        DT::isDelta(ls.posType), false, false, 0));
    float uC = rng.uniform(), uD0 = rng.uniform(), uD1 = rng.uniform();
    (void)uC;
    V3 dirLocal;
    float dirPDF;
    uint32_t sampledType;
    SS<N> Le1 = edf.sample(uD0, uD1, &dirLocal, &dirPDF, &sampledType);
    V3 rayOrg = offsetRayOrigin(ls.sp.position, ls.sp.geometricNormal);
    if (ls.sp.atInfinity) {
        rayOrg = s.center + 1.1f * s.worldRadius * ls.sp.position + ls.sp.frame.x * dirLocal.x +
                 ls.sp.frame.y * dirLocal.y;
        dirLocal.x = 0;
        dirLocal.y = 0;
    }
    V3 rayDir = ls.sp.frame.fromLocal(dirLocal);
    float cosTerm = ls.sp.calcCosTerm(rayDir);
    alpha *= Le1 * (cosTerm / dirPDF);

    // payload state
    float roDirPDF = dirPDF, roCosTerm = cosTerm, roPrevRevAreaPDF = 1;
    bool roSecondPrevDelta = false, roPrevDelta = DT::isDelta(ls.posType);
    uint32_t roSampledType = sampledType;
    float probDensity = probDensity0, prevProbDensity = 1.0f;
    float secondPrevPartialDenomMisWeight = 0, secondPrevProbRatioToFirst = 1;
    bool single = false;
    bool originIsPoint = ls.sp.isPoint, originIsInfinity = ls.sp.atInfinity;
    uint32_t pathLength = 0;
    while (true) {
        bool terminate = true;
        ++pathLength;
        HitRec hit;
        ++st.closestRays;
        bool found = traceRay(s, rayOrg, rayDir, 0.0f, FLT_MAX, &hit, false, true);
        V3 nextOrg, nextDir;
        float woDirPDF = 0, woCosTerm = 0, woRevAreaPDF = 0;
        uint32_t woSampledType = 0;
        if (found) {
            // closest-hit program for light paths (lvc_bpt.cu:246-334)
            WLS<N> wls = wlsCommon;
            if (single) wls.single = true;
            SurfPt sp;
            float hypAreaPDF;
            calcSurfacePoint(s, hit, &sp, &hypAreaPDF);
            const GeomInst &g = s.geomInsts[hit.geomInst];
            BSDF<N> bsdf = setupBSDF<N>(s.materials[g.material], wls);
            V3 dirInLocal = sp.frame.toLocal(-rayDir);
            V3 gnLocal = sp.frame.toLocal(sp.geometricNormal);
            BSDFQuery<N> q{dirInLocal, gnLocal, DT::All, TransportMode::Importance, wls.sel};
            bool secondLastValid = (!roPrevDelta && !roSecondPrevDelta) && pathLength > 2;
            float probRatio = roPrevRevAreaPDF / prevProbDensity;
            secondPrevPartialDenomMisWeight = pow2(probRatio) * (secondPrevPartialDenomMisWeight + (secondLastValid ? 1 : 0));
            secondPrevProbRatioToFirst *= probRatio;
            prevProbDensity = probDensity;
            float lastDist2 = sqDistance(rayOrg, sp.position);
            if (originIsInfinity) lastDist2 = 1;
            probDensity = roDirPDF * absDot(dirInLocal, gnLocal) / lastDist2;
            cache->push_back(makeLightVertex<N>(sp, alpha, dirInLocal, roCosTerm / lastDist2, probDensity,
                                                prevProbDensity, secondPrevPartialDenomMisWeight,
                                                secondPrevProbRatioToFirst, DT::isDelta(roSampledType), roPrevDelta,
                                                single, pathLength));
            float sC = rng.uniform(), s0 = rng.uniform(), s1 = rng.uniform();
            BSDFResult<N> res;
            RevResult<N> rev;
            SS<N> fs = bsdf.sample(q, sC, s0, s1, &res, &rev);
            if (!(fs == SS<N>(0.0f) || res.dirPDF == 0.0f)) {
                if (DT::isDispersive(res.sampledType) && !single) single = true;
                float cosT = dot(res.dirLocal, gnLocal);
                SS<N> localThroughput = fs * (std::fabs(cosT) / res.dirPDF);
                alpha *= localThroughput;
                float continueProb = std::fmin(localThroughput.importance(wlsCommon.sel), 1.0f);
                if (rng.uniform() < continueProb) {
                    alpha /= continueProb;
                    terminate = false;
                    V3 dirOut = sp.frame.fromLocal(res.dirLocal);
                    nextOrg = offsetRayOrigin(sp.position, cosT > 0.0f ? sp.geometricNormal : -sp.geometricNormal);
                    nextDir = dirOut;
                    woDirPDF = res.dirPDF;
                    woCosTerm = std::fabs(cosT);
                    woRevAreaPDF = rev.dirPDF * roCosTerm / lastDist2;
                    woSampledType = res.sampledType;
                    woDirPDF *= continueProb;
                    woRevAreaPDF *= rrProbability<N>(rev.value, q.dirLocal, rev.dirPDF, gnLocal, wlsCommon.sel);
                    if (originIsPoint) woRevAreaPDF = 0;
                    originIsPoint = false;
                    originIsInfinity = false;
                }
            }
        }
        if (pathLength >= bs.maxPathLength) terminate = true;
        if (terminate) break;
        rayOrg = nextOrg;
        rayDir = nextDir;
        roDirPDF = woDirPDF;
        roCosTerm = woCosTerm;
        roPrevRevAreaPDF = woRevAreaPDF;
        roSecondPrevDelta = roPrevDelta;
        roPrevDelta = DT::isDelta(roSampledType);
        roSampledType = woSampledType;
    }
}

// One connection between an eye vertex E and a randomly chosen cache vertex
// (lvc_bpt.cu:383-478 for the lens vertex, 662-803 for surface vertices).
template <int N>
struct EyeConnection {
    bool valid = false;
    SS<N> value;      // unweighted contribution divided by nothing yet (flux * conTerm)
    float partialL = 0, partialE = 0;
    V3 conRayDir;
};

template <int N>
static SS<N> eyePathContribution(const Scene &s, const WLS<N> &wlsCommon, float wavelengthProbability,
                                 const BPTSettings &bs, const std::vector<LightVertex<N>> &cache, PCG32 &rng,
                                 uint32_t px, uint32_t py, uint32_t W, uint32_t H, std::vector<float> *splat,
                                 OracleRenderStats &st) {
    constexpr int C = N == 4 ? 16 : 3;
    const uint32_t numLightVertices = static_cast<uint32_t>(cache.size());
    float pX = px + rng.uniform();
    float pY = py + rng.uniform();
    const PerspectiveCamera &cam = s.camera;
    M3 rot = cam.orientation.toMatrix3x3();
    float lx, ly;
    float ul0 = rng.uniform(), ul1 = rng.uniform();
    concentricSampleDisk(ul0, ul1, &lx, &ly);
    V3 orgLocal(cam.lensRadius * lx, cam.lensRadius * ly, 0.0f);
    SurfPt spE0;
    spE0.frame.z = normalize(rot * V3(0, 0, 1));
    spE0.frame.x = normalize(rot * V3(1, 0, 0));
    spE0.frame.y = cross(spE0.frame.z, spE0.frame.x);
    spE0.position = rot * orgLocal + cam.position;
    spE0.geometricNormal = spE0.frame.z;
    spE0.isPoint = false;
    spE0.atInfinity = false;
    spE0.u = lx;
    spE0.v = ly;
    float lensAreaPDF = cam.lensRadius > 0.0f ? 1.0f / (kPi * cam.lensRadius * cam.lensRadius) : 1.0f;
    uint32_t lensPosType = cam.lensRadius > 0.0f ? uint32_t(DT::LowFreq) : uint32_t(DT::Delta0D);
    LensIDF idf{cam.sensitivity, cam.lensRadius * lx, cam.lensRadius * ly, cam.opWidth, cam.opHeight,
                cam.imgPlaneDistance, cam.objPlaneDistance, cam.imgPlaneArea};
    SS<N> We0(cam.sensitivity);
    SS<N> alpha = We0 / (lensAreaPDF * wavelengthProbability * W * H);
    const float vertexProb = 1.0f / numLightVertices;

    // connection of the lens vertex (t = 1), splatted to the pixel it projects to
    if (numLightVertices > 0) {
        uint32_t idx = std::min<uint32_t>(static_cast<uint32_t>(numLightVertices * rng.uniform()), numLightVertices - 1);
        const LightVertex<N> &vtx = cache[idx];
        SurfPt spL;
        int matL = decodeLightVertex<N>(s, vtx, &spL);
        V3 conRayDir;
        float sqDist;
        if (testVisibility(s, spE0, spL, &conRayDir, &sqDist, &st.shadowRays)) {
            float recSq = 1.0f / sqDist;
            VertexBSDF<N> bL;
            bL.fromEDF = vtx.pathLength == 0;
            if (bL.fromEDF) bL.edf = setupEDF<N>(s, s.materials[matL], spL, wlsCommon);
            else bL.bsdf = setupBSDF<N>(s.materials[matL], wlsCommon);
            if (bL.hasNonDelta()) {
                V3 gnLocalL = spL.frame.toLocal(spL.geometricNormal);
                BSDFQuery<N> qL{vtx.dirInLocal, gnLocalL, DT::All, TransportMode::Importance, wlsCommon.sel};
                V3 conLocalL = spL.frame.toLocal(-conRayDir);
                V3 conLocalE = spE0.frame.toLocal(conRayDir);
                float cosL = spL.calcCosTerm(conRayDir);
                float cosE = spE0.calcCosTerm(conRayDir);
                float G = cosL * cosE * recSq;
                float scalarConTerm = G * 1.0f / vertexProb;
                if (vtx.wlSelected) scalarConTerm *= N;
                SS<N> backwardFsL;
                SS<N> forwardFsL = bL.evaluate(qL, conLocalL, &backwardFsL);
                float backwardDirDensityL;
                bL.evaluatePDF(qL, conLocalL, &backwardDirDensityL);
                backwardDirDensityL *= rrProbability<N>(backwardFsL, qL.dirLocal, backwardDirDensityL, gnLocalL, wlsCommon.sel);
                float backwardAreaDensityL = backwardDirDensityL * vtx.backwardConversionFactor;
                if (vtx.prevDeltaSampled) backwardAreaDensityL = 0;
                SS<N> forwardFsE(idf.directionalImportance(conLocalE));
                float forwardDirDensityE = idf.pdf(conLocalE);
                float forwardAreaDensityE = forwardDirDensityE * cosL * recSq;
                if (spL.isPoint) forwardAreaDensityE = 0;
                float u0, u1;
                bool onScreen = idf.project(conLocalE, &u0, &u1);
                float partialL;
                {
                    bool lastValid = (!vtx.deltaSampled && !vtx.prevDeltaSampled) && vtx.pathLength > 1;
                    float lastToSecondLastProbRatio = backwardAreaDensityL / vtx.prevProbDensity;
                    partialL = pow2(lastToSecondLastProbRatio) * (vtx.secondPrevPartialDenomMisWeight + (lastValid ? 1 : 0));
                    float probRatioToFirst = vtx.secondPrevProbRatioToFirst;
                    if (vtx.pathLength > 0) probRatioToFirst *= lastToSecondLastProbRatio;
                    bool curValid = !vtx.deltaSampled && vtx.pathLength > 0;
                    float curToLastProbRatio = forwardAreaDensityE / vtx.probDensity;
                    partialL = pow2(curToLastProbRatio) * (partialL + (curValid ? 1 : 0));
                    probRatioToFirst *= curToLastProbRatio;
                    partialL += pow2(probRatioToFirst / vertexProb);
                }
                SS<N> conTerm = forwardFsL * scalarConTerm * forwardFsE;
                SS<N> unweighted = vtx.flux * conTerm * alpha;
                float misWeight = 1.0f / (1.0f + partialL);
                SS<N> contribution = misWeight * unweighted;
                if (contribution.allFinite() && onScreen) {
                    uint32_t ipx = static_cast<uint32_t>(u0 * W), ipy = static_cast<uint32_t>(u1 * H);
                    if (ipx < W && ipy < H) {
                        float *dst = splat->data() + (size_t(ipy) * W + ipx) * C;
                        if constexpr (N == 4) {
                            SS<16> b = binSpectral(wlsCommon, contribution);
                            for (int c = 0; c < 16; ++c) {
#pragma omp atomic
                                dst[c] += b[c];
                            }
                        } else {
                            for (int c = 0; c < 3; ++c) {
#pragma omp atomic
                                dst[c] += contribution[c];
                            }
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
    float dirPDF = pow2(cam.imgPlaneDistance) / (pow3(dirLocal.z) * cam.imgPlaneArea);
    V3 rayOrg = spE0.position;
    V3 rayDir = spE0.frame.fromLocal(dirLocal);
    float cosTerm = spE0.calcCosTerm(rayDir);
    alpha *= SS<N>(1.0f) * (cosTerm / dirPDF);

    float roDirPDF = dirPDF, roCosTerm = cosTerm, roPrevRevAreaPDF = 1;
    bool roSecondPrevDelta = false, roPrevDelta = DT::isDelta(lensPosType);
    uint32_t roSampledType = DT::Acquisition | DT::LowFreq;
    float probDensity = lensAreaPDF, prevProbDensity = 1.0f, secondPrevPartialDenomMisWeight = 0;
    bool single = false;
    uint32_t pathLength = 0;
    SS<N> contribution(0.0f);
    while (true) {
        bool terminate = true;
        ++pathLength;
        HitRec hit;
        ++st.closestRays;
        bool found = traceRay(s, rayOrg, rayDir, 0.0f, FLT_MAX, &hit, false, true);
        WLS<N> wls = wlsCommon;
        if (single) wls.single = true;
        V3 nextOrg, nextDir;
        float woDirPDF = 0, woCosTerm = 0, woRevAreaPDF = 0;
        uint32_t woSampledType = 0;
        if (!found) {
            // miss program (lvc_bpt.cu:841-971)
            if (s.hasEnv && s.envImportance > 0) {
                SurfPt spE;
                float theta, posPhi;
                rayDir.toPolarYUp(&theta, &posPhi);
                float sinPhi = std::sin(posPhi), cosPhi = std::cos(posPhi);
                V3 tc0 = normalize(V3(-cosPhi, 0.0f, -sinPhi));
                spE.frame.x = tc0;
                spE.frame.z = -rayDir;
                spE.frame.y = cross(spE.frame.z, spE.frame.x);
                spE.position = rayDir;
                spE.isPoint = false;
                spE.atInfinity = true;
                spE.geometricNormal = -rayDir;
                spE.u = posPhi;
                spE.v = theta;
                float phi = posPhi + s.envRotation;
                phi = phi - std::floor(phi / (2 * kPi)) * 2 * kPi;
                spE.texU = phi / (2 * kPi);
                spE.texV = theta / kPi;
                float uvPDF = s.envImportanceMap.evaluatePDF(phi / (2 * kPi), theta / kPi);
                float hypAreaPDF = uvPDF / (2 * kPi * kPi * std::sin(theta));
                EDF<N> edf = setupEDF<N>(s, s.materials[s.envMaterial], spE, wls);
                V3 dirOutLocalE = spE.frame.toLocal(-rayDir);
                bool secondLastValid = (!roPrevDelta && !roSecondPrevDelta) && pathLength > 2;
                secondPrevPartialDenomMisWeight = pow2(roPrevRevAreaPDF / prevProbDensity) *
                                                  (secondPrevPartialDenomMisWeight + (secondLastValid ? 1 : 0));
                prevProbDensity = probDensity;
                float lastDist2 = 1.0f;
                probDensity = roDirPDF / lastDist2;
                SS<N> spEmittance = edf.evaluateEmittance();
                if (spEmittance.hasNonZero() && edf.matches(DT::Emission | DT::NonDelta)) {
                    SS<N> Le = spEmittance * edf.evaluate(dirOutLocalE);
                    SS<N> unweighted = alpha * Le;
                    if (single) unweighted *= float(N);
                    float instProb = 1.0f / s.lightInstDist.integral;
                    float geomInstProb = 1.0f;
                    float forwardAreaDensityL = bs.numLightPaths * instProb * geomInstProb * hypAreaPDF;
                    float backwardDirDensityE = edf.evaluatePDF(dirOutLocalE);
                    float backwardAreaDensityE = backwardDirDensityE * roCosTerm / lastDist2;
                    float partialE;
                    {
                        bool lastValid = (!DT::isDelta(roSampledType) && !roPrevDelta) && pathLength > 1;
                        float r = backwardAreaDensityE / prevProbDensity;
                        partialE = pow2(r) * (secondPrevPartialDenomMisWeight + (lastValid ? 1 : 0));
                        bool curValid = !DT::isDelta(roSampledType);
                        float r2 = forwardAreaDensityL / probDensity;
                        partialE = pow2(r2) * (partialE + (curValid ? 1 : 0));
                        partialE *= pow2(vertexProb);
                    }
                    SS<N> c = (1.0f / (1.0f + partialE)) * unweighted;
                    if (c.allFinite()) contribution += c;
                }
            }
            break;
        }
        // closest-hit program for eye paths (lvc_bpt.cu:570-837)
        SurfPt spE;
        float hypAreaPDF;
        calcSurfacePoint(s, hit, &spE, &hypAreaPDF);
        const GeomInst &g = s.geomInsts[hit.geomInst];
        const Material &matE = s.materials[g.material];
        BSDF<N> bsdfE = setupBSDF<N>(matE, wls);
        EDF<N> edf = setupEDF<N>(s, matE, spE, wls);
        V3 dirOutLocalE = spE.frame.toLocal(-rayDir);
        V3 gnLocalE = spE.frame.toLocal(spE.geometricNormal);
        BSDFQuery<N> qE{dirOutLocalE, gnLocalE, DT::All, TransportMode::Radiance, wls.sel};
        bool secondLastValidE = (!roPrevDelta && !roSecondPrevDelta) && pathLength > 2;
        secondPrevPartialDenomMisWeight = pow2(roPrevRevAreaPDF / prevProbDensity) *
                                          (secondPrevPartialDenomMisWeight + (secondLastValidE ? 1 : 0));
        prevProbDensity = probDensity;
        float lastDist2 = sqDistance(rayOrg, spE.position);
        probDensity = roDirPDF * absDot(dirOutLocalE, gnLocalE) / lastDist2;

        SS<N> spEmittance = edf.evaluateEmittance();
        if (spEmittance.hasNonZero() && edf.matches(DT::Emission | DT::NonDelta)) {
            SS<N> Le = spEmittance * edf.evaluate(dirOutLocalE);
            SS<N> unweighted = alpha * Le;
            if (single) unweighted *= float(N);
            const Instance &in = s.instances[hit.inst];
            float instProb = in.lightDist.integral / s.lightInstDist.integral;
            float geomInstProb = g.importance / in.lightDist.integral;
            float forwardAreaDensityL = bs.numLightPaths * instProb * geomInstProb * hypAreaPDF;
            float backwardDirDensityE = edf.evaluatePDF(dirOutLocalE);
            float backwardAreaDensityE = backwardDirDensityE * roCosTerm / lastDist2;
            float partialE;
            {
                bool lastValid = (!DT::isDelta(roSampledType) && !roPrevDelta) && pathLength > 1;
                float r = backwardAreaDensityE / prevProbDensity;
                partialE = pow2(r) * (secondPrevPartialDenomMisWeight + (lastValid ? 1 : 0));
                bool curValid = !DT::isDelta(roSampledType);
                float r2 = forwardAreaDensityL / probDensity;
                partialE = pow2(r2) * (partialE + (curValid ? 1 : 0));
                partialE *= pow2(vertexProb);
            }
            SS<N> c = (1.0f / (1.0f + partialE)) * unweighted;
            if (c.allFinite()) contribution += c;
        }

        if (bsdfE.hasNonDelta() && numLightVertices > 0) {
            uint32_t idx = std::min<uint32_t>(static_cast<uint32_t>(numLightVertices * rng.uniform()), numLightVertices - 1);
            const LightVertex<N> &vtx = cache[idx];
            SurfPt spL;
            int matL = decodeLightVertex<N>(s, vtx, &spL);
            V3 conRayDir;
            float sqDist;
            if (testVisibility(s, spE, spL, &conRayDir, &sqDist, &st.shadowRays)) {
                float recSq = 1.0f / sqDist;
                VertexBSDF<N> bL;
                bL.fromEDF = vtx.pathLength == 0;
                if (bL.fromEDF) bL.edf = setupEDF<N>(s, s.materials[matL], spL, wls);
                else bL.bsdf = setupBSDF<N>(s.materials[matL], wls);
                if (bL.hasNonDelta()) {
                    V3 gnLocalL = spL.frame.toLocal(spL.geometricNormal);
                    BSDFQuery<N> qL{vtx.dirInLocal, gnLocalL, DT::All, TransportMode::Importance, wls.sel};
                    V3 conLocalL = spL.frame.toLocal(-conRayDir);
                    V3 conLocalE = spE.frame.toLocal(conRayDir);
                    float cosL = spL.calcCosTerm(conRayDir);
                    float cosE = spE.calcCosTerm(conRayDir);
                    float G = cosL * cosE * recSq;
                    float scalarConTerm = G * 1.0f / vertexProb;
                    if (vtx.wlSelected || single) scalarConTerm *= N;
                    SS<N> backwardFsL;
                    SS<N> forwardFsL = bL.evaluate(qL, conLocalL, &backwardFsL);
                    float backwardDirDensityL;
                    float forwardDirDensityL = bL.evaluatePDF(qL, conLocalL, &backwardDirDensityL);
                    if (vtx.pathLength > 0)
                        forwardDirDensityL *= rrProbability<N>(forwardFsL, conLocalL, forwardDirDensityL, gnLocalL, wlsCommon.sel);
                    backwardDirDensityL *= rrProbability<N>(backwardFsL, qL.dirLocal, backwardDirDensityL, gnLocalL, wlsCommon.sel);
                    float forwardAreaDensityL = forwardDirDensityL * cosE * recSq;
                    float backwardAreaDensityL = backwardDirDensityL * vtx.backwardConversionFactor;
                    if (vtx.prevDeltaSampled) backwardAreaDensityL = 0;

                    SS<N> backwardFsE;
                    SS<N> forwardFsE = bsdfE.evaluate(qE, conLocalE, &backwardFsE);
                    float backwardDirDensityE;
                    float forwardDirDensityE = bsdfE.evaluatePDF(qE, conLocalE, &backwardDirDensityE);
                    forwardDirDensityE *= rrProbability<N>(forwardFsE, conLocalE, forwardDirDensityE, gnLocalE, wlsCommon.sel);
                    backwardDirDensityE *= rrProbability<N>(backwardFsE, qE.dirLocal, backwardDirDensityE, gnLocalE, wlsCommon.sel);
                    float forwardAreaDensityE = forwardDirDensityE * cosL * recSq;
                    if (spL.isPoint) forwardAreaDensityE = 0;
                    float backwardAreaDensityE = backwardDirDensityE * roCosTerm / lastDist2;

                    float partialL;
                    {
                        bool lastValid = (!vtx.deltaSampled && !vtx.prevDeltaSampled) && vtx.pathLength > 1;
                        float r = backwardAreaDensityL / vtx.prevProbDensity;
                        partialL = pow2(r) * (vtx.secondPrevPartialDenomMisWeight + (lastValid ? 1 : 0));
                        float probRatioToFirst = vtx.secondPrevProbRatioToFirst;
                        if (vtx.pathLength > 0) probRatioToFirst *= r;
                        bool curValid = !vtx.deltaSampled && vtx.pathLength > 0;
                        float r2 = forwardAreaDensityE / vtx.probDensity;
                        partialL = pow2(r2) * (partialL + (curValid ? 1 : 0));
                        probRatioToFirst *= r2;
                        partialL += pow2(probRatioToFirst / vertexProb);
                    }
                    float partialE;
                    {
                        bool lastValid = !DT::isDelta(roSampledType) && !roPrevDelta;
                        float r = backwardAreaDensityE / prevProbDensity;
                        partialE = pow2(r) * (secondPrevPartialDenomMisWeight + (lastValid ? 1 : 0));
                        bool curValid = !DT::isDelta(roSampledType);
                        float r2 = forwardAreaDensityL / probDensity;
                        partialE = pow2(r2) * (partialE + (curValid ? 1 : 0));
                    }
                    SS<N> conTerm = forwardFsL * scalarConTerm * forwardFsE;
                    SS<N> unweighted = vtx.flux * conTerm * alpha;
                    SS<N> c = (1.0f / (1.0f + partialL + partialE)) * unweighted;
                    if (c.allFinite()) contribution += c;
                }
            }
        }

        float sC = rng.uniform(), s0 = rng.uniform(), s1 = rng.uniform();
        BSDFResult<N> res;
        RevResult<N> rev;
        SS<N> fs = bsdfE.sample(qE, sC, s0, s1, &res, &rev);
        if (!(fs == SS<N>(0.0f) || res.dirPDF == 0.0f)) {
            if (DT::isDispersive(res.sampledType) && !single) single = true;
            float cosT = dot(res.dirLocal, gnLocalE);
            SS<N> localThroughput = fs * (std::fabs(cosT) / res.dirPDF);
            alpha *= localThroughput;
            float continueProb = std::fmin(localThroughput.importance(wlsCommon.sel), 1.0f);
            if (rng.uniform() < continueProb) {
                alpha /= continueProb;
                terminate = false;
                V3 dirInE = spE.frame.fromLocal(res.dirLocal);
                nextOrg = offsetRayOrigin(spE.position, cosT > 0.0f ? spE.geometricNormal : -spE.geometricNormal);
                nextDir = dirInE;
                woDirPDF = res.dirPDF * continueProb;
                woCosTerm = std::fabs(cosT);
                woRevAreaPDF = rev.dirPDF * roCosTerm / lastDist2;
                woRevAreaPDF *= rrProbability<N>(rev.value, qE.dirLocal, rev.dirPDF, gnLocalE, wlsCommon.sel);
                woSampledType = res.sampledType;
            }
        }
        if (pathLength >= bs.maxPathLength) terminate = true;
        if (terminate) break;
        rayOrg = nextOrg;
        rayDir = nextDir;
        roDirPDF = woDirPDF;
        roCosTerm = woCosTerm;
        roPrevRevAreaPDF = woRevAreaPDF;
        roSecondPrevDelta = roPrevDelta;
        roPrevDelta = DT::isDelta(roSampledType);
        roSampledType = woSampledType;
    }
    return contribution;
}

// context.cpp:1464-1494: per frame, common wavelengths from the context's mt19937 (two
// uniform floats), the light-path pass into a fresh cache, the eye pass, then the splats.
template <int N, int C>
static void renderBPTImpl(const Scene &s, uint32_t W, uint32_t H, uint32_t numFrames, uint32_t maxPathLength,
                          uint32_t numLightPaths, int numThreads, float *outRGBA, uint64_t *rngState,
                          uint64_t *lightRngState, float *accumOut, OracleRenderStats *stats) {
    std::vector<uint64_t> rngs(size_t(W) * H), lrngs(numLightPaths);
    if (rngState) std::memcpy(rngs.data(), rngState, rngs.size() * 8);
    else {
        std::mt19937_64 seedRng(591842031321323413ULL);
        for (auto &r : rngs) r = seedRng();
    }
    if (lightRngState) std::memcpy(lrngs.data(), lightRngState, lrngs.size() * 8);
    else {
        std::mt19937_64 seedRng(459182033132123413ULL);
        for (auto &r : lrngs) r = seedRng();
    }
    std::mt19937 wlRng;
    std::uniform_real_distribution<float> u01;
    BPTSettings bs{numLightPaths, maxPathLength};
    std::vector<Kahan<SS<C>>> accum(size_t(W) * H);
    for (auto &a : accum) a.reset();
    OracleRenderStats total = {};
    if (numThreads > 0) omp_set_num_threads(numThreads);
    for (uint32_t frame = 1; frame <= numFrames; ++frame) {
        float wlOff = u01(wlRng), wlSel = u01(wlRng);
        float wavelengthProbability;
        WLS<N> wls = WLS<N>::createWithEqualOffsets(wlOff, wlSel, &wavelengthProbability);
        // light paths: each light path appends its vertices in order; the cache is the paths'
        // vertex lists concatenated in light-path order
        std::vector<std::vector<LightVertex<N>>> perPath(numLightPaths);
#pragma omp parallel
        {
            OracleRenderStats st = {};
#pragma omp for schedule(dynamic, 256)
            for (int64_t i = 0; i < int64_t(numLightPaths); ++i) {
                PCG32 rng{lrngs[i]};
                traceLightPath<N>(s, wls, bs, rng, &perPath[i], st);
                lrngs[i] = rng.state;
            }
#pragma omp critical
            {
                total.closestRays += st.closestRays;
            }
        }
        std::vector<LightVertex<N>> cache;
        for (auto &v : perPath) cache.insert(cache.end(), v.begin(), v.end());
        std::vector<float> splat(size_t(W) * H * C, 0.0f);
        std::vector<SS<C>> eye(size_t(W) * H);
        std::vector<uint8_t> eyeOk(size_t(W) * H, 0);
#pragma omp parallel
        {
            OracleRenderStats st = {};
#pragma omp for schedule(dynamic, 64)
            for (int64_t p = 0; p < int64_t(W) * H; ++p) {
                uint32_t x = uint32_t(p % W), y = uint32_t(p / W);
                PCG32 rng{rngs[p]};
                SS<N> c = eyePathContribution<N>(s, wls, wavelengthProbability, bs, cache, rng, x, y, W, H, &splat, st);
                rngs[p] = rng.state;
                ++st.paths;
                if (!c.allFinite()) {
                    ++st.droppedNonFinite;
                    continue;
                }
                eyeOk[p] = 1;
                if constexpr (N == 3) eye[p] = c;
                else eye[p] = binSpectral(wls, c);
            }
#pragma omp critical
            {
                total.closestRays += st.closestRays;
                total.shadowRays += st.shadowRays;
                total.paths += st.paths;
                total.droppedNonFinite += st.droppedNonFinite;
            }
        }
        for (size_t p = 0; p < accum.size(); ++p) {
            if (eyeOk[p]) {
                if (frame == 1) accum[p].reset();
                accum[p].add(eye[p]);
            } else if (frame == 1) {
                accum[p].reset();
            }
            SS<C> sv;
            for (int c = 0; c < C; ++c) sv[c] = splat[p * C + c];
            accum[p].add(sv);
        }
    }
    if (rngState) std::memcpy(rngState, rngs.data(), rngs.size() * 8);
    if (lightRngState) std::memcpy(lightRngState, lrngs.data(), lrngs.size() * 8);
    const float recN = 1.0f / numFrames;
    const float resCorrection = float(W) * float(H);
    for (size_t p = 0; p < accum.size(); ++p) {
        float XYZ[3];
        if constexpr (N == 3) {
            float RGB[3] = {accum[p].result[0], accum[p].result[1], accum[p].result[2]};
            transformTristimulus(mat_Rec709_D65_to_XYZ, RGB, XYZ);
        } else {
            strataToXYZ(accum[p].result, XYZ);
        }
        XYZ[0] *= recN; XYZ[1] *= recN; XYZ[2] *= recN;
        float RGB[3];
        transformTristimulus(mat_XYZ_to_Rec709_D65, XYZ, RGB);
        if (outRGBA) {
            outRGBA[4 * p + 0] = RGB[0] * resCorrection;
            outRGBA[4 * p + 1] = RGB[1] * resCorrection;
            outRGBA[4 * p + 2] = RGB[2] * resCorrection;
            outRGBA[4 * p + 3] = 1.0f;
        }
        if (accumOut) {
            for (int c = 0; c < C; ++c) {
                accumOut[(p * C + c) * 2 + 0] = accum[p].result[c];
                accumOut[(p * C + c) * 2 + 1] = accum[p].comp[c];
            }
        }
    }
    if (stats) *stats = total;
}

template <int N, int C>
static void renderImpl(const Scene &s, uint32_t W, uint32_t H, uint32_t numFrames, uint32_t maxPathLength,
                       int numThreads, float *outRGBA, uint64_t *rngState, float *accumOut,
                       const int32_t *range, OracleRenderStats *stats, float *albedoRGBA = nullptr,
                       float *normalRGBA = nullptr) {
    const bool guides = albedoRGBA || normalRGBA;
    std::vector<SS<C>> accumAlbedo(guides ? size_t(W) * H : 0, SS<C>(0.0f));
    std::vector<V3> accumNormal(guides ? size_t(W) * H : 0, V3(0.0f, 0.0f, 0.0f));
    std::vector<uint64_t> rngs(size_t(W) * H);
    if (rngState) {
        std::memcpy(rngs.data(), rngState, rngs.size() * 8);
    } else {
        std::mt19937_64 seedRng(591842031321323413ULL);
        for (auto &r : rngs) r = seedRng();
    }
    int32_t x0 = 0, y0 = 0, x1 = W, y1 = H;
    if (range) {
        x0 = range[0]; y0 = range[1]; x1 = range[2]; y1 = range[3];
    }
    std::vector<Kahan<SS<C>>> accum(size_t(W) * H);
    for (auto &a : accum) a.reset();
    const uint32_t tilesX = (x1 - x0 + 15) / 16, tilesY = (y1 - y0 + 15) / 16;
    OracleRenderStats total = {};
    if (numThreads > 0) omp_set_num_threads(numThreads);
    for (uint32_t frame = 1; frame <= numFrames; ++frame) {
#pragma omp parallel
        {
            OracleRenderStats st = {};
#pragma omp for schedule(dynamic)
            for (int64_t tile = 0; tile < int64_t(tilesX) * tilesY; ++tile) {
                int32_t tx = int32_t(tile % tilesX), ty = int32_t(tile / tilesX);
                for (int32_t y = y0 + ty * 16; y < std::min(y0 + ty * 16 + 16, y1); ++y) {
                    for (int32_t x = x0 + tx * 16; x < std::min(x0 + tx * 16 + 16, x1); ++x) {
                        size_t p = size_t(y) * W + x;
                        PCG32 rng{rngs[p]};
                        WLS<N> wls;
                        FirstHit<N> fh;
                        SS<N> c = tracePath<N>(s, rng, x, y, W, H, maxPathLength, &wls, st, guides ? &fh : nullptr);
                        rngs[p] = rng.state;
                        ++st.paths;
                        if (guides) {
                            // accumulated every sample, finite or not (path_tracing.cu:95-108)
                            if constexpr (N == 3) accumAlbedo[p] += fh.albedo;
                            else accumAlbedo[p] += binSpectral(wls, fh.albedo);
                            accumNormal[p] = accumNormal[p] + fh.normal;
                        }
                        if (!c.allFinite()) {
                            ++st.droppedNonFinite;
                            continue;
                        }
                        if constexpr (N == 3) {
                            accum[p].add(c);
                        } else {
                            accum[p].add(binSpectral(wls, c));
                        }
                    }
                }
            }
#pragma omp critical
            {
                total.closestRays += st.closestRays;
                total.shadowRays += st.shadowRays;
                total.paths += st.paths;
                total.droppedNonFinite += st.droppedNonFinite;
            }
        }
    }
    // convertToRGB (post_process.cu:93-163)
    const float recN = 1.0f / numFrames;
    const float resCorrection = float(W) * float(H);
    for (size_t p = 0; p < accum.size(); ++p) {
        float XYZ[3];
        if constexpr (N == 3) {
            float RGB[3] = {accum[p].result[0], accum[p].result[1], accum[p].result[2]};
            transformTristimulus(mat_Rec709_D65_to_XYZ, RGB, XYZ);
        } else {
            strataToXYZ(accum[p].result, XYZ);
        }
        XYZ[0] *= recN; XYZ[1] *= recN; XYZ[2] *= recN;
        float RGB[3];
        transformTristimulus(mat_XYZ_to_Rec709_D65, XYZ, RGB);
        if (outRGBA) {
            outRGBA[4 * p + 0] = RGB[0] * resCorrection;
            outRGBA[4 * p + 1] = RGB[1] * resCorrection;
            outRGBA[4 * p + 2] = RGB[2] * resCorrection;
            outRGBA[4 * p + 3] = 1.0f;
        }
        if (accumOut) {
            for (int c = 0; c < C; ++c) {
                accumOut[(p * C + c) * 2 + 0] = accum[p].result[c];
                accumOut[(p * C + c) * 2 + 1] = accum[p].comp[c];
            }
        }
        if (guides) {
            // copyBuffers (post_process.cu:70-90)
            float aXYZ[3];
            if constexpr (N == 3) {
                float a[3] = {accumAlbedo[p][0], accumAlbedo[p][1], accumAlbedo[p][2]};
                transformTristimulus(mat_Rec709_D65_to_XYZ, a, aXYZ);
            } else {
                strataToXYZ(accumAlbedo[p], aXYZ);
            }
            aXYZ[0] *= recN; aXYZ[1] *= recN; aXYZ[2] *= recN;
            float aRGB[3];
            transformTristimulus(mat_XYZ_to_Rec709_D65, aXYZ, aRGB);
            const M3 rot = s.camera.orientation.toMatrix3x3();
            const V3 n = accumNormal[p];
            // the inverse orientation: the transposed rotation
            V3 nc(dot(rot.c0, n), dot(rot.c1, n), dot(rot.c2, n));
            nc.x *= -1;
            if (nc.x != 0 || nc.y != 0 || nc.z != 0) nc = normalize(nc);
            if (albedoRGBA) {
                albedoRGBA[4 * p + 0] = aRGB[0];
                albedoRGBA[4 * p + 1] = aRGB[1];
                albedoRGBA[4 * p + 2] = aRGB[2];
                albedoRGBA[4 * p + 3] = 1.0f;
            }
            if (normalRGBA) {
                normalRGBA[4 * p + 0] = nc.x;
                normalRGBA[4 * p + 1] = nc.y;
                normalRGBA[4 * p + 2] = nc.z;
                normalRGBA[4 * p + 3] = 1.0f;
            }
        }
    }
    if (rngState) std::memcpy(rngState, rngs.data(), rngs.size() * 8);
    if (stats) *stats = total;
}

} // namespace ov

using namespace ov;

// ---- light tracing (light_tracing.cu:49-180 ray generation, 182-263 closest hit) ----
// One light path per launch index with its own wavelengths; every non-delta vertex (the light
// vertex included) is connected to a sampled lens point and splatted to the pixel it projects
// to.  The path ends at a miss, at a zero BSDF sample, by Russian roulette on the throughput's
// importance relative to its initial value, or at maxPathLength segments.
template <int N>
static void splatToPixel(const WLS<N> &wls, const SS<N> &contribution, float u0, float u1, uint32_t W, uint32_t H,
                         std::vector<float> *splat) {
    constexpr int C = N == 4 ? 16 : 3;
    // atomicAddToBuffer (light_tracing.cu:32-46): off-image and non-finite values are dropped
    uint32_t ipx = static_cast<uint32_t>(u0 * W), ipy = static_cast<uint32_t>(u1 * H);
    if (ipx >= W || ipy >= H || !contribution.allFinite()) return;
    float *dst = splat->data() + (size_t(ipy) * W + ipx) * C;
    if constexpr (N == 4) {
        SS<16> b = binSpectral(wls, contribution);
        for (int c = 0; c < 16; ++c) {
#pragma omp atomic
            dst[c] += b[c];
        }
    } else {
        for (int c = 0; c < 3; ++c) {
#pragma omp atomic
            dst[c] += contribution[c];
        }
    }
}

// a sampled lens point of the perspective camera (cameras.cu:43-76)
static void sampleLensPoint(const PerspectiveCamera &cam, float ul0, float ul1, SurfPt *spE0, LensIDF *idf,
                            float *lensAreaPDF) {
    M3 rot = cam.orientation.toMatrix3x3();
    float lx, ly;
    concentricSampleDisk(ul0, ul1, &lx, &ly);
    V3 orgLocal(cam.lensRadius * lx, cam.lensRadius * ly, 0.0f);
    spE0->frame.z = normalize(rot * V3(0, 0, 1));
    spE0->frame.x = normalize(rot * V3(1, 0, 0));
    spE0->frame.y = cross(spE0->frame.z, spE0->frame.x);
    spE0->position = rot * orgLocal + cam.position;
    spE0->geometricNormal = spE0->frame.z;
    spE0->isPoint = false;
    spE0->atInfinity = false;
    spE0->u = lx;
    spE0->v = ly;
    *lensAreaPDF = cam.lensRadius > 0.0f ? 1.0f / (kPi * cam.lensRadius * cam.lensRadius) : 1.0f;
    *idf = LensIDF{cam.sensitivity, cam.lensRadius * lx, cam.lensRadius * ly, cam.opWidth, cam.opHeight,
                   cam.imgPlaneDistance, cam.objPlaneDistance, cam.imgPlaneArea};
}

template <int N>
static void lightTracingPath(const Scene &s, const BPTSettings &bs, PCG32 &rng, uint32_t W, uint32_t H,
                             std::vector<float> *splat, OracleRenderStats &st) {
    float wlOff = rng.uniform(), wlSel = rng.uniform();
    float selectWLPDF;
    WLS<N> wls = WLS<N>::createWithEqualOffsets(wlOff, wlSel, &selectWLPDF);
    if (s.lightInstDist.integral <= 0.0f) return;
    float uLight = rng.uniform();
    float uP0 = rng.uniform(), uP1 = rng.uniform();
    LightSample ls;
    float lightProb;
    bool valid;
    sampleLight(s, uLight, uP0, uP1, &ls, &lightProb, &valid);
    const Material &lmat = s.materials[ls.material];
    EDF<N> edf = setupEDF<N>(s, lmat, ls.sp, wls);
    SS<N> Le0 = edf.evaluateEmittance();
    SS<N> alpha = Le0 / (bs.numLightPaths * lightProb * ls.areaPDF * selectWLPDF);

    // lens connection of the light vertex
    {
        float ul0 = rng.uniform(), ul1 = rng.uniform();
        SurfPt spE0;
        LensIDF idf;
        float lensPDF;
        sampleLensPoint(s.camera, ul0, ul1, &spE0, &idf, &lensPDF);
        SS<N> We0(idf.sensitivity);
        V3 dir;
        float dist2;
        if (We0.hasNonZero() && testVisibility(s, spE0, ls.sp, &dir, &dist2, &st.shadowRays)) {
            V3 dirLens = spE0.frame.toLocal(dir);
            V3 dirSn = ls.sp.frame.toLocal(-dir);
            SS<N> We = We0 * idf.directionalImportance(dirLens);
            float u0 = 0, u1 = 0;
            bool onScreen = idf.project(dirLens, &u0, &u1);
            SS<N> Le1 = edf.evaluate(dirSn);
            float cosLens = spE0.calcCosTerm(dir);
            float cosLight = ls.sp.calcCosTerm(-dir);
            float G = cosLight * cosLens / dist2;
            float scalarCoeff = G / lensPDF;
            SS<N> contribution = alpha * We * Le1 * scalarCoeff;
            if (onScreen) splatToPixel<N>(wls, contribution, u0, u1, W, H, splat);
        }
    }

    float uC = rng.uniform(), uD0 = rng.uniform(), uD1 = rng.uniform();
    (void)uC;
    V3 dirLocal;
    float dirPDF;
    uint32_t sampledType;
    SS<N> Le1 = edf.sample(uD0, uD1, &dirLocal, &dirPDF, &sampledType);
    V3 rayOrg = offsetRayOrigin(ls.sp.position, ls.sp.geometricNormal);
    if (ls.sp.atInfinity) {
        rayOrg = s.center + 1.1f * s.worldRadius * ls.sp.position + ls.sp.frame.x * dirLocal.x +
                 ls.sp.frame.y * dirLocal.y;
        dirLocal.x = 0;
        dirLocal.y = 0;
    }
    V3 rayDir = ls.sp.frame.fromLocal(dirLocal);
    alpha *= Le1 * (ls.sp.calcCosTerm(rayDir) / dirPDF);
    const float initImportance = alpha.importance(wls.sel);

    // The ray-generation program hands a copy of its generator to the hit program
    // (rwPayload.rng = rng, light_tracing.cu:138) and keeps drawing its Russian-roulette numbers
    // from its own copy (light_tracing.cu:165-169); the hit program's copy is the state stored
    // at the end (light_tracing.cu:176).  `rng` plays the hit program's copy here.
    PCG32 rrRng = rng;
    uint32_t pathLength = 0;
    while (true) {
        ++pathLength;
        if (pathLength >= bs.maxPathLength) break;
        HitRec hit;
        ++st.closestRays;
        if (!traceRay(s, rayOrg, rayDir, 0.0f, FLT_MAX, &hit, false, true)) break;
        SurfPt sp;
        float hypAreaPDF;
        calcSurfacePoint(s, hit, &sp, &hypAreaPDF);
        const GeomInst &g = s.geomInsts[hit.geomInst];
        BSDF<N> bsdf = setupBSDF<N>(s.materials[g.material], wls);
        V3 dirInLocal = sp.frame.toLocal(-rayDir);
        V3 gnLocal = sp.frame.toLocal(sp.geometricNormal);
        BSDFQuery<N> q{dirInLocal, gnLocal, DT::All, TransportMode::Importance, wls.sel};
        if (bsdf.hasNonDelta()) {
            float ul0 = rng.uniform(), ul1 = rng.uniform();
            SurfPt spE0;
            LensIDF idf;
            float lensPDF;
            sampleLensPoint(s.camera, ul0, ul1, &spE0, &idf, &lensPDF);
            SS<N> We0(idf.sensitivity);
            V3 dir;
            float dist2;
            if (We0.hasNonZero() && testVisibility(s, spE0, sp, &dir, &dist2, &st.shadowRays)) {
                V3 dirLens = spE0.frame.toLocal(dir);
                V3 dirSn = sp.frame.toLocal(-dir);
                SS<N> We = We0 * idf.directionalImportance(dirLens);
                float u0 = 0, u1 = 0;
                bool onScreen = idf.project(dirLens, &u0, &u1);
                SS<N> fs = bsdf.evaluate(q, dirSn);
                float cosLens = spE0.calcCosTerm(dir);
                float G = absDot(dirSn, gnLocal) * cosLens / dist2;
                float scalarCoeff = G / lensPDF;
                SS<N> contribution = alpha * We * fs * scalarCoeff;
                if (onScreen) splatToPixel<N>(wls, contribution, u0, u1, W, H, splat);
            }
        }
        float sC = rng.uniform(), s0 = rng.uniform(), s1 = rng.uniform();
        BSDFResult<N> res;
        SS<N> fs = bsdf.sample(q, sC, s0, s1, &res);
        if (fs == SS<N>(0.0f) || res.dirPDF == 0.0f) break;
        if (DT::isDispersive(res.sampledType) && !wls.single) {
            res.dirPDF /= N;
            wls.single = true;
        }
        float cosFactor = dot(res.dirLocal, gnLocal);
        alpha *= fs * (std::fabs(cosFactor) / res.dirPDF);
        rayOrg = offsetRayOrigin(sp.position, cosFactor > 0.0f ? sp.geometricNormal : -sp.geometricNormal);
        rayDir = sp.frame.fromLocal(res.dirLocal);
        // Russian roulette (light_tracing.cu:165-169), from the ray-generation program's generator
        float continueProb = std::fmin(alpha.importance(wls.sel) / initImportance, 1.0f);
        if (rrRng.uniform() >= continueProb) break;
        alpha /= continueProb;
    }
}

// Frames of light tracing: per frame the splats of every light path, added to the Kahan
// accumulation (reset on the first frame), as context.cpp:1444-1463 and post_process.cu:18-31.
template <int N, int C>
static void renderLTImpl(const Scene &s, uint32_t W, uint32_t H, uint32_t numFrames, uint32_t maxPathLength,
                         uint32_t numLightPaths, int numThreads, float *outRGBA, uint64_t *lightRngState,
                         float *accumOut, OracleRenderStats *stats) {
    std::vector<uint64_t> lrngs(numLightPaths);
    if (lightRngState) std::memcpy(lrngs.data(), lightRngState, lrngs.size() * 8);
    else {
        std::mt19937_64 seedRng(459182033132123413ULL);
        for (auto &r : lrngs) r = seedRng();
    }
    BPTSettings bs{numLightPaths, maxPathLength};
    std::vector<Kahan<SS<C>>> accum(size_t(W) * H);
    for (auto &a : accum) a.reset();
    OracleRenderStats total = {};
    if (numThreads > 0) omp_set_num_threads(numThreads);
    for (uint32_t frame = 1; frame <= numFrames; ++frame) {
        std::vector<float> splat(size_t(W) * H * C, 0.0f);
#pragma omp parallel
        {
            OracleRenderStats st = {};
#pragma omp for schedule(dynamic, 256)
            for (int64_t i = 0; i < int64_t(numLightPaths); ++i) {
                PCG32 rng{lrngs[i]};
                lightTracingPath<N>(s, bs, rng, W, H, &splat, st);
                lrngs[i] = rng.state;
                ++st.paths;
            }
#pragma omp critical
            {
                total.closestRays += st.closestRays;
                total.shadowRays += st.shadowRays;
                total.paths += st.paths;
            }
        }
        for (size_t p = 0; p < accum.size(); ++p) {
            if (frame == 1) accum[p].reset();
            SS<C> sv;
            for (int c = 0; c < C; ++c) sv[c] = splat[p * C + c];
            accum[p].add(sv);
        }
    }
    if (lightRngState) std::memcpy(lightRngState, lrngs.data(), lrngs.size() * 8);
    const float recN = 1.0f / numFrames;
    const float resCorrection = float(W) * float(H);
    for (size_t p = 0; p < accum.size(); ++p) {
        float XYZ[3];
        if constexpr (N == 3) {
            float RGB[3] = {accum[p].result[0], accum[p].result[1], accum[p].result[2]};
            transformTristimulus(mat_Rec709_D65_to_XYZ, RGB, XYZ);
        } else {
            strataToXYZ(accum[p].result, XYZ);
        }
        XYZ[0] *= recN; XYZ[1] *= recN; XYZ[2] *= recN;
        float RGB[3];
        transformTristimulus(mat_XYZ_to_Rec709_D65, XYZ, RGB);
        if (outRGBA) {
            outRGBA[4 * p + 0] = RGB[0] * resCorrection;
            outRGBA[4 * p + 1] = RGB[1] * resCorrection;
            outRGBA[4 * p + 2] = RGB[2] * resCorrection;
            outRGBA[4 * p + 3] = 1.0f;
        }
        if (accumOut) {
            for (int c = 0; c < C; ++c) {
                accumOut[(p * C + c) * 2 + 0] = accum[p].result[c];
                accumOut[(p * C + c) * 2 + 1] = accum[p].comp[c];
            }
        }
    }
    if (stats) *stats = total;
}

extern "C" {

int oracle_load_tables(const char *path) { return g_tables.load(path) ? 0 : 1; }

void *oracle_scene_create(const OracleScene *desc, int spectral) {
    if (!g_tables.loaded) return nullptr;
    return buildScene(desc, spectral != 0);
}

void oracle_scene_destroy(void *s) { delete static_cast<Scene *>(s); }

void oracle_scene_bounds(void *sp, float out[5]) {
    Scene *s = static_cast<Scene *>(sp);
    out[0] = s->center.x; out[1] = s->center.y; out[2] = s->center.z;
    out[3] = s->worldRadius; out[4] = s->worldDiscArea;
}

int oracle_trace_closest(void *sp, const OracleRay *rays, uint32_t n, OracleHit *hits, int useBVH, int numThreads) {
    const Scene &s = *static_cast<Scene *>(sp);
    if (numThreads > 0) omp_set_num_threads(numThreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        HitRec h;
        const OracleRay &r = rays[i];
        bool f = traceRay(s, V3(r.org[0], r.org[1], r.org[2]), V3(r.dir[0], r.dir[1], r.dir[2]), r.tmin, r.tmax, &h,
                          false, useBVH != 0);
        hits[i].instIndex = f ? h.inst : -1;
        hits[i].geomInstIndex = f ? h.geomInst : -1;
        hits[i].primIndex = f ? h.prim : -1;
        hits[i].t = f ? h.t : 0.0f;
        hits[i].b1 = f ? h.b1 : 0.0f;
        hits[i].b2 = f ? h.b2 : 0.0f;
    }
    return 0;
}

int oracle_trace_any(void *sp, const OracleRay *rays, uint32_t n, uint8_t *occ, int useBVH, int numThreads) {
    const Scene &s = *static_cast<Scene *>(sp);
    if (numThreads > 0) omp_set_num_threads(numThreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        const OracleRay &r = rays[i];
        occ[i] = traceRay(s, V3(r.org[0], r.org[1], r.org[2]), V3(r.dir[0], r.dir[1], r.dir[2]), r.tmin, r.tmax,
                          nullptr, true, useBVH != 0) ? 1 : 0;
    }
    return 0;
}

int oracle_render_lt(void *sp, uint32_t W, uint32_t H, uint32_t numFrames, uint32_t maxPathLength,
                     uint32_t numLightPaths, int numThreads, float *outRGBA, uint64_t *lightRngState, float *accumOut,
                     OracleRenderStats *stats) {
    Scene *s = static_cast<Scene *>(sp);
    if (!s || numLightPaths == 0) return 1;
    if (s->spectral)
        renderLTImpl<4, 16>(*s, W, H, numFrames, maxPathLength, numLightPaths, numThreads, outRGBA, lightRngState,
                            accumOut, stats);
    else
        renderLTImpl<3, 3>(*s, W, H, numFrames, maxPathLength, numLightPaths, numThreads, outRGBA, lightRngState,
                           accumOut, stats);
    return 0;
}

int oracle_render_bpt(void *sp, uint32_t W, uint32_t H, uint32_t numFrames, uint32_t maxPathLength,
                      uint32_t numLightPaths, int numThreads, float *outRGBA, uint64_t *rngState,
                      uint64_t *lightRngState, float *accumOut, OracleRenderStats *stats) {
    Scene *s = static_cast<Scene *>(sp);
    if (!s || numLightPaths == 0) return 1;
    if (s->spectral)
        renderBPTImpl<4, 16>(*s, W, H, numFrames, maxPathLength, numLightPaths, numThreads, outRGBA, rngState,
                             lightRngState, accumOut, stats);
    else
        renderBPTImpl<3, 3>(*s, W, H, numFrames, maxPathLength, numLightPaths, numThreads, outRGBA, rngState,
                            lightRngState, accumOut, stats);
    return 0;
}

int oracle_render(void *sp, uint32_t W, uint32_t H, uint32_t numFrames, uint32_t maxPathLength, int numThreads,
                  float *outRGBA, uint64_t *rngState, float *accumOut, const int32_t *range,
                  OracleRenderStats *stats) {
    const Scene &s = *static_cast<Scene *>(sp);
    if (numFrames == 0 || W == 0 || H == 0) return 1;
    if (s.spectral)
        renderImpl<4, 16>(s, W, H, numFrames, maxPathLength, numThreads, outRGBA, rngState, accumOut, range, stats);
    else
        renderImpl<3, 3>(s, W, H, numFrames, maxPathLength, numThreads, outRGBA, rngState, accumOut, range, stats);
    return 0;
}

// The path tracer's denoiser guides: first-hit albedo and normal accumulated over numFrames
// frames and converted as copyBuffers does (linear Rec.709 albedo; camera-space normal with x
// negated, normalised), W*H RGBA each; outRGBA, rngState and stats as oracle_render.
int oracle_render_guides(void *sp, uint32_t W, uint32_t H, uint32_t numFrames, uint32_t maxPathLength,
                         int numThreads, float *outRGBA, uint64_t *rngState, float *albedoRGBA, float *normalRGBA,
                         OracleRenderStats *stats) {
    const Scene &s = *static_cast<Scene *>(sp);
    if (numFrames == 0 || W == 0 || H == 0) return 1;
    if (s.spectral)
        renderImpl<4, 16>(s, W, H, numFrames, maxPathLength, numThreads, outRGBA, rngState, nullptr, nullptr, stats,
                          albedoRGBA, normalRGBA);
    else
        renderImpl<3, 3>(s, W, H, numFrames, maxPathLength, numThreads, outRGBA, rngState, nullptr, nullptr, stats,
                         albedoRGBA, normalRGBA);
    return 0;
}

} // extern "C"

template <int N>
static void bsdfQueries(const OracleMaterial *mat, const OracleBSDFQuery *q, uint32_t n, OracleBSDFResult *out) {
    OracleScene d = {};
    d.materials = mat;
    d.numMaterials = 1;
    d.envMaterial = -1;
    d.camera.aspect = 1; d.camera.fovY = 1; d.camera.opDistance = 1; d.camera.orientation[3] = 1;
    std::unique_ptr<Scene> s(buildScene(&d, N == 4));
    for (uint32_t i = 0; i < n; ++i) {
        float pdf;
        WLS<N> wls = WLS<N>::createWithEqualOffsets(q[i].wlOffset, q[i].wlSelect, &pdf);
        wls.single = q[i].singleIsSelected != 0;
        BSDF<N> b = setupBSDF<N>(s->materials[0], wls);
        BSDFQuery<N> bq{V3(q[i].dirLocal[0], q[i].dirLocal[1], q[i].dirLocal[2]),
                        V3(q[i].geomNormalLocal[0], q[i].geomNormalLocal[1], q[i].geomNormalLocal[2]), DT::All,
                        q[i].transportMode ? TransportMode::Importance : TransportMode::Radiance, wls.sel};
        BSDFResult<N> r;
        RevResult<N> sr;
        SS<N> f = b.sample(bq, q[i].uComponent, q[i].uDir[0], q[i].uDir[1], &r, &sr);
        OracleBSDFResult &o = out[i];
        std::memset(&o, 0, sizeof(o));
        o.sampledDir[0] = r.dirLocal.x; o.sampledDir[1] = r.dirLocal.y; o.sampledDir[2] = r.dirLocal.z;
        o.sampledPDF = r.dirPDF;
        o.sampledType = r.sampledType;
        for (int c = 0; c < N; ++c) o.sampledValue[c] = f[c];
        V3 ed(q[i].evalDir[0], q[i].evalDir[1], q[i].evalDir[2]);
        SS<N> rv;
        SS<N> e = b.evaluate(bq, ed, &rv);
        float revPDF = 0;
        o.evalPDF = b.evaluatePDF(bq, ed, &revPDF);
        o.revPDF = revPDF;
        for (int c = 0; c < N; ++c) { o.evalValue[c] = e[c]; o.revValue[c] = rv[c]; }
        o.sampledRevPDF = r.dirPDF > 0.0f ? sr.dirPDF : 0.0f;
        for (int c = 0; c < N; ++c) o.sampledRevValue[c] = r.dirPDF > 0.0f ? sr.value[c] : 0.0f;
    }
}

extern "C" {

int oracle_bsdf_queries(const OracleMaterial *mat, int spectral, const OracleBSDFQuery *q, uint32_t n,
                        OracleBSDFResult *out) {
    if (!g_tables.loaded) return 1;
    if (spectral) bsdfQueries<4>(mat, q, n, out);
    else bsdfQueries<3>(mat, q, n, out);
    return 0;
}

uint32_t oracle_pcg32_stream(uint64_t state, uint32_t n, uint32_t *out, uint64_t *finalState) {
    PCG32 r{state};
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next();
    if (finalState) *finalState = r.state;
    return n;
}

void oracle_seed_states(uint32_t W, uint32_t H, uint64_t *out) {
    std::mt19937_64 seedRng(591842031321323413ULL);
    for (size_t i = 0; i < size_t(W) * H; ++i) out[i] = seedRng();
}

void oracle_sample_lights(void *sp, uint32_t n, const float *u, OracleLightSample *out) {
    const Scene &s = *static_cast<Scene *>(sp);
    for (uint32_t i = 0; i < n; ++i) {
        LightSample ls;
        float lightProb = 0;
        bool valid = false;
        OracleLightSample &o = out[i];
        std::memset(&o, 0, sizeof(o));
        if (s.lightInstDist.integral <= 0.0f) continue;
        sampleLight(s, u[3 * i], u[3 * i + 1], u[3 * i + 2], &ls, &lightProb, &valid);
        o.position[0] = ls.sp.position.x; o.position[1] = ls.sp.position.y; o.position[2] = ls.sp.position.z;
        o.normal[0] = ls.sp.geometricNormal.x; o.normal[1] = ls.sp.geometricNormal.y;
        o.normal[2] = ls.sp.geometricNormal.z;
        o.areaPDF = ls.areaPDF;
        o.lightProb = lightProb;
        o.instIndex = ls.sp.atInfinity ? -1 : static_cast<int32_t>(ls.sp.instIndex);
        o.primIndex = static_cast<int32_t>(ls.sp.primIndex);
        o.atInfinity = ls.sp.atInfinity ? 1 : 0;
        o.valid = valid ? 1 : 0;
    }
}

void oracle_env_lookup(void *sp, uint32_t n, const float *dirs, float *out) {
    const Scene &s = *static_cast<Scene *>(sp);
    for (uint32_t i = 0; i < n; ++i) {
        V3 d(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
        float phi, theta;
        d.toPolarYUp(&theta, &phi);
        phi += s.envRotation;
        phi = phi - std::floor(phi / (2 * kPi)) * 2 * kPi;
        float u = phi / (2 * kPi), v = theta / kPi;
        out[4 * i + 0] = u;
        out[4 * i + 1] = v;
        out[4 * i + 2] = s.hasEnv ? s.envImportanceMap.evaluatePDF(u, v) / (2 * kPi * kPi * std::sin(theta)) : 0.0f;
        float tex[4] = {0, 0, 0, 0};
        if (!s.envTex.empty()) envFetch(s, u, v, tex);
        out[4 * i + 3] = tex[0];
    }
}

void oracle_pcg32_floats(uint64_t state, uint32_t n, float *out) {
    PCG32 r{state};
    for (uint32_t i = 0; i < n; ++i) out[i] = r.uniform();
}

void oracle_mt19937_64(uint64_t seed, uint32_t n, uint64_t *out) {
    std::mt19937_64 g(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = g();
}

void oracle_meng_eval(int spectrumType, int colorSpace, float e0, float e1, float e2, const float lambdas[4],
                      float out[4]) {
    Meng m = Meng::create(static_cast<SpectrumType>(spectrumType), static_cast<ColorSpace>(colorSpace), e0, e1, e2);
    WLS<4> w;
    for (int i = 0; i < 4; ++i) w.lambdas[i] = lambdas[i];
    w.sel = 0;
    w.single = false;
    SS<4> r = m.evaluate(w);
    for (int i = 0; i < 4; ++i) out[i] = r[i];
}

void oracle_meng_uvs_eval(float u, float v, float scale, const float lambdas[4], float out[4]) {
    Meng m = Meng::fromUVS(u, v, scale);
    WLS<4> w;
    for (int i = 0; i < 4; ++i) w.lambdas[i] = lambdas[i];
    w.sel = 0;
    w.single = false;
    SS<4> r = m.evaluate(w);
    for (int i = 0; i < 4; ++i) out[i] = r[i];
}

void oracle_cmf_bins(float x[16], float y[16], float z[16], float *integral) {
    for (int i = 0; i < 16; ++i) { x[i] = g_tables.binX[i]; y[i] = g_tables.binY[i]; z[i] = g_tables.binZ[i]; }
    *integral = g_tables.integralCMF16;
}

void oracle_triplet_rgb(int spectrumType, int colorSpace, float e0, float e1, float e2, float out[3]) {
    Triplet t = Triplet::create(static_cast<SpectrumType>(spectrumType), static_cast<ColorSpace>(colorSpace), e0, e1, e2);
    out[0] = t.rgb[0]; out[1] = t.rgb[1]; out[2] = t.rgb[2];
}

} // extern "C"
