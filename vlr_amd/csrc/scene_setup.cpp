// Scene flattening and device upload (replaces Scene::prepareSetup / Scene::setup and the
// OptiX GAS/IAS builds of libVLR/scene.cpp:1149-1402).
//
//   instances      one per parent node that holds surface nodes, depth-first in child order,
//                  with the composed local-to-world transform (scene.cpp:1234-1240)
//   geometry inst. one per (surface node, material group), importance 1 for emitters
//                  (scene.cpp:284, 432, 508)
//   light dists    per instance over its geometry instances, scene-wide over
//                  [environment, instances...] (scene.cpp:1249-1281, 1380-1392); prim
//                  distributions over triangle areas for emitters (scene.cpp:209-219)
//   scene bounds   union of transformed instance boxes (setup_scene.cu:7-92)
//   BVHs           binned-SAH BVH2 per geometry group (BLAS) and over instances (TLAS),
//                  collapsed into 4-wide nodes
//   environment    RGBA16F texels (RGB build) or (u, v, s) halves (spectral build,
//                  image.cpp:542-565), importance map = 4x box-shrunk luminance times
//                  sin(theta) (shader_nodes.cpp:1992-2010)
#include "runtime.h"

#include <cmath>
#include <limits>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>

namespace vlra {

DeviceBuffer::~DeviceBuffer() { release(); }
void DeviceBuffer::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}
void DeviceBuffer::reserve(size_t n) {
    if (n <= bytes && ptr) return;
    release();
    size_t alloc = std::max<size_t>(n, 256);
    VLRA_HIP(hipMalloc(&ptr, alloc));
    bytes = alloc;
}
void DeviceBuffer::upload(const void *src, size_t n) {
    reserve(n);
    if (n) VLRA_HIP(hipMemcpy(ptr, src, n, hipMemcpyHostToDevice));
}

void throwHip(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const float *p) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    void grow(const Box &b) {
        grow(b.lo);
        grow(b.hi);
    }
    bool empty() const { return lo[0] > hi[0]; }
    float area() const {
        if (empty()) return 0.0f;
        float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
        return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
    }
    float centroid(int a) const { return 0.5f * (lo[a] + hi[a]); }
};

// Binned-SAH BVH2 builder writing the 64-byte node layout.
class BvhBuilder {
public:
    BvhBuilder(const std::vector<Box> &boxes, uint32_t maxLeaf) : boxes_(boxes), maxLeaf_(maxLeaf) {
        order.resize(boxes.size());
        for (size_t i = 0; i < boxes.size(); ++i) order[i] = static_cast<uint32_t>(i);
        cent_.resize(boxes.size() * 3);
        for (size_t i = 0; i < boxes.size(); ++i)
            for (int a = 0; a < 3; ++a) cent_[3 * i + a] = boxes[i].centroid(a);
    }
    std::vector<BVHNode> nodes;
    std::vector<uint32_t> order;
    uint32_t maxDepth = 0;

    void build() {
        nodes.clear();
        nodes.reserve(boxes_.size() + 1);
        nodes.push_back(BVHNode());
        uint32_t n = static_cast<uint32_t>(boxes_.size());
        if (n <= maxLeaf_) {
            // root with a single leaf child; child index 0 (the root) marks an empty slot
            Box b;
            for (uint32_t i = 0; i < n; ++i) b.grow(boxes_[i]);
            setChild(0, 0, b, n ? leafCode(0, n) : 0);
            setChild(0, 1, Box(), 0);
            return;
        }
        splitInto(0, 0, n, 1);
    }

private:
    const std::vector<Box> &boxes_;
    uint32_t maxLeaf_;
    std::vector<float> cent_;

    static int32_t leafCode(uint32_t first, uint32_t count) {
        return static_cast<int32_t>(~((first << kLeafShift) | (count - 1)));
    }
    void setChild(uint32_t node, int slot, const Box &b, int32_t code) {
        BVHNode &N = nodes[node];
        float *lo = slot == 0 ? N.lo0 : N.lo1;
        float *hi = slot == 0 ? N.hi0 : N.hi1;
        for (int a = 0; a < 3; ++a) {
            lo[a] = b.empty() ? 0.0f : b.lo[a];
            hi[a] = b.empty() ? 0.0f : b.hi[a];
        }
        N.child[slot] = code;
        N.pad[0] = N.pad[1] = 0;
    }
    Box rangeBox(uint32_t b, uint32_t e) const {
        Box r;
        for (uint32_t i = b; i < e; ++i) r.grow(boxes_[order[i]]);
        return r;
    }
    uint32_t partition(uint32_t b, uint32_t e, uint32_t depth) {
        const uint32_t n = e - b;
        float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (uint32_t i = b; i < e; ++i)
            for (int a = 0; a < 3; ++a) {
                clo[a] = std::min(clo[a], cent_[3 * order[i] + a]);
                chi[a] = std::max(chi[a], cent_[3 * order[i] + a]);
            }
        constexpr int NB = 16;
        float best = FLT_MAX;
        int bestAxis = -1, bestSplit = -1;
        if (depth < 40) {
            for (int a = 0; a < 3; ++a) {
                float ext = chi[a] - clo[a];
                if (!(ext > 0.0f)) continue;
                Box bb[NB];
                uint32_t bc[NB] = {};
                for (uint32_t i = b; i < e; ++i) {
                    int k = std::min(NB - 1, static_cast<int>(NB * (cent_[3 * order[i] + a] - clo[a]) / ext));
                    bb[k].grow(boxes_[order[i]]);
                    ++bc[k];
                }
                Box left[NB];
                uint32_t nl[NB];
                Box acc;
                uint32_t cnt = 0;
                for (int k = 0; k < NB; ++k) {
                    if (bc[k]) acc.grow(bb[k]);
                    cnt += bc[k];
                    left[k] = acc;
                    nl[k] = cnt;
                }
                Box racc;
                uint32_t rc = 0;
                for (int k = NB - 1; k >= 1; --k) {
                    if (bc[k]) racc.grow(bb[k]);
                    rc += bc[k];
                    uint32_t lcount = nl[k - 1];
                    if (!lcount || !rc) continue;
                    float c = left[k - 1].area() * lcount + racc.area() * rc;
                    if (c < best) {
                        best = c;
                        bestAxis = a;
                        bestSplit = k;
                    }
                }
            }
        }
        uint32_t mid = b + n / 2;
        if (bestAxis >= 0) {
            float ext = chi[bestAxis] - clo[bestAxis];
            auto it = std::partition(order.begin() + b, order.begin() + e, [&](uint32_t p) {
                int k = std::min(NB - 1, static_cast<int>(NB * (cent_[3 * p + bestAxis] - clo[bestAxis]) / ext));
                return k < bestSplit;
            });
            mid = static_cast<uint32_t>(it - order.begin());
        }
        if (mid == b || mid == e) {
            // median split on the widest centroid axis
            int a = 0;
            for (int k = 1; k < 3; ++k)
                if (chi[k] - clo[k] > chi[a] - clo[a]) a = k;
            mid = b + n / 2;
            std::nth_element(order.begin() + b, order.begin() + mid, order.begin() + e,
                             [&](uint32_t x, uint32_t y) { return cent_[3 * x + a] < cent_[3 * y + a]; });
        }
        return mid;
    }
    // node `idx` gets children for [b, m) and [m, e)
    void splitInto(uint32_t idx, uint32_t b, uint32_t e, uint32_t depth) {
        maxDepth = std::max(maxDepth, depth);
        uint32_t m = partition(b, e, depth);
        uint32_t ranges[2][2] = {{b, m}, {m, e}};
        for (int s = 0; s < 2; ++s) {
            uint32_t rb = ranges[s][0], re = ranges[s][1];
            Box box = rangeBox(rb, re);
            if (re - rb <= maxLeaf_) {
                setChild(idx, s, box, leafCode(rb, re - rb));
            } else {
                uint32_t child = static_cast<uint32_t>(nodes.size());
                nodes.push_back(BVHNode());
                setChild(idx, s, box, static_cast<int32_t>(child));
                splitInto(child, rb, re, depth + 1);
            }
        }
    }
};

// Collapses a BVH2 into 4-wide nodes: starting from a node's two children, the inner child
// with the largest box is replaced by its own two children until four slots are filled or
// only leaves remain.  Leaf codes and child order carry over.
class Bvh4Collapser {
public:
    explicit Bvh4Collapser(const std::vector<BVHNode> &bin) : bin_(bin) {}
    std::vector<BVH4Node> nodes;
    void run() {
        nodes.clear();
        if (bin_.empty()) return;
        emit(0);
    }

private:
    struct Slot {
        Box box;
        int32_t code;
    };
    const std::vector<BVHNode> &bin_;
    Slot slotOf(uint32_t n, int s) const {
        const BVHNode &N = bin_[n];
        Slot r;
        const float *lo = s == 0 ? N.lo0 : N.lo1, *hi = s == 0 ? N.hi0 : N.hi1;
        for (int a = 0; a < 3; ++a) {
            r.box.lo[a] = lo[a];
            r.box.hi[a] = hi[a];
        }
        r.code = N.child[s];
        return r;
    }
    uint32_t emit(uint32_t binNode) {
        uint32_t idx = static_cast<uint32_t>(nodes.size());
        nodes.push_back(BVH4Node());
        Slot slots[4];
        int n = 0;
        for (int s = 0; s < 2; ++s) {
            Slot c = slotOf(binNode, s);
            if (c.code != 0) slots[n++] = c;
        }
        while (n < 4) {
            int best = -1;
            float bestArea = -1.0f;
            for (int k = 0; k < n; ++k)
                if (slots[k].code > 0 && slots[k].box.area() > bestArea) {
                    best = k;
                    bestArea = slots[k].box.area();
                }
            if (best < 0) break;
            uint32_t inner = static_cast<uint32_t>(slots[best].code);
            Slot a = slotOf(inner, 0), b = slotOf(inner, 1);
            // keep the order: the replaced slot's children take its place
            Slot tmp[4];
            int m = 0;
            for (int k = 0; k < n; ++k) {
                if (k == best) {
                    if (a.code != 0) tmp[m++] = a;
                    if (b.code != 0) tmp[m++] = b;
                } else {
                    tmp[m++] = slots[k];
                }
            }
            for (int k = 0; k < m; ++k) slots[k] = tmp[k];
            n = m;
        }
        int32_t codes[4] = {0, 0, 0, 0};
        for (int k = 0; k < n; ++k) codes[k] = slots[k].code > 0 ? static_cast<int32_t>(emit(static_cast<uint32_t>(slots[k].code)))
                                                                  : slots[k].code;
        BVH4Node &N = nodes[idx];
        for (int k = 0; k < 4; ++k) {
            const bool used = k < n;
            N.loX[k] = used ? slots[k].box.lo[0] : 0.0f;
            N.loY[k] = used ? slots[k].box.lo[1] : 0.0f;
            N.loZ[k] = used ? slots[k].box.lo[2] : 0.0f;
            N.hiX[k] = used ? slots[k].box.hi[0] : 0.0f;
            N.hiY[k] = used ? slots[k].box.hi[1] : 0.0f;
            N.hiZ[k] = used ? slots[k].box.hi[2] : 0.0f;
            N.child[k] = codes[k];
            N.pad[k] = 0;
        }
        return idx;
    }
};

std::vector<BVH4Node> collapseToBvh4(const std::vector<BVHNode> &bin) {
    Bvh4Collapser c(bin);
    c.run();
    return std::move(c.nodes);
}

// One quantised node: per axis the grid step is the power of two 2^k with 255 * 2^k >= the
// node extent; each child plane is rounded outwards and checked with the decoding arithmetic
// the traversal uses (origin + float(q) * step, no fused multiply-add).
BVH4QNode quantizeNode(const BVH4Node &f) {
    BVH4QNode q;
    std::memset(&q, 0, sizeof(q));
    const float *lo[3] = {f.loX, f.loY, f.loZ}, *hi[3] = {f.hiX, f.hiY, f.hiZ};
    for (int k = 0; k < 4; ++k) q.child[k] = f.child[k];
    for (int a = 0; a < 3; ++a) {
        float blo = std::numeric_limits<float>::infinity(), bhi = -blo;
        for (int k = 0; k < 4; ++k)
            if (f.child[k] != 0) {
                blo = std::min(blo, lo[a][k]);
                bhi = std::max(bhi, hi[a][k]);
            }
        if (!(blo <= bhi)) blo = bhi = 0.0f;   // no child
        q.origin[a] = blo;
        int e;
        std::frexp(std::max((bhi - blo) / 255.0f, 1e-30f), &e);   // (ext / 255) < 2^e
        e = std::max(e, -126);
        for (;; ++e) {
            const float step = std::ldexp(1.0f, e);
            uint32_t wl = 0, wh = 0;
            bool ok = true;
            for (int k = 0; k < 4; ++k) {
                if (f.child[k] == 0) continue;
                float ql = std::floor((lo[a][k] - blo) / step), qh = std::ceil((hi[a][k] - blo) / step);
                int il = static_cast<int>(std::min(std::max(ql, 0.0f), 255.0f));
                int ih = static_cast<int>(std::min(std::max(qh, 0.0f), 255.0f));
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

std::vector<BVH4QNode> quantizeBvh4(const std::vector<BVH4Node> &nodes) {
    std::vector<BVH4QNode> out(nodes.size());
    for (size_t i = 0; i < nodes.size(); ++i) out[i] = quantizeNode(nodes[i]);
    return out;
}

void composeTransform(const float a[16], const float b[16], float out[16]) {
    // column-major 4x4 product out = a * b
    float r[16];
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += a[k * 4 + row] * b[c * 4 + k];
            r[c * 4 + row] = s;
        }
    std::memcpy(out, r, sizeof(r));
}

struct FlatInstance {
    float xform[16];
    std::vector<const SurfaceNode *> surfaces;
};

void flatten(const ParentNode &p, const float parent[16], std::vector<FlatInstance> *out) {
    float xf[16];
    if (p.transform) composeTransform(parent, p.transform->mat, xf);
    else std::memcpy(xf, parent, sizeof(xf));
    FlatInstance fi;
    std::memcpy(fi.xform, xf, sizeof(xf));
    for (const Node *c : p.children)
        if (c->isSurfaceNode()) fi.surfaces.push_back(static_cast<const SurfaceNode *>(c));
    if (!fi.surfaces.empty()) out->push_back(std::move(fi));
    for (const Node *c : p.children)
        if (c->isParentNode()) flatten(*static_cast<const ParentNode *>(c), xf, out);
}

// DiscreteDistribution1D::initialize (context.cpp:1797-1822): PMF then CDF into the pool
float appendDiscrete(std::vector<float> &pool, const std::vector<float> &values) {
    const size_t n = values.size();
    size_t base = pool.size();
    pool.resize(base + 2 * n + 1);
    float *pmf = pool.data() + base, *cdf = pmf + n;
    std::copy(values.begin(), values.end(), pmf);
    float sum = 0.0f, comp = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        cdf[i] = sum;
        float cIn = pmf[i] - comp, t = sum + cIn;
        comp = (t - sum) - cIn;
        sum = t;
    }
    for (size_t i = 0; i < n; ++i) {
        pmf[i] /= sum;
        cdf[i] /= sum;
    }
    cdf[n] = 1.0f;
    return sum;
}

// RegularConstantContinuousDistribution1D::initialize (context.cpp:1846-1868)
float appendContinuous(std::vector<float> &pool, const float *values, size_t n) {
    size_t base = pool.size();
    pool.resize(base + 2 * n + 1);
    float *pdf = pool.data() + base, *cdf = pdf + n;
    std::copy(values, values + n, pdf);
    float sum = 0.0f, comp = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        cdf[i] = sum;
        float cIn = pdf[i] / n - comp, t = sum + cIn;
        comp = (t - sum) - cIn;
        sum = t;
    }
    for (size_t i = 0; i < n; ++i) {
        pdf[i] /= sum;
        cdf[i] /= sum;
    }
    cdf[n] = 1.0f;
    return sum;
}

} // namespace

void buildSceneGPU(Context &ctx, const Scene &scene, bool spectral, SceneGPU *g) {
    if (!loadSpectralTables()) throw std::runtime_error("spectral tables not found next to the library");
    std::vector<FlatInstance> flat;
    float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    flatten(scene, ident, &flat);

    // materials
    std::vector<DevMaterial> mats;
    std::vector<float> cacheMats;   // 8 floats per material
    std::map<const SurfaceMaterial *, uint32_t> matIndex;
    const Image2D *envImage = nullptr;
    auto matOf = [&](const SurfaceMaterial *m) -> uint32_t {
        auto it = matIndex.find(m);
        if (it != matIndex.end()) return it->second;
        if (!m->supportedOnRenderPath())
            throw ApiError{4 /* IncompatibleNodeType */,
                           std::string("material type not on the render path: ") + m->typeName()};
        bool usesTex;
        const Image2D *img;
        DevMaterial d = m->toDevice(spectral, &usesTex, &img);
        if (usesTex) envImage = img;
        uint32_t idx = static_cast<uint32_t>(mats.size());
        mats.push_back(d);
        float cf[8];
        m->cacheFeatures(cf);
        cacheMats.insert(cacheMats.end(), cf, cf + 8);
        matIndex[m] = idx;
        return idx;
    };

    // geometry instances, vertices, indices
    std::vector<DevGeomInst> geoms;
    std::vector<DevVertex> verts;
    std::vector<uint32_t> indices;
    std::vector<float> pool;
    std::map<std::pair<const SurfaceNode *, uint32_t>, uint32_t> geomIndex;
    std::map<const SurfaceNode *, uint32_t> vertexBase;
    auto geomOf = [&](const SurfaceNode *s, uint32_t gIdx) -> uint32_t {
        auto key = std::make_pair(s, gIdx);
        auto it = geomIndex.find(key);
        if (it != geomIndex.end()) return it->second;
        auto vb = vertexBase.find(s);
        uint32_t vbase;
        if (vb == vertexBase.end()) {
            vbase = static_cast<uint32_t>(verts.size());
            for (const VLRVertex &v : s->vertices) {
                DevVertex d;
                d.p[0] = v.position.x; d.p[1] = v.position.y; d.p[2] = v.position.z; d.u = v.texCoord.u;
                d.n[0] = v.normal.x; d.n[1] = v.normal.y; d.n[2] = v.normal.z; d.v = v.texCoord.v;
                d.t[0] = v.tc0Direction.x; d.t[1] = v.tc0Direction.y; d.t[2] = v.tc0Direction.z; d.pad = 0;
                verts.push_back(d);
            }
            vertexBase[s] = vbase;
        } else {
            vbase = vb->second;
        }
        const MaterialGroup &mg = s->groups[gIdx];
        DevGeomInst G;
        std::memset(&G, 0, sizeof(G));
        G.material = matOf(mg.material);
        G.vertexBase = vbase;
        G.indexBase = static_cast<uint32_t>(indices.size());
        G.isPoints = s->type == ObjType::PointSurfaceNode ? 1 : 0;
        G.numPrims = G.isPoints ? static_cast<uint32_t>(mg.indices.size()) : static_cast<uint32_t>(mg.indices.size() / 3);
        for (uint32_t k = 0; k < (G.isPoints ? G.numPrims : 3 * G.numPrims); ++k) {
            if (mg.indices[k] >= s->vertices.size())
                throw ApiError{3 /* InvalidArgument */, "vertex index out of range in a material group"};
            indices.push_back(mg.indices[k]);
        }
        bool emit = mg.material->isEmitting();
        G.importance = emit ? 1.0f : 0.0f;
        if (emit && G.numPrims) {
            std::vector<float> w(G.numPrims, 1.0f);
            if (!G.isPoints) {
                for (uint32_t t = 0; t < G.numPrims; ++t) {
                    const VLRVertex &a = s->vertices[mg.indices[3 * t]];
                    const VLRVertex &b = s->vertices[mg.indices[3 * t + 1]];
                    const VLRVertex &c = s->vertices[mg.indices[3 * t + 2]];
                    float e1[3] = {b.position.x - a.position.x, b.position.y - a.position.y, b.position.z - a.position.z};
                    float e2[3] = {c.position.x - a.position.x, c.position.y - a.position.y, c.position.z - a.position.z};
                    float cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                          cz = e1[0] * e2[1] - e1[1] * e2[0];
                    w[t] = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
                }
            }
            G.distBase = static_cast<uint32_t>(pool.size());
            G.distIntegral = appendDiscrete(pool, w);
        }
        uint32_t idx = static_cast<uint32_t>(geoms.size());
        geoms.push_back(G);
        geomIndex[key] = idx;
        return idx;
    };

    // instances + BLAS per distinct surface-node list
    std::vector<DevInstance> insts;
    std::vector<uint32_t> instGeoms;
    std::vector<BVH4Node> blasNodes;
    std::vector<TriRecord> blasTris;
    std::map<std::vector<const SurfaceNode *>, std::pair<uint32_t, uint32_t>> blasCache;
    std::vector<Box> instBoxes;
    std::vector<uint32_t> tlasInst;
    Box sceneBox;
    std::vector<float> lightImp;
    lightImp.push_back(0.0f);   // environment slot, set below
    uint64_t numTris = 0;
    for (const FlatInstance &fi : flat) {
        DevInstance I;
        std::memset(&I, 0, sizeof(I));
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) I.o2w[r * 4 + c] = fi.xform[c * 4 + r];
        affineInverse(fi.xform, I.w2o);
        I.geomInstListBase = static_cast<uint32_t>(instGeoms.size());
        std::vector<float> imp;
        std::vector<std::pair<uint32_t, const SurfaceNode *>> triGeoms;   // (global geomInst, node)
        Box child;
        for (const SurfaceNode *s : fi.surfaces) {
            for (uint32_t gi = 0; gi < s->groups.size(); ++gi) {
                uint32_t g = geomOf(s, gi);
                instGeoms.push_back(g);
                imp.push_back(geoms[g].importance);
                if (!geoms[g].isPoints) {
                    triGeoms.emplace_back(g, s);
                    for (uint32_t k : s->groups[gi].indices) {
                        const VLRVertex &v = s->vertices[k];
                        child.grow(&v.position.x);
                    }
                }
            }
        }
        I.numGeomInsts = static_cast<uint32_t>(imp.size());
        float sumImp = 0;
        for (float v : imp) sumImp += v;
        if (!imp.empty()) {
            I.lightDistBase = static_cast<uint32_t>(pool.size());
            I.lightDistIntegral = appendDiscrete(pool, imp);
        }
        lightImp.push_back(sumImp > 0.0f ? 1.0f : 0.0f);
        uint32_t instIdx = static_cast<uint32_t>(insts.size());
        if (!triGeoms.empty()) {
            auto cacheIt = blasCache.find(fi.surfaces);
            if (cacheIt == blasCache.end()) {
                std::vector<Box> boxes;
                std::vector<TriRecord> recs;
                for (auto &tg : triGeoms) {
                    const DevGeomInst &G = geoms[tg.first];
                    for (uint32_t t = 0; t < G.numPrims; ++t) {
                        TriRecord r;
                        std::memset(&r, 0, sizeof(r));
                        const DevVertex &a = verts[G.vertexBase + indices[G.indexBase + 3 * t]];
                        const DevVertex &b = verts[G.vertexBase + indices[G.indexBase + 3 * t + 1]];
                        const DevVertex &c = verts[G.vertexBase + indices[G.indexBase + 3 * t + 2]];
                        std::memcpy(r.v0, a.p, 12);
                        std::memcpy(r.v1, b.p, 12);
                        std::memcpy(r.v2, c.p, 12);
                        r.geomInst = tg.first;
                        r.prim = t;
                        Box bb;
                        bb.grow(a.p);
                        bb.grow(b.p);
                        bb.grow(c.p);
                        boxes.push_back(bb);
                        recs.push_back(r);
                    }
                }
                BvhBuilder bld(boxes, 4);
                bld.build();
                uint32_t nodeBase = static_cast<uint32_t>(blasNodes.size());
                uint32_t triBase = static_cast<uint32_t>(blasTris.size());
                std::vector<BVH4Node> wide = collapseToBvh4(bld.nodes);
                blasNodes.insert(blasNodes.end(), wide.begin(), wide.end());
                for (uint32_t o : bld.order) blasTris.push_back(recs[o]);
                numTris += recs.size();
                cacheIt = blasCache.emplace(fi.surfaces, std::make_pair(nodeBase, triBase)).first;
            } else {
                // shared geometry group: same BLAS
            }
            I.blasNodeBase = cacheIt->second.first;
            I.blasTriBase = cacheIt->second.second;
            I.hasGeometry = 1;
            // world box of the instance (8 transformed corners)
            Box wb;
            for (int c = 0; c < 8; ++c) {
                float p[3] = {(c & 1) ? child.hi[0] : child.lo[0], (c & 2) ? child.hi[1] : child.lo[1],
                              (c & 4) ? child.hi[2] : child.lo[2]};
                float q[3];
                for (int r = 0; r < 3; ++r)
                    q[r] = I.o2w[r * 4 + 0] * p[0] + I.o2w[r * 4 + 1] * p[1] + I.o2w[r * 4 + 2] * p[2] + I.o2w[r * 4 + 3];
                wb.grow(q);
            }
            instBoxes.push_back(wb);
            tlasInst.push_back(instIdx);
            sceneBox.grow(wb);
        }
        insts.push_back(I);
    }

    // TLAS (one instance per leaf); leaf code carries the instance index
    std::vector<BVH4Node> tlasNodes;
    if (!instBoxes.empty()) {
        BvhBuilder bld(instBoxes, 1);
        bld.build();
        std::vector<BVHNode> bin = bld.nodes;
        for (BVHNode &n : bin) {
            for (int s = 0; s < 2; ++s) {
                if (n.child[s] < 0) {
                    uint32_t first = (~static_cast<uint32_t>(n.child[s])) >> kLeafShift;
                    n.child[s] = static_cast<int32_t>(~(tlasInst[bld.order[first]] << kLeafShift));
                }
            }
        }
        tlasNodes = collapseToBvh4(bin);
    }

    // environment
    DevScene &D = g->dev;
    std::memset(&D, 0, sizeof(D));
    std::vector<float4> envTex;
    if (scene.envMaterial) {
        uint32_t em = matOf(scene.envMaterial);
        D.hasEnv = 1;
        D.envMaterial = em;
        D.envImportance = scene.envMaterial->isEmitting() ? 1.0f : 0.0f;
        D.envRotation = scene.envRotation;
        lightImp[0] = D.envImportance > 0.0f ? 1.0f : 0.0f;
        std::vector<float> rgba;
        uint32_t W = 0, H = 0;
        if (mats[em].useEnvTexture && envImage && envImage->toRGBA32F(&rgba)) {
            W = envImage->width;
            H = envImage->height;
        }
        if (W && H) {
            envTex.resize(size_t(W) * H);
            for (size_t p = 0; p < envTex.size(); ++p) {
                float RGB[3] = {halfToFloat(floatToHalf(rgba[4 * p])), halfToFloat(floatToHalf(rgba[4 * p + 1])),
                                halfToFloat(floatToHalf(rgba[4 * p + 2]))};
                if (!spectral) {
                    envTex[p] = make_float4(RGB[0], RGB[1], RGB[2], 1.0f);
                } else {
                    const float m[9] = {0.4124564f, 0.2126729f, 0.0193339f, 0.3575761f, 0.7151522f,
                                        0.1191920f, 0.1804375f, 0.0721750f, 0.9503041f};
                    float X = m[0] * RGB[0] + m[3] * RGB[1] + m[6] * RGB[2];
                    float Y = m[1] * RGB[0] + m[4] * RGB[1] + m[7] * RGB[2];
                    float Z = m[2] * RGB[0] + m[5] * RGB[1] + m[8] * RGB[2];
                    float b = X + Y + Z;
                    float xy[2] = {b > 0.0f ? X / b : (1.0f / 3.0f), b > 0.0f ? Y / b : (1.0f / 3.0f)};
                    float uv[2];
                    mengXYtoUV(xy, uv);
                    envTex[p] = make_float4(halfToFloat(floatToHalf(uv[0])), halfToFloat(floatToHalf(uv[1])),
                                            halfToFloat(floatToHalf(b)), 1.0f);
                }
            }
            uint32_t mw = std::max<uint32_t>(1, W / 4), mh = std::max<uint32_t>(1, H / 4);
            uint32_t fx = W / mw, fy = H / mh;
            std::vector<float> lum(size_t(mw) * mh);
            for (uint32_t y = 0; y < mh; ++y) {
                float theta = 3.14159265358979323846f * (y + 0.5f) / mh;
                for (uint32_t x = 0; x < mw; ++x) {
                    float acc[3] = {0, 0, 0};
                    for (uint32_t yy = 0; yy < fy; ++yy)
                        for (uint32_t xx = 0; xx < fx; ++xx) {
                            const float4 &t = envTex[size_t(y * fy + yy) * W + x * fx + xx];
                            acc[0] += t.x;
                            acc[1] += t.y;
                            acc[2] += t.z;
                        }
                    for (int c = 0; c < 3; ++c) acc[c] /= float(fx * fy);
                    float Yl;
                    if (!spectral)
                        Yl = 0.2126729f * acc[0] + 0.7151522f * acc[1] + 0.0721750f * acc[2];
                    else
                        Yl = (0.022853819546830627f * acc[0] + 0.05077639329371236f * acc[1] - 0.006895157122499944f) * acc[2];
                    lum[size_t(y) * mw + x] = Yl * std::sin(theta);
                }
            }
            D.envW = W;
            D.envH = H;
            D.envMapW = mw;
            D.envMapH = mh;
            D.envRowBase = static_cast<uint32_t>(pool.size());
            std::vector<float> rowIntegrals(mh);
            for (uint32_t y = 0; y < mh; ++y) rowIntegrals[y] = appendContinuous(pool, lum.data() + size_t(y) * mw, mw);
            D.envTopBase = static_cast<uint32_t>(pool.size());
            appendContinuous(pool, rowIntegrals.data(), mh);
        } else {
            float one = 1.0f;
            D.envMapW = D.envMapH = 1;
            D.envRowBase = static_cast<uint32_t>(pool.size());
            appendContinuous(pool, &one, 1);
            D.envTopBase = static_cast<uint32_t>(pool.size());
            appendContinuous(pool, &one, 1);
        }
    }
    D.lightInstBase = static_cast<uint32_t>(pool.size());
    D.numLightInsts = static_cast<uint32_t>(lightImp.size());
    D.lightInstIntegral = appendDiscrete(pool, lightImp);

    // scene bounds (setup_scene.cu:82-92)
    if (!sceneBox.empty()) {
        for (int a = 0; a < 3; ++a) D.center[a] = 0.5f * (sceneBox.lo[a] + sceneBox.hi[a]);
        float d0 = sceneBox.hi[0] - D.center[0], d1 = sceneBox.hi[1] - D.center[1], d2 = sceneBox.hi[2] - D.center[2];
        D.worldRadius = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        D.worldDiscArea = 3.14159265358979323846f * (D.worldRadius * D.worldRadius);
    }

    // upload
    const std::vector<BVH4QNode> tlasQ = quantizeBvh4(tlasNodes), blasQ = quantizeBvh4(blasNodes);
    g->tlasNodes.upload(tlasQ.data(), tlasQ.size() * sizeof(BVH4QNode));
    g->blasNodes.upload(blasQ.data(), blasQ.size() * sizeof(BVH4QNode));
    g->blasTris.upload(blasTris.data(), blasTris.size() * sizeof(TriRecord));
    g->instances.upload(insts.data(), insts.size() * sizeof(DevInstance));
    g->instGeomInsts.upload(instGeoms.data(), instGeoms.size() * sizeof(uint32_t));
    g->geomInsts.upload(geoms.data(), geoms.size() * sizeof(DevGeomInst));
    g->vertices.upload(verts.data(), verts.size() * sizeof(DevVertex));
    g->indices.upload(indices.data(), indices.size() * sizeof(uint32_t));
    g->distPool.upload(pool.data(), pool.size() * sizeof(float));
    g->materials.upload(mats.data(), mats.size() * sizeof(DevMaterial));
    g->cacheMaterials.upload(cacheMats.data(), cacheMats.size() * sizeof(float));
    g->envTex.upload(envTex.data(), envTex.size() * sizeof(float4));
    g->mengGrid.upload(mengGrid().data(), mengGrid().size() * sizeof(MengCellDev));
    g->mengPoints.upload(mengPoints().data(), mengPoints().size() * sizeof(float));

    D.tlasNodes = static_cast<const BVH4QNode *>(g->tlasNodes.ptr);
    D.blasNodes = static_cast<const BVH4QNode *>(g->blasNodes.ptr);
    D.blasTris = static_cast<const TriRecord *>(g->blasTris.ptr);
    D.instances = static_cast<const DevInstance *>(g->instances.ptr);
    D.instGeomInsts = static_cast<const uint32_t *>(g->instGeomInsts.ptr);
    D.geomInsts = static_cast<const DevGeomInst *>(g->geomInsts.ptr);
    D.vertices = static_cast<const DevVertex *>(g->vertices.ptr);
    D.indices = static_cast<const uint32_t *>(g->indices.ptr);
    D.distPool = static_cast<const float *>(g->distPool.ptr);
    D.materials = static_cast<const DevMaterial *>(g->materials.ptr);
    D.cacheMaterials = static_cast<const float4 *>(g->cacheMaterials.ptr);
    D.envTex = envTex.empty() ? nullptr : static_cast<const float4 *>(g->envTex.ptr);
    D.mengGrid = static_cast<const MengCellDev *>(g->mengGrid.ptr);
    D.mengPoints = static_cast<const float *>(g->mengPoints.ptr);
    D.numInstances = static_cast<uint32_t>(insts.size());
    D.hasTlas = tlasNodes.empty() ? 0 : 1;
    cmfBins(D.cmfX, D.cmfY, D.cmfZ, &D.integralCMF);
    g->spectral = spectral;
    g->valid = true;
    g->numInstances = static_cast<uint32_t>(insts.size());
    g->numGeomInsts = static_cast<uint32_t>(geoms.size());
    g->numTriangles = numTris;
    (void)ctx;
}

} // namespace vlra
