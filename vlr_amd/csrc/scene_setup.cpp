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
#include "bvh.h"

#include <chrono>
#include <tuple>

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

// TLAS refit: the 4-wide tree over the instances keeps its topology; each node's child boxes
// are recomputed bottom-up from the instances' world boxes and quantised again.  Returns node
// n's box and adds the areas of the inner nodes' boxes to *sumArea (the tree's SAH-style cost
// up to a constant).  Leaf codes carry the instance index.
static Box refitTlasNode(std::vector<BVH4QNode> &nodes, uint32_t n, const std::vector<Box> &boxOfInst, double *sumArea) {
    BVH4Node f;
    std::memset(&f, 0, sizeof(f));
    Box all;
    for (int k = 0; k < 4; ++k) {
        const int32_t code = nodes[n].child[k];
        f.child[k] = code;
        if (code == 0) continue;
        Box b = code > 0 ? refitTlasNode(nodes, static_cast<uint32_t>(code), boxOfInst, sumArea)
                         : boxOfInst[static_cast<uint32_t>(~code) >> kLeafShift];
        f.loX[k] = bvhNoNegZero(b.lo[0]);
        f.loY[k] = bvhNoNegZero(b.lo[1]);
        f.loZ[k] = bvhNoNegZero(b.lo[2]);
        f.hiX[k] = bvhNoNegZero(b.hi[0]);
        f.hiY[k] = bvhNoNegZero(b.hi[1]);
        f.hiZ[k] = bvhNoNegZero(b.hi[2]);
        all.grow(b);
    }
    nodes[n] = quantizeBvh4Node(f);
    *sumArea += all.area();
    return all;
}

// relative cost of a TLAS over the current instance boxes (rewrites `nodes` with refitted ones)
static double refitTlas(std::vector<BVH4QNode> &nodes, const std::vector<Box> &boxOfInst) {
    double sum = 0.0;
    const Box root = refitTlasNode(nodes, 0, boxOfInst, &sum);
    return root.area() > 0.0f ? sum / root.area() : sum;
}

void buildSceneGPU(Context &ctx, const Scene &scene, bool spectral, SceneGPU *g) {
    if (!loadSpectralTables()) throw std::runtime_error("spectral tables not found next to the library");
    std::vector<FlatInstance> flat;
    float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    flatten(scene, ident, &flat);

    // shader nodes reachable from materials and material groups
    ShaderTable shaders;
    bool hasAlpha = false;
    shaders.spectral = spectral;

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
        std::vector<std::pair<PlugRef, uint32_t>> roots;
        DevMaterial d = m->toDevice(spectral, &usesTex, &img, &roots);
        if (!roots.empty()) std::tie(d.progBase, d.progCount) = shaders.program(roots);
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
        if (!G.isPoints && (mg.nodeNormal.valid() || mg.nodeTangent.valid())) {
            // normal (output vector 0) and tangent (vector 1) nodes; alpha nodes are not evaluated
            std::vector<std::pair<PlugRef, uint32_t>> roots;
            if (mg.nodeNormal.valid()) roots.push_back({mg.nodeNormal, 8u});
            if (mg.nodeTangent.valid()) roots.push_back({mg.nodeTangent, 16u});
            std::tie(G.progBase, G.progCount) = shaders.program(roots);
        }
        if (!G.isPoints && mg.nodeAlpha.valid()) {
            // the alpha node (renderer_common.h:532-548), its own program: traversal runs it at
            // candidate hits
            std::tie(G.alphaBase, G.alphaCount) = shaders.program({{mg.nodeAlpha, 8u}});
            hasAlpha = true;
        }
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
    struct TriGeom {
        uint32_t geomInst, group;
        const SurfaceNode *node;
    };
    std::vector<DevInstance> insts;
    std::vector<uint32_t> instGeoms;
    // distinct geometry groups: their triangle records, in the order of first use
    struct BlasInput {
        uint32_t triBase, numTris, nodeBase;
    };
    std::vector<BlasInput> blasList;
    std::vector<std::vector<TriGeom>> blasGroups;   // per BLAS: its geometry groups
    std::vector<uint32_t> instBlas;   // per instance: index into blasList or ~0
    std::map<std::vector<const SurfaceNode *>, uint32_t> blasCache;
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
        std::vector<TriGeom> triGeoms;
        Box child;
        for (const SurfaceNode *s : fi.surfaces) {
            for (uint32_t gi = 0; gi < s->groups.size(); ++gi) {
                uint32_t g = geomOf(s, gi);
                instGeoms.push_back(g);
                imp.push_back(geoms[g].importance);
                if (!geoms[g].isPoints) {
                    triGeoms.push_back({g, gi, s});
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
        uint32_t blasIdx = ~0u;
        uint64_t instTris = 0;   // material groups without triangles make no BLAS
        for (auto &tg : triGeoms) instTris += geoms[tg.geomInst].numPrims;
        if (instTris > 0) {
            auto cacheIt = blasCache.find(fi.surfaces);
            if (cacheIt == blasCache.end()) {
                BlasInput in;
                in.triBase = static_cast<uint32_t>(numTris);
                in.numTris = static_cast<uint32_t>(instTris);
                in.nodeBase = 0;
                numTris += instTris;
                cacheIt = blasCache.emplace(fi.surfaces, static_cast<uint32_t>(blasList.size())).first;
                blasList.push_back(in);
                blasGroups.push_back(triGeoms);
            }
            blasIdx = cacheIt->second;
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
        instBlas.push_back(blasIdx);
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

    // acceleration structures: one BLAS per distinct geometry group (leaves index its triangle
    // records, child nodes count from its first node), the TLAS over the instance boxes with the
    // instance index in each leaf (scene.cpp:1319-1329 rebuilds the GASes and the IAS)
    AccelStats &as = ctx.accelStats;
    as = AccelStats();
    as.numBlas = static_cast<uint32_t>(blasList.size());
    as.numTlasInstances = static_cast<uint32_t>(instBoxes.size());
    as.numTriangles = numTris;
    std::vector<float> tlasBoxes(instBoxes.size() * 6);
    for (size_t i = 0; i < instBoxes.size(); ++i)
        for (int a = 0; a < 3; ++a) {
            tlasBoxes[6 * i + a] = instBoxes[i].lo[a];
            tlasBoxes[6 * i + 3 + a] = instBoxes[i].hi[a];
        }
    const auto tStart = std::chrono::steady_clock::now();
    auto msSince = [](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t).count();
    };
    uint32_t blasNodeCount = 0, tlasNodeCount = 0;
    // BLASes are kept while no geometry group changed (Scene::setup rebuilds only the dirty
    // GASes, scene.cpp:1319-1323)
    std::vector<uint64_t> blasKey;
    for (const auto &groups : blasGroups) {
        blasKey.push_back(groups.size());
        for (const TriGeom &tg : groups) {
            blasKey.push_back(reinterpret_cast<uintptr_t>(tg.node));
            blasKey.push_back(tg.node->version);
            blasKey.push_back((uint64_t(tg.group) << 32) | tg.geomInst);
            blasKey.push_back(geoms[tg.geomInst].numPrims);
        }
    }
    const bool reuseBlas = g->valid && g->blasBuilder == ctx.accelBuilder && g->blasKey == blasKey &&
                           g->blasLayout.size() == 3 * blasList.size();
    std::vector<TriRecord> blasRecs;
    if (!reuseBlas) {
        g->blasKey.clear();   // set again once the new BLASes are complete
        blasRecs.reserve(numTris);
        for (const auto &groups : blasGroups)
            for (const TriGeom &tg : groups) {
                const DevGeomInst &G = geoms[tg.geomInst];
                for (uint32_t t = 0; t < G.numPrims; ++t) {
                    TriRecord r;
                    std::memset(&r, 0, sizeof(r));
                    const DevVertex &a = verts[G.vertexBase + indices[G.indexBase + 3 * t]];
                    const DevVertex &b = verts[G.vertexBase + indices[G.indexBase + 3 * t + 1]];
                    const DevVertex &c = verts[G.vertexBase + indices[G.indexBase + 3 * t + 2]];
                    std::memcpy(r.v0, a.p, 12);
                    std::memcpy(r.v1, b.p, 12);
                    std::memcpy(r.v2, c.p, 12);
                    r.geomInst = tg.geomInst;
                    r.prim = t;
                    blasRecs.push_back(r);
                }
            }
    }
    uint32_t blasDepth = 0, tlasDepth = 0;
    if (reuseBlas) {
        for (size_t k = 0; k < blasList.size(); ++k) blasList[k].nodeBase = g->blasLayout[3 * k + 2];
        blasNodeCount = g->numBlasNodes;
        blasDepth = g->blasDepth;
        as.blasReused = 1;
        as.msBlas = msSince(tStart);
    }
    const bool onDevice = ctx.accelBuilder == kAccelBuilderDevice;
    if (!reuseBlas && onDevice) {
        // all BLASes in one build: segment k (BLAS k) has its root at node k
        GpuBvhBuilder &B = ctx.bvhBuilder();
        size_t cap = 0;
        std::vector<uint32_t> segFirst, segCount;
        for (BlasInput &b : blasList) {
            cap += GpuBvhBuilder::nodeCapacity(b.numTris);
            b.nodeBase = static_cast<uint32_t>(segFirst.size());
            segFirst.push_back(b.triBase);
            segCount.push_back(b.numTris);
        }
        DeviceBuffer recsDev;
        recsDev.upload(blasRecs.data(), blasRecs.size() * sizeof(TriRecord));
        g->blasNodes.reserve(std::max<size_t>(cap, 1) * sizeof(BVH4QNode));
        g->blasTris.reserve(std::max<size_t>(blasRecs.size(), 1) * sizeof(TriRecord));
        const auto *recs = static_cast<const TriRecord *>(recsDev.ptr);
        BvhBuildTimes info;
        blasNodeCount = B.buildTriangles(recs, segFirst, segCount, 4, static_cast<BVH4QNode *>(g->blasNodes.ptr), nullptr,
                                         &info);
        blasDepth = info.collapseLevels;
        gatherTriangles(recs, B.order(), static_cast<uint32_t>(blasRecs.size()),
                        static_cast<TriRecord *>(g->blasTris.ptr), nullptr);
        VLRA_HIP(hipStreamSynchronize(nullptr));
        as.msBlas = msSince(tStart);
    } else if (!reuseBlas) {
        // one host build per BLAS, stored one after the other
        std::vector<BVH4QNode> blasQ;
        std::vector<TriRecord> ordered(blasRecs.size());
        std::vector<float> boxes;
        HostBvh4 hb;
        for (BlasInput &b : blasList) {
            boxes.resize(size_t(b.numTris) * 6);
            for (uint32_t t = 0; t < b.numTris; ++t) {
                const TriRecord &r = blasRecs[b.triBase + t];
                for (int a = 0; a < 3; ++a) {
                    boxes[6 * size_t(t) + a] = std::min(std::min(r.v0[a], r.v1[a]), r.v2[a]);
                    boxes[6 * size_t(t) + 3 + a] = std::max(std::max(r.v0[a], r.v1[a]), r.v2[a]);
                }
            }
            hostBuildBvh4(boxes.data(), b.numTris, 4, nullptr, &hb);
            blasDepth = std::max(blasDepth, hb.depth);
            b.nodeBase = static_cast<uint32_t>(blasQ.size());
            blasQ.insert(blasQ.end(), hb.nodes.begin(), hb.nodes.end());
            for (uint32_t t = 0; t < b.numTris; ++t) ordered[b.triBase + t] = blasRecs[b.triBase + hb.order[t]];
        }
        blasNodeCount = static_cast<uint32_t>(blasQ.size());
        g->blasNodes.upload(blasQ.data(), blasQ.size() * sizeof(BVH4QNode));
        g->blasTris.upload(ordered.data(), ordered.size() * sizeof(TriRecord));
        as.msBlas = msSince(tStart);
    }
    // the TLAS is rebuilt on every setup (scene.cpp:1325-1327)
    const auto tTlas = std::chrono::steady_clock::now();
    const uint32_t numTlasPrims = static_cast<uint32_t>(instBoxes.size());
    // refit instead of a rebuild when only instance transforms changed since the last build
    // (the same instances over the same BLASes), unless the refitted tree's cost grew past
    // twice the cost it had when it was built
    std::vector<uint64_t> tlasKey(tlasInst.begin(), tlasInst.end());
    tlasKey.insert(tlasKey.end(), instBlas.begin(), instBlas.end());
    std::vector<Box> boxOfInst(insts.size());
    for (size_t p = 0; p < tlasInst.size(); ++p) boxOfInst[tlasInst[p]] = instBoxes[p];
    bool refitted = false;
    if (ctx.tlasRefit && reuseBlas && numTlasPrims && g->tlasKey == tlasKey && !g->tlasHost.empty()) {
        std::vector<BVH4QNode> nodes = g->tlasHost;
        const double cost = refitTlas(nodes, boxOfInst);
        if (cost <= 2.0 * g->tlasBuildCost) {
            g->tlasNodes.upload(nodes.data(), nodes.size() * sizeof(BVH4QNode));
            g->tlasHost = std::move(nodes);
            tlasNodeCount = g->numTlasNodes;
            tlasDepth = g->tlasDepth;
            refitted = true;
        }
    }
    if (refitted) {
    } else if (onDevice) {
        if (numTlasPrims) {
            DeviceBuffer boxDev, remapDev;
            boxDev.upload(tlasBoxes.data(), tlasBoxes.size() * sizeof(float));
            remapDev.upload(tlasInst.data(), tlasInst.size() * sizeof(uint32_t));
            g->tlasNodes.reserve(GpuBvhBuilder::nodeCapacity(numTlasPrims) * sizeof(BVH4QNode));
            BvhBuildTimes info;
            tlasNodeCount = ctx.bvhBuilder().build(static_cast<const float *>(boxDev.ptr), numTlasPrims, 1,
                                                   static_cast<const uint32_t *>(remapDev.ptr),
                                                   static_cast<BVH4QNode *>(g->tlasNodes.ptr), nullptr, &info);
            tlasDepth = info.collapseLevels;
        }
    } else {
        HostBvh4 hb;
        hostBuildBvh4(tlasBoxes.data(), numTlasPrims, 1, tlasInst.data(), &hb);
        tlasDepth = hb.depth;
        tlasNodeCount = static_cast<uint32_t>(hb.nodes.size());
        g->tlasNodes.upload(hb.nodes.data(), hb.nodes.size() * sizeof(BVH4QNode));
    }
    if (!refitted) {
        // the host copy a later refit starts from, and the cost it is judged against
        g->tlasHost.assign(tlasNodeCount, BVH4QNode{});
        if (tlasNodeCount)
            VLRA_HIP(hipMemcpy(g->tlasHost.data(), g->tlasNodes.ptr, tlasNodeCount * sizeof(BVH4QNode),
                               hipMemcpyDeviceToHost));
        std::vector<BVH4QNode> probe = g->tlasHost;
        g->tlasBuildCost = tlasNodeCount ? refitTlas(probe, boxOfInst) : 0.0;
        g->tlasKey = std::move(tlasKey);
        g->tlasDepth = tlasDepth;
    }
    as.tlasRefit = refitted ? 1 : 0;
    as.msTlas = msSince(tTlas);
    // every entry a traversal pushes must fit its stack (travStackBound)
    as.tlasDepth = tlasDepth;
    as.blasDepth = blasDepth;
    g->blasDepth = blasDepth;
    if (travStackBound(tlasDepth, blasDepth) > kTravStackCapacity)
        throw ApiError{3, "acceleration structure too deep for the traversal stack: TLAS depth " +
                              std::to_string(tlasDepth) + " + BLAS depth " + std::to_string(blasDepth) +
                              " needs " + std::to_string(travStackBound(tlasDepth, blasDepth)) + " of " +
                              std::to_string(kTravStackCapacity) + " entries"};
    g->blasKey = std::move(blasKey);
    g->blasBuilder = ctx.accelBuilder;
    g->blasLayout.clear();
    for (const BlasInput &b : blasList) {
        g->blasLayout.push_back(b.triBase);
        g->blasLayout.push_back(b.numTris);
        g->blasLayout.push_back(b.nodeBase);
    }
    as.numBlasNodes = blasNodeCount;
    as.numTlasNodes = tlasNodeCount;
    g->numBlasNodes = blasNodeCount;
    g->numTlasNodes = tlasNodeCount;
    for (size_t i = 0; i < insts.size(); ++i) {
        if (instBlas[i] == ~0u) continue;
        insts[i].blasNodeBase = blasList[instBlas[i]].nodeBase;
        insts[i].blasTriBase = blasList[instBlas[i]].triBase;
    }

    // upload
    g->instances.upload(insts.data(), insts.size() * sizeof(DevInstance));
    g->instGeomInsts.upload(instGeoms.data(), instGeoms.size() * sizeof(uint32_t));
    g->geomInsts.upload(geoms.data(), geoms.size() * sizeof(DevGeomInst));
    g->vertices.upload(verts.data(), verts.size() * sizeof(DevVertex));
    g->indices.upload(indices.data(), indices.size() * sizeof(uint32_t));
    g->distPool.upload(pool.data(), pool.size() * sizeof(float));
    g->materials.upload(mats.data(), mats.size() * sizeof(DevMaterial));
    if (shaders.nodes.empty()) shaders.nodes.emplace_back();
    if (shaders.textures.empty()) shaders.textures.emplace_back();
    if (shaders.texels.empty()) shaders.texels.push_back(make_float4(0, 0, 0, 0));
    if (shaders.pool.empty()) shaders.pool.push_back(0.0f);
    const bool hasPrograms = !shaders.steps.empty();
    as.numNodeProgramSteps = static_cast<uint32_t>(shaders.steps.size());
    if (shaders.steps.empty()) shaders.steps.emplace_back();
    g->shaderNodes.upload(shaders.nodes.data(), shaders.nodes.size() * sizeof(DevShaderNode));
    g->nodeSteps.upload(shaders.steps.data(), shaders.steps.size() * sizeof(DevNodeStep));
    g->textures.upload(shaders.textures.data(), shaders.textures.size() * sizeof(DevTexture));
    g->texels.upload(shaders.texels.data(), shaders.texels.size() * sizeof(float4));
    g->nodePool.upload(shaders.pool.data(), shaders.pool.size() * sizeof(float));
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
    D.nodes = static_cast<const DevShaderNode *>(g->shaderNodes.ptr);
    D.nodeSteps = static_cast<const DevNodeStep *>(g->nodeSteps.ptr);
    D.hasNodePrograms = hasPrograms ? 1 : 0;
    D.hasAlpha = hasAlpha ? 1 : 0;
    D.textures = static_cast<const DevTexture *>(g->textures.ptr);
    D.texels = static_cast<const float4 *>(g->texels.ptr);
    D.nodePool = static_cast<const float *>(g->nodePool.ptr);
    D.cacheMaterials = static_cast<const float4 *>(g->cacheMaterials.ptr);
    D.envTex = envTex.empty() ? nullptr : static_cast<const float4 *>(g->envTex.ptr);
    D.mengGrid = static_cast<const MengCellDev *>(g->mengGrid.ptr);
    D.mengPoints = static_cast<const float *>(g->mengPoints.ptr);
    D.numInstances = static_cast<uint32_t>(insts.size());
    D.hasTlas = instBoxes.empty() ? 0 : 1;
    cmfBins(D.cmfX, D.cmfY, D.cmfZ, &D.integralCMF);
    {
        ImmSpectrum white;
        white.cs = ColorSpace::Rec709_D65;
        white.e[0] = white.e[1] = white.e[2] = 1.0f;
        D.whitePoint = makeDevSpectrum(SpectrumType::LightSource, white, spectral);
    }
    g->spectral = spectral;
    g->valid = true;
    g->numInstances = static_cast<uint32_t>(insts.size());
    g->numGeomInsts = static_cast<uint32_t>(geoms.size());
    g->numTriangles = numTris;
    (void)ctx;
}

} // namespace vlra
