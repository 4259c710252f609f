// Host runtime behind the VLR C API on MI355X.
//
// Object model mirrors the reference's scene graph (libVLR/scene.h:137-530, materials.h,
// shader_nodes.h, image.h, queryable.h): surface nodes own vertices and material groups,
// parent nodes (internal nodes and the scene) own transforms and children, materials and
// shader nodes are queryable by parameter name.  Context::render flattens the graph into
// the device layout of dev_types.h (instances, geometry instances, BVHs, light
// distributions) and launches the wavefront kernels of kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <functional>
#include <cstdint>
#include <map>
#include <memory>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "dev_types.h"
#include "bvh.h"
#include "color.h"
#include "launch.h"

#include "../../include/vlr/public_types.h"

namespace vlra {

class Context;

// ---- errors ----
struct ApiError {
    int code;   // VLRResult value
    std::string msg;
};

// ---- colour ----
enum class SpectrumType : uint32_t { Reflectance = 0, LightSource = 1, IndexOfRefraction = 2, NA = 3 };
enum class ColorSpace : uint32_t { Rec709_D65_sRGBGamma = 0, Rec709_D65 = 1, XYZ = 2, xyY = 3 };

struct ImmSpectrum {
    ColorSpace cs = ColorSpace::Rec709_D65;
    float e[3] = {0, 0, 0};
};

// spectral_host.cpp
bool loadSpectralTables();
const std::vector<MengCellDev> &mengGrid();
const std::vector<float> &mengPoints();
void cmfBins(float x[16], float y[16], float z[16], float *integral);
// device spectrum parameter for an immediate spectrum (RGB build: r,g,b; spectral: Meng)
DevSpectrum makeDevSpectrum(SpectrumType type, const ImmSpectrum &s, bool spectral);
MengTables hostMengTables();
void sampledSpectrumToRGB(SpectrumType type, bool regular, float minLambda, float maxLambda, const float *lambdas,
                          const float *values, uint32_t n, float rgb[3]);
bool colorSpaceFromString(const char *s, ColorSpace *cs);
const char *colorSpaceToString(ColorSpace cs);
bool spectrumTypeFromString(const char *s, SpectrumType *t);
uint16_t floatToHalf(float f);
float halfToFloat(uint16_t h);

// ---- objects ----
enum class ObjType : uint32_t {
    Context = 1,
    ParameterInfo,
    LinearImage2D,
    BlockCompressedImage2D,
    ShaderNode,
    SurfaceMaterial,
    StaticTransform,
    TriangleMeshSurfaceNode,
    PointSurfaceNode,
    InternalNode,
    Scene,
    Camera,
};

constexpr uint32_t kObjMagic = 0x564C5241u;   // "VLRA"

struct ParamInfo;

class Object {
public:
    const uint32_t magic = kObjMagic;
    const ObjType type;
    Context *ctx;
    Object(ObjType t, Context *c) : type(t), ctx(c) {}
    virtual ~Object();
    virtual const char *typeName() const = 0;
    bool isA(const char *abstractName) const;
    bool isQueryable() const;
    bool isImage() const { return type == ObjType::LinearImage2D || type == ObjType::BlockCompressedImage2D; }
    bool isNode() const {
        return type == ObjType::TriangleMeshSurfaceNode || type == ObjType::PointSurfaceNode ||
               type == ObjType::InternalNode || type == ObjType::Scene;
    }
    bool isSurfaceNode() const { return type == ObjType::TriangleMeshSurfaceNode || type == ObjType::PointSurfaceNode; }
    bool isParentNode() const { return type == ObjType::InternalNode || type == ObjType::Scene; }
};

// Queryable parameter description (queryable.h:25-70)
struct ParamInfo : public Object {
    std::string name;
    uint32_t formFlags;
    std::string typeStr;
    uint32_t tupleSize;
    ParamInfo(const char *n, uint32_t form, const char *t, uint32_t tuple = 1)
        : Object(ObjType::ParameterInfo, nullptr), name(n), formFlags(form), typeStr(t), tupleSize(tuple) {}
    const char *typeName() const override { return "ParameterInfo"; }
};

class ShaderNode;
class Image2D;
class SurfaceMaterial;

struct PlugRef {
    const ShaderNode *node = nullptr;
    uint32_t plugType = 0;
    uint32_t option = 0;
    bool valid() const { return node != nullptr; }
};

// The value type a socket reads; a plug connects to it where the implicit conversion is
// defined (NodeTypeInfo<T>::ConversionIsDefinedFrom, shared.h:390-470).
enum class SocketType : uint32_t { Float, Float3, Point3D, Vector3D, Normal3D, Spectrum };
bool plugConvertible(SocketType socket, uint32_t plugType);

class Queryable : public Object {
public:
    using Object::Object;
    virtual const std::vector<ParamInfo *> &params() const = 0;
    // generic parameter access; returns false when the name/form does not apply
    virtual bool getEnum(const char *, const char **) const { return false; }
    virtual bool setEnum(const char *, const char *) { return false; }
    virtual bool getPoint(const char *, float *) const { return false; }
    virtual bool setPoint(const char *, const float *) { return false; }
    virtual bool getVector(const char *, float *) const { return false; }
    virtual bool setVector(const char *, const float *) { return false; }
    virtual bool getNormal(const char *, float *) const { return false; }
    virtual bool setNormal(const char *, const float *) { return false; }
    virtual bool getQuaternion(const char *, float *) const { return false; }
    virtual bool setQuaternion(const char *, const float *) { return false; }
    virtual bool getFloatTuple(const char *, float *, uint32_t) const { return false; }
    virtual bool setFloatTuple(const char *, const float *, uint32_t) { return false; }
    virtual bool getFloatArray(const char *, const float **, uint32_t *) const { return false; }
    virtual bool getImage(const char *, const Image2D **) const { return false; }
    virtual bool setImage(const char *, const Image2D *) { return false; }
    virtual bool getSpectrum(const char *, ImmSpectrum *) const { return false; }
    virtual bool setSpectrum(const char *, const ImmSpectrum &) { return false; }
    virtual bool getMaterial(const char *, const SurfaceMaterial **) const { return false; }
    virtual bool setMaterial(const char *, const SurfaceMaterial *) { return false; }
    virtual bool getPlug(const char *, PlugRef *) const { return false; }
    // returns 0 ok, 1 unknown name, 2 incompatible node
    virtual int setPlug(const char *, const PlugRef &) { return 1; }
};

// ---- images (image.h) ----
class Image2D : public Queryable {
public:
    uint32_t width = 0, height = 0;
    std::string formatStr;
    SpectrumType spectrumType = SpectrumType::NA;
    ColorSpace colorSpace = ColorSpace::Rec709_D65;
    bool hasAlpha = false;
    using Queryable::Queryable;
    const std::vector<ParamInfo *> &params() const override;
    virtual uint32_t stride() const = 0;
    // RGBA linear floats for environment use; false when the format cannot be decoded
    virtual bool toRGBA32F(std::vector<float> *out) const = 0;
};

class LinearImage2D : public Image2D {
public:
    std::vector<uint8_t> data;
    LinearImage2D(Context *c) : Image2D(ObjType::LinearImage2D, c) {}
    const char *typeName() const override { return "LinearImage2D"; }
    uint32_t stride() const override;
    bool toRGBA32F(std::vector<float> *out) const override;
};

// BC1-BC7 (signed variants included) decoded into RGBA texels (bc_decode.cpp)
bool bcFormatDecodable(const std::string &format);
bool decodeBlockCompressed(const std::string &format, const uint8_t *data, size_t size, uint32_t width, uint32_t height,
                           std::vector<float4> *texels);

class BlockCompressedImage2D : public Image2D {
public:
    std::vector<std::vector<uint8_t>> mips;
    BlockCompressedImage2D(Context *c) : Image2D(ObjType::BlockCompressedImage2D, c) {}
    const char *typeName() const override { return "BlockCompressedImage2D"; }
    uint32_t stride() const override { return 0; }
    bool toRGBA32F(std::vector<float> *) const override { return false; }
};

// ---- shader nodes (shader_nodes.h) ----
enum class NodeKind : uint32_t {
    Geometry, Tangent, Float2, Float3, Float4, ScaleAndOffsetFloat, TripletSpectrum, RegularSampledSpectrum,
    IrregularSampledSpectrum, Float3ToSpectrum, ScaleAndOffsetUVTextureMap2D, Image2DTexture, EnvironmentTexture,
};

class ShaderNode : public Queryable {
public:
    NodeKind kind;
    std::map<std::string, std::vector<float>> floats;     // immediate float tuples / arrays
    std::map<std::string, PlugRef> plugs;
    std::map<std::string, std::string> enums;
    const Image2D *image = nullptr;
    ShaderNode(Context *c, NodeKind k);
    const char *typeName() const override;
    const std::vector<ParamInfo *> &params() const override;
    bool providesPlug(uint32_t plugType) const;
    bool getEnum(const char *, const char **) const override;
    bool setEnum(const char *, const char *) override;
    bool getFloatTuple(const char *, float *, uint32_t) const override;
    bool setFloatTuple(const char *, const float *, uint32_t) override;
    bool getFloatArray(const char *, const float **, uint32_t *) const override;
    bool getImage(const char *, const Image2D **) const override;
    bool setImage(const char *, const Image2D *) override;
    bool getPlug(const char *, PlugRef *) const override;
    int setPlug(const char *, const PlugRef &) override;
    // constant folding of a plug into an immediate value (returns false if not constant)
    bool constantFloat(uint32_t option, float *v) const;
    bool constantSpectrum(ImmSpectrum *s, SpectrumType *t) const;
};

// ---- materials (materials.h) ----
enum class MatKind : uint32_t {
    Matte, SpecularReflection, SpecularScattering, MicrofacetReflection, MicrofacetScattering, LambertianScattering,
    UE4, OldStyle, DiffuseEmitter, DirectionalEmitter, PointEmitter, Multi, EnvironmentEmitter,
};

class SurfaceMaterial : public Queryable {
public:
    MatKind kind;
    // spectrum slots (names per kind), float slots, node plugs per parameter name
    std::vector<std::string> specNames;
    std::vector<ImmSpectrum> specs;
    std::vector<std::string> floatNames;
    std::vector<float> floatVals;
    std::map<std::string, PlugRef> plugs;
    const SurfaceMaterial *sub[4] = {nullptr, nullptr, nullptr, nullptr};
    float direction[3] = {0, 0, 1};
    SurfaceMaterial(Context *c, MatKind k);
    const char *typeName() const override;
    const std::vector<ParamInfo *> &params() const override;
    bool isEmitting() const {
        return kind == MatKind::DiffuseEmitter || kind == MatKind::DirectionalEmitter ||
               kind == MatKind::PointEmitter || kind == MatKind::EnvironmentEmitter;
    }
    bool supportedOnRenderPath() const;
    bool getSpectrum(const char *, ImmSpectrum *) const override;
    bool setSpectrum(const char *, const ImmSpectrum &) override;
    bool getFloatTuple(const char *, float *, uint32_t) const override;
    bool setFloatTuple(const char *, const float *, uint32_t) override;
    bool getVector(const char *, float *) const override;
    bool setVector(const char *, const float *) override;
    bool getMaterial(const char *, const SurfaceMaterial **) const override;
    bool setMaterial(const char *, const SurfaceMaterial *) override;
    bool getPlug(const char *, PlugRef *) const override;
    int setPlug(const char *, const PlugRef &) override;
    // device descriptor (immediate values with constant node plugs folded in)
    // roots (may be null) receives the parameters connected to nodes that do not fold into
    // immediates, with their node-program output bit (1 << k: spectrum k, 8: float3 parameter)
    DevMaterial toDevice(bool spectral, bool *usesEnvTexture, const Image2D **envImage,
                         std::vector<std::pair<PlugRef, uint32_t>> *roots) const;
    // radiance-cache inputs: diffuse RGB, roughness, specular RGB, 0
    void cacheFeatures(float out[8]) const;
};

// ---- transforms ----
class StaticTransform : public Object {
public:
    float mat[16];      // column-major
    float invMat[16];
    StaticTransform(Context *c, const float m[16]);
    const char *typeName() const override { return "StaticTransform"; }
};

// ---- nodes (scene.h) ----
class ParentNode;

class Node : public Object {
public:
    std::string name;
    std::set<ParentNode *> parents;
    using Object::Object;
};

struct MaterialGroup {
    std::vector<uint32_t> indices;
    const SurfaceMaterial *material = nullptr;
    PlugRef nodeNormal, nodeTangent, nodeAlpha;
};

// Process-wide counter: a surface node takes a new value at creation and on every geometry
// change, so (node, version) never repeats even when a node's address is reused.
uint64_t nextGeometryVersion();

class SurfaceNode : public Node {
public:
    std::vector<VLRVertex> vertices;
    std::vector<MaterialGroup> groups;
    uint64_t version = nextGeometryVersion();
    using Node::Node;
};

class TriangleMeshSurfaceNode : public SurfaceNode {
public:
    TriangleMeshSurfaceNode(Context *c) : SurfaceNode(ObjType::TriangleMeshSurfaceNode, c) {}
    const char *typeName() const override { return "TriangleMeshSurfaceNode"; }
};

class PointSurfaceNode : public SurfaceNode {
public:
    PointSurfaceNode(Context *c) : SurfaceNode(ObjType::PointSurfaceNode, c) {}
    const char *typeName() const override { return "PointSurfaceNode"; }
};

class ParentNode : public Node {
public:
    const StaticTransform *transform = nullptr;
    std::vector<Node *> children;
    using Node::Node;
};

class InternalNode : public ParentNode {
public:
    InternalNode(Context *c) : ParentNode(ObjType::InternalNode, c) {}
    const char *typeName() const override { return "InternalNode"; }
};

class Scene : public ParentNode {
public:
    const SurfaceMaterial *envMaterial = nullptr;
    float envRotation = 0.0f;
    Scene(Context *c) : ParentNode(ObjType::Scene, c) {}
    const char *typeName() const override { return "Scene"; }
};

// ---- camera ----
class Camera : public Queryable {
public:
    bool perspective = true;
    float position[3] = {0, 0, 0};
    float orientation[4] = {0, 0, 0, 1};
    float sensitivity = 1.0f, aspect = 1.0f, fovY = 0.8726646f, lensRadius = 0.0f, opDistance = 1.0f;
    float phiAngle = 6.28318530718f, thetaAngle = 3.14159265359f;
    Camera(Context *c, bool persp) : Queryable(ObjType::Camera, c), perspective(persp) {}
    const char *typeName() const override { return perspective ? "PerspectiveCamera" : "EquirectangularCamera"; }
    const std::vector<ParamInfo *> &params() const override;
    bool getPoint(const char *, float *) const override;
    bool setPoint(const char *, const float *) override;
    bool getQuaternion(const char *, float *) const override;
    bool setQuaternion(const char *, const float *) override;
    bool getFloatTuple(const char *, float *, uint32_t) const override;
    bool setFloatTuple(const char *, const float *, uint32_t) override;
    DevCamera toDevice() const;
};

// ---- shader nodes and textures of a scene as device records (shader_table.cpp) ----
struct ShaderTable {
    bool spectral = false;
    std::vector<DevShaderNode> nodes;
    std::vector<DevTexture> textures;
    std::vector<float4> texels;
    std::vector<float> pool;
    std::map<const ShaderNode *, uint32_t> nodeIndex;
    std::map<std::pair<const Image2D *, bool>, uint32_t> texIndex;
    std::vector<DevNodeStep> steps;
    uint32_t addNode(const ShaderNode *n);
    // envRead: the texels as EnvironmentTextureShaderNode's sampler reads them (no sRGB read,
    // and a magenta 1 x 1 light-source image where none is set)
    uint32_t addTexture(const Image2D *img, bool envRead = false);
    // One node program over the given roots (plug, output bits); returns (first step, count).
    std::pair<uint32_t, uint32_t> program(const std::vector<std::pair<PlugRef, uint32_t>> &roots);
};

// ---- device scene (scene_setup.cpp) ----
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    ~DeviceBuffer();
    void upload(const void *src, size_t n);
    void reserve(size_t n);
    void release();
};

struct SceneGPU {
    DeviceBuffer tlasNodes, blasNodes, blasTris, instances, instGeomInsts, geomInsts, vertices, indices, distPool,
        materials, envTex, mengGrid, mengPoints, cacheMaterials, shaderNodes, nodeSteps, textures, texels, nodePool;
    DevScene dev = {};
    bool spectral = false;
    bool valid = false;
    uint64_t sceneVersion = ~0ull;
    const Scene *scene = nullptr;
    const Camera *camera = nullptr;
    uint32_t numInstances = 0, numGeomInsts = 0;
    uint64_t numTriangles = 0;
    uint32_t numBlasNodes = 0, numTlasNodes = 0;
    uint32_t blasDepth = 0;
    // the TLAS as last built or refitted (host copy), its inputs' key (instances and their
    // BLASes) and the relative cost it had when built; a transform-only update refits it
    std::vector<BVH4QNode> tlasHost;
    std::vector<uint64_t> tlasKey;
    double tlasBuildCost = 0.0;
    uint32_t tlasDepth = 0;
    // BLAS inputs of the last build (per geometry group: node, version, group, geometry
    // instance, triangle count) and their layout; an equal key reuses the BLASes
    std::vector<uint64_t> blasKey;
    std::vector<uint32_t> blasLayout;   // triBase, numTris, nodeBase per BLAS
    uint32_t blasBuilder = ~0u;
};

// Acceleration-structure builds of the last scene setup.
constexpr uint32_t kAccelBuilderDevice = 0, kAccelBuilderHost = 1;
struct AccelStats {
    uint32_t numBlas = 0, numTlasInstances = 0;
    uint64_t numTriangles = 0;
    uint32_t numBlasNodes = 0, numTlasNodes = 0;
    float msBlas = 0.0f, msTlas = 0.0f;
    uint32_t blasReused = 0;
    uint32_t numNodeProgramSteps = 0;
    uint32_t tlasDepth = 0, blasDepth = 0;   // levels of 4-wide nodes
    uint32_t tlasRefit = 0;                  // 1: the TLAS was refitted, not rebuilt
};

void buildSceneGPU(Context &ctx, const Scene &scene, bool spectral, SceneGPU *out);

// ---- batches of render jobs ----
struct BatchJobDesc {
    const Scene *scene;
    const Camera *camera;
    uint32_t width, height, spp, maxPathLength;
};
struct BatchState {
    std::map<const Scene *, std::unique_ptr<SceneGPU>> scenes;
    uint64_t sceneVersion = ~0ull;
    bool spectral = false;
    DeviceBuffer rng, alpha, contrib, meta, rayOrg[2], rayDir[2], hitA, hitB, shOrg, shDir, shVal, jobs, jobOffsets,
        activeOffsets;
    std::vector<std::unique_ptr<DeviceBuffer>> accum, output;
    std::vector<uint32_t> widths, heights;
    hipStream_t stream = nullptr;
    uint32_t wavefrontDepth = 0, tailBlocks = 0;   // schedule of the last batch
    ~BatchState();
};

// ---- context ----
class Context : public Object {
public:
    int device = -1;
    uint64_t version = 1;     // bumped on every scene-affecting change
    uint32_t accelBuilder = kAccelBuilderDevice;
    bool tlasRefit = true;             // transform-only updates refit the TLAS
    AccelStats accelStats;
    std::unique_ptr<GpuBvhBuilder> gpuBvh;
    GpuBvhBuilder &bvhBuilder() {
        if (!gpuBvh) gpuBvh.reset(new GpuBvhBuilder());
        return *gpuBvh;
    }
    std::set<Object *> objects;
    Scene *scene = nullptr;
    uint32_t renderer = 0;
    uint32_t debugAttr = 0;
    int32_t probeX = 0, probeY = 0;
    uint32_t maxPathLength = 25;
    bool spectralRequested = false;
    bool spectralActive = false;
    bool kernelTiming = false;
    bool traversalStats = false;
    uint32_t wavefrontDepth = kDefaultWavefrontDepth;
    uint32_t tailBlocks = kDefaultTailBlocks;
    bool shadeSort = false;          // material sort of the per-depth shade queues
    DeviceBuffer sortKey, sortPerm, sortCounts;
    bool autotune = true;            // cleared when the caller sets the schedule explicitly
    ScheduleTuner tuner;
    hipEvent_t tuneEvents[2] = {nullptr, nullptr};
    // kernel timing: an event before and after every launch, read back lazily, so that timed
    // frames are not synchronised; timers[1] times the second half of a split frame
    struct EventTimer : public LaunchTimer {
        static constexpr size_t kMaxPending = size_t(1) << 14;   // events in flight before a read-back
        hipStream_t stream = nullptr;
        std::vector<hipEvent_t> pool;
        std::vector<int> stages;
        size_t used = 0;
        ~EventTimer();
        hipEvent_t next();
        void begin(int) override;
        void end(int stage) override;
        void collect(double *ms, uint64_t *launches);
    };
    EventTimer timers[2];
    EventTimer *timer(int k, hipStream_t s);   // null while kernel timing is off
    void collectTimers();                      // adds the finished launches to stageMs / launches
    uint32_t width = 0, height = 0;
    uint32_t numAccumFrames = 0;
    FrameParams lastFrame = {};
    DevCamera frameCamera = {};
    bool resumePending = false;   // accumulation written: the next frame reads the camera
    // device buffers
    DeviceBuffer rng, alpha, contrib, meta, rayOrg[2], rayDir[2], hitA, hitB, shOrg, shDir, shVal, counters, accum,
        output, scratch;
    uint32_t accumChannels = 3;
    // LVC-BPT state: light-path PCG32 states seeded by mt19937_64(459182033132123413)
    // (context.cpp:618-626) and the common-wavelength generator (context.h:226, a default
    // std::mt19937 that two uniform floats are drawn from per frame, context.cpp:1464-1467)
    uint32_t numLightPaths = 500000;
    uint32_t lightRngPaths = 0;      // light paths the light RNG buffer was seeded for
    std::mt19937 bptWlRng;
    uint64_t bptWlFrames = 0;        // BPT frames that drew their wavelengths from bptWlRng
    void bptWavelengths(float *offset, float *select);
    void setBptWavelengthFrames(uint64_t frames);
    DeviceBuffer lightRng, lvSlots, lvCounts, lvOffsets, lvBlockSums, lvCache, lvNum, splat;
    size_t splatLayout = 0;          // bytes of the splat planes last cleared
    // radiance cache (no counterpart in the reference)
    struct CacheSettings {
        float learningRate = 1e-2f;
        float beta1 = 0.9f, beta2 = 0.99f, epsilon = 1e-15f;
        uint32_t trainingPaths = 21845;   // x kCacheTrainVerts vertices <= rc::kMaxBatch
        float threshold = 0.01f;
        uint32_t warmupSteps = 16;
        uint32_t seed = 1;
        bool fullRadiance = true;
    } cacheSettings;
    bool cacheEnabled = false;
    bool cacheInitialized = false;
    uint32_t cacheStep = 0;          // Adam steps taken since the last reset
    uint32_t frameStamp = 0;
    DeviceBuffer rcW, rcWT, rcBias, rcMaster, rcM, rcV, rcGrad, rcPartial, rcAct, rcDelta, rcLossPartial, rcLoss;
    DeviceBuffer cqRaw, cqThroughput, cqStamp, cSpread, ctRaw, ctThroughput, ctContrib, ctStamp, ctAfterStamp, cbRaw,
        cbTarget, cbCount, csFlags, csOffsets, csBlockSums, csTargets;
    SceneGPU gpu;
    double stageMs[NUM_STAGES] = {};
    uint64_t launches[NUM_STAGES] = {};
    uint64_t framesRendered = 0, pathsStarted = 0;

    Context(hipCtx_t hctx);
    ~Context() override;
    const char *typeName() const override { return "Context"; }
    void touch() { ++version; }
    void ensureDevice();
    void bindOutputBuffer(uint32_t w, uint32_t h);
    hipArray_t outputArray = nullptr;   // made by vlrContextGetOutputBuffer, then updated every frame
    hipArray_t outputArrayHandle();
    void updateOutputArray(hipStream_t stream);
    void render(hipStream_t stream, const Camera *camera, uint32_t shrinkCoeff, bool firstFrame, uint32_t limit,
                uint32_t *numAccum);
    void readOutput(float *dst);
    void ensureScene(const Scene &s);
    void resizeAccum();
    void traceRays(hipStream_t stream, const Scene &s, const void *rays, uint32_t n, void *hits, bool anyHit);
    // acceleration structure of `s` as set up for rendering: 0 TLAS nodes, 1 BLAS nodes,
    // 2 BLAS triangle records, 3 per-instance (first BLAS node, first triangle record)
    void readAccel(const Scene &s, uint32_t which, void *dst, size_t capacity, size_t *bytes);
    void setAccelBuilder(uint32_t builder);
    // one node plug evaluated at device-resident surface points (vlrContextEvaluateShaderNodePlug)
    void evaluateNodePlug(hipStream_t stream, const PlugRef &plug, bool spectral, const float *points, uint32_t n,
                          float *values);
    void evaluateBSDFs(hipStream_t stream, const SurfaceMaterial &m, bool spectral, const void *q, uint32_t n, void *r);
    // trav: nodes, triangles, instances of closest-hit then of shadow traversals
    // wavefront: extension and shadow rays traced by the per-depth kernels
    void statistics(uint64_t *closest, uint64_t *shadow, uint64_t *dropped, uint64_t trav[6], uint64_t wavefront[2]);
    void resetStats();
    void queueCounts(uint32_t *closest, uint32_t *shadow, uint32_t maxDepth);
    // The device keeps the Kahan pairs as planes (sum of channel k at plane 2k, compensation at
    // 2k + 1, width*height floats each); the API exchanges them per pixel (channel, pair).
    void readAccumulation(float *perPixel);
    // denoiser guides of the path tracer (first-hit albedo and shading normal per sample,
    // path_tracing.cu:95-108), converted on read as copyBuffers does (post_process.cu:70-90)
    bool guidesEnabled = false;
    DeviceBuffer guideAlbedo, guideNormal;
    uint32_t guideFrames = 0;        // samples per pixel summed in the guides
    size_t guideBytes = 0;
    void readGuides(float *albedoRGBA, float *normalRGBA);
    void writeAccumulation(const float *perPixel);
    // LVC-BPT buffers for the current light-path count, path length and output size
    void ensureBpt(uint32_t slotsPerPath);
    void setNumLightPaths(uint32_t n);
    void renderBptFrame(hipStream_t stream, const FrameParams &fp, bool firstFrame, bool lightTracing);
    void prepareGuides(FrameParams &fp, FrameBuffers &fb);
    void readLightRngStates(uint64_t *dst);
    void writeLightRngStates(const uint64_t *src);
    // frames rendered as two pixel halves on two streams (one half's deep paths overlap the
    // other half's per-depth kernels)
    bool splitFrames = false;
    bool lastFrameSplit = false;
    bool halvesPending = false;       // the halves' last frame has not been joined into a caller's stream
    hipStream_t partStreams[2] = {nullptr, nullptr};
    hipEvent_t partEvents[3] = {nullptr, nullptr, nullptr};
    hipEvent_t wavefrontEvents[2] = {nullptr, nullptr};   // each half's per-depth kernels of its last frame
    bool wavefrontRecorded[2] = {false, false};
    DeviceBuffer counters2;           // queue counters and statistics of the second half
    // batches
    BatchState batch;
    bool batchGraphs = true;
    void renderBatch(hipStream_t stream, const std::vector<BatchJobDesc> &jobs, float *completionMs);
    void readBatchOutput(uint32_t job, float *rgba);
    // radiance cache
    void ensureCacheNet();
    void resetCache();
    void writeCacheWeights(const float *master);   // also clears the optimiser state
    void readCacheWeights(float *master);
    RcDevice rcDevice();
    void ensureCacheFrameBuffers();
    void cacheInfer(hipStream_t stream, const float *records, uint32_t n, float *out);
    void cacheTrain(hipStream_t stream, const float *records, const float *targets, uint32_t n, uint32_t numOut);
    FrameBuffers frameBuffers();
    void renderSplitFrame(hipStream_t stream, const FrameParams &fp, const FrameBuffers &fb, bool join);
    void joinHalves(hipStream_t stream);
};

void throwHip(hipError_t e, const char *what);
#define VLRA_HIP(x) ::vlra::throwHip((x), #x)

bool paramNameEquals(const std::string &a, const std::string &b);
void affineInverse(const float m[16], float inv[12]);

} // namespace vlra
