// Device data layout of the light-transport path (shared by host set-up code and kernels).
//
// HBM layout, all arrays 16-byte aligned:
//   BVH nodes     128 B, 4-wide: the four children's boxes stored in the parent (SoA), child
//                  index > 0 -> inner node, < 0 -> leaf code ~((first << 3) | (count - 1)),
//                  0 -> empty slot (index 0 is the root, never a child).
//   BLAS triangles 48 B: three float4 = object-space vertices; .w lanes carry the global
//                  geometry-instance index and the triangle index inside it.
//   Instances     128 B: world->object and object->world 3x4 rows plus BLAS bases and
//                  light-distribution offsets.
//   Vertices       48 B: (position, u) (normal, v) (tc0Direction, 0).
//   Path state     per pixel SoA: rng u64, throughput float4, contribution float4,
//                  meta float4 (wavelength offset, previous dir pdf, initial importance, flags).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define VLRA_HD __host__ __device__
#else
#define VLRA_HD
#endif

namespace vlra {

constexpr uint32_t kLeafShift = 3;          // leaf code: (first << 3) | (count - 1)
constexpr uint32_t kMaxLeafPrims = 8;
constexpr int32_t kInvalidIndex = -1;

struct alignas(16) BVHNode {
    float lo0[3], hi0[3];
    float lo1[3], hi1[3];
    int32_t child[2];
    uint32_t pad[2];
};
static_assert(sizeof(BVHNode) == 64, "node layout");

// 4-wide node with full-precision child boxes (the build format): the four child boxes as
// structure of arrays, then the four child codes (> 0 inner node, < 0 leaf code, 0 empty slot).
// Built by collapsing the binned-SAH BVH2 (two levels into one where it pays).
struct alignas(16) BVH4Node {
    float loX[4], hiX[4], loY[4], hiY[4], loZ[4], hiZ[4];
    int32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(BVH4Node) == 128, "4-wide node layout");

// The layout the traversal reads (64 B): child boxes quantised to 8 bits per plane on a grid of
// the node box, origin + q * 2^(e - 127) per axis, rounded outwards so every decoded box
// contains the exact one (and the decoding, one exact product and one rounded sum, is done the
// same way on the host that checks it).  Byte k of qlo[a] / qhi[a] belongs to child k.
struct alignas(16) BVH4QNode {
    float origin[3];
    uint32_t scaleExp;        // bytes 0, 1, 2: biased exponents of the x, y, z grid steps
    uint32_t qlo[3], qhi[3];
    int32_t child[4];
    uint32_t pad[2];
};
static_assert(sizeof(BVH4QNode) == 64, "quantised 4-wide node layout");

struct alignas(16) TriRecord {
    float v0[3]; uint32_t geomInst;
    float v1[3]; uint32_t prim;
    float v2[3]; uint32_t pad;
};
static_assert(sizeof(TriRecord) == 48, "tri layout");

struct alignas(16) DevVertex {
    float p[3]; float u;
    float n[3]; float v;
    float t[3]; float pad;
};
static_assert(sizeof(DevVertex) == 48, "vertex layout");

struct alignas(16) DevInstance {
    float w2o[12];
    float o2w[12];
    uint32_t blasNodeBase;
    uint32_t blasTriBase;
    uint32_t geomInstListBase;
    uint32_t numGeomInsts;
    uint32_t lightDistBase;   // PMF at base, CDF at base + numGeomInsts
    float lightDistIntegral;
    uint32_t hasGeometry;
    uint32_t pad;
};
static_assert(sizeof(DevInstance) == 128, "instance layout");

enum DevMaterialType : uint32_t {
    MAT_NULL = 0,
    MAT_MATTE,
    MAT_SPECULAR_REFLECTION,
    MAT_SPECULAR_SCATTERING,
    MAT_MICROFACET_REFLECTION,
    MAT_UE4,
    MAT_DIFFUSE_EMITTER,
    MAT_POINT_EMITTER,
    MAT_ENVIRONMENT_EMITTER,
    MAT_MICROFACET_SCATTERING,
};

// A spectrum parameter: RGB build -> (r, g, b, 0); spectral build -> Meng coefficients
// (adjacent-point indices packed in .x bits, s | t << 16 in .y bits, scale in .z).
struct alignas(16) DevSpectrum {
    float x, y, z, w;
};

// Material record.  spec[k] / scalars are the immediate parameters.  Parameters connected to
// shader nodes are computed per hit by the node program of progCount steps at progBase
// (outputs: spec[0..2] and the float3 parameter: roughness / anisotropy / rotation or
// occlusion / roughness / metallic, whose immediates are raw[0..2]).
struct alignas(16) DevMaterial {
    uint32_t type;
    uint32_t useEnvTexture;
    uint32_t isEmitter;
    uint32_t pad;
    float scalars[4];
    DevSpectrum spec[3];
    uint32_t progBase, progCount, pad2[2];
    float raw[4];
};
static_assert(sizeof(DevMaterial) == 112, "material layout");

// ---- shader nodes (shader_nodes.cu:20-483) ----
// A plug: 0 = not connected, else (node index + 1) << 8 | plug type << 4 | option, plug types
// as VLRShaderNodePlugType (float1..4, Point3D, Vector3D, Normal3D, Spectrum, Alpha,
// TextureCoordinates).
enum DevPlugType : uint32_t {
    PLUG_FLOAT1 = 0, PLUG_FLOAT2, PLUG_FLOAT3, PLUG_FLOAT4, PLUG_POINT3D, PLUG_VECTOR3D, PLUG_NORMAL3D, PLUG_SPECTRUM,
    PLUG_ALPHA, PLUG_TEXCOORD,
};
__host__ __device__ inline uint32_t makePlug(uint32_t node, uint32_t type, uint32_t option) {
    return ((node + 1) << 8) | (type << 4) | option;
}
__host__ __device__ inline uint32_t plugNode(uint32_t p) { return (p >> 8) - 1; }
__host__ __device__ inline uint32_t plugType(uint32_t p) { return (p >> 4) & 0xF; }
__host__ __device__ inline uint32_t plugOption(uint32_t p) { return p & 0xF; }

// Node kinds in the order of the C API's type names.
enum DevNodeKind : uint32_t {
    NODE_GEOMETRY = 0, NODE_TANGENT, NODE_FLOAT2, NODE_FLOAT3, NODE_FLOAT4, NODE_SCALE_OFFSET_FLOAT,
    NODE_TRIPLET_SPECTRUM, NODE_REGULAR_SPECTRUM, NODE_IRREGULAR_SPECTRUM, NODE_FLOAT3_TO_SPECTRUM, NODE_UV_MAP,
    NODE_IMAGE_TEXTURE, NODE_ENV_TEXTURE,
};

// One node (64 B): its kind and immediates (the inputs it reads are in the node-program
// steps).  Per kind:
//   TANGENT             a[0] tangent type (0 TC0 direction, 1..3 radial about x, y, z)
//   FLOAT2/3/4          f[k]: component k's immediate (input k)
//   SCALE_OFFSET_FLOAT  inputs value, scale, offset; f[1], f[2] immediates of scale, offset
//   TRIPLET_SPECTRUM    f[0..3] the spectrum parameter (DevSpectrum) of the render mode
//   REGULAR_SPECTRUM    spectral build: a[0] pool offset of the values, a[1] count, f[0..1]
//                       wavelength range; RGB build: f[0..3] the converted triplet, a[2] = 1
//   IRREGULAR_SPECTRUM  spectral build: a[0] pool offset of wavelengths then values, a[1]
//                       count; RGB build as REGULAR
//   FLOAT3_TO_SPECTRUM  input 0 float3, f[0..2] its immediate, a[0] spectrum type, a[1] colour space
//   UV_MAP              f[0..1] scale, f[2..3] offset
//   IMAGE_TEXTURE       input 0 texture coordinates, a[0] texture, a[1] filter | wrap u << 2 |
//                       wrap v << 4 | bump type << 8, f[0] bump coefficient
struct alignas(16) DevShaderNode {
    uint32_t kind;
    uint32_t pad[4];
    uint32_t a[3];
    float f[8];
};
static_assert(sizeof(DevShaderNode) == 64, "shader node layout");

// One step of a node program: evaluates `node`'s program for plug type (plug & 15) and option
// (plug >> 4) into slot i of the program, reading its inputs from earlier slots (in[k] = slot
// + 1, 0 = not connected: the node's immediate).  out: bit k (k < 3) sends the step's spectrum
// to output spectrum k, bit 3 / 4 its value to output vector 0 / 1.  Programs list the nodes of
// a graph in post-order, at most kNodeProgramSlots steps.
constexpr uint32_t kNodeProgramSlots = 8;
struct alignas(16) DevNodeStep {
    uint32_t node, plug, out, pad;
    uint32_t in[4];
};
static_assert(sizeof(DevNodeStep) == 32, "node step layout");

// A texture: width x height float4 texels from `offset` in the texel pool, prepared on the
// host for the render mode (linear RGBA, or (u, v, s, a) grid coordinates of upsampled
// spectra: format 1 from 8-bit sources as u / 12, v / 14, s / 3, format 2 from half sources).
constexpr uint32_t kTexRGBA = 0, kTexUVS8 = 1, kTexUVS16 = 2;
struct alignas(16) DevTexture {
    uint32_t offset, width, height, format;
    uint32_t gray;            // one-channel source: the spectrum uses x for all three channels
    uint32_t spectrumType, colorSpace, pad;
};
static_assert(sizeof(DevTexture) == 32, "texture layout");

struct alignas(16) DevGeomInst {
    uint32_t material;
    uint32_t vertexBase;
    uint32_t indexBase;     // into the index pool (3 per triangle, 1 per point)
    uint32_t numPrims;
    uint32_t distBase;      // prim distribution: PMF at base, CDF at base + numPrims
    float distIntegral;
    float importance;
    uint32_t isPoints;
    // node program of the material group's normal (output vec 0) and tangent (vec 1) nodes
    uint32_t progBase, progCount;
    // node program of its alpha node (output vec 0, a float); alphaCount 0: opaque
    uint32_t alphaBase, alphaCount;
};
static_assert(sizeof(DevGeomInst) == 48, "geomInst layout");

struct MengCellDev { uint8_t inside, numPoints, idx[6]; };

struct DevCamera {
    float position[3];
    float rot[9];            // column-major rotation from the orientation quaternion
    float sensitivity;
    float lensRadius;
    float opWidth, opHeight;
    float imgPlaneDistance, objPlaneDistance, imgPlaneArea;
};

struct DevScene {
    const BVH4QNode* tlasNodes;
    const BVH4QNode* blasNodes;
    const TriRecord* blasTris;
    const DevInstance* instances;
    const uint32_t* instGeomInsts;
    const DevGeomInst* geomInsts;
    const DevVertex* vertices;
    const uint32_t* indices;
    const float* distPool;
    const DevMaterial* materials;
    const float4* envTex;
    const MengCellDev* mengGrid;
    const float* mengPoints;     // 186 x 99 floats: xy*(2) uv(2) spectrum(95)
    const float4* cacheMaterials; // per material: (diffuse rgb, roughness), (specular rgb, 0) for the radiance cache
    const DevShaderNode* nodes;   // shader nodes reachable from the scene's materials and groups
    const DevNodeStep* nodeSteps; // node programs of materials and material groups
    const DevTexture* textures;
    const float4* texels;
    const float* nodePool;        // sampled-spectrum tables of nodes
    uint32_t numInstances;
    uint32_t hasTlas;
    uint32_t hasNodePrograms;    // some material or material group has a node program
    uint32_t hasAlpha;           // some triangle group has an alpha node (stochastic alpha test)
    // light selection over [environment, instances...]
    uint32_t lightInstBase;
    uint32_t numLightInsts;
    float lightInstIntegral;
    // environment
    uint32_t hasEnv;
    uint32_t envMaterial;
    float envImportance;
    uint32_t envW, envH;
    float envRotation;
    uint32_t envMapW, envMapH;   // importance map size
    uint32_t envRowBase;         // row r: PDF at envRowBase + r*(2*mapW+1), CDF after it
    uint32_t envTopBase;         // PDF[mapH], CDF[mapH+1]
    // scene bounds (setup_scene.cu:82-92)
    float center[3];
    float worldRadius, worldDiscArea;
    DevCamera camera;
    float cmfX[16], cmfY[16], cmfZ[16];
    float integralCMF;
    DevSpectrum whitePoint;      // light-source (1, 1, 1) in Rec.709 (D65): the guides' white
};

// Wavefront queues and per-pixel path state.
struct PathState {
    uint64_t* rng;
    float4* alpha;
    float4* contrib;
    float4* meta;     // x wlOffset, y prevDirPDF, z initImportance, w flags (as bits)
};

struct RayQueue {
    float4* org;      // xyz, w = path index (bits)
    float4* dir;      // xyz, w = tmax
};

struct HitBuffer {
    int4* a;          // inst, geomInst, prim, t (bits)
    float2* b;        // b1, b2
};

struct ShadowQueue {
    float4* org;      // xyz, w = path index (bits)
    float4* dir;      // xyz, w = tmax
    float4* value;    // contribution added when unoccluded
};

// counters layout (uint32): [2*d] = rays at depth d, [2*d + 1] = shadow rays at depth d,
// [kStatBase + ...] = statistics
constexpr uint32_t kMaxDepthCounters = 64;
constexpr uint32_t kStatClosest = 2 * kMaxDepthCounters + 0;   // 64-bit counters as 2 x u32
constexpr uint32_t kStatShadow = 2 * kMaxDepthCounters + 2;
constexpr uint32_t kStatDropped = 2 * kMaxDepthCounters + 4;
// traversal work (64-bit each: inner nodes fetched, triangles tested, instances entered)
constexpr uint32_t kStatTravClosest = 2 * kMaxDepthCounters + 6;
constexpr uint32_t kStatTravShadow = 2 * kMaxDepthCounters + 12;
constexpr uint32_t kTailWork = 2 * kMaxDepthCounters + 18;   // next queue entry for the tail kernel
// rays traced by the per-depth kernels (64-bit: extension, shadow); the rest went through the tail
constexpr uint32_t kStatWavefront = 2 * kMaxDepthCounters + 20;
constexpr uint32_t kNumCounters = 2 * kMaxDepthCounters + 24;

// path flag bits in meta.w
constexpr uint32_t kFlagSelMask = 0x3;
constexpr uint32_t kFlagSingle = 1u << 2;
constexpr uint32_t kFlagPrevTypeShift = 8;   // 7 bits of DirectionType
constexpr uint32_t kFlagPathLenShift = 16;

} // namespace vlra
