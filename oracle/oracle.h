/* Oracle (test infrastructure only): C interface of the CPU restatement of the reference's
 * light-transport path (libVLR/GPU_kernels/path_tracing.cu:30-337 and the shared code it
 * calls).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * The scene is handed over flat, in terms of the reference's own parameters (material
 * type names and immediate spectra, vertices, per-instance 4x4 transforms), so the oracle
 * does its own material/light set-up and never shares code with the product.
 */
#ifndef VLR_ORACLE_H
#define VLR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OracleVertex {
    float position[3];
    float normal[3];
    float tc0Direction[3];
    float texCoord[2];
} OracleVertex; /* same 44-byte layout as VLRVertex */

enum OracleMaterialType {
    ORACLE_MAT_MATTE = 0,                /* spectra[0]=albedo */
    ORACLE_MAT_SPECULAR_REFLECTION = 1,  /* spectra: coeff, eta, k */
    ORACLE_MAT_SPECULAR_SCATTERING = 2,  /* spectra: coeff, eta ext, eta int */
    ORACLE_MAT_MICROFACET_REFLECTION = 3,/* spectra: eta, k; scalars: roughness, anisotropy, rotation */
    ORACLE_MAT_UE4 = 4,                  /* spectra[0]=base color; scalars: occlusion, roughness, metallic */
    ORACLE_MAT_DIFFUSE_EMITTER = 5,      /* spectra[0]=emittance; scalars[0]=scale */
    ORACLE_MAT_POINT_EMITTER = 6,        /* spectra[0]=intensity; scalars[0]=scale */
    ORACLE_MAT_ENVIRONMENT_EMITTER = 7,  /* spectra[0]=emittance or the scene texture; scalars[0]=scale */
    ORACLE_MAT_MICROFACET_SCATTERING = 8 /* spectra: coeff, eta ext, eta int; scalars: roughness, anisotropy, rotation */
};

/* colour spaces: 0 "Rec709(D65) sRGB Gamma", 1 "Rec709(D65)", 2 "XYZ", 3 "xyY" */
typedef struct OracleMaterial {
    int32_t type;
    int32_t colorSpace[3];
    float spectra[3][3];
    float scalars[4];
    int32_t useEnvTexture;
} OracleMaterial;

typedef struct OracleGeomInst {
    const OracleVertex* vertices;
    uint32_t numVertices;
    const uint32_t* indices;   /* 3 per triangle, or 1 per point when isPoints */
    uint32_t numIndices;
    int32_t material;
    int32_t isPoints;
    float alpha;               /* constant alpha of the group's alpha node; < 0: no alpha node */
} OracleGeomInst;

typedef struct OracleInstance {
    float transform[16];       /* column-major local-to-world */
    const uint32_t* geomInsts;
    uint32_t numGeomInsts;
} OracleInstance;

typedef struct OracleCamera {
    float position[3];
    float orientation[4];      /* quaternion x, y, z, w */
    float sensitivity, aspect, fovY, lensRadius, opDistance;
} OracleCamera;

typedef struct OracleScene {
    const OracleMaterial* materials; uint32_t numMaterials;
    const OracleGeomInst* geomInsts; uint32_t numGeomInsts;
    const OracleInstance* instances; uint32_t numInstances;
    int32_t envMaterial;       /* -1: no environment */
    const uint16_t* envTextureHalf; /* RGBA16F, "Light Source", Rec709(D65) linear */
    uint32_t envWidth, envHeight;
    float envRotation;
    OracleCamera camera;
} OracleScene;

typedef struct OracleRay { float org[3]; float tmin; float dir[3]; float tmax; } OracleRay;
typedef struct OracleHit { int32_t instIndex; int32_t geomInstIndex; int32_t primIndex; float t, b1, b2; } OracleHit;

typedef struct OracleRenderStats {
    uint64_t closestRays;
    uint64_t shadowRays;
    uint64_t paths;
    uint64_t droppedNonFinite;
} OracleRenderStats;

/* Returns 0 on success. tablesPath: vlr_amd/data/spectral_tables.bin */
int oracle_load_tables(const char* tablesPath);

void* oracle_scene_create(const OracleScene* desc, int spectral);
void oracle_scene_destroy(void* scene);
/* Scene bounds as the reference computes them (setup_scene.cu:51-92): center xyz, radius, discArea */
void oracle_scene_bounds(void* scene, float out[5]);

/* Closest hit for each ray.  useBVH=0 tests every triangle of every instance (brute force). */
int oracle_trace_closest(void* scene, const OracleRay* rays, uint32_t numRays, OracleHit* hits,
                         int useBVH, int numThreads);
/* Any hit in [tmin, tmax): occluded[i] = 1 if blocked. */
int oracle_trace_any(void* scene, const OracleRay* rays, uint32_t numRays, uint8_t* occluded,
                     int useBVH, int numThreads);

/* Renders numFrames frames (one sample per pixel each, like numFrames calls of
 * Context::render) into outRGBA (W*H*4 floats, the reference's linear output buffer).
 * rngState (W*H uint64) is read and written when non-null; otherwise the reference's
 * seeding (mt19937_64(591842031321323413), row-major) is used.
 * accumOut (optional, W*H*channels*2 floats, channels = 3 RGB / 16 spectral) receives the
 * Kahan (sum, compensation) pairs.  pixelRange (optional, 4 ints x0 y0 x1 y1) limits work
 * to a sub-rectangle for timing samples. */
int oracle_render(void* scene, uint32_t width, uint32_t height, uint32_t numFrames, uint32_t maxPathLength,
                  int numThreads, float* outRGBA, uint64_t* rngState, float* accumOut,
                  const int32_t* pixelRange, OracleRenderStats* stats);

/* LVC-BPT render (lvc_bpt.cu:144-971, context.cpp:1464-1494).  lightRngState: numLightPaths
 * PCG32 states of the light paths (seeded from mt19937_64(459182033132123413) when null). */
int oracle_render_bpt(void* scene, uint32_t width, uint32_t height, uint32_t numFrames, uint32_t maxPathLength,
                      uint32_t numLightPaths, int numThreads, float* outRGBA, uint64_t* rngState,
                      uint64_t* lightRngState, float* accumOut, OracleRenderStats* stats);
/* Light tracing (light_tracing.cu:49-263): numLightPaths light paths per frame, each splatting
 * its lens connections; lightRngState as for oracle_render_bpt. */
int oracle_render_lt(void* scene, uint32_t width, uint32_t height, uint32_t numFrames, uint32_t maxPathLength,
                     uint32_t numLightPaths, int numThreads, float* outRGBA, uint64_t* lightRngState,
                     float* accumOut, OracleRenderStats* stats);

/* BSDF parity helper: builds the BSDF of `mat` (immediate values only) for the given
 * wavelength set and evaluates sample / evaluate / evaluatePDF in the shading frame. */
typedef struct OracleBSDFQuery {
    float dirLocal[3];
    float geomNormalLocal[3];
    float wlOffset, wlSelect;    /* createWithEqualOffsets inputs */
    float uComponent, uDir[2];   /* sample inputs */
    float evalDir[3];            /* direction for evaluate / evaluatePDF */
    int32_t singleIsSelected;
    int32_t transportMode;       /* 0 radiance, 1 importance */
} OracleBSDFQuery;
typedef struct OracleBSDFResult {
    float sampledDir[3];
    float sampledPDF;
    uint32_t sampledType;
    float sampledValue[4];
    float evalValue[4];
    float evalPDF;
    float revValue[4];           /* bidirectional reverse quantities of evaluate/evaluatePDF */
    float revPDF;
    float sampledRevValue[4];    /* reverse quantities of the sample */
    float sampledRevPDF;
} OracleBSDFResult;
int oracle_bsdf_queries(const OracleMaterial* mat, int spectral, const OracleBSDFQuery* q, uint32_t n,
                        OracleBSDFResult* out);

/* Reference helpers exposed for golden-vector checks. */
uint32_t oracle_pcg32_stream(uint64_t state, uint32_t n, uint32_t* out, uint64_t* finalState);
void oracle_seed_states(uint32_t width, uint32_t height, uint64_t* out);
/* Light sampling as next-event estimation does it (light_transport_common.h:179-192,
 * triangle.cu:113-177, point.cu:38-80, infinite_sphere.cu:60-114): u holds 3 numbers per
 * sample (uLight, uPos0, uPos1).  areaPDF is per unit area (per unit solid angle / sin-free
 * for the environment, where position is the direction). */
typedef struct OracleLightSample {
    float position[3];
    float normal[3];
    float areaPDF;
    float lightProb;
    int32_t instIndex;      /* -1 for the environment */
    int32_t primIndex;
    int32_t atInfinity;
    int32_t valid;
} OracleLightSample;
void oracle_sample_lights(void* scene, uint32_t n, const float* u, OracleLightSample* out);
/* The environment along world directions (3 floats each): texture u, v, the solid-angle pdf
 * of environment light sampling, and the first channel of the bilinear texture fetch. */
void oracle_env_lookup(void* scene, uint32_t n, const float* dirs, float* out);
/* getFloat0cTo1o stream (random_distributions.h:23-26) */
void oracle_pcg32_floats(uint64_t state, uint32_t n, float* out);
/* the seeding generator of context.cpp:1182-1187 with an arbitrary seed */
void oracle_mt19937_64(uint64_t seed, uint32_t n, uint64_t* out);
void oracle_meng_eval(int spectrumType, int colorSpace, float e0, float e1, float e2,
                      const float lambdas[4], float out[4]);
/* UpsampledSpectrum(u, v, scale) (grid coordinates) evaluated at 4 wavelengths */
void oracle_meng_uvs_eval(float u, float v, float scale, const float lambdas[4], float out[4]);
void oracle_cmf_bins(float x[16], float y[16], float z[16], float* integral);
void oracle_triplet_rgb(int spectrumType, int colorSpace, float e0, float e1, float e2, float out[3]);

#ifdef __cplusplus
}
#endif
#endif
