/* A C caller of libvlr_amd.so: builds the Cornell box through the VLR C API (vlr.h), renders
 * it with vlrContextRender and writes the RGBA32F output as raw floats.  It also reads the
 * output through the array that vlrContextGetOutputBuffer returns and checks that it holds
 * the same values as vlrContextReadOutputBuffer.
 *
 *     cornell <out.f32> <width> <height> <frames> <max path length> <angle short box> <angle tall box>
 *
 * The scene is the one of vlr_amd.scenes.cornell_box with the two box rotations given (radians):
 * 5 wall quads (white floor, back wall and ceiling, red left, green right), a 0.25 x 0.25 light
 * quad of emittance pi (17, 12, 4) and two white boxes; camera at (0, 1.5, 6) turned by pi about
 * +Y, vertical field of view 40 degrees (HostProgram/scene.cpp:538-571, 789-801 of the
 * reference). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>
#include "vlr/vlr.h"

#define CHECK(call)                                                                         \
    do {                                                                                    \
        VLRResult r_ = (call);                                                              \
        if (r_ != VLRResult_NoError) {                                                      \
            fprintf(stderr, "%s failed: %s (%s:%d)\n", #call, vlrGetErrorMessage(r_), __FILE__, \
                    __LINE__);                                                              \
            exit(1);                                                                        \
        }                                                                                   \
    } while (0)

static void quad(VLRVertex *v, const float p[4][3], const float n[3], const float t[3]) {
    static const float uv[4][2] = {{0, 1}, {1, 1}, {1, 0}, {0, 0}};
    for (int i = 0; i < 4; ++i) {
        v[i].position.x = p[i][0]; v[i].position.y = p[i][1]; v[i].position.z = p[i][2];
        v[i].normal.x = n[0]; v[i].normal.y = n[1]; v[i].normal.z = n[2];
        v[i].tc0Direction.x = t[0]; v[i].tc0Direction.y = t[1]; v[i].tc0Direction.z = t[2];
        v[i].texCoord.u = uv[i][0]; v[i].texCoord.v = uv[i][1];
    }
}

static void quadIndices(uint32_t *idx, int numQuads) {
    for (int q = 0; q < numQuads; ++q) {
        const uint32_t b = 4 * q, k[6] = {b, b + 1, b + 2, b, b + 2, b + 3};
        memcpy(idx + 6 * q, k, sizeof(k));
    }
}

/* axis-aligned box [-sx/2, sx/2] x [0, sy] x [-sz/2, sz/2], 6 quads */
static void boxMesh(VLRVertex *v, float sx, float sy, float sz) {
    const float hx = sx / 2, hz = sz / 2;
    const float top[4][3] = {{-hx, sy, -hz}, {-hx, sy, hz}, {hx, sy, hz}, {hx, sy, -hz}};
    const float bottom[4][3] = {{-hx, 0, hz}, {-hx, 0, -hz}, {hx, 0, -hz}, {hx, 0, hz}};
    const float pz[4][3] = {{-hx, 0, hz}, {hx, 0, hz}, {hx, sy, hz}, {-hx, sy, hz}};
    const float nz[4][3] = {{hx, 0, -hz}, {-hx, 0, -hz}, {-hx, sy, -hz}, {hx, sy, -hz}};
    const float px[4][3] = {{hx, 0, hz}, {hx, 0, -hz}, {hx, sy, -hz}, {hx, sy, hz}};
    const float nx[4][3] = {{-hx, 0, -hz}, {-hx, 0, hz}, {-hx, sy, hz}, {-hx, sy, -hz}};
    const float up[3] = {0, 1, 0}, down[3] = {0, -1, 0}, fwd[3] = {0, 0, 1}, back[3] = {0, 0, -1};
    const float right[3] = {1, 0, 0}, left[3] = {-1, 0, 0};
    quad(v + 0, top, up, right);
    quad(v + 4, bottom, down, right);
    quad(v + 8, pz, fwd, right);
    quad(v + 12, nz, back, left);
    quad(v + 16, px, right, back);
    quad(v + 20, nx, left, fwd);
}

/* column-major translate(t) * rotateY(a), computed in double and rounded once */
static void boxTransform(float m[16], double tx, double ty, double tz, double a) {
    const double c = cos(a), s = sin(a);
    const double r[16] = {c, 0, -s, 0, 0, 1, 0, 0, s, 0, c, 0, tx, ty, tz, 1};
    for (int i = 0; i < 16; ++i) m[i] = (float)r[i];
}

static VLRSurfaceMaterial matte(VLRContext ctx, float r, float g, float b) {
    VLRSurfaceMaterial m;
    CHECK(vlrSurfaceMaterialCreate(ctx, "Matte", &m));
    VLRImmediateSpectrum s = {"Rec709(D65)", r, g, b};
    CHECK(vlrQueryableSetImmediateSpectrum((VLRQueryable)m, "albedo", &s));
    return m;
}

int main(int argc, char **argv) {
    if (argc != 8) {
        fprintf(stderr, "usage: %s out.f32 width height frames maxPathLength angle0 angle1\n", argv[0]);
        return 2;
    }
    const uint32_t W = (uint32_t)atoi(argv[2]), H = (uint32_t)atoi(argv[3]), frames = (uint32_t)atoi(argv[4]);
    const uint32_t maxLen = (uint32_t)atoi(argv[5]);
    const double a0 = strtod(argv[6], NULL), a1 = strtod(argv[7], NULL);
    const double pi = 3.14159265358979323846;

    VLRContext ctx;
    CHECK(vlrCreateContext(NULL, false, 8, &ctx));
    CHECK(vlrContextSetMaxPathLength(ctx, maxLen));

    VLRSurfaceMaterial white = matte(ctx, 0.73f, 0.73f, 0.73f), red = matte(ctx, 0.65f, 0.05f, 0.05f),
                       green = matte(ctx, 0.12f, 0.45f, 0.15f);
    VLRSurfaceMaterial light;
    CHECK(vlrSurfaceMaterialCreate(ctx, "DiffuseEmitter", &light));
    VLRImmediateSpectrum e = {"Rec709(D65)", (float)(pi * 17), (float)(pi * 12), (float)(pi * 4)};
    CHECK(vlrQueryableSetImmediateSpectrum((VLRQueryable)light, "emittance", &e));
    CHECK(vlrQueryableSetFloat((VLRQueryable)light, "scale", 1.0f));

    VLRVertex walls[20];
    {
        const float floorP[4][3] = {{-1.5f, 0, -1.5f}, {-1.5f, 0, 1.5f}, {1.5f, 0, 1.5f}, {1.5f, 0, -1.5f}};
        const float backP[4][3] = {{-1.5f, 0, -1.5f}, {1.5f, 0, -1.5f}, {1.5f, 3, -1.5f}, {-1.5f, 3, -1.5f}};
        const float ceilP[4][3] = {{-1.5f, 3, -1.5f}, {1.5f, 3, -1.5f}, {1.5f, 3, 1.5f}, {-1.5f, 3, 1.5f}};
        const float leftP[4][3] = {{-1.5f, 0, 1.5f}, {-1.5f, 0, -1.5f}, {-1.5f, 3, -1.5f}, {-1.5f, 3, 1.5f}};
        const float rightP[4][3] = {{1.5f, 0, -1.5f}, {1.5f, 0, 1.5f}, {1.5f, 3, 1.5f}, {1.5f, 3, -1.5f}};
        const float up[3] = {0, 1, 0}, down[3] = {0, -1, 0}, fwd[3] = {0, 0, 1}, xp[3] = {1, 0, 0};
        const float xn[3] = {-1, 0, 0}, zn[3] = {0, 0, -1}, zp[3] = {0, 0, 1};
        quad(walls + 0, floorP, up, xp);
        quad(walls + 4, backP, fwd, xp);
        quad(walls + 8, ceilP, down, xp);
        quad(walls + 12, leftP, xp, zn);
        quad(walls + 16, rightP, xn, zp);
    }
    uint32_t wallIdx[30];
    quadIndices(wallIdx, 5);
    VLRShaderNodePlug none;
    memset(&none, 0, sizeof(none));

    VLRTriangleMeshSurfaceNode room, lightMesh, shortBox, tallBox;
    CHECK(vlrTriangleMeshSurfaceNodeCreate(ctx, "CornellBox", &room));
    CHECK(vlrTriangleMeshSurfaceNodeSetVertices(room, walls, 20));
    CHECK(vlrTriangleMeshSurfaceNodeAddMaterialGroup(room, wallIdx, 18, white, none, none, none));
    CHECK(vlrTriangleMeshSurfaceNodeAddMaterialGroup(room, wallIdx + 18, 6, red, none, none, none));
    CHECK(vlrTriangleMeshSurfaceNodeAddMaterialGroup(room, wallIdx + 24, 6, green, none, none, none));

    VLRVertex lv[4];
    {
        const float h = 0.125f;
        const float p[4][3] = {{-h, 2.99f, -h}, {h, 2.99f, -h}, {h, 2.99f, h}, {-h, 2.99f, h}};
        const float down[3] = {0, -1, 0}, xp[3] = {1, 0, 0};
        quad(lv, p, down, xp);
    }
    uint32_t lightIdx[6];
    quadIndices(lightIdx, 1);
    CHECK(vlrTriangleMeshSurfaceNodeCreate(ctx, "Light", &lightMesh));
    CHECK(vlrTriangleMeshSurfaceNodeSetVertices(lightMesh, lv, 4));
    CHECK(vlrTriangleMeshSurfaceNodeAddMaterialGroup(lightMesh, lightIdx, 6, light, none, none, none));

    VLRVertex sb[24], tb[24];
    uint32_t boxIdx[36];
    boxMesh(sb, 0.6f, 0.6f, 0.6f);
    boxMesh(tb, 0.6f, 1.2f, 0.6f);
    quadIndices(boxIdx, 6);
    CHECK(vlrTriangleMeshSurfaceNodeCreate(ctx, "ShortBox", &shortBox));
    CHECK(vlrTriangleMeshSurfaceNodeSetVertices(shortBox, sb, 24));
    CHECK(vlrTriangleMeshSurfaceNodeAddMaterialGroup(shortBox, boxIdx, 36, white, none, none, none));
    CHECK(vlrTriangleMeshSurfaceNodeCreate(ctx, "TallBox", &tallBox));
    CHECK(vlrTriangleMeshSurfaceNodeSetVertices(tallBox, tb, 24));
    CHECK(vlrTriangleMeshSurfaceNodeAddMaterialGroup(tallBox, boxIdx, 36, white, none, none, none));

    float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}, ms[16], mt[16];
    boxTransform(ms, 0.55, 0.0, 0.45, -a0);
    boxTransform(mt, -0.5, 0.0, -0.4, a1);
    VLRStaticTransform tIdent, tShort, tTall;
    CHECK(vlrStaticTransformCreate(ctx, ident, &tIdent));
    CHECK(vlrStaticTransformCreate(ctx, ms, &tShort));
    CHECK(vlrStaticTransformCreate(ctx, mt, &tTall));
    VLRInternalNode roomInst, shortInst, tallInst;
    CHECK(vlrInternalNodeCreate(ctx, "Room", (VLRTransformConst)tIdent, &roomInst));
    CHECK(vlrParentNodeAddChild((VLRParentNode)roomInst, (VLRNode)room));
    CHECK(vlrParentNodeAddChild((VLRParentNode)roomInst, (VLRNode)lightMesh));
    CHECK(vlrInternalNodeCreate(ctx, "ShortBoxInst", (VLRTransformConst)tShort, &shortInst));
    CHECK(vlrParentNodeAddChild((VLRParentNode)shortInst, (VLRNode)shortBox));
    CHECK(vlrInternalNodeCreate(ctx, "TallBoxInst", (VLRTransformConst)tTall, &tallInst));
    CHECK(vlrParentNodeAddChild((VLRParentNode)tallInst, (VLRNode)tallBox));
    VLRScene scene;
    CHECK(vlrSceneCreate(ctx, (VLRTransformConst)tIdent, &scene));
    CHECK(vlrParentNodeAddChild((VLRParentNode)scene, (VLRNode)roomInst));
    CHECK(vlrParentNodeAddChild((VLRParentNode)scene, (VLRNode)shortInst));
    CHECK(vlrParentNodeAddChild((VLRParentNode)scene, (VLRNode)tallInst));

    VLRCamera camera;
    CHECK(vlrCameraCreate(ctx, "Perspective", &camera));
    VLRPoint3D pos = {0.0f, 1.5f, 6.0f};
    VLRQuaternion ori = {0.0f, (float)sin(pi / 2), 0.0f, (float)cos(pi / 2)};
    CHECK(vlrQueryableSetPoint3D((VLRQueryable)camera, "position", &pos));
    CHECK(vlrQueryableSetQuaternion((VLRQueryable)camera, "orientation", &ori));
    CHECK(vlrQueryableSetFloat((VLRQueryable)camera, "aspect", 1.0f));
    CHECK(vlrQueryableSetFloat((VLRQueryable)camera, "sensitivity", 1.0f));
    CHECK(vlrQueryableSetFloat((VLRQueryable)camera, "fovy", (float)(40.0 * (pi / 180.0))));
    CHECK(vlrQueryableSetFloat((VLRQueryable)camera, "lens radius", 0.0f));
    CHECK(vlrQueryableSetFloat((VLRQueryable)camera, "op distance", 1.0f));

    CHECK(vlrContextBindOutputBuffer(ctx, W, H, 0));
    CHECK(vlrContextSetScene(ctx, scene));
    CHECK(vlrContextSetRenderer(ctx, VLRRenderer_PathTracing));
    /* the output as a hipArray, where the GPU has arrays (image support) */
    int imageSupport = 0;
    if (hipDeviceGetAttribute(&imageSupport, hipDeviceAttributeImageSupport, 0) != hipSuccess) imageSupport = 0;
    printf("image support: %d\n", imageSupport);
    hipArray_t array = NULL;
    if (imageSupport) {
        CHECK(vlrContextGetOutputBuffer(ctx, &array));
    } else if (vlrContextGetOutputBuffer(ctx, &array) != VLRResult_InternalError) {
        fprintf(stderr, "vlrContextGetOutputBuffer should fail without image support\n");
        return 1;
    }
    uint32_t accumulated = 0;
    for (uint32_t f = 0; f < frames; ++f)
        CHECK(vlrContextRender(ctx, NULL, camera, false, 1, f == 0, frames, &accumulated));

    float *rgba = (float *)malloc(sizeof(float) * 4 * W * H), *viaArray = (float *)malloc(sizeof(float) * 4 * W * H);
    CHECK(vlrContextReadOutputBuffer(ctx, rgba));
    if (imageSupport) {
        if (hipMemcpy2DFromArray(viaArray, sizeof(float) * 4 * W, array, 0, 0, sizeof(float) * 4 * W, H,
                                 hipMemcpyDeviceToHost) != hipSuccess) {
            fprintf(stderr, "reading the output array failed\n");
            return 1;
        }
        if (memcmp(rgba, viaArray, sizeof(float) * 4 * W * H) != 0) {
            fprintf(stderr, "the output array differs from the output buffer\n");
            return 1;
        }
    }
    FILE *f = fopen(argv[1], "wb");
    if (!f || fwrite(rgba, sizeof(float) * 4, (size_t)W * H, f) != (size_t)W * H) {
        fprintf(stderr, "cannot write %s\n", argv[1]);
        return 1;
    }
    fclose(f);
    if (imageSupport)
        printf("frames %u, output array matches the output buffer\n", accumulated);
    else
        printf("frames %u, output array refused (no image support)\n", accumulated);
    free(rgba);
    free(viaArray);
    CHECK(vlrDestroyContext(ctx));
    return 0;
}
