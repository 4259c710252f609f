// A C++ caller of libvlr_amd.so through the vlr::Context wrapper (vlr/vlrcpp.h): the same
// Cornell box as cornell.c, built with the wrapper's objects (shared references, exceptions
// on errors), rendered with Context::render and written as raw RGBA32F floats.
//
//     cornell_cpp <out.f32> <width> <height> <frames> <max path length> <angle short box> <angle tall box>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "vlr/vlrcpp.h"

using namespace vlr;

static void quad(std::vector<Vertex> &v, const Point3D (&p)[4], const Normal3D &n, const Vector3D &t) {
    const TexCoord2D uv[4] = {TexCoord2D(0, 1), TexCoord2D(1, 1), TexCoord2D(1, 0), TexCoord2D(0, 0)};
    for (int i = 0; i < 4; ++i) v.push_back(Vertex{p[i], n, t, uv[i]});
}

static std::vector<uint32_t> quadIndices(int numQuads) {
    std::vector<uint32_t> idx;
    for (uint32_t q = 0; q < static_cast<uint32_t>(numQuads); ++q)
        for (uint32_t k : {0u, 1u, 2u, 0u, 2u, 3u}) idx.push_back(4 * q + k);
    return idx;
}

static std::vector<Vertex> boxMesh(float sx, float sy, float sz) {
    const float hx = sx / 2, hz = sz / 2;
    std::vector<Vertex> v;
    quad(v, {Point3D(-hx, sy, -hz), Point3D(-hx, sy, hz), Point3D(hx, sy, hz), Point3D(hx, sy, -hz)},
         Normal3D(0, 1, 0), Vector3D(1, 0, 0));
    quad(v, {Point3D(-hx, 0, hz), Point3D(-hx, 0, -hz), Point3D(hx, 0, -hz), Point3D(hx, 0, hz)},
         Normal3D(0, -1, 0), Vector3D(1, 0, 0));
    quad(v, {Point3D(-hx, 0, hz), Point3D(hx, 0, hz), Point3D(hx, sy, hz), Point3D(-hx, sy, hz)},
         Normal3D(0, 0, 1), Vector3D(1, 0, 0));
    quad(v, {Point3D(hx, 0, -hz), Point3D(-hx, 0, -hz), Point3D(-hx, sy, -hz), Point3D(hx, sy, -hz)},
         Normal3D(0, 0, -1), Vector3D(-1, 0, 0));
    quad(v, {Point3D(hx, 0, hz), Point3D(hx, 0, -hz), Point3D(hx, sy, -hz), Point3D(hx, sy, hz)},
         Normal3D(1, 0, 0), Vector3D(0, 0, -1));
    quad(v, {Point3D(-hx, 0, -hz), Point3D(-hx, 0, hz), Point3D(-hx, sy, hz), Point3D(-hx, sy, -hz)},
         Normal3D(-1, 0, 0), Vector3D(0, 0, 1));
    return v;
}

// column-major translate(t) * rotateY(a), computed in double and rounded once
static StaticTransformRef boxTransform(const ContextRef &ctx, double tx, double ty, double tz, double a) {
    const double c = std::cos(a), s = std::sin(a);
    const double r[16] = {c, 0, -s, 0, 0, 1, 0, 0, s, 0, c, 0, tx, ty, tz, 1};
    float m[16];
    for (int i = 0; i < 16; ++i) m[i] = static_cast<float>(r[i]);
    return ctx->createStaticTransform(m);
}

int main(int argc, char **argv) {
    if (argc != 8) {
        std::fprintf(stderr, "usage: %s out.f32 width height frames maxPathLength angle0 angle1\n", argv[0]);
        return 2;
    }
    const uint32_t W = std::atoi(argv[2]), H = std::atoi(argv[3]), frames = std::atoi(argv[4]);
    const uint32_t maxLen = std::atoi(argv[5]);
    const double a0 = std::strtod(argv[6], nullptr), a1 = std::strtod(argv[7], nullptr);
    const double pi = 3.14159265358979323846;
    try {
        ContextRef context = Context::create(nullptr, false);
        context->enableAllExceptions();
        if (vlrContextSetMaxPathLength(context->get(), maxLen) != VLRResult_NoError) return 1;

        auto matte = [&](float r, float g, float b) {
            SurfaceMaterialRef m = context->createSurfaceMaterial("Matte");
            m->set("albedo", VLRImmediateSpectrum{"Rec709(D65)", r, g, b});
            return m;
        };
        SurfaceMaterialRef white = matte(0.73f, 0.73f, 0.73f), red = matte(0.65f, 0.05f, 0.05f),
                           green = matte(0.12f, 0.45f, 0.15f);
        SurfaceMaterialRef light = context->createSurfaceMaterial("DiffuseEmitter");
        light->set("emittance", VLRImmediateSpectrum{"Rec709(D65)", static_cast<float>(pi * 17),
                                                     static_cast<float>(pi * 12), static_cast<float>(pi * 4)});
        light->set("scale", 1.0f);

        std::vector<Vertex> walls;
        quad(walls, {Point3D(-1.5f, 0, -1.5f), Point3D(-1.5f, 0, 1.5f), Point3D(1.5f, 0, 1.5f), Point3D(1.5f, 0, -1.5f)},
             Normal3D(0, 1, 0), Vector3D(1, 0, 0));
        quad(walls, {Point3D(-1.5f, 0, -1.5f), Point3D(1.5f, 0, -1.5f), Point3D(1.5f, 3, -1.5f), Point3D(-1.5f, 3, -1.5f)},
             Normal3D(0, 0, 1), Vector3D(1, 0, 0));
        quad(walls, {Point3D(-1.5f, 3, -1.5f), Point3D(1.5f, 3, -1.5f), Point3D(1.5f, 3, 1.5f), Point3D(-1.5f, 3, 1.5f)},
             Normal3D(0, -1, 0), Vector3D(1, 0, 0));
        quad(walls, {Point3D(-1.5f, 0, 1.5f), Point3D(-1.5f, 0, -1.5f), Point3D(-1.5f, 3, -1.5f), Point3D(-1.5f, 3, 1.5f)},
             Normal3D(1, 0, 0), Vector3D(0, 0, -1));
        quad(walls, {Point3D(1.5f, 0, -1.5f), Point3D(1.5f, 0, 1.5f), Point3D(1.5f, 3, 1.5f), Point3D(1.5f, 3, -1.5f)},
             Normal3D(-1, 0, 0), Vector3D(0, 0, 1));
        std::vector<uint32_t> wallIdx = quadIndices(5);
        TriangleMeshSurfaceNodeRef room = context->createTriangleMeshSurfaceNode("CornellBox");
        room->setVertices(walls.data(), static_cast<uint32_t>(walls.size()));
        room->addMaterialGroup(wallIdx.data(), 18, white);
        room->addMaterialGroup(wallIdx.data() + 18, 6, red);
        room->addMaterialGroup(wallIdx.data() + 24, 6, green);

        std::vector<Vertex> lv;
        const float h = 0.125f;
        quad(lv, {Point3D(-h, 2.99f, -h), Point3D(h, 2.99f, -h), Point3D(h, 2.99f, h), Point3D(-h, 2.99f, h)},
             Normal3D(0, -1, 0), Vector3D(1, 0, 0));
        std::vector<uint32_t> lightIdx = quadIndices(1);
        TriangleMeshSurfaceNodeRef lightMesh = context->createTriangleMeshSurfaceNode("Light");
        lightMesh->setVertices(lv.data(), 4);
        lightMesh->addMaterialGroup(lightIdx.data(), 6, light);

        std::vector<Vertex> sb = boxMesh(0.6f, 0.6f, 0.6f), tb = boxMesh(0.6f, 1.2f, 0.6f);
        std::vector<uint32_t> boxIdx = quadIndices(6);
        TriangleMeshSurfaceNodeRef shortBox = context->createTriangleMeshSurfaceNode("ShortBox");
        shortBox->setVertices(sb.data(), 24);
        shortBox->addMaterialGroup(boxIdx.data(), 36, white);
        TriangleMeshSurfaceNodeRef tallBox = context->createTriangleMeshSurfaceNode("TallBox");
        tallBox->setVertices(tb.data(), 24);
        tallBox->addMaterialGroup(boxIdx.data(), 36, white);

        InternalNodeRef roomInst = context->createInternalNode("Room");
        roomInst->addChild(room);
        roomInst->addChild(lightMesh);
        InternalNodeRef shortInst = context->createInternalNode("ShortBoxInst", boxTransform(context, 0.55, 0.0, 0.45, -a0));
        shortInst->addChild(shortBox);
        InternalNodeRef tallInst = context->createInternalNode("TallBoxInst", boxTransform(context, -0.5, 0.0, -0.4, a1));
        tallInst->addChild(tallBox);
        SceneRef scene = context->createScene();
        scene->addChild(roomInst);
        scene->addChild(shortInst);
        scene->addChild(tallInst);

        CameraRef camera = context->createCamera("Perspective");
        camera->set("position", Point3D(0.0f, 1.5f, 6.0f));
        camera->set("orientation", Quaternion(0.0f, static_cast<float>(std::sin(pi / 2)), 0.0f,
                                              static_cast<float>(std::cos(pi / 2))));
        camera->set("aspect", 1.0f);
        camera->set("sensitivity", 1.0f);
        camera->set("fovy", static_cast<float>(40.0 * (pi / 180.0)));
        camera->set("lens radius", 0.0f);
        camera->set("op distance", 1.0f);

        context->bindOutputBuffer(W, H, 0);
        context->setScene(scene);
        context->setRenderer(VLRRenderer_PathTracing);
        uint32_t accumulated = 0;
        for (uint32_t f = 0; f < frames; ++f) context->render(nullptr, camera, false, 1, f == 0, frames, &accumulated);
        std::vector<float> rgba(size_t(4) * W * H);
        context->readOutputBuffer(rgba.data());
        FILE *out = std::fopen(argv[1], "wb");
        if (!out || std::fwrite(rgba.data(), sizeof(float) * 4, size_t(W) * H, out) != size_t(W) * H) {
            std::fprintf(stderr, "cannot write %s\n", argv[1]);
            return 1;
        }
        std::fclose(out);
        // an invalid call raises once its code is enabled as an exception
        bool threw = false;
        try {
            context->createSurfaceMaterial("NoSuchMaterial");
        } catch (const std::runtime_error &) {
            threw = true;
        }
        std::printf("frames %u, invalid material %s\n", accumulated, threw ? "raised" : "did not raise");
        return threw ? 0 : 1;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
