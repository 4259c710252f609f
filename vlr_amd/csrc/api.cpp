// extern "C" entry points of include/vlr/vlr.h and include/vlr/vlr_amd.h.
//
// Error semantics follow the reference's libVLR/vlr.cpp:23-68: a null handle returns
// VLRResult_InvalidArgument, a handle of the wrong type VLRResult_InvalidInstance, any
// exception VLRResult_InternalError; unsupported parameter names return InvalidArgument.
#include "runtime.h"

#include "../../include/vlr/vlr.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

using namespace vlra;

namespace {

thread_local std::string g_lastError;

template <typename H>
inline Object *obj(H h) {
    return reinterpret_cast<Object *>(const_cast<void *>(static_cast<const void *>(h)));
}

inline bool alive(const Object *o) { return o && o->magic == kObjMagic; }

#define VLRA_TRY try {
#define VLRA_CATCH                                                   \
    }                                                                \
    catch (const ApiError &e) {                                      \
        g_lastError = e.msg;                                         \
        return static_cast<VLRResult>(e.code);                       \
    }                                                                \
    catch (const std::exception &e) {                                \
        g_lastError = e.what();                                      \
        return VLRResult_InternalError;                              \
    }                                                                \
    catch (...) {                                                    \
        return VLRResult_InternalError;                              \
    }

#define CHECK(h, pred)                                               \
    do {                                                             \
        if ((h) == nullptr) return VLRResult_InvalidArgument;        \
        Object *o_ = obj(h);                                         \
        if (!alive(o_) || !(pred)) return VLRResult_InvalidInstance; \
    } while (0)

#define IS(h, T) (obj(h)->type == ObjType::T)
#define CTX(h) static_cast<Context *>(obj(h))
#define QRY(h) static_cast<Queryable *>(obj(h))

template <typename T>
T *track(Context *c, T *o) {
    c->objects.insert(o);
    return o;
}

void untrack(Object *o) {
    if (o->ctx) o->ctx->objects.erase(o);
}

} // namespace

namespace {
// queryable.cpp:31-80
const std::vector<std::pair<std::string, std::vector<const char *>>> &enumTables() {
    static std::vector<std::pair<std::string, std::vector<const char *>>> t = {
        {"SpectrumType", {"Reflectance", "Transmittance", "Light Source", "NA"}},
        {"ColorSpace", {"Rec709(D65) sRGB Gamma", "Rec709(D65)", "XYZ", "xyY"}},
        {"DataFormat", {"RGB8x3", "RGB_8x4", "RGBA8x4", "RGBA16Fx4", "RGBA32Fx4", "RG32Fx2", "Gray32F", "Gray8",
                        "GrayA8x2", "BC1", "BC2", "BC3", "BC4", "BC4_Signed", "BC5", "BC5_Signed", "BC6H",
                        "BC6H_Signed", "BC7"}},
        {"BumpType", {"Normal Map (DirectX)", "Normal Map (OpenGL)", "Height Map"}},
        {"TextureFilter", {"Nearest", "Linear", "None"}},
        {"TextureWrapMode", {"Repeat", "Clamp to Edge", "Mirror", "Clamp to Border"}},
        {"TangentType", {"TC0 Direction", "Radial X", "Radial Y", "Radial Z"}},
    };
    return t;
}
} // namespace

extern "C" {

VLR_API const char *vlrGetErrorMessage(VLRResult code) {
    switch (code) {
    case VLRResult_NoError: return "No Error";
    case VLRResult_InvalidContext: return "Invalid Context";
    case VLRResult_InvalidInstance: return "Invalid Instance";
    case VLRResult_InvalidArgument: return "Invalid Argument";
    case VLRResult_IncompatibleNodeType: return "Incompatible Node Type";
    case VLRResult_InternalError: return "Internal Error";
    default: return "";
    }
}

// ------------------------------------------------------------------------------------------ Context
VLR_API VLRResult vlrCreateContext(hipCtx_t hipContext, bool logging, uint32_t maxCallableDepth, VLRContext *context) {
    VLRA_TRY
    (void)logging;
    (void)maxCallableDepth;
    if (!context) return VLRResult_InvalidArgument;
    *context = reinterpret_cast<VLRContext>(new Context(hipContext));
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrDestroyContext(VLRContext context) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    delete CTX(context);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextGetCUcontext(VLRContext context, hipCtx_t *hipContext) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!hipContext) return VLRResult_InvalidArgument;
    *hipContext = nullptr;
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextBindOutputBuffer(VLRContext context, uint32_t width, uint32_t height, uint32_t glTexID) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (glTexID != 0) return VLRResult_InvalidArgument;
    CTX(context)->bindOutputBuffer(width, height);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextGetOutputBufferSize(VLRContext context, uint32_t *width, uint32_t *height) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!width || !height) return VLRResult_InvalidArgument;
    *width = CTX(context)->width;
    *height = CTX(context)->height;
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextGetOutputBuffer(VLRContext context, hipArray_t *array) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!array || !CTX(context)->width) return VLRResult_InvalidArgument;
    *array = CTX(context)->outputArrayHandle();
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextReadOutputBuffer(VLRContext context, float *data) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!data) return VLRResult_InvalidArgument;
    CTX(context)->readOutput(data);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextSetScene(VLRContext context, VLRScene scene) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CHECK(scene, IS(scene, Scene));
    CTX(context)->scene = static_cast<Scene *>(obj(scene));
    CTX(context)->touch();
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextSetRenderer(VLRContext context, VLRRenderer renderer) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->renderer = static_cast<uint32_t>(renderer);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextSetDebugRenderingAttribute(VLRContext context, VLRDebugRenderingMode attr) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->debugAttr = static_cast<uint32_t>(attr);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextSetProbePixel(VLRContext context, int32_t x, int32_t y) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->probeX = x;
    CTX(context)->probeY = y;
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextRender(VLRContext context, hipStream_t stream, VLRCameraConst camera, bool denoise,
                                   uint32_t shrinkCoeff, bool firstFrame, uint32_t limitNumAccumFrames,
                                   uint32_t *numAccumFrames) {
    VLRA_TRY
    (void)denoise;
    CHECK(context, IS(context, Context));
    CHECK(camera, IS(camera, Camera));
    if (!numAccumFrames) return VLRResult_InvalidArgument;
    CTX(context)->render(stream, static_cast<const Camera *>(obj(camera)), shrinkCoeff, firstFrame,
                         limitNumAccumFrames, numAccumFrames);
    return VLRResult_NoError;
    VLRA_CATCH
}

// ------------------------------------------------------------------------------------------ Object / enums
VLR_API VLRResult vlrObjectGetType(VLRObjectConst object, const char **typeName) {
    VLRA_TRY
    CHECK(object, true);
    if (!typeName) return VLRResult_InvalidArgument;
    *typeName = obj(object)->typeName();
    return VLRResult_NoError;
    VLRA_CATCH
}

#define PI(h) static_cast<const ParamInfo *>(obj(h))
VLR_API VLRResult vlrParameterInfoGetName(VLRParameterInfoConst p, const char **name) {
    VLRA_TRY
    CHECK(p, IS(p, ParameterInfo));
    if (!name) return VLRResult_InvalidArgument;
    *name = PI(p)->name.c_str();
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrParameterInfoGetSocketForm(VLRParameterInfoConst p, VLRParameterFormFlag *form) {
    VLRA_TRY
    CHECK(p, IS(p, ParameterInfo));
    if (!form) return VLRResult_InvalidArgument;
    *form = static_cast<VLRParameterFormFlag>(PI(p)->formFlags);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrParameterInfoGetType(VLRParameterInfoConst p, const char **type) {
    VLRA_TRY
    CHECK(p, IS(p, ParameterInfo));
    if (!type) return VLRResult_InvalidArgument;
    *type = PI(p)->typeStr.c_str();
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrParameterInfoGetTupleSize(VLRParameterInfoConst p, uint32_t *size) {
    VLRA_TRY
    CHECK(p, IS(p, ParameterInfo));
    if (!size) return VLRResult_InvalidArgument;
    *size = PI(p)->tupleSize;
    return VLRResult_NoError;
    VLRA_CATCH
}


VLR_API VLRResult vlrGetNumEnumMembers(const char *typeName, uint32_t *numMembers) {
    VLRA_TRY
    if (!typeName || !numMembers) return VLRResult_InvalidArgument;
    for (auto &e : enumTables())
        if (e.first == typeName) {
            *numMembers = static_cast<uint32_t>(e.second.size());
            return VLRResult_NoError;
        }
    return VLRResult_InvalidArgument;
    VLRA_CATCH
}

VLR_API VLRResult vlrGetEnumMember(const char *typeName, uint32_t index, const char **value) {
    VLRA_TRY
    if (!typeName || !value) return VLRResult_InvalidArgument;
    for (auto &e : enumTables())
        if (e.first == typeName) {
            if (index >= e.second.size()) return VLRResult_InvalidArgument;
            *value = e.second[index];
            return VLRResult_NoError;
        }
    return VLRResult_InvalidArgument;
    VLRA_CATCH
}

// ------------------------------------------------------------------------------------------ Queryable
VLR_API VLRResult vlrQueryableGetNumParameters(VLRQueryableConst q, uint32_t *numParams) {
    VLRA_TRY
    CHECK(q, obj(q)->isQueryable());
    if (!numParams) return VLRResult_InvalidArgument;
    *numParams = static_cast<uint32_t>(QRY(q)->params().size());
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrQueryableGetParameterInfo(VLRQueryableConst q, uint32_t index, VLRParameterInfoConst *info) {
    VLRA_TRY
    CHECK(q, obj(q)->isQueryable());
    if (!info) return VLRResult_InvalidArgument;
    const auto &ps = QRY(q)->params();
    if (index >= ps.size()) return VLRResult_InvalidArgument;
    *info = reinterpret_cast<VLRParameterInfoConst>(ps[index]);
    return VLRResult_NoError;
    VLRA_CATCH
}

#define QGET(fn, call)                                         \
    VLRA_TRY                                                   \
    CHECK(q, obj(q)->isQueryable());                           \
    if (!paramName || !out) return VLRResult_InvalidArgument;  \
    if (!(call)) return VLRResult_InvalidArgument;             \
    return VLRResult_NoError;                                  \
    VLRA_CATCH

VLR_API VLRResult vlrQueryableGetEnumValue(VLRQueryableConst q, const char *paramName, const char **out) {
    QGET(getEnum, QRY(q)->getEnum(paramName, out))
}
VLR_API VLRResult vlrQueryableGetPoint3D(VLRQueryableConst q, const char *paramName, VLRPoint3D *out) {
    QGET(getPoint, QRY(q)->getPoint(paramName, &out->x))
}
VLR_API VLRResult vlrQueryableGetVector3D(VLRQueryableConst q, const char *paramName, VLRVector3D *out) {
    QGET(getVector, QRY(q)->getVector(paramName, &out->x))
}
VLR_API VLRResult vlrQueryableGetNormal3D(VLRQueryableConst q, const char *paramName, VLRNormal3D *out) {
    QGET(getNormal, QRY(q)->getNormal(paramName, &out->x))
}
VLR_API VLRResult vlrQueryableGetQuaternion(VLRQueryableConst q, const char *paramName, VLRQuaternion *out) {
    QGET(getQuaternion, QRY(q)->getQuaternion(paramName, &out->x))
}
VLR_API VLRResult vlrQueryableGetFloat(VLRQueryableConst q, const char *paramName, float *out) {
    QGET(getFloat, QRY(q)->getFloatTuple(paramName, out, 1))
}
VLR_API VLRResult vlrQueryableGetFloatTuple(VLRQueryableConst q, const char *paramName, float *out, uint32_t length) {
    QGET(getFloatTuple, QRY(q)->getFloatTuple(paramName, out, length))
}
VLR_API VLRResult vlrQueryableGetFloatArray(VLRQueryableConst q, const char *paramName, const float **out,
                                            uint32_t *length) {
    VLRA_TRY
    CHECK(q, obj(q)->isQueryable());
    if (!paramName || !out || !length) return VLRResult_InvalidArgument;
    if (!QRY(q)->getFloatArray(paramName, out, length)) return VLRResult_InvalidArgument;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrQueryableGetImage2D(VLRQueryableConst q, const char *paramName, VLRImage2DConst *out) {
    VLRA_TRY
    CHECK(q, obj(q)->isQueryable());
    if (!paramName || !out) return VLRResult_InvalidArgument;
    const Image2D *img = nullptr;
    if (!QRY(q)->getImage(paramName, &img)) return VLRResult_InvalidArgument;
    *out = reinterpret_cast<VLRImage2DConst>(img);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrQueryableGetImmediateSpectrum(VLRQueryableConst q, const char *paramName, VLRImmediateSpectrum *out) {
    VLRA_TRY
    CHECK(q, obj(q)->isQueryable());
    if (!paramName || !out) return VLRResult_InvalidArgument;
    ImmSpectrum s;
    if (!QRY(q)->getSpectrum(paramName, &s)) return VLRResult_InvalidArgument;
    out->colorSpace = colorSpaceToString(s.cs);
    out->e0 = s.e[0];
    out->e1 = s.e[1];
    out->e2 = s.e[2];
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrQueryableGetSurfaceMaterial(VLRQueryableConst q, const char *paramName, VLRSurfaceMaterialConst *out) {
    VLRA_TRY
    CHECK(q, obj(q)->isQueryable());
    if (!paramName || !out) return VLRResult_InvalidArgument;
    const SurfaceMaterial *m = nullptr;
    if (!QRY(q)->getMaterial(paramName, &m)) return VLRResult_InvalidArgument;
    *out = reinterpret_cast<VLRSurfaceMaterialConst>(m);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrQueryableGetShaderNodePlug(VLRQueryableConst q, const char *paramName, VLRShaderNodePlug *out) {
    VLRA_TRY
    CHECK(q, obj(q)->isQueryable());
    if (!paramName || !out) return VLRResult_InvalidArgument;
    PlugRef p;
    if (!QRY(q)->getPlug(paramName, &p)) return VLRResult_InvalidArgument;
    out->nodeRef = reinterpret_cast<uintptr_t>(p.node);
    out->info = (p.plugType & 0xFF) | (p.option << 8);
    return VLRResult_NoError;
    VLRA_CATCH
}

#define QSET(call)                                           \
    VLRA_TRY                                                 \
    CHECK(q, obj(q)->isQueryable());                         \
    if (!paramName) return VLRResult_InvalidArgument;        \
    if (!(call)) return VLRResult_InvalidArgument;           \
    return VLRResult_NoError;                                \
    VLRA_CATCH

VLR_API VLRResult vlrQueryableSetEnumValue(VLRQueryable q, const char *paramName, const char *value) {
    QSET(value && QRY(q)->setEnum(paramName, value))
}
VLR_API VLRResult vlrQueryableSetPoint3D(VLRQueryable q, const char *paramName, const VLRPoint3D *value) {
    QSET(value && QRY(q)->setPoint(paramName, &value->x))
}
VLR_API VLRResult vlrQueryableSetVector3D(VLRQueryable q, const char *paramName, const VLRVector3D *value) {
    QSET(value && QRY(q)->setVector(paramName, &value->x))
}
VLR_API VLRResult vlrQueryableSetNormal3D(VLRQueryable q, const char *paramName, const VLRNormal3D *value) {
    QSET(value && QRY(q)->setNormal(paramName, &value->x))
}
VLR_API VLRResult vlrQueryableSetQuaternion(VLRQueryable q, const char *paramName, const VLRQuaternion *value) {
    QSET(value && QRY(q)->setQuaternion(paramName, &value->x))
}
VLR_API VLRResult vlrQueryableSetFloat(VLRQueryable q, const char *paramName, float value) {
    QSET(QRY(q)->setFloatTuple(paramName, &value, 1))
}
VLR_API VLRResult vlrQueryableSetFloatTuple(VLRQueryable q, const char *paramName, const float *values, uint32_t length) {
    QSET(values && QRY(q)->setFloatTuple(paramName, values, length))
}
VLR_API VLRResult vlrQueryableSetImage2D(VLRQueryable q, const char *paramName, VLRImage2DConst image) {
    VLRA_TRY
    CHECK(q, obj(q)->isQueryable());
    if (!paramName) return VLRResult_InvalidArgument;
    if (image) CHECK(image, obj(image)->isImage());
    if (!QRY(q)->setImage(paramName, image ? static_cast<const Image2D *>(obj(image)) : nullptr))
        return VLRResult_InvalidArgument;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrQueryableSetImmediateSpectrum(VLRQueryable q, const char *paramName, const VLRImmediateSpectrum *value) {
    VLRA_TRY
    CHECK(q, obj(q)->isQueryable());
    if (!paramName || !value) return VLRResult_InvalidArgument;
    ImmSpectrum s;
    if (!colorSpaceFromString(value->colorSpace, &s.cs)) return VLRResult_InvalidArgument;
    s.e[0] = value->e0;
    s.e[1] = value->e1;
    s.e[2] = value->e2;
    if (!QRY(q)->setSpectrum(paramName, s)) return VLRResult_InvalidArgument;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrQueryableSetSurfaceMaterial(VLRQueryable q, const char *paramName, VLRSurfaceMaterialConst value) {
    VLRA_TRY
    CHECK(q, obj(q)->isQueryable());
    if (!paramName) return VLRResult_InvalidArgument;
    if (value) CHECK(value, IS(value, SurfaceMaterial));
    if (!QRY(q)->setMaterial(paramName, value ? static_cast<const SurfaceMaterial *>(obj(value)) : nullptr))
        return VLRResult_InvalidArgument;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrQueryableSetShaderNodePlug(VLRQueryable q, const char *paramName, VLRShaderNodePlug plug) {
    VLRA_TRY
    CHECK(q, obj(q)->isQueryable());
    if (!paramName) return VLRResult_InvalidArgument;
    PlugRef p;
    if (plug.nodeRef) {
        Object *n = reinterpret_cast<Object *>(plug.nodeRef);
        if (!alive(n) || n->type != ObjType::ShaderNode) return VLRResult_InvalidInstance;
        p.node = static_cast<const ShaderNode *>(n);
        p.plugType = plug.info & 0xFF;
        p.option = plug.info >> 8;
    }
    int rc = QRY(q)->setPlug(paramName, p);
    if (rc == 1) return VLRResult_InvalidArgument;
    if (rc == 2) return VLRResult_IncompatibleNodeType;
    return VLRResult_NoError;
    VLRA_CATCH
}

// ------------------------------------------------------------------------------------------ Image2D
#define IMG(h) static_cast<const Image2D *>(obj(h))
VLR_API VLRResult vlrImage2DGetWidth(VLRImage2DConst image, uint32_t *width) {
    VLRA_TRY
    CHECK(image, obj(image)->isImage());
    if (!width) return VLRResult_InvalidArgument;
    *width = IMG(image)->width;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrImage2DGetHeight(VLRImage2DConst image, uint32_t *height) {
    VLRA_TRY
    CHECK(image, obj(image)->isImage());
    if (!height) return VLRResult_InvalidArgument;
    *height = IMG(image)->height;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrImage2DGetStride(VLRImage2DConst image, uint32_t *stride) {
    VLRA_TRY
    CHECK(image, obj(image)->isImage());
    if (!stride) return VLRResult_InvalidArgument;
    *stride = IMG(image)->stride();
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrImage2DGetOriginalDataFormat(VLRImage2DConst image, const char **format) {
    VLRA_TRY
    CHECK(image, obj(image)->isImage());
    if (!format) return VLRResult_InvalidArgument;
    *format = IMG(image)->formatStr.c_str();
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrImage2DOriginalHasAlpha(VLRImage2DConst image, bool *hasAlpha) {
    VLRA_TRY
    CHECK(image, obj(image)->isImage());
    if (!hasAlpha) return VLRResult_InvalidArgument;
    *hasAlpha = IMG(image)->hasAlpha;
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrLinearImage2DCreate(VLRContext context, uint8_t *linearData, uint32_t width, uint32_t height,
                                         const char *format, const char *spectrumType, const char *colorSpace,
                                         VLRLinearImage2D *image) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!linearData || !format || !spectrumType || !colorSpace || !image || !width || !height)
        return VLRResult_InvalidArgument;
    auto *img = new LinearImage2D(CTX(context));
    img->width = width;
    img->height = height;
    img->formatStr = format;
    uint32_t st = img->stride();
    if (!st || !spectrumTypeFromString(spectrumType, &img->spectrumType) ||
        !colorSpaceFromString(colorSpace, &img->colorSpace)) {
        delete img;
        return VLRResult_InvalidArgument;
    }
    img->hasAlpha = img->formatStr == "RGBA8x4" || img->formatStr == "RGBA16Fx4" || img->formatStr == "RGBA32Fx4" ||
                    img->formatStr == "GrayA8x2";
    img->data.assign(linearData, linearData + size_t(st) * width * height);
    *image = reinterpret_cast<VLRLinearImage2D>(track(CTX(context), img));
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrLinearImage2DDestroy(VLRContext context, VLRLinearImage2D image) {
    VLRA_TRY
    CHECK(image, IS(image, LinearImage2D));
    untrack(obj(image));
    delete obj(image);
    (void)context;
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrBlockCompressedImage2DCreate(VLRContext context, uint8_t **data, size_t *sizes, uint32_t mipCount,
                                                  uint32_t width, uint32_t height, const char *dataFormat,
                                                  const char *spectrumType, const char *colorSpace,
                                                  VLRBlockCompressedImage2D *image) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!data || !sizes || !dataFormat || !spectrumType || !colorSpace || !image || mipCount == 0)
        return VLRResult_InvalidArgument;
    if (!bcFormatDecodable(dataFormat)) {
        g_lastError = "not a block-compressed format (BC1-BC7 and the signed variants are)";
        return VLRResult_InvalidArgument;
    }
    auto *img = new BlockCompressedImage2D(CTX(context));
    img->width = width;
    img->height = height;
    img->formatStr = dataFormat;
    if (!spectrumTypeFromString(spectrumType, &img->spectrumType) || !colorSpaceFromString(colorSpace, &img->colorSpace)) {
        delete img;
        return VLRResult_InvalidArgument;
    }
    for (uint32_t m = 0; m < mipCount; ++m) img->mips.emplace_back(data[m], data[m] + sizes[m]);
    *image = reinterpret_cast<VLRBlockCompressedImage2D>(track(CTX(context), img));
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrBlockCompressedImage2DDestroy(VLRContext context, VLRBlockCompressedImage2D image) {
    VLRA_TRY
    CHECK(image, IS(image, BlockCompressedImage2D));
    untrack(obj(image));
    delete obj(image);
    (void)context;
    return VLRResult_NoError;
    VLRA_CATCH
}

// ------------------------------------------------------------------------------------------ ShaderNode
VLR_API VLRResult vlrShaderNodeCreate(VLRContext context, const char *typeName, VLRShaderNode *node) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!typeName || !node) return VLRResult_InvalidArgument;
    static const std::pair<const char *, NodeKind> kinds[] = {
        {"Geometry", NodeKind::Geometry}, {"Tangent", NodeKind::Tangent}, {"Float2", NodeKind::Float2},
        {"Float3", NodeKind::Float3}, {"Float4", NodeKind::Float4}, {"ScaleAndOffsetFloat", NodeKind::ScaleAndOffsetFloat},
        {"TripletSpectrum", NodeKind::TripletSpectrum}, {"RegularSampledSpectrum", NodeKind::RegularSampledSpectrum},
        {"IrregularSampledSpectrum", NodeKind::IrregularSampledSpectrum},
        {"Float3ToSpectrum", NodeKind::Float3ToSpectrum},
        {"ScaleAndOffsetUVTextureMap2D", NodeKind::ScaleAndOffsetUVTextureMap2D},
        {"Image2DTexture", NodeKind::Image2DTexture}, {"EnvironmentTexture", NodeKind::EnvironmentTexture}};
    for (auto &k : kinds) {
        if (paramNameEquals(typeName, k.first)) {
            *node = reinterpret_cast<VLRShaderNode>(track(CTX(context), new ShaderNode(CTX(context), k.second)));
            return VLRResult_NoError;
        }
    }
    return VLRResult_InvalidArgument;
    VLRA_CATCH
}

VLR_API VLRResult vlrShaderNodeDestroy(VLRContext context, VLRShaderNode node) {
    VLRA_TRY
    CHECK(node, IS(node, ShaderNode));
    untrack(obj(node));
    delete obj(node);
    (void)context;
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrShaderNodeGetPlug(VLRShaderNodeConst node, VLRShaderNodePlugType plugType, uint32_t option,
                                       VLRShaderNodePlug *plug) {
    VLRA_TRY
    CHECK(node, IS(node, ShaderNode));
    if (!plug) return VLRResult_InvalidArgument;
    const ShaderNode *n = static_cast<const ShaderNode *>(obj(node));
    if (!n->providesPlug(static_cast<uint32_t>(plugType))) {
        plug->nodeRef = 0;
        plug->info = 0;
        return VLRResult_NoError;
    }
    plug->nodeRef = reinterpret_cast<uintptr_t>(n);
    plug->info = (static_cast<uint32_t>(plugType) & 0xFF) | (option << 8);
    return VLRResult_NoError;
    VLRA_CATCH
}

// ------------------------------------------------------------------------------------------ SurfaceMaterial
VLR_API VLRResult vlrSurfaceMaterialCreate(VLRContext context, const char *typeName, VLRSurfaceMaterial *material) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!typeName || !material) return VLRResult_InvalidArgument;
    // vlr.cpp:955-991
    static const std::pair<const char *, MatKind> kinds[] = {
        {"Matte", MatKind::Matte}, {"SpecularReflection", MatKind::SpecularReflection},
        {"SpecularScattering", MatKind::SpecularScattering}, {"MicrofacetReflection", MatKind::MicrofacetReflection},
        {"MicrofacetScattering", MatKind::MicrofacetScattering}, {"LambertianScattering", MatKind::LambertianScattering},
        {"UE4", MatKind::UE4}, {"OldStyle", MatKind::OldStyle}, {"DiffuseEmitter", MatKind::DiffuseEmitter},
        {"DirectionalEmitter", MatKind::DirectionalEmitter}, {"PointEmitter", MatKind::PointEmitter},
        {"Multi", MatKind::Multi}, {"EnvironmentEmitter", MatKind::EnvironmentEmitter}};
    for (auto &k : kinds) {
        if (paramNameEquals(typeName, k.first)) {
            *material = reinterpret_cast<VLRSurfaceMaterial>(
                track(CTX(context), new SurfaceMaterial(CTX(context), k.second)));
            return VLRResult_NoError;
        }
    }
    return VLRResult_InvalidArgument;
    VLRA_CATCH
}

VLR_API VLRResult vlrSurfaceMaterialDestroy(VLRContext context, VLRSurfaceMaterial material) {
    VLRA_TRY
    CHECK(material, IS(material, SurfaceMaterial));
    untrack(obj(material));
    delete obj(material);
    (void)context;
    return VLRResult_NoError;
    VLRA_CATCH
}

// ------------------------------------------------------------------------------------------ Transform
VLR_API VLRResult vlrStaticTransformCreate(VLRContext context, const float mat[16], VLRStaticTransform *transform) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!mat || !transform) return VLRResult_InvalidArgument;
    *transform = reinterpret_cast<VLRStaticTransform>(track(CTX(context), new StaticTransform(CTX(context), mat)));
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrStaticTransformDestroy(VLRContext context, VLRStaticTransform transform) {
    VLRA_TRY
    CHECK(transform, IS(transform, StaticTransform));
    untrack(obj(transform));
    delete obj(transform);
    (void)context;
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrStaticTransformGetArrays(VLRStaticTransformConst transform, float mat[16], float invMat[16]) {
    VLRA_TRY
    CHECK(transform, IS(transform, StaticTransform));
    if (!mat || !invMat) return VLRResult_InvalidArgument;
    const StaticTransform *t = static_cast<const StaticTransform *>(obj(transform));
    std::memcpy(mat, t->mat, 64);
    std::memcpy(invMat, t->invMat, 64);
    return VLRResult_NoError;
    VLRA_CATCH
}

// ------------------------------------------------------------------------------------------ Nodes
#define NODE_(h) static_cast<Node *>(obj(h))
VLR_API VLRResult vlrNodeSetName(VLRNode node, const char *name) {
    VLRA_TRY
    CHECK(node, obj(node)->isNode());
    if (!name) return VLRResult_InvalidArgument;
    NODE_(node)->name = name;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrNodeGetName(VLRNodeConst node, const char **name) {
    VLRA_TRY
    CHECK(node, obj(node)->isNode());
    if (!name) return VLRResult_InvalidArgument;
    *name = NODE_(node)->name.c_str();
    return VLRResult_NoError;
    VLRA_CATCH
}

static void detachNode(Node *n) {
    for (ParentNode *p : std::set<ParentNode *>(n->parents)) {
        auto &ch = p->children;
        ch.erase(std::remove(ch.begin(), ch.end(), n), ch.end());
    }
    if (n->isParentNode()) {
        for (Node *c : static_cast<ParentNode *>(n)->children) c->parents.erase(static_cast<ParentNode *>(n));
    }
    n->ctx->touch();
}

VLR_API VLRResult vlrTriangleMeshSurfaceNodeCreate(VLRContext context, const char *name,
                                                   VLRTriangleMeshSurfaceNode *surfaceNode) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!name || !surfaceNode) return VLRResult_InvalidArgument;
    auto *n = new TriangleMeshSurfaceNode(CTX(context));
    n->name = name;
    *surfaceNode = reinterpret_cast<VLRTriangleMeshSurfaceNode>(track(CTX(context), n));
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrTriangleMeshSurfaceNodeDestroy(VLRContext context, VLRTriangleMeshSurfaceNode surfaceNode) {
    VLRA_TRY
    CHECK(surfaceNode, IS(surfaceNode, TriangleMeshSurfaceNode));
    detachNode(NODE_(surfaceNode));
    untrack(obj(surfaceNode));
    delete obj(surfaceNode);
    (void)context;
    return VLRResult_NoError;
    VLRA_CATCH
}

static VLRResult setVertices(SurfaceNode *n, const VLRVertex *vertices, uint32_t numVertices) {
    if (!vertices && numVertices) return VLRResult_InvalidArgument;
    n->vertices.assign(vertices, vertices + numVertices);
    n->version = nextGeometryVersion();
    n->ctx->touch();
    return VLRResult_NoError;
}

VLR_API VLRResult vlrTriangleMeshSurfaceNodeSetVertices(VLRTriangleMeshSurfaceNode surfaceNode, const VLRVertex *vertices,
                                                       uint32_t numVertices) {
    VLRA_TRY
    CHECK(surfaceNode, IS(surfaceNode, TriangleMeshSurfaceNode));
    return setVertices(static_cast<SurfaceNode *>(obj(surfaceNode)), vertices, numVertices);
    VLRA_CATCH
}

static PlugRef toPlug(VLRShaderNodePlug p) {
    PlugRef r;
    if (p.nodeRef) {
        Object *n = reinterpret_cast<Object *>(p.nodeRef);
        if (alive(n) && n->type == ObjType::ShaderNode) {
            r.node = static_cast<const ShaderNode *>(n);
            r.plugType = p.info & 0xFF;
            r.option = p.info >> 8;
        }
    }
    return r;
}

VLR_API VLRResult vlrTriangleMeshSurfaceNodeAddMaterialGroup(VLRTriangleMeshSurfaceNode surfaceNode,
                                                             const uint32_t *indices, uint32_t numIndices,
                                                             VLRSurfaceMaterialConst material,
                                                             VLRShaderNodePlug nodeNormal, VLRShaderNodePlug nodeTangent,
                                                             VLRShaderNodePlug nodeAlpha) {
    VLRA_TRY
    CHECK(surfaceNode, IS(surfaceNode, TriangleMeshSurfaceNode));
    CHECK(material, IS(material, SurfaceMaterial));
    if (!indices || numIndices % 3 != 0) return VLRResult_InvalidArgument;
    auto *n = static_cast<SurfaceNode *>(obj(surfaceNode));
    MaterialGroup g;
    g.indices.assign(indices, indices + numIndices);
    g.material = static_cast<const SurfaceMaterial *>(obj(material));
    // scene.cpp:225-246: a normal or tangent plug of an inconvertible type is ignored
    g.nodeNormal = toPlug(nodeNormal);
    if (g.nodeNormal.valid() && !plugConvertible(SocketType::Normal3D, g.nodeNormal.plugType)) g.nodeNormal = PlugRef();
    g.nodeTangent = toPlug(nodeTangent);
    if (g.nodeTangent.valid() && !plugConvertible(SocketType::Vector3D, g.nodeTangent.plugType)) g.nodeTangent = PlugRef();
    g.nodeAlpha = toPlug(nodeAlpha);
    if (g.nodeAlpha.valid() && !plugConvertible(SocketType::Float, g.nodeAlpha.plugType)) g.nodeAlpha = PlugRef();
    n->groups.push_back(std::move(g));
    n->version = nextGeometryVersion();
    n->ctx->touch();
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrPointSurfaceNodeCreate(VLRContext context, const char *name, VLRPointSurfaceNode *surfaceNode) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!name || !surfaceNode) return VLRResult_InvalidArgument;
    auto *n = new PointSurfaceNode(CTX(context));
    n->name = name;
    *surfaceNode = reinterpret_cast<VLRPointSurfaceNode>(track(CTX(context), n));
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrPointSurfaceNodeDestroy(VLRContext context, VLRPointSurfaceNode surfaceNode) {
    VLRA_TRY
    CHECK(surfaceNode, IS(surfaceNode, PointSurfaceNode));
    detachNode(NODE_(surfaceNode));
    untrack(obj(surfaceNode));
    delete obj(surfaceNode);
    (void)context;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrPointSurfaceNodeSetVertices(VLRPointSurfaceNode surfaceNode, const VLRVertex *vertices,
                                                 uint32_t numVertices) {
    VLRA_TRY
    CHECK(surfaceNode, IS(surfaceNode, PointSurfaceNode));
    return setVertices(static_cast<SurfaceNode *>(obj(surfaceNode)), vertices, numVertices);
    VLRA_CATCH
}
VLR_API VLRResult vlrPointSurfaceNodeAddMaterialGroup(VLRPointSurfaceNode surfaceNode, const uint32_t *indices,
                                                      uint32_t numIndices, VLRSurfaceMaterialConst material) {
    VLRA_TRY
    CHECK(surfaceNode, IS(surfaceNode, PointSurfaceNode));
    CHECK(material, IS(material, SurfaceMaterial));
    if (!indices) return VLRResult_InvalidArgument;
    auto *n = static_cast<SurfaceNode *>(obj(surfaceNode));
    MaterialGroup g;
    g.indices.assign(indices, indices + numIndices);
    g.material = static_cast<const SurfaceMaterial *>(obj(material));
    n->groups.push_back(std::move(g));
    n->version = nextGeometryVersion();
    n->ctx->touch();
    return VLRResult_NoError;
    VLRA_CATCH
}

#define PARENT(h) static_cast<ParentNode *>(obj(h))
VLR_API VLRResult vlrParentNodeSetTransform(VLRParentNode node, VLRTransformConst localToWorld) {
    VLRA_TRY
    CHECK(node, obj(node)->isParentNode());
    CHECK(localToWorld, IS(localToWorld, StaticTransform));
    PARENT(node)->transform = static_cast<const StaticTransform *>(obj(localToWorld));
    PARENT(node)->ctx->touch();
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrParentNodeGetTransform(VLRParentNodeConst node, VLRTransformConst *localToWorld) {
    VLRA_TRY
    CHECK(node, obj(node)->isParentNode());
    if (!localToWorld) return VLRResult_InvalidArgument;
    *localToWorld = reinterpret_cast<VLRTransformConst>(PARENT(node)->transform);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrParentNodeAddChild(VLRParentNode node, VLRNode child) {
    VLRA_TRY
    CHECK(node, obj(node)->isParentNode());
    CHECK(child, obj(child)->isNode() && !IS(child, Scene));
    ParentNode *p = PARENT(node);
    Node *c = NODE_(child);
    if (std::find(p->children.begin(), p->children.end(), c) != p->children.end()) return VLRResult_NoError;
    p->children.push_back(c);
    c->parents.insert(p);
    p->ctx->touch();
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrParentNodeRemoveChild(VLRParentNode node, VLRNode child) {
    VLRA_TRY
    CHECK(node, obj(node)->isParentNode());
    CHECK(child, obj(child)->isNode());
    ParentNode *p = PARENT(node);
    Node *c = NODE_(child);
    p->children.erase(std::remove(p->children.begin(), p->children.end(), c), p->children.end());
    c->parents.erase(p);
    p->ctx->touch();
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrParentNodeGetNumChildren(VLRParentNodeConst node, uint32_t *numChildren) {
    VLRA_TRY
    CHECK(node, obj(node)->isParentNode());
    if (!numChildren) return VLRResult_InvalidArgument;
    *numChildren = static_cast<uint32_t>(PARENT(node)->children.size());
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrParentNodeGetChildren(VLRParentNodeConst node, VLRNode *children) {
    VLRA_TRY
    CHECK(node, obj(node)->isParentNode());
    if (!children) return VLRResult_InvalidArgument;
    const auto &ch = PARENT(node)->children;
    for (size_t i = 0; i < ch.size(); ++i) children[i] = reinterpret_cast<VLRNode>(ch[i]);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrParentNodeGetChildAt(VLRParentNodeConst node, uint32_t index, VLRNode *child) {
    VLRA_TRY
    CHECK(node, obj(node)->isParentNode());
    if (!child) return VLRResult_InvalidArgument;
    const auto &ch = PARENT(node)->children;
    if (index >= ch.size()) return VLRResult_InvalidArgument;
    *child = reinterpret_cast<VLRNode>(ch[index]);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrInternalNodeCreate(VLRContext context, const char *name, VLRTransformConst transform,
                                        VLRInternalNode *node) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CHECK(transform, IS(transform, StaticTransform));
    if (!name || !node) return VLRResult_InvalidArgument;
    auto *n = new InternalNode(CTX(context));
    n->name = name;
    n->transform = static_cast<const StaticTransform *>(obj(transform));
    *node = reinterpret_cast<VLRInternalNode>(track(CTX(context), n));
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrInternalNodeDestroy(VLRContext context, VLRInternalNode node) {
    VLRA_TRY
    CHECK(node, IS(node, InternalNode));
    detachNode(NODE_(node));
    untrack(obj(node));
    delete obj(node);
    (void)context;
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrSceneCreate(VLRContext context, VLRTransformConst transform, VLRScene *scene) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CHECK(transform, IS(transform, StaticTransform));
    if (!scene) return VLRResult_InvalidArgument;
    auto *s = new Scene(CTX(context));
    s->name = "Scene";
    s->transform = static_cast<const StaticTransform *>(obj(transform));
    *scene = reinterpret_cast<VLRScene>(track(CTX(context), s));
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrSceneDestroy(VLRContext context, VLRScene scene) {
    VLRA_TRY
    CHECK(scene, IS(scene, Scene));
    Context *c = obj(scene)->ctx;
    detachNode(NODE_(scene));
    if (c->scene == static_cast<Scene *>(obj(scene))) c->scene = nullptr;
    if (c->gpu.scene == static_cast<Scene *>(obj(scene))) c->gpu.valid = false;
    untrack(obj(scene));
    delete obj(scene);
    (void)context;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrSceneSetEnvironment(VLRScene scene, VLRSurfaceMaterial material) {
    VLRA_TRY
    CHECK(scene, IS(scene, Scene));
    CHECK(material, IS(material, SurfaceMaterial));
    auto *m = static_cast<const SurfaceMaterial *>(obj(material));
    if (m->kind != MatKind::EnvironmentEmitter) return VLRResult_InvalidArgument;
    static_cast<Scene *>(obj(scene))->envMaterial = m;
    obj(scene)->ctx->touch();
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrSceneSetEnvironmentRotation(VLRScene scene, float rotationPhi) {
    VLRA_TRY
    CHECK(scene, IS(scene, Scene));
    static_cast<Scene *>(obj(scene))->envRotation = rotationPhi;
    obj(scene)->ctx->touch();
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrCameraCreate(VLRContext context, const char *typeName, VLRCamera *camera) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!typeName || !camera) return VLRResult_InvalidArgument;
    bool persp;
    if (paramNameEquals(typeName, "Perspective")) persp = true;
    else if (paramNameEquals(typeName, "Equirectangular")) persp = false;
    else return VLRResult_InvalidArgument;
    *camera = reinterpret_cast<VLRCamera>(track(CTX(context), new Camera(CTX(context), persp)));
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrCameraDestroy(VLRContext context, VLRCamera camera) {
    VLRA_TRY
    CHECK(camera, IS(camera, Camera));
    untrack(obj(camera));
    delete obj(camera);
    (void)context;
    return VLRResult_NoError;
    VLRA_CATCH
}

// ------------------------------------------------------------------------------------------ extensions
VLR_API VLRResult vlrContextSetMaxPathLength(VLRContext context, uint32_t maxPathLength) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (maxPathLength == 0 || maxPathLength >= kMaxDepthCounters) return VLRResult_InvalidArgument;
    CTX(context)->maxPathLength = maxPathLength;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetMaxPathLength(VLRContextConst context, uint32_t *maxPathLength) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!maxPathLength) return VLRResult_InvalidArgument;
    *maxPathLength = CTX(context)->maxPathLength;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextSetSpectralRendering(VLRContext context, bool enable) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->spectralRequested = enable;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetSpectralRendering(VLRContextConst context, bool *enabled) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!enabled) return VLRResult_InvalidArgument;
    *enabled = CTX(context)->spectralRequested;
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextGetRenderStats(VLRContextConst context, VLRRenderStats *stats) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!stats) return VLRResult_InvalidArgument;
    Context *c = CTX(context);
    std::memset(stats, 0, sizeof(*stats));
    uint64_t trav[6], wavefront[2];
    c->statistics(&stats->closestRays, &stats->shadowRays, &stats->droppedNonFinite, trav, wavefront);
    stats->closestRaysWavefront = wavefront[0];
    stats->shadowRaysWavefront = wavefront[1];
    stats->nodesClosest = trav[0];
    stats->trianglesClosest = trav[1];
    stats->instancesClosest = trav[2];
    stats->nodesShadow = trav[3];
    stats->trianglesShadow = trav[4];
    stats->instancesShadow = trav[5];
    stats->numFrames = c->framesRendered;
    stats->paths = c->pathsStarted;
    c->collectTimers();
    stats->msGenerate = c->stageMs[STAGE_GENERATE];
    stats->msIntersect = c->stageMs[STAGE_INTERSECT];
    stats->msShade = c->stageMs[STAGE_SHADE];
    stats->msShadow = c->stageMs[STAGE_SHADOW];
    stats->msAccumulate = c->stageMs[STAGE_ACCUMULATE];
    stats->msConvert = c->stageMs[STAGE_CONVERT];
    stats->launchesIntersect = c->launches[STAGE_INTERSECT];
    stats->launchesShadow = c->launches[STAGE_SHADOW];
    stats->launchesShade = c->launches[STAGE_SHADE];
    stats->msTail = c->stageMs[STAGE_TAIL];
    stats->launchesTail = c->launches[STAGE_TAIL];
    stats->msLightPaths = c->stageMs[STAGE_LIGHT_PATHS];
    stats->msEyePaths = c->stageMs[STAGE_EYE_PATHS];
    stats->msCacheQuery = c->stageMs[STAGE_CACHE_QUERY];
    stats->msCacheTrain = c->stageMs[STAGE_CACHE_TRAIN];
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextResetRenderStats(VLRContext context) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->resetStats();
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextSetKernelTiming(VLRContext context, bool enable) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->collectTimers();
    CTX(context)->kernelTiming = enable;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextSetDenoiserAuxBuffers(VLRContext context, bool enable) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    Context *c = CTX(context);
    if (enable && !c->guidesEnabled) c->guideFrames = 0;
    c->guidesEnabled = enable;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextReadDenoiserAuxBuffers(VLRContextConst context, float *albedoRGBA, float *normalRGBA) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    Context *c = CTX(context);
    if (!c->width) return VLRResult_InvalidArgument;
    c->readGuides(albedoRGBA, normalRGBA);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextSetWavefrontDepth(VLRContext context, uint32_t depth) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->wavefrontDepth = depth;
    CTX(context)->autotune = false;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextSetShadeSorting(VLRContext context, bool enable) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->shadeSort = enable;
    CTX(context)->autotune = false;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetShadeSorting(VLRContextConst context, bool *enabled) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!enabled) return VLRResult_InvalidArgument;
    *enabled = CTX(context)->lastFrame.sortShade != 0;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetWavefrontDepth(VLRContextConst context, uint32_t *depth) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!depth) return VLRResult_InvalidArgument;
    *depth = CTX(context)->wavefrontDepth;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextSetScheduleAutotuning(VLRContext context, bool enable) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    Context *c = CTX(context);
    c->autotune = enable;
    if (enable) c->tuner = ScheduleTuner();
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetSchedule(VLRContextConst context, uint32_t *wavefrontDepth, uint32_t *tailBlocks,
                                        bool *tuned) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!wavefrontDepth || !tailBlocks || !tuned) return VLRResult_InvalidArgument;
    const Context *c = CTX(context);
    *wavefrontDepth = c->lastFrame.wavefrontDepth;
    *tailBlocks = c->lastFrame.tailBlocks;
    *tuned = c->autotune && c->tuner.done();
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetQueueCounts(VLRContextConst context, uint32_t *closest, uint32_t *shadow,
                                           uint32_t maxDepth) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!closest || !shadow) return VLRResult_InvalidArgument;
    CTX(context)->queueCounts(closest, shadow, maxDepth);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextSetTraversalStatistics(VLRContext context, bool enable) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->traversalStats = enable;
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextReadRNGStates(VLRContextConst context, uint64_t *states) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    Context *c = CTX(context);
    if (!states || !c->width) return VLRResult_InvalidArgument;
    VLRA_HIP(hipDeviceSynchronize());
    VLRA_HIP(hipMemcpy(states, c->rng.ptr, size_t(c->width) * c->height * 8, hipMemcpyDeviceToHost));
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextWriteRNGStates(VLRContext context, const uint64_t *states) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    Context *c = CTX(context);
    if (!states || !c->width) return VLRResult_InvalidArgument;
    VLRA_HIP(hipDeviceSynchronize());
    VLRA_HIP(hipMemcpy(c->rng.ptr, states, size_t(c->width) * c->height * 8, hipMemcpyHostToDevice));
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextReadLightRNGStates(VLRContextConst context, uint64_t *states) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!states) return VLRResult_InvalidArgument;
    CTX(context)->readLightRngStates(states);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextWriteLightRNGStates(VLRContext context, const uint64_t *states) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!states) return VLRResult_InvalidArgument;
    CTX(context)->writeLightRngStates(states);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextSetNumLightPaths(VLRContext context, uint32_t numLightPaths) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->setNumLightPaths(numLightPaths);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetNumLightPaths(VLRContextConst context, uint32_t *numLightPaths) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!numLightPaths) return VLRResult_InvalidArgument;
    *numLightPaths = CTX(context)->numLightPaths;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextSetSplitFrames(VLRContext context, bool enable) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->splitFrames = enable;
    return VLRResult_NoError;
    VLRA_CATCH
}

// ---- batches ----
VLR_API VLRResult vlrContextRenderBatch(VLRContext context, hipStream_t stream, const VLRBatchJob *jobs,
                                        uint32_t numJobs, float *completionMs) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (numJobs && !jobs) return VLRResult_InvalidArgument;
    std::vector<BatchJobDesc> descs(numJobs);
    for (uint32_t k = 0; k < numJobs; ++k) {
        CHECK(jobs[k].scene, IS(jobs[k].scene, Scene));
        CHECK(jobs[k].camera, IS(jobs[k].camera, Camera));
        descs[k].scene = static_cast<const Scene *>(obj(jobs[k].scene));
        descs[k].camera = static_cast<const Camera *>(obj(jobs[k].camera));
        descs[k].width = jobs[k].width;
        descs[k].height = jobs[k].height;
        descs[k].spp = jobs[k].numSamples;
        descs[k].maxPathLength = jobs[k].maxPathLength;
    }
    CTX(context)->renderBatch(stream, descs, completionMs);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextReadBatchOutput(VLRContextConst context, uint32_t job, float *rgba) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!rgba) return VLRResult_InvalidArgument;
    CTX(context)->readBatchOutput(job, rgba);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextSetBatchGraphs(VLRContext context, bool enable) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->batchGraphs = enable;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetBatchSchedule(VLRContextConst context, uint32_t *wavefrontDepth, uint32_t *tailBlocks) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (wavefrontDepth) *wavefrontDepth = CTX(context)->batch.wavefrontDepth;
    if (tailBlocks) *tailBlocks = CTX(context)->batch.tailBlocks;
    return VLRResult_NoError;
    VLRA_CATCH
}

// ---- radiance cache ----
VLR_API VLRResult vlrContextSetRadianceCache(VLRContext context, bool enable) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->cacheEnabled = enable;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetRadianceCache(VLRContextConst context, bool *enabled) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!enabled) return VLRResult_InvalidArgument;
    *enabled = CTX(context)->cacheEnabled;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextSetRadianceCacheSettings(VLRContext context, const VLRRadianceCacheSettings *st) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!st) return VLRResult_InvalidArgument;
    if (st->trainingPaths < 1 || st->trainingPaths > rc::kMaxBatch / kCacheTrainVerts) return VLRResult_InvalidArgument;
    if (!(st->learningRate > 0.0f) || !(st->beta1 >= 0.0f && st->beta1 < 1.0f) || !(st->beta2 >= 0.0f && st->beta2 < 1.0f) ||
        !(st->epsilon > 0.0f) || !(st->terminationThreshold >= 0.0f))
        return VLRResult_InvalidArgument;
    Context::CacheSettings &c = CTX(context)->cacheSettings;
    c.learningRate = st->learningRate;
    c.beta1 = st->beta1;
    c.beta2 = st->beta2;
    c.epsilon = st->epsilon;
    c.trainingPaths = st->trainingPaths;
    c.threshold = st->terminationThreshold;
    c.warmupSteps = st->warmupSteps;
    c.seed = st->seed;
    c.fullRadiance = st->fullRadiance != 0;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetRadianceCacheSettings(VLRContextConst context, VLRRadianceCacheSettings *st) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!st) return VLRResult_InvalidArgument;
    const Context::CacheSettings &c = CTX(context)->cacheSettings;
    st->learningRate = c.learningRate;
    st->beta1 = c.beta1;
    st->beta2 = c.beta2;
    st->epsilon = c.epsilon;
    st->trainingPaths = c.trainingPaths;
    st->terminationThreshold = c.threshold;
    st->warmupSteps = c.warmupSteps;
    st->seed = c.seed;
    st->fullRadiance = c.fullRadiance ? 1u : 0u;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextResetRadianceCache(VLRContext context) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->resetCache();
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrRadianceCacheNumParameters(uint32_t *numParameters) {
    if (!numParameters) return VLRResult_InvalidArgument;
    *numParameters = rc::kNumParams;
    return VLRResult_NoError;
}
VLR_API VLRResult vlrContextReadRadianceCacheWeights(VLRContextConst context, float *parameters) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!parameters) return VLRResult_InvalidArgument;
    CTX(context)->readCacheWeights(parameters);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextWriteRadianceCacheWeights(VLRContext context, const float *parameters) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!parameters) return VLRResult_InvalidArgument;
    Context *c = CTX(context);
    c->ensureCacheNet();
    VLRA_HIP(hipDeviceSynchronize());
    c->writeCacheWeights(parameters);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextReadRadianceCacheGradient(VLRContextConst context, float *gradient) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!gradient) return VLRResult_InvalidArgument;
    Context *c = CTX(context);
    c->ensureCacheNet();
    VLRA_HIP(hipDeviceSynchronize());
    VLRA_HIP(hipMemcpy(gradient, c->rcGrad.ptr, rc::kNumParams * sizeof(float), hipMemcpyDeviceToHost));
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetRadianceCacheLoss(VLRContextConst context, float *loss, uint32_t *steps) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    Context *c = CTX(context);
    if (loss) {
        *loss = 0.0f;
        if (c->cacheInitialized) {
            VLRA_HIP(hipDeviceSynchronize());
            VLRA_HIP(hipMemcpy(loss, c->rcLoss.ptr, sizeof(float), hipMemcpyDeviceToHost));
        }
    }
    if (steps) *steps = c->cacheStep;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextRadianceCacheInfer(VLRContext context, hipStream_t stream, const float *records,
                                               uint32_t numRecords, float *outputs) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if ((!records || !outputs) && numRecords) return VLRResult_InvalidArgument;
    CTX(context)->cacheInfer(stream, records, numRecords, outputs);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextRadianceCacheTrain(VLRContext context, hipStream_t stream, const float *records,
                                               const float *targets, uint32_t numRecords, uint32_t numOutputs) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!records || !targets) return VLRResult_InvalidArgument;
    CTX(context)->cacheTrain(stream, records, targets, numRecords, numOutputs);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextGetAccumulationChannels(VLRContextConst context, uint32_t *channels) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!channels) return VLRResult_InvalidArgument;
    *channels = CTX(context)->spectralRequested ? 16u : 3u;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextReadAccumulation(VLRContextConst context, float *data, size_t numFloats) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    Context *c = CTX(context);
    if (!data || !c->width) return VLRResult_InvalidArgument;
    // the accumulation in place belongs to the other mode until a first frame switches it
    if (c->spectralActive != c->spectralRequested) return VLRResult_InvalidArgument;
    if (numFloats != size_t(c->width) * c->height * c->accumChannels * 2) return VLRResult_InvalidArgument;
    c->readAccumulation(data);
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextWriteAccumulation(VLRContext context, const float *data, size_t numFloats,
                                              uint32_t numAccumFrames) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    Context *c = CTX(context);
    if (!data || !c->width) return VLRResult_InvalidArgument;
    const uint32_t channels = c->spectralRequested ? 16u : 3u;
    if (numFloats != size_t(c->width) * c->height * channels * 2) return VLRResult_InvalidArgument;
    // a resumed render continues in the requested transport (RGB or spectral channels)
    if (c->spectralActive != c->spectralRequested) {
        c->spectralActive = c->spectralRequested;
        c->resizeAccum();
    }
    c->writeAccumulation(data);
    c->numAccumFrames = numAccumFrames;
    c->resumePending = true;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetBptWavelengthFrames(VLRContextConst context, uint64_t *frames) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!frames) return VLRResult_InvalidArgument;
    *frames = CTX(context)->bptWlFrames;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextSetBptWavelengthFrames(VLRContext context, uint64_t frames) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->setBptWavelengthFrames(frames);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextTraceRays(VLRContext context, hipStream_t stream, VLRScene scene, const VLRRay *rays,
                                      uint32_t numRays, VLRHit *hits, bool anyHit) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CHECK(scene, IS(scene, Scene));
    if ((!rays || !hits) && numRays) return VLRResult_InvalidArgument;
    CTX(context)->traceRays(stream, *static_cast<const Scene *>(obj(scene)), rays, numRays, hits, anyHit);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextSetAccelBuilder(VLRContext context, uint32_t builder) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->setAccelBuilder(builder);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextGetAccelBuilder(VLRContextConst context, uint32_t *builder) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!builder) return VLRResult_InvalidArgument;
    *builder = CTX(context)->accelBuilder;
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextSetTlasRefit(VLRContext context, bool enable) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CTX(context)->tlasRefit = enable;
    return VLRResult_NoError;
    VLRA_CATCH
}
VLR_API VLRResult vlrContextGetAccelStats(VLRContextConst context, VLRAccelStats *stats) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if (!stats) return VLRResult_InvalidArgument;
    const AccelStats &a = CTX(context)->accelStats;
    stats->numBlas = a.numBlas;
    stats->numTlasInstances = a.numTlasInstances;
    stats->numTriangles = a.numTriangles;
    stats->numBlasNodes = a.numBlasNodes;
    stats->numTlasNodes = a.numTlasNodes;
    stats->msBlas = a.msBlas;
    stats->msTlas = a.msTlas;
    stats->blasReused = a.blasReused;
    stats->numNodeProgramSteps = a.numNodeProgramSteps;
    stats->tlasDepth = a.tlasDepth;
    stats->blasDepth = a.blasDepth;
    stats->tlasRefit = a.tlasRefit;
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextReadAccel(VLRContext context, VLRScene scene, uint32_t which, void *dst, size_t capacity,
                                      size_t *bytes) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CHECK(scene, IS(scene, Scene));
    if (!bytes) return VLRResult_InvalidArgument;
    CTX(context)->readAccel(*static_cast<const Scene *>(obj(scene)), which, dst, capacity, bytes);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrBvh4NodeCapacity(uint32_t numPrims, uint32_t *capacity) {
    if (!capacity) return VLRResult_InvalidArgument;
    *capacity = GpuBvhBuilder::nodeCapacity(numPrims);
    return VLRResult_NoError;
}

VLR_API VLRResult vlrBuildBvh4Host(const float *boxes, uint32_t numPrims, uint32_t maxLeafPrims,
                                   const uint32_t *leafRemap, void *nodes, uint32_t nodeCapacity, uint32_t *order,
                                   uint32_t *numNodes) {
    VLRA_TRY
    if ((!boxes || !nodes || !order) && numPrims) return VLRResult_InvalidArgument;
    if (!numNodes || nodeCapacity < GpuBvhBuilder::nodeCapacity(numPrims)) return VLRResult_InvalidArgument;
    if (maxLeafPrims < 1 || maxLeafPrims > kMaxLeafPrims || numPrims >= kBvhMaxPrims) return VLRResult_InvalidArgument;
    if (leafRemap && maxLeafPrims != 1) return VLRResult_InvalidArgument;
    HostBvh4 hb;
    hostBuildBvh4(boxes, numPrims, maxLeafPrims, leafRemap, &hb);
    if (!hb.nodes.empty()) std::memcpy(nodes, hb.nodes.data(), hb.nodes.size() * sizeof(BVH4QNode));
    if (!hb.order.empty()) std::memcpy(order, hb.order.data(), hb.order.size() * sizeof(uint32_t));
    *numNodes = static_cast<uint32_t>(hb.nodes.size());
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextBuildBvh4(VLRContext context, hipStream_t stream, const float *boxes, uint32_t numPrims,
                                      uint32_t maxLeafPrims, const uint32_t *leafRemap, void *nodes,
                                      uint32_t nodeCapacity, uint32_t *order, uint32_t *numNodes) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if ((!boxes || !nodes || !order) && numPrims) return VLRResult_InvalidArgument;
    if (!numNodes || nodeCapacity < GpuBvhBuilder::nodeCapacity(numPrims)) return VLRResult_InvalidArgument;
    if (maxLeafPrims < 1 || maxLeafPrims > kMaxLeafPrims || numPrims >= kBvhMaxPrims) return VLRResult_InvalidArgument;
    if (leafRemap && maxLeafPrims != 1) return VLRResult_InvalidArgument;
    Context *ctx = CTX(context);
    ctx->ensureDevice();
    GpuBvhBuilder &B = ctx->bvhBuilder();
    *numNodes = B.build(boxes, numPrims, maxLeafPrims, leafRemap, static_cast<BVH4QNode *>(nodes), stream);
    if (numPrims) {
        VLRA_HIP(hipMemcpyAsync(order, B.order(), size_t(numPrims) * 4, hipMemcpyDeviceToDevice, stream));
        VLRA_HIP(hipStreamSynchronize(stream));
    }
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrBuildBvh4NoAlloc(hipStream_t stream, const float *boxes, uint32_t numPrims,
                                      uint32_t maxLeafPrims, const uint32_t *leafRemap, void *nodes,
                                      uint32_t nodeCapacity, uint32_t *order, uint32_t *numNodes) {
    static_assert(VLR_BVH4_NO_ALLOC_MAX_PRIMS == kNoAllocMaxPrims, "header and library agree on the limit");
    VLRA_TRY
    if ((!boxes || !nodes || !order) && numPrims) return VLRResult_InvalidArgument;
    if (!numNodes || nodeCapacity < GpuBvhBuilder::nodeCapacity(numPrims)) return VLRResult_InvalidArgument;
    if (maxLeafPrims < 1 || maxLeafPrims > kMaxLeafPrims) return VLRResult_InvalidArgument;
    if (leafRemap && maxLeafPrims != 1) return VLRResult_InvalidArgument;
    if (numPrims > kNoAllocMaxPrims) {
        g_lastError = "vlrBuildBvh4NoAlloc: numPrims " + std::to_string(numPrims) + " exceeds the maximum " +
                      std::to_string(kNoAllocMaxPrims) + " (VLR_BVH4_NO_ALLOC_MAX_PRIMS)";
        return VLRResult_InvalidArgument;
    }
    *numNodes = buildBvh4NoAlloc(boxes, numPrims, maxLeafPrims, leafRemap, static_cast<BVH4QNode *>(nodes), order,
                                 stream);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextEvaluateShaderNodePlug(VLRContext context, hipStream_t stream, VLRShaderNodePlug plug,
                                                   bool spectral, const float *points, uint32_t numPoints,
                                                   float *values) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    if ((!points || !values) && numPoints) return VLRResult_InvalidArgument;
    const PlugRef r = toPlug(plug);
    if (!r.valid()) return VLRResult_InvalidArgument;
    CTX(context)->evaluateNodePlug(stream, r, spectral, points, numPoints, values);
    return VLRResult_NoError;
    VLRA_CATCH
}

VLR_API VLRResult vlrContextEvaluateBSDFs(VLRContext context, hipStream_t stream, VLRSurfaceMaterialConst material,
                                          bool spectral, const VLRBSDFQuery *queries, uint32_t numQueries,
                                          VLRBSDFResult *results) {
    VLRA_TRY
    CHECK(context, IS(context, Context));
    CHECK(material, IS(material, SurfaceMaterial));
    if ((!queries || !results) && numQueries) return VLRResult_InvalidArgument;
    CTX(context)->evaluateBSDFs(stream, *static_cast<const SurfaceMaterial *>(obj(material)), spectral, queries,
                                numQueries, results);
    return VLRResult_NoError;
    VLRA_CATCH
}

} // extern "C"
