/* VLR C API on MI355X (gfx950).
 *
 * Every entry point of the reference's libVLR/include/vlr/vlr.h is declared here with the
 * same name, argument list and VLRResult semantics; each comment cites the reference
 * declaration it replaces.  The only difference is the device-handle types: the reference
 * takes CUcontext / CUstream / CUarray from <cuda.h> (vlr.h:4); this library takes
 * hipCtx_t / hipStream_t / hipArray_t from HIP's runtime API.  A null hipCtx_t means "the
 * calling thread's current HIP device".
 *
 * Render path: vlrContextRender runs a wavefront path tracer made of hand-written HIP
 * kernels (generate, intersect, shade + next-event estimation, shadow, accumulate,
 * convert) over the library's own two-level BVH.
 */
#pragma once

#include "public_types.h"
#include <hip/hip_runtime_api.h>

#if defined(__GNUC__)
#   define VLR_API __attribute__((visibility("default")))
#else
#   define VLR_API
#endif

#if defined(__cplusplus)
extern "C" {
#endif

/* vlr.h:28-40 */
enum VLRResult {
    VLRResult_NoError = 0,
    VLRResult_InvalidContext,
    VLRResult_InvalidInstance,
    VLRResult_InvalidArgument,
    VLRResult_IncompatibleNodeType,
    VLRResult_InternalError,
    VLRResult_NumErrors
};
#if !defined(__cplusplus)
typedef enum VLRResult VLRResult;
#endif

/* vlr.h:46-60: opaque handles */
#if !defined(VLR_TYPEDEF_DONE)
#   define VLR_PROCESS_CLASS(name) \
        typedef struct VLR ## name ## _API* VLR ## name; \
        typedef const struct VLR ## name ## _API* VLR ## name ## Const
VLR_PROCESS_CLASS(Context);
VLR_PROCESS_CLASS_LIST();
#   undef VLR_PROCESS_CLASS
#   define VLR_TYPEDEF_DONE
#endif

/* vlr.h:64 */
VLR_API const char* vlrGetErrorMessage(VLRResult code);

/* ---- Context ---- */
/* vlr.h:70-72 */
VLR_API VLRResult vlrCreateContext(hipCtx_t hipContext, bool logging, uint32_t maxCallableDepth, VLRContext* context);
/* vlr.h:73-74 */
VLR_API VLRResult vlrDestroyContext(VLRContext context);
/* vlr.h:75-77 */
VLR_API VLRResult vlrContextGetCUcontext(VLRContext context, hipCtx_t* hipContext);
/* vlr.h:79-81; glTexID must be 0 (no OpenGL interop on this platform) */
VLR_API VLRResult vlrContextBindOutputBuffer(VLRContext context, uint32_t width, uint32_t height, uint32_t glTexID);
/* vlr.h:82-84 */
VLR_API VLRResult vlrContextGetOutputBufferSize(VLRContext context, uint32_t* width, uint32_t* height);
/* vlr.h:85-87: a width x height hipArray of float4 (RGBA32F) that holds the converted output;
 * made on the first call, then written on the render stream after every rendered frame; owned
 * by the context (freed when the output is rebound at another size or the context is
 * destroyed).  A GPU without image support (hipDeviceAttributeImageSupport 0, as gfx950
 * reports) has no arrays: the call returns VLRResult_InternalError and the output is read with
 * vlrContextReadOutputBuffer. */
VLR_API VLRResult vlrContextGetOutputBuffer(VLRContext context, hipArray_t* array);
/* vlr.h:88-90: synchronous copy of width*height RGBA floats */
VLR_API VLRResult vlrContextReadOutputBuffer(VLRContext context, float* data);
/* vlr.h:91-93 */
VLR_API VLRResult vlrContextSetScene(VLRContext context, VLRScene scene);
/* vlr.h:94-96 */
VLR_API VLRResult vlrContextSetRenderer(VLRContext context, VLRRenderer renderer);
/* vlr.h:97-99 */
VLR_API VLRResult vlrContextSetDebugRenderingAttribute(VLRContext context, VLRDebugRenderingMode attr);
/* vlr.h:100-102 */
VLR_API VLRResult vlrContextSetProbePixel(VLRContext context, int32_t x, int32_t y);
/* vlr.h:103-106: one sample per pixel per call while *numAccumFrames < limit; denoise is
 * accepted and ignored (the reference's denoiser is a closed NVIDIA component) */
VLR_API VLRResult vlrContextRender(VLRContext context, hipStream_t stream, VLRCameraConst camera, bool denoise,
                                   uint32_t shrinkCoeff, bool firstFrame, uint32_t limitNumAccumFrames,
                                   uint32_t* numAccumFrames);

/* ---- Object / ParameterInfo / enums ---- */
/* vlr.h:110-112 */
VLR_API VLRResult vlrObjectGetType(VLRObjectConst object, const char** typeName);
/* vlr.h:116-130 */
VLR_API VLRResult vlrParameterInfoGetName(VLRParameterInfoConst paramInfo, const char** name);
VLR_API VLRResult vlrParameterInfoGetSocketForm(VLRParameterInfoConst paramInfo, VLRParameterFormFlag* form);
VLR_API VLRResult vlrParameterInfoGetType(VLRParameterInfoConst paramInfo, const char** type);
VLR_API VLRResult vlrParameterInfoGetTupleSize(VLRParameterInfoConst paramInfo, uint32_t* size);
/* vlr.h:131-136 */
VLR_API VLRResult vlrGetNumEnumMembers(const char* typeName, uint32_t* numMembers);
VLR_API VLRResult vlrGetEnumMember(const char* typeName, uint32_t index, const char** value);

/* ---- Queryable (vlr.h:143-235) ---- */
VLR_API VLRResult vlrQueryableGetNumParameters(VLRQueryableConst queryable, uint32_t* numParams);
VLR_API VLRResult vlrQueryableGetParameterInfo(VLRQueryableConst queryable, uint32_t index, VLRParameterInfoConst* paramInfo);
VLR_API VLRResult vlrQueryableGetEnumValue(VLRQueryableConst queryable, const char* paramName, const char** value);
VLR_API VLRResult vlrQueryableGetPoint3D(VLRQueryableConst queryable, const char* paramName, VLRPoint3D* value);
VLR_API VLRResult vlrQueryableGetVector3D(VLRQueryableConst queryable, const char* paramName, VLRVector3D* value);
VLR_API VLRResult vlrQueryableGetNormal3D(VLRQueryableConst queryable, const char* paramName, VLRNormal3D* value);
VLR_API VLRResult vlrQueryableGetQuaternion(VLRQueryableConst queryable, const char* paramName, VLRQuaternion* value);
VLR_API VLRResult vlrQueryableGetFloat(VLRQueryableConst queryable, const char* paramName, float* value);
VLR_API VLRResult vlrQueryableGetFloatTuple(VLRQueryableConst queryable, const char* paramName, float* values, uint32_t length);
VLR_API VLRResult vlrQueryableGetFloatArray(VLRQueryableConst queryable, const char* paramName, const float** values, uint32_t* length);
VLR_API VLRResult vlrQueryableGetImage2D(VLRQueryableConst queryable, const char* paramName, VLRImage2DConst* image);
VLR_API VLRResult vlrQueryableGetImmediateSpectrum(VLRQueryableConst queryable, const char* paramName, VLRImmediateSpectrum* value);
VLR_API VLRResult vlrQueryableGetSurfaceMaterial(VLRQueryableConst queryable, const char* paramName, VLRSurfaceMaterialConst* value);
VLR_API VLRResult vlrQueryableGetShaderNodePlug(VLRQueryableConst queryable, const char* paramName, VLRShaderNodePlug* plug);
VLR_API VLRResult vlrQueryableSetEnumValue(VLRQueryable queryable, const char* paramName, const char* value);
VLR_API VLRResult vlrQueryableSetPoint3D(VLRQueryable queryable, const char* paramName, const VLRPoint3D* value);
VLR_API VLRResult vlrQueryableSetVector3D(VLRQueryable queryable, const char* paramName, const VLRVector3D* value);
VLR_API VLRResult vlrQueryableSetNormal3D(VLRQueryable queryable, const char* paramName, const VLRNormal3D* value);
VLR_API VLRResult vlrQueryableSetQuaternion(VLRQueryable queryable, const char* paramName, const VLRQuaternion* value);
VLR_API VLRResult vlrQueryableSetFloat(VLRQueryable queryable, const char* paramName, float value);
VLR_API VLRResult vlrQueryableSetFloatTuple(VLRQueryable queryable, const char* paramName, const float* values, uint32_t length);
VLR_API VLRResult vlrQueryableSetImage2D(VLRQueryable queryable, const char* paramName, VLRImage2DConst image);
VLR_API VLRResult vlrQueryableSetImmediateSpectrum(VLRQueryable queryable, const char* paramName, const VLRImmediateSpectrum* value);
VLR_API VLRResult vlrQueryableSetSurfaceMaterial(VLRQueryable queryable, const char* paramName, VLRSurfaceMaterialConst value);
VLR_API VLRResult vlrQueryableSetShaderNodePlug(VLRQueryable queryable, const char* paramName, VLRShaderNodePlug plug);

/* ---- Image2D (vlr.h:238-282) ---- */
VLR_API VLRResult vlrImage2DGetWidth(VLRImage2DConst image, uint32_t* width);
VLR_API VLRResult vlrImage2DGetHeight(VLRImage2DConst image, uint32_t* height);
VLR_API VLRResult vlrImage2DGetStride(VLRImage2DConst image, uint32_t* stride);
VLR_API VLRResult vlrImage2DGetOriginalDataFormat(VLRImage2DConst image, const char** format);
VLR_API VLRResult vlrImage2DOriginalHasAlpha(VLRImage2DConst image, bool* hasAlpha);
VLR_API VLRResult vlrLinearImage2DCreate(VLRContext context, uint8_t* linearData, uint32_t width, uint32_t height,
                                         const char* format, const char* spectrumType, const char* colorSpace,
                                         VLRLinearImage2D* image);
VLR_API VLRResult vlrLinearImage2DDestroy(VLRContext context, VLRLinearImage2D image);
/* vlr.h:271-279: block-compressed images are accepted for API compatibility; their
 * pixels are not sampled by the render path (texture-driven nodes are not on it) */
VLR_API VLRResult vlrBlockCompressedImage2DCreate(VLRContext context, uint8_t** data, size_t* sizes, uint32_t mipCount,
                                                  uint32_t width, uint32_t height, const char* dataFormat,
                                                  const char* spectrumType, const char* colorSpace,
                                                  VLRBlockCompressedImage2D* image);
VLR_API VLRResult vlrBlockCompressedImage2DDestroy(VLRContext context, VLRBlockCompressedImage2D image);

/* ---- ShaderNode (vlr.h:284-299) ---- */
VLR_API VLRResult vlrShaderNodeCreate(VLRContext context, const char* typeName, VLRShaderNode* node);
VLR_API VLRResult vlrShaderNodeDestroy(VLRContext context, VLRShaderNode node);
VLR_API VLRResult vlrShaderNodeGetPlug(VLRShaderNodeConst node, VLRShaderNodePlugType plugType, uint32_t option,
                                       VLRShaderNodePlug* plug);

/* ---- SurfaceMaterial (vlr.h:301-311) ---- */
VLR_API VLRResult vlrSurfaceMaterialCreate(VLRContext context, const char* typeName, VLRSurfaceMaterial* material);
VLR_API VLRResult vlrSurfaceMaterialDestroy(VLRContext context, VLRSurfaceMaterial material);

/* ---- Transform (vlr.h:313-326), matrices column-major ---- */
VLR_API VLRResult vlrStaticTransformCreate(VLRContext context, const float mat[16], VLRStaticTransform* transform);
VLR_API VLRResult vlrStaticTransformDestroy(VLRContext context, VLRStaticTransform transform);
VLR_API VLRResult vlrStaticTransformGetArrays(VLRStaticTransformConst transform, float mat[16], float invMat[16]);

/* ---- Node (vlr.h:330-337) ---- */
VLR_API VLRResult vlrNodeSetName(VLRNode node, const char* name);
VLR_API VLRResult vlrNodeGetName(VLRNodeConst node, const char** name);

/* ---- TriangleMeshSurfaceNode (vlr.h:341-355) ---- */
VLR_API VLRResult vlrTriangleMeshSurfaceNodeCreate(VLRContext context, const char* name, VLRTriangleMeshSurfaceNode* surfaceNode);
VLR_API VLRResult vlrTriangleMeshSurfaceNodeDestroy(VLRContext context, VLRTriangleMeshSurfaceNode surfaceNode);
VLR_API VLRResult vlrTriangleMeshSurfaceNodeSetVertices(VLRTriangleMeshSurfaceNode surfaceNode, const VLRVertex* vertices, uint32_t numVertices);
VLR_API VLRResult vlrTriangleMeshSurfaceNodeAddMaterialGroup(VLRTriangleMeshSurfaceNode surfaceNode, const uint32_t* indices,
                                                             uint32_t numIndices, VLRSurfaceMaterialConst material,
                                                             VLRShaderNodePlug nodeNormal, VLRShaderNodePlug nodeTangent,
                                                             VLRShaderNodePlug nodeAlpha);

/* ---- PointSurfaceNode (vlr.h:361-374) ---- */
VLR_API VLRResult vlrPointSurfaceNodeCreate(VLRContext context, const char* name, VLRPointSurfaceNode* surfaceNode);
VLR_API VLRResult vlrPointSurfaceNodeDestroy(VLRContext context, VLRPointSurfaceNode surfaceNode);
VLR_API VLRResult vlrPointSurfaceNodeSetVertices(VLRPointSurfaceNode surfaceNode, const VLRVertex* vertices, uint32_t numVertices);
VLR_API VLRResult vlrPointSurfaceNodeAddMaterialGroup(VLRPointSurfaceNode surfaceNode, const uint32_t* indices,
                                                      uint32_t numIndices, VLRSurfaceMaterialConst material);

/* ---- ParentNode (vlr.h:381-405) ---- */
VLR_API VLRResult vlrParentNodeSetTransform(VLRParentNode node, VLRTransformConst localToWorld);
VLR_API VLRResult vlrParentNodeGetTransform(VLRParentNodeConst node, VLRTransformConst* localToWorld);
VLR_API VLRResult vlrParentNodeAddChild(VLRParentNode node, VLRNode child);
VLR_API VLRResult vlrParentNodeRemoveChild(VLRParentNode node, VLRNode child);
VLR_API VLRResult vlrParentNodeGetNumChildren(VLRParentNodeConst node, uint32_t* numChildren);
VLR_API VLRResult vlrParentNodeGetChildren(VLRParentNodeConst node, VLRNode* children);
VLR_API VLRResult vlrParentNodeGetChildAt(VLRParentNodeConst node, uint32_t index, VLRNode* child);

/* ---- InternalNode (vlr.h:408-415) ---- */
VLR_API VLRResult vlrInternalNodeCreate(VLRContext context, const char* name, VLRTransformConst transform, VLRInternalNode* node);
VLR_API VLRResult vlrInternalNodeDestroy(VLRContext context, VLRInternalNode node);

/* ---- Scene (vlr.h:420-432) ---- */
VLR_API VLRResult vlrSceneCreate(VLRContext context, VLRTransformConst transform, VLRScene* scene);
VLR_API VLRResult vlrSceneDestroy(VLRContext context, VLRScene scene);
VLR_API VLRResult vlrSceneSetEnvironment(VLRScene scene, VLRSurfaceMaterial material);
VLR_API VLRResult vlrSceneSetEnvironmentRotation(VLRScene scene, float rotationPhi);

/* ---- Camera (vlr.h:438-444) ---- */
VLR_API VLRResult vlrCameraCreate(VLRContext context, const char* typeName, VLRCamera* camera);
VLR_API VLRResult vlrCameraDestroy(VLRContext context, VLRCamera camera);

#if defined(__cplusplus)
}
#endif

#include "vlr_amd.h"
