/* C++ wrapper of the VLR C API (vlr.h) with the class and method names of the reference's
 * wrapper (libVLR/include/vlr/vlrcpp.h:15-835), so that C++ callers of vlr::Context compile
 * against this library unchanged apart from the device handle types (hipCtx_t, hipStream_t,
 * hipArray_t for CUcontext, CUstream, CUarray).
 *
 * Every API object is held by a shared_ptr "Ref" whose holder destroys the object when the
 * last reference goes; holders keep the context and the objects they use (materials of a
 * mesh, children of a node, the node behind a plug) alive.  Calls return through
 * Context::errorCheck, which throws std::runtime_error for the result codes enabled as
 * exceptions (by default InvalidContext, InvalidInstance and InternalError, as the reference's
 * wrapper, vlrcpp.h:684-701). */
#pragma once

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "vlr.h"
#include "basic_types.h"

namespace vlr {

class Context;
using ContextRef = std::shared_ptr<Context>;
using ContextConstRef = std::shared_ptr<const Context>;

#define VLR_PROCESS_CLASS(Name) \
    class Name##Holder;         \
    using Name##Ref = std::shared_ptr<Name##Holder>
VLR_PROCESS_CLASS_LIST();
#undef VLR_PROCESS_CLASS

static inline VLRContext getRawContext(const ContextConstRef &context);
static inline VLRResult checkResult(const ContextConstRef &context, VLRResult code);

class ObjectHolder : public std::enable_shared_from_this<ObjectHolder> {
protected:
    ContextConstRef m_context;
    VLRObject m_raw;
    VLRResult errorCheck(VLRResult code) const { return checkResult(m_context, code); }

public:
    explicit ObjectHolder(const ContextConstRef &context) : m_context(context), m_raw(nullptr) {}
    virtual ~ObjectHolder() {}
    std::string getType() const {
        const char *name = nullptr;
        errorCheck(vlrObjectGetType(m_raw, &name));
        return name ? name : "";
    }
    template <typename VLRType>
    VLRType getRaw() const {
        return (VLRType)m_raw;
    }
};

// a plug of a shader node, keeping the node alive
struct ShaderNodePlug {
    ShaderNodeRef node;
    VLRShaderNodePlug plug;
    ShaderNodePlug() { std::memset(&plug, 0, sizeof(plug)); }
    ShaderNodePlug(const ShaderNodeRef &n, const VLRShaderNodePlug &p) : node(n), plug(p) {}
};

struct ParameterInfo {
    ContextConstRef context;
    VLRParameterInfoConst raw;
    ParameterInfo(const ContextConstRef &c, VLRParameterInfoConst r) : context(c), raw(r) {}
    const char *getName(VLRResult *error = nullptr) const {
        const char *v = nullptr;
        VLRResult e = checkResult(context, vlrParameterInfoGetName(raw, &v));
        if (error) *error = e;
        return v;
    }
    const char *getType(VLRResult *error = nullptr) const {
        const char *v = nullptr;
        VLRResult e = checkResult(context, vlrParameterInfoGetType(raw, &v));
        if (error) *error = e;
        return v;
    }
    uint32_t getTupleSize(VLRResult *error = nullptr) const {
        uint32_t v = 0;
        VLRResult e = checkResult(context, vlrParameterInfoGetTupleSize(raw, &v));
        if (error) *error = e;
        return v;
    }
    VLRParameterFormFlag getFormFlags(VLRResult *error = nullptr) const {
        VLRParameterFormFlag v = VLRParameterFormFlag_ImmediateValue;
        VLRResult e = checkResult(context, vlrParameterInfoGetSocketForm(raw, &v));
        if (error) *error = e;
        return v;
    }
};

// objects with named parameters (images, shader nodes, materials, cameras)
class QueryableHolder : public ObjectHolder {
    // objects a parameter refers to, kept alive while they are set
    std::unordered_map<std::string, std::shared_ptr<ObjectHolder>> m_refs;

public:
    explicit QueryableHolder(const ContextConstRef &context) : ObjectHolder(context) {}

    bool get(const char *name, const char **enumValue) const {
        return errorCheck(vlrQueryableGetEnumValue(getRaw<VLRQueryable>(), name, enumValue)) == VLRResult_NoError;
    }
    bool get(const char *name, Point3D *value) const {
        return errorCheck(vlrQueryableGetPoint3D(getRaw<VLRQueryable>(), name, reinterpret_cast<VLRPoint3D *>(value))) ==
               VLRResult_NoError;
    }
    bool get(const char *name, Vector3D *value) const {
        return errorCheck(vlrQueryableGetVector3D(getRaw<VLRQueryable>(), name,
                                                  reinterpret_cast<VLRVector3D *>(value))) == VLRResult_NoError;
    }
    bool get(const char *name, Normal3D *value) const {
        return errorCheck(vlrQueryableGetNormal3D(getRaw<VLRQueryable>(), name,
                                                  reinterpret_cast<VLRNormal3D *>(value))) == VLRResult_NoError;
    }
    bool get(const char *name, Quaternion *value) const {
        return errorCheck(vlrQueryableGetQuaternion(getRaw<VLRQueryable>(), name,
                                                    reinterpret_cast<VLRQuaternion *>(value))) == VLRResult_NoError;
    }
    bool get(const char *name, float *value) const {
        return errorCheck(vlrQueryableGetFloat(getRaw<VLRQueryable>(), name, value)) == VLRResult_NoError;
    }
    bool get(const char *name, float *values, uint32_t length) const {
        return errorCheck(vlrQueryableGetFloatTuple(getRaw<VLRQueryable>(), name, values, length)) ==
               VLRResult_NoError;
    }
    bool get(const char *name, const float **values, uint32_t *length) const {
        return errorCheck(vlrQueryableGetFloatArray(getRaw<VLRQueryable>(), name, values, length)) ==
               VLRResult_NoError;
    }
    bool get(const char *name, VLRImmediateSpectrum *spectrum) const {
        return errorCheck(vlrQueryableGetImmediateSpectrum(getRaw<VLRQueryable>(), name, spectrum)) ==
               VLRResult_NoError;
    }
    bool get(const char *name, Image2DRef *image) const;
    bool get(const char *name, SurfaceMaterialRef *material) const;
    bool get(const char *name, ShaderNodePlug *plug) const;

    bool set(const char *name, const char *enumValue) const {
        return errorCheck(vlrQueryableSetEnumValue(getRaw<VLRQueryable>(), name, enumValue)) == VLRResult_NoError;
    }
    bool set(const char *name, const Point3D &value) const {
        return errorCheck(vlrQueryableSetPoint3D(getRaw<VLRQueryable>(), name,
                                                 reinterpret_cast<const VLRPoint3D *>(&value))) == VLRResult_NoError;
    }
    bool set(const char *name, const Vector3D &value) const {
        return errorCheck(vlrQueryableSetVector3D(getRaw<VLRQueryable>(), name,
                                                  reinterpret_cast<const VLRVector3D *>(&value))) == VLRResult_NoError;
    }
    bool set(const char *name, const Normal3D &value) const {
        return errorCheck(vlrQueryableSetNormal3D(getRaw<VLRQueryable>(), name,
                                                  reinterpret_cast<const VLRNormal3D *>(&value))) == VLRResult_NoError;
    }
    bool set(const char *name, const Quaternion &value) const {
        return errorCheck(vlrQueryableSetQuaternion(getRaw<VLRQueryable>(), name,
                                                    reinterpret_cast<const VLRQuaternion *>(&value))) ==
               VLRResult_NoError;
    }
    bool set(const char *name, float value) const {
        return errorCheck(vlrQueryableSetFloat(getRaw<VLRQueryable>(), name, value)) == VLRResult_NoError;
    }
    bool set(const char *name, const float *values, uint32_t length) const {
        return errorCheck(vlrQueryableSetFloatTuple(getRaw<VLRQueryable>(), name, values, length)) ==
               VLRResult_NoError;
    }
    bool set(const char *name, const VLRImmediateSpectrum &spectrum) const {
        return errorCheck(vlrQueryableSetImmediateSpectrum(getRaw<VLRQueryable>(), name, &spectrum)) ==
               VLRResult_NoError;
    }
    bool set(const char *name, const Image2DRef &image);
    bool set(const char *name, const SurfaceMaterialRef &material);
    bool set(const char *name, const ShaderNodePlug &plug);

    uint32_t getNumParameters() const {
        uint32_t n = 0;
        errorCheck(vlrQueryableGetNumParameters(getRaw<VLRQueryable>(), &n));
        return n;
    }
    ParameterInfo getParameterInfo(uint32_t index) const {
        VLRParameterInfoConst info = nullptr;
        errorCheck(vlrQueryableGetParameterInfo(getRaw<VLRQueryable>(), index, &info));
        return ParameterInfo(m_context, info);
    }

protected:
    void keep(const char *name, const std::shared_ptr<ObjectHolder> &ref) {
        if (ref)
            m_refs[name] = ref;
        else
            m_refs.erase(name);
    }
    std::shared_ptr<ObjectHolder> kept(const char *name) const {
        auto it = m_refs.find(name);
        return it == m_refs.end() ? nullptr : it->second;
    }
};

class Image2DHolder : public QueryableHolder {
public:
    explicit Image2DHolder(const ContextConstRef &context) : QueryableHolder(context) {}
    uint32_t getWidth() const {
        uint32_t v = 0;
        errorCheck(vlrImage2DGetWidth(getRaw<VLRImage2D>(), &v));
        return v;
    }
    uint32_t getHeight() const {
        uint32_t v = 0;
        errorCheck(vlrImage2DGetHeight(getRaw<VLRImage2D>(), &v));
        return v;
    }
    uint32_t getStride() const {
        uint32_t v = 0;
        errorCheck(vlrImage2DGetStride(getRaw<VLRImage2D>(), &v));
        return v;
    }
    const char *getOriginalDataFormat() const {
        const char *v = nullptr;
        errorCheck(vlrImage2DGetOriginalDataFormat(getRaw<VLRImage2D>(), &v));
        return v;
    }
    bool originalHasAlpha() const {
        bool v = false;
        errorCheck(vlrImage2DOriginalHasAlpha(getRaw<VLRImage2D>(), &v));
        return v;
    }
};

class LinearImage2DHolder : public Image2DHolder {
public:
    LinearImage2DHolder(const ContextConstRef &context, const uint8_t *linearData, uint32_t width, uint32_t height,
                        const char *format, const char *spectrumType, const char *colorSpace)
        : Image2DHolder(context) {
        errorCheck(vlrLinearImage2DCreate(getRawContext(m_context), const_cast<uint8_t *>(linearData), width, height,
                                          format, spectrumType, colorSpace,
                                          reinterpret_cast<VLRLinearImage2D *>(&m_raw)));
    }
    ~LinearImage2DHolder() {
        if (m_raw) vlrLinearImage2DDestroy(getRawContext(m_context), getRaw<VLRLinearImage2D>());
    }
};

class BlockCompressedImage2DHolder : public Image2DHolder {
public:
    BlockCompressedImage2DHolder(const ContextConstRef &context, uint8_t **data, size_t *sizes, uint32_t mipCount,
                                 uint32_t width, uint32_t height, const char *dataFormat, const char *spectrumType,
                                 const char *colorSpace)
        : Image2DHolder(context) {
        errorCheck(vlrBlockCompressedImage2DCreate(getRawContext(m_context), data, sizes, mipCount, width, height,
                                                   dataFormat, spectrumType, colorSpace,
                                                   reinterpret_cast<VLRBlockCompressedImage2D *>(&m_raw)));
    }
    ~BlockCompressedImage2DHolder() {
        if (m_raw) vlrBlockCompressedImage2DDestroy(getRawContext(m_context), getRaw<VLRBlockCompressedImage2D>());
    }
};

class ShaderNodeHolder : public QueryableHolder {
public:
    ShaderNodeHolder(const ContextConstRef &context, const char *typeName) : QueryableHolder(context) {
        errorCheck(vlrShaderNodeCreate(getRawContext(m_context), typeName, reinterpret_cast<VLRShaderNode *>(&m_raw)));
    }
    ~ShaderNodeHolder() {
        if (m_raw) vlrShaderNodeDestroy(getRawContext(m_context), getRaw<VLRShaderNode>());
    }
    ShaderNodePlug getPlug(VLRShaderNodePlugType plugType, uint32_t option) {
        VLRShaderNodePlug plug;
        errorCheck(vlrShaderNodeGetPlug(getRaw<VLRShaderNode>(), plugType, option, &plug));
        return ShaderNodePlug(std::static_pointer_cast<ShaderNodeHolder>(shared_from_this()), plug);
    }
};

class SurfaceMaterialHolder : public QueryableHolder {
public:
    SurfaceMaterialHolder(const ContextConstRef &context, const char *typeName) : QueryableHolder(context) {
        errorCheck(vlrSurfaceMaterialCreate(getRawContext(m_context), typeName,
                                            reinterpret_cast<VLRSurfaceMaterial *>(&m_raw)));
    }
    ~SurfaceMaterialHolder() {
        if (m_raw) vlrSurfaceMaterialDestroy(getRawContext(m_context), getRaw<VLRSurfaceMaterial>());
    }
};

inline bool QueryableHolder::set(const char *name, const Image2DRef &image) {
    VLRResult r = errorCheck(vlrQueryableSetImage2D(getRaw<VLRQueryable>(), name,
                                                    image ? image->getRaw<VLRImage2DConst>() : nullptr));
    if (r == VLRResult_NoError) keep(name, image);
    return r == VLRResult_NoError;
}
inline bool QueryableHolder::set(const char *name, const SurfaceMaterialRef &material) {
    VLRResult r = errorCheck(vlrQueryableSetSurfaceMaterial(
        getRaw<VLRQueryable>(), name, material ? material->getRaw<VLRSurfaceMaterialConst>() : nullptr));
    if (r == VLRResult_NoError) keep(name, material);
    return r == VLRResult_NoError;
}
inline bool QueryableHolder::set(const char *name, const ShaderNodePlug &plug) {
    VLRResult r = errorCheck(vlrQueryableSetShaderNodePlug(getRaw<VLRQueryable>(), name, plug.plug));
    if (r == VLRResult_NoError) keep(name, plug.node);
    return r == VLRResult_NoError;
}
inline bool QueryableHolder::get(const char *name, Image2DRef *image) const {
    VLRImage2DConst raw = nullptr;
    if (errorCheck(vlrQueryableGetImage2D(getRaw<VLRQueryable>(), name, &raw)) != VLRResult_NoError) return false;
    *image = raw ? std::static_pointer_cast<Image2DHolder>(kept(name)) : nullptr;
    return true;
}
inline bool QueryableHolder::get(const char *name, SurfaceMaterialRef *material) const {
    VLRSurfaceMaterialConst raw = nullptr;
    if (errorCheck(vlrQueryableGetSurfaceMaterial(getRaw<VLRQueryable>(), name, &raw)) != VLRResult_NoError)
        return false;
    *material = raw ? std::static_pointer_cast<SurfaceMaterialHolder>(kept(name)) : nullptr;
    return true;
}
inline bool QueryableHolder::get(const char *name, ShaderNodePlug *plug) const {
    VLRShaderNodePlug raw;
    if (errorCheck(vlrQueryableGetShaderNodePlug(getRaw<VLRQueryable>(), name, &raw)) != VLRResult_NoError)
        return false;
    *plug = raw.nodeRef ? ShaderNodePlug(std::static_pointer_cast<ShaderNodeHolder>(kept(name)), raw)
                        : ShaderNodePlug();
    return true;
}

class TransformHolder : public ObjectHolder {
public:
    explicit TransformHolder(const ContextConstRef &context) : ObjectHolder(context) {}
};

class StaticTransformHolder : public TransformHolder {
public:
    StaticTransformHolder(const ContextConstRef &context, const float mat[16]) : TransformHolder(context) {
        errorCheck(vlrStaticTransformCreate(getRawContext(m_context), mat,
                                            reinterpret_cast<VLRStaticTransform *>(&m_raw)));
    }
    StaticTransformHolder(const ContextConstRef &context, const Matrix4x4 &mat)
        : StaticTransformHolder(context, mat.m) {}
    ~StaticTransformHolder() {
        if (m_raw) vlrStaticTransformDestroy(getRawContext(m_context), getRaw<VLRStaticTransform>());
    }
    void getArrays(float mat[16], float invMat[16]) const {
        errorCheck(vlrStaticTransformGetArrays(getRaw<VLRStaticTransformConst>(), mat, invMat));
    }
    void getMatrices(Matrix4x4 *mat, Matrix4x4 *invMat) const {
        float a[16], b[16];
        getArrays(a, b);
        if (mat) *mat = Matrix4x4(a);
        if (invMat) *invMat = Matrix4x4(b);
    }
};

class NodeHolder : public ObjectHolder {
public:
    explicit NodeHolder(const ContextConstRef &context) : ObjectHolder(context) {}
    void setName(const std::string &name) const { errorCheck(vlrNodeSetName(getRaw<VLRNode>(), name.c_str())); }
    const char *getName() const {
        const char *v = nullptr;
        errorCheck(vlrNodeGetName(getRaw<VLRNodeConst>(), &v));
        return v;
    }
};

class SurfaceNodeHolder : public NodeHolder {
public:
    explicit SurfaceNodeHolder(const ContextConstRef &context) : NodeHolder(context) {}
};

class TriangleMeshSurfaceNodeHolder : public SurfaceNodeHolder {
    std::vector<std::shared_ptr<ObjectHolder>> m_used;   // materials and plug nodes of the groups

public:
    TriangleMeshSurfaceNodeHolder(const ContextConstRef &context, const char *name) : SurfaceNodeHolder(context) {
        errorCheck(vlrTriangleMeshSurfaceNodeCreate(getRawContext(m_context), name,
                                                    reinterpret_cast<VLRTriangleMeshSurfaceNode *>(&m_raw)));
    }
    ~TriangleMeshSurfaceNodeHolder() {
        if (m_raw) vlrTriangleMeshSurfaceNodeDestroy(getRawContext(m_context), getRaw<VLRTriangleMeshSurfaceNode>());
    }
    void setVertices(Vertex *vertices, uint32_t numVertices) {
        errorCheck(vlrTriangleMeshSurfaceNodeSetVertices(getRaw<VLRTriangleMeshSurfaceNode>(),
                                                         reinterpret_cast<const VLRVertex *>(vertices), numVertices));
    }
    void addMaterialGroup(uint32_t *indices, uint32_t numIndices, const SurfaceMaterialRef &material,
                          const ShaderNodePlug &nodeNormal = ShaderNodePlug(),
                          const ShaderNodePlug &nodeTangent = ShaderNodePlug(),
                          const ShaderNodePlug &nodeAlpha = ShaderNodePlug()) {
        errorCheck(vlrTriangleMeshSurfaceNodeAddMaterialGroup(
            getRaw<VLRTriangleMeshSurfaceNode>(), indices, numIndices, material->getRaw<VLRSurfaceMaterialConst>(),
            nodeNormal.plug, nodeTangent.plug, nodeAlpha.plug));
        m_used.push_back(material);
        for (const ShaderNodePlug *p : {&nodeNormal, &nodeTangent, &nodeAlpha})
            if (p->node) m_used.push_back(p->node);
    }
};

class PointSurfaceNodeHolder : public SurfaceNodeHolder {
    std::vector<std::shared_ptr<ObjectHolder>> m_used;

public:
    PointSurfaceNodeHolder(const ContextConstRef &context, const char *name) : SurfaceNodeHolder(context) {
        errorCheck(vlrPointSurfaceNodeCreate(getRawContext(m_context), name,
                                             reinterpret_cast<VLRPointSurfaceNode *>(&m_raw)));
    }
    ~PointSurfaceNodeHolder() {
        if (m_raw) vlrPointSurfaceNodeDestroy(getRawContext(m_context), getRaw<VLRPointSurfaceNode>());
    }
    void setVertices(Vertex *vertices, uint32_t numVertices) {
        errorCheck(vlrPointSurfaceNodeSetVertices(getRaw<VLRPointSurfaceNode>(),
                                                  reinterpret_cast<const VLRVertex *>(vertices), numVertices));
    }
    void addMaterialGroup(uint32_t *indices, uint32_t numIndices, const SurfaceMaterialRef &material) {
        errorCheck(vlrPointSurfaceNodeAddMaterialGroup(getRaw<VLRPointSurfaceNode>(), indices, numIndices,
                                                       material->getRaw<VLRSurfaceMaterialConst>()));
        m_used.push_back(material);
    }
};

class ParentNodeHolder : public NodeHolder {
    StaticTransformRef m_transform;
    std::vector<NodeRef> m_children;   // in the order they were added

public:
    ParentNodeHolder(const ContextConstRef &context, const StaticTransformRef &transform)
        : NodeHolder(context), m_transform(transform) {}
    void setTransform(const StaticTransformRef &transform) {
        errorCheck(vlrParentNodeSetTransform(getRaw<VLRParentNode>(), transform->getRaw<VLRTransformConst>()));
        m_transform = transform;
    }
    TransformRef getTransform() const { return m_transform; }
    void addChild(const ParentNodeRef &child) {
        errorCheck(vlrParentNodeAddChild(getRaw<VLRParentNode>(), child->getRaw<VLRNode>()));
        m_children.push_back(child);
    }
    void removeChild(const ParentNodeRef &child) {
        errorCheck(vlrParentNodeRemoveChild(getRaw<VLRParentNode>(), child->getRaw<VLRNode>()));
        forget(child);
    }
    void addChild(const SurfaceNodeRef &child) {
        errorCheck(vlrParentNodeAddChild(getRaw<VLRParentNode>(), child->getRaw<VLRNode>()));
        m_children.push_back(child);
    }
    void removeChild(const SurfaceNodeRef &child) {
        errorCheck(vlrParentNodeRemoveChild(getRaw<VLRParentNode>(), child->getRaw<VLRNode>()));
        forget(child);
    }
    uint32_t getNumChildren() const {
        uint32_t n = 0;
        errorCheck(vlrParentNodeGetNumChildren(getRaw<VLRParentNodeConst>(), &n));
        return n;
    }
    void getChildren(uint32_t numChildren, NodeRef *children) const {
        for (uint32_t i = 0; i < numChildren && i < m_children.size(); ++i) children[i] = m_children[i];
    }
    NodeRef getChildAt(uint32_t index) const { return index < m_children.size() ? m_children[index] : nullptr; }

private:
    void forget(const NodeRef &child) {
        for (auto it = m_children.begin(); it != m_children.end(); ++it)
            if (*it == child) {
                m_children.erase(it);
                return;
            }
    }
};

class InternalNodeHolder : public ParentNodeHolder {
public:
    InternalNodeHolder(const ContextConstRef &context, const char *name, const StaticTransformRef &transform)
        : ParentNodeHolder(context, transform) {
        errorCheck(vlrInternalNodeCreate(getRawContext(m_context), name, transform->getRaw<VLRTransformConst>(),
                                         reinterpret_cast<VLRInternalNode *>(&m_raw)));
    }
    ~InternalNodeHolder() {
        if (m_raw) vlrInternalNodeDestroy(getRawContext(m_context), getRaw<VLRInternalNode>());
    }
};

class SceneHolder : public ParentNodeHolder {
    SurfaceMaterialRef m_environment;

public:
    SceneHolder(const ContextConstRef &context, const StaticTransformRef &transform)
        : ParentNodeHolder(context, transform) {
        errorCheck(vlrSceneCreate(getRawContext(m_context), transform->getRaw<VLRTransformConst>(),
                                  reinterpret_cast<VLRScene *>(&m_raw)));
    }
    ~SceneHolder() {
        if (m_raw) vlrSceneDestroy(getRawContext(m_context), getRaw<VLRScene>());
    }
    void setEnvironment(const SurfaceMaterialRef &matEnv, float rotationPhi) {
        errorCheck(vlrSceneSetEnvironment(getRaw<VLRScene>(), matEnv->getRaw<VLRSurfaceMaterial>()));
        errorCheck(vlrSceneSetEnvironmentRotation(getRaw<VLRScene>(), rotationPhi));
        m_environment = matEnv;
    }
    void setEnvironmentRotation(float rotationPhi) {
        errorCheck(vlrSceneSetEnvironmentRotation(getRaw<VLRScene>(), rotationPhi));
    }
};

class CameraHolder : public QueryableHolder {
public:
    CameraHolder(const ContextConstRef &context, const char *typeName) : QueryableHolder(context) {
        errorCheck(vlrCameraCreate(getRawContext(m_context), typeName, reinterpret_cast<VLRCamera *>(&m_raw)));
    }
    ~CameraHolder() {
        if (m_raw) vlrCameraDestroy(getRawContext(m_context), getRaw<VLRCamera>());
    }
};

class Context : public std::enable_shared_from_this<Context> {
    VLRContext m_rawContext = nullptr;
    hipCtx_t m_hipContext = nullptr;
    bool m_throws[VLRResult_NumErrors] = {};
    mutable StaticTransformRef m_identity;
    SceneRef m_scene;

    Context() {
        m_throws[VLRResult_InvalidContext] = true;
        m_throws[VLRResult_InvalidInstance] = true;
        m_throws[VLRResult_InternalError] = true;
    }
    void initialize(hipCtx_t hipContext, bool logging, uint32_t maxCallableDepth) {
        m_hipContext = hipContext;
        errorCheck(vlrCreateContext(hipContext, logging, maxCallableDepth, &m_rawContext));
    }

public:
    static ContextRef create(hipCtx_t hipContext, bool logging, uint32_t maxCallableDepth = 8) {
        ContextRef c(new Context());
        c->initialize(hipContext, logging, maxCallableDepth);
        return c;
    }
    ~Context() {
        m_scene.reset();
        m_identity.reset();
        if (m_rawContext) vlrDestroyContext(m_rawContext);
    }

    void enableException(VLRResult code) { m_throws[code] = true; }
    void disableException(VLRResult code) { m_throws[code] = false; }
    void enableAllExceptions() {
        for (int i = 0; i < VLRResult_NumErrors; ++i) m_throws[i] = i != VLRResult_NoError;
    }
    void disableAllExceptions() {
        for (bool &t : m_throws) t = false;
    }
    VLRResult errorCheck(VLRResult code) const {
        if (code != VLRResult_NoError && code < VLRResult_NumErrors && m_throws[code])
            throw std::runtime_error(vlrGetErrorMessage(code));
        return code;
    }

    VLRContext get() const { return m_rawContext; }
    hipCtx_t getCUcontext() const { return m_hipContext; }

    void bindOutputBuffer(uint32_t width, uint32_t height, uint32_t glTexID) const {
        errorCheck(vlrContextBindOutputBuffer(m_rawContext, width, height, glTexID));
    }
    void getOutputBufferSize(uint32_t *width, uint32_t *height) const {
        errorCheck(vlrContextGetOutputBufferSize(m_rawContext, width, height));
    }
    void getOutputBuffer(hipArray_t *array) const { errorCheck(vlrContextGetOutputBuffer(m_rawContext, array)); }
    void readOutputBuffer(float *data) const { errorCheck(vlrContextReadOutputBuffer(m_rawContext, data)); }
    void setScene(const SceneRef &scene) {
        errorCheck(vlrContextSetScene(m_rawContext, scene->getRaw<VLRScene>()));
        m_scene = scene;
    }
    void setRenderer(VLRRenderer renderer) const { errorCheck(vlrContextSetRenderer(m_rawContext, renderer)); }
    void setDebugRenderingAttribute(VLRDebugRenderingMode attr) const {
        errorCheck(vlrContextSetDebugRenderingAttribute(m_rawContext, attr));
    }
    void setProbePixel(int32_t x, int32_t y) const { errorCheck(vlrContextSetProbePixel(m_rawContext, x, y)); }
    void render(hipStream_t stream, const CameraRef &camera, bool denoise, uint32_t shrinkCoeff, bool firstFrame,
                uint32_t limitNumAccumFrames, uint32_t *numAccumFrames) const {
        errorCheck(vlrContextRender(m_rawContext, stream, camera->getRaw<VLRCameraConst>(), denoise, shrinkCoeff,
                                    firstFrame, limitNumAccumFrames, numAccumFrames));
    }

    LinearImage2DRef createLinearImage2D(const uint8_t *linearData, uint32_t width, uint32_t height, const char *format,
                                         const char *spectrumType, const char *colorSpace) const {
        return std::make_shared<LinearImage2DHolder>(shared_from_this(), linearData, width, height, format,
                                                     spectrumType, colorSpace);
    }
    BlockCompressedImage2DRef createBlockCompressedImage2D(uint8_t **data, size_t *sizes, uint32_t mipCount,
                                                           uint32_t width, uint32_t height, const char *dataFormat,
                                                           const char *spectrumType, const char *colorSpace) const {
        return std::make_shared<BlockCompressedImage2DHolder>(shared_from_this(), data, sizes, mipCount, width,
                                                              height, dataFormat, spectrumType, colorSpace);
    }
    ShaderNodeRef createShaderNode(const char *typeName) const {
        return std::make_shared<ShaderNodeHolder>(shared_from_this(), typeName);
    }
    SurfaceMaterialRef createSurfaceMaterial(const char *typeName) const {
        return std::make_shared<SurfaceMaterialHolder>(shared_from_this(), typeName);
    }
    StaticTransformRef createStaticTransform(const float mat[16]) const {
        return std::make_shared<StaticTransformHolder>(shared_from_this(), mat);
    }
    StaticTransformRef createStaticTransform(const Matrix4x4 &mat) const {
        return std::make_shared<StaticTransformHolder>(shared_from_this(), mat);
    }
    const StaticTransformRef &getIdentityTransform() const {
        if (!m_identity) m_identity = createStaticTransform(Matrix4x4());
        return m_identity;
    }
    TriangleMeshSurfaceNodeRef createTriangleMeshSurfaceNode(const char *name) const {
        return std::make_shared<TriangleMeshSurfaceNodeHolder>(shared_from_this(), name);
    }
    PointSurfaceNodeRef createPointSurfaceNode(const char *name) const {
        return std::make_shared<PointSurfaceNodeHolder>(shared_from_this(), name);
    }
    InternalNodeRef createInternalNode(const char *name, const StaticTransformRef &transform = nullptr) const {
        return std::make_shared<InternalNodeHolder>(shared_from_this(), name,
                                                    transform ? transform : getIdentityTransform());
    }
    SceneRef createScene(const StaticTransformRef &transform = nullptr) const {
        return std::make_shared<SceneHolder>(shared_from_this(), transform ? transform : getIdentityTransform());
    }
    CameraRef createCamera(const char *typeName) const {
        return std::make_shared<CameraHolder>(shared_from_this(), typeName);
    }
};

static inline VLRContext getRawContext(const ContextConstRef &context) { return context->get(); }
static inline VLRResult checkResult(const ContextConstRef &context, VLRResult code) {
    return context->errorCheck(code);
}

} // namespace vlr
