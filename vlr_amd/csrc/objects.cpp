// Queryable objects of the VLR API: images, shader nodes, surface materials, transforms,
// cameras.  Parameter names, forms and defaults follow the reference's tables:
//   materials      libVLR/materials.cpp:247-2179 (defaults in each constructor)
//   shader nodes   libVLR/shader_nodes.cpp:165-1866
//   cameras        libVLR/scene.cpp:1490-1496, 1624-1663, 1684-1688
//   enum strings   libVLR/queryable.cpp:34-80
#include "runtime.h"

#include <atomic>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>

namespace vlra {

uint64_t nextGeometryVersion() {
    static std::atomic<uint64_t> counter{0};
    return ++counter;
}

bool paramNameEquals(const std::string &a, const std::string &b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (std::tolower(static_cast<unsigned char>(a[i])) != std::tolower(static_cast<unsigned char>(b[i])))
            return false;
    return true;
}

Object::~Object() {}

bool Object::isQueryable() const {
    return isImage() || type == ObjType::ShaderNode || type == ObjType::SurfaceMaterial || type == ObjType::Camera;
}

bool Object::isA(const char *n) const {
    std::string s = n;
    if (s == "Object") return true;
    if (s == "Queryable") return isQueryable();
    if (s == "Image2D") return isImage();
    if (s == "Node") return isNode();
    if (s == "SurfaceNode") return isSurfaceNode();
    if (s == "ParentNode") return isParentNode();
    if (s == "Transform") return type == ObjType::StaticTransform;
    return false;
}

namespace {
constexpr uint32_t IMM = VLRParameterFormFlag_ImmediateValue;
constexpr uint32_t NODE = VLRParameterFormFlag_Node;
constexpr uint32_t BOTH = VLRParameterFormFlag_Both;

std::vector<ParamInfo *> mk(std::initializer_list<ParamInfo> l) {
    std::vector<ParamInfo *> v;
    for (const auto &p : l) v.push_back(new ParamInfo(p.name.c_str(), p.formFlags, p.typeStr.c_str(), p.tupleSize));
    return v;
}

struct MatSpec {
    const char *typeName;
    std::vector<ParamInfo *> params;
    std::vector<std::pair<const char *, ImmSpectrum>> specs;   // immediate spectra with defaults
    std::vector<std::pair<const char *, float>> floats;       // immediate floats with defaults
    const char *tupleNode;                                    // node-only float3 plug name (or null)
};

ImmSpectrum imm(float a, float b, float c) {
    ImmSpectrum s;
    s.cs = ColorSpace::Rec709_D65;
    s.e[0] = a; s.e[1] = b; s.e[2] = c;
    return s;
}

const float kPiHost = 3.14159265358979323846f;

const MatSpec &matSpec(MatKind k) {
    static MatSpec table[13] = {
        {"MatteSurfaceMaterial", mk({{"albedo", BOTH, "Spectrum"}}), {{"albedo", imm(0.18f, 0.18f, 0.18f)}}, {}, nullptr},
        {"SpecularReflectionSurfaceMaterial",
         mk({{"coeff", BOTH, "Spectrum"}, {"eta", BOTH, "Spectrum"}, {"k", BOTH, "Spectrum"}}),
         {{"coeff", imm(0.8f, 0.8f, 0.8f)}, {"eta", imm(1, 1, 1)}, {"k", imm(0, 0, 0)}}, {}, nullptr},
        {"SpecularScatteringSurfaceMaterial",
         mk({{"coeff", BOTH, "Spectrum"}, {"eta ext", BOTH, "Spectrum"}, {"eta int", BOTH, "Spectrum"}}),
         {{"coeff", imm(0.8f, 0.8f, 0.8f)}, {"eta ext", imm(1, 1, 1)}, {"eta int", imm(1.5f, 1.5f, 1.5f)}}, {}, nullptr},
        {"MicrofacetReflectionSurfaceMaterial",
         mk({{"eta", BOTH, "Spectrum"}, {"k", BOTH, "Spectrum"}, {"roughness/anisotropy/rotation", NODE, "Float", 3},
             {"roughness", IMM, "Float"}, {"anisotropy", IMM, "Float"}, {"rotation", IMM, "Float"}}),
         {{"eta", imm(1, 1, 1)}, {"k", imm(0, 0, 0)}},
         {{"roughness", 0.1f}, {"anisotropy", 0.0f}, {"rotation", 0.0f}}, "roughness/anisotropy/rotation"},
        {"MicrofacetScatteringSurfaceMaterial",
         mk({{"coeff", BOTH, "Spectrum"}, {"eta ext", BOTH, "Spectrum"}, {"eta Int", BOTH, "Spectrum"},
             {"roughness/anisotropy/rotation", NODE, "Float", 3}, {"roughness", IMM, "Float"},
             {"anisotropy", IMM, "Float"}, {"rotation", IMM, "Float"}}),
         {{"coeff", imm(0.8f, 0.8f, 0.8f)}, {"eta ext", imm(1, 1, 1)}, {"eta Int", imm(1.5f, 1.5f, 1.5f)}},
         {{"roughness", 0.1f}, {"anisotropy", 0.0f}, {"rotation", 0.0f}}, "roughness/anisotropy/rotation"},
        {"LambertianScatteringSurfaceMaterial", mk({{"coeff", BOTH, "Spectrum"}, {"f0", BOTH, "Float"}}),
         {{"coeff", imm(0.8f, 0.8f, 0.8f)}}, {{"f0", 0.04f}}, nullptr},
        {"UE4SurfaceMaterial",
         mk({{"base color", BOTH, "Spectrum"}, {"occlusion/roughness/metallic", NODE, "Float", 3},
             {"occlusion", IMM, "Float"}, {"roughness", IMM, "Float"}, {"metallic", IMM, "Float"}}),
         {{"base color", imm(0.18f, 0.18f, 0.18f)}}, {{"occlusion", 0.0f}, {"roughness", 0.1f}, {"metallic", 0.0f}},
         "occlusion/roughness/metallic"},
        {"OldStyleSurfaceMaterial",
         mk({{"diffuse", BOTH, "Spectrum"}, {"specular", BOTH, "Spectrum"}, {"glossiness", BOTH, "Float"}}),
         {{"diffuse", imm(0.18f, 0.18f, 0.18f)}, {"specular", imm(0.04f, 0.04f, 0.04f)}}, {{"glossiness", 0.6f}}, nullptr},
        {"DiffuseEmitterSurfaceMaterial", mk({{"emittance", BOTH, "Spectrum"}, {"scale", IMM, "Float"}}),
         {{"emittance", imm(kPiHost, kPiHost, kPiHost)}}, {{"scale", 1.0f}}, nullptr},
        {"DirectionalEmitterSurfaceMaterial",
         mk({{"emittance", BOTH, "Spectrum"}, {"scale", IMM, "Float"}, {"direction", BOTH, "Vector3D"}}),
         {{"emittance", imm(kPiHost, kPiHost, kPiHost)}}, {{"scale", 1.0f}}, nullptr},
        {"PointEmitterSurfaceMaterial", mk({{"intensity", BOTH, "Spectrum"}, {"scale", IMM, "Float"}}),
         {{"intensity", imm(kPiHost, kPiHost, kPiHost)}}, {{"scale", 1.0f}}, nullptr},
        {"MultiSurfaceMaterial",
         mk({{"0", NODE, "SurfaceMaterial"}, {"1", NODE, "SurfaceMaterial"}, {"2", NODE, "SurfaceMaterial"},
             {"3", NODE, "SurfaceMaterial"}}), {}, {}, nullptr},
        {"EnvironmentEmitterSurfaceMaterial", mk({{"emittance", BOTH, "Spectrum"}, {"scale", IMM, "Float"}}),
         {{"emittance", imm(kPiHost, kPiHost, kPiHost)}}, {{"scale", 1.0f}}, nullptr},
    };
    return table[static_cast<uint32_t>(k)];
}

const char *kNodeTypeNames[] = {
    "GeometryShaderNode", "TangentShaderNode", "Float2ShaderNode", "Float3ShaderNode", "Float4ShaderNode",
    "ScaleAndOffsetFloatShaderNode", "TripletSpectrumShaderNode", "RegularSampledSpectrumShaderNode",
    "IrregularSampledSpectrumShaderNode", "Float3ToSpectrumShaderNode", "ScaleAndOffsetUVTextureMap2DShaderNode",
    "Image2DTextureShaderNode", "EnvironmentTextureShaderNode"};

const std::vector<ParamInfo *> &nodeParams(NodeKind k) {
    static std::vector<ParamInfo *> tables[13] = {
        mk({}),
        mk({{"tangent type", IMM, "TangentType"}}),
        mk({{"0", BOTH, "Float"}, {"1", BOTH, "Float"}}),
        mk({{"0", BOTH, "Float"}, {"1", BOTH, "Float"}, {"2", BOTH, "Float"}}),
        mk({{"0", BOTH, "Float"}, {"1", BOTH, "Float"}, {"2", BOTH, "Float"}, {"3", BOTH, "Float"}}),
        mk({{"value", NODE, "Float"}, {"scale", BOTH, "Float"}, {"offset", BOTH, "Float"}}),
        mk({{"spectrum type", IMM, "SpectrumType"}, {"color space", IMM, "ColorSpace"}, {"triplet", IMM, "Float", 3}}),
        mk({{"spectrum type", IMM, "SpectrumType"}, {"min wavelength", IMM, "Float"}, {"max wavelength", IMM, "Float"},
            {"values", IMM, "Float", 0}}),
        mk({{"spectrum type", IMM, "SpectrumType"}, {"wavelengths", IMM, "Float", 0}, {"values", IMM, "Float", 0}}),
        mk({{"spectrum type", IMM, "SpectrumType"}, {"color space", IMM, "ColorSpace"}, {"value", BOTH, "Float", 3}}),
        mk({{"scale", IMM, "Float", 2}, {"offset", IMM, "Float", 2}}),
        mk({{"image", NODE, "Image"}, {"bump type", IMM, "BumpType"}, {"bump coeff", IMM, "Float"},
            {"min filter", IMM, "TextureFilter"}, {"mag filter", IMM, "TextureFilter"},
            {"wrap u", IMM, "TextureWrapMode"}, {"wrap v", IMM, "TextureWrapMode"},
            {"texcoord", NODE, "TextureCoordinates"}}),
        mk({{"image", NODE, "Image"}, {"min filter", IMM, "TextureFilter"}, {"mag filter", IMM, "TextureFilter"}}),
    };
    return tables[static_cast<uint32_t>(k)];
}
} // namespace

// ---------------------------------------------------------------------------------------------
// images
const std::vector<ParamInfo *> &Image2D::params() const {
    static std::vector<ParamInfo *> none;
    return none;
}

static uint32_t formatStride(const std::string &f) {
    if (f == "RGB8x3") return 3;
    if (f == "RGB_8x4" || f == "RGBA8x4") return 4;
    if (f == "RGBA16Fx4") return 8;
    if (f == "RGBA32Fx4") return 16;
    if (f == "RG32Fx2") return 8;
    if (f == "Gray32F") return 4;
    if (f == "Gray8") return 1;
    if (f == "GrayA8x2") return 2;
    return 0;
}

uint32_t LinearImage2D::stride() const { return formatStride(formatStr); }

bool LinearImage2D::toRGBA32F(std::vector<float> *out) const {
    size_t n = size_t(width) * height;
    out->assign(n * 4, 0.0f);
    const uint32_t st = stride();
    if (!st || data.size() < n * st) return false;
    for (size_t p = 0; p < n; ++p) {
        const uint8_t *px = data.data() + p * st;
        float *o = out->data() + 4 * p;
        if (formatStr == "RGBA32Fx4") {
            std::memcpy(o, px, 16);
        } else if (formatStr == "RGBA16Fx4") {
            uint16_t h[4];
            std::memcpy(h, px, 8);
            for (int c = 0; c < 4; ++c) o[c] = halfToFloat(h[c]);
        } else if (formatStr == "RGB8x3" || formatStr == "RGB_8x4" || formatStr == "RGBA8x4") {
            for (int c = 0; c < 3; ++c) o[c] = px[c] / 255.0f;
            o[3] = formatStr == "RGBA8x4" ? px[3] / 255.0f : 1.0f;
        } else if (formatStr == "Gray32F") {
            float g;
            std::memcpy(&g, px, 4);
            o[0] = o[1] = o[2] = g;
            o[3] = 1.0f;
        } else if (formatStr == "Gray8") {
            o[0] = o[1] = o[2] = px[0] / 255.0f;
            o[3] = 1.0f;
        } else {
            return false;
        }
    }
    return true;
}

// ---------------------------------------------------------------------------------------------
// shader nodes
ShaderNode::ShaderNode(Context *c, NodeKind k) : Queryable(ObjType::ShaderNode, c), kind(k) {
    switch (k) {
    case NodeKind::Float2: floats["0"] = {0}; floats["1"] = {0}; break;
    case NodeKind::Float3: floats["0"] = {0}; floats["1"] = {0}; floats["2"] = {0}; break;
    case NodeKind::Float4: floats["0"] = {0}; floats["1"] = {0}; floats["2"] = {0}; floats["3"] = {0}; break;
    case NodeKind::ScaleAndOffsetFloat: floats["scale"] = {1}; floats["offset"] = {0}; break;
    case NodeKind::TripletSpectrum:
        enums["spectrum type"] = "Reflectance";
        enums["color space"] = "Rec709(D65)";
        floats["triplet"] = {0.18f, 0.18f, 0.18f};
        break;
    case NodeKind::Float3ToSpectrum:
        enums["spectrum type"] = "Reflectance";
        enums["color space"] = "Rec709(D65)";
        floats["value"] = {0, 0, 0};
        break;
    case NodeKind::ScaleAndOffsetUVTextureMap2D: floats["scale"] = {1, 1}; floats["offset"] = {0, 0}; break;
    case NodeKind::Image2DTexture:
        enums["bump type"] = "Normal Map (DirectX)";
        floats["bump coeff"] = {1};
        enums["min filter"] = enums["mag filter"] = "Linear";
        enums["wrap u"] = enums["wrap v"] = "Repeat";
        break;
    case NodeKind::EnvironmentTexture: enums["min filter"] = enums["mag filter"] = "Linear"; break;
    case NodeKind::Tangent: enums["tangent type"] = "TC0 Direction"; break;
    default: break;
    }
}

const char *ShaderNode::typeName() const { return kNodeTypeNames[static_cast<uint32_t>(kind)]; }
const std::vector<ParamInfo *> &ShaderNode::params() const { return nodeParams(kind); }

static const ParamInfo *findParam(const std::vector<ParamInfo *> &ps, const char *name) {
    for (auto *p : ps)
        if (paramNameEquals(p->name, name)) return p;
    return nullptr;
}

bool ShaderNode::providesPlug(uint32_t t) const {
    switch (kind) {
    case NodeKind::Geometry: return t == VLRShaderNodePlugType_Point3D || t == VLRShaderNodePlugType_Normal3D ||
                                    t == VLRShaderNodePlugType_Vector3D || t == VLRShaderNodePlugType_TextureCoordinates;
    case NodeKind::Tangent: return t == VLRShaderNodePlugType_Vector3D;
    case NodeKind::Float2: return t == VLRShaderNodePlugType_float1 || t == VLRShaderNodePlugType_float2;
    case NodeKind::Float3: return t <= VLRShaderNodePlugType_float3;
    case NodeKind::Float4: return t <= VLRShaderNodePlugType_float4;
    case NodeKind::ScaleAndOffsetFloat: return t == VLRShaderNodePlugType_float1;
    case NodeKind::TripletSpectrum:
    case NodeKind::RegularSampledSpectrum:
    case NodeKind::IrregularSampledSpectrum:
    case NodeKind::Float3ToSpectrum:
    case NodeKind::EnvironmentTexture: return t == VLRShaderNodePlugType_Spectrum;
    case NodeKind::ScaleAndOffsetUVTextureMap2D: return t == VLRShaderNodePlugType_TextureCoordinates;
    case NodeKind::Image2DTexture: return true;
    }
    return false;
}

bool ShaderNode::getEnum(const char *n, const char **v) const {
    for (auto &kv : enums)
        if (paramNameEquals(kv.first, n)) { *v = kv.second.c_str(); return true; }
    return false;
}
bool ShaderNode::setEnum(const char *n, const char *v) {
    const ParamInfo *p = findParam(params(), n);
    if (!p || !(p->formFlags & IMM) || p->typeStr == "Float") return false;
    enums[p->name] = v;
    ctx->touch();
    return true;
}
bool ShaderNode::getFloatTuple(const char *n, float *v, uint32_t len) const {
    const ParamInfo *p = findParam(params(), n);
    if (!p) return false;
    auto it = floats.find(p->name);
    if (it == floats.end() || it->second.size() != len) return false;
    std::copy(it->second.begin(), it->second.end(), v);
    return true;
}
bool ShaderNode::setFloatTuple(const char *n, const float *v, uint32_t len) {
    const ParamInfo *p = findParam(params(), n);
    if (!p || p->typeStr != "Float" || !(p->formFlags & IMM)) return false;
    if (p->tupleSize != 0 && p->tupleSize != len) return false;
    floats[p->name].assign(v, v + len);
    ctx->touch();
    return true;
}
bool ShaderNode::getFloatArray(const char *n, const float **v, uint32_t *len) const {
    const ParamInfo *p = findParam(params(), n);
    if (!p) return false;
    auto it = floats.find(p->name);
    if (it == floats.end()) return false;
    *v = it->second.data();
    *len = static_cast<uint32_t>(it->second.size());
    return true;
}
bool ShaderNode::getImage(const char *n, const Image2D **img) const {
    if (!paramNameEquals(n, "image") || (kind != NodeKind::Image2DTexture && kind != NodeKind::EnvironmentTexture))
        return false;
    *img = image;
    return true;
}
bool ShaderNode::setImage(const char *n, const Image2D *img) {
    if (!paramNameEquals(n, "image") || (kind != NodeKind::Image2DTexture && kind != NodeKind::EnvironmentTexture))
        return false;
    image = img;
    ctx->touch();
    return true;
}
bool ShaderNode::getPlug(const char *n, PlugRef *p) const {
    const ParamInfo *pi = findParam(params(), n);
    if (!pi || !(pi->formFlags & NODE)) return false;
    auto it = plugs.find(pi->name);
    *p = it == plugs.end() ? PlugRef() : it->second;
    return true;
}
bool plugConvertible(SocketType socket, uint32_t plugType) {
    // the value type of each plug type: float1 and Alpha are float, TextureCoordinates Point3D
    const bool scalar = plugType == VLRShaderNodePlugType_float1 || plugType == VLRShaderNodePlugType_Alpha;
    const bool triple = plugType == VLRShaderNodePlugType_float3 || plugType == VLRShaderNodePlugType_Point3D ||
                        plugType == VLRShaderNodePlugType_TextureCoordinates ||
                        plugType == VLRShaderNodePlugType_Vector3D || plugType == VLRShaderNodePlugType_Normal3D;
    switch (socket) {
    case SocketType::Float: return scalar;
    case SocketType::Float3: return scalar || triple;
    case SocketType::Point3D:
    case SocketType::Vector3D:
    case SocketType::Normal3D: return triple;
    case SocketType::Spectrum: return plugType == VLRShaderNodePlugType_Spectrum;
    }
    return false;
}

namespace {
// the socket type of a parameter that takes a plug
bool socketOf(const ParamInfo *pi, SocketType *s) {
    if (pi->typeStr == "Spectrum") *s = SocketType::Spectrum;
    else if (pi->typeStr == "Float") *s = pi->tupleSize == 3 ? SocketType::Float3 : SocketType::Float;
    else if (pi->typeStr == "TextureCoordinates") *s = SocketType::Point3D;
    else if (pi->typeStr == "Vector3D") *s = SocketType::Vector3D;
    else return false;
    return true;
}
} // namespace

int ShaderNode::setPlug(const char *n, const PlugRef &p) {
    const ParamInfo *pi = findParam(params(), n);
    if (!pi || !(pi->formFlags & NODE)) return 1;
    SocketType socket;
    if (!socketOf(pi, &socket)) return 1;
    // an inconvertible plug is refused as the reference's set() refuses it (shader_nodes.cpp)
    if (p.valid() && !plugConvertible(socket, p.plugType)) return 1;
    plugs[pi->name] = p;
    ctx->touch();
    return 0;
}

bool ShaderNode::constantFloat(uint32_t option, float *v) const {
    const char *keys[] = {"0", "1", "2", "3"};
    switch (kind) {
    case NodeKind::Float2:
    case NodeKind::Float3:
    case NodeKind::Float4: {
        if (option > 3) return false;
        auto pit = plugs.find(keys[option]);
        if (pit != plugs.end() && pit->second.valid())
            return pit->second.node->constantFloat(pit->second.option, v);
        auto it = floats.find(keys[option]);
        if (it == floats.end() || it->second.empty()) return false;
        *v = it->second[0];
        return true;
    }
    case NodeKind::ScaleAndOffsetFloat: {
        auto pv = plugs.find("value");
        if (pv == plugs.end() || !pv->second.valid()) return false;
        float x;
        if (!pv->second.node->constantFloat(pv->second.option, &x)) return false;
        *v = floats.at("scale")[0] * x + floats.at("offset")[0];
        return true;
    }
    default:
        return false;
    }
}

bool ShaderNode::constantSpectrum(ImmSpectrum *s, SpectrumType *t) const {
    if (kind != NodeKind::TripletSpectrum) return false;
    const auto &tr = floats.at("triplet");
    ColorSpace cs;
    if (!colorSpaceFromString(enums.at("color space").c_str(), &cs)) return false;
    s->cs = cs;
    s->e[0] = tr[0]; s->e[1] = tr[1]; s->e[2] = tr[2];
    spectrumTypeFromString(enums.at("spectrum type").c_str(), t);
    return true;
}

// ---------------------------------------------------------------------------------------------
// surface materials
SurfaceMaterial::SurfaceMaterial(Context *c, MatKind k) : Queryable(ObjType::SurfaceMaterial, c), kind(k) {
    const MatSpec &ms = matSpec(k);
    for (auto &s : ms.specs) {
        specNames.push_back(s.first);
        specs.push_back(s.second);
    }
    for (auto &f : ms.floats) {
        floatNames.push_back(f.first);
        floatVals.push_back(f.second);
    }
}
const char *SurfaceMaterial::typeName() const { return matSpec(kind).typeName; }
const std::vector<ParamInfo *> &SurfaceMaterial::params() const { return matSpec(kind).params; }

bool SurfaceMaterial::supportedOnRenderPath() const {
    switch (kind) {
    case MatKind::Matte: case MatKind::SpecularReflection: case MatKind::SpecularScattering:
    case MatKind::MicrofacetReflection: case MatKind::MicrofacetScattering: case MatKind::UE4:
    case MatKind::DiffuseEmitter:
    case MatKind::PointEmitter: case MatKind::EnvironmentEmitter:
        return true;
    default:
        return false;
    }
}

bool SurfaceMaterial::getSpectrum(const char *n, ImmSpectrum *s) const {
    for (size_t i = 0; i < specNames.size(); ++i)
        if (paramNameEquals(specNames[i], n)) { *s = specs[i]; return true; }
    return false;
}
bool SurfaceMaterial::setSpectrum(const char *n, const ImmSpectrum &s) {
    for (size_t i = 0; i < specNames.size(); ++i)
        if (paramNameEquals(specNames[i], n)) { specs[i] = s; ctx->touch(); return true; }
    return false;
}
bool SurfaceMaterial::getFloatTuple(const char *n, float *v, uint32_t len) const {
    if (len != 1) return false;
    for (size_t i = 0; i < floatNames.size(); ++i)
        if (paramNameEquals(floatNames[i], n)) { *v = floatVals[i]; return true; }
    return false;
}
bool SurfaceMaterial::setFloatTuple(const char *n, const float *v, uint32_t len) {
    if (len != 1) return false;
    for (size_t i = 0; i < floatNames.size(); ++i)
        if (paramNameEquals(floatNames[i], n)) { floatVals[i] = v[0]; ctx->touch(); return true; }
    return false;
}
bool SurfaceMaterial::getVector(const char *n, float *v) const {
    if (kind != MatKind::DirectionalEmitter || !paramNameEquals(n, "direction")) return false;
    std::memcpy(v, direction, 12);
    return true;
}
bool SurfaceMaterial::setVector(const char *n, const float *v) {
    if (kind != MatKind::DirectionalEmitter || !paramNameEquals(n, "direction")) return false;
    std::memcpy(direction, v, 12);
    ctx->touch();
    return true;
}
bool SurfaceMaterial::getMaterial(const char *n, const SurfaceMaterial **m) const {
    if (kind != MatKind::Multi) return false;
    for (int i = 0; i < 4; ++i)
        if (paramNameEquals(std::to_string(i), n)) { *m = sub[i]; return true; }
    return false;
}
bool SurfaceMaterial::setMaterial(const char *n, const SurfaceMaterial *m) {
    if (kind != MatKind::Multi) return false;
    for (int i = 0; i < 4; ++i)
        if (paramNameEquals(std::to_string(i), n)) { sub[i] = m; ctx->touch(); return true; }
    return false;
}
bool SurfaceMaterial::getPlug(const char *n, PlugRef *p) const {
    const ParamInfo *pi = findParam(params(), n);
    if (!pi || !(pi->formFlags & NODE)) return false;
    auto it = plugs.find(pi->name);
    *p = it == plugs.end() ? PlugRef() : it->second;
    return true;
}
int SurfaceMaterial::setPlug(const char *n, const PlugRef &p) {
    const ParamInfo *pi = findParam(params(), n);
    if (!pi || !(pi->formFlags & NODE)) return 1;
    SocketType socket;
    if (!socketOf(pi, &socket)) return 1;
    // materials.cpp's set(): a plug whose type does not convert to the socket's is refused
    if (p.valid() && !plugConvertible(socket, p.plugType)) return 1;
    plugs[pi->name] = p;
    ctx->touch();
    return 0;
}

DevMaterial SurfaceMaterial::toDevice(bool spectral, bool *usesEnvTexture, const Image2D **envImage,
                                      std::vector<std::pair<PlugRef, uint32_t>> *roots) const {
    DevMaterial d;
    std::memset(&d, 0, sizeof(d));
    *usesEnvTexture = false;
    *envImage = nullptr;
    // A spectrum slot connected to a node: a TripletSpectrum node folds into the immediate (its
    // own spectrum type decides the colour conversion); any other node is evaluated per hit
    // by the material's node program, with the immediate as its default.  Only a Spectrum plug
    // converts to a spectrum (shared.h:430); another plug type leaves the immediate.
    auto spectrumOf = [&](size_t slot, SpectrumType type) {
        ImmSpectrum s = specs[slot];
        auto it = plugs.find(specNames[slot]);
        if (it != plugs.end() && it->second.valid()) {
            SpectrumType t;
            if (it->second.node->constantSpectrum(&s, &t)) return makeDevSpectrum(t, s, spectral);
            // the environment emitter reads its environment texture with its own lookup, not
            // through a node program
            if (roots && it->second.plugType == VLRShaderNodePlugType_Spectrum &&
                !(kind == MatKind::EnvironmentEmitter && it->second.node->kind == NodeKind::EnvironmentTexture))
                roots->push_back({it->second, 1u << slot});
        }
        return makeDevSpectrum(type, s, spectral);
    };
    auto floatOf = [&](const char *name) {
        for (size_t i = 0; i < floatNames.size(); ++i)
            if (paramNameEquals(floatNames[i], name)) return floatVals[i];
        return 0.0f;
    };
    // The float3 parameter: folded when its node is constant, else node-program output 8 with
    // the immediates in raw[0..2] as its default.
    auto tupleNode = [&](const char *plugName, const char *n0, const char *n1, const char *n2) {
        d.raw[0] = floatOf(n0);
        d.raw[1] = floatOf(n1);
        d.raw[2] = floatOf(n2);
        auto it = plugs.find(plugName);
        if (it == plugs.end() || !it->second.valid()) return;
        float v[3];
        bool constant = true;
        for (int k = 0; k < 3; ++k) constant = constant && it->second.node->constantFloat(static_cast<uint32_t>(k), &v[k]);
        if (constant) {
            for (int k = 0; k < 3; ++k) d.raw[k] = v[k];
        } else if (roots) {
            roots->push_back({it->second, 8u});
        }
    };
    auto tupleOf = [&](const char *, int component, float) { return d.raw[component]; };
    switch (kind) {
    case MatKind::Matte:
        d.type = MAT_MATTE;
        d.spec[0] = spectrumOf(0, SpectrumType::Reflectance);
        break;
    case MatKind::SpecularReflection:
        d.type = MAT_SPECULAR_REFLECTION;
        d.spec[0] = spectrumOf(0, SpectrumType::Reflectance);
        d.spec[1] = spectrumOf(1, SpectrumType::IndexOfRefraction);
        d.spec[2] = spectrumOf(2, SpectrumType::IndexOfRefraction);
        break;
    case MatKind::SpecularScattering:
        d.type = MAT_SPECULAR_SCATTERING;
        d.spec[0] = spectrumOf(0, SpectrumType::Reflectance);
        d.spec[1] = spectrumOf(1, SpectrumType::IndexOfRefraction);
        d.spec[2] = spectrumOf(2, SpectrumType::IndexOfRefraction);
        break;
    case MatKind::MicrofacetReflection: {
        // materials.cu:1226-1236
        d.type = MAT_MICROFACET_REFLECTION;
        d.spec[0] = spectrumOf(0, SpectrumType::IndexOfRefraction);
        d.spec[1] = spectrumOf(1, SpectrumType::IndexOfRefraction);
        tupleNode("roughness/anisotropy/rotation", "roughness", "anisotropy", "rotation");
        float r = tupleOf("roughness/anisotropy/rotation", 0, floatOf("roughness"));
        float an = tupleOf("roughness/anisotropy/rotation", 1, floatOf("anisotropy"));
        float rot = tupleOf("roughness/anisotropy/rotation", 2, floatOf("rotation"));
        float alpha = r * r;
        float aspect = std::sqrt(1.0f - 0.9f * an);
        d.scalars[0] = std::fmax(0.001f, alpha / aspect);
        d.scalars[1] = std::fmax(0.001f, alpha * aspect);
        d.scalars[2] = 2 * kPiHost * rot;
        break;
    }
    case MatKind::MicrofacetScattering: {
        // materials.cu:1516-1536
        d.type = MAT_MICROFACET_SCATTERING;
        d.spec[0] = spectrumOf(0, SpectrumType::Reflectance);
        d.spec[1] = spectrumOf(1, SpectrumType::IndexOfRefraction);
        d.spec[2] = spectrumOf(2, SpectrumType::IndexOfRefraction);
        tupleNode("roughness/anisotropy/rotation", "roughness", "anisotropy", "rotation");
        float r = tupleOf("roughness/anisotropy/rotation", 0, floatOf("roughness"));
        float an = tupleOf("roughness/anisotropy/rotation", 1, floatOf("anisotropy"));
        float rot = tupleOf("roughness/anisotropy/rotation", 2, floatOf("rotation"));
        float alpha = r * r;
        float aspect = std::sqrt(1.0f - 0.9f * an);
        d.scalars[0] = std::fmax(0.001f, alpha / aspect);
        d.scalars[1] = std::fmax(0.001f, alpha * aspect);
        d.scalars[2] = 2 * kPiHost * rot;
        break;
    }
    case MatKind::UE4:
        d.type = MAT_UE4;
        d.spec[0] = spectrumOf(0, SpectrumType::Reflectance);
        tupleNode("occlusion/roughness/metallic", "occlusion", "roughness", "metallic");
        d.scalars[0] = tupleOf("occlusion/roughness/metallic", 0, floatOf("occlusion"));
        d.scalars[1] = tupleOf("occlusion/roughness/metallic", 1, floatOf("roughness"));
        d.scalars[2] = tupleOf("occlusion/roughness/metallic", 2, floatOf("metallic"));
        break;
    case MatKind::DiffuseEmitter:
        d.type = MAT_DIFFUSE_EMITTER;
        d.isEmitter = 1;
        d.spec[0] = spectrumOf(0, SpectrumType::LightSource);
        d.scalars[0] = floatOf("scale");
        break;
    case MatKind::PointEmitter:
        d.type = MAT_POINT_EMITTER;
        d.isEmitter = 1;
        d.spec[0] = spectrumOf(0, SpectrumType::LightSource);
        d.scalars[0] = floatOf("scale");
        break;
    case MatKind::EnvironmentEmitter: {
        d.type = MAT_ENVIRONMENT_EMITTER;
        d.isEmitter = 1;
        d.spec[0] = spectrumOf(0, SpectrumType::LightSource);
        d.scalars[0] = floatOf("scale");
        auto it = plugs.find("emittance");
        if (it != plugs.end() && it->second.valid() && it->second.node->kind == NodeKind::EnvironmentTexture &&
            it->second.node->image) {
            d.useEnvTexture = 1;
            *usesEnvTexture = true;
            *envImage = it->second.node->image;
        }
        break;
    }
    default:
        d.type = MAT_NULL;
        break;
    }
    return d;
}

// Radiance-cache material inputs.  These only steer the network (no parity with a reference
// quantity): the RGB triple of the colour parameter as given, split into a diffuse and a
// specular part by material kind, and the roughness.
void SurfaceMaterial::cacheFeatures(float out[8]) const {
    auto rgbOf = [&](size_t slot, float rgb[3]) {
        rgb[0] = rgb[1] = rgb[2] = 0.0f;
        if (slot >= specs.size()) return;
        ImmSpectrum s = specs[slot];
        auto it = plugs.find(specNames[slot]);
        if (it != plugs.end() && it->second.valid()) {
            SpectrumType t;
            it->second.node->constantSpectrum(&s, &t);
        }
        for (int k = 0; k < 3; ++k) rgb[k] = std::fmin(std::fmax(s.e[k], 0.0f), 1.0f);
    };
    auto floatOf = [&](const char *name, float fallback) {
        for (size_t i = 0; i < floatNames.size(); ++i)
            if (paramNameEquals(floatNames[i], name)) return floatVals[i];
        return fallback;
    };
    float d[3] = {0, 0, 0}, sp[3] = {0, 0, 0}, rough = 1.0f;
    switch (kind) {
    case MatKind::Matte:
        rgbOf(0, d);
        break;
    case MatKind::SpecularReflection:
    case MatKind::SpecularScattering:
        rgbOf(0, sp);
        rough = 0.0f;
        break;
    case MatKind::MicrofacetReflection:
        sp[0] = sp[1] = sp[2] = 1.0f;
        rough = floatOf("roughness", 0.5f);
        break;
    case MatKind::MicrofacetScattering:
        rgbOf(0, sp);
        rough = floatOf("roughness", 0.5f);
        break;
    case MatKind::UE4: {
        float base[3];
        rgbOf(0, base);
        const float metallic = floatOf("metallic", 0.0f);
        rough = floatOf("roughness", 0.5f);
        for (int k = 0; k < 3; ++k) {
            d[k] = base[k] * (1.0f - metallic);
            sp[k] = 0.04f + (base[k] - 0.04f) * metallic;
        }
        break;
    }
    default:
        rough = 0.0f;
        break;
    }
    out[0] = d[0];
    out[1] = d[1];
    out[2] = d[2];
    out[3] = rough;
    out[4] = sp[0];
    out[5] = sp[1];
    out[6] = sp[2];
    out[7] = 0.0f;
}

// ---------------------------------------------------------------------------------------------
// transforms: the inverse of an affine matrix is computed in double from the float matrix by
// cofactors and rounded once (affineInverse), the definition shared with the oracle so that
// object-space rays agree bit for bit.
void affineInverse(const float m[16], float inv[12]) {
    double M[3][3], T[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) M[r][c] = m[c * 4 + r];
        T[r] = m[12 + r];
    }
    double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1])
               - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
               + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    double id = 1.0 / det;
    double I[3][3];
    I[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) * id;
    I[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * id;
    I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * id;
    I[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) * id;
    I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * id;
    I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * id;
    I[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) * id;
    I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * id;
    I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * id;
    for (int r = 0; r < 3; ++r) {
        double tr = -(I[r][0] * T[0] + I[r][1] * T[1] + I[r][2] * T[2]);
        for (int c = 0; c < 3; ++c) inv[r * 4 + c] = static_cast<float>(I[r][c]);
        inv[r * 4 + 3] = static_cast<float>(tr);
    }
}

StaticTransform::StaticTransform(Context *c, const float m[16]) : Object(ObjType::StaticTransform, c) {
    std::memcpy(mat, m, sizeof(mat));
    float inv[12];
    affineInverse(m, inv);
    for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 4; ++col) invMat[col * 4 + r] = inv[r * 4 + col];
    invMat[3] = invMat[7] = invMat[11] = 0.0f;
    invMat[15] = 1.0f;
}

// ---------------------------------------------------------------------------------------------
// cameras
const std::vector<ParamInfo *> &Camera::params() const {
    static std::vector<ParamInfo *> persp = mk({{"position", IMM, "Point3D"}, {"orientation", IMM, "Quaternion"},
                                                {"aspect", IMM, "Float"}, {"sensitivity", IMM, "Float"},
                                                {"fovy", IMM, "Float"}, {"lens radius", IMM, "Float"},
                                                {"op distance", IMM, "Float"}});
    static std::vector<ParamInfo *> equi = mk({{"position", IMM, "Point3D"}, {"orientation", IMM, "Quaternion"},
                                               {"sensitivity", IMM, "Float"}, {"h angle", IMM, "Float"},
                                               {"v angle", IMM, "Float"}});
    return perspective ? persp : equi;
}
bool Camera::getPoint(const char *n, float *v) const {
    if (!paramNameEquals(n, "position")) return false;
    std::memcpy(v, position, 12);
    return true;
}
bool Camera::setPoint(const char *n, const float *v) {
    if (!paramNameEquals(n, "position")) return false;
    std::memcpy(position, v, 12);
    return true;
}
bool Camera::getQuaternion(const char *n, float *v) const {
    if (!paramNameEquals(n, "orientation")) return false;
    std::memcpy(v, orientation, 16);
    return true;
}
bool Camera::setQuaternion(const char *n, const float *v) {
    if (!paramNameEquals(n, "orientation")) return false;
    std::memcpy(orientation, v, 16);
    return true;
}
bool Camera::getFloatTuple(const char *n, float *v, uint32_t len) const {
    if (len != 1) return false;
    if (paramNameEquals(n, "sensitivity")) { *v = sensitivity; return true; }
    if (perspective) {
        if (paramNameEquals(n, "aspect")) { *v = aspect; return true; }
        if (paramNameEquals(n, "fovy")) { *v = fovY; return true; }
        if (paramNameEquals(n, "lens radius")) { *v = lensRadius; return true; }
        if (paramNameEquals(n, "op distance")) { *v = opDistance; return true; }
    } else {
        if (paramNameEquals(n, "h angle")) { *v = phiAngle; return true; }
        if (paramNameEquals(n, "v angle")) { *v = thetaAngle; return true; }
    }
    return false;
}
// value clamps as scene.cpp:1638-1662
bool Camera::setFloatTuple(const char *n, const float *v, uint32_t len) {
    if (len != 1) return false;
    float x = v[0];
    if (paramNameEquals(n, "sensitivity")) { sensitivity = std::max(0.0f, std::isfinite(x) ? x : 1.0f); return true; }
    if (perspective) {
        if (paramNameEquals(n, "aspect")) { aspect = std::max(0.001f, x); return true; }
        if (paramNameEquals(n, "fovy")) { fovY = std::min(std::max(x, 0.0001f), kPiHost * 0.999f); return true; }
        if (paramNameEquals(n, "lens radius")) { lensRadius = std::max(0.0f, x); return true; }
        if (paramNameEquals(n, "op distance")) { opDistance = std::max(0.0f, x); return true; }
    } else {
        if (paramNameEquals(n, "h angle")) { phiAngle = x; return true; }
        if (paramNameEquals(n, "v angle")) { thetaAngle = x; return true; }
    }
    return false;
}

DevCamera Camera::toDevice() const {
    DevCamera c;
    std::memcpy(c.position, position, 12);
    // Quaternion::toMatrix3x3 (basic_types.h:1386-1393), columns
    float x = orientation[0], y = orientation[1], z = orientation[2], w = orientation[3];
    float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, zx = z * x, xw = x * w, yw = y * w, zw = z * w;
    float m[9] = {1 - 2 * (yy + zz), 2 * (xy + zw), 2 * (zx - yw), 2 * (xy - zw), 1 - 2 * (xx + zz), 2 * (yz + xw),
                  2 * (zx + yw), 2 * (yz - xw), 1 - 2 * (xx + yy)};
    std::memcpy(c.rot, m, sizeof(m));
    c.sensitivity = sensitivity;
    c.lensRadius = lensRadius;
    // PerspectiveCamera::setImagePlaneArea (shared.h:897-902)
    c.objPlaneDistance = opDistance;
    c.opHeight = 2.0f * opDistance * std::tan(fovY * 0.5f);
    c.opWidth = c.opHeight * aspect;
    c.imgPlaneDistance = 1.0f;
    float r = c.imgPlaneDistance / c.objPlaneDistance;
    c.imgPlaneArea = c.opWidth * c.opHeight * (r * r);
    return c;
}

} // namespace vlra
