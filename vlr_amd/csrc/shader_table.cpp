// Shader nodes and textures of a scene as device records (the node descriptors and texture
// objects of shader_nodes.cpp, setupNodeDescriptor, and the image conversions of
// image.cpp:410-790).
//
// Texels are prepared once per render mode, as the reference's LinearImage2D stores them:
//   spectral build, spectrum type other than NA, RGB sources: per texel the Meng grid
//       coordinates (u, v) and the brightness s = X + Y + Z of the (degammed) colour, quantised
//       as uvsA8x4 (u / 12, v / 14, s / 3 in 8 bits) for 8-bit sources or uvsA16Fx4 (halves)
//       for half and float sources;
//   otherwise linear RGBA: 8-bit sRGB sources are degammed per texel before filtering (the
//       hardware sRGB read of the reference), half / float sRGB sources at creation in the RGB
//       build; gray formats fill x (and y = alpha for GrayA8x2).
// A one-channel texel reads (v, 0, 0, 1) and a two-channel one (r, g, 0, 1).
#include "runtime.h"

#include <cmath>
#include <cstring>
#include <functional>
#include <tuple>

namespace vlra {

namespace {
uint32_t enumIndex(const std::map<std::string, std::string> &e, const char *key, std::initializer_list<const char *> names,
                   uint32_t dflt) {
    auto it = e.find(key);
    if (it == e.end()) return dflt;
    uint32_t k = 0;
    for (const char *n : names) {
        if (it->second == n) return k;
        ++k;
    }
    return dflt;
}

float floatAt(const std::map<std::string, std::vector<float>> &f, const char *key, size_t i, float dflt) {
    auto it = f.find(key);
    if (it == f.end() || it->second.size() <= i) return dflt;
    return it->second[i];
}
} // namespace

std::pair<uint32_t, uint32_t> ShaderTable::program(const std::vector<std::pair<PlugRef, uint32_t>> &roots) {
    const uint32_t base = static_cast<uint32_t>(steps.size());
    std::map<std::tuple<const ShaderNode *, uint32_t, uint32_t>, uint32_t> memo;   // -> slot + 1
    // Post-order: a node's inputs come first; an input that does not fit in the program's
    // kNodeProgramSlots slots counts as not connected (its node's immediate is used).
    std::function<uint32_t(const PlugRef &)> emit = [&](const PlugRef &p) -> uint32_t {
        if (!p.valid()) return 0;
        const auto key = std::make_tuple(p.node, p.plugType, p.option);
        auto m = memo.find(key);
        if (m != memo.end()) return m->second;
        DevNodeStep st;
        std::memset(&st, 0, sizeof(st));
        st.node = addNode(p.node);
        st.plug = p.plugType | (p.option << 4);
        auto inputPlug = [&](const char *name) {
            auto it = p.node->plugs.find(name);
            return it == p.node->plugs.end() ? PlugRef() : it->second;
        };
        switch (p.node->kind) {
        case NodeKind::Float2:
        case NodeKind::Float3:
        case NodeKind::Float4: {
            const char *keys[] = {"0", "1", "2", "3"};
            const uint32_t width = p.plugType == VLRShaderNodePlugType_float1 ? 1
                                   : p.plugType == VLRShaderNodePlugType_float2 ? 2
                                   : p.plugType == VLRShaderNodePlugType_float3 ? 3 : 4;
            for (uint32_t k = p.option; k < p.option + width && k < 4; ++k) st.in[k] = emit(inputPlug(keys[k]));
            break;
        }
        case NodeKind::ScaleAndOffsetFloat:
            st.in[0] = emit(inputPlug("value"));
            st.in[1] = emit(inputPlug("scale"));
            st.in[2] = emit(inputPlug("offset"));
            break;
        case NodeKind::Float3ToSpectrum:
            st.in[0] = emit(inputPlug("value"));
            break;
        case NodeKind::Image2DTexture:
            st.in[0] = emit(inputPlug("texcoord"));
            break;
        default:
            break;
        }
        if (steps.size() - base >= kNodeProgramSlots) return 0;
        steps.push_back(st);
        const uint32_t slot = static_cast<uint32_t>(steps.size() - base);
        memo[key] = slot;
        return slot;
    };
    for (const auto &r : roots) {
        const uint32_t s = emit(r.first);
        if (s) steps[base + s - 1].out |= r.second;
    }
    return {base, static_cast<uint32_t>(steps.size()) - base};
}

uint32_t ShaderTable::addNode(const ShaderNode *n) {
    auto found = nodeIndex.find(n);
    if (found != nodeIndex.end()) return found->second;
    const uint32_t idx = static_cast<uint32_t>(nodes.size());
    nodeIndex[n] = idx;
    nodes.emplace_back();
    DevShaderNode d;
    std::memset(&d, 0, sizeof(d));
    d.kind = static_cast<uint32_t>(n->kind);
    SpectrumType st = SpectrumType::Reflectance;
    auto se = n->enums.find("spectrum type");
    if (se != n->enums.end()) spectrumTypeFromString(se->second.c_str(), &st);
    ColorSpace cs = ColorSpace::Rec709_D65;
    auto ce = n->enums.find("color space");
    if (ce != n->enums.end()) colorSpaceFromString(ce->second.c_str(), &cs);
    switch (n->kind) {
    case NodeKind::Tangent:
        d.a[0] = enumIndex(n->enums, "tangent type", {"TC0 Direction", "Radial X", "Radial Y", "Radial Z"}, 0);
        break;
    case NodeKind::Float2:
    case NodeKind::Float3:
    case NodeKind::Float4: {
        const int count = n->kind == NodeKind::Float2 ? 2 : (n->kind == NodeKind::Float3 ? 3 : 4);
        const char *keys[] = {"0", "1", "2", "3"};
        for (int k = 0; k < count; ++k) {
            d.f[k] = floatAt(n->floats, keys[k], 0, 0.0f);
        }
        break;
    }
    case NodeKind::ScaleAndOffsetFloat:
        d.f[1] = floatAt(n->floats, "scale", 0, 1.0f);
        d.f[2] = floatAt(n->floats, "offset", 0, 0.0f);
        break;
    case NodeKind::TripletSpectrum: {
        ImmSpectrum is;
        is.cs = cs;
        for (int k = 0; k < 3; ++k) is.e[k] = floatAt(n->floats, "triplet", k, 0.0f);
        const DevSpectrum s = makeDevSpectrum(st, is, spectral);
        d.f[0] = s.x;
        d.f[1] = s.y;
        d.f[2] = s.z;
        d.f[3] = s.w;
        break;
    }
    case NodeKind::RegularSampledSpectrum:
    case NodeKind::IrregularSampledSpectrum: {
        const bool regular = n->kind == NodeKind::RegularSampledSpectrum;
        auto vit = n->floats.find("values");
        std::vector<float> values = vit == n->floats.end() ? std::vector<float>{0.0f} : vit->second;
        if (values.empty()) values.push_back(0.0f);
        std::vector<float> lambdas(values.size());
        float lo = 0.0f, hi = 0.0f;
        if (regular) {
            lo = floatAt(n->floats, "min wavelength", 0, 360.0f);
            hi = floatAt(n->floats, "max wavelength", 0, 830.0f);
            d.f[0] = lo;
            d.f[1] = hi;
        } else {
            auto lit = n->floats.find("wavelengths");
            if (lit != n->floats.end())
                for (size_t i = 0; i < values.size() && i < lit->second.size(); ++i) lambdas[i] = lit->second[i];
        }
        if (spectral) {
            d.a[0] = static_cast<uint32_t>(pool.size());
            d.a[1] = static_cast<uint32_t>(values.size());
            if (!regular) pool.insert(pool.end(), lambdas.begin(), lambdas.end());
            pool.insert(pool.end(), values.begin(), values.end());
        } else {
            float rgb[3];
            sampledSpectrumToRGB(st, regular, lo, hi, lambdas.data(), values.data(), static_cast<uint32_t>(values.size()),
                                 rgb);
            d.f[0] = rgb[0];
            d.f[1] = rgb[1];
            d.f[2] = rgb[2];
            d.f[3] = 0.0f;
            d.a[2] = 1;
        }
        break;
    }
    case NodeKind::Float3ToSpectrum:
        for (int k = 0; k < 3; ++k) d.f[k] = floatAt(n->floats, "value", k, 0.0f);
        d.a[0] = static_cast<uint32_t>(st);
        d.a[1] = static_cast<uint32_t>(cs);
        break;
    case NodeKind::ScaleAndOffsetUVTextureMap2D:
        d.f[0] = floatAt(n->floats, "scale", 0, 1.0f);
        d.f[1] = floatAt(n->floats, "scale", 1, 1.0f);
        d.f[2] = floatAt(n->floats, "offset", 0, 0.0f);
        d.f[3] = floatAt(n->floats, "offset", 1, 0.0f);
        break;
    case NodeKind::Image2DTexture: {
        d.a[0] = addTexture(n->image);
        // the reference's sampler has one filter for both magnification and minification
        const uint32_t filter = enumIndex(n->enums, "mag filter", {"Nearest", "Linear", "None"}, 1) == 1 ? 1u : 0u;
        const uint32_t wu = enumIndex(n->enums, "wrap u", {"Repeat", "Clamp to Edge", "Mirror", "Clamp to Border"}, 0);
        const uint32_t wv = enumIndex(n->enums, "wrap v", {"Repeat", "Clamp to Edge", "Mirror", "Clamp to Border"}, 0);
        const uint32_t bump =
            enumIndex(n->enums, "bump type", {"Normal Map (DirectX)", "Normal Map (OpenGL)", "Height Map"}, 0);
        d.a[1] = filter | (wu << 2) | (wv << 4) | (bump << 8);
        // 5-bit coefficient (shader_nodes.cpp:1685-1688, shared.h:743-746)
        const float coeff = floatAt(n->floats, "bump coeff", 0, 1.0f);
        const uint32_t q = std::min<uint32_t>(static_cast<uint32_t>(std::round(31.0f * coeff * 0.5f)), 31u);
        d.f[0] = static_cast<float>(q + 1) / 16.0f;
        break;
    }
    case NodeKind::EnvironmentTexture:
        // repeat wrap on both axes, filter as set (shader_nodes.cpp:1910-1921)
        d.a[0] = addTexture(n->image, true);
        d.a[1] = enumIndex(n->enums, "mag filter", {"Nearest", "Linear", "None"}, 1) == 1 ? 1u : 0u;
        break;
    default:
        break;
    }
    nodes[idx] = d;
    return idx;
}

uint32_t ShaderTable::addTexture(const Image2D *img, bool envRead) {
    auto found = texIndex.find({img, envRead});
    if (found != texIndex.end()) return found->second;
    const uint32_t idx = static_cast<uint32_t>(textures.size());
    texIndex[{img, envRead}] = idx;
    DevTexture T;
    std::memset(&T, 0, sizeof(T));
    T.offset = static_cast<uint32_t>(texels.size());
    if (envRead && !img) {
        // the node's null image: 1 x 1 RGBA16F (1, 0, 1, 1), light source, Rec.709 D65
        // (shader_nodes.cpp:1880-1881), stored as the spectral build stores such an image
        T.width = T.height = 1;
        T.spectrumType = kSpecLightSource;
        T.colorSpace = kCsRec709;
        float4 t = make_float4(1.0f, 0.0f, 1.0f, 1.0f);
        if (spectral) {
            const float rgb[3] = {1.0f, 0.0f, 1.0f};
            float XYZ[3];
            colorTristim(0, rgb, XYZ);
            const float b = XYZ[0] + XYZ[1] + XYZ[2];
            const float xy[2] = {XYZ[0] / b, XYZ[1] / b};
            float uv[2];
            mengXYtoUV(xy, uv);
            auto h = [](float v) { return halfToFloat(floatToHalf(v)); };
            t = make_float4(h(uv[0]), h(uv[1]), h(b), 1.0f);
            T.format = kTexUVS16;
        } else {
            T.format = kTexRGBA;
        }
        texels.push_back(t);
        textures.push_back(T);
        return idx;
    }
    if (img && img->type == ObjType::BlockCompressedImage2D) {
        // mip level 0 decoded; the texture unit's sRGB read of BC1-BC4 and BC7 (image.cpp:336-346)
        // is a per-texel degamma of the colour channels, done here
        const auto *bi = static_cast<const BlockCompressedImage2D *>(img);
        std::vector<float4> dec;
        if (!bi->mips.empty() &&
            decodeBlockCompressed(bi->formatStr, bi->mips[0].data(), bi->mips[0].size(), bi->width, bi->height, &dec)) {
            const std::string &f = bi->formatStr;
            const bool degamma = !envRead && bi->colorSpace == ColorSpace::Rec709_D65_sRGBGamma &&
                                 (f == "BC1" || f == "BC2" || f == "BC3" || f == "BC4" || f == "BC7");
            T.width = bi->width;
            T.height = bi->height;
            T.format = kTexRGBA;
            T.gray = 0;
            T.spectrumType = static_cast<uint32_t>(bi->spectrumType);
            T.colorSpace = static_cast<uint32_t>(degamma ? ColorSpace::Rec709_D65 : bi->colorSpace);
            for (float4 &t : dec) {
                if (degamma) {
                    t.x = colorDegamma(t.x);
                    if (f != "BC4") {
                        t.y = colorDegamma(t.y);
                        t.z = colorDegamma(t.z);
                    }
                }
                texels.push_back(t);
            }
            textures.push_back(T);
            return idx;
        }
    }
    const LinearImage2D *li =
        img && img->type == ObjType::LinearImage2D ? static_cast<const LinearImage2D *>(img) : nullptr;
    const uint32_t st = li ? li->stride() : 0;
    if (!li || !st || li->data.size() < size_t(li->width) * li->height * st || !li->width || !li->height) {
        // no image (the reference binds a 1 x 1 null image) or a format this path does not read
        T.width = T.height = 1;
        T.spectrumType = kSpecNA;
        T.colorSpace = kCsRec709;
        texels.push_back(make_float4(0, 0, 0, 0));
        textures.push_back(T);
        return idx;
    }
    const std::string &f = li->formatStr;
    const bool rgb8 = f == "RGB8x3" || f == "RGB_8x4" || f == "RGBA8x4";
    const bool half = f == "RGBA16Fx4", full = f == "RGBA32Fx4";
    const bool srgb = li->colorSpace == ColorSpace::Rec709_D65_sRGBGamma;
    const uint32_t stype = static_cast<uint32_t>(li->spectrumType);
    const bool uvs = spectral && stype != kSpecNA && (rgb8 || half || full);
    T.width = li->width;
    T.height = li->height;
    T.format = uvs ? (rgb8 ? kTexUVS8 : kTexUVS16) : kTexRGBA;
    T.gray = (f == "Gray32F" || f == "Gray8" || f == "GrayA8x2") ? 1 : 0;
    T.spectrumType = stype;
    // texels read through the hardware sRGB path count as degammed (shader_nodes.cpp:1678-1683);
    // the environment texture's sampler reads without it (shader_nodes.cpp:1915-1919)
    const bool hwDegamma = !envRead && srgb && !uvs && (rgb8 || f == "Gray8");
    const bool createDegamma = srgb && !uvs && ((!spectral && (half || full)) || f == "RG32Fx2" || f == "Gray32F");
    const bool grayA8Degamma = srgb && f == "GrayA8x2";
    T.colorSpace = static_cast<uint32_t>((hwDegamma || createDegamma) ? ColorSpace::Rec709_D65 : li->colorSpace);
    const size_t n = size_t(li->width) * li->height;
    texels.reserve(texels.size() + n);
    for (size_t p = 0; p < n; ++p) {
        const uint8_t *px = li->data.data() + p * st;
        float c[4] = {0, 0, 0, 1};
        if (rgb8) {
            for (int k = 0; k < 3; ++k) c[k] = px[k] / 255.0f;
            c[3] = f == "RGBA8x4" ? px[3] / 255.0f : 1.0f;
        } else if (half) {
            uint16_t h[4];
            std::memcpy(h, px, 8);
            for (int k = 0; k < 4; ++k) c[k] = halfToFloat(h[k]);
        } else if (full) {
            std::memcpy(c, px, 16);
        } else if (f == "RG32Fx2") {
            std::memcpy(c, px, 8);
            c[2] = 0.0f;
        } else if (f == "Gray32F") {
            std::memcpy(c, px, 4);
        } else if (f == "Gray8") {
            c[0] = px[0] / 255.0f;
        } else if (f == "GrayA8x2") {
            c[0] = px[0] / 255.0f;
            c[1] = px[1] / 255.0f;
        }
        float4 t;
        if (uvs) {
            // image.cpp:510-555 (perPixelFunc to uvsA8x4 / uvsA16Fx4)
            float rgb[3] = {c[0], c[1], c[2]};
            if (srgb)
                for (float &v : rgb) v = colorDegamma(v);
            float XYZ[3];
            const bool refl = stype == kSpecReflectance || stype == kSpecIndexOfRefraction;
            colorTristim(refl ? 2 : 0, rgb, XYZ);
            const float b = XYZ[0] + XYZ[1] + XYZ[2];
            const float xy[2] = {b > 0.0f ? XYZ[0] / b : (1.0f / 3.0f), b > 0.0f ? XYZ[1] / b : (1.0f / 3.0f)};
            float uv[2];
            mengXYtoUV(xy, uv);
            if (rgb8) {
                auto q8 = [](float v) { return static_cast<float>(static_cast<uint8_t>(255 * std::fmin(std::fmax(v, 0.0f), 1.0f))); };
                t = make_float4(q8(uv[0] / 12.0f) / 255.0f, q8(uv[1] / 14.0f) / 255.0f, q8(b / 3.0f) / 255.0f, c[3]);
            } else {
                auto h = [](float v) { return halfToFloat(floatToHalf(v)); };
                t = make_float4(h(uv[0]), h(uv[1]), h(b), half ? c[3] : h(c[3]));
            }
        } else {
            if (hwDegamma) {
                for (int k = 0; k < 3; ++k) c[k] = colorDegamma(c[k]);
                if (f == "Gray8") c[1] = c[2] = 0.0f;
            } else if (createDegamma) {
                const int chans = f == "RG32Fx2" ? 2 : (f == "Gray32F" ? 1 : 3);
                for (int k = 0; k < chans; ++k) {
                    c[k] = colorDegamma(c[k]);
                    if (half) c[k] = halfToFloat(floatToHalf(c[k]));
                }
            } else if (grayA8Degamma) {
                // uint8 degamma at creation (image.cpp:445-450)
                const float v = std::min(255.0f, std::floor(256.0f * colorDegamma(c[0])));
                c[0] = static_cast<float>(static_cast<uint8_t>(v)) / 255.0f;
            }
            t = make_float4(c[0], c[1], c[2], c[3]);
        }
        texels.push_back(t);
    }
    textures.push_back(T);
    return idx;
}

} // namespace vlra
