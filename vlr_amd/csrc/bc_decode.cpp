// Block-compressed images (BlockCompressedImage2D, image.cpp:1059-1068) decoded on the host into
// the float4 texels the device samples.  The reference hands the blocks to the texture unit;
// this build decodes mip level 0 (the level tex2DLod(..., 0) reads) once at scene set-up,
// following the Direct3D block-compression specification:
//   BC1  two RGB565 endpoints, 2-bit indices; c0 > c1: four colours (thirds), otherwise three
//        (half) and transparent black
//   BC2  explicit 4-bit alpha per texel, then a four-colour BC1 block
//   BC3  a BC4-style alpha block, then a four-colour BC1 block
//   BC4  two 8-bit endpoints, 3-bit indices; e0 > e1: eight values (sevenths), otherwise six
//        (fifths) plus 0 and 1; the signed variant on [-1, 1] with -128 read as -127
//   BC5  two BC4 blocks (red, green)
// BC6H and BC7 are not decoded: creating such an image fails.
#include "runtime.h"

#include <cstring>

namespace vlra {

namespace {
void rgb565(uint16_t c, float rgb[3]) {
    rgb[0] = static_cast<float>((c >> 11) & 31) / 31.0f;
    rgb[1] = static_cast<float>((c >> 5) & 63) / 63.0f;
    rgb[2] = static_cast<float>(c & 31) / 31.0f;
}

// colour block (8 bytes) into 16 RGBA texels
void colorBlock(const uint8_t *b, bool allowThreeColor, float4 out[16]) {
    uint16_t c0, c1;
    uint32_t idx;
    std::memcpy(&c0, b, 2);
    std::memcpy(&c1, b + 2, 2);
    std::memcpy(&idx, b + 4, 4);
    float p[4][4];
    rgb565(c0, p[0]);
    rgb565(c1, p[1]);
    p[0][3] = p[1][3] = 1.0f;
    if (c0 > c1 || !allowThreeColor) {
        for (int k = 0; k < 3; ++k) {
            p[2][k] = (2.0f * p[0][k] + p[1][k]) / 3.0f;
            p[3][k] = (p[0][k] + 2.0f * p[1][k]) / 3.0f;
        }
        p[2][3] = p[3][3] = 1.0f;
    } else {
        for (int k = 0; k < 3; ++k) {
            p[2][k] = (p[0][k] + p[1][k]) / 2.0f;
            p[3][k] = 0.0f;
        }
        p[2][3] = 1.0f;
        p[3][3] = 0.0f;
    }
    for (int t = 0; t < 16; ++t) {
        const uint32_t i = (idx >> (2 * t)) & 3u;
        out[t] = make_float4(p[i][0], p[i][1], p[i][2], p[i][3]);
    }
}

// single-channel block (8 bytes) into 16 values
void channelBlock(const uint8_t *b, bool isSigned, float out[16]) {
    float e0, e1;
    bool eightValues;
    if (isSigned) {
        int8_t s0 = static_cast<int8_t>(b[0]), s1 = static_cast<int8_t>(b[1]);
        eightValues = s0 > s1;
        e0 = static_cast<float>(s0 == -128 ? -127 : s0) / 127.0f;
        e1 = static_cast<float>(s1 == -128 ? -127 : s1) / 127.0f;
    } else {
        eightValues = b[0] > b[1];
        e0 = static_cast<float>(b[0]) / 255.0f;
        e1 = static_cast<float>(b[1]) / 255.0f;
    }
    float p[8];
    p[0] = e0;
    p[1] = e1;
    if (eightValues) {
        for (int i = 2; i < 8; ++i) p[i] = (static_cast<float>(8 - i) * e0 + static_cast<float>(i - 1) * e1) / 7.0f;
    } else {
        for (int i = 2; i < 6; ++i) p[i] = (static_cast<float>(6 - i) * e0 + static_cast<float>(i - 1) * e1) / 5.0f;
        p[6] = isSigned ? -1.0f : 0.0f;
        p[7] = 1.0f;
    }
    uint64_t idx = 0;
    for (int k = 0; k < 6; ++k) idx |= static_cast<uint64_t>(b[2 + k]) << (8 * k);
    for (int t = 0; t < 16; ++t) out[t] = p[(idx >> (3 * t)) & 7u];
}
} // namespace

bool bcFormatDecodable(const std::string &f) {
    return f == "BC1" || f == "BC2" || f == "BC3" || f == "BC4" || f == "BC4_Signed" || f == "BC5" ||
           f == "BC5_Signed";
}

size_t bcBlockBytes(const std::string &f) { return (f == "BC1" || f == "BC4" || f == "BC4_Signed") ? 8 : 16; }

bool decodeBlockCompressed(const std::string &f, const uint8_t *data, size_t size, uint32_t width, uint32_t height,
                           std::vector<float4> *texels) {
    if (!bcFormatDecodable(f) || width == 0 || height == 0) return false;
    const uint32_t bw = (width + 3) / 4, bh = (height + 3) / 4;
    const size_t blockBytes = bcBlockBytes(f);
    if (size < size_t(bw) * bh * blockBytes) return false;
    texels->assign(size_t(width) * height, make_float4(0, 0, 0, 1));
    for (uint32_t by = 0; by < bh; ++by) {
        for (uint32_t bx = 0; bx < bw; ++bx) {
            const uint8_t *b = data + (size_t(by) * bw + bx) * blockBytes;
            float4 t[16];
            if (f == "BC1") {
                colorBlock(b, true, t);
            } else if (f == "BC2") {
                colorBlock(b + 8, false, t);
                uint64_t a;
                std::memcpy(&a, b, 8);
                for (int k = 0; k < 16; ++k) t[k].w = static_cast<float>((a >> (4 * k)) & 15u) / 15.0f;
            } else if (f == "BC3") {
                colorBlock(b + 8, false, t);
                float a[16];
                channelBlock(b, false, a);
                for (int k = 0; k < 16; ++k) t[k].w = a[k];
            } else if (f == "BC4" || f == "BC4_Signed") {
                float r[16];
                channelBlock(b, f == "BC4_Signed", r);
                for (int k = 0; k < 16; ++k) t[k] = make_float4(r[k], 0, 0, 1);
            } else {
                float r[16], g[16];
                channelBlock(b, f == "BC5_Signed", r);
                channelBlock(b + 8, f == "BC5_Signed", g);
                for (int k = 0; k < 16; ++k) t[k] = make_float4(r[k], g[k], 0, 1);
            }
            for (int k = 0; k < 16; ++k) {
                const uint32_t x = bx * 4 + (k & 3), y = by * 4 + (k >> 2);
                if (x < width && y < height) (*texels)[size_t(y) * width + x] = t[k];
            }
        }
    }
    return true;
}

} // namespace vlra
