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
//   BC6H 14 modes of one or two regions: endpoints of 6-16 bits (two-region modes and three
//        one-region modes store the other endpoints as deltas from the first), unquantized to
//        16 bits, interpolated with 3- or 4-bit indices, scaled by 31/64 (signed: 31/32) into
//        half-float bits; alpha 1
//   BC7  8 modes of one to three subsets: RGB(A) endpoints of 4-8 bits with optional p-bits,
//        expanded to 8 bits, interpolated with 2-4-bit indices (modes 4 and 5 keep separate
//        colour and alpha indices), optionally with alpha swapped into a colour channel;
//        texels are the 8-bit values / 255
// Reserved modes of BC6H and BC7 decode to zeros, as the specification says.
#include "runtime.h"

#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

namespace vlra {

namespace {
// ---- BC6H / BC7 tables (Direct3D 11 "BC6H Format", "BC7 Format") ----
// subset of texel t: bit t of the 2-subset mask, digit t of the 3-subset row
const uint16_t kPart2[64] = {
    0xCCCC, 0x8888, 0xEEEE, 0xECC8, 0xC880, 0xFEEC, 0xFEC8, 0xEC80, 0xC800, 0xFFEC, 0xFE80, 0xE800, 0xFFE8,
    0xFF00, 0xFFF0, 0xF000, 0xF710, 0x008E, 0x7100, 0x08CE, 0x008C, 0x7310, 0x3100, 0x8CCE, 0x088C, 0x3110,
    0x6666, 0x366C, 0x17E8, 0x0FF0, 0x718E, 0x399C, 0xAAAA, 0xF0F0, 0x5A5A, 0x33CC, 0x3C3C, 0x55AA, 0x9696,
    0xA55A, 0x73CE, 0x13C8, 0x324C, 0x3BDC, 0x6996, 0xC33C, 0x9966, 0x0660, 0x0272, 0x04E4, 0x4E40, 0x2720,
    0xC936, 0x936C, 0x39C6, 0x639C, 0x9336, 0x9CC6, 0x817E, 0xE718, 0xCCF0, 0x0FCC, 0x7744, 0xEE22};
const char *const kPart3[64] = {
    "0011001102212222", "0001001122112221", "0000200122112211", "0222002200110111",
    "0000000011221122", "0011001100220022", "0022002211111111", "0011001122112211",
    "0000000011112222", "0000111111112222", "0000111122222222", "0012001200120012",
    "0112011201120112", "0122012201220122", "0011011211221222", "0011200122002220",
    "0001001101121122", "0111001120012200", "0000112211221122", "0022002200221111",
    "0111011102220222", "0001000122212221", "0000001101220122", "0000110022102210",
    "0122012200110000", "0012001211222222", "0110122112210110", "0000011012211221",
    "0022110211020022", "0110011020022222", "0011012201220011", "0000200022112221",
    "0000000211221222", "0222002200120011", "0011001200220222", "0120012001200120",
    "0000111122220000", "0120120120120120", "0120201212010120", "0011220011220011",
    "0011112222000011", "0101010122222222", "0000000021212121", "0022112200221122",
    "0022001100220011", "0220122102201221", "0101222222220101", "0000212121212121",
    "0101010101012222", "0222011102220111", "0002111200021112", "0000211221122112",
    "0222011101110222", "0002111211120002", "0110011001102222", "0000000021122112",
    "0110011022222222", "0022001100110022", "0022112211220022", "0000000000002112",
    "0002000100020001", "0222122202221222", "0101222222222222", "0111201122012220"};
// anchor texel of subset 1 (2 subsets), of subsets 1 and 2 (3 subsets): its index has one bit fewer
const uint8_t kAnchor2[64] = {15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 2, 8, 2, 2, 8,
                              8,  15, 2,  8,  2,  2,  8,  8,  2,  2,  15, 15, 6,  8,  2,  8,  15, 15, 2, 8, 2, 2,
                              2,  15, 15, 6,  6,  2,  6,  8,  15, 15, 2,  2,  15, 15, 15, 15, 15, 2,  2, 15};
const uint8_t kAnchor3a[64] = {3, 3, 15, 15, 8, 3, 15, 15, 8, 8, 6,  6,  6, 5,  3, 3,  3, 3,  8,  15, 3, 3,
                               6, 10, 5, 8,  8, 6, 8,  5,  15, 15, 8, 15, 3, 5, 6, 10, 8, 15, 15, 3, 15, 5,
                               15, 15, 15, 15, 3, 15, 5, 5, 5, 8, 5, 10, 5, 10, 8, 13, 15, 12, 3, 3};
const uint8_t kAnchor3b[64] = {15, 8, 8, 3,  15, 15, 3,  8,  15, 15, 15, 15, 15, 15, 15, 8,  15, 8,  15, 3, 15, 8,
                               15, 8, 3, 15, 6,  10, 15, 15, 10, 8,  15, 3,  15, 10, 10, 8,  9,  10, 6,  15, 8, 15,
                               3,  6, 6, 8,  15, 3,  15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 3,  15, 15, 8};
const uint8_t kWeights2[4] = {0, 21, 43, 64};
const uint8_t kWeights3[8] = {0, 9, 18, 27, 37, 46, 55, 64};
const uint8_t kWeights4[16] = {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64};
const uint8_t *weightsFor(int bits) { return bits == 2 ? kWeights2 : (bits == 3 ? kWeights3 : kWeights4); }

// the 128 bits of a block, read from bit 0 up
struct BitReader {
    uint64_t lo, hi;
    int pos = 0;
    explicit BitReader(const uint8_t *b) {
        std::memcpy(&lo, b, 8);
        std::memcpy(&hi, b + 8, 8);
    }
    uint32_t get(int n) {
        uint32_t r = 0;
        for (int k = 0; k < n; ++k, ++pos) {
            const uint64_t w = pos < 64 ? lo >> pos : hi >> (pos - 64);
            r |= static_cast<uint32_t>(w & 1u) << k;
        }
        return r;
    }
};

int subsetOf(int ns, int part, int t) {
    if (ns == 1) return 0;
    if (ns == 2) return (kPart2[part] >> t) & 1;
    return kPart3[part][t] - '0';
}

bool isAnchor(int ns, int part, int t) {
    if (t == 0) return true;
    if (ns == 2) return t == kAnchor2[part];
    if (ns == 3) return t == kAnchor3a[part] || t == kAnchor3b[part];
    return false;
}

// BC7: subsets, partition / rotation / index-selection bits, colour and alpha bits, per-endpoint
// and per-subset p-bits, index and secondary index bits
struct Bc7Mode {
    int ns, pb, rb, isb, cb, ab, epb, spb, ib, ib2;
};
const Bc7Mode kBc7Modes[8] = {{3, 4, 0, 0, 4, 0, 1, 0, 3, 0}, {2, 6, 0, 0, 6, 0, 0, 1, 3, 0},
                              {3, 6, 0, 0, 5, 0, 0, 0, 2, 0}, {2, 6, 0, 0, 7, 0, 1, 0, 2, 0},
                              {1, 0, 2, 1, 5, 6, 0, 0, 2, 3}, {1, 0, 2, 0, 7, 8, 0, 0, 2, 2},
                              {1, 0, 0, 0, 7, 7, 1, 0, 4, 0}, {2, 6, 0, 0, 5, 5, 1, 0, 2, 0}};

void bc7Block(const uint8_t *b, float4 out[16]) {
    BitReader br(b);
    int mode = 0;
    while (mode < 8 && br.get(1) == 0) ++mode;
    if (mode == 8) {
        for (int t = 0; t < 16; ++t) out[t] = make_float4(0, 0, 0, 0);
        return;
    }
    const Bc7Mode &M = kBc7Modes[mode];
    const int part = static_cast<int>(br.get(M.pb)), rot = static_cast<int>(br.get(M.rb));
    const int isel = static_cast<int>(br.get(M.isb));
    uint32_t ep[3][2][4] = {};
    for (int c = 0; c < 3; ++c)
        for (int s = 0; s < M.ns; ++s)
            for (int e = 0; e < 2; ++e) ep[s][e][c] = br.get(M.cb);
    if (M.ab)
        for (int s = 0; s < M.ns; ++s)
            for (int e = 0; e < 2; ++e) ep[s][e][3] = br.get(M.ab);
    int cprec = M.cb, aprec = M.ab;
    if (M.epb || M.spb) {
        uint32_t p[3][2];
        for (int s = 0; s < M.ns; ++s) {
            p[s][0] = br.get(1);
            p[s][1] = M.epb ? br.get(1) : p[s][0];
        }
        for (int s = 0; s < M.ns; ++s)
            for (int e = 0; e < 2; ++e)
                for (int c = 0; c < 4; ++c) ep[s][e][c] = (ep[s][e][c] << 1) | p[s][e];
        ++cprec;
        if (M.ab) ++aprec;
    }
    auto expand = [](uint32_t v, int prec) {   // to 8 bits, the top bits repeated below
        v <<= 8 - prec;
        return v | (v >> prec);
    };
    for (int s = 0; s < M.ns; ++s)
        for (int e = 0; e < 2; ++e) {
            for (int c = 0; c < 3; ++c) ep[s][e][c] = expand(ep[s][e][c], cprec);
            ep[s][e][3] = M.ab ? expand(ep[s][e][3], aprec) : 255u;
        }
    uint32_t idx1[16], idx2[16];
    for (int t = 0; t < 16; ++t) idx1[t] = br.get(isAnchor(M.ns, part, t) ? M.ib - 1 : M.ib);
    if (M.ib2)
        for (int t = 0; t < 16; ++t) idx2[t] = br.get(t == 0 ? M.ib2 - 1 : M.ib2);
    for (int t = 0; t < 16; ++t) {
        const int s = subsetOf(M.ns, part, t);
        uint32_t ci = idx1[t], ai = idx1[t];
        int cbits = M.ib, abits = M.ib;
        if (M.ib2) {
            if (isel) {
                ci = idx2[t];
                cbits = M.ib2;
                ai = idx1[t];
                abits = M.ib;
            } else {
                ai = idx2[t];
                abits = M.ib2;
            }
        }
        uint32_t px[4];
        for (int c = 0; c < 4; ++c) {
            const uint32_t w = c < 3 ? weightsFor(cbits)[ci] : weightsFor(abits)[ai];
            px[c] = ((64u - w) * ep[s][0][c] + w * ep[s][1][c] + 32u) >> 6;
        }
        if (rot) std::swap(px[rot - 1], px[3]);
        out[t] = make_float4(px[0] / 255.0f, px[1] / 255.0f, px[2] / 255.0f, px[3] / 255.0f);
    }
}

// BC6H: mode value and its field width, transformed endpoints, endpoint bits, delta bits,
// regions, and where each endpoint bit lies after the mode field, in stream order (x[a:b] puts
// bit b first, then on towards a; endpoints 0-3 are region 0's two, then region 1's two)
struct Bc6Mode {
    uint32_t value;
    int width;
    bool transformed;
    int eb, db[3], regions;
    const char *layout;
};
const Bc6Mode kBc6Modes[14] = {
    {0b00, 2, true, 10, {5, 5, 5}, 2, "g2[4] b2[4] b3[4] r0[9:0] g0[9:0] b0[9:0] r1[4:0] g3[4] g2[3:0] g1[4:0] b3[0] g3[3:0] b1[4:0] b3[1] b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3]"},
    {0b01, 2, true, 7, {6, 6, 6}, 2, "g2[5] g3[4] g3[5] r0[6:0] b3[0] b3[1] b2[4] g0[6:0] b2[5] b3[2] g2[4] b0[6:0] b3[3] b3[5] b3[4] r1[5:0] g2[3:0] g1[5:0] g3[3:0] b1[5:0] b2[3:0] r2[5:0] r3[5:0]"},
    {0b00010, 5, true, 11, {5, 4, 4}, 2, "r0[9:0] g0[9:0] b0[9:0] r1[4:0] r0[10] g2[3:0] g1[3:0] g0[10] b3[0] g3[3:0] b1[3:0] b0[10] b3[1] b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3]"},
    {0b00110, 5, true, 11, {4, 5, 4}, 2, "r0[9:0] g0[9:0] b0[9:0] r1[3:0] r0[10] g3[4] g2[3:0] g1[4:0] g0[10] g3[3:0] b1[3:0] b0[10] b3[1] b2[3:0] r2[3:0] b3[0] b3[2] r3[3:0] g2[4] b3[3]"},
    {0b01010, 5, true, 11, {4, 4, 5}, 2, "r0[9:0] g0[9:0] b0[9:0] r1[3:0] r0[10] b2[4] g2[3:0] g1[3:0] g0[10] b3[0] g3[3:0] b1[4:0] b0[10] b2[3:0] r2[3:0] b3[1] b3[2] r3[3:0] b3[4] b3[3]"},
    {0b01110, 5, true, 9, {5, 5, 5}, 2, "r0[8:0] b2[4] g0[8:0] g2[4] b0[8:0] b3[4] r1[4:0] g3[4] g2[3:0] g1[4:0] b3[0] g3[3:0] b1[4:0] b3[1] b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3]"},
    {0b10010, 5, true, 8, {6, 5, 5}, 2, "r0[7:0] g3[4] b2[4] g0[7:0] b3[2] g2[4] b0[7:0] b3[3] b3[4] r1[5:0] g2[3:0] g1[4:0] b3[0] g3[3:0] b1[4:0] b3[1] b2[3:0] r2[5:0] r3[5:0]"},
    {0b10110, 5, true, 8, {5, 6, 5}, 2, "r0[7:0] b3[0] b2[4] g0[7:0] g2[5] g2[4] b0[7:0] g3[5] b3[4] r1[4:0] g3[4] g2[3:0] g1[5:0] g3[3:0] b1[4:0] b3[1] b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3]"},
    {0b11010, 5, true, 8, {5, 5, 6}, 2, "r0[7:0] b3[1] b2[4] g0[7:0] b2[5] g2[4] b0[7:0] b3[5] b3[4] r1[4:0] g3[4] g2[3:0] g1[4:0] b3[0] g3[3:0] b1[5:0] b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3]"},
    {0b11110, 5, false, 6, {6, 6, 6}, 2, "r0[5:0] g3[4] b3[0] b3[1] b2[4] g0[5:0] g2[5] b2[5] b3[2] g2[4] b0[5:0] g3[5] b3[3] b3[5] b3[4] r1[5:0] g2[3:0] g1[5:0] g3[3:0] b1[5:0] b2[3:0] r2[5:0] r3[5:0]"},
    {0b00011, 5, false, 10, {10, 10, 10}, 1, "r0[9:0] g0[9:0] b0[9:0] r1[9:0] g1[9:0] b1[9:0]"},
    {0b00111, 5, true, 11, {9, 9, 9}, 1, "r0[9:0] g0[9:0] b0[9:0] r1[8:0] r0[10] g1[8:0] g0[10] b1[8:0] b0[10]"},
    {0b01011, 5, true, 12, {8, 8, 8}, 1, "r0[9:0] g0[9:0] b0[9:0] r1[7:0] r0[10:11] g1[7:0] g0[10:11] b1[7:0] b0[10:11]"},
    {0b01111, 5, true, 16, {4, 4, 4}, 1, "r0[9:0] g0[9:0] b0[9:0] r1[3:0] r0[10:15] g1[3:0] g0[10:15] b1[3:0] b0[10:15]"},
};

struct LayoutBit {
    uint8_t channel, endpoint, bit;
};
// the layout strings as (channel, endpoint, bit) in stream order, parsed once
const std::vector<LayoutBit> &bc6Layout(int mode) {
    static const std::vector<std::vector<LayoutBit>> layouts = [] {
        std::vector<std::vector<LayoutBit>> all;
        for (const Bc6Mode &M : kBc6Modes) {
            std::vector<LayoutBit> bits;
            const char *p = M.layout;
            while (*p) {
                while (*p == ' ') ++p;
                if (!*p) break;
                const uint8_t ch = static_cast<uint8_t>(*p == 'r' ? 0 : (*p == 'g' ? 1 : 2));
                const uint8_t ep = static_cast<uint8_t>(p[1] - '0');
                p += 3;   // past "x0["
                int a = 0, b = 0;
                while (*p >= '0' && *p <= '9') a = a * 10 + (*p++ - '0');
                b = a;
                if (*p == ':') {
                    ++p;
                    b = 0;
                    while (*p >= '0' && *p <= '9') b = b * 10 + (*p++ - '0');
                }
                ++p;      // past ']'
                const int step = a >= b ? 1 : -1;
                for (int k = b;; k += step) {
                    bits.push_back({ch, ep, static_cast<uint8_t>(k)});
                    if (k == a) break;
                }
            }
            all.push_back(std::move(bits));
        }
        return all;
    }();
    return layouts[mode];
}

int32_t signExtend(uint32_t v, int n) {
    v &= (1u << n) - 1u;
    return (v & (1u << (n - 1))) ? static_cast<int32_t>(v) - (1 << n) : static_cast<int32_t>(v);
}

float halfToFloat(uint16_t h) {
    const uint32_t sign = (h >> 15) & 1u, exp = (h >> 10) & 31u, man = h & 1023u;
    float v;
    if (exp == 0) v = std::ldexp(static_cast<float>(man), -24);
    else if (exp == 31) v = man ? NAN : INFINITY;
    else v = std::ldexp(static_cast<float>(man | 1024u), static_cast<int>(exp) - 25);
    return sign ? -v : v;
}

void bc6Block(const uint8_t *b, bool isSigned, float4 out[16]) {
    BitReader br(b);
    uint32_t m = br.get(2);
    if (m & 2u) m |= br.get(3) << 2;
    const int width = (m & 2u) ? 5 : 2;
    int mode = -1;
    for (int i = 0; i < 14; ++i)
        if (kBc6Modes[i].value == m && kBc6Modes[i].width == width) mode = i;
    if (mode < 0) {
        for (int t = 0; t < 16; ++t) out[t] = make_float4(0, 0, 0, 1);
        return;
    }
    const Bc6Mode &M = kBc6Modes[mode];
    uint32_t raw[4][3] = {};
    for (const LayoutBit &L : bc6Layout(mode)) raw[L.endpoint][L.channel] |= br.get(1) << L.bit;
    const int part = M.regions == 2 ? static_cast<int>(br.get(5)) : 0;
    const int numEp = 2 * M.regions;
    int32_t e[4][3];
    for (int c = 0; c < 3; ++c) {
        e[0][c] = isSigned ? signExtend(raw[0][c], M.eb) : static_cast<int32_t>(raw[0][c]);
        for (int i = 1; i < numEp; ++i) {
            if (M.transformed) {   // base + sign-extended delta, wrapped to eb bits
                const uint32_t v = static_cast<uint32_t>(e[0][c] + signExtend(raw[i][c], M.db[c])) & ((1u << M.eb) - 1u);
                e[i][c] = isSigned ? signExtend(v, M.eb) : static_cast<int32_t>(v);
            } else {
                e[i][c] = isSigned ? signExtend(raw[i][c], M.eb) : static_cast<int32_t>(raw[i][c]);
            }
        }
    }
    auto unquantize = [&](int32_t c) -> int32_t {
        if (!isSigned) {
            if (M.eb >= 15) return c;
            if (c == 0) return 0;
            if (c == (1 << M.eb) - 1) return 0xFFFF;
            return ((c << 16) + 0x8000) >> M.eb;
        }
        if (M.eb >= 16) return c;
        const bool neg = c < 0;
        if (neg) c = -c;
        int32_t q;
        if (c == 0) q = 0;
        else if (c >= (1 << (M.eb - 1)) - 1) q = 0x7FFF;
        else q = ((c << 15) + 0x4000) >> (M.eb - 1);
        return neg ? -q : q;
    };
    int32_t u[4][3];
    for (int i = 0; i < numEp; ++i)
        for (int c = 0; c < 3; ++c) u[i][c] = unquantize(e[i][c]);
    const int ib = M.regions == 2 ? 3 : 4;
    const uint8_t *wts = weightsFor(ib);
    uint32_t idx[16];
    for (int t = 0; t < 16; ++t) idx[t] = br.get(isAnchor(M.regions, part, t) ? ib - 1 : ib);
    for (int t = 0; t < 16; ++t) {
        const int s = M.regions == 2 ? (kPart2[part] >> t) & 1 : 0;
        const int32_t w = wts[idx[t]];
        float v[3];
        for (int c = 0; c < 3; ++c) {
            const int32_t q = (u[2 * s][c] * (64 - w) + u[2 * s + 1][c] * w + 32) >> 6;
            uint16_t h;
            if (isSigned)   // magnitude scaled by 31/32, sign in bit 15
                h = q < 0 ? static_cast<uint16_t>(0x8000u | static_cast<uint32_t>(((-q) * 31) >> 5))
                          : static_cast<uint16_t>((q * 31) >> 5);
            else
                h = static_cast<uint16_t>((q * 31) >> 6);
            v[c] = halfToFloat(h);
        }
        out[t] = make_float4(v[0], v[1], v[2], 1.0f);
    }
}


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
           f == "BC5_Signed" || f == "BC6H" || f == "BC6H_Signed" || f == "BC7";
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
            } else if (f == "BC7") {
                bc7Block(b, t);
            } else if (f == "BC6H" || f == "BC6H_Signed") {
                bc6Block(b, f == "BC6H_Signed", t);
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
