// Host-side colour handling for material set-up.
//
// Immediate spectra become device parameters exactly as the reference's
// createTripletSpectrum does on the host (libVLR/shared/shared.h:60-118 for the RGB build,
// UpsampledSpectrum's constructor, shared/spectrum_types.cpp:163-228, for the spectral build).
// Tables: CIE 1931 2-degree colour-matching functions and the Meng et al. 2015 grid, read
// from vlr_amd/data/spectral_tables.bin (published data, layout in oracle/extract_tables.py).
#include "runtime.h"

#include <dlfcn.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

namespace vlra {

namespace {
struct Tables {
    float xbar[471], ybar[471], zbar[471];
    std::vector<MengCellDev> grid;
    std::vector<float> points;   // 186 x 99
    float binX[16], binY[16], binZ[16], integral;
    bool ok = false;
};
Tables g_tables;
std::once_flag g_once;

std::string tablesPath() {
    if (const char *env = std::getenv("VLR_AMD_SPECTRAL_TABLES")) return env;
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&tablesPath), &info) && info.dli_fname) {
        std::string so = info.dli_fname;
        size_t slash = so.rfind('/');
        std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
        return dir + "/data/spectral_tables.bin";
    }
    return "vlr_amd/data/spectral_tables.bin";
}

// CMF binning into 16 strata (spectrum_types.cpp:624-670)
void binCMF(Tables &T) {
    uint32_t sBin = 0;
    float nextP = float(sBin + 1) / 16;
    float xSum = 0, xPrev = T.xbar[0], ySum = 0, yPrev = T.ybar[0], zSum = 0, zPrev = T.zbar[0];
    const float interval = 1;
    for (int i = 1; i < 471; ++i) {
        float curP = float(i) / (471 - 1);
        float width = interval;
        float xCur = T.xbar[i], yCur = T.ybar[i], zCur = T.zbar[i];
        if (curP >= nextP) {
            width = (curP - nextP) * (830.0f - 360.0f);
            float t = 1 - width / interval;
            float xIn = xPrev * (1 - t) + xCur * t;
            float yIn = yPrev * (1 - t) + yCur * t;
            float zIn = zPrev * (1 - t) + zCur * t;
            xSum += (xPrev + xIn) * (interval - width) * 0.5f;
            ySum += (yPrev + yIn) * (interval - width) * 0.5f;
            zSum += (zPrev + zIn) * (interval - width) * 0.5f;
            T.binX[sBin] = xSum;
            T.binY[sBin] = ySum;
            T.binZ[sBin] = zSum;
            xSum = ySum = zSum = 0;
            xPrev = xIn;
            yPrev = yIn;
            zPrev = zIn;
            ++sBin;
            nextP = float(sBin + 1) / 16;
        }
        xSum += (xPrev + xCur) * width * 0.5f;
        ySum += (yPrev + yCur) * width * 0.5f;
        zSum += (zPrev + zCur) * width * 0.5f;
        xPrev = xCur;
        yPrev = yCur;
        zPrev = zCur;
    }
    T.integral = 0.0f;
    for (int i = 0; i < 16; ++i) T.integral += T.binY[i];
}

void loadOnce() {
    Tables &T = g_tables;
    std::string path = tablesPath();
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return;
    char magic[8];
    uint32_t n = 0, gw = 0, gh = 0, np = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "VLRSPEC1", 8) == 0;
    ok = ok && std::fread(&n, 4, 1, f) == 1 && n == 471;
    ok = ok && std::fread(T.xbar, 4, 471, f) == 471 && std::fread(T.ybar, 4, 471, f) == 471 &&
         std::fread(T.zbar, 4, 471, f) == 471;
    ok = ok && std::fread(&gw, 4, 1, f) == 1 && std::fread(&gh, 4, 1, f) == 1 && gw == 12 && gh == 14;
    T.grid.resize(168);
    ok = ok && std::fread(T.grid.data(), sizeof(MengCellDev), 168, f) == 168;
    ok = ok && std::fread(&np, 4, 1, f) == 1 && np == 186;
    T.points.resize(186 * 99);
    ok = ok && std::fread(T.points.data(), 4, 186 * 99, f) == 186 * 99;
    std::fclose(f);
    if (!ok) return;
    binCMF(T);
    T.ok = true;
}

const float mat_Rec709_D65_to_XYZ[9] = {0.4124564f, 0.2126729f, 0.0193339f, 0.3575761f, 0.7151522f,
                                        0.1191920f, 0.1804375f, 0.0721750f, 0.9503041f};
const float mat_XYZ_to_Rec709_D65[9] = {3.2404542f, -0.9692660f, 0.0556434f, -1.5371385f, 1.8760108f,
                                        -0.2040259f, -0.4985314f, 0.0415560f, 1.0572252f};
const float mat_Rec709_E_to_XYZ[9] = {0.4969f, 0.2562f, 0.0233f, 0.3391f, 0.6782f, 0.1130f, 0.1640f, 0.0656f, 0.8637f};
const float mat_XYZ_to_Rec709_E[9] = {2.6897f, -1.0221f, 0.0612f, -1.2759f, 1.9783f, -0.2245f, -0.4138f, 0.0438f, 1.1633f};

inline void tristim(const float m[9], const float s[3], float d[3]) {
    d[0] = m[0] * s[0] + m[3] * s[1] + m[6] * s[2];
    d[1] = m[1] * s[0] + m[4] * s[1] + m[7] * s[2];
    d[2] = m[2] * s[0] + m[5] * s[1] + m[8] * s[2];
}

inline float degamma(float v) { return v <= 0.04045f ? v / 12.92f : std::pow((v + 0.055f) / 1.055f, 2.4f); }

// Meng grid lookup (spectrum_types.cpp:11-91)
void mengAdj(float u, float v, uint32_t *adj, uint16_t *s, uint16_t *t) {
    const Tables &T = g_tables;
    u = std::fmin(std::fmax(u, 0.0f), 12.0f);
    v = std::fmin(std::fmax(v, 0.0f), 14.0f);
    int32_t ui = std::min(static_cast<int32_t>(u), 11), vi = std::min(static_cast<int32_t>(v), 13);
    const MengCellDev &cell = T.grid[ui + 12 * vi];
    const uint8_t *ix = cell.idx;
    const float *P = T.points.data();
    *adj = 0;
    *s = *t = 0;
    if (cell.inside) {
        float ss = u - ui, tt = v - vi;
        *adj = (uint32_t(ix[3]) << 24) | (uint32_t(ix[2]) << 16) | (uint32_t(ix[1]) << 8) | uint32_t(ix[0]);
        *s = static_cast<uint16_t>(std::fmax(ss * (UINT16_MAX - 1), 0.0f));
        *t = static_cast<uint16_t>(std::fmax(tt * (UINT16_MAX - 1), 0.0f));
        return;
    }
    // cells with no data points lie outside the spectral locus; keep adj = 0
    if (cell.numPoints < 3) return;
    const float ex = u - P[ix[0] * 99 + 2], ey = v - P[ix[0] * 99 + 3];
    float e0x = P[ix[1] * 99 + 2] - P[ix[0] * 99 + 2];
    float e0y = P[ix[1] * 99 + 3] - P[ix[0] * 99 + 3];
    float uu = e0x * ey - ex * e0y;
    for (int i = 1; i < cell.numPoints; ++i) {
        uint32_t idx = ix[i % (cell.numPoints - 1) + 1];
        float e1x = P[idx * 99 + 2] - P[ix[0] * 99 + 2];
        float e1y = P[idx * 99 + 3] - P[ix[0] * 99 + 3];
        float vv = ex * e1y - e1x * ey;
        const float area = e0x * e1y - e1x * e0y;
        const float b0 = uu / area, b1 = vv / area;
        float b2 = 1.0f - b0 - b1;
        if (b0 < -1e-6f || b1 < -1e-6f || b2 < -1e-6f) {
            uu = -vv;
            e0x = e1x;
            e0y = e1y;
            continue;
        }
        *adj = (uint32_t(UINT8_MAX) << 24) | (uint32_t(ix[0]) << 16) | (uint32_t(ix[i]) << 8) | idx;
        *s = static_cast<uint16_t>(std::fmax(b0 * (UINT16_MAX - 1), 0.0f));
        *t = static_cast<uint16_t>(std::fmax(b1 * (UINT16_MAX - 1), 0.0f));
        break;
    }
}
} // namespace

bool loadSpectralTables() {
    std::call_once(g_once, loadOnce);
    return g_tables.ok;
}
const std::vector<MengCellDev> &mengGrid() { return g_tables.grid; }
const std::vector<float> &mengPoints() { return g_tables.points; }
void cmfBins(float x[16], float y[16], float z[16], float *integral) {
    for (int i = 0; i < 16; ++i) {
        x[i] = g_tables.binX[i];
        y[i] = g_tables.binY[i];
        z[i] = g_tables.binZ[i];
    }
    *integral = g_tables.integral;
}

void mengXYtoUV(const float xy[2], float uv[2]) {
    uv[0] = 16.730260708356887f * xy[0] + 7.7801960340706f * xy[1] - 2.170152247475828f;
    uv[1] = -7.530081094743006f * xy[0] + 16.192422314095225f * xy[1] + 1.1125529268825947f;
}

DevSpectrum makeDevSpectrum(SpectrumType type, const ImmSpectrum &is, bool spectral) {
    float e0 = is.e[0], e1 = is.e[1], e2 = is.e[2];
    DevSpectrum out = {0, 0, 0, 0};
    const bool light = type == SpectrumType::LightSource;
    if (!spectral) {
        float XYZ[3] = {0, 0, 0};
        switch (is.cs) {
        case ColorSpace::Rec709_D65_sRGBGamma:
            e0 = degamma(e0);
            e1 = degamma(e1);
            e2 = degamma(e2);
            [[fallthrough]];
        case ColorSpace::Rec709_D65: {
            float RGB[3] = {e0, e1, e2};
            tristim(light ? mat_Rec709_D65_to_XYZ : mat_Rec709_E_to_XYZ, RGB, XYZ);
            break;
        }
        case ColorSpace::XYZ:
            XYZ[0] = e0; XYZ[1] = e1; XYZ[2] = e2;
            break;
        case ColorSpace::xyY:
            if (e1 != 0) {
                float z = 1 - (e0 + e1), b = e2 / e1;
                XYZ[0] = e0 * b; XYZ[1] = e2; XYZ[2] = z * b;
            }
            break;
        }
        float RGB[3];
        tristim(light ? mat_XYZ_to_Rec709_D65 : mat_XYZ_to_Rec709_E, XYZ, RGB);
        out.x = RGB[0]; out.y = RGB[1]; out.z = RGB[2];
        return out;
    }
    float xy[2] = {0, 0}, brightness = 0;
    switch (is.cs) {
    case ColorSpace::Rec709_D65_sRGBGamma:
        e0 = degamma(e0);
        e1 = degamma(e1);
        e2 = degamma(e2);
        [[fallthrough]];
    case ColorSpace::Rec709_D65: {
        float RGB[3] = {e0, e1, e2}, XYZ[3];
        tristim(light ? mat_Rec709_D65_to_XYZ : mat_Rec709_E_to_XYZ, RGB, XYZ);
        e0 = XYZ[0]; e1 = XYZ[1]; e2 = XYZ[2];
        [[fallthrough]];
    }
    case ColorSpace::XYZ: {
        float Y = e1, b = e0 + e1 + e2;
        e0 = e0 / b;
        e1 = e1 / b;
        e2 = Y;
        [[fallthrough]];
    }
    case ColorSpace::xyY:
        xy[0] = e0;
        xy[1] = e1;
        brightness = e2 / e1;
        if (e2 == 0) {
            xy[0] = 0.3333f;
            xy[1] = 0.3333f;
            brightness = 0;
        }
        break;
    }
    float scale = brightness / 0.009355121400914532f;
    float uv[2];
    mengXYtoUV(xy, uv);
    uint32_t adj;
    uint16_t s, t;
    mengAdj(uv[0], uv[1], &adj, &s, &t);
    uint32_t st = uint32_t(s) | (uint32_t(t) << 16);
    std::memcpy(&out.x, &adj, 4);
    std::memcpy(&out.y, &st, 4);
    out.z = scale;
    return out;
}

static const char *kColorSpaces[] = {"Rec709(D65) sRGB Gamma", "Rec709(D65)", "XYZ", "xyY"};

bool colorSpaceFromString(const char *s, ColorSpace *cs) {
    if (!s) return false;
    for (uint32_t i = 0; i < 4; ++i) {
        if (paramNameEquals(s, kColorSpaces[i])) {
            *cs = static_cast<ColorSpace>(i);
            return true;
        }
    }
    return false;
}
const char *colorSpaceToString(ColorSpace cs) { return kColorSpaces[static_cast<uint32_t>(cs) & 3]; }

bool spectrumTypeFromString(const char *s, SpectrumType *t) {
    if (!s) return false;
    if (paramNameEquals(s, "Reflectance") || paramNameEquals(s, "Transmittance")) { *t = SpectrumType::Reflectance; return true; }
    if (paramNameEquals(s, "Light Source")) { *t = SpectrumType::LightSource; return true; }
    if (paramNameEquals(s, "NA")) { *t = SpectrumType::NA; return true; }
    return false;
}

uint16_t floatToHalf(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000;
    uint32_t absx = x & 0x7FFFFFFF;
    if (absx >= 0x7F800000) return uint16_t(sign | 0x7C00 | (absx > 0x7F800000 ? 0x200 : 0));
    if (absx >= 0x477FF000) return uint16_t(sign | 0x7C00);
    if (absx < 0x38800000) {
        if (absx < 0x33000000) return uint16_t(sign);
        uint32_t e = absx >> 23;
        uint32_t m = (absx & 0x7FFFFF) | 0x800000;
        uint32_t shift = 126 - e;
        uint32_t hm = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1);
        uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (hm & 1))) ++hm;
        return uint16_t(sign | hm);
    }
    uint32_t e = (absx >> 23) - 127 + 15;
    uint32_t m = absx & 0x7FFFFF;
    uint32_t h = (e << 10) | (m >> 13);
    uint32_t rem = m & 0x1FFF;
    if (rem > 0x1000 || (rem == 0x1000 && (h & 1))) ++h;
    return uint16_t(sign | h);
}

float halfToFloat(uint16_t h) {
    uint32_t sign = uint32_t(h & 0x8000) << 16;
    uint32_t exp = (h >> 10) & 0x1F, man = h & 0x3FF, bits;
    if (exp == 0) {
        if (man == 0) {
            bits = sign;
        } else {
            int e = -1;
            do { ++e; man <<= 1; } while ((man & 0x400) == 0);
            man &= 0x3FF;
            bits = sign | uint32_t(127 - 15 - e) << 23 | (man << 13);
        }
    } else if (exp == 31) {
        bits = sign | 0x7F800000u | (man << 13);
    } else {
        bits = sign | ((exp - 15 + 127) << 23) | (man << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

} // namespace vlra
