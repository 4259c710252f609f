// Host-side colour handling for material set-up.
//
// Immediate spectra become device parameters exactly as the reference's
// createTripletSpectrum does on the host (libVLR/shared/shared.h:60-118 for the RGB build,
// UpsampledSpectrum's constructor, shared/spectrum_types.cpp:163-228, for the spectral build).
// Tables: CIE 1931 2-degree colour-matching functions and the Meng et al. 2015 grid, read
// from vlr_amd/data/spectral_tables.bin (published data, layout in oracle/extract_tables.py).
#include "runtime.h"
#include "color.h"

#include <dlfcn.h>
#include <algorithm>
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

MengTables hostMengTables() { return MengTables{g_tables.grid.data(), g_tables.points.data()}; }

// A sampled spectrum as a rendering-RGB triplet for the RGB build, as
// Regular/IrregularSampledSpectrum::toXYZ and transformToRenderingRGB do
// (shared/spectrum_types.cpp:414-488, 521-591; shared/spectrum_base.h:138-152): XYZ integrated
// over the merged break points of the 1 nm CMF table (360-830 nm) and of the spectrum's samples,
// from min(360 nm, first sample) to 830 nm, each interval adding
// mean(value) x (cmf_a + cmf_b) x width / 2 to compensated (Kahan) sums; values are
// interpolated linearly and held constant beyond the end samples, the CMF is zero below 360 nm
// and interpolated between its samples; the sums are divided by the integral of ybar (the same
// trapezoid rule over its 471 samples, shared/spectrum_base.cpp:196-205). Then the reflectance
// (E) or light-source (D65) matrix, and negative channels clamped to zero. A regular spectrum
// places sample k at minLambda + k * (maxLambda - minLambda) / (n - 1); a spectrum of one sample
// is that value everywhere. An irregular spectrum's wavelengths must ascend (the walk stops
// after 471 + 2n + 2 steps, which an ascending list never reaches).
namespace {
struct Kahan {
    float sum = 0.0f, comp = 0.0f;
    void add(float v) {
        const float in = v - comp;
        const float t = sum + in;
        comp = (t - sum) - in;
        sum = t;
    }
};
}  // namespace

void sampledSpectrumToRGB(SpectrumType type, bool regular, float minLambda, float maxLambda, const float *lambdas,
                          const float *values, uint32_t n, float rgb[3]) {
    const Tables &T = g_tables;
    constexpr uint32_t kCmf = 471;
    constexpr float kLow = 360.0f, kBin = 1.0f;
    Kahan yInt;
    for (uint32_t i = 1; i < kCmf; ++i) yInt.add((T.ybar[i - 1] + T.ybar[i]) * 1 * 0.5f);
    float XYZ[3];
    if (n == 1) {
        Kahan c[3];
        for (uint32_t i = 1; i < kCmf; ++i) {
            c[0].add(values[0] * (T.xbar[i - 1] + T.xbar[i]) * 0.5f);
            c[1].add(values[0] * (T.ybar[i - 1] + T.ybar[i]) * 0.5f);
            c[2].add(values[0] * (T.zbar[i - 1] + T.zbar[i]) * 0.5f);
        }
        for (int k = 0; k < 3; ++k) XYZ[k] = c[k].sum / yInt.sum;
    } else {
        const float first = regular ? minLambda : lambdas[0];
        const float last = regular ? maxLambda : lambdas[n - 1];
        const float binWidth = regular ? (maxLambda - minLambda) / (n - 1) : 0.0f;
        auto sampleAt = [&](uint32_t k) { return regular ? minLambda + k * binWidth : lambdas[k]; };
        uint32_t cmfIdx = 0, baseIdx = 0;
        float wl = std::fmin(kLow, first);
        float prevCmf[3] = {0.0f, 0.0f, 0.0f}, prevValue = 0.0f, halfWidth = 0.0f;
        Kahan sum[3];
        // each step passes a CMF or a sample point; the bound only stops unsorted wavelengths
        for (uint32_t step = 0; step < kCmf + 2 * n + 2; ++step) {
            float cmf[3];
            if (wl < kLow) {
                cmf[0] = cmf[1] = cmf[2] = 0.0f;
            } else if (wl == kLow + cmfIdx * kBin) {
                cmf[0] = T.xbar[cmfIdx];
                cmf[1] = T.ybar[cmfIdx];
                cmf[2] = T.zbar[cmfIdx];
                ++cmfIdx;
            } else {
                const uint32_t idx = std::min(static_cast<uint32_t>((wl - kLow) / kBin), kCmf - 1);
                const float t = (wl - (kLow + idx * kBin)) / kBin;
                cmf[0] = (1 - t) * T.xbar[idx] + t * T.xbar[idx + 1];
                cmf[1] = (1 - t) * T.ybar[idx] + t * T.ybar[idx + 1];
                cmf[2] = (1 - t) * T.zbar[idx] + t * T.zbar[idx + 1];
            }
            float value;
            if (wl < first) {
                value = values[0];
            } else if (wl > last) {
                value = values[n - 1];
            } else if (wl == sampleAt(baseIdx)) {
                value = values[baseIdx];
                ++baseIdx;
            } else if (regular) {
                const uint32_t idx = std::min(static_cast<uint32_t>((wl - minLambda) / binWidth), n - 2);
                const float t = (wl - (minLambda + idx * binWidth)) / binWidth;
                value = (1 - t) * values[idx] + t * values[idx + 1];
            } else {
                // the last sample at or before wl, searched from the one before baseIdx
                const float *lb = std::lower_bound(lambdas + (baseIdx > 0 ? baseIdx - 1 : 0), lambdas + n, wl);
                const uint32_t idx = static_cast<uint32_t>(std::max<long>(long(lb - lambdas) - 1, 0));
                const float t = (wl - lambdas[idx]) / (lambdas[idx + 1] - lambdas[idx]);
                value = (1 - t) * values[idx] + t * values[idx + 1];
            }
            const float avg = (prevValue + value) * 0.5f;
            for (int k = 0; k < 3; ++k) {
                sum[k].add(avg * (prevCmf[k] + cmf[k]) * halfWidth);
                prevCmf[k] = cmf[k];
            }
            prevValue = value;
            const float prevWl = wl;
            wl = std::fmin(kLow + cmfIdx * kBin, baseIdx < n ? sampleAt(baseIdx) : INFINITY);
            halfWidth = (wl - prevWl) * 0.5f;
            if (cmfIdx == kCmf) break;
        }
        for (int k = 0; k < 3; ++k) XYZ[k] = sum[k].sum / yInt.sum;
    }
    colorTristim(type == SpectrumType::LightSource ? 1 : 3, XYZ, rgb);
    for (int k = 0; k < 3; ++k) rgb[k] = std::fmax(0.0f, rgb[k]);
}

DevSpectrum makeDevSpectrum(SpectrumType type, const ImmSpectrum &is, bool spectral) {
    return colorToDevSpectrum(static_cast<uint32_t>(type), static_cast<uint32_t>(is.cs), is.e[0], is.e[1], is.e[2],
                              spectral, hostMengTables());
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
