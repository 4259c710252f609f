// Oracle (test infrastructure only): CPU restatement of the reference's RGB and spectral
// types.  Sources followed:
//   RGB mode:       libVLR/shared/rgb_spectrum_types.h:6-273, shared.h:60-118
//   spectral mode:  libVLR/shared/spectrum_types.h:6-704, spectrum_types.cpp:11-336, 596-670
//   colour spaces:  libVLR/shared/spectrum_base.h:109-150
//   Kahan sum:      libVLR/shared/common_internal.h:350-368
#pragma once
#include "ov_math.h"
#include <cstdio>
#include <string>

namespace ov {

enum class SpectrumType { Reflectance = 0, LightSource, IndexOfRefraction, NA };
enum class ColorSpace { Rec709_D65_sRGBGamma = 0, Rec709_D65, XYZ, xyY };

constexpr float kWLLow = 360.0f;
constexpr float kWLHigh = 830.0f;
constexpr int kNumCMF = 471;
constexpr int kNumStrata = 16;

// spectrum_base.h:109-128 (column-major)
static const float mat_Rec709_D65_to_XYZ[9] = {0.4124564f, 0.2126729f, 0.0193339f, 0.3575761f, 0.7151522f,
                                               0.1191920f, 0.1804375f, 0.0721750f, 0.9503041f};
static const float mat_XYZ_to_Rec709_D65[9] = {3.2404542f, -0.9692660f, 0.0556434f, -1.5371385f, 1.8760108f,
                                               -0.2040259f, -0.4985314f, 0.0415560f, 1.0572252f};
static const float mat_Rec709_E_to_XYZ[9] = {0.4969f, 0.2562f, 0.0233f, 0.3391f, 0.6782f,
                                             0.1130f, 0.1640f, 0.0656f, 0.8637f};
static const float mat_XYZ_to_Rec709_E[9] = {2.6897f, -1.0221f, 0.0612f, -1.2759f, 1.9783f,
                                             -0.2245f, -0.4138f, 0.0438f, 1.1633f};

inline void transformTristimulus(const float m[9], const float s[3], float d[3]) {
    d[0] = m[0] * s[0] + m[3] * s[1] + m[6] * s[2];
    d[1] = m[1] * s[0] + m[4] * s[1] + m[7] * s[2];
    d[2] = m[2] * s[0] + m[5] * s[1] + m[8] * s[2];
}

// spectrum_base.h:42-48
inline float sRGB_degamma(float v) {
    if (v <= 0.04045f)
        return v / 12.92f;
    return std::pow((v + 0.055f) / 1.055f, 2.4f);
}

// ---------------------------------------------------------------------------------------
// Tables (published data written by oracle/extract_tables.py)
struct MengCell { uint8_t inside, numPoints, idx[6]; };
struct MengPoint { float xystar[2], uv[2], spectrum[95]; };

struct SpectralTables {
    float xbar[kNumCMF], ybar[kNumCMF], zbar[kNumCMF];
    MengCell grid[12 * 14];
    MengPoint points[186];
    // DiscretizedSpectrum CMF bins (spectrum_types.cpp:624-670)
    float binX[kNumStrata], binY[kNumStrata], binZ[kNumStrata];
    float integralCMF16;
    bool loaded = false;

    bool load(const char *path) {
        FILE *f = std::fopen(path, "rb");
        if (!f)
            return false;
        char magic[8];
        uint32_t n = 0, gw = 0, gh = 0, np = 0;
        bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "VLRSPEC1", 8) == 0;
        ok = ok && std::fread(&n, 4, 1, f) == 1 && n == kNumCMF;
        ok = ok && std::fread(xbar, 4, kNumCMF, f) == kNumCMF;
        ok = ok && std::fread(ybar, 4, kNumCMF, f) == kNumCMF;
        ok = ok && std::fread(zbar, 4, kNumCMF, f) == kNumCMF;
        ok = ok && std::fread(&gw, 4, 1, f) == 1 && std::fread(&gh, 4, 1, f) == 1 && gw == 12 && gh == 14;
        ok = ok && std::fread(grid, sizeof(MengCell), 168, f) == 168;
        ok = ok && std::fread(&np, 4, 1, f) == 1 && np == 186;
        ok = ok && std::fread(points, sizeof(MengPoint), 186, f) == 186;
        std::fclose(f);
        if (!ok)
            return false;
        initBins();
        loaded = true;
        return true;
    }

    // spectrum_types.cpp:624-670, trapezoidal integration of the 1 nm CMF into 16 strata.
    void initBins() {
        uint32_t sBin = 0;
        float nextP = float(sBin + 1) / kNumStrata;
        float xSum = 0, xPrev = xbar[0];
        float ySum = 0, yPrev = ybar[0];
        float zSum = 0, zPrev = zbar[0];
        const float interval = 1;
        for (int i = 1; i < kNumCMF; ++i) {
            float curP = float(i) / (kNumCMF - 1);
            float width = interval;
            float xCur = xbar[i], yCur = ybar[i], zCur = zbar[i];
            if (curP >= nextP) {
                width = (curP - nextP) * (kWLHigh - kWLLow);
                float t = 1 - width / interval;
                float xIn = xPrev * (1 - t) + xCur * t;
                float yIn = yPrev * (1 - t) + yCur * t;
                float zIn = zPrev * (1 - t) + zCur * t;
                xSum += (xPrev + xIn) * (interval - width) * 0.5f;
                ySum += (yPrev + yIn) * (interval - width) * 0.5f;
                zSum += (zPrev + zIn) * (interval - width) * 0.5f;
                binX[sBin] = xSum;
                binY[sBin] = ySum;
                binZ[sBin] = zSum;
                xSum = ySum = zSum = 0;
                xPrev = xIn;
                yPrev = yIn;
                zPrev = zIn;
                ++sBin;
                nextP = float(sBin + 1) / kNumStrata;
            }
            xSum += (xPrev + xCur) * width * 0.5f;
            ySum += (yPrev + yCur) * width * 0.5f;
            zSum += (zPrev + zCur) * width * 0.5f;
            xPrev = xCur;
            yPrev = yCur;
            zPrev = zCur;
        }
        integralCMF16 = 0.0f;
        for (int i = 0; i < kNumStrata; ++i)
            integralCMF16 += binY[i];
    }
};

SpectralTables &tables();

// ---------------------------------------------------------------------------------------
// Sampled spectra: N = 3 (RGB build) or 4 (spectral build, 4 hero wavelengths)
template <int N>
struct SS {
    float v[N];
    SS() { for (int i = 0; i < N; ++i) v[i] = 0; }
    explicit SS(float s) { for (int i = 0; i < N; ++i) v[i] = s; }
    float &operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
#define OV_SS_BINOP(op) \
    SS operator op(const SS &o) const { SS r; for (int i = 0; i < N; ++i) r.v[i] = v[i] op o.v[i]; return r; } \
    SS &operator op##=(const SS &o) { for (int i = 0; i < N; ++i) v[i] op##= o.v[i]; return *this; }
    OV_SS_BINOP(+)
    OV_SS_BINOP(-)
    OV_SS_BINOP(*)
    OV_SS_BINOP(/)
#undef OV_SS_BINOP
    SS operator*(float s) const { SS r; for (int i = 0; i < N; ++i) r.v[i] = v[i] * s; return r; }
    SS operator/(float s) const { float rc = 1.0f / s; SS r; for (int i = 0; i < N; ++i) r.v[i] = v[i] * rc; return r; }
    SS &operator*=(float s) { for (int i = 0; i < N; ++i) v[i] *= s; return *this; }
    SS &operator/=(float s) { float rc = 1.0f / s; for (int i = 0; i < N; ++i) v[i] *= rc; return *this; }
    SS operator+(float s) const { SS r; for (int i = 0; i < N; ++i) r.v[i] = v[i] + s; return r; }
    SS operator-(float s) const { SS r; for (int i = 0; i < N; ++i) r.v[i] = v[i] - s; return r; }
    bool operator==(const SS &o) const { for (int i = 0; i < N; ++i) if (v[i] != o.v[i]) return false; return true; }
    bool hasNonZero() const { for (int i = 0; i < N; ++i) if (v[i] != 0.0f) return true; return false; }
    bool allFinite() const { for (int i = 0; i < N; ++i) if (!std::isfinite(v[i])) return false; return true; }
    // rgb_spectrum_types.h:145-150 and spectrum_types.h:250-264
    float importance(uint32_t sel) const {
        float sum = 0;
        for (int i = 0; i < N; ++i) sum += v[i];
        const float primary = 0.9f;
        const float marginal = (1 - primary) / (N - 1);
        return sum * marginal + v[sel] * (primary - marginal);
    }
};
template <int N> inline SS<N> operator*(float s, const SS<N> &c) { SS<N> r; for (int i = 0; i < N; ++i) r.v[i] = s * c.v[i]; return r; }
template <int N> inline SS<N> operator-(float s, const SS<N> &c) { SS<N> r; for (int i = 0; i < N; ++i) r.v[i] = s - c.v[i]; return r; }
template <int N> inline SS<N> lerpS(const SS<N> &a, const SS<N> &b, float t) { return (1 - t) * a + t * b; }

// Wavelength samples.  RGB: rgb_spectrum_types.h:6-44.  Spectral: spectrum_types.h:6-69.
template <int N>
struct WLS {
    float lambdas[N];
    uint32_t sel;
    bool single;
    static WLS createWithEqualOffsets(float offset, float uLambda, float *pdf) {
        WLS w;
        if (N == 4) {
            for (int i = 0; i < N; ++i)
                w.lambdas[i] = kWLLow + (kWLHigh - kWLLow) * (i + offset) / N;
            *pdf = N / (kWLHigh - kWLLow);
        } else {
            for (int i = 0; i < N; ++i) w.lambdas[i] = 0.0f;
            *pdf = 1;
        }
        w.sel = std::min<uint16_t>(static_cast<uint16_t>(N * uLambda), N - 1);
        w.single = false;
        return w;
    }
};

// ---------------------------------------------------------------------------------------
// Meng et al. 2015 upsampled spectrum (spectrum_types.cpp:11-91, 163-228, 273-336)
struct Meng {
    uint32_t adj = 0;
    uint16_t s = 0, t = 0;
    float scale = 0;
    static constexpr float EER = 0.009355121400914532f;

    static void xy_to_uv(const float xy[2], float uv[2]) {
        uv[0] = 16.730260708356887f * xy[0] + 7.7801960340706f * xy[1] - 2.170152247475828f;
        uv[1] = -7.530081094743006f * xy[0] + 16.192422314095225f * xy[1] + 1.1125529268825947f;
    }

    void computeAdjacents(float u, float v) {
        const SpectralTables &T = tables();
        u = clampv(u, 0.0f, 12.0f);
        v = clampv(v, 0.0f, 14.0f);
        // u == 12 or v == 14 (clamped input) would index past the grid: keep the last cell
        int32_t ui = std::min(static_cast<int32_t>(u), 11), vi = std::min(static_cast<int32_t>(v), 13);
        const MengCell &cell = T.grid[ui + 12 * vi];
        const uint8_t *indices = cell.idx;
        const uint8_t numPoints = cell.numPoints;
        if (cell.inside) {
            float ss = u - ui, tt = v - vi;
            adj = (uint32_t(indices[3]) << 24) | (uint32_t(indices[2]) << 16) | (uint32_t(indices[1]) << 8) | uint32_t(indices[0]);
            s = static_cast<uint16_t>(std::fmax(ss * (UINT16_MAX - 1), 0.0f));
            t = static_cast<uint16_t>(std::fmax(tt * (UINT16_MAX - 1), 0.0f));
        } else if (numPoints >= 3) {
            // cells with no data points lie outside the spectral locus; keep adj = 0
            const float ex = u - T.points[indices[0]].uv[0];
            const float ey = v - T.points[indices[0]].uv[1];
            float e0x = T.points[indices[1]].uv[0] - T.points[indices[0]].uv[0];
            float e0y = T.points[indices[1]].uv[1] - T.points[indices[0]].uv[1];
            float uu = e0x * ey - ex * e0y;
            for (int i = 1; i < numPoints; ++i) {
                uint32_t idx = indices[i % (numPoints - 1) + 1];
                float e1x = T.points[idx].uv[0] - T.points[indices[0]].uv[0];
                float e1y = T.points[idx].uv[1] - T.points[indices[0]].uv[1];
                float vv = ex * e1y - e1x * ey;
                const float area = e0x * e1y - e1x * e0y;
                const float b0 = uu / area;
                const float b1 = vv / area;
                float b2 = 1.0f - b0 - b1;
                if (b0 < -1e-6f || b1 < -1e-6f || b2 < -1e-6f) {
                    uu = -vv;
                    e0x = e1x;
                    e0y = e1y;
                    continue;
                }
                adj = (uint32_t(UINT8_MAX) << 24) | (uint32_t(indices[0]) << 16) | (uint32_t(indices[i]) << 8) | idx;
                s = static_cast<uint16_t>(std::fmax(b0 * (UINT16_MAX - 1), 0.0f));
                t = static_cast<uint16_t>(std::fmax(b1 * (UINT16_MAX - 1), 0.0f));
                break;
            }
        }
    }

    static Meng fromUVS(float u, float v, float sc) {
        Meng m;
        m.computeAdjacents(u, v);
        m.scale = sc;
        return m;
    }

    static Meng create(SpectrumType spType, ColorSpace space, float e0, float e1, float e2) {
        float xy[2] = {0.0f, 0.0f};
        float brightness = 0;
        switch (space) {
        case ColorSpace::Rec709_D65_sRGBGamma:
            e0 = sRGB_degamma(e0);
            e1 = sRGB_degamma(e1);
            e2 = sRGB_degamma(e2);
            [[fallthrough]];
        case ColorSpace::Rec709_D65: {
            float RGB[3] = {e0, e1, e2}, XYZ[3];
            if (spType == SpectrumType::LightSource)
                transformTristimulus(mat_Rec709_D65_to_XYZ, RGB, XYZ);
            else
                transformTristimulus(mat_Rec709_E_to_XYZ, RGB, XYZ);
            e0 = XYZ[0];
            e1 = XYZ[1];
            e2 = XYZ[2];
            [[fallthrough]];
        }
        case ColorSpace::XYZ: {
            float Y = e1;
            float b = e0 + e1 + e2;
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
        Meng m;
        m.scale = brightness / EER;
        float uv[2];
        xy_to_uv(xy, uv);
        m.computeAdjacents(uv[0], uv[1]);
        return m;
    }

    SS<4> evaluate(const WLS<4> &wls) const {
        const SpectralTables &T = tables();
        uint8_t a[4] = {uint8_t(adj & 0xFF), uint8_t((adj >> 8) & 0xFF), uint8_t((adj >> 16) & 0xFF), uint8_t((adj >> 24) & 0xFF)};
        float sf = static_cast<float>(s) / (UINT16_MAX - 1);
        float tf = static_cast<float>(t) / (UINT16_MAX - 1);
        int numAdj;
        float w[4];
        if (a[3] != UINT8_MAX) {
            w[0] = (1 - sf) * (1 - tf);
            w[1] = sf * (1 - tf);
            w[2] = (1 - sf) * tf;
            w[3] = sf * tf;
            numAdj = 4;
        } else {
            w[0] = sf;
            w[1] = tf;
            w[2] = 1.0f - sf - tf;
            numAdj = 3;
        }
        SS<4> ret(0.0f);
        for (int i = 0; i < 4; ++i) {
            float p = (wls.lambdas[i] - 360.0f) / (830.0f - 360.0f);
            p = clampv(p, 0.0f, 1.0f);
            float sBinF = p * (95 - 1);
            uint32_t sBin = std::min<uint32_t>(static_cast<uint32_t>(sBinF), 95 - 1);
            uint32_t sBinNext = std::min<uint32_t>(sBin + 1, 95 - 1);
            float tt = sBinF - sBin;
            for (int j = 0; j < numAdj; ++j) {
                const float *sp = T.points[a[j]].spectrum;
                ret[i] += w[j] * (sp[sBin] * (1 - tt) + sp[sBinNext] * tt);
            }
        }
        return ret * scale;
    }
};

// A triplet parameter: an RGB value in the RGB build, a Meng spectrum in the spectral build.
struct Triplet {
    float rgb[3] = {0, 0, 0};
    Meng meng;
    // shared.h:60-118 (RGB) / UpsampledSpectrum ctor (spectral)
    static Triplet create(SpectrumType spType, ColorSpace cs, float e0, float e1, float e2) {
        Triplet t;
        t.meng = Meng::create(spType, cs, e0, e1, e2);
        float XYZ[3] = {0.0f, 0.0f, 0.0f};
        switch (cs) {
        case ColorSpace::Rec709_D65_sRGBGamma:
            e0 = sRGB_degamma(e0);
            e1 = sRGB_degamma(e1);
            e2 = sRGB_degamma(e2);
            [[fallthrough]];
        case ColorSpace::Rec709_D65: {
            float RGB[3] = {e0, e1, e2};
            if (spType == SpectrumType::LightSource)
                transformTristimulus(mat_Rec709_D65_to_XYZ, RGB, XYZ);
            else
                transformTristimulus(mat_Rec709_E_to_XYZ, RGB, XYZ);
            break;
        }
        case ColorSpace::XYZ:
            XYZ[0] = e0; XYZ[1] = e1; XYZ[2] = e2;
            break;
        case ColorSpace::xyY: {
            if (e1 == 0) {
                XYZ[0] = XYZ[1] = XYZ[2] = 0;
                break;
            }
            float z = 1 - (e0 + e1);
            float b = e2 / e1;
            XYZ[0] = e0 * b;
            XYZ[1] = e2;
            XYZ[2] = z * b;
            break;
        }
        }
        if (spType == SpectrumType::LightSource)
            transformTristimulus(mat_XYZ_to_Rec709_D65, XYZ, t.rgb);
        else
            transformTristimulus(mat_XYZ_to_Rec709_E, XYZ, t.rgb);
        return t;
    }
    template <int N> SS<N> eval(const WLS<N> &wls) const;
};
template <> inline SS<3> Triplet::eval<3>(const WLS<3> &) const {
    SS<3> r; r[0] = rgb[0]; r[1] = rgb[1]; r[2] = rgb[2]; return r;
}
template <> inline SS<4> Triplet::eval<4>(const WLS<4> &wls) const { return meng.evaluate(wls); }

// ---------------------------------------------------------------------------------------
// Per-pixel storage with Kahan summation.
template <typename T>
struct Kahan {
    T result, comp;
    void reset() { result = T(0.0f); comp = T(0.0f); }
    void add(const T &value) {
        T cInput = value - comp;
        T sumTemp = result + cInput;
        comp = (sumTemp - result) - cInput;
        result = sumTemp;
    }
};

// SpectrumStorageTemplate::add (spectrum_types.h:674-684)
inline SS<16> binSpectral(const WLS<4> &wls, const SS<4> &val) {
    const float recBinWidth = kNumStrata / (kWLHigh - kWLLow);
    SS<16> addend(0.0f);
    for (int i = 0; i < 4; ++i) {
        uint32_t sBin = std::min<uint32_t>(static_cast<uint32_t>((wls.lambdas[i] - kWLLow) / (kWLHigh - kWLLow) * kNumStrata), kNumStrata - 1);
        addend[sBin] += val[i] * recBinWidth;
    }
    return addend;
}

// DiscretizedSpectrum::toXYZ (spectrum_types.cpp:605-616)
inline void strataToXYZ(const SS<16> &s, float XYZ[3]) {
    const SpectralTables &T = tables();
    XYZ[0] = XYZ[1] = XYZ[2] = 0;
    for (int i = 0; i < kNumStrata; ++i) {
        XYZ[0] += T.binX[i] * s[i];
        XYZ[1] += T.binY[i] * s[i];
        XYZ[2] += T.binZ[i] * s[i];
    }
    XYZ[0] /= T.integralCMF16;
    XYZ[1] /= T.integralCMF16;
    XYZ[2] /= T.integralCMF16;
}

} // namespace ov
