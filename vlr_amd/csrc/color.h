// Colour conversions shared by the host (immediate spectra at material set-up) and the device
// (per-hit spectra from shader nodes): RGB / XYZ / xyY triplets to the RGB build's rendering
// RGB (libVLR/shared/shared.h:60-118) or to the Meng et al. 2015 upsampled spectrum of the
// spectral build (UpsampledSpectrum's constructor and computeAdjacents,
// shared/spectrum_types.cpp:11-91, 163-228).
//
// A spectral result is a DevSpectrum: the grid-point indices packed in x's bits, the
// barycentric or bilinear weights as 16-bit fractions in y's bits, the scale in z.
#pragma once

#include <hip/hip_runtime.h>

#include "dev_types.h"

namespace vlra {

// spectrum types and colour spaces as the C API enumerates them
constexpr uint32_t kSpecReflectance = 0, kSpecLightSource = 1, kSpecIndexOfRefraction = 2, kSpecNA = 3;
constexpr uint32_t kCsRec709Gamma = 0, kCsRec709 = 1, kCsXYZ = 2, kCsxyY = 3;

struct MengTables {
    const MengCellDev *grid;   // 12 x 14 cells
    const float *points;       // 99 floats per point: x, y, u, v, then 95 spectrum samples
};

__host__ __device__ inline void colorTristim(uint32_t which, const float s[3], float d[3]) {
    // 0 Rec.709 D65 -> XYZ, 1 XYZ -> Rec.709 D65, 2 Rec.709 E -> XYZ, 3 XYZ -> Rec.709 E
    // (spectrum_base.h:109-128, column-major)
    const float m[4][9] = {{0.4124564f, 0.2126729f, 0.0193339f, 0.3575761f, 0.7151522f, 0.1191920f, 0.1804375f,
                            0.0721750f, 0.9503041f},
                           {3.2404542f, -0.9692660f, 0.0556434f, -1.5371385f, 1.8760108f, -0.2040259f, -0.4985314f,
                            0.0415560f, 1.0572252f},
                           {0.4969f, 0.2562f, 0.0233f, 0.3391f, 0.6782f, 0.1130f, 0.1640f, 0.0656f, 0.8637f},
                           {2.6897f, -1.0221f, 0.0612f, -1.2759f, 1.9783f, -0.2245f, -0.4138f, 0.0438f, 1.1633f}};
    const float *M = m[which];
    d[0] = M[0] * s[0] + M[3] * s[1] + M[6] * s[2];
    d[1] = M[1] * s[0] + M[4] * s[1] + M[7] * s[2];
    d[2] = M[2] * s[0] + M[5] * s[1] + M[8] * s[2];
}

__host__ __device__ inline float colorDegamma(float v) {
    return v <= 0.04045f ? v / 12.92f : powf((v + 0.055f) / 1.055f, 2.4f);
}

__host__ __device__ inline void mengXYtoUV(const float xy[2], float uv[2]) {
    uv[0] = 16.730260708356887f * xy[0] + 7.7801960340706f * xy[1] - 2.170152247475828f;
    uv[1] = -7.530081094743006f * xy[0] + 16.192422314095225f * xy[1] + 1.1125529268825947f;
}

// Grid lookup of (u, v): adjacent points and weights (spectrum_types.cpp:11-91).
__host__ __device__ inline void mengAdjacents(const MengTables &T, float u, float v, uint32_t *adj, uint16_t *s,
                                              uint16_t *t) {
    u = fminf(fmaxf(u, 0.0f), 12.0f);
    v = fminf(fmaxf(v, 0.0f), 14.0f);
    int32_t ui = static_cast<int32_t>(u), vi = static_cast<int32_t>(v);
    ui = ui < 11 ? ui : 11;
    vi = vi < 13 ? vi : 13;
    const MengCellDev &cell = T.grid[ui + 12 * vi];
    const uint8_t *ix = cell.idx;
    const float *P = T.points;
    *adj = 0;
    *s = *t = 0;
    if (cell.inside) {
        const float ss = u - ui, tt = v - vi;
        *adj = (uint32_t(ix[3]) << 24) | (uint32_t(ix[2]) << 16) | (uint32_t(ix[1]) << 8) | uint32_t(ix[0]);
        *s = static_cast<uint16_t>(fmaxf(ss * (65535 - 1), 0.0f));
        *t = static_cast<uint16_t>(fmaxf(tt * (65535 - 1), 0.0f));
        return;
    }
    // cells with no data points lie outside the spectral locus; keep adj = 0
    if (cell.numPoints < 3) return;
    const float ex = u - P[ix[0] * 99 + 2], ey = v - P[ix[0] * 99 + 3];
    float e0x = P[ix[1] * 99 + 2] - P[ix[0] * 99 + 2];
    float e0y = P[ix[1] * 99 + 3] - P[ix[0] * 99 + 3];
    float uu = e0x * ey - ex * e0y;
    for (int i = 1; i < cell.numPoints; ++i) {
        const uint32_t idx = ix[i % (cell.numPoints - 1) + 1];
        const float e1x = P[idx * 99 + 2] - P[ix[0] * 99 + 2];
        const float e1y = P[idx * 99 + 3] - P[ix[0] * 99 + 3];
        const float vv = ex * e1y - e1x * ey;
        const float area = e0x * e1y - e1x * e0y;
        const float b0 = uu / area, b1 = vv / area;
        const float b2 = 1.0f - b0 - b1;
        if (b0 < -1e-6f || b1 < -1e-6f || b2 < -1e-6f) {
            uu = -vv;
            e0x = e1x;
            e0y = e1y;
            continue;
        }
        *adj = (uint32_t(255) << 24) | (uint32_t(ix[0]) << 16) | (uint32_t(ix[i]) << 8) | idx;
        *s = static_cast<uint16_t>(fmaxf(b0 * (65535 - 1), 0.0f));
        *t = static_cast<uint16_t>(fmaxf(b1 * (65535 - 1), 0.0f));
        break;
    }
}

// Upsampled spectrum of grid coordinates (u, v) and scale (UpsampledSpectrum(u, v, s)).
__host__ __device__ inline DevSpectrum mengFromUV(const MengTables &T, float u, float v, float scale) {
    uint32_t adj;
    uint16_t s, t;
    mengAdjacents(T, u, v, &adj, &s, &t);
    const uint32_t st = uint32_t(s) | (uint32_t(t) << 16);
    DevSpectrum out;
    union {
        uint32_t u;
        float f;
    } a, b;
    a.u = adj;
    b.u = st;
    out.x = a.f;
    out.y = b.f;
    out.z = scale;
    out.w = 0.0f;
    return out;
}

// A colour triplet as a device spectrum parameter: rendering RGB in the RGB build, an
// upsampled spectrum in the spectral build.
__host__ __device__ inline DevSpectrum colorToDevSpectrum(uint32_t type, uint32_t cs, float e0, float e1, float e2,
                                                          bool spectral, const MengTables &T) {
    DevSpectrum out = {0, 0, 0, 0};
    const bool light = type == kSpecLightSource;
    if (!spectral) {
        float XYZ[3] = {0, 0, 0};
        if (cs == kCsRec709Gamma) {
            e0 = colorDegamma(e0);
            e1 = colorDegamma(e1);
            e2 = colorDegamma(e2);
        }
        if (cs == kCsRec709Gamma || cs == kCsRec709) {
            const float RGB[3] = {e0, e1, e2};
            colorTristim(light ? 0 : 2, RGB, XYZ);
        } else if (cs == kCsXYZ) {
            XYZ[0] = e0;
            XYZ[1] = e1;
            XYZ[2] = e2;
        } else if (e1 != 0) {
            const float z = 1 - (e0 + e1), b = e2 / e1;
            XYZ[0] = e0 * b;
            XYZ[1] = e2;
            XYZ[2] = z * b;
        }
        float RGB[3];
        colorTristim(light ? 1 : 3, XYZ, RGB);
        out.x = RGB[0];
        out.y = RGB[1];
        out.z = RGB[2];
        return out;
    }
    if (cs == kCsRec709Gamma) {
        e0 = colorDegamma(e0);
        e1 = colorDegamma(e1);
        e2 = colorDegamma(e2);
    }
    if (cs == kCsRec709Gamma || cs == kCsRec709) {
        const float RGB[3] = {e0, e1, e2};
        float XYZ[3];
        colorTristim(light ? 0 : 2, RGB, XYZ);
        e0 = XYZ[0];
        e1 = XYZ[1];
        e2 = XYZ[2];
    }
    if (cs != kCsxyY) {
        const float Y = e1, b = e0 + e1 + e2;
        e0 = e0 / b;
        e1 = e1 / b;
        e2 = Y;
    }
    float xy[2] = {e0, e1};
    float brightness = e2 / e1;
    if (e2 == 0) {
        xy[0] = 0.3333f;
        xy[1] = 0.3333f;
        brightness = 0;
    }
    const float scale = brightness / 0.009355121400914532f;
    float uv[2];
    mengXYtoUV(xy, uv);
    return mengFromUV(T, uv[0], uv[1], scale);
}

} // namespace vlra
