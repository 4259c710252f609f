// Device spectra, Meng upsampling and the flattened BSDF / EDF switch (gfx950).
//
// Each material type of the reference's callable tables becomes one case of a switch that
// the shade kernel inlines; the per-type arithmetic follows libVLR/GPU_kernels/materials.cu:
//   Fresnel conductor/dielectric :13-129, GGX :133-228, MatteBRDF :584-663,
//   SpecularBRDF :667-789, SpecularBSDF :793-1060, MicrofacetBRDF :1064-1242,
//   DiffuseAndSpecularBRDF (UE4) :1692-1994, DiffuseEDF :2040-2096, PointEDF :2194-2250,
//   EnvironmentEDF :2830-2896; shading-normal correction: shared/renderer_common.h:88-163.
// Spectra: RGB build N = 3, spectral build N = 4 hero wavelengths in [360, 830) nm
// (shared/spectrum_types.h:55-65), importance weighting 0.9 / 0.1 (spectrum_types.h:250-264).
#pragma once
#include "dev_math.h"
#include "dev_types.h"

namespace vlra {

template <int N>
struct Spec {
    float v[N];
    __device__ __forceinline__ Spec() {}
    __device__ __forceinline__ explicit Spec(float s) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = s;
    }
    __device__ __forceinline__ float &operator[](int i) { return v[i]; }
    __device__ __forceinline__ float operator[](int i) const { return v[i]; }
#define VLRA_SPEC_OP(op)                                                                               \
    __device__ __forceinline__ Spec operator op(const Spec &o) const {                                 \
        Spec r;                                                                                        \
        _Pragma("unroll") for (int i = 0; i < N; ++i) r.v[i] = v[i] op o.v[i];                         \
        return r;                                                                                      \
    }                                                                                                  \
    __device__ __forceinline__ Spec &operator op##=(const Spec &o) {                                   \
        _Pragma("unroll") for (int i = 0; i < N; ++i) v[i] op##= o.v[i];                               \
        return *this;                                                                                  \
    }
    VLRA_SPEC_OP(+)
    VLRA_SPEC_OP(-)
    VLRA_SPEC_OP(*)
    VLRA_SPEC_OP(/)
#undef VLRA_SPEC_OP
    __device__ __forceinline__ Spec operator*(float s) const {
        Spec r;
#pragma unroll
        for (int i = 0; i < N; ++i) r.v[i] = v[i] * s;
        return r;
    }
    __device__ __forceinline__ Spec operator/(float s) const { return (*this) * (1.0f / s); }
    __device__ __forceinline__ Spec operator+(float s) const {
        Spec r;
#pragma unroll
        for (int i = 0; i < N; ++i) r.v[i] = v[i] + s;
        return r;
    }
    __device__ __forceinline__ Spec operator-(float s) const {
        Spec r;
#pragma unroll
        for (int i = 0; i < N; ++i) r.v[i] = v[i] - s;
        return r;
    }
    __device__ __forceinline__ bool isZero() const {
        bool z = true;
#pragma unroll
        for (int i = 0; i < N; ++i) z = z && (v[i] == 0.0f);
        return z;
    }
    __device__ __forceinline__ bool allFinite() const {
        bool f = true;
#pragma unroll
        for (int i = 0; i < N; ++i) f = f && isfinite(v[i]);
        return f;
    }
    __device__ __forceinline__ float importance(uint32_t sel) const {
        float sum = 0;
        float picked = v[0];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            sum += v[i];
            if (i == static_cast<int>(sel)) picked = v[i];
        }
        const float primary = 0.9f;
        const float marginal = (1 - primary) / (N - 1);
        return sum * marginal + picked * (primary - marginal);
    }
    __device__ __forceinline__ float get(uint32_t sel) const {
        float r = v[0];
#pragma unroll
        for (int i = 1; i < N; ++i) if (i == static_cast<int>(sel)) r = v[i];
        return r;
    }
    __device__ __forceinline__ void set(uint32_t sel, float x) {
#pragma unroll
        for (int i = 0; i < N; ++i) if (i == static_cast<int>(sel)) v[i] = x;
    }
};
template <int N> __device__ __forceinline__ Spec<N> operator*(float s, const Spec<N> &c) { return c * s; }
template <int N> __device__ __forceinline__ Spec<N> lerpSpec(const Spec<N> &a, const Spec<N> &b, float t) {
    return a * (1 - t) + b * t;
}

template <int N>
struct Wls {
    float offset;
    uint32_t sel;
    bool single;
    __device__ __forceinline__ float lambda(int i) const { return 360.0f + (830.0f - 360.0f) * (i + offset) / N; }
};

// Meng et al. 2015 evaluation (spectrum_types.cpp:273-336)
__device__ __forceinline__ Spec<4> mengEvaluate(const DevScene &sc, uint32_t adj, uint32_t st, float scale,
                                                const Wls<4> &wls) {
    uint32_t a0 = adj & 0xFF, a1 = (adj >> 8) & 0xFF, a2 = (adj >> 16) & 0xFF, a3 = (adj >> 24) & 0xFF;
    float sf = static_cast<float>(st & 0xFFFF) / (65535 - 1);
    float tf = static_cast<float>(st >> 16) / (65535 - 1);
    float w0, w1, w2, w3;
    bool quad = a3 != 0xFF;
    if (quad) {
        w0 = (1 - sf) * (1 - tf);
        w1 = sf * (1 - tf);
        w2 = (1 - sf) * tf;
        w3 = sf * tf;
    } else {
        w0 = sf;
        w1 = tf;
        w2 = 1.0f - sf - tf;
        w3 = 0.0f;
    }
    const float *P = sc.mengPoints;
    Spec<4> ret(0.0f);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float p = (wls.lambda(i) - 360.0f) / (830.0f - 360.0f);
        p = fminf(fmaxf(p, 0.0f), 1.0f);
        float sBinF = p * (95 - 1);
        uint32_t sBin = min(static_cast<uint32_t>(sBinF), 94u);
        uint32_t sNext = min(sBin + 1, 94u);
        float t = sBinF - sBin;
        const float *s0 = P + a0 * 99 + 4, *s1 = P + a1 * 99 + 4, *s2 = P + a2 * 99 + 4;
        float acc = 0.0f;
        acc += w0 * (s0[sBin] * (1 - t) + s0[sNext] * t);
        acc += w1 * (s1[sBin] * (1 - t) + s1[sNext] * t);
        acc += w2 * (s2[sBin] * (1 - t) + s2[sNext] * t);
        if (quad) {
            const float *s3 = P + a3 * 99 + 4;
            acc += w3 * (s3[sBin] * (1 - t) + s3[sNext] * t);
        }
        ret[i] = acc;
    }
    return ret * scale;
}

// Meng grid lookup (spectrum_types.cpp:11-91) for per-texel (u, v, s) environment values.
__device__ __forceinline__ void mengAdjacents(const DevScene &sc, float u, float v, uint32_t *adjOut, uint32_t *stOut) {
    u = fminf(fmaxf(u, 0.0f), 12.0f);
    v = fminf(fmaxf(v, 0.0f), 14.0f);
    // u == 12 or v == 14 would index one cell past the grid; stay on the last cell
    int32_t ui = min(static_cast<int32_t>(u), 11), vi = min(static_cast<int32_t>(v), 13);
    const MengCellDev cell = sc.mengGrid[ui + 12 * vi];
    uint32_t adj = 0, st = 0;
    if (cell.inside) {
        float s = u - ui, t = v - vi;
        adj = (uint32_t(cell.idx[3]) << 24) | (uint32_t(cell.idx[2]) << 16) | (uint32_t(cell.idx[1]) << 8) | cell.idx[0];
        uint32_t ss = static_cast<uint16_t>(fmaxf(s * (65535 - 1), 0.0f));
        uint32_t tt = static_cast<uint16_t>(fmaxf(t * (65535 - 1), 0.0f));
        st = ss | (tt << 16);
    } else if (cell.numPoints >= 3) {
        // cells with no data points lie outside the spectral locus; their (u, v) keeps adj = 0
        const float *P = sc.mengPoints;
        const uint32_t i0 = cell.idx[0];
        const float ex = u - P[i0 * 99 + 2];
        const float ey = v - P[i0 * 99 + 3];
        float e0x = P[cell.idx[1] * 99 + 2] - P[i0 * 99 + 2];
        float e0y = P[cell.idx[1] * 99 + 3] - P[i0 * 99 + 3];
        float uu = e0x * ey - ex * e0y;
        for (int i = 1; i < cell.numPoints; ++i) {
            uint32_t idx = cell.idx[i % (cell.numPoints - 1) + 1];
            float e1x = P[idx * 99 + 2] - P[i0 * 99 + 2];
            float e1y = P[idx * 99 + 3] - P[i0 * 99 + 3];
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
            adj = (0xFFu << 24) | (i0 << 16) | (uint32_t(cell.idx[i]) << 8) | idx;
            uint32_t ss = static_cast<uint16_t>(fmaxf(b0 * (65535 - 1), 0.0f));
            uint32_t tt = static_cast<uint16_t>(fmaxf(b1 * (65535 - 1), 0.0f));
            st = ss | (tt << 16);
            break;
        }
    }
    *adjOut = adj;
    *stOut = st;
}

template <int N>
__device__ __forceinline__ Spec<N> evalSpectrum(const DevScene &sc, const DevSpectrum &p, const Wls<N> &wls) {
    if constexpr (N == 3) {
        Spec<3> r;
        r[0] = p.x; r[1] = p.y; r[2] = p.z;
        return r;
    } else {
        return mengEvaluate(sc, __float_as_uint(p.x), __float_as_uint(p.y), p.z, wls);
    }
}

// ---------------------------------------------------------------------------------------------
namespace dt {
constexpr uint32_t LowFreq = 1u << 0, HighFreq = 1u << 1, Delta0D = 1u << 2, Delta1D = 1u << 3;
constexpr uint32_t NonDelta = LowFreq | HighFreq, Delta = Delta0D | Delta1D, AllFreq = NonDelta | Delta;
constexpr uint32_t Reflection = 1u << 4, Transmission = 1u << 5, WholeSphere = Reflection | Transmission;
constexpr uint32_t All = AllFreq | WholeSphere, Dispersive = 1u << 6;
__device__ __forceinline__ bool matches(uint32_t self, uint32_t t) {
    uint32_t r = self & t;
    return (r & WholeSphere) && (r & AllFreq);
}
__device__ __forceinline__ bool isDelta(uint32_t v) { return (v & Delta) && !(v & NonDelta); }
} // namespace dt

// random_distributions.h:80-124
__device__ __forceinline__ void concentricSampleDisk(float u0, float u1, float *dx, float *dy) {
    float r, theta;
    float sx = 2 * u0 - 1;
    float sy = 2 * u1 - 1;
    if (sx == 0 && sy == 0) {
        *dx = 0;
        *dy = 0;
        return;
    }
    if (sx >= -sy) {
        if (sx > sy) { r = sx; theta = sy / sx; }
        else { r = sy; theta = 2 - sx / sy; }
    } else {
        if (sx > sy) { r = -sy; theta = 6 + sx / sy; }
        else { r = -sx; theta = 4 + sy / sx; }
    }
    theta *= kPiF / 4;
    float s, c;
    sincosf(theta, &s, &c);
    *dx = r * c;
    *dy = r * s;
}
__device__ __forceinline__ V3 cosineSampleHemisphere(float u0, float u1) {
    float x, y;
    concentricSampleDisk(u0, u1, &x, &y);
    return V3(x, y, sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y)));
}
// random_distributions.h: uniform direction on the sphere
__device__ __forceinline__ V3 uniformSampleSphere(float u0, float u1) {
    float phi = 2 * kPiF * u1;
    float theta = acosf(1 - 2 * u0);
    return fromPolarZUp(phi, theta);
}
__device__ __forceinline__ void uniformSampleTriangle(float u0, float u1, float *b0, float *b1) {
    *b0 = 0.5f * u0;
    *b1 = 0.5f * u1;
    float offset = *b1 - *b0;
    if (offset > 0) *b1 += offset;
    else *b0 -= offset;
}

template <int N>
__device__ __forceinline__ Spec<N> fresnelConductor(const Spec<N> &eta, const Spec<N> &k, float cosEnter) {
    cosEnter = fabsf(cosEnter);
    float c2 = cosEnter * cosEnter;
    Spec<N> r;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float e2c = 2.0f * eta[i] * cosEnter;
        float tf = eta[i] * eta[i] + k[i] * k[i];
        float tmp = tf * c2;
        float Rparl2 = (tmp - e2c + 1.0f) / (tmp + e2c + 1.0f);
        float Rperp2 = (tf - e2c + c2) / (tf + e2c + c2);
        r[i] = (Rparl2 + Rperp2) / 2.0f;
    }
    return r;
}

__device__ __forceinline__ float fresnelDielectricF(float etaEnter, float etaExit, float cosEnter, float cosExit) {
    float Rparl = ((etaExit * cosEnter) - (etaEnter * cosExit)) / ((etaExit * cosEnter) + (etaEnter * cosExit));
    float Rperp = ((etaEnter * cosEnter) - (etaExit * cosExit)) / ((etaEnter * cosEnter) + (etaExit * cosExit));
    return (Rparl * Rparl + Rperp * Rperp) / 2.0f;
}

// FresnelDielectric::evaluate with eEnter/eExit already chosen for a positive cosine
template <int N>
__device__ __forceinline__ Spec<N> fresnelDielectric(const Spec<N> &etaExt, const Spec<N> &etaInt, float cosEnter) {
    cosEnter = fminf(fmaxf(cosEnter, -1.0f), 1.0f);
    bool entering = cosEnter > 0.0f;
    float sinEnter = sqrtf(fmaxf(0.0f, 1.0f - cosEnter * cosEnter));
    cosEnter = fabsf(cosEnter);
    Spec<N> r;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float eEnter = entering ? etaExt[i] : etaInt[i];
        float eExit = entering ? etaInt[i] : etaExt[i];
        float sinExit = eEnter / eExit * sinEnter;
        if (sinExit >= 1.0f) {
            r[i] = 1.0f;
        } else {
            float cosExit = sqrtf(fmaxf(0.0f, 1.0f - sinExit * sinExit));
            r[i] = fresnelDielectricF(eEnter, eExit, cosEnter, cosExit);
        }
    }
    return r;
}

// FresnelDielectric::evaluate(cosEnter, wlIdx) for one wavelength (materials.cu:79-95)
__device__ __forceinline__ float fresnelDielectric1(float etaExt, float etaInt, float cosEnter) {
    cosEnter = fminf(fmaxf(cosEnter, -1.0f), 1.0f);
    bool entering = cosEnter > 0.0f;
    float eEnter = entering ? etaExt : etaInt;
    float eExit = entering ? etaInt : etaExt;
    float sinExit = eEnter / eExit * sqrtf(fmaxf(0.0f, 1.0f - cosEnter * cosEnter));
    cosEnter = fabsf(cosEnter);
    if (sinExit >= 1.0f) return 1.0f;
    float cosExit = sqrtf(fmaxf(0.0f, 1.0f - sinExit * sinExit));
    return fresnelDielectricF(eEnter, eExit, cosEnter, cosExit);
}

struct Ggx {
    float ax, ay, cr, sr;
    __device__ __forceinline__ Ggx(float alphaX, float alphaY, float rotation) : ax(alphaX), ay(alphaY) {
        sincosf(rotation, &sr, &cr);
    }
    __device__ __forceinline__ V3 rot(const V3 &m) const { return V3(cr * m.x + sr * m.y, -sr * m.x + cr * m.y, m.z); }
    __device__ __forceinline__ float D(const V3 &m) const {
        V3 mr = rot(m);
        if (mr.z <= 0) return 0.0f;
        float temp = sq(mr.x / ax) + sq(mr.y / ay) + sq(mr.z);
        return 1.0f / (kPiF * ax * ay * sq(temp));
    }
    __device__ __forceinline__ float lambda(const V3 &v) const {
        V3 vr = rot(v);
        float a2t2 = (sq(vr.x * ax) + sq(vr.y * ay)) / sq(vr.z);
        return (-1 + sqrtf(1 + a2t2)) / 2;
    }
    __device__ __forceinline__ float G1(const V3 &v, const V3 &m) const {
        float chi = (dot(v, m) / v.z) > 0 ? 1.0f : 0.0f;
        return chi / (1 + lambda(v));
    }
    __device__ __forceinline__ float Gcorr(const V3 &v1, const V3 &v2, const V3 &m) const {
        float chi1 = (dot(v1, m) / v1.z) > 0 ? 1.0f : 0.0f;
        float chi2 = (dot(v2, m) / v2.z) > 0 ? 1.0f : 0.0f;
        return chi1 * chi2 / (1 + lambda(v1) + lambda(v2));
    }
    // visible-normal sampling (materials.cu:184-223)
    __device__ __forceinline__ float sample(const V3 &v, float u0, float u1, V3 *m, float *normalPDF) const {
        V3 vr = rot(v);
        V3 sv = normalize(V3(ax * vr.x, ay * vr.y, vr.z));
        float dist2D = sqrtf(sv.x * sv.x + sv.y * sv.y);
        float rd = 1.0f / dist2D;
        V3 T1 = (sv.z < 0.9999f) ? V3(sv.y * rd, -sv.x * rd, 0) : V3(1, 0, 0);
        V3 T2(T1.y * sv.z, -T1.x * sv.z, dist2D);
        float a = 1.0f / (1.0f + sv.z);
        float r = sqrtf(u0);
        float phi = kPiF * ((u1 < a) ? u1 / a : 1 + (u1 - a) / (1.0f - a));
        float sp, cp;
        sincosf(phi, &sp, &cp);
        float P1 = r * cp;
        float P2 = r * sp * ((u1 < a) ? 1.0f : sv.z);
        V3 mr = P1 * T1 + P2 * T2 + sqrtf(1.0f - P1 * P1 - P2 * P2) * sv;
        mr = normalize(V3(ax * mr.x, ay * mr.y, mr.z));
        float d = D(mr);
        float chi = (dot(vr, mr) / vr.z) > 0 ? 1.0f : 0.0f;
        *normalPDF = chi / (1 + lambda(vr)) * absDot(vr, mr) * d / fabsf(vr.z);
        *m = V3(cr * mr.x - sr * mr.y, sr * mr.x + cr * mr.y, mr.z);
        return d;
    }
    __device__ __forceinline__ float pdf(const V3 &v, const V3 &m) const {
        return G1(v, m) * absDot(v, m) * D(m) / fabsf(v.z);
    }
};

enum BsdfKind : uint32_t {
    BSDF_NULL = 0, BSDF_MATTE, BSDF_SPEC_BRDF, BSDF_SPEC_BSDF, BSDF_MICRO_BRDF, BSDF_DIFFSPEC, BSDF_MICRO_BSDF,
    // an emitter's directional distribution used as a BSDF at the first light-path vertex
    // (the EDF_as_BSDF_* callables of materials.cu:415-481); a0 = world disc area for the environment
    BSDF_EDF_DIFFUSE, BSDF_EDF_POINT, BSDF_EDF_ENV
};

// Reverse (adjoint-direction) quantities of a BSDF query, used by bidirectional estimators.
template <int N>
struct BsdfRev {
    Spec<N> value;
    float pdf;
};

template <int N>
struct Bsdf {
    uint32_t kind;
    Spec<N> c0, c1, c2;
    float a0, a1, a2;
    bool dispersive;

    __device__ __forceinline__ uint32_t typeMask() const {
        switch (kind) {
        case BSDF_MATTE: return dt::Reflection | dt::LowFreq;
        case BSDF_SPEC_BRDF: return dt::Reflection | dt::Delta0D;
        case BSDF_SPEC_BSDF: return dt::WholeSphere | dt::Delta0D;
        case BSDF_MICRO_BRDF: return dt::Reflection | dt::HighFreq;
        case BSDF_DIFFSPEC: return dt::Reflection | dt::LowFreq | dt::HighFreq;
        case BSDF_MICRO_BSDF: return dt::WholeSphere | dt::HighFreq;
        case BSDF_EDF_DIFFUSE: case BSDF_EDF_POINT: case BSDF_EDF_ENV: return dt::Reflection | dt::LowFreq;
        default: return 0u;
        }
    }
    __device__ __forceinline__ bool matches(uint32_t flags) const { return dt::matches(typeMask(), flags); }
    __device__ __forceinline__ bool hasNonDelta() const { return matches(dt::WholeSphere | dt::NonDelta); }
    // getBaseColor of each BSDF (materials.cu:537-539, 592-594, 677-679, 805-807, 1077-1081,
    // 1551-1553, 1702-1704; an emitter's EDF as a BSDF, 416-419): the denoiser's albedo guide
    __device__ __forceinline__ Spec<N> baseColor() const {
        switch (kind) {
        case BSDF_MATTE: case BSDF_SPEC_BRDF: case BSDF_SPEC_BSDF: case BSDF_MICRO_BSDF: return c0;
        case BSDF_MICRO_BRDF: return fresnelConductor<N>(c0, c1, 1.0f);
        case BSDF_DIFFSPEC: return c0 + c1;
        default: return Spec<N>(0.0f);
        }
    }

    // UE4 selection weights (materials.cu:1724-1734) for a query direction of cosine qz
    __device__ __forceinline__ void diffSpecWeights(float qz, float vz, uint32_t wl, float *wd, float *ws) const {
        const float rough = a0;
        float eF = 0.5f * rough + 2 * rough * sq(qz);
        float om5 = pow5f(1 - vz);
        float edf = lerpf(1.0f, eF, om5);
        *wd = c0.importance(wl) * sq(edf) * lerpf(1.0f, 1.0f / 1.51f, rough);
        *ws = lerpf(c1.importance(wl), 1.0f, om5);
    }
    // reverse pdf of the UE4 lobes for the light direction l
    __device__ __forceinline__ float diffSpecRevPDF(const Ggx &g, const V3 &v, const V3 &l, const V3 &m, float dotLH,
                                                     uint32_t wl) const {
        float revDiffuse = v.z / kPiF;
        float revSpecular = g.pdf(l, m) / (4 * dotLH);
        float rF_D90 = 0.5f * a0 + 2 * a0 * sq(l.z);
        float rDiffFresnel = lerpf(1.0f, rF_D90, pow5f(1 - l.z));
        float rDiffW = c0.importance(wl) * sq(rDiffFresnel) * lerpf(1.0f, 1.0f / 1.51f, a0);
        float rSpecW = lerpf(c1.importance(wl), 1.0f, pow5f(1 - l.z));
        return (revDiffuse * rDiffW + revSpecular * rSpecW) / (rDiffW + rSpecW);
    }
    // one wavelength of the rough-dielectric transmission term (materials.cu:1352-1371)
    __device__ __forceinline__ float roughTransmission(const Spec<N> &eEnter, const Spec<N> &eExit,
                                                       const Spec<N> &fEnter, const Spec<N> &fExit, const Ggx &g,
                                                       const V3 &v, const V3 &l, bool entering, int i) const {
        V3 m = normalize(-(eEnter[i] * v + eExit[i] * l) * (entering ? 1.0f : -1.0f));
        float dotHV = dot(v, m), dotHL = dot(l, m);
        float F = fresnelDielectric1(fEnter[i], fExit[i], dotHV);
        float G = g.G1(v, m) * g.G1(l, m);
        return fabsf(dotHV * dotHL) * (1 - F) * G * g.D(m) / sq(eEnter[i] * dotHV + eExit[i] * dotHL);
    }

    __device__ Spec<N> sampleInternal(const V3 &q, uint32_t wl, bool importance, float uComponent, float uD0,
                                      float uD1, V3 *dir, float *dirPDF, uint32_t *type, BsdfRev<N> *rev) const {
        *dirPDF = 0.0f;
        *type = 0;
        if (rev) rev->pdf = 0.0f;
        switch (kind) {
        case BSDF_MATTE: {
            V3 d = cosineSampleHemisphere(uD0, uD1);
            *dirPDF = d.z / kPiF;
            *type = dt::Reflection | dt::LowFreq;
            d.z *= q.z >= 0 ? 1.0f : -1.0f;
            *dir = d;
            Spec<N> fs = c0 / kPiF;
            if (rev) {
                rev->value = fs;
                rev->pdf = fabsf(q.z) / kPiF;
            }
            return fs;
        }
        case BSDF_SPEC_BRDF: {
            *dir = V3(-q.x, -q.y, q.z);
            *dirPDF = 1.0f;
            *type = dt::Reflection | dt::Delta0D;
            Spec<N> ret = c0 * fresnelConductor<N>(c1, c2, q.z) * (1.0f / fabsf(q.z));
            if (rev) {
                rev->value = ret;
                rev->pdf = 1.0f;
            }
            return ret;
        }
        case BSDF_SPEC_BSDF: {
            bool entering = q.z >= 0.0f;
            const Spec<N> &eEnter = entering ? c1 : c2;
            const Spec<N> &eExit = entering ? c2 : c1;
            V3 v = entering ? q : -q;
            Spec<N> F = fresnelDielectric<N>(eEnter, eExit, v.z);
            float reflectProb = F.importance(wl);
            if (uComponent < reflectProb) {
                if (v.z == 0.0f) return Spec<N>(0.0f);
                V3 l(-v.x, -v.y, v.z);
                *dir = entering ? l : -l;
                *dirPDF = reflectProb;
                *type = dt::Reflection | dt::Delta0D;
                Spec<N> ret = c0 * F * (1.0f / fabsf(v.z));
                if (rev) {
                    rev->value = ret;
                    rev->pdf = reflectProb;
                }
                return ret;
            }
            float sinEnter2 = 1.0f - v.z * v.z;
            float rr = eEnter.get(wl) / eExit.get(wl);
            float sinExit2 = rr * rr * sinEnter2;
            if (sinExit2 >= 1.0f) return Spec<N>(0.0f);
            float cosExit = sqrtf(fmaxf(0.0f, 1.0f - sinExit2));
            V3 l(rr * -v.x, rr * -v.y, -cosExit);
            *dir = entering ? l : -l;
            *dirPDF = 1.0f - reflectProb;
            *type = dt::Transmission | dt::Delta0D | (dispersive ? dt::Dispersive : 0u);
            Spec<N> ret(0.0f);
            ret.set(wl, c0.get(wl) * (1.0f - F.get(wl)));
            if (rev) {
                float revSqueeze = importance ? sq(eExit.get(wl) / eEnter.get(wl)) : 1.0f;
                Spec<N> revRet = ret;
                revRet.set(wl, revRet.get(wl) * (revSqueeze / fabsf(q.z)));
                rev->value = revRet;
                rev->pdf = revSqueeze * *dirPDF;
            }
            float squeeze = importance ? 1.0f : sq(eEnter.get(wl) / eExit.get(wl));
            ret.set(wl, ret.get(wl) * (squeeze / fabsf(cosExit)));
            *dirPDF *= squeeze;
            return ret;
        }
        case BSDF_MICRO_BRDF: {
            bool entering = q.z >= 0.0f;
            Ggx g(a0, a1, a2);
            V3 v = entering ? q : -q;
            V3 m;
            float mPDF;
            float D = g.sample(v, uD0, uD1, &m, &mPDF);
            float dotHV = dot(v, m);
            if (dotHV <= 0) return Spec<N>(0.0f);
            V3 l = 2 * dotHV * m - v;
            *dir = entering ? l : -l;
            if (l.z * v.z <= 0) return Spec<N>(0.0f);
            float common = 1.0f / (4 * dotHV);
            *dirPDF = common * mPDF;
            *type = dt::Reflection | dt::HighFreq;
            float G = g.G1(v, m) * g.G1(l, m);
            Spec<N> fs = fresnelConductor<N>(c0, c1, dotHV) * D * G / (4 * v.z * l.z);
            if (rev) {
                rev->value = fs;
                rev->pdf = common * g.pdf(l, m);
            }
            return fs;
        }
        case BSDF_DIFFSPEC: {
            Ggx g(sq(a0), sq(a0), 0.0f);
            bool entering = q.z >= 0.0f;
            V3 v = entering ? q : -q;
            float wd, ws;
            diffSpecWeights(q.z, v.z, wl, &wd, &ws);
            // sampleDiscrete over {wd, ws} (random_distributions.h:53-76)
            float sumW = wd + ws;
            float su = uComponent * sumW;
            uint32_t comp = (su < wd) ? 0u : ((su < wd + ws) ? 1u : 0u);
            V3 l, m;
            float dotLH, D, pd, ps;
            if (comp == 0) {
                *type = dt::Reflection | dt::LowFreq;
                l = cosineSampleHemisphere(uD0, uD1);
                pd = l.z / kPiF;
                m = halfVector(l, v);
                dotLH = dot(l, m);
                ps = g.pdf(v, m) / (4 * dotLH);
                D = g.D(m);
            } else {
                *type = dt::Reflection | dt::HighFreq;
                float mPDF;
                D = g.sample(v, uD0, uD1, &m, &mPDF);
                dotLH = dot(v, m);
                l = 2 * dotLH * m - v;
                if (l.z * v.z <= 0) return Spec<N>(0.0f);
                ps = mPDF / (4 * dotLH);
                pd = l.z / kPiF;
            }
            float om5 = pow5f(1 - dotLH);
            float G = g.Gcorr(l, v, m);
            Spec<N> F = lerpSpec<N>(c1, Spec<N>(1.0f), om5);
            Spec<N> specular = F * ((D * G) / (4 * l.z * v.z));
            float FD90 = 0.5f * a0 + 2 * a0 * dotLH * dotLH;
            float fo = lerpf(1.0f, FD90, pow5f(1 - v.z));
            float fi = lerpf(1.0f, FD90, pow5f(1 - l.z));
            Spec<N> diffuse = c0 * (fo * fi * lerpf(1.0f, 1.0f / 1.51f, a0) / kPiF);
            *dir = entering ? l : -l;
            *dirPDF = (pd * wd + ps * ws) / sumW;
            Spec<N> ret = diffuse + specular;
            if (rev) {
                rev->value = ret;
                rev->pdf = diffSpecRevPDF(g, v, l, m, dotLH, wl);
            }
            return ret;
        }
        case BSDF_MICRO_BSDF: {
            // materials.cu:1270-1368
            bool entering = q.z >= 0.0f;
            const Spec<N> &eEnter = entering ? c1 : c2;
            const Spec<N> &eExit = entering ? c2 : c1;
            Ggx g(a0, a1, a2);
            V3 v = entering ? q : -q;
            V3 m;
            float mPDF;
            float D = g.sample(v, uD0, uD1, &m, &mPDF);
            float dotHV = dot(v, m);
            if (dotHV <= 0 || isnan(D)) return Spec<N>(0.0f);
            Spec<N> F = fresnelDielectric<N>(eEnter, eExit, dotHV);
            float reflectProb = F.importance(wl);
            if (uComponent < reflectProb) {
                V3 l = 2 * dotHV * m - v;
                *dir = entering ? l : -l;
                if (l.z * v.z <= 0) return Spec<N>(0.0f);
                float common = reflectProb / (4 * dotHV);
                *dirPDF = common * mPDF;
                *type = dt::Reflection | dt::HighFreq;
                float G = g.G1(v, m) * g.G1(l, m);
                Spec<N> ret = c0 * F * D * G / (4 * v.z * l.z);
                if (rev) {
                    rev->value = ret;
                    rev->pdf = common * g.pdf(l, m);
                }
                return ret;
            }
            const float eEw = eEnter.get(wl), eXw = eExit.get(wl);
            float rr = eEw / eXw;
            float innerRoot = 1 + rr * rr * (dotHV * dotHV - 1);
            if (innerRoot < 0) return Spec<N>(0.0f);
            V3 l = (rr * dotHV - sqrtf(innerRoot)) * m - rr * v;
            *dir = entering ? l : -l;
            if (l.z * v.z >= 0) return Spec<N>(0.0f);
            float dotHL = dot(l, m);
            float common = (1 - reflectProb) / sq(eEw * dotHV + eXw * dotHL);
            *dirPDF = common * mPDF * sq(eXw) * fabsf(dotHL);
            *type = dt::Transmission | dt::HighFreq;
            Spec<N> ret;
#pragma unroll
            for (int i = 0; i < N; ++i) ret[i] = roughTransmission(eEnter, eExit, eEnter, eExit, g, v, l, entering, i);
            ret = ret / fabsf(v.z * l.z);
            ret *= c0;
            ret *= importance ? eExit * eExit : eEnter * eEnter;
            if (rev) {
                rev->value = ret;
                rev->pdf = common * g.pdf(l, m) * sq(eEw) * fabsf(dotHV);
            }
            return ret;
        }
        case BSDF_EDF_DIFFUSE: {
            *dir = cosineSampleHemisphere(uD0, uD1);
            *dirPDF = dir->z / kPiF;
            *type = dt::Reflection | dt::LowFreq;
            if (rev) rev->value = Spec<N>(0.0f);
            return Spec<N>(1.0f / kPiF);
        }
        case BSDF_EDF_POINT: {
            *dir = uniformSampleSphere(uD0, uD1);
            *dirPDF = 1.0f / (4 * kPiF);
            *type = dt::Reflection | dt::LowFreq;
            if (rev) rev->value = Spec<N>(0.0f);
            return Spec<N>(1.0f / (4 * kPiF));
        }
        case BSDF_EDF_ENV: {
            float dx, dy;
            concentricSampleDisk(uD0, uD1, &dx, &dy);
            *dir = V3(a1 * dx, a1 * dy, 1.0f);
            *dirPDF = 1.0f / a0;
            *type = dt::Reflection | dt::LowFreq;
            if (rev) rev->value = Spec<N>(0.0f);
            return Spec<N>(1.0f / kPiF);
        }
        default:
            return Spec<N>(0.0f);
        }
    }

    __device__ Spec<N> evaluateInternal(const V3 &q, const V3 &d, uint32_t wl, bool importance,
                                        Spec<N> *revValue) const {
        Spec<N> ret(0.0f);
        switch (kind) {
        case BSDF_MATTE:
            if (q.z * d.z > 0.0f) ret = c0 / kPiF;
            break;
        case BSDF_MICRO_BRDF: {
            bool entering = q.z >= 0.0f;
            V3 v = entering ? q : -q;
            V3 l = entering ? d : -d;
            float nn = l.z * v.z;
            if (nn <= 0) return Spec<N>(0.0f);   // the reference leaves the reverse value unset here
            Ggx g(a0, a1, a2);
            V3 m = halfVector(v, l);
            float dotHV = dot(v, m);
            float G = g.G1(v, m) * g.G1(l, m);
            ret = fresnelConductor<N>(c0, c1, dotHV) * g.D(m) * G / (4 * nn);
            break;
        }
        case BSDF_DIFFSPEC: {
            if (d.z * q.z <= 0) return Spec<N>(0.0f);
            Ggx g(sq(a0), sq(a0), 0.0f);
            bool entering = q.z >= 0.0f;
            V3 v = entering ? q : -q;
            V3 l = entering ? d : -d;
            V3 m = halfVector(l, v);
            float dotLH = dot(l, m);
            Spec<N> F = lerpSpec<N>(c1, Spec<N>(1.0f), pow5f(1 - dotLH));
            Spec<N> specular = F * ((g.D(m) * g.Gcorr(l, v, m)) / (4 * l.z * v.z));
            float FD90 = 0.5f * a0 + 2 * a0 * sq(dotLH);
            float fo = lerpf(1.0f, FD90, pow5f(1 - v.z));
            float fi = lerpf(1.0f, FD90, pow5f(1 - l.z));
            ret = c0 * (fo * fi * lerpf(1.0f, 1.0f / 1.51f, a0) / kPiF) + specular;
            break;
        }
        case BSDF_MICRO_BSDF: {
            bool entering = q.z >= 0.0f;
            const Spec<N> &eEnter = entering ? c1 : c2;
            const Spec<N> &eExit = entering ? c2 : c1;
            Ggx g(a0, a1, a2);
            V3 v = entering ? q : -q;
            V3 l = entering ? d : -d;
            float nn = l.z * v.z;
            if (nn > 0) {
                V3 m = halfVector(v, l);
                float dotHV = dot(v, m);
                Spec<N> F = fresnelDielectric<N>(eEnter, eExit, dotHV);
                float G = g.G1(v, m) * g.G1(l, m);
                ret = c0 * F * g.D(m) * G / (4 * nn);
            } else if (nn < 0) {
#pragma unroll
                for (int i = 0; i < N; ++i) ret[i] = roughTransmission(eEnter, eExit, eEnter, eExit, g, v, l, entering, i);
                ret = ret / fabsf(nn);
                ret *= c0;
                ret *= importance ? eExit * eExit : eEnter * eEnter;
            }
            break;
        }
        case BSDF_EDF_DIFFUSE:
        case BSDF_EDF_ENV:
            if (revValue) *revValue = Spec<N>(0.0f);
            return Spec<N>(d.z > 0.0f ? 1.0f / kPiF : 0.0f);
        case BSDF_EDF_POINT:
            if (revValue) *revValue = Spec<N>(0.0f);
            return Spec<N>(1.0f / (4 * kPiF));
        default:
            break;
        }
        if (revValue) *revValue = ret;
        return ret;
    }

    __device__ float evaluatePDFInternal(const V3 &q, const V3 &d, uint32_t wl, float *revPDF) const {
        if (revPDF) *revPDF = 0.0f;
        switch (kind) {
        case BSDF_MATTE:
            if (q.z * d.z <= 0.0f) return 0.0f;
            if (revPDF) *revPDF = fabsf(q.z) / kPiF;
            return fabsf(d.z) / kPiF;
        case BSDF_MICRO_BRDF: {
            bool entering = q.z >= 0.0f;
            V3 v = entering ? q : -q;
            V3 l = entering ? d : -d;
            if (l.z * v.z <= 0.0f) return 0.0f;
            V3 m = halfVector(v, l);
            float dotHV = dot(v, m);
            if (dotHV <= 0) return 0.0f;
            Ggx g(a0, a1, a2);
            float common = 1.0f / (4 * dotHV);
            if (revPDF) *revPDF = common * g.pdf(l, m);
            return common * g.pdf(v, m);
        }
        case BSDF_DIFFSPEC: {
            Ggx g(sq(a0), sq(a0), 0.0f);
            bool entering = q.z >= 0.0f;
            V3 v = entering ? q : -q;
            V3 l = entering ? d : -d;
            V3 m = halfVector(l, v);
            float dotLH = dot(l, m);
            float wd, ws;
            diffSpecWeights(q.z, v.z, wl, &wd, &ws);
            float pd = fmaxf(l.z / kPiF, 0.0f);
            float ps = g.pdf(v, m) / (4 * dotLH);
            if (revPDF) *revPDF = diffSpecRevPDF(g, v, l, m, dotLH, wl);
            return (pd * wd + ps * ws) / (wd + ws);
        }
        case BSDF_MICRO_BSDF: {
            // materials.cu:1462-1510; the transmission half vector carries no sign flip for
            // queries from inside, as in the reference
            bool entering = q.z >= 0.0f;
            const Spec<N> &eEnter = entering ? c1 : c2;
            const Spec<N> &eExit = entering ? c2 : c1;
            V3 v = entering ? q : -q;
            V3 l = entering ? d : -d;
            float nn = l.z * v.z;
            if (nn == 0) return 0.0f;
            const float eEw = eEnter.get(wl), eXw = eExit.get(wl);
            V3 m = nn > 0 ? halfVector(v, l) : normalize(-(eEw * v + eXw * l));
            float dotHV = dot(v, m);
            if (dotHV <= 0) return 0.0f;
            Ggx g(a0, a1, a2);
            float mPDF = g.pdf(v, m);
            Spec<N> F = fresnelDielectric<N>(eEnter, eExit, dotHV);
            float reflectProb = F.importance(wl);
            if (nn > 0) {
                float common = reflectProb / (4 * dotHV);
                if (revPDF) *revPDF = common * g.pdf(l, m);
                return common * mPDF;
            }
            float dotHL = dot(l, m);
            float common = (1 - reflectProb) / sq(eEw * dotHV + eXw * dotHL);
            if (revPDF) *revPDF = common * g.pdf(l, m) * sq(eEw) * fabsf(dotHV);
            return common * mPDF * sq(eXw) * fabsf(dotHL);
        }
        case BSDF_EDF_DIFFUSE:
            return d.z <= 0.0f ? 0.0f : d.z / kPiF;
        case BSDF_EDF_POINT:
            return 1.0f / (4 * kPiF);
        case BSDF_EDF_ENV:
            return d.z <= 0.0f ? 0.0f : 1.0f / a0;
        default:
            return 0.0f;
        }
    }

    // wrappers with the shading-normal correction (renderer_common.h:107-115, 130-136, 224-264):
    // |out.ns / out.ng| for radiance transport, |q.ns / q.ng| for importance transport
    __device__ Spec<N> sample(const V3 &q, const V3 &gnL, uint32_t wl, float uC, float u0, float u1, V3 *dir,
                              float *pdf, uint32_t *type, bool importance = false, BsdfRev<N> *rev = nullptr) const {
        *pdf = 0.0f;
        *type = 0;
        if (rev) rev->pdf = 0.0f;
        if (!matches(dt::All)) return Spec<N>(0.0f);
        Spec<N> f = sampleInternal(q, wl, importance, uC, u0, u1, dir, pdf, type, rev);
        if (*pdf == 0.0f) return Spec<N>(0.0f);
        float sn = importance ? fabsf(q.z / dot(q, gnL)) : fabsf(dir->z / dot(*dir, gnL));
        if (!isfinite(sn)) return Spec<N>(0.0f);
        if (rev) rev->value = rev->value * sn;
        return f * sn;
    }
    __device__ Spec<N> evaluate(const V3 &q, const V3 &gnL, const V3 &d, uint32_t wl = 0, bool importance = false,
                                Spec<N> *revValue = nullptr) const {
        float sn = importance ? fabsf(q.z / dot(q, gnL)) : fabsf(d.z / dot(d, gnL));
        if (revValue) *revValue = Spec<N>(0.0f);
        if (!isfinite(sn)) return Spec<N>(0.0f);
        Spec<N> r = evaluateInternal(q, d, wl, importance, revValue) * sn;
        if (revValue) *revValue = *revValue * sn;
        return r;
    }
    __device__ float evaluatePDF(const V3 &q, const V3 &d, uint32_t wl, float *revPDF = nullptr) const {
        if (revPDF) *revPDF = 0.0f;
        if (!matches(dt::All)) return 0.0f;
        return evaluatePDFInternal(q, d, wl, revPDF);
    }
};

// ---------------------------------------------------------------------------------------------
// surface points
struct SurfPt {
    V3 position, gn;
    Frame frame;
    float texU, texV;
    bool isPoint, atInfinity;
    // hit identity and surface parameters (barycentrics b0, b1; phi, theta at infinity)
    uint32_t instIndex, geomInstIndex, primIndex;
    float u, v;
    __device__ __forceinline__ float cosTerm(const V3 &w) const { return isPoint ? 1.0f : absDot(w, gn); }
};

} // namespace vlra
