// Oracle (test infrastructure only): CPU restatement of the reference's BSDF/EDF/IDF code.
// Sources followed (libVLR/GPU_kernels/materials.cu unless stated):
//   DirectionType             shared/kernel_common.h:209-283
//   sampling routines         shared/random_distributions.h:80-170
//   FresnelConductor/Dielectric  :13-129
//   GGXMicrofacetDistribution :133-228
//   MatteBRDF                 :584-663   (setup :652-661)
//   SpecularBRDF              :667-789
//   SpecularBSDF              :793-1060  (setup :1045-1058)
//   MicrofacetBRDF            :1064-1242 (setup :1220-1240)
//   DiffuseAndSpecularBRDF    :1692-1994 (UE4 setup :1959-1978)
//   DiffuseEDF :2040-2096, PointEDF :2194-2250, EnvironmentEDF :2830-2896
//   BSDF wrapper (shading-normal correction)  shared/renderer_common.h:44-163
//   PerspectiveCamera sample/IDF  GPU_kernels/cameras.cu:43-76, 124-222
#pragma once
#include "ov_spectrum.h"

namespace ov {

// kernel_common.h:209-283
namespace DT {
enum : uint32_t {
    LowFreq = 1 << 0, HighFreq = 1 << 1, Delta0D = 1 << 2, Delta1D = 1 << 3,
    NonDelta = LowFreq | HighFreq, Delta = Delta0D | Delta1D, AllFreq = NonDelta | Delta,
    Reflection = 1 << 4, Transmission = 1 << 5, Emission = Reflection, Acquisition = Reflection,
    WholeSphere = Reflection | Transmission, All = AllFreq | WholeSphere, Dispersive = 1 << 6,
};
inline bool matches(uint32_t self, uint32_t t) { uint32_t r = self & t; return (r & WholeSphere) && (r & AllFreq); }
inline bool isDelta(uint32_t v) { return (v & Delta) && !(v & NonDelta); }
inline bool isReflection(uint32_t v) { return (v & Reflection) && !(v & Transmission); }
inline bool isTransmission(uint32_t v) { return !(v & Reflection) && (v & Transmission); }
inline bool isDispersive(uint32_t v) { return v & Dispersive; }
} // namespace DT

// random_distributions.h:80-114
inline void concentricSampleDisk(float u0, float u1, float *dx, float *dy) {
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
    theta *= kPi / 4;
    *dx = r * std::cos(theta);
    *dy = r * std::sin(theta);
}
// random_distributions.h:116-124
inline V3 cosineSampleHemisphere(float u0, float u1) {
    float x, y;
    concentricSampleDisk(u0, u1, &x, &y);
    return V3(x, y, std::sqrt(std::fmax(0.0f, 1.0f - x * x - y * y)));
}
// random_distributions.h:140-145
inline V3 uniformSampleSphere(float u0, float u1) {
    float phi = 2 * kPi * u1;
    float theta = std::acos(1 - 2 * u0);
    return V3::fromPolarZUp(phi, theta);
}
// random_distributions.h:154-170 (Heitz's low-distortion square-to-triangle map)
inline void uniformSampleTriangle(float u0, float u1, float *b0, float *b1) {
    *b0 = 0.5f * u0;
    *b1 = 0.5f * u1;
    float offset = *b1 - *b0;
    if (offset > 0)
        *b1 += offset;
    else
        *b0 -= offset;
}
// random_distributions.h:53-76
inline uint32_t sampleDiscrete(const float *imp, uint32_t n, float u, float *prob, float *sumImp, float *remapped) {
    float sum = 0;
    for (uint32_t i = 0; i < n; ++i) sum += imp[i];
    *sumImp = sum;
    float base = 0, su = u * sum, cum = 0;
    for (uint32_t i = 0; i < n; ++i) {
        base = cum;
        cum += imp[i];
        if (su < cum) {
            *prob = imp[i] / sum;
            *remapped = (su - base) / imp[i];
            return i;
        }
    }
    *prob = imp[0] / sum;
    return 0;
}

enum class TransportMode { Radiance = 0, Importance = 1 };

template <int N>
struct BSDFQuery {
    V3 dirLocal;
    V3 geometricNormalLocal;
    uint32_t dirTypeFilter;
    TransportMode transportMode;
    uint32_t wlHint;
};

template <int N>
struct BSDFResult {
    V3 dirLocal;
    float dirPDF = 0;
    uint32_t sampledType = 0;
};

template <int N>
struct RevResult {
    SS<N> value;
    float dirPDF = 0;
};

// materials.cu:13-46
template <int N>
struct FresnelConductor {
    SS<N> eta, k;
    SS<N> evaluate(float cosEnter) const {
        cosEnter = std::fabs(cosEnter);
        float c2 = cosEnter * cosEnter;
        SS<N> _2EtaCos = 2.0f * eta * cosEnter;
        SS<N> tmp_f = eta * eta + k * k;
        SS<N> tmp = tmp_f * c2;
        SS<N> Rparl2 = (tmp - _2EtaCos + 1.0f) / (tmp + _2EtaCos + 1.0f);
        SS<N> Rperp2 = (tmp_f - _2EtaCos + c2) / (tmp_f + _2EtaCos + c2);
        return (Rparl2 + Rperp2) / 2.0f;
    }
};

// materials.cu:50-110
inline float fresnelDielectricEvalF(float etaEnter, float etaExit, float cosEnter, float cosExit) {
    float Rparl = ((etaExit * cosEnter) - (etaEnter * cosExit)) / ((etaExit * cosEnter) + (etaEnter * cosExit));
    float Rperp = ((etaEnter * cosEnter) - (etaExit * cosExit)) / ((etaEnter * cosEnter) + (etaExit * cosExit));
    return (Rparl * Rparl + Rperp * Rperp) / 2.0f;
}
template <int N>
struct FresnelDielectric {
    SS<N> etaExt, etaInt;
    SS<N> evaluate(float cosEnter) const {
        cosEnter = clampv(cosEnter, -1.0f, 1.0f);
        bool entering = cosEnter > 0.0f;
        const SS<N> &eEnter = entering ? etaExt : etaInt;
        const SS<N> &eExit = entering ? etaInt : etaExt;
        SS<N> sinExit = eEnter / eExit * std::sqrt(std::fmax(0.0f, 1.0f - cosEnter * cosEnter));
        SS<N> ret(0.0f);
        cosEnter = std::fabs(cosEnter);
        for (int i = 0; i < N; ++i) {
            if (sinExit[i] >= 1.0f) {
                ret[i] = 1.0f;
            } else {
                float cosExit = std::sqrt(std::fmax(0.0f, 1.0f - sinExit[i] * sinExit[i]));
                ret[i] = fresnelDielectricEvalF(eEnter[i], eExit[i], cosEnter, cosExit);
            }
        }
        return ret;
    }
    float evaluate1(float cosEnter, int i) const {
        cosEnter = clampv(cosEnter, -1.0f, 1.0f);
        bool entering = cosEnter > 0.0f;
        float eEnter = entering ? etaExt[i] : etaInt[i];
        float eExit = entering ? etaInt[i] : etaExt[i];
        float sinExit = eEnter / eExit * std::sqrt(std::fmax(0.0f, 1.0f - cosEnter * cosEnter));
        cosEnter = std::fabs(cosEnter);
        if (sinExit >= 1.0f)
            return 1.0f;
        float cosExit = std::sqrt(std::fmax(0.0f, 1.0f - sinExit * sinExit));
        return fresnelDielectricEvalF(eEnter, eExit, cosEnter, cosExit);
    }
};

// materials.cu:133-228
struct GGX {
    float ax, ay, cosRt, sinRt;
    GGX(float alphaX, float alphaY, float rotation) : ax(alphaX), ay(alphaY) {
        sinRt = std::sin(rotation);
        cosRt = std::cos(rotation);
    }
    float evaluate(const V3 &m) const {
        V3 mr(cosRt * m.x + sinRt * m.y, -sinRt * m.x + cosRt * m.y, m.z);
        if (mr.z <= 0)
            return 0.0f;
        float temp = pow2(mr.x / ax) + pow2(mr.y / ay) + pow2(mr.z);
        return 1.0f / (kPi * ax * ay * pow2(temp));
    }
    float evaluateSmithG1(const V3 &v, const V3 &m) const {
        V3 vr(cosRt * v.x + sinRt * v.y, -sinRt * v.x + cosRt * v.y, v.z);
        float a2t2 = (pow2(vr.x * ax) + pow2(vr.y * ay)) / pow2(vr.z);
        float Lambda = (-1 + std::sqrt(1 + a2t2)) / 2;
        float chi = (dot(v, m) / v.z) > 0 ? 1.0f : 0.0f;
        return chi / (1 + Lambda);
    }
    float evaluateHeightCorrelatedSmithG(const V3 &v1, const V3 &v2, const V3 &m) const {
        V3 v1r(cosRt * v1.x + sinRt * v1.y, -sinRt * v1.x + cosRt * v1.y, v1.z);
        V3 v2r(cosRt * v2.x + sinRt * v2.y, -sinRt * v2.x + cosRt * v2.y, v2.z);
        float t1 = (pow2(v1r.x * ax) + pow2(v1r.y * ay)) / pow2(v1r.z);
        float t2 = (pow2(v2r.x * ax) + pow2(v2r.y * ay)) / pow2(v2r.z);
        float L1 = (-1 + std::sqrt(1 + t1)) / 2;
        float L2 = (-1 + std::sqrt(1 + t2)) / 2;
        float chi1 = (dot(v1, m) / v1.z) > 0 ? 1.0f : 0.0f;
        float chi2 = (dot(v2, m) / v2.z) > 0 ? 1.0f : 0.0f;
        return chi1 * chi2 / (1 + L1 + L2);
    }
    float sample(const V3 &v, float u0, float u1, V3 *m, float *normalPDF) const {
        V3 vr(cosRt * v.x + sinRt * v.y, -sinRt * v.x + cosRt * v.y, v.z);
        V3 sv = normalize(V3(ax * vr.x, ay * vr.y, vr.z));
        float distIn2D = std::sqrt(sv.x * sv.x + sv.y * sv.y);
        float recDistIn2D = 1.0f / distIn2D;
        V3 T1 = (sv.z < 0.9999f) ? V3(sv.y * recDistIn2D, -sv.x * recDistIn2D, 0) : V3(1, 0, 0);
        V3 T2(T1.y * sv.z, -T1.x * sv.z, distIn2D);
        float a = 1.0f / (1.0f + sv.z);
        float r = std::sqrt(u0);
        float phi = kPi * ((u1 < a) ? u1 / a : 1 + (u1 - a) / (1.0f - a));
        float sinPhi = std::sin(phi), cosPhi = std::cos(phi);
        float P1 = r * cosPhi;
        float P2 = r * sinPhi * ((u1 < a) ? 1.0f : sv.z);
        V3 mr = P1 * T1 + P2 * T2 + std::sqrt(1.0f - P1 * P1 - P2 * P2) * sv;
        mr = normalize(V3(ax * mr.x, ay * mr.y, mr.z));
        float D = evaluate(mr);
        *normalPDF = evaluateSmithG1(vr, mr) * absDot(vr, mr) * D / std::fabs(vr.z);
        *m = V3(cosRt * mr.x - sinRt * mr.y, sinRt * mr.x + cosRt * mr.y, mr.z);
        return D;
    }
    float evaluatePDF(const V3 &v, const V3 &m) const {
        return evaluateSmithG1(v, m) * absDot(v, m) * evaluate(m) / std::fabs(v.z);
    }
};

enum class BSDFKind { Null = 0, Matte, SpecularBRDF, SpecularBSDF, MicrofacetBRDF, DiffuseAndSpecular, MicrofacetBSDF };

// One BSDF after material setup at a surface point (the content of the reference's 32-dword
// parameter block).
template <int N>
struct BSDF {
    BSDFKind kind = BSDFKind::Null;
    SS<N> c0, c1, c2;   // Matte: albedo | SpecBRDF: coeffR, eta, k | SpecBSDF: coeff, etaExt, etaInt
                        // MicrofacetBRDF: eta, k | DiffSpec: diffuseColor, specularF0Color
                        // MicrofacetBSDF: coeff, etaExt, etaInt
    float a0 = 0, a1 = 0, a2 = 0;   // Microfacet*: alphaX, alphaY, rotation | DiffSpec: roughness
    bool dispersive = false;

    bool matches(uint32_t flags) const {
        uint32_t t = 0;
        switch (kind) {
        case BSDFKind::Null: return false;
        case BSDFKind::Matte: t = DT::Reflection | DT::LowFreq; break;
        case BSDFKind::SpecularBRDF: t = DT::Reflection | DT::Delta0D; break;
        case BSDFKind::SpecularBSDF: t = DT::WholeSphere | DT::Delta0D; break;
        case BSDFKind::MicrofacetBRDF: t = DT::Reflection | DT::HighFreq; break;
        case BSDFKind::DiffuseAndSpecular: t = DT::Reflection | DT::LowFreq | DT::HighFreq; break;
        case BSDFKind::MicrofacetBSDF: t = DT::WholeSphere | DT::HighFreq; break;
        }
        return DT::matches(t, flags);
    }
    bool hasNonDelta() const { return matches(DT::WholeSphere | DT::NonDelta); }

    SS<N> getBaseColor() const {
        switch (kind) {
        case BSDFKind::Matte: return c0;
        case BSDFKind::SpecularBRDF: return c0;
        case BSDFKind::SpecularBSDF: return c0;
        case BSDFKind::MicrofacetBRDF: return FresnelConductor<N>{c0, c1}.evaluate(1.0f);
        case BSDFKind::DiffuseAndSpecular: return c0 + c1;
        case BSDFKind::MicrofacetBSDF: return c0;
        default: return SS<N>(0.0f);
        }
    }

    // ---------------- internal (shading-normal frame) functions ----------------
    SS<N> sampleInternal(const BSDFQuery<N> &q, float uComponent, const float uDir[2], BSDFResult<N> *res,
                         RevResult<N> *rev) const {
        switch (kind) {
        case BSDFKind::Matte: {
            res->dirLocal = cosineSampleHemisphere(uDir[0], uDir[1]);
            res->dirPDF = res->dirLocal.z / kPi;
            res->sampledType = DT::Reflection | DT::LowFreq;
            res->dirLocal.z *= q.dirLocal.z >= 0 ? 1 : -1;
            SS<N> fs = c0 / kPi;
            if (rev) {
                rev->value = fs;
                rev->dirPDF = std::fabs(q.dirLocal.z) / kPi;
            }
            return fs;
        }
        case BSDFKind::SpecularBRDF: {
            FresnelConductor<N> fr{c1, c2};
            res->dirLocal = V3(-q.dirLocal.x, -q.dirLocal.y, q.dirLocal.z);
            res->dirPDF = 1.0f;
            res->sampledType = DT::Reflection | DT::Delta0D;
            SS<N> ret = c0 * fr.evaluate(q.dirLocal.z) * (1.0f / std::fabs(q.dirLocal.z));
            if (rev) {
                rev->value = ret;
                rev->dirPDF = 1.0f;
            }
            return ret;
        }
        case BSDFKind::SpecularBSDF: return sampleSpecularBSDF(q, uComponent, res, rev);
        case BSDFKind::MicrofacetBRDF: return sampleMicrofacetBRDF(q, uDir, res, rev);
        case BSDFKind::DiffuseAndSpecular: return sampleDiffSpec(q, uComponent, uDir, res, rev);
        case BSDFKind::MicrofacetBSDF: return sampleRoughDielectric(q, uComponent, uDir, res, rev);
        default:
            res->dirPDF = 0.0f;
            if (rev) rev->dirPDF = 0.0f;
            return SS<N>(0.0f);
        }
    }

    SS<N> evaluateInternal(const BSDFQuery<N> &q, const V3 &dir, SS<N> *revValue) const {
        switch (kind) {
        case BSDFKind::Matte: {
            if (q.dirLocal.z * dir.z <= 0.0f) {
                if (revValue) *revValue = SS<N>(0.0f);
                return SS<N>(0.0f);
            }
            SS<N> fs = c0 / kPi;
            if (revValue) *revValue = fs;
            return fs;
        }
        case BSDFKind::MicrofacetBRDF: {
            bool entering = q.dirLocal.z >= 0.0f;
            FresnelConductor<N> fr{c0, c1};
            GGX ggx(a0, a1, a2);
            V3 dirV = entering ? q.dirLocal : -q.dirLocal;
            V3 dirL = entering ? dir : -dir;
            float dotNVdotNL = dirL.z * dirV.z;
            if (dotNVdotNL <= 0)
                return SS<N>(0.0f);
            V3 m = halfVector(dirV, dirL);
            float dotHV = dot(dirV, m);
            float D = ggx.evaluate(m);
            SS<N> F = fr.evaluate(dotHV);
            float G = ggx.evaluateSmithG1(dirV, m) * ggx.evaluateSmithG1(dirL, m);
            SS<N> fs = F * D * G / (4 * dotNVdotNL);
            if (revValue) *revValue = fs;
            return fs;
        }
        case BSDFKind::DiffuseAndSpecular: return evaluateDiffSpec(q, dir, revValue);
        case BSDFKind::MicrofacetBSDF: return evaluateRoughDielectric(q, dir, revValue);
        default:
            if (revValue) *revValue = SS<N>(0.0f);
            return SS<N>(0.0f);
        }
    }

    float evaluatePDFInternal(const BSDFQuery<N> &q, const V3 &dir, float *revValue) const {
        switch (kind) {
        case BSDFKind::Matte: {
            if (q.dirLocal.z * dir.z <= 0.0f) {
                if (revValue) *revValue = 0.0f;
                return 0.0f;
            }
            if (revValue) *revValue = std::fabs(q.dirLocal.z) / kPi;
            return std::fabs(dir.z) / kPi;
        }
        case BSDFKind::MicrofacetBRDF: {
            if (revValue) *revValue = 0.0f;
            bool entering = q.dirLocal.z >= 0.0f;
            GGX ggx(a0, a1, a2);
            V3 dirV = entering ? q.dirLocal : -q.dirLocal;
            V3 dirL = entering ? dir : -dir;
            if (dirL.z * dirV.z <= 0.0f)
                return 0.0f;
            V3 m = halfVector(dirV, dirL);
            float dotHV = dot(dirV, m);
            if (dotHV <= 0)
                return 0.0f;
            float mPDF = ggx.evaluatePDF(dirV, m);
            float common = 1.0f / (4 * dotHV);
            if (revValue) *revValue = common * ggx.evaluatePDF(dirL, m);
            return common * mPDF;
        }
        case BSDFKind::DiffuseAndSpecular: return evaluatePDFDiffSpec(q, dir, revValue);
        case BSDFKind::MicrofacetBSDF: return evaluatePDFRoughDielectric(q, dir, revValue);
        default:
            if (revValue) *revValue = 0.0f;
            return 0.0f;
        }
    }

    // ---------------- wrapper with the shading-normal correction ----------------
    SS<N> sample(const BSDFQuery<N> &q, float uComponent, float uDir0, float uDir1, BSDFResult<N> *res,
                 RevResult<N> *rev = nullptr) const {
        res->dirPDF = 0.0f;
        res->sampledType = 0;
        if (rev) rev->dirPDF = 0.0f;
        SS<N> ret(0.0f);
        if (!matches(q.dirTypeFilter))
            return ret;
        const float uDir[2] = {uDir0, uDir1};
        SS<N> fs_sn = sampleInternal(q, uComponent, uDir, res, rev);
        float snCorrection;
        if (q.transportMode == TransportMode::Radiance)
            snCorrection = std::fabs(res->dirLocal.z / dot(res->dirLocal, q.geometricNormalLocal));
        else
            snCorrection = std::fabs(q.dirLocal.z / dot(q.dirLocal, q.geometricNormalLocal));
        if (!std::isfinite(snCorrection))
            return ret;
        ret = fs_sn * snCorrection;
        if (rev) rev->value *= snCorrection;
        return ret;
    }
    SS<N> evaluate(const BSDFQuery<N> &q, const V3 &dir, SS<N> *revValue = nullptr) const {
        SS<N> ret(0.0f);
        float snCorrection;
        if (q.transportMode == TransportMode::Radiance)
            snCorrection = std::fabs(dir.z / dot(dir, q.geometricNormalLocal));
        else
            snCorrection = std::fabs(q.dirLocal.z / dot(q.dirLocal, q.geometricNormalLocal));
        if (!std::isfinite(snCorrection))
            return ret;
        SS<N> fs_sn = evaluateInternal(q, dir, revValue);
        ret = fs_sn * snCorrection;
        if (revValue) *revValue *= snCorrection;
        return ret;
    }
    float evaluatePDF(const BSDFQuery<N> &q, const V3 &dir, float *revValue = nullptr) const {
        if (!matches(q.dirTypeFilter)) {
            if (revValue) *revValue = 0.0f;
            return 0;
        }
        return evaluatePDFInternal(q, dir, revValue);
    }

private:
    // Rough dielectric (materials.cu:1246-1538): GGX half-vector sampling, Fresnel-weighted
    // choice between reflection and refraction at the selected wavelength, per-wavelength
    // half vectors when evaluating transmission.
    float roughTransmission(const SS<N> &eEnter, const SS<N> &eExit, const FresnelDielectric<N> &fresnel,
                            const GGX &ggx, const V3 &dirV, const V3 &dirL, bool entering, int i) const {
        V3 m = normalize(-(eEnter[i] * dirV + eExit[i] * dirL) * (entering ? 1.0f : -1.0f));
        float dotHV = dot(dirV, m), dotHL = dot(dirL, m);
        float F = fresnel.evaluate1(dotHV, i);
        float G = ggx.evaluateSmithG1(dirV, m) * ggx.evaluateSmithG1(dirL, m);
        float D = ggx.evaluate(m);
        return std::fabs(dotHV * dotHL) * (1 - F) * G * D / pow2(eEnter[i] * dotHV + eExit[i] * dotHL);
    }
    SS<N> sampleRoughDielectric(const BSDFQuery<N> &q, float uComponent, const float uDir[2], BSDFResult<N> *res,
                                RevResult<N> *rev) const {
        if (rev) rev->dirPDF = 0.0f;
        bool entering = q.dirLocal.z >= 0.0f;
        const SS<N> &eEnter = entering ? c1 : c2;
        const SS<N> &eExit = entering ? c2 : c1;
        FresnelDielectric<N> fresnel{eEnter, eExit};
        GGX ggx(a0, a1, a2);
        V3 dirV = entering ? q.dirLocal : -q.dirLocal;
        V3 m;
        float mPDF;
        float D = ggx.sample(dirV, uDir[0], uDir[1], &m, &mPDF);
        float dotHV = dot(dirV, m);
        if (dotHV <= 0 || std::isnan(D)) {
            res->dirPDF = 0.0f;
            return SS<N>(0.0f);
        }
        SS<N> F = fresnel.evaluate(dotHV);
        float reflectProb = F.importance(q.wlHint);
        if (DT::isReflection(q.dirTypeFilter)) reflectProb = 1.0f;
        if (DT::isTransmission(q.dirTypeFilter)) reflectProb = 0.0f;
        if (uComponent < reflectProb) {
            V3 dirL = 2 * dotHV * m - dirV;
            res->dirLocal = entering ? dirL : -dirL;
            if (dirL.z * dirV.z <= 0) {
                res->dirPDF = 0.0f;
                return SS<N>(0.0f);
            }
            float common = reflectProb / (4 * dotHV);
            res->dirPDF = common * mPDF;
            res->sampledType = DT::Reflection | DT::HighFreq;
            float G = ggx.evaluateSmithG1(dirV, m) * ggx.evaluateSmithG1(dirL, m);
            SS<N> ret = c0 * F * D * G / (4 * dirV.z * dirL.z);
            if (rev) {
                rev->value = ret;
                rev->dirPDF = common * ggx.evaluatePDF(dirL, m);
            }
            return ret;
        }
        const int w = static_cast<int>(q.wlHint);
        float recRelIOR = eEnter[w] / eExit[w];
        float innerRoot = 1 + recRelIOR * recRelIOR * (dotHV * dotHV - 1);
        if (innerRoot < 0) {
            res->dirPDF = 0.0f;
            return SS<N>(0.0f);
        }
        V3 dirL = (recRelIOR * dotHV - std::sqrt(innerRoot)) * m - recRelIOR * dirV;
        res->dirLocal = entering ? dirL : -dirL;
        if (dirL.z * dirV.z >= 0) {
            res->dirPDF = 0.0f;
            return SS<N>(0.0f);
        }
        float dotHL = dot(dirL, m);
        float common = (1 - reflectProb) / pow2(eEnter[w] * dotHV + eExit[w] * dotHL);
        res->dirPDF = common * mPDF * pow2(eExit[w]) * std::fabs(dotHL);
        res->sampledType = DT::Transmission | DT::HighFreq;
        SS<N> ret(0.0f);
        for (int i = 0; i < N; ++i) ret[i] = roughTransmission(eEnter, eExit, fresnel, ggx, dirV, dirL, entering, i);
        ret /= std::fabs(dirV.z * dirL.z);
        ret *= c0;
        ret *= q.transportMode == TransportMode::Radiance ? eEnter * eEnter : eExit * eExit;
        if (rev) {
            rev->value = ret;
            rev->dirPDF = common * ggx.evaluatePDF(dirL, m) * pow2(eEnter[w]) * std::fabs(dotHV);
        }
        return ret;
    }
    SS<N> evaluateRoughDielectric(const BSDFQuery<N> &q, const V3 &dir, SS<N> *revValue) const {
        bool entering = q.dirLocal.z >= 0.0f;
        const SS<N> &eEnter = entering ? c1 : c2;
        const SS<N> &eExit = entering ? c2 : c1;
        FresnelDielectric<N> fresnel{eEnter, eExit};
        GGX ggx(a0, a1, a2);
        V3 dirV = entering ? q.dirLocal : -q.dirLocal;
        V3 dirL = entering ? dir : -dir;
        float dotNVdotNL = dirL.z * dirV.z;
        if (dotNVdotNL > 0 && DT::matches(DT::Reflection | DT::AllFreq, q.dirTypeFilter)) {
            V3 m = halfVector(dirV, dirL);
            float dotHV = dot(dirV, m);
            float D = ggx.evaluate(m);
            SS<N> F = fresnel.evaluate(dotHV);
            float G = ggx.evaluateSmithG1(dirV, m) * ggx.evaluateSmithG1(dirL, m);
            SS<N> ret = c0 * F * D * G / (4 * dotNVdotNL);
            if (revValue) *revValue = ret;
            return ret;
        }
        if (dotNVdotNL < 0 && DT::matches(DT::Transmission | DT::AllFreq, q.dirTypeFilter)) {
            SS<N> ret(0.0f);
            for (int i = 0; i < N; ++i) ret[i] = roughTransmission(eEnter, eExit, fresnel, ggx, dirV, dirL, entering, i);
            ret /= std::fabs(dotNVdotNL);
            ret *= c0;
            ret *= q.transportMode == TransportMode::Radiance ? eEnter * eEnter : eExit * eExit;
            if (revValue) *revValue = ret;
            return ret;
        }
        if (revValue) *revValue = SS<N>(0.0f);
        return SS<N>(0.0f);
    }
    float evaluatePDFRoughDielectric(const BSDFQuery<N> &q, const V3 &dir, float *revValue) const {
        if (revValue) *revValue = 0.0f;
        bool entering = q.dirLocal.z >= 0.0f;
        const SS<N> &eEnter = entering ? c1 : c2;
        const SS<N> &eExit = entering ? c2 : c1;
        FresnelDielectric<N> fresnel{eEnter, eExit};
        GGX ggx(a0, a1, a2);
        V3 dirV = entering ? q.dirLocal : -q.dirLocal;
        V3 dirL = entering ? dir : -dir;
        float dotNVdotNL = dirL.z * dirV.z;
        if (dotNVdotNL == 0)
            return 0.0f;
        const int w = static_cast<int>(q.wlHint);
        V3 m = dotNVdotNL > 0 ? halfVector(dirV, dirL) : normalize(-(eEnter[w] * dirV + eExit[w] * dirL));
        float dotHV = dot(dirV, m);
        if (dotHV <= 0)
            return 0.0f;
        float mPDF = ggx.evaluatePDF(dirV, m);
        SS<N> F = fresnel.evaluate(dotHV);
        float reflectProb = F.importance(q.wlHint);
        if (DT::isReflection(q.dirTypeFilter)) reflectProb = 1.0f;
        if (DT::isTransmission(q.dirTypeFilter)) reflectProb = 0.0f;
        if (dotNVdotNL > 0) {
            float common = reflectProb / (4 * dotHV);
            if (revValue) *revValue = common * ggx.evaluatePDF(dirL, m);
            return common * mPDF;
        }
        float dotHL = dot(dirL, m);
        float common = (1 - reflectProb) / pow2(eEnter[w] * dotHV + eExit[w] * dotHL);
        if (revValue) *revValue = common * ggx.evaluatePDF(dirL, m) * pow2(eEnter[w]) * std::fabs(dotHV);
        return common * mPDF * pow2(eExit[w]) * std::fabs(dotHL);
    }

    // materials.cu:817-910
    SS<N> sampleSpecularBSDF(const BSDFQuery<N> &q, float uComponent, BSDFResult<N> *res, RevResult<N> *rev) const {
        if (rev) rev->dirPDF = 0.0f;
        bool entering = q.dirLocal.z >= 0.0f;
        const SS<N> &eEnter = entering ? c1 : c2;
        const SS<N> &eExit = entering ? c2 : c1;
        FresnelDielectric<N> fresnel{eEnter, eExit};
        V3 dirV = entering ? q.dirLocal : -q.dirLocal;
        SS<N> F = fresnel.evaluate(dirV.z);
        float reflectProb = F.importance(q.wlHint);
        if (DT::isReflection(q.dirTypeFilter)) reflectProb = 1.0f;
        if (DT::isTransmission(q.dirTypeFilter)) reflectProb = 0.0f;
        if (uComponent < reflectProb) {
            if (dirV.z == 0.0f) {
                res->dirPDF = 0.0f;
                return SS<N>(0.0f);
            }
            V3 dirL(-dirV.x, -dirV.y, dirV.z);
            res->dirLocal = entering ? dirL : -dirL;
            res->dirPDF = reflectProb;
            res->sampledType = DT::Reflection | DT::Delta0D;
            SS<N> ret = c0 * F * (1.0f / std::fabs(dirV.z));
            if (rev) {
                rev->value = ret;
                rev->dirPDF = res->dirPDF;
            }
            return ret;
        }
        float sinEnter2 = 1.0f - dirV.z * dirV.z;
        float recRelIOR = eEnter[q.wlHint] / eExit[q.wlHint];
        float sinExit2 = recRelIOR * recRelIOR * sinEnter2;
        if (sinExit2 >= 1.0f) {
            res->dirPDF = 0.0f;
            return SS<N>(0.0f);
        }
        float cosExit = std::sqrt(std::fmax(0.0f, 1.0f - sinExit2));
        V3 dirL(recRelIOR * -dirV.x, recRelIOR * -dirV.y, -cosExit);
        res->dirLocal = entering ? dirL : -dirL;
        res->dirPDF = (1.0f - reflectProb);
        res->sampledType = DT::Transmission | DT::Delta0D | (dispersive ? uint32_t(DT::Dispersive) : 0u);
        SS<N> ret(0.0f);
        ret[q.wlHint] = c0[q.wlHint] * (1.0f - F[q.wlHint]);
        if (rev) {
            float revSqueeze = 1.0f;
            if (q.transportMode == TransportMode::Importance)
                revSqueeze = pow2(eExit[q.wlHint] / eEnter[q.wlHint]);
            SS<N> revRet = ret;
            revRet[q.wlHint] *= revSqueeze / std::fabs(q.dirLocal.z);
            rev->value = revRet;
            rev->dirPDF = revSqueeze * res->dirPDF;
        }
        float squeeze = 1.0f;
        if (q.transportMode == TransportMode::Radiance)
            squeeze *= pow2(eEnter[q.wlHint] / eExit[q.wlHint]);
        ret[q.wlHint] *= squeeze / std::fabs(cosExit);
        res->dirPDF *= squeeze;
        return ret;
    }

    // materials.cu:1089-1140
    SS<N> sampleMicrofacetBRDF(const BSDFQuery<N> &q, const float uDir[2], BSDFResult<N> *res, RevResult<N> *rev) const {
        if (rev) rev->dirPDF = 0.0f;
        bool entering = q.dirLocal.z >= 0.0f;
        FresnelConductor<N> fr{c0, c1};
        GGX ggx(a0, a1, a2);
        V3 dirV = entering ? q.dirLocal : -q.dirLocal;
        V3 m;
        float mPDF;
        float D = ggx.sample(dirV, uDir[0], uDir[1], &m, &mPDF);
        float dotHV = dot(dirV, m);
        if (dotHV <= 0) {
            res->dirPDF = 0.0f;
            return SS<N>(0.0f);
        }
        V3 dirL = 2 * dotHV * m - dirV;
        res->dirLocal = entering ? dirL : -dirL;
        if (dirL.z * dirV.z <= 0) {
            res->dirPDF = 0.0f;
            return SS<N>(0.0f);
        }
        float common = 1.0f / (4 * dotHV);
        res->dirPDF = common * mPDF;
        res->sampledType = DT::Reflection | DT::HighFreq;
        SS<N> F = fr.evaluate(dotHV);
        float G = ggx.evaluateSmithG1(dirV, m) * ggx.evaluateSmithG1(dirL, m);
        SS<N> fs = F * D * G / (4 * dirV.z * dirL.z);
        if (rev) {
            rev->value = fs;
            rev->dirPDF = common * ggx.evaluatePDF(dirL, m);
        }
        return fs;
    }

    // materials.cu:1713-1832
    SS<N> sampleDiffSpec(const BSDFQuery<N> &q, float uComponent, const float uDir[2], BSDFResult<N> *res,
                         RevResult<N> *rev) const {
        if (rev) rev->dirPDF = 0.0f;
        const float rough = a0;
        float alpha = pow2(rough);
        GGX ggx(alpha, alpha, 0.0f);
        bool entering = q.dirLocal.z >= 0.0f;
        V3 dirL;
        V3 dirV = entering ? q.dirLocal : -q.dirLocal;
        float expectedF_D90 = 0.5f * rough + 2 * rough * pow2(q.dirLocal.z);
        float oneMinusDotVN5 = pow5(1 - dirV.z);
        float expectedDiffuseFresnel = lerp(1.0f, expectedF_D90, oneMinusDotVN5);
        float iBaseColor = c0.importance(q.wlHint) * pow2(expectedDiffuseFresnel) * lerp(1.0f, 1.0f / 1.51f, rough);
        float expectedOneMinusDotVH5 = pow5(1 - dirV.z);
        float iSpecularF0 = c1.importance(q.wlHint);
        float diffuseWeight = iBaseColor;
        float specularWeight = lerp(iSpecularF0, 1.0f, expectedOneMinusDotVH5);
        float weights[] = {diffuseWeight, specularWeight};
        float probSelection, sumWeights = 0.0f;
        uint32_t component = sampleDiscrete(weights, 2, uComponent, &probSelection, &sumWeights, &uComponent);
        float diffuseDirPDF = 0, specularDirPDF = 0;
        V3 m;
        float dotLH = 0, D = 0;
        if (component == 0) {
            res->sampledType = DT::Reflection | DT::LowFreq;
            dirL = cosineSampleHemisphere(uDir[0], uDir[1]);
            diffuseDirPDF = dirL.z / kPi;
            m = halfVector(dirL, dirV);
            dotLH = dot(dirL, m);
            float common = 1.0f / (4 * dotLH);
            specularDirPDF = common * ggx.evaluatePDF(dirV, m);
            D = ggx.evaluate(m);
        } else if (component == 1) {
            res->sampledType = DT::Reflection | DT::HighFreq;
            float mPDF;
            D = ggx.sample(dirV, uDir[0], uDir[1], &m, &mPDF);
            float dotVH = dot(dirV, m);
            dotLH = dotVH;
            dirL = 2 * dotVH * m - dirV;
            if (dirL.z * dirV.z <= 0) {
                res->dirPDF = 0.0f;
                return SS<N>(0.0f);
            }
            float common = 1.0f / (4 * dotLH);
            specularDirPDF = common * mPDF;
            diffuseDirPDF = dirL.z / kPi;
        }
        float oneMinusDotLH5 = pow5(1 - dotLH);
        float G = ggx.evaluateHeightCorrelatedSmithG(dirL, dirV, m);
        SS<N> F = lerpS(c1, SS<N>(1.0f), oneMinusDotLH5);
        float microfacetDenom = 4 * dirL.z * dirV.z;
        SS<N> specularValue = F * ((D * G) / microfacetDenom);
        float F_D90 = 0.5f * rough + 2 * rough * dotLH * dotLH;
        float oneMinusDotLN5 = pow5(1 - dirL.z);
        float diffuseFresnelOut = lerp(1.0f, F_D90, oneMinusDotVN5);
        float diffuseFresnelIn = lerp(1.0f, F_D90, oneMinusDotLN5);
        SS<N> diffuseValue = c0 * (diffuseFresnelOut * diffuseFresnelIn * lerp(1.0f, 1.0f / 1.51f, rough) / kPi);
        SS<N> ret = diffuseValue + specularValue;
        res->dirLocal = entering ? dirL : -dirL;
        res->dirPDF = (diffuseDirPDF * diffuseWeight + specularDirPDF * specularWeight) / sumWeights;
        if (rev) {
            float revDiffuseDirPDF = dirV.z / kPi;
            float common = 1.0f / (4 * dotLH);
            float revSpecularDirPDF = common * ggx.evaluatePDF(dirL, m);
            float rF_D90 = 0.5f * rough + 2 * rough * pow2(dirL.z);
            float rOneMinusDotVN5 = pow5(1 - dirL.z);
            float rDiffFresnel = lerp(1.0f, rF_D90, rOneMinusDotVN5);
            float rIBase = c0.importance(q.wlHint) * pow2(rDiffFresnel) * lerp(1.0f, 1.0f / 1.51f, rough);
            float rOneMinusDotVH5 = pow5(1 - dirL.z);
            float rISpecF0 = c1.importance(q.wlHint);
            float rDiffW = rIBase;
            float rSpecW = lerp(rISpecF0, 1.0f, rOneMinusDotVH5);
            rev->value = ret;
            rev->dirPDF = (revDiffuseDirPDF * rDiffW + revSpecularDirPDF * rSpecW) / (rDiffW + rSpecW);
        }
        return ret;
    }

    // materials.cu:1834-1875
    SS<N> evaluateDiffSpec(const BSDFQuery<N> &q, const V3 &dir, SS<N> *revValue) const {
        const float rough = a0;
        float alpha = pow2(rough);
        GGX ggx(alpha, alpha, 0.0f);
        if (dir.z * q.dirLocal.z <= 0)
            return SS<N>(0.0f);
        bool entering = q.dirLocal.z >= 0.0f;
        V3 dirV = entering ? q.dirLocal : -q.dirLocal;
        V3 dirL = entering ? dir : -dir;
        V3 m = halfVector(dirL, dirV);
        float dotLH = dot(dirL, m);
        float oneMinusDotLH5 = pow5(1 - dotLH);
        float D = ggx.evaluate(m);
        float G = ggx.evaluateHeightCorrelatedSmithG(dirL, dirV, m);
        SS<N> F = lerpS(c1, SS<N>(1.0f), oneMinusDotLH5);
        float microfacetDenom = 4 * dirL.z * dirV.z;
        SS<N> specularValue = F * ((D * G) / microfacetDenom);
        float F_D90 = 0.5f * rough + 2 * rough * pow2(dotLH);
        float oneMinusDotVN5 = pow5(1 - dirV.z);
        float oneMinusDotLN5 = pow5(1 - dirL.z);
        float diffuseFresnelOut = lerp(1.0f, F_D90, oneMinusDotVN5);
        float diffuseFresnelIn = lerp(1.0f, F_D90, oneMinusDotLN5);
        SS<N> diffuseValue = c0 * (diffuseFresnelOut * diffuseFresnelIn * lerp(1.0f, 1.0f / 1.51f, rough) / kPi);
        SS<N> ret = diffuseValue + specularValue;
        if (revValue) *revValue = ret;
        return ret;
    }

    // materials.cu:1877-1932
    float evaluatePDFDiffSpec(const BSDFQuery<N> &q, const V3 &dir, float *revValue) const {
        if (revValue) *revValue = 0.0f;
        const float rough = a0;
        float alpha = pow2(rough);
        GGX ggx(alpha, alpha, 0.0f);
        bool entering = q.dirLocal.z >= 0.0f;
        V3 dirV = entering ? q.dirLocal : -q.dirLocal;
        V3 dirL = entering ? dir : -dir;
        V3 m = halfVector(dirL, dirV);
        float dotLH = dot(dirL, m);
        float common = 1.0f / (4 * dotLH);
        float expectedF_D90 = 0.5f * rough + 2 * rough * q.dirLocal.z * q.dirLocal.z;
        float oneMinusDotVN5 = pow5(1 - dirV.z);
        float expectedDiffuseFresnel = lerp(1.0f, expectedF_D90, oneMinusDotVN5);
        float iBaseColor = c0.importance(q.wlHint) * expectedDiffuseFresnel * expectedDiffuseFresnel * lerp(1.0f, 1.0f / 1.51f, rough);
        float expectedOneMinusDotVH5 = pow5(1 - dirV.z);
        float iSpecularF0 = c1.importance(q.wlHint);
        float diffuseWeight = iBaseColor;
        float specularWeight = lerp(iSpecularF0, 1.0f, expectedOneMinusDotVH5);
        float sumWeights = diffuseWeight + specularWeight;
        float diffuseDirPDF = std::max(dirL.z / kPi, 0.0f);
        float specularDirPDF = common * ggx.evaluatePDF(dirV, m);
        float ret = (diffuseDirPDF * diffuseWeight + specularDirPDF * specularWeight) / sumWeights;
        if (revValue) {
            float revDiffuseDirPDF = dirV.z / kPi;
            float revSpecularDirPDF = common * ggx.evaluatePDF(dirL, m);
            float rF_D90 = 0.5f * rough + 2 * rough * pow2(dirL.z);
            float rOneMinusDotVN5 = pow5(1 - dirL.z);
            float rDiffFresnel = lerp(1.0f, rF_D90, rOneMinusDotVN5);
            float rIBase = c0.importance(q.wlHint) * pow2(rDiffFresnel) * lerp(1.0f, 1.0f / 1.51f, rough);
            float rOneMinusDotVH5 = pow5(1 - dirL.z);
            float rISpecF0 = c1.importance(q.wlHint);
            float rDiffW = rIBase;
            float rSpecW = lerp(rISpecF0, 1.0f, rOneMinusDotVH5);
            *revValue = (revDiffuseDirPDF * rDiffW + revSpecularDirPDF * rSpecW) / (rDiffW + rSpecW);
        }
        return ret;
    }
};

enum class EDFKind { Null = 0, Diffuse, Point, Environment };

// EDFs (materials.cu:2040-2096, 2194-2250, 2830-2896)
template <int N>
struct EDF {
    EDFKind kind = EDFKind::Null;
    SS<N> emittance;
    float worldRadius = 0, worldDiscArea = 0;

    bool matches(uint32_t flags) const {
        if (kind == EDFKind::Null) return false;
        return DT::matches(DT::Emission | DT::LowFreq, flags);
    }
    SS<N> evaluateEmittance() const {
        if (kind == EDFKind::Environment) return kPi * emittance;
        if (kind == EDFKind::Null) return SS<N>(0.0f);
        return emittance;
    }
    SS<N> evaluate(const V3 &dirLocal) const {
        switch (kind) {
        case EDFKind::Diffuse:
        case EDFKind::Environment: return SS<N>(dirLocal.z > 0.0f ? 1.0f / kPi : 0.0f);
        case EDFKind::Point: return SS<N>(1.0f / (4 * kPi));
        default: return SS<N>(0.0f);
        }
    }
    float evaluatePDF(const V3 &dirLocal) const {
        switch (kind) {
        case EDFKind::Diffuse: return dirLocal.z <= 0.0f ? 0.0f : dirLocal.z / kPi;
        case EDFKind::Point: return 1.0f / (4 * kPi);
        case EDFKind::Environment: return dirLocal.z <= 0.0f ? 0.0f : 1.0f / worldDiscArea;
        default: return 0.0f;
        }
    }
    SS<N> sample(float uDir0, float uDir1, V3 *dirLocal, float *dirPDF, uint32_t *sampledType) const {
        *sampledType = DT::Emission | DT::LowFreq;
        switch (kind) {
        case EDFKind::Diffuse:
            *dirLocal = cosineSampleHemisphere(uDir0, uDir1);
            *dirPDF = dirLocal->z / kPi;
            return SS<N>(1.0f / kPi);
        case EDFKind::Point:
            *dirLocal = uniformSampleSphere(uDir0, uDir1);
            *dirPDF = 1.0f / (4 * kPi);
            return SS<N>(1.0f / (4 * kPi));
        case EDFKind::Environment: {
            float dx, dy;
            concentricSampleDisk(uDir0, uDir1, &dx, &dy);
            *dirLocal = V3(worldRadius * dx, worldRadius * dy, 1);
            *dirPDF = 1.0f / worldDiscArea;
            return SS<N>(1 / kPi);
        }
        default:
            *dirPDF = 0;
            *sampledType = 0;
            return SS<N>(0.0f);
        }
    }
};

// shared.h:880-903 and cameras.cu:43-76, 124-222
struct PerspectiveCamera {
    V3 position;
    Quat orientation;
    float sensitivity, aspect, fovY, lensRadius, imgPlaneDistance, objPlaneDistance;
    float opWidth, opHeight, imgPlaneArea;
    void setImagePlaneArea() {
        opHeight = 2.0f * objPlaneDistance * std::tan(fovY * 0.5f);
        opWidth = opHeight * aspect;
        imgPlaneDistance = 1.0f;
        imgPlaneArea = opWidth * opHeight * pow2(imgPlaneDistance / objPlaneDistance);
    }
};

} // namespace ov
