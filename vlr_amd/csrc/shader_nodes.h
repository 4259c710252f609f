// Shader nodes evaluated per hit (calcNode, kernel_common.h:658-737, with the implicit plug
// conversions of shared.h:390-470; the node programs of shader_nodes.cu:20-483) and the
// material set-up at a surface point (the *_setupBSDF callables of materials.cu), which
// evaluates every parameter connected to a node.
//
// A material's connected parameters (and a material group's normal and tangent nodes) are
// compiled on the host into one node program: the graph's nodes in post-order, each step
// reading its inputs from earlier steps' slots.  The device runs it in one loop
// (runNodeProgram), without calls or recursion.  Kernels come in two variants (NODES): scenes
// without node programs run code that contains none of this.
// Textures are float4 texel arrays prepared on the host for the render mode and sampled
// here with CUDA's normalised-coordinate conventions: linear filtering
// at u * width - 0.5, the four wrap modes applied to texel indices, gathers in the order
// (i0, j1), (i1, j1), (i1, j0), (i0, j0).
#pragma once

#include "color.h"
#include "dev_shading.h"

namespace vlra {

// value classes of the plug types (the C++ type a node program returns)
enum : uint32_t { VC_F1 = 0, VC_F2, VC_F3, VC_F4, VC_P3, VC_V3, VC_N3, VC_SPEC };
__host__ __device__ __forceinline__ uint32_t valueClass(uint32_t plugType) {
    switch (plugType) {
    case PLUG_FLOAT1: case PLUG_ALPHA: return VC_F1;
    case PLUG_FLOAT2: return VC_F2;
    case PLUG_FLOAT3: return VC_F3;
    case PLUG_FLOAT4: return VC_F4;
    case PLUG_POINT3D: case PLUG_TEXCOORD: return VC_P3;
    case PLUG_VECTOR3D: return VC_V3;
    case PLUG_NORMAL3D: return VC_N3;
    default: return VC_SPEC;
    }
}

// NodeTypeInfo<want>::convertFrom<src> (shared.h:400-436); false where it is not defined
__device__ __forceinline__ bool convertValue(uint32_t want, uint32_t src, const float4 &v, float4 *out) {
    auto norm3 = [](const float4 &a) {
        const float l = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
        return make_float4(a.x / l, a.y / l, a.z / l, 0.0f);
    };
    switch (want) {
    case VC_F1:
        if (src == VC_F1) { *out = v; return true; }
        return false;
    case VC_F2:
        if (src == VC_F2) { *out = v; return true; }
        if (src == VC_F1) { *out = make_float4(v.x, v.x, 0, 0); return true; }
        return false;
    case VC_F3:
        if (src == VC_F3 || src == VC_P3 || src == VC_V3 || src == VC_N3) { *out = make_float4(v.x, v.y, v.z, 0); return true; }
        if (src == VC_F1) { *out = make_float4(v.x, v.x, v.x, 0); return true; }
        return false;
    case VC_F4:
        if (src == VC_F4) { *out = v; return true; }
        if (src == VC_F1) { *out = make_float4(v.x, v.x, v.x, v.x); return true; }
        return false;
    case VC_P3:
    case VC_V3:
        if (src == VC_P3 || src == VC_V3 || src == VC_N3 || src == VC_F3) { *out = make_float4(v.x, v.y, v.z, 0); return true; }
        return false;
    case VC_N3:
        if (src == VC_N3) { *out = make_float4(v.x, v.y, v.z, 0); return true; }
        if (src == VC_F3 || src == VC_P3 || src == VC_V3) { *out = norm3(v); return true; }
        return false;
    default:
        return false;
    }
}

// ---- textures ----
// wrap modes: 0 repeat, 1 clamp to edge, 2 mirror, 3 clamp to border (texel outside -> 0)
__device__ __forceinline__ int wrapTexel(int i, int n, uint32_t mode, bool *outside) {
    switch (mode) {
    case 0: {
        int m = i % n;
        return m < 0 ? m + n : m;
    }
    case 2: {
        int p = 2 * n, m = i % p;
        if (m < 0) m += p;
        return m < n ? m : p - 1 - m;
    }
    case 3:
        if (i < 0 || i >= n) *outside = true;
        return i < 0 ? 0 : (i >= n ? n - 1 : i);
    default:
        return i < 0 ? 0 : (i >= n ? n - 1 : i);
    }
}

__device__ __forceinline__ float4 texelAt(const DevScene &sc, const DevTexture &T, int x, int y, uint32_t wrapU,
                                          uint32_t wrapV) {
    bool outside = false;
    const int ix = wrapTexel(x, static_cast<int>(T.width), wrapU, &outside);
    const int iy = wrapTexel(y, static_cast<int>(T.height), wrapV, &outside);
    if (outside) return make_float4(0, 0, 0, 0);
    return sc.texels[T.offset + static_cast<uint32_t>(iy) * T.width + static_cast<uint32_t>(ix)];
}

// sampler bits: filter (0 nearest, 1 linear) | wrap u << 2 | wrap v << 4
__device__ __forceinline__ float4 sampleTexture(const DevScene &sc, const DevTexture &T, uint32_t sampler, float u,
                                                float v) {
    const uint32_t wu = (sampler >> 2) & 3, wv = (sampler >> 4) & 3;
    if ((sampler & 3) == 0) {
        const int x = static_cast<int>(floorf(u * T.width)), y = static_cast<int>(floorf(v * T.height));
        return texelAt(sc, T, x, y, wu, wv);
    }
    const float fx = u * T.width - 0.5f, fy = v * T.height - 0.5f;
    const float x0f = floorf(fx), y0f = floorf(fy);
    const float a = fx - x0f, b = fy - y0f;
    const int x0 = static_cast<int>(x0f), y0 = static_cast<int>(y0f);
    const float4 t00 = texelAt(sc, T, x0, y0, wu, wv), t10 = texelAt(sc, T, x0 + 1, y0, wu, wv);
    const float4 t01 = texelAt(sc, T, x0, y0 + 1, wu, wv), t11 = texelAt(sc, T, x0 + 1, y0 + 1, wu, wv);
    const float w00 = (1 - a) * (1 - b), w10 = a * (1 - b), w01 = (1 - a) * b, w11 = a * b;
    return make_float4(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
                       w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
                       w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z,
                       w00 * t00.w + w10 * t10.w + w01 * t01.w + w11 * t11.w);
}

// tex2Dgather of component `comp`: x (i0, j1), y (i1, j1), z (i1, j0), w (i0, j0)
__device__ __forceinline__ float4 gatherTexture(const DevScene &sc, const DevTexture &T, uint32_t sampler, float u,
                                                float v, uint32_t comp) {
    const uint32_t wu = (sampler >> 2) & 3, wv = (sampler >> 4) & 3;
    const int x0 = static_cast<int>(floorf(u * T.width - 0.5f)), y0 = static_cast<int>(floorf(v * T.height - 0.5f));
    auto c = [&](const float4 &t) { return comp == 0 ? t.x : (comp == 1 ? t.y : (comp == 2 ? t.z : t.w)); };
    return make_float4(c(texelAt(sc, T, x0, y0 + 1, wu, wv)), c(texelAt(sc, T, x0 + 1, y0 + 1, wu, wv)),
                       c(texelAt(sc, T, x0 + 1, y0, wu, wv)), c(texelAt(sc, T, x0, y0, wu, wv)));
}

__device__ __forceinline__ float comp4(const float4 &t, uint32_t k) {
    return k == 0 ? t.x : (k == 1 ? t.y : (k == 2 ? t.z : t.w));
}

// ---- spectra ----
template <int N>
__device__ __forceinline__ Spec<N> specFromTriplet(const DevScene &sc, uint32_t type, uint32_t cs, float r, float g,
                                                   float b, const Wls<N> &wls) {
    if constexpr (N == 3) {
        Spec<3> s;
        s[0] = r;
        s[1] = g;
        s[2] = b;
        return s;
    } else {
        const DevSpectrum d = colorToDevSpectrum(type, cs, r, g, b, true, MengTables{sc.mengGrid, sc.mengPoints});
        return evalSpectrum<N>(sc, d, wls);
    }
}

// RegularSampledSpectrum / IrregularSampledSpectrum::evaluate (spectrum_types.cpp:394-412, 496-518)
template <int N>
__device__ __forceinline__ Spec<N> sampledSpectrum(const DevScene &sc, const DevShaderNode &nd, const Wls<N> &wls) {
    Spec<N> r(0.0f);
    const float *vals = sc.nodePool + nd.a[0];
    const uint32_t n = nd.a[1];
    if (nd.kind == NODE_REGULAR_SPECTRUM) {
        for (int i = 0; i < N; ++i) {
            const float binF = (wls.lambda(i) - nd.f[0]) / (nd.f[1] - nd.f[0]) * (n - 1);
            if (binF <= 0.0f) {
                r[i] = vals[0];
            } else if (binF >= n - 1) {
                r[i] = vals[n - 1];
            } else {
                const int bin = static_cast<int>(binF);
                const float t = binF - bin;
                r[i] = (1 - t) * vals[bin] + t * vals[bin + 1];
            }
        }
        return r;
    }
    const float *lambdas = vals, *values = vals + n;
    uint32_t base = 0;
    for (int i = 0; i < N; ++i) {
        const float wl = wls.lambda(i);
        uint32_t p2 = 1;
        while (p2 < n - base) p2 <<= 1;
        for (uint32_t delta = p2 >> 1; delta > 0; delta >>= 1)
            if (base + delta < n && lambdas[base + delta] <= wl) base += delta;
        if (base >= n - 1) {
            r[i] = values[n - 1];
            continue;
        }
        const float t = (wl - lambdas[base]) / (lambdas[base + 1] - lambdas[base]);
        r[i] = t <= 0.0f ? values[base] : (1 - t) * values[base] + t * values[base + 1];
    }
    return r;
}

// ---- node programs ----
// Outputs of a program: up to three spectra and two vectors (converted to the classes the
// caller asks for: a material's float3 parameter, a group's normal and tangent).
template <int N>
struct NodeOutputs {
    Spec<N> spec[3];
    float4 vec[2];
    uint32_t has;   // bits 0..2 spectra, 3..4 vectors
};

// Runs the program of `count` steps at `base` (the node programs of shader_nodes.cu:20-483,
// with calcNode's conversions for every input and output).  SPEC = false skips spectrum steps
// (programs of normals and tangents have none).
template <int N, bool SPEC>
__device__ __forceinline__ void runNodeProgram(const DevScene &sc, uint32_t base, uint32_t count, const SurfPt &sp,
                                               const Wls<N> &wls, uint32_t wantVec0, uint32_t wantVec1,
                                               NodeOutputs<N> *out) {
    float4 slot[kNodeProgramSlots];
    uint32_t cls[kNodeProgramSlots];
    out->has = 0;
    count = count < kNodeProgramSlots ? count : kNodeProgramSlots;
    for (uint32_t s = 0; s < count; ++s) {
        const DevNodeStep st = sc.nodeSteps[base + s];
        const DevShaderNode &nd = sc.nodes[st.node];
        const uint32_t type = st.plug & 0xF, opt = st.plug >> 4;
        // input k converted to class `want`, or `dflt` where unconnected or not convertible
        auto input = [&](int k, uint32_t want, const float4 &dflt) {
            const uint32_t ref = st.in[k];
            if (ref == 0 || ref > s) return dflt;
            float4 o;
            return convertValue(want, cls[ref - 1], slot[ref - 1], &o) ? o : dflt;
        };
        const float4 zero = make_float4(0, 0, 0, 0);
        float4 v = zero;
        cls[s] = valueClass(type);
        if (type == PLUG_SPECTRUM) {
            if constexpr (SPEC) {
                Spec<N> r(0.0f);
                bool ok = true;
                switch (nd.kind) {
                case NODE_TRIPLET_SPECTRUM:
                    r = evalSpectrum<N>(sc, DevSpectrum{nd.f[0], nd.f[1], nd.f[2], nd.f[3]}, wls);
                    break;
                case NODE_REGULAR_SPECTRUM:
                case NODE_IRREGULAR_SPECTRUM:
                    if (N == 3 || nd.a[2]) r = evalSpectrum<N>(sc, DevSpectrum{nd.f[0], nd.f[1], nd.f[2], nd.f[3]}, wls);
                    else r = sampledSpectrum<N>(sc, nd, wls);
                    break;
                case NODE_FLOAT3_TO_SPECTRUM: {
                    // shader_nodes.cu:254-274
                    const float4 f3 = input(0, VC_F3, make_float4(nd.f[0], nd.f[1], nd.f[2], 0));
                    auto c = [](float x) { return fminf(fmaxf(0.5f * x + 0.5f, 0.0f), 1.0f); };
                    r = specFromTriplet<N>(sc, nd.a[0], nd.a[1], c(f3.x), c(f3.y), c(f3.z), wls);
                    break;
                }
                case NODE_IMAGE_TEXTURE: {
                    // shader_nodes.cu:413-452
                    const DevTexture &T = sc.textures[nd.a[0]];
                    const float4 tc = input(0, VC_P3, make_float4(sp.texU, sp.texV, 0, 0));
                    float4 t = sampleTexture(sc, T, nd.a[1], tc.x, tc.y);
                    if (T.gray) t.y = t.z = t.x;
                    if (N == 4 && (T.format == kTexUVS8 || T.format == kTexUVS16)) {
                        float u = t.x, vv = t.y, sc3 = t.z;
                        if (T.format == kTexUVS8) {
                            u *= 12.0f;
                            vv *= 14.0f;
                            sc3 *= 3.0f;
                        }
                        sc3 /= 0.009355121400914532f;
                        r = evalSpectrum<N>(sc, mengFromUV(MengTables{sc.mengGrid, sc.mengPoints}, u, vv, sc3), wls);
                    } else {
                        r = specFromTriplet<N>(sc, T.spectrumType, T.colorSpace, t.x, t.y, t.z, wls);
                    }
                    break;
                }
                case NODE_ENV_TEXTURE: {
                    // shader_nodes.cu:457-483: read at the surface's texture coordinates; only a
                    // uvsA16Fx4 image is taken as Meng coordinates, anything else as a light-source
                    // triplet in the image's colour space, and the RGB build takes the texel as is
                    const DevTexture &T = sc.textures[nd.a[0]];
                    const float4 t = sampleTexture(sc, T, nd.a[1], sp.texU, sp.texV);
                    if constexpr (N == 3) {
                        r[0] = t.x;
                        r[1] = t.y;
                        r[2] = t.z;
                    } else if (T.format == kTexUVS16) {
                        r = evalSpectrum<N>(
                            sc, mengFromUV(MengTables{sc.mengGrid, sc.mengPoints}, t.x, t.y, t.z / 0.009355121400914532f),
                            wls);
                    } else {
                        r = specFromTriplet<N>(sc, kSpecLightSource, T.colorSpace, t.x, t.y, t.z, wls);
                    }
                    break;
                }
                default:
                    ok = false;
                    break;
                }
                if (ok)
                    for (int k = 0; k < 3; ++k)
                        if (st.out & (1u << k)) {
                            out->spec[k] = r;
                            out->has |= 1u << k;
                        }
            }
            slot[s] = zero;
            continue;
        }
        switch (nd.kind) {
        case NODE_GEOMETRY:
            if (type == PLUG_POINT3D) {
                v = make_float4(sp.position.x, sp.position.y, sp.position.z, 0);
            } else if (type == PLUG_NORMAL3D) {
                const V3 n = opt == 0 ? sp.gn : (opt == 1 ? sp.frame.z : V3(0.0f));
                v = make_float4(n.x, n.y, n.z, 0);
            } else if (type == PLUG_VECTOR3D) {
                const V3 t = opt == 0 ? sp.frame.x : (opt == 1 ? sp.frame.y : V3(0.0f));
                v = make_float4(t.x, t.y, t.z, 0);
            } else if (type == PLUG_TEXCOORD) {
                v = make_float4(sp.texU, sp.texV, 0, 0);
            }
            break;
        case NODE_TANGENT: {
            // shader_nodes.cu:56-87
            if (nd.a[0] == 0) {
                v = make_float4(sp.frame.x.x, sp.frame.x.y, sp.frame.x.z, 0);
            } else {
                const DevInstance &I = sc.instances[sp.instIndex];
                const V3 lp = xfmPoint(I.w2o, sp.position);
                // radial about x, y or z: the axis cross the local position, computed without
                // branches (gfx950 code for the three-way branch over a per-lane type lost the
                // Radial Z case's first component under divergence)
                const V3 axis(nd.a[0] == 1 ? 1.0f : 0.0f, nd.a[0] == 2 ? 1.0f : 0.0f, nd.a[0] == 3 ? 1.0f : 0.0f);
                const V3 lt = cross(axis, lp);
                const V3 t = xfmVector(I.o2w, lt);
                v = make_float4(t.x, t.y, t.z, 0);
            }
            break;
        }
        case NODE_FLOAT2:
        case NODE_FLOAT3:
        case NODE_FLOAT4: {
            // shader_nodes.cu:91-219: component opt .. opt + width - 1
            const uint32_t n = nd.kind == NODE_FLOAT2 ? 2 : (nd.kind == NODE_FLOAT3 ? 3 : 4);
            const uint32_t width = type == PLUG_FLOAT1 ? 1 : (type == PLUG_FLOAT2 ? 2 : (type == PLUG_FLOAT3 ? 3 : 4));
            if (opt + width <= n) {
                float c[4] = {0, 0, 0, 0};
                for (uint32_t k = 0; k < width; ++k)
                    c[k] = input(static_cast<int>(opt + k), VC_F1, make_float4(nd.f[opt + k], 0, 0, 0)).x;
                v = make_float4(c[0], c[1], c[2], c[3]);
            }
            break;
        }
        case NODE_SCALE_OFFSET_FLOAT: {
            // shader_nodes.cu:223-231
            const float value = input(0, VC_F1, zero).x;
            const float scale = input(1, VC_F1, make_float4(nd.f[1], 0, 0, 0)).x;
            const float offset = input(2, VC_F1, make_float4(nd.f[2], 0, 0, 0)).x;
            v = make_float4(scale * value + offset, 0, 0, 0);
            break;
        }
        case NODE_UV_MAP:
            // shader_nodes.cu:278-285
            v = make_float4(nd.f[0] * sp.texU + nd.f[2], nd.f[1] * sp.texV + nd.f[3], 0.0f, 0);
            break;
        case NODE_IMAGE_TEXTURE: {
            // shader_nodes.cu:289-411, 456-474
            const DevTexture &T = sc.textures[nd.a[0]];
            const float4 tc = input(0, VC_P3, make_float4(sp.texU, sp.texV, 0, 0));
            const uint32_t bump = (nd.a[1] >> 8) & 3;
            if (type == PLUG_NORMAL3D) {
                V3 n(0.0f, 0.0f, 1.0f);
                if (bump != 2) {
                    const float4 t = sampleTexture(sc, T, nd.a[1], tc.x, tc.y);
                    if (opt < 2) {
                        n = opt == 0 ? V3(t.x, t.y, t.z) : V3(t.y, t.z, t.w);
                        n = V3(2 * n.x - 1.0f, 2 * n.y - 1.0f, 2 * n.z - 1.0f);
                        if (bump == 0) n.y *= -1;
                    }
                } else {
                    const float4 t = gatherTexture(sc, T, nd.a[1], tc.x, tc.y, opt);
                    const float coeff = 5.0f / 1024;
                    const float dhdu = (coeff * T.width) * (t.y - t.x);
                    const float dhdv = (coeff * T.height) * (t.x - t.w);
                    n = V3(-dhdu, dhdv, 1);
                }
                n.x *= nd.f[0];
                n.y *= nd.f[0];
                n = normalize(n);
                v = make_float4(n.x, n.y, n.z, 0);
            } else {
                const float4 t = sampleTexture(sc, T, nd.a[1], tc.x, tc.y);
                if ((type == PLUG_FLOAT1 || type == PLUG_ALPHA) && opt < 4) v = make_float4(comp4(t, opt), 0, 0, 0);
                else if (type == PLUG_FLOAT2 && opt < 3) v = make_float4(comp4(t, opt), comp4(t, opt + 1), 0, 0);
                else if (type == PLUG_FLOAT3 && opt < 2)
                    v = make_float4(comp4(t, opt), comp4(t, opt + 1), comp4(t, opt + 2), 0);
                else if (type == PLUG_FLOAT4) v = t;
            }
            break;
        }
        default:
            break;
        }
        slot[s] = v;
        for (int k = 0; k < 2; ++k)
            if (st.out & (8u << k)) {
                float4 o;
                if (convertValue(k == 0 ? wantVec0 : wantVec1, cls[s], v, &o)) {
                    out->vec[k] = o;
                    out->has |= 8u << k;
                }
            }
    }
}

// ---- surface point: normal and tangent nodes of the material group (renderer_common.h:554-626) ----
__device__ __forceinline__ void applyBumpMapping(const V3 &n, SurfPt *sp) {
    if (n.x == 0.0f && n.y == 0.0f) return;
    const float projLength = sqrtf(n.x * n.x + n.y * n.y);
    const float tiltAngle = atanf(projLength / n.z);
    float qSin, qCos;
    sincosf(tiltAngle / 2, &qSin, &qCos);
    const float qX = (-n.y / projLength) * qSin;
    const float qY = (n.x / projLength) * qSin;
    const float qW = qCos;
    const V3 mt(1 - 2 * qY * qY, 2 * qX * qY, -2 * qY * qW);
    const V3 mb(2 * qX * qY, 1 - 2 * qX * qX, 2 * qX * qW);
    const Frame f = sp->frame;
    sp->frame.x = f.fromLocal(mt);
    sp->frame.y = f.fromLocal(mb);
    sp->frame.z = f.fromLocal(n);
}

__device__ __forceinline__ void modifyTangent(const V3 &t, SurfPt *sp) {
    if (t.x == sp->frame.x.x && t.y == sp->frame.x.y && t.z == sp->frame.x.z) return;
    const float dotNT = dot(sp->frame.z, t);
    const V3 proj = t - sp->frame.z * dotNT;
    const float lx = dot(sp->frame.x, proj), ly = dot(sp->frame.y, proj);
    const float angle = atan2f(ly, lx);
    float s, c;
    sincosf(angle, &s, &c);
    const Frame f = sp->frame;
    sp->frame.x = normalize(f.fromLocal(V3(c, s, 0)));
    sp->frame.y = normalize(f.fromLocal(V3(-s, c, 0)));
}

__device__ __forceinline__ void applyGroupNodes(const DevScene &sc, const DevGeomInst &G, SurfPt *sp) {
    if (G.progCount == 0) return;
    NodeOutputs<3> o;
    Wls<3> none;
    runNodeProgram<3, false>(sc, G.progBase, G.progCount, *sp, none, VC_N3, VC_V3, &o);
    // calcSurfacePoint: the normal node's default (0, 0, 1) leaves the frame as it is
    if (o.has & 8u) applyBumpMapping(V3(o.vec[0].x, o.vec[0].y, o.vec[0].z), sp);
    if (o.has & 16u) modifyTangent(V3(o.vec[1].x, o.vec[1].y, o.vec[1].z), sp);
}

// A material's node outputs at a surface point (sp may be null: immediates only).
template <int N>
__device__ __forceinline__ void materialNodes(const DevScene &sc, const DevMaterial &m, const SurfPt *sp,
                                              const Wls<N> &wls, NodeOutputs<N> *o) {
    o->has = 0;
    if (m.progCount && sp) runNodeProgram<N, true>(sc, m.progBase, m.progCount, *sp, wls, VC_F3, VC_F3, o);
}

template <int N>
__device__ __forceinline__ Spec<N> materialSpectrum(const DevScene &sc, const DevMaterial &m, int k,
                                                    const NodeOutputs<N> &o, const Wls<N> &wls) {
    return (o.has & (1u << k)) ? o.spec[k] : evalSpectrum<N>(sc, m.spec[k], wls);
}

// GGX alphas and rotation from roughness / anisotropy / rotation (materials.cu:1230-1236)
__device__ __forceinline__ void ggxParams(const float4 &rar, float *ax, float *ay, float *rot) {
    const float alpha = rar.x * rar.x;
    const float aspect = sqrtf(1.0f - 0.9f * rar.y);
    *ax = fmaxf(0.001f, alpha / aspect);
    *ay = fmaxf(0.001f, alpha * aspect);
    *rot = 2 * kPiF * rar.z;
}

// Material set-up at a surface point (the *_setupBSDF callables of materials.cu)
template <int N, bool NODES = true>
__device__ __forceinline__ Bsdf<N> setupBsdf(const DevScene &sc, const DevMaterial &m, const Wls<N> &wls,
                                             const SurfPt *sp) {
    NodeOutputs<N> o;
    if constexpr (NODES) materialNodes<N>(sc, m, sp, wls, &o);
    else o.has = 0;
    Bsdf<N> b;
    b.kind = BSDF_NULL;
    b.a0 = b.a1 = b.a2 = 0.0f;
    b.dispersive = false;
    switch (m.type) {
    case MAT_MATTE:
        b.kind = BSDF_MATTE;
        b.c0 = materialSpectrum<N>(sc, m, 0, o, wls);
        break;
    case MAT_SPECULAR_REFLECTION:
        b.kind = BSDF_SPEC_BRDF;
        b.c0 = materialSpectrum<N>(sc, m, 0, o, wls);
        b.c1 = materialSpectrum<N>(sc, m, 1, o, wls);
        b.c2 = materialSpectrum<N>(sc, m, 2, o, wls);
        break;
    case MAT_SPECULAR_SCATTERING:
        b.kind = BSDF_SPEC_BSDF;
        b.c0 = materialSpectrum<N>(sc, m, 0, o, wls);
        b.c1 = materialSpectrum<N>(sc, m, 1, o, wls);
        b.c2 = materialSpectrum<N>(sc, m, 2, o, wls);
        b.dispersive = !wls.single;
        break;
    case MAT_MICROFACET_REFLECTION:
        // materials.cu:1226-1236; alphas derived on the host unless a node supplies the parameters
        b.kind = BSDF_MICRO_BRDF;
        b.c0 = materialSpectrum<N>(sc, m, 0, o, wls);
        b.c1 = materialSpectrum<N>(sc, m, 1, o, wls);
        if (o.has & 8u) {
            ggxParams(o.vec[0], &b.a0, &b.a1, &b.a2);
        } else {
            b.a0 = m.scalars[0];
            b.a1 = m.scalars[1];
            b.a2 = m.scalars[2];
        }
        break;
    case MAT_MICROFACET_SCATTERING:
        // materials.cu:1516-1536
        b.kind = BSDF_MICRO_BSDF;
        b.c0 = materialSpectrum<N>(sc, m, 0, o, wls);
        b.c1 = materialSpectrum<N>(sc, m, 1, o, wls);
        b.c2 = materialSpectrum<N>(sc, m, 2, o, wls);
        if (o.has & 8u) {
            ggxParams(o.vec[0], &b.a0, &b.a1, &b.a2);
        } else {
            b.a0 = m.scalars[0];
            b.a1 = m.scalars[1];
            b.a2 = m.scalars[2];
        }
        break;
    case MAT_UE4: {
        // materials.cu:1959-1978
        b.kind = BSDF_DIFFSPEC;
        Spec<N> base = materialSpectrum<N>(sc, m, 0, o, wls);
        float roughness = m.scalars[1], metallic = m.scalars[2];
        if (o.has & 8u) {
            roughness = o.vec[0].y;
            metallic = o.vec[0].z;
        }
        roughness = fmaxf(0.01f, roughness);
        const float specular = 0.5f;
        b.c0 = base * (1 - metallic);
        b.c1 = lerpSpec<N>(Spec<N>(0.08f * specular), base, metallic);
        b.a0 = roughness;
        break;
    }
    default:
        break;
    }
    return b;
}

} // namespace vlra
