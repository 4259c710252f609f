// Device math for the light-transport kernels (gfx950).
//
// Restates the reference's vector/frame helpers (libVLR/include/vlr/basic_types.h,
// libVLR/shared/kernel_common.h:147-175, renderer_common.h:502-528) for wave64 device code.
// Scalar division is a reciprocal multiply as in basic_types.h:26,32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vlra {

constexpr float kPiF = 3.14159265358979323846f;

struct V3 {
    float x, y, z;
    __device__ __forceinline__ V3() {}
    __device__ __forceinline__ constexpr V3(float v) : x(v), y(v), z(v) {}
    __device__ __forceinline__ constexpr V3(float a, float b, float c) : x(a), y(b), z(c) {}
    __device__ __forceinline__ V3 operator-() const { return V3(-x, -y, -z); }
    __device__ __forceinline__ V3 operator+(const V3 &o) const { return V3(x + o.x, y + o.y, z + o.z); }
    __device__ __forceinline__ V3 operator-(const V3 &o) const { return V3(x - o.x, y - o.y, z - o.z); }
    __device__ __forceinline__ V3 operator*(float s) const { return V3(x * s, y * s, z * s); }
    __device__ __forceinline__ V3 operator/(float s) const { float r = 1.0f / s; return V3(x * r, y * r, z * r); }
    __device__ __forceinline__ float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    __device__ __forceinline__ float sqLength() const { return x * x + y * y + z * z; }
    __device__ __forceinline__ float length() const { return sqrtf(x * x + y * y + z * z); }
    __device__ __forceinline__ bool allFinite() const { return isfinite(x) && isfinite(y) && isfinite(z); }
};
__device__ __forceinline__ V3 operator*(float s, const V3 &v) { return V3(s * v.x, s * v.y, s * v.z); }
__device__ __forceinline__ float dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float absDot(const V3 &a, const V3 &b) { return fabsf(dot(a, b)); }
__device__ __forceinline__ V3 cross(const V3 &a, const V3 &b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 normalize(const V3 &v) { return v / v.length(); }
__device__ __forceinline__ V3 halfVector(const V3 &a, const V3 &b) { return normalize(a + b); }
__device__ __forceinline__ V3 ld3(const float *p) { return V3(p[0], p[1], p[2]); }
__device__ __forceinline__ V3 xyz(const float4 &v) { return V3(v.x, v.y, v.z); }

template <typename T> __device__ __forceinline__ T sq(T x) { return x * x; }
template <typename T> __device__ __forceinline__ T cube(T x) { return x * x * x; }
template <typename T> __device__ __forceinline__ T pow5f(T x) { return x * x * x * x * x; }
__device__ __forceinline__ float lerpf(float a, float b, float t) { return (1 - t) * a + t * b; }

// 3x4 affine rows [r | t]
__device__ __forceinline__ V3 xfmPoint(const float *a, const V3 &p) {
    return V3(a[0] * p.x + a[1] * p.y + a[2] * p.z + a[3],
              a[4] * p.x + a[5] * p.y + a[6] * p.z + a[7],
              a[8] * p.x + a[9] * p.y + a[10] * p.z + a[11]);
}
__device__ __forceinline__ V3 xfmVector(const float *a, const V3 &v) {
    return V3(a[0] * v.x + a[1] * v.y + a[2] * v.z,
              a[4] * v.x + a[5] * v.y + a[6] * v.z,
              a[8] * v.x + a[9] * v.y + a[10] * v.z);
}
// normal: transpose of the inverse (w2o) rows
__device__ __forceinline__ V3 xfmNormal(const float *inv, const V3 &n) {
    return V3(inv[0] * n.x + inv[4] * n.y + inv[8] * n.z,
              inv[1] * n.x + inv[5] * n.y + inv[9] * n.z,
              inv[2] * n.x + inv[6] * n.y + inv[10] * n.z);
}

struct Frame {
    V3 x, y, z;
    __device__ __forceinline__ Frame() {}
    __device__ __forceinline__ Frame(const V3 &t, const V3 &n) : x(t), y(cross(n, t)), z(n) {}
    __device__ __forceinline__ V3 toLocal(const V3 &v) const { return V3(dot(x, v), dot(y, v), dot(z, v)); }
    __device__ __forceinline__ V3 fromLocal(const V3 &v) const {
        return V3(dot(V3(x.x, y.x, z.x), v), dot(V3(x.y, y.y, z.y), v), dot(V3(x.z, y.z, z.z), v));
    }
};

// Duff et al. orthonormal basis (basic_types.h:58-64)
__device__ __forceinline__ void makeCoordinateSystem(const V3 &n, V3 *vx, V3 *vy) {
    float sign = n.z >= 0 ? 1.0f : -1.0f;
    const float a = -1 / (sign + n.z);
    const float b = n.x * n.y * a;
    *vx = V3(1 + sign * n.x * n.x * a, sign * b, -sign * n.x);
    *vy = V3(b, sign + n.y * n.y * a, -n.y);
}

__device__ __forceinline__ V3 fromPolarYUp(float phi, float theta) {
    float sp, cp, st, ct;
    sincosf(phi, &sp, &cp);
    sincosf(theta, &st, &ct);
    return V3(-sp * st, ct, cp * st);
}
__device__ __forceinline__ V3 fromPolarZUp(float phi, float theta) {
    float sp, cp, st, ct;
    sincosf(phi, &sp, &cp);
    sincosf(theta, &st, &ct);
    return V3(cp * st, sp * st, ct);
}

// Wächter & Binder offset (renderer_common.h:502-528)
__device__ __forceinline__ V3 offsetRayOrigin(const V3 &p, const V3 &gn) {
    constexpr float kOrigin = 1.0f / 32.0f;
    constexpr float kFloatScale = 1.0f / 65536.0f;
    constexpr float kIntScale = 256.0f;
    int32_t ox = static_cast<int32_t>(kIntScale * gn.x);
    int32_t oy = static_cast<int32_t>(kIntScale * gn.y);
    int32_t oz = static_cast<int32_t>(kIntScale * gn.z);
    V3 p1(__int_as_float(__float_as_int(p.x) + (p.x < 0 ? -ox : ox)),
          __int_as_float(__float_as_int(p.y) + (p.y < 0 ? -oy : oy)),
          __int_as_float(__float_as_int(p.z) + (p.z < 0 ? -oz : oz)));
    V3 p2 = p + kFloatScale * gn;
    return V3(fabsf(p.x) < kOrigin ? p2.x : p1.x, fabsf(p.y) < kOrigin ? p2.y : p1.y,
              fabsf(p.z) < kOrigin ? p2.z : p1.z);
}

// PCG32, XSH-RR on the old state, increment 1 (random_distributions.h:6-27)
struct Pcg32 {
    uint64_t s;
    __device__ __forceinline__ uint32_t next() {
        uint64_t old = s;
        s = old * 6364136223846793005ULL + 1ULL;
        uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((-static_cast<int32_t>(rot)) & 31));
    }
    __device__ __forceinline__ float uniform() {
        return __uint_as_float((next() >> 9) | 0x3f800000u) - 1.0f;
    }
};

} // namespace vlra
