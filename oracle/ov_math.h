// Oracle (test infrastructure only): CPU restatement of the reference's vector math.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only
// as the checker or as the timed CPU baseline; the product never links it.
//
// Follows libVLR/include/vlr/basic_types.h (Vector3D/Point3D/Normal3D/Quaternion/Matrix)
// and libVLR/include/vlr/common.h:144-173 (sincos).  Division by a scalar is a reciprocal
// multiply as in basic_types.h:26,32.  Compiled with -ffp-contract=off.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>

namespace ov {

constexpr float kPi = 3.14159265358979323846f;

struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float v) : x(v), y(v), z(v) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    V3 operator-() const { return V3(-x, -y, -z); }
    V3 operator+(const V3 &o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3 &o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(const V3 &o) const { return V3(x * o.x, y * o.y, z * o.z); }
    V3 operator*(float s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(float s) const { float r = 1 / s; return V3(x * r, y * r, z * r); }
    V3 &operator+=(const V3 &o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3 &operator*=(float s) { x *= s; y *= s; z *= s; return *this; }
    V3 &operator/=(float s) { float r = 1 / s; x *= r; y *= r; z *= r; return *this; }
    float operator[](int i) const { return (&x)[i]; }
    float &operator[](int i) { return (&x)[i]; }
    float sqLength() const { return x * x + y * y + z * z; }
    float length() const { return std::sqrt(x * x + y * y + z * z); }
    bool allFinite() const { return std::isfinite(x) && std::isfinite(y) && std::isfinite(z); }
    bool operator==(const V3 &o) const { return x == o.x && y == o.y && z == o.z; }
    // basic_types.h:58-64 (Duff et al., "Building an Orthonormal Basis, Revisited")
    void makeCoordinateSystem(V3 *vx, V3 *vy) const {
        float sign = z >= 0 ? 1.0f : -1.0f;
        const float a = -1 / (sign + z);
        const float b = x * y * a;
        *vx = V3(1 + sign * x * x * a, sign * b, -sign * x);
        *vy = V3(b, sign + y * y * a, -y);
    }
    // basic_types.h:72-78
    static V3 fromPolarYUp(float phi, float theta) {
        float sp = std::sin(phi), cp = std::cos(phi);
        float st = std::sin(theta), ct = std::cos(theta);
        return V3(-sp * st, ct, cp * st);
    }
    // basic_types.h:65-71
    static V3 fromPolarZUp(float phi, float theta) {
        float sp = std::sin(phi), cp = std::cos(phi);
        float st = std::sin(theta), ct = std::cos(theta);
        return V3(cp * st, sp * st, ct);
    }
    // basic_types.h:84-88
    void toPolarYUp(float *theta, float *phi) const {
        *theta = std::acos(std::min(std::max(y, -1.0f), 1.0f));
        *phi = std::fmod(static_cast<float>(std::atan2(-x, z) + 2 * kPi), static_cast<float>(2 * kPi));
    }
};
inline V3 operator*(float s, const V3 &v) { return V3(s * v.x, s * v.y, s * v.z); }
inline float dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float absDot(const V3 &a, const V3 &b) { return std::fabs(dot(a, b)); }
inline V3 cross(const V3 &a, const V3 &b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline V3 normalize(const V3 &v) { return v / v.length(); }
inline V3 halfVector(const V3 &a, const V3 &b) { return normalize(a + b); }
inline float sqDistance(const V3 &a, const V3 &b) { return (a - b).sqLength(); }
inline float distance(const V3 &a, const V3 &b) { return (a - b).length(); }

template <typename T> inline T pow2(T x) { return x * x; }
template <typename T> inline T pow3(T x) { return x * x * x; }
template <typename T> inline T pow5(T x) { return x * x * x * x * x; }
template <typename T> inline T lerp(T a, T b, float t) { return (1 - t) * a + t * b; }
template <typename T> inline T clampv(T v, T lo, T hi) { return std::min(std::max(v, lo), hi); }

// Column-major 3x3 (basic_types.h Matrix3x3Template: columns c0, c1, c2).
struct M3 {
    V3 c0, c1, c2;
    M3() {}
    M3(const V3 &a, const V3 &b, const V3 &c) : c0(a), c1(b), c2(c) {}
    V3 operator*(const V3 &v) const {
        return V3(c0.x * v.x + c1.x * v.y + c2.x * v.z,
                  c0.y * v.x + c1.y * v.y + c2.y * v.z,
                  c0.z * v.x + c1.z * v.y + c2.z * v.z);
    }
};

struct Quat {
    float x, y, z, w;
    // basic_types.h:1386-1393
    M3 toMatrix3x3() const {
        float xx = x * x, yy = y * y, zz = z * z;
        float xy = x * y, yz = y * z, zx = z * x;
        float xw = x * w, yw = y * w, zw = z * w;
        return M3(V3(1 - 2 * (yy + zz), 2 * (xy + zw), 2 * (zx - yw)),
                  V3(2 * (xy - zw), 1 - 2 * (xx + zz), 2 * (yz + xw)),
                  V3(2 * (zx + yw), 2 * (yz - xw), 1 - 2 * (xx + yy)));
    }
};

// Affine transform kept as 3x4 row-major rows for the forward and inverse maps.
// The inverse is computed in double from the float forward matrix and rounded once;
// the product's host code follows the same definition, so object-space rays agree bitwise.
struct Xfm {
    float m[12];    // rows: [r0 | t0], [r1 | t1], [r2 | t2]
    float inv[12];
    static Xfm fromColumnMajor4x4(const float a[16]) {
        Xfm t;
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 4; ++c)
                t.m[r * 4 + c] = a[c * 4 + r];
        }
        double M[3][3], T[3];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) M[r][c] = t.m[r * 4 + c];
            T[r] = t.m[r * 4 + 3];
        }
        double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1])
                   - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
                   + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
        double id = 1.0 / det;
        double I[3][3];
        I[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) * id;
        I[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * id;
        I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * id;
        I[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) * id;
        I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * id;
        I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * id;
        I[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) * id;
        I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * id;
        I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * id;
        for (int r = 0; r < 3; ++r) {
            double tr = -(I[r][0] * T[0] + I[r][1] * T[1] + I[r][2] * T[2]);
            for (int c = 0; c < 3; ++c) t.inv[r * 4 + c] = static_cast<float>(I[r][c]);
            t.inv[r * 4 + 3] = static_cast<float>(tr);
        }
        return t;
    }
    static V3 mulP(const float *a, const V3 &p) {
        return V3(a[0] * p.x + a[1] * p.y + a[2] * p.z + a[3],
                  a[4] * p.x + a[5] * p.y + a[6] * p.z + a[7],
                  a[8] * p.x + a[9] * p.y + a[10] * p.z + a[11]);
    }
    static V3 mulV(const float *a, const V3 &v) {
        return V3(a[0] * v.x + a[1] * v.y + a[2] * v.z,
                  a[4] * v.x + a[5] * v.y + a[6] * v.z,
                  a[8] * v.x + a[9] * v.y + a[10] * v.z);
    }
    V3 point(const V3 &p) const { return mulP(m, p); }
    V3 vector(const V3 &v) const { return mulV(m, v); }
    // Normal transform by the inverse transpose (shared.h:317-322).
    V3 normal(const V3 &n) const {
        return V3(inv[0] * n.x + inv[4] * n.y + inv[8] * n.z,
                  inv[1] * n.x + inv[5] * n.y + inv[9] * n.z,
                  inv[2] * n.x + inv[6] * n.y + inv[10] * n.z);
    }
    V3 invPoint(const V3 &p) const { return mulP(inv, p); }
    V3 invVector(const V3 &v) const { return mulV(inv, v); }
};

// renderer_common.h:502-528 (Wächter & Binder, Ray Tracing Gems ch. 6)
inline float asFloat(int32_t i) { float f; std::memcpy(&f, &i, 4); return f; }
inline int32_t asInt(float f) { int32_t i; std::memcpy(&i, &f, 4); return i; }
inline V3 offsetRayOrigin(const V3 &p, const V3 &gn) {
    constexpr float kOrigin = 1.0f / 32.0f;
    constexpr float kFloatScale = 1.0f / 65536.0f;
    constexpr float kIntScale = 256.0f;
    int32_t oi[3] = {static_cast<int32_t>(kIntScale * gn.x), static_cast<int32_t>(kIntScale * gn.y),
                     static_cast<int32_t>(kIntScale * gn.z)};
    V3 p1(asFloat(asInt(p.x) + (p.x < 0 ? -1 : 1) * oi[0]),
          asFloat(asInt(p.y) + (p.y < 0 ? -1 : 1) * oi[1]),
          asFloat(asInt(p.z) + (p.z < 0 ? -1 : 1) * oi[2]));
    V3 p2 = p + kFloatScale * gn;
    return V3(std::fabs(p.x) < kOrigin ? p2.x : p1.x,
              std::fabs(p.y) < kOrigin ? p2.y : p1.y,
              std::fabs(p.z) < kOrigin ? p2.z : p1.z);
}

// kernel_common.h:147-175 ReferenceFrame
struct Frame {
    V3 x, y, z;
    Frame() {}
    Frame(const V3 &t, const V3 &n) : x(t), y(cross(n, t)), z(n) {}
    Frame(const V3 &t, const V3 &b, const V3 &n) : x(t), y(b), z(n) {}
    V3 toLocal(const V3 &v) const { return V3(dot(x, v), dot(y, v), dot(z, v)); }
    V3 fromLocal(const V3 &v) const {
        return V3(dot(V3(x.x, y.x, z.x), v), dot(V3(x.y, y.y, z.y), v), dot(V3(x.z, y.z, z.z), v));
    }
};

} // namespace ov
