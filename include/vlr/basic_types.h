/* Geometry types of the C++ wrapper (vlrcpp.h): points, vectors, normals, texture coordinates,
 * vertices, quaternions and 4x4 matrices with the names and member layouts the reference's
 * wrapper uses (libVLR/include/vlr/basic_types.h), so that callers of vlr::Context compile
 * unchanged.  Each type has the layout of the matching C API struct (public_types.h), which the
 * wrapper passes through as is.  Only what a scene description needs is here: construction,
 * arithmetic, normalisation, and affine matrices (column-major storage, as vlrStaticTransformCreate
 * reads them). */
#pragma once

#include <cmath>
#include <cstdint>

#include "public_types.h"

namespace vlr {

struct Vector3D {
    float x, y, z;
    constexpr Vector3D(float v = 0.0f) : x(v), y(v), z(v) {}
    constexpr Vector3D(float xx, float yy, float zz) : x(xx), y(yy), z(zz) {}
    Vector3D operator+(const Vector3D &o) const { return Vector3D(x + o.x, y + o.y, z + o.z); }
    Vector3D operator-(const Vector3D &o) const { return Vector3D(x - o.x, y - o.y, z - o.z); }
    Vector3D operator-() const { return Vector3D(-x, -y, -z); }
    Vector3D operator*(float s) const { return Vector3D(x * s, y * s, z * s); }
    Vector3D operator/(float s) const { return Vector3D(x / s, y / s, z / s); }
    float length() const { return std::sqrt(x * x + y * y + z * z); }
    Vector3D &normalize() {
        const float l = length();
        x /= l;
        y /= l;
        z /= l;
        return *this;
    }
};

struct Normal3D {
    float x, y, z;
    constexpr Normal3D(float v = 0.0f) : x(v), y(v), z(v) {}
    constexpr Normal3D(float xx, float yy, float zz) : x(xx), y(yy), z(zz) {}
    explicit Normal3D(const Vector3D &v) : x(v.x), y(v.y), z(v.z) {}
    float length() const { return std::sqrt(x * x + y * y + z * z); }
    Normal3D &normalize() {
        const float l = length();
        x /= l;
        y /= l;
        z /= l;
        return *this;
    }
};

struct Point3D {
    float x, y, z;
    constexpr Point3D(float v = 0.0f) : x(v), y(v), z(v) {}
    constexpr Point3D(float xx, float yy, float zz) : x(xx), y(yy), z(zz) {}
    Point3D operator+(const Vector3D &v) const { return Point3D(x + v.x, y + v.y, z + v.z); }
    Vector3D operator-(const Point3D &o) const { return Vector3D(x - o.x, y - o.y, z - o.z); }
};

struct TexCoord2D {
    float u, v;
    constexpr TexCoord2D(float w = 0.0f) : u(w), v(w) {}
    constexpr TexCoord2D(float uu, float vv) : u(uu), v(vv) {}
};

struct Vertex {
    Point3D position;
    Normal3D normal;
    Vector3D tc0Direction;
    TexCoord2D texCoord;
};

// unit quaternion (x, y, z: vector part; w: scalar part)
struct Quaternion {
    float x, y, z, w;
    constexpr Quaternion() : x(0.0f), y(0.0f), z(0.0f), w(1.0f) {}
    constexpr Quaternion(float xx, float yy, float zz, float ww) : x(xx), y(yy), z(zz), w(ww) {}
    Quaternion operator*(const Quaternion &q) const {
        return Quaternion(w * q.x + x * q.w + y * q.z - z * q.y, w * q.y + y * q.w + z * q.x - x * q.z,
                          w * q.z + z * q.w + x * q.y - y * q.x, w * q.w - x * q.x - y * q.y - z * q.z);
    }
};

inline Quaternion qRotate(float angle, const Vector3D &axis) {
    Vector3D a = axis;
    a.normalize();
    const float s = std::sin(angle / 2);
    return Quaternion(a.x * s, a.y * s, a.z * s, std::cos(angle / 2));
}
inline Quaternion qRotateX(float angle) { return qRotate(angle, Vector3D(1, 0, 0)); }
inline Quaternion qRotateY(float angle) { return qRotate(angle, Vector3D(0, 1, 0)); }
inline Quaternion qRotateZ(float angle) { return qRotate(angle, Vector3D(0, 0, 1)); }

// 4x4 matrix, column-major: m[4 * column + row]
struct Matrix4x4 {
    float m[16];
    Matrix4x4() {
        for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    }
    explicit Matrix4x4(const float colMajor[16]) {
        for (int i = 0; i < 16; ++i) m[i] = colMajor[i];
    }
    float &at(int row, int col) { return m[4 * col + row]; }
    float at(int row, int col) const { return m[4 * col + row]; }
    Matrix4x4 operator*(const Matrix4x4 &b) const {
        Matrix4x4 r;
        for (int row = 0; row < 4; ++row)
            for (int col = 0; col < 4; ++col) {
                float s = 0.0f;
                for (int k = 0; k < 4; ++k) s += at(row, k) * b.at(k, col);
                r.at(row, col) = s;
            }
        return r;
    }
    void getArray(float colMajor[16]) const {
        for (int i = 0; i < 16; ++i) colMajor[i] = m[i];
    }
};

inline Matrix4x4 translate(float tx, float ty, float tz) {
    Matrix4x4 r;
    r.at(0, 3) = tx;
    r.at(1, 3) = ty;
    r.at(2, 3) = tz;
    return r;
}
inline Matrix4x4 translate(const Vector3D &t) { return translate(t.x, t.y, t.z); }
inline Matrix4x4 scale(float sx, float sy, float sz) {
    Matrix4x4 r;
    r.at(0, 0) = sx;
    r.at(1, 1) = sy;
    r.at(2, 2) = sz;
    return r;
}
inline Matrix4x4 scale(float s) { return scale(s, s, s); }
// rotation by `angle` radians about a unit axis (right-handed)
inline Matrix4x4 rotate(float angle, const Vector3D &axis) {
    Vector3D a = axis;
    a.normalize();
    const float c = std::cos(angle), s = std::sin(angle), t = 1 - c;
    Matrix4x4 r;
    r.at(0, 0) = t * a.x * a.x + c;
    r.at(0, 1) = t * a.x * a.y - s * a.z;
    r.at(0, 2) = t * a.x * a.z + s * a.y;
    r.at(1, 0) = t * a.x * a.y + s * a.z;
    r.at(1, 1) = t * a.y * a.y + c;
    r.at(1, 2) = t * a.y * a.z - s * a.x;
    r.at(2, 0) = t * a.x * a.z - s * a.y;
    r.at(2, 1) = t * a.y * a.z + s * a.x;
    r.at(2, 2) = t * a.z * a.z + c;
    return r;
}
inline Matrix4x4 rotateX(float angle) { return rotate(angle, Vector3D(1, 0, 0)); }
inline Matrix4x4 rotateY(float angle) { return rotate(angle, Vector3D(0, 1, 0)); }
inline Matrix4x4 rotateZ(float angle) { return rotate(angle, Vector3D(0, 0, 1)); }

static_assert(sizeof(Point3D) == sizeof(VLRPoint3D), "Point3D layout");
static_assert(sizeof(Vector3D) == sizeof(VLRVector3D), "Vector3D layout");
static_assert(sizeof(Normal3D) == sizeof(VLRNormal3D), "Normal3D layout");
static_assert(sizeof(Quaternion) == sizeof(VLRQuaternion), "Quaternion layout");
static_assert(sizeof(Vertex) == sizeof(VLRVertex), "Vertex layout");

} // namespace vlr
