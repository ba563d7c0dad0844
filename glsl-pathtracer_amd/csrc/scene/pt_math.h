// Host-side math types for the scene builder.
//
// The arithmetic here feeds the BVH builder, whose output must be bit-identical
// to the reference's flattened arrays, so every operation keeps the reference's
// float evaluation order:
//   - Vec3 component ops and Min/Max follow src/math/Vec3.h:95-160
//     (Min/Max use std::min/std::max argument order, which fixes NaN handling).
//   - Mat4 stores data[i] as one GLSL column (translation in data[3]),
//     src/math/Mat4.h:31-45; operator* and QuatToMatrix follow :80-140.
//   - BBox follows thirdparty/RadeonRays/bbox.h:40-103 and bbox.cpp:27-93.
// The library is compiled with -ffp-contract=off so that no FMA contraction
// changes a rounding.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>

namespace glslpt {

constexpr float kPi = 3.14159265358979323846f;  // PI of src/math/MathUtils.h:27

struct Vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
    Vec3() = default;
    Vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}

    Vec3 operator+(const Vec3& b) const { return Vec3(x + b.x, y + b.y, z + b.z); }
    Vec3 operator-(const Vec3& b) const { return Vec3(x - b.x, y - b.y, z - b.z); }
    Vec3 operator*(const Vec3& b) const { return Vec3(x * b.x, y * b.y, z * b.z); }
    Vec3 operator*(float b) const { return Vec3(x * b, y * b, z * b); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }

    static Vec3 Min(const Vec3& a, const Vec3& b) {
        return Vec3(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z));
    }
    static Vec3 Max(const Vec3& a, const Vec3& b) {
        return Vec3(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z));
    }
    static Vec3 Cross(const Vec3& a, const Vec3& b) {
        return Vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
    }
    static float Dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    static float Length(const Vec3& a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
    static float Distance(const Vec3& a, const Vec3& b) { return Length(a - b); }
    static Vec3 Normalize(const Vec3& a) {
        float l = Length(a);
        return Vec3(a.x / l, a.y / l, a.z / l);
    }
};

struct Vec4 {
    float x = 0.f, y = 0.f, z = 0.f, w = 0.f;
    Vec4() = default;
    Vec4(float x_, float y_, float z_, float w_) : x(x_), y(y_), z(z_), w(w_) {}
};

// 4x4 matrix, data[c] is column c (GLSL convention), translation in data[3].
struct Mat4 {
    float data[4][4];
    Mat4() {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) data[i][j] = (i == j) ? 1.f : 0.f;
    }
    float (&operator[](int i))[4] { return data[i]; }
    const float (&operator[](int i) const)[4] { return data[i]; }

    Mat4 operator*(const Mat4& b) const {
        Mat4 out;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                out.data[i][j] = data[i][0] * b.data[0][j] + data[i][1] * b.data[1][j] +
                                 data[i][2] * b.data[2][j] + data[i][3] * b.data[3][j];
        return out;
    }
    static Mat4 Translate(const Vec3& a) {
        Mat4 m;
        m.data[3][0] = a.x; m.data[3][1] = a.y; m.data[3][2] = a.z;
        return m;
    }
    static Mat4 Scale(const Vec3& a) {
        Mat4 m;
        m.data[0][0] = a.x; m.data[1][1] = a.y; m.data[2][2] = a.z;
        return m;
    }
    static Mat4 QuatToMatrix(float x, float y, float z, float w) {
        Mat4 o;
        const float x2 = x + x, y2 = y + y, z2 = z + z;
        const float xx = x * x2, xy = x * y2, xz = x * z2;
        const float yy = y * y2, yz = y * z2, zz = z * z2;
        const float wx = w * x2, wy = w * y2, wz = w * z2;
        o.data[0][0] = 1.0f - (yy + zz); o.data[0][1] = xy + wz; o.data[0][2] = xz - wy; o.data[0][3] = 0.f;
        o.data[1][0] = xy - wz; o.data[1][1] = 1.0f - (xx + zz); o.data[1][2] = yz + wx; o.data[1][3] = 0.f;
        o.data[2][0] = xz + wy; o.data[2][1] = yz - wx; o.data[2][2] = 1.0f - (xx + yy); o.data[2][3] = 0.f;
        o.data[3][0] = 0.f; o.data[3][1] = 0.f; o.data[3][2] = 0.f; o.data[3][3] = 1.0f;
        return o;
    }
};

// Axis-aligned box; an empty box is (+max, -max) so that growing by anything
// yields that thing.
struct BBox {
    Vec3 pmin{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
              std::numeric_limits<float>::max()};
    Vec3 pmax{-std::numeric_limits<float>::max(), -std::numeric_limits<float>::max(),
              -std::numeric_limits<float>::max()};

    Vec3 center() const { return (pmax + pmin) * 0.5f; }
    Vec3 extents() const { return pmax - pmin; }
    float surface_area() const {
        Vec3 e = extents();
        return 2.f * (e.x * e.y + e.x * e.z + e.y * e.z);
    }
    int maxdim() const {
        Vec3 e = extents();
        if (e.x >= e.y && e.x >= e.z) return 0;
        if (e.y >= e.x && e.y >= e.z) return 1;
        if (e.z >= e.x && e.z >= e.y) return 2;
        return 0;
    }
    void grow(const Vec3& p) { pmin = Vec3::Min(pmin, p); pmax = Vec3::Max(pmax, p); }
    void grow(const BBox& b) { pmin = Vec3::Min(pmin, b.pmin); pmax = Vec3::Max(pmax, b.pmax); }
};

}  // namespace glslpt
