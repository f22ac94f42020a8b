#pragma once
// Public value types of the wr:: API (libWR/include/wr/types.h:1-51).
//
// libWR aliases Vec3/Vec4/Mat3/Mat4 to glm::vec3/vec4/mat3/mat4. This
// library has no GLM dependency: they are POD structs with GLM's layout
// (12/16 bytes; matrices column-major, m[column][row], 36/64 bytes) and the
// few helpers libWR's callers use (construction, +, -, *, normalize, cross,
// radians, matrix-vector and matrix-matrix products). Camera keeps libWR's
// fields, defaults and helpers, including the view and projection matrices.
#include <cmath>
#include <cstdint>

namespace wr {

struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    constexpr Vec3() = default;
    constexpr explicit Vec3(float s) : x(s), y(s), z(s) {}
    constexpr Vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
static_assert(sizeof(Vec3) == 12, "Vec3 must match glm::vec3 layout");

struct Vec4 {
    float x = 0.0f, y = 0.0f, z = 0.0f, w = 0.0f;
    constexpr Vec4() = default;
    constexpr Vec4(float x_, float y_, float z_, float w_) : x(x_), y(y_), z(z_), w(w_) {}
    constexpr Vec4(const Vec3& v, float w_) : x(v.x), y(v.y), z(v.z), w(w_) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
};
static_assert(sizeof(Vec4) == 16, "Vec4 must match glm::vec4 layout");

// 3x3 matrix, three columns (glm::mat3 layout). Mat3(d) puts d on the
// diagonal; the default is the identity.
struct Mat3 {
    Vec3 col[3];
    constexpr Mat3() : Mat3(1.0f) {}
    constexpr explicit Mat3(float d) : col{Vec3(d, 0.0f, 0.0f), Vec3(0.0f, d, 0.0f), Vec3(0.0f, 0.0f, d)} {}
    constexpr Mat3(const Vec3& c0, const Vec3& c1, const Vec3& c2) : col{c0, c1, c2} {}
    Vec3& operator[](int c) { return col[c]; }
    const Vec3& operator[](int c) const { return col[c]; }
};
static_assert(sizeof(Mat3) == 36, "Mat3 must match glm::mat3 layout");

// 4x4 matrix, four columns (glm::mat4 layout): m[c][r] is row r of column c.
struct Mat4 {
    Vec4 col[4];
    constexpr Mat4() : Mat4(1.0f) {}
    constexpr explicit Mat4(float d)
        : col{Vec4(d, 0.0f, 0.0f, 0.0f), Vec4(0.0f, d, 0.0f, 0.0f), Vec4(0.0f, 0.0f, d, 0.0f),
              Vec4(0.0f, 0.0f, 0.0f, d)} {}
    Vec4& operator[](int c) { return col[c]; }
    const Vec4& operator[](int c) const { return col[c]; }
};
static_assert(sizeof(Mat4) == 64, "Mat4 must match glm::mat4 layout");

inline Vec3 operator*(const Mat3& m, const Vec3& v) {
    return {m[0].x * v.x + m[1].x * v.y + m[2].x * v.z, m[0].y * v.x + m[1].y * v.y + m[2].y * v.z,
            m[0].z * v.x + m[1].z * v.y + m[2].z * v.z};
}
inline Vec4 operator*(const Mat4& m, const Vec4& v) {
    Vec4 r;
    for (int i = 0; i < 4; ++i) r[i] = m[0][i] * v.x + m[1][i] * v.y + m[2][i] * v.z + m[3][i] * v.w;
    return r;
}
inline Mat4 operator*(const Mat4& a, const Mat4& b) {
    Mat4 r(0.0f);
    for (int c = 0; c < 4; ++c) r[c] = a * b[c];
    return r;
}

inline Vec3 operator+(const Vec3& a, const Vec3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(const Vec3& a, const Vec3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(const Vec3& a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator*(float s, const Vec3& a) { return {a.x * s, a.y * s, a.z * s}; }
// glm::dot for vec3: (x*x + y*y) + z*z
inline float dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// glm::cross
inline Vec3 cross(const Vec3& a, const Vec3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// glm::normalize: v * inversesqrt(dot(v, v)), inversesqrt = 1/sqrt
inline Vec3 normalize(const Vec3& v) {
    float inv = 1.0f / std::sqrt(dot(v, v));
    return v * inv;
}
constexpr float radians(float degrees) { return degrees * 0.01745329251994329576923690768489f; }

// Right-handed view matrix (glm::lookAt's default convention): the camera
// looks down -z with +x to the right (s = normalize(f x up)) and +y up.
inline Mat4 lookAt(const Vec3& eye, const Vec3& center, const Vec3& up) {
    const Vec3 f = normalize(center - eye);
    const Vec3 s = normalize(cross(f, up));
    const Vec3 u = cross(s, f);
    Mat4 m(1.0f);
    m[0][0] = s.x; m[1][0] = s.y; m[2][0] = s.z;
    m[0][1] = u.x; m[1][1] = u.y; m[2][1] = u.z;
    m[0][2] = -f.x; m[1][2] = -f.y; m[2][2] = -f.z;
    m[3][0] = -dot(s, eye);
    m[3][1] = -dot(u, eye);
    m[3][2] = dot(f, eye);
    return m;
}

// Right-handed perspective projection with depth mapped to [-1, 1]
// (glm::perspective's default convention, the gluPerspective matrix).
inline Mat4 perspective(float fovY, float aspect, float zNear, float zFar) {
    const float t = std::tan(fovY / 2.0f);
    Mat4 m(0.0f);
    m[0][0] = 1.0f / (aspect * t);
    m[1][1] = 1.0f / t;
    m[2][2] = -(zFar + zNear) / (zFar - zNear);
    m[2][3] = -1.0f;
    m[3][2] = -(2.0f * zFar * zNear) / (zFar - zNear);
    return m;
}

// Camera configuration (libWR/include/wr/types.h:16-51)
struct Camera {
    Vec3 position{0.0f, 0.0f, 0.0f};
    Vec3 target{0.0f, 0.0f, -1.0f};
    Vec3 up{0.0f, 1.0f, 0.0f};
    float fovY{radians(45.0f)};
    float aspect{1.0f};

    // Depth of Field (Thin Lens Model)
    float focalDistance{0.0f};  // 0 = infinite focus (no DOF)
    float lensRadius{0.0f};     // 0 = pinhole camera (no DOF)

    Mat4 getViewMatrix() const { return lookAt(position, target, up); }
    Mat4 getProjectionMatrix(float nearPlane = 0.1f, float farPlane = 100.0f) const {
        return perspective(fovY, aspect, nearPlane, farPlane);
    }
    Vec3 getForward() const { return normalize(target - position); }
    Vec3 getRight() const { return normalize(cross(getForward(), up)); }
    Vec3 getUp() const { return normalize(cross(getRight(), getForward())); }
};

}  // namespace wr
