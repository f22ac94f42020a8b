#pragma once
// Public value types of the wr:: API (libWR/include/wr/types.h:1-51).
//
// libWR aliases Vec3 to glm::vec3. This library has no GLM dependency:
// Vec3/Vec4 are 12/16-byte POD structs with the same layout and the few
// helpers libWR's callers use (construction, +, -, *, normalize, cross,
// radians). Camera keeps libWR's fields, defaults and basis helpers.
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
};

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

    Vec3 getForward() const { return normalize(target - position); }
    Vec3 getRight() const { return normalize(cross(getForward(), up)); }
    Vec3 getUp() const { return normalize(cross(getRight(), getForward())); }
};

}  // namespace wr
