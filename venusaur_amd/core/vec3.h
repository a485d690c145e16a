// vec3.h -- the small host vector type of the C++ API (replaces the glm
// vec3 the reference's camera and scene code use, Core/camera.h:3-4).
// Operations are single float operations in the order written, so the
// camera frame is reproducible bit for bit.
#pragma once
#include <cmath>

struct vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    vec3() = default;
    vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    vec3 operator+(const vec3& b) const { return vec3(x + b.x, y + b.y, z + b.z); }
    vec3 operator-(const vec3& b) const { return vec3(x - b.x, y - b.y, z - b.z); }
    vec3 operator*(float s) const { return vec3(x * s, y * s, z * s); }
    vec3 operator*(const vec3& b) const { return vec3(x * b.x, y * b.y, z * b.z); }
    vec3 operator-() const { return vec3(-x, -y, -z); }
    vec3& operator+=(const vec3& b)
    {
        x += b.x;
        y += b.y;
        z += b.z;
        return *this;
    }
    vec3& operator*=(float s)
    {
        x *= s;
        y *= s;
        z *= s;
        return *this;
    }
};

inline float dot(const vec3& a, const vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(const vec3& a, const vec3& b)
{
    return vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length(const vec3& v) { return std::sqrt(dot(v, v)); }
inline vec3 normalize(const vec3& v)
{
    float inv = 1.0f / std::sqrt(dot(v, v));
    return v * inv;
}
// Rotation of v by `degrees` about the unit axis (Rodrigues' formula).
inline vec3 rotate(const vec3& v, const vec3& axis, float degrees)
{
    float a = degrees * 0.01745329251994329577f;
    float c = std::cos(a), s = std::sin(a);
    return v * c + cross(axis, v) * s + axis * (dot(axis, v) * (1.0f - c));
}
