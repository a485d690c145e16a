// camera.h -- Camera with the reference's interface (Core/camera.h:9-100,
// Core/Camera.cpp:1-38), on the API's own vec3 instead of glm.  UVWFrame
// produces exactly the frame vzCameraLookAt does for the same inputs.
#pragma once
#include <cmath>

#include "vec3.h"

class Camera {
public:
    Camera() { UpdateUVW(); }
    Camera(const vec3& origin, float vfov, float aspect, float aperture, float focalLength)
        : m_position(origin), m_vfov(vfov), m_aspect(aspect), m_aperture(aperture), m_focalLength(focalLength)
    {
        UpdateUVW();
    }

    // U, V, W of the thin-lens camera (Camera.cpp:16-37): W = forward*focal,
    // U = normalize(W x up)*focal*aspect*2h, V = normalize(U x W)*focal*2h,
    // h = tan(vfov/2).
    void UVWFrame(vec3& U, vec3& V, vec3& W)
    {
        UpdateUVW();
        U = m_u;
        V = m_v;
        W = m_w;
    }

    float GetLensRadius() const { return m_aperture * 0.5f; }
    const vec3& GetPosition() const { return m_position; }
    void SetPosition(const vec3& p)
    {
        m_position = p;
        m_changed = true;
    }
    void SetForward(const vec3& direction)
    {
        m_forward = normalize(direction);
        m_changed = true;
    }
    const vec3& GetForward() const { return m_forward; }
    float GetAspect() const { return m_aspect; }
    void SetAspect(float a)
    {
        if (m_aspect != a) {
            m_aspect = a;
            m_changed = true;
        }
    }
    float GetFocalLength() const { return m_focalLength; }
    void SetFocalLength(float length)
    {
        if (m_focalLength != length) {
            m_focalLength = length;
            m_changed = true;
        }
    }

    // Interactive moves (camera.h:36-78).  The reference rotates with
    // glm's v * q, i.e. by -angle about the (normalised) axis.  As there,
    // the axes and the move directions are the frame of the last
    // UpdateUVW (construction or UVWFrame): several moves between two
    // frames all use that frame.
    void MoveForward(float speed)
    {
        m_position += m_forward * speed;
        m_changed = true;
    }
    void MoveRight(float speed)
    {
        m_position += normalize(m_u) * speed;
        m_changed = true;
    }
    void MoveUp(float speed)
    {
        m_position += normalize(m_v) * speed;
        m_changed = true;
    }
    void Pitch(float speed) { Turn(m_u, speed); }
    void Yaw(float speed) { Turn(m_v, speed); }
    void Roll(float speed) { Turn(m_w, speed); }

    // True once after any change (camera.h:80-85); the renderer restarts
    // its progressive accumulation when it sees it.
    bool Changed()
    {
        bool c = m_changed;
        m_changed = false;
        return c;
    }

private:
    void Turn(const vec3& axis, float degrees)
    {
        m_forward = normalize(rotate(m_forward, normalize(axis), -degrees));
        m_changed = true;
    }

    void UpdateUVW()
    {
        const vec3 up(0.0f, 1.0f, 0.0f);
        m_w = m_forward * m_focalLength;
        m_u = normalize(cross(m_w, up));
        m_v = normalize(cross(m_u, m_w));
        float theta = m_vfov * 0.01745329251994329577f;
        float h = std::tan(theta * 0.5f);
        float vh = 2.0f * h;
        float vw = m_aspect * vh;
        m_u *= m_focalLength * vw;
        m_v *= m_focalLength * vh;
    }

    vec3 m_position = vec3(0.0f, 0.0f, 0.0f);
    vec3 m_forward = vec3(0.0f, 0.0f, -1.0f);
    float m_vfov = 45.0f;
    float m_aspect = 1.6f;
    float m_aperture = 0.0f;
    float m_focalLength = 1.0f;
    vec3 m_u, m_v, m_w;
    bool m_changed = true;
};
