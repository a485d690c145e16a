// Scene.h, material.h and sphere.h of the reference (Core/Scene.h:10-112,
// Core/material.h:7-33, Core/sphere.h:5-36) in one header.  The default
// constructor builds the book final scene through vzGenerateBookScene (the
// reference's generator with an explicit x->y->z draw order); Scene::Box
// builds the random-box scene.
#pragma once
#include <cstdint>
#include <vector>

#include "Exception.h"
#include "vec3.h"

class Material {
public:
    enum Type { Lambertian = 0, Metal, Dielectric };

    explicit Material(Type type = Lambertian, const vec3& albedo = vec3(0.0f, 0.0f, 0.0f), float fuzz = 0.0f,
                      float ir = 0.0f)
        : m_type(type), m_albedo(albedo), m_fuzz(fuzz), m_ir(ir)
    {
    }
    Type GetType() const { return m_type; }
    vec3 GetAlbedo() const { return m_albedo; }
    float GetFuzz() const { return m_fuzz; }
    float GetIR() const { return m_ir; }

private:
    Type m_type;
    vec3 m_albedo;
    float m_fuzz;
    float m_ir;
};

struct Aabb {  // OptixAabb's layout: min xyz, max xyz
    float minX, minY, minZ, maxX, maxY, maxZ;
};

class Sphere {
public:
    Sphere(const vec3& center, float radius, const Material& material)
        : m_center(center), m_radius(radius), m_material(material)
    {
    }
    vec3 GetCenter() const { return m_center; }
    float GetRadius() const { return m_radius; }
    const Material& GetMaterial() const { return m_material; }
    void SetMaterial(const Material& m) { m_material = m; }
    // c -+ |r| (sphere.h:17-28 computes |r| but then uses r; |r| is meant)
    Aabb GetAABB() const
    {
        float r = m_radius < 0.0f ? -m_radius : m_radius;
        return {m_center.x - r, m_center.y - r, m_center.z - r, m_center.x + r, m_center.y + r, m_center.z + r};
    }

private:
    vec3 m_center;
    float m_radius;
    Material m_material;
};

class Scene {
public:
    // The book final scene from std::mt19937(seed) (Scene.h:13-80).
    explicit Scene(uint32_t seed = 5489)
    {
        uint32_t n = 0;
        std::vector<vzSphere> s(1 + 22 * 22 + 3);  // ground, grid, three large spheres
        std::vector<vzMaterial> m(s.size());
        VZ_CHECK(vzGenerateBookScene(seed, s.data(), m.data(), (uint32_t)s.size(), &n));
        s.resize(n);
        m.resize(n);
        Assign(s, m);
    }

    // Ground sphere plus n_small random spheres (the 1M-sphere workload).
    static Scene Box(uint32_t n_small, uint32_t seed)
    {
        std::vector<vzSphere> s(n_small + 1);
        std::vector<vzMaterial> m(n_small + 1);
        VZ_CHECK(vzGenerateBoxScene(n_small, seed, s.data(), m.data()));
        Scene sc(Empty{});
        sc.Assign(s, m);
        return sc;
    }

    void Add(const Sphere& s)
    {
        m_spheres.push_back(s);
        m_aabbs.push_back(s.GetAABB());
        m_indices.push_back((uint32_t)m_indices.size());
    }

    // Flattened arrays for the C ABI.
    void ToArrays(std::vector<vzSphere>& s, std::vector<vzMaterial>& m) const
    {
        s.resize(m_spheres.size());
        m.resize(m_spheres.size());
        for (size_t i = 0; i < m_spheres.size(); i++) {
            const Sphere& sp = m_spheres[i];
            vec3 c = sp.GetCenter();
            s[i] = vzSphere{{c.x, c.y, c.z}, sp.GetRadius()};
            const Material& mt = sp.GetMaterial();
            vec3 a = mt.GetAlbedo();
            m[i] = vzMaterial{(uint32_t)mt.GetType(), {a.x, a.y, a.z}, mt.GetFuzz(), mt.GetIR()};
        }
    }

    std::vector<Sphere> m_spheres;
    std::vector<Aabb> m_aabbs;
    std::vector<uint32_t> m_indices;

private:
    struct Empty {};
    explicit Scene(Empty) {}

    void Assign(const std::vector<vzSphere>& s, const std::vector<vzMaterial>& m)
    {
        m_spheres.clear();
        m_aabbs.clear();
        m_indices.clear();
        for (size_t i = 0; i < s.size(); i++) {
            Material mat((Material::Type)m[i].type, vec3(m[i].albedo[0], m[i].albedo[1], m[i].albedo[2]), m[i].fuzz,
                         m[i].ir);
            Add(Sphere(vec3(s[i].center[0], s[i].center[1], s[i].center[2]), s[i].radius, mat));
        }
    }
};
