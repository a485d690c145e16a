// vz_wire.cpp -- scene and request records, the input of the serving form
// (layout in include/vz_api.h).  The reference keeps its scene in code
// (Core/Scene.h:13-80) and its per-launch state in Params
// (Core/RayTracer.h:3-17); these records carry the same content between
// processes.  Host code only: encoding and decoding never touch the GPU.
#include <cmath>
#include <cstring>
#include <string>

#include "vz_internal.h"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "the records are little-endian");
static_assert(sizeof(vzSphere) == 16 && sizeof(vzMaterial) == 24, "record element sizes");

namespace vz {
namespace {

constexpr char kSceneMagic[4] = {'V', 'Z', 'S', 'C'};
constexpr char kRequestMagic[4] = {'V', 'Z', 'R', 'Q'};
constexpr uint32_t kVersion = 1;
constexpr uint32_t kMaxSpheres = (1u << 30) - 1;  // vzSceneCreate's limit
constexpr size_t kSceneHeader = VZ_WIRE_SCENE_HEADER, kRequestHeader = VZ_WIRE_REQUEST_HEADER;
constexpr size_t kPerSphere = sizeof(vzSphere) + sizeof(vzMaterial);

// FNV-1a, 64-bit (Fowler, Noll, Vo): offset basis 14695981039346656037,
// prime 1099511628211.
uint64_t fnv1a64(const uint8_t* p, size_t n)
{
    uint64_t h = 14695981039346656037ull;
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

template <typename T>
T get(const uint8_t* p, size_t off)
{
    T v;
    std::memcpy(&v, p + off, sizeof(T));
    return v;
}
template <typename T>
void put(uint8_t* p, size_t off, T v)
{
    std::memcpy(p + off, &v, sizeof(T));
}

vzStatus check_materials(const vzMaterial* m, uint32_t n)
{
    for (uint32_t i = 0; i < n; i++) {
        if (m[i].type > VZ_DIELECTRIC)
            return set_error(VZ_ERROR_INVALID_VALUE, "material " + std::to_string(i) + " has an unknown type");
        if (!std::isfinite(m[i].albedo[0]) || !std::isfinite(m[i].albedo[1]) || !std::isfinite(m[i].albedo[2]) ||
            !std::isfinite(m[i].fuzz) || !std::isfinite(m[i].ir))
            return set_error(VZ_ERROR_INVALID_VALUE, "material " + std::to_string(i) + " has a non-finite value");
    }
    return VZ_SUCCESS;
}

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

// Everything vzServerSubmit would reject, plus the record's own fields.
vzStatus check_request(const vzRequestDesc* r)
{
    if (r->flags & ~VZ_ACCUM_FP64) return set_error(VZ_ERROR_INVALID_VALUE, "request flags other than VZ_ACCUM_FP64");
    if (r->subframe_index < 1) return set_error(VZ_ERROR_INVALID_VALUE, "subframe_index must be >= 1");
    if (!finite3(r->lookfrom) || !finite3(r->lookat))
        return set_error(VZ_ERROR_INVALID_VALUE, "camera position is not finite");
    if (r->lookfrom[0] == r->lookat[0] && r->lookfrom[1] == r->lookat[1] && r->lookfrom[2] == r->lookat[2])
        return set_error(VZ_ERROR_INVALID_VALUE, "lookfrom equals lookat");
    if (!(r->vfov_deg > 0.0f && r->vfov_deg < 180.0f)) return set_error(VZ_ERROR_INVALID_VALUE, "vfov must be in (0, 180)");
    if (!(r->focus_dist > 0.0f) || !std::isfinite(r->focus_dist))
        return set_error(VZ_ERROR_INVALID_VALUE, "focus_dist must be positive and finite");
    if (!(r->aperture >= 0.0f) || !std::isfinite(r->aperture))
        return set_error(VZ_ERROR_INVALID_VALUE, "aperture must be >= 0 and finite");
    if (!(r->arrival_ms >= 0.0f) || !std::isfinite(r->arrival_ms))
        return set_error(VZ_ERROR_INVALID_VALUE, "arrival_ms must be >= 0 and finite");
    vzLaunchParams p;
    // check_launch_params wants an accum buffer; any non-NULL address stands in
    VZ_TRY(vzRequestLaunchParams(r, nullptr, reinterpret_cast<void*>(uintptr_t(256)), &p));
    return VZ_SUCCESS;
}

}  // namespace
}  // namespace vz

using namespace vz;

extern "C" {

size_t vzWireSceneSize(uint32_t n) { return kSceneHeader + kPerSphere * (size_t)n; }

vzStatus vzWireEncodeScene(const vzSphere* spheres, const vzMaterial* materials, uint32_t n, void* buf, size_t cap,
                           size_t* written)
{
    if (!spheres || !materials || !buf) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    if (n == 0 || n > kMaxSpheres) return set_error(VZ_ERROR_INVALID_VALUE, "sphere count must be in [1, 2^30)");
    const size_t need = vzWireSceneSize(n);
    if (cap < need) return set_error(VZ_ERROR_INVALID_VALUE, "buffer smaller than the scene record");
    VZ_TRY(check_spheres(spheres, n));
    VZ_TRY(check_materials(materials, n));
    uint8_t* p = static_cast<uint8_t*>(buf);
    std::memcpy(p, kSceneMagic, 4);
    put<uint32_t>(p, 4, kVersion);
    put<uint32_t>(p, 8, n);
    put<uint32_t>(p, 12, 0u);
    std::memcpy(p + kSceneHeader, spheres, sizeof(vzSphere) * (size_t)n);
    std::memcpy(p + kSceneHeader + sizeof(vzSphere) * (size_t)n, materials, sizeof(vzMaterial) * (size_t)n);
    put<uint64_t>(p, 16, fnv1a64(p + kSceneHeader, need - kSceneHeader));
    if (written) *written = need;
    return VZ_SUCCESS;
}

vzStatus vzWireDecodeScene(const void* buf, size_t size, vzSphere* spheres, vzMaterial* materials, uint32_t cap,
                           uint32_t* n, size_t* consumed)
{
    if (!buf || !n) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    if ((spheres == nullptr) != (materials == nullptr))
        return set_error(VZ_ERROR_INVALID_VALUE, "spheres and materials must both be given or both be NULL");
    const uint8_t* p = static_cast<const uint8_t*>(buf);
    if (size < kSceneHeader) return set_error(VZ_ERROR_INVALID_VALUE, "scene record truncated (header)");
    if (std::memcmp(p, kSceneMagic, 4) != 0) return set_error(VZ_ERROR_INVALID_VALUE, "not a scene record");
    if (get<uint32_t>(p, 4) != kVersion) return set_error(VZ_ERROR_INVALID_VALUE, "unsupported scene record version");
    const uint32_t count = get<uint32_t>(p, 8);
    if (count == 0 || count > kMaxSpheres) return set_error(VZ_ERROR_INVALID_VALUE, "scene record sphere count out of range");
    if (get<uint32_t>(p, 12) != 0u) return set_error(VZ_ERROR_INVALID_VALUE, "scene record reserved word is not 0");
    const size_t need = vzWireSceneSize(count);
    if (size < need) return set_error(VZ_ERROR_INVALID_VALUE, "scene record truncated (payload)");
    if (fnv1a64(p + kSceneHeader, need - kSceneHeader) != get<uint64_t>(p, 16))
        return set_error(VZ_ERROR_INVALID_VALUE, "scene record checksum mismatch");
    // values are checked in chunks copied out of the (possibly unaligned) buffer
    const uint8_t* ps = p + kSceneHeader;
    const uint8_t* pm = ps + sizeof(vzSphere) * (size_t)count;
    constexpr uint32_t kChunk = 1024;
    vzSphere s[kChunk];
    vzMaterial m[kChunk];
    for (uint32_t i = 0; i < count; i += kChunk) {
        const uint32_t k = count - i < kChunk ? count - i : kChunk;
        std::memcpy(s, ps + sizeof(vzSphere) * (size_t)i, sizeof(vzSphere) * k);
        std::memcpy(m, pm + sizeof(vzMaterial) * (size_t)i, sizeof(vzMaterial) * k);
        if (check_spheres(s, k) != VZ_SUCCESS || check_materials(m, k) != VZ_SUCCESS)
            return set_error(VZ_ERROR_INVALID_VALUE, "scene record element " + std::to_string(i) + "+: " + vzGetLastError());
    }
    if (spheres) {
        if (cap < count) return set_error(VZ_ERROR_INVALID_VALUE, "output arrays smaller than the record's sphere count");
        std::memcpy(spheres, ps, sizeof(vzSphere) * (size_t)count);
        std::memcpy(materials, pm, sizeof(vzMaterial) * (size_t)count);
    }
    *n = count;
    if (consumed) *consumed = need;
    return VZ_SUCCESS;
}

vzStatus vzWireEncodeRequest(const vzRequestDesc* req, const vzSphere* spheres, const vzMaterial* materials,
                             uint32_t n, void* buf, size_t cap, size_t* written)
{
    if (!req || !buf) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    VZ_TRY(check_request(req));
    const size_t scene = n ? vzWireSceneSize(n) : 0;
    if (cap < kRequestHeader + scene) return set_error(VZ_ERROR_INVALID_VALUE, "buffer smaller than the request record");
    if (scene > 0xFFFFFFFFull) return set_error(VZ_ERROR_INVALID_VALUE, "inline scene larger than 4 GiB");
    uint8_t* p = static_cast<uint8_t*>(buf);
    if (n) VZ_TRY(vzWireEncodeScene(spheres, materials, n, p + kRequestHeader, cap - kRequestHeader, nullptr));
    std::memcpy(p, kRequestMagic, 4);
    put<uint32_t>(p, 4, kVersion);
    put<uint64_t>(p, 8, req->tag);
    put<uint64_t>(p, 16, req->scene_key);
    put<uint32_t>(p, 24, req->width);
    put<uint32_t>(p, 28, req->height);
    put<uint32_t>(p, 32, req->samples_per_pixel);
    put<uint32_t>(p, 36, req->max_depth);
    put<uint32_t>(p, 40, req->flags);
    put<uint32_t>(p, 44, req->subframe_index);
    for (int k = 0; k < 3; k++) {
        put<float>(p, 48 + 4 * k, req->lookfrom[k]);
        put<float>(p, 60 + 4 * k, req->lookat[k]);
    }
    put<float>(p, 72, req->vfov_deg);
    put<float>(p, 76, req->aperture);
    put<float>(p, 80, req->focus_dist);
    put<float>(p, 84, req->arrival_ms);
    put<uint32_t>(p, 88, (uint32_t)scene);
    put<uint32_t>(p, 92, 0u);
    if (written) *written = kRequestHeader + scene;
    return VZ_SUCCESS;
}

vzStatus vzWireDecodeRequest(const void* buf, size_t size, vzRequestDesc* out, size_t* scene_bytes, size_t* consumed)
{
    if (!buf || !out) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    const uint8_t* p = static_cast<const uint8_t*>(buf);
    if (size < kRequestHeader) return set_error(VZ_ERROR_INVALID_VALUE, "request record truncated (header)");
    if (std::memcmp(p, kRequestMagic, 4) != 0) return set_error(VZ_ERROR_INVALID_VALUE, "not a request record");
    if (get<uint32_t>(p, 4) != kVersion) return set_error(VZ_ERROR_INVALID_VALUE, "unsupported request record version");
    if (get<uint32_t>(p, 92) != 0u) return set_error(VZ_ERROR_INVALID_VALUE, "request record reserved word is not 0");
    vzRequestDesc r;
    r.tag = get<uint64_t>(p, 8);
    r.scene_key = get<uint64_t>(p, 16);
    r.width = get<uint32_t>(p, 24);
    r.height = get<uint32_t>(p, 28);
    r.samples_per_pixel = get<uint32_t>(p, 32);
    r.max_depth = get<uint32_t>(p, 36);
    r.flags = get<uint32_t>(p, 40);
    r.subframe_index = get<uint32_t>(p, 44);
    for (int k = 0; k < 3; k++) {
        r.lookfrom[k] = get<float>(p, 48 + 4 * k);
        r.lookat[k] = get<float>(p, 60 + 4 * k);
    }
    r.vfov_deg = get<float>(p, 72);
    r.aperture = get<float>(p, 76);
    r.focus_dist = get<float>(p, 80);
    r.arrival_ms = get<float>(p, 84);
    const size_t scene = get<uint32_t>(p, 88);
    VZ_TRY(check_request(&r));
    if (scene) {
        if (size - kRequestHeader < scene) return set_error(VZ_ERROR_INVALID_VALUE, "request record truncated (inline scene)");
        uint32_t n = 0;
        size_t used = 0;
        VZ_TRY(vzWireDecodeScene(p + kRequestHeader, scene, nullptr, nullptr, 0, &n, &used));
        if (used != scene) return set_error(VZ_ERROR_INVALID_VALUE, "inline scene size does not match its record");
    }
    *out = r;
    if (scene_bytes) *scene_bytes = scene;
    if (consumed) *consumed = kRequestHeader + scene;
    return VZ_SUCCESS;
}

vzStatus vzRequestLaunchParams(const vzRequestDesc* req, void* image, void* accum, vzLaunchParams* out)
{
    if (!req || !out) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    vzLaunchParams p;
    std::memset(&p, 0, sizeof(p));
    p.image = image;
    p.accum = accum;
    p.width = req->width;
    p.height = req->height;
    p.samples_per_pixel = req->samples_per_pixel;
    p.subframe_index = req->subframe_index;
    p.max_depth = req->max_depth;
    p.flags = req->flags;
    const float aspect = (float)req->width / (float)req->height;
    VZ_TRY(vzCameraLookAt(req->lookfrom, req->lookat, req->vfov_deg, aspect, req->aperture, req->focus_dist, &p));
    VZ_TRY(check_launch_params(&p));
    *out = p;
    return VZ_SUCCESS;
}

}  // extern "C"
