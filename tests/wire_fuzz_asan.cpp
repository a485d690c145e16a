// Host-only fuzz harness for the wire decoders (venusaur_amd/csrc/vz_wire.cpp),
// built with g++ under AddressSanitizer and UBSan by tests/test_wire_asan.py.
// It links vz_wire.cpp alone: the library's own validators and error store,
// defined in vz_api.cpp next to the HIP code, are replaced by the minimal
// stubs below.  Truncated, extended and bit-flipped copies of a valid scene
// record and a valid request record (with an inline scene) go through every
// decoder, each from an exact-size heap buffer so that any read past the end
// is caught.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>
#include "../venusaur_amd/csrc/vz_internal.h"
namespace vz {
vzStatus set_error(vzStatus s, const std::string&) { return s; }
vzStatus check_launch_params(const vzLaunchParams* p) { return p ? VZ_SUCCESS : VZ_ERROR_INVALID_VALUE; }
vzStatus check_spheres(const vzSphere* s, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) if (!(s[i].radius == s[i].radius)) return VZ_ERROR_INVALID_VALUE;
    return VZ_SUCCESS;
}
}
extern "C" vzStatus vzCameraLookAt(const float*, const float*, float, float, float, float, vzLaunchParams*) { return VZ_SUCCESS; }
extern "C" const char* vzGetLastError(void) { return ""; }
int main() {
    std::mt19937 rng(7);
    // a valid scene record and request record to mutate
    std::vector<vzSphere> sph(5); std::vector<vzMaterial> mat(5);
    for (int i = 0; i < 5; i++) { sph[i] = vzSphere{{float(i), 0.f, 0.f}, 0.5f}; std::memset(&mat[i], 0, sizeof(vzMaterial)); }
    std::vector<uint8_t> scene(vzWireSceneSize(5));
    if (vzWireEncodeScene(sph.data(), mat.data(), 5, scene.data(), scene.size(), nullptr) != VZ_SUCCESS) { std::puts("encode scene failed"); }
    vzRequestDesc d; std::memset(&d, 0, sizeof d);
    d.width = 8; d.height = 8; d.samples_per_pixel = 1; d.max_depth = 4; d.subframe_index = 1;
    d.lookfrom[0] = 1; d.vfov_deg = 40; d.focus_dist = 1;
    std::vector<uint8_t> req(VZ_WIRE_REQUEST_HEADER + scene.size());
    size_t used = 0;
    vzStatus es = vzWireEncodeRequest(&d, sph.data(), mat.data(), 5, req.data(), req.size(), &used);
    std::printf("encode request: %d, %zu bytes\n", (int)es, used);
    long ok = 0, bad = 0;
    for (int it = 0; it < 200000; it++) {
        const std::vector<uint8_t>& base = (it & 1) ? scene : req;
        size_t len = rng() % (base.size() + 8);
        std::vector<uint8_t> b(len);
        for (size_t i = 0; i < len; i++) b[i] = i < base.size() ? base[i] : (uint8_t)rng();
        int flips = rng() % 4;
        for (int f = 0; f < flips && len; f++) b[rng() % len] ^= (uint8_t)(1u << (rng() % 8));
        // exact-size heap copy so any over-read is caught
        uint8_t* h = (uint8_t*)std::malloc(len ? len : 1); if (len) std::memcpy(h, b.data(), len);
        vzStatus s;
        if (it & 1) {
            uint32_t n = 0; size_t cons = 0;
            s = vzWireDecodeScene(h, len, nullptr, nullptr, 0, &n, &cons);
            if (s == VZ_SUCCESS && n <= 64) {
                std::vector<vzSphere> o(n); std::vector<vzMaterial> m(n);
                s = vzWireDecodeScene(h, len, o.data(), m.data(), n, &n, &cons);
                if (s == VZ_SUCCESS && cons > len) { std::puts("scene consumed beyond the buffer"); return 1; }
            }
        } else {
            vzRequestDesc o; size_t sb = 0, cons = 0;
            s = vzWireDecodeRequest(h, len, &o, &sb, &cons);
            if (s == VZ_SUCCESS && cons > len) { std::puts("consumed beyond the buffer"); return 1; }
            if (s == VZ_SUCCESS && sb) {
                uint32_t n = 0; size_t c2 = 0;
                s = vzWireDecodeScene(h + VZ_WIRE_REQUEST_HEADER, len - VZ_WIRE_REQUEST_HEADER, nullptr, nullptr, 0, &n, &c2);
                if (s == VZ_SUCCESS && n <= 64) {
                    std::vector<vzSphere> o2(n); std::vector<vzMaterial> m2(n);
                    s = vzWireDecodeScene(h + VZ_WIRE_REQUEST_HEADER, len - VZ_WIRE_REQUEST_HEADER, o2.data(), m2.data(), n, &n, &c2);
                }
            }
        }
        std::free(h);
        (s == VZ_SUCCESS ? ok : bad)++;
    }
    std::printf("fuzz done: %ld accepted, %ld rejected\n", ok, bad);
    return 0;
}
