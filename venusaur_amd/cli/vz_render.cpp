// vz_render -- headless replacement of the reference's windowed main
// (Core/Core.cpp:239-438): builds a scene, calls Renderer::Init, then
// Renderer::Draw for N progressive frames into a HIPOutputBuffer, and writes
// the result as PPM (8-bit sRGB, make_color) and/or PFM (linear radiance).
//
//   vz_render [--scene book|box] [--spheres N] [--seed S] [--width W]
//             [--height H] [--spp S] [--depth D] [--frames F] [--fp64]
//             [--buffer device|zero_copy|p2p] [--move-after K --move-to X,Y,Z]
//             [--ppm file] [--pfm file] [--accum file]
//
// --buffer picks the HIPOutputBuffer type the frames are drawn into (the
// reference picks CUDA_DEVICE when headless, Core/Core.cpp:251-252).
// --move-after K --move-to X,Y,Z moves the camera there (still looking at the
// scene's target) after K frames, as the interactive loop does on input
// (Core/Core.cpp:33-82): the next Draw restarts the progressive average.
//
// Rows are stored bottom-up on the device (y = 0 is the bottom row, the GL
// texture convention of Core.cpp:141): PPM output flips them, PFM (which is
// bottom-up by definition) does not.  Prints one JSON line of statistics.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../core/Renderer.h"

namespace {

struct Args {
    std::string scene = "book", ppm, pfm, accum, buffer = "device";
    uint32_t spheres = 1000000, seed = 5489, width = 1200, height = 800, spp = 16, depth = 4, frames = 1;
    uint32_t move_after = 0;
    float move_to[3] = {0.0f, 0.0f, 0.0f};
    bool fp64 = false;
};

Args parse(int argc, char** argv)
{
    Args a;
    for (int i = 1; i < argc; i++) {
        std::string k = argv[i];
        auto val = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", k.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        if (k == "--scene") a.scene = val();
        else if (k == "--spheres") a.spheres = (uint32_t)std::stoul(val());
        else if (k == "--seed") a.seed = (uint32_t)std::stoul(val());
        else if (k == "--width") a.width = (uint32_t)std::stoul(val());
        else if (k == "--height") a.height = (uint32_t)std::stoul(val());
        else if (k == "--spp") a.spp = (uint32_t)std::stoul(val());
        else if (k == "--depth") a.depth = (uint32_t)std::stoul(val());
        else if (k == "--frames") a.frames = (uint32_t)std::stoul(val());
        else if (k == "--fp64") a.fp64 = true;
        else if (k == "--ppm") a.ppm = val();
        else if (k == "--pfm") a.pfm = val();
        else if (k == "--accum") a.accum = val();
        else if (k == "--buffer") a.buffer = val();
        else if (k == "--move-after") a.move_after = (uint32_t)std::stoul(val());
        else if (k == "--move-to") {
            const std::string v = val();
            if (std::sscanf(v.c_str(), "%f,%f,%f", &a.move_to[0], &a.move_to[1], &a.move_to[2]) != 3) {
                std::fprintf(stderr, "--move-to wants X,Y,Z\n");
                std::exit(2);
            }
        }
        else {
            std::fprintf(stderr, "unknown option %s\n", k.c_str());
            std::exit(2);
        }
    }
    return a;
}

}  // namespace

int main(int argc, char** argv)
{
    Args a = parse(argc, argv);
    try {
        Scene scene = a.scene == "box" ? Scene::Box(a.spheres, a.seed) : Scene(a.seed);
        Renderer renderer;
        auto t0 = std::chrono::steady_clock::now();
        renderer.Init(scene, "");
        auto t1 = std::chrono::steady_clock::now();
        vzAccelInfo ai{};
        VZ_CHECK(vzAccelInfoGet(renderer.SceneHandle(), &ai));
        renderer.SetSamplesPerPixel(a.spp);
        renderer.SetMaxDepth(a.depth);
        renderer.SetAccumFP64(a.fp64);
        const float aspect = (float)a.width / (float)a.height;
        Camera camera;
        if (a.scene == "box") {
            camera = Camera(vec3(0.0f, 15.0f, -120.0f), 40.0f, aspect, 0.0f, 10.0f);
            camera.SetForward(vec3(0.0f, 0.0f, 0.0f) - vec3(0.0f, 15.0f, -120.0f));
        } else {
            camera = Camera(vec3(13.0f, 2.0f, 3.0f), 20.0f, aspect, 0.1f, 10.0f);
            camera.SetForward(vec3(0.0f, 0.0f, 0.0f) - vec3(13.0f, 2.0f, 3.0f));
        }
        HIPOutputBufferType btype = HIPOutputBufferType::HIP_DEVICE;
        if (a.buffer == "zero_copy") btype = HIPOutputBufferType::ZERO_COPY;
        else if (a.buffer == "p2p") btype = HIPOutputBufferType::HIP_P2P;
        else if (a.buffer != "device") throw Exception("--buffer must be device, zero_copy or p2p");
        HIPOutputBuffer<uchar4> out(btype, (int32_t)a.width, (int32_t)a.height);
        unsigned long long segments = 0;
        double gpu_ms = 0.0;
        auto t2 = std::chrono::steady_clock::now();
        const vec3 target(0.0f, 0.0f, 0.0f);
        for (uint32_t f = 0; f < a.frames; f++) {
            if (a.move_after && f == a.move_after) {
                const vec3 to(a.move_to[0], a.move_to[1], a.move_to[2]);
                camera.SetPosition(to);
                camera.SetForward(target - to);
            }
            renderer.Draw(camera, out);
            segments += renderer.LastStats().segments;
            gpu_ms += renderer.LastStats().render_ms;
        }
        auto t3 = std::chrono::steady_clock::now();
        const size_t npx = (size_t)a.width * a.height;
        std::vector<double> lin(npx * 3);
        if (a.fp64) {
            std::vector<double> acc(npx * 4);
            HIP_CHECK(hipMemcpy(acc.data(), renderer.AccumDevicePointer(), npx * 32, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < npx; i++)
                for (int k = 0; k < 3; k++) lin[3 * i + k] = acc[4 * i + k];
            if (!a.accum.empty()) {
                FILE* f = std::fopen(a.accum.c_str(), "wb");
                if (!f) throw Exception("cannot write " + a.accum);
                std::fwrite(acc.data(), 8, acc.size(), f);
                std::fclose(f);
            }
        } else {
            std::vector<float> acc(npx * 4);
            HIP_CHECK(hipMemcpy(acc.data(), renderer.AccumDevicePointer(), npx * 16, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < npx; i++)
                for (int k = 0; k < 3; k++) lin[3 * i + k] = acc[4 * i + k];
            if (!a.accum.empty()) {
                FILE* f = std::fopen(a.accum.c_str(), "wb");
                if (!f) throw Exception("cannot write " + a.accum);
                std::fwrite(acc.data(), 4, acc.size(), f);
                std::fclose(f);
            }
        }
        if (!a.ppm.empty()) {
            const uchar4* px = out.getHostPointer();
            FILE* f = std::fopen(a.ppm.c_str(), "wb");
            if (!f) throw Exception("cannot write " + a.ppm);
            std::fprintf(f, "P6\n%u %u\n255\n", a.width, a.height);
            for (int y = (int)a.height - 1; y >= 0; y--)
                for (uint32_t x = 0; x < a.width; x++) {
                    const uchar4 p = px[(size_t)y * a.width + x];
                    unsigned char rgb[3] = {p.x, p.y, p.z};
                    std::fwrite(rgb, 1, 3, f);
                }
            std::fclose(f);
        }
        if (!a.pfm.empty()) {
            FILE* f = std::fopen(a.pfm.c_str(), "wb");
            if (!f) throw Exception("cannot write " + a.pfm);
            std::fprintf(f, "PF\n%u %u\n-1.0\n", a.width, a.height);  // little endian, bottom row first
            std::vector<float> row(3 * a.width);
            for (uint32_t y = 0; y < a.height; y++) {
                for (size_t i = 0; i < 3 * (size_t)a.width; i++) row[i] = (float)lin[3 * (size_t)y * a.width + i];
                std::fwrite(row.data(), 4, row.size(), f);
            }
            std::fclose(f);
        }
        const double wall = std::chrono::duration<double>(t3 - t2).count();
        std::printf("{\"scene\": \"%s\", \"spheres\": %u, \"width\": %u, \"height\": %u, \"spp\": %u, "
                    "\"depth\": %u, \"frames\": %u, \"segments\": %llu, \"wall_s\": %.6f, \"gpu_ms\": %.3f, "
                    "\"mray_segments_per_s\": %.2f, \"lbvh_build_ms\": %.3f, \"init_s\": %.4f, \"bvh_depth\": %u}\n",
                    a.scene.c_str(), (unsigned)scene.m_spheres.size(), a.width, a.height, a.spp, a.depth, a.frames,
                    segments, wall, gpu_ms, segments / wall / 1e6, ai.build_ms,
                    std::chrono::duration<double>(t1 - t0).count(), ai.depth);
        renderer.Cleanup();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "vz_render: %s\n", e.what());
        return 1;
    }
    return 0;
}
