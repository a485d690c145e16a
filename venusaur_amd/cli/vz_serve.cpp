// vz_serve -- headless front end of the serving form: reads a stream of
// request records (vz_api.h wire format), keeps each scene resident under its
// key, submits the requests to a vzServer (continuous batching, admission by
// HBM footprint) and reports per-request latency and p50/p99.
//
//   vz_serve --in FILE [--budget-mb N] [--max-active N] [--realtime]
//            [--accum-dir DIR] [--ppm-dir DIR]
//
// FILE is a regular file of concatenated request records (its size bounds
// every record's inline scene).  Every record is decoded once before any GPU
// work, so a malformed stream fails before the first request runs.  Each
// request gets a cleared accumulator, so a subframe_index above 1 blends into
// zeros.
// --realtime submits each request at its arrival_ms after the start; without
// it every request is submitted as soon as it is read.  A request whose key is
// not resident must carry its scene inline; an inline scene for a resident
// key is ignored (the resident one is used).  --accum-dir writes each
// request's raw accumulator (<tag>.accum: float4 or double4 per pixel, rows
// bottom-up), --ppm-dir its 8-bit image (<tag>.ppm, rows top-down).
//
// Prints one JSON line per request, in submission order, then a summary line.
// Exit status 0 when every request succeeded, 1 otherwise, 2 on bad usage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vz_api.h"

namespace {

struct Args {
    std::string in, accum_dir, ppm_dir;
    uint64_t budget_mb = 64 * 1024;
    uint32_t max_active = 256;
    bool realtime = false;
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
        if (k == "--in") a.in = val();
        else if (k == "--budget-mb") a.budget_mb = std::stoull(val());
        else if (k == "--max-active") a.max_active = (uint32_t)std::stoul(val());
        else if (k == "--realtime") a.realtime = true;
        else if (k == "--accum-dir") a.accum_dir = val();
        else if (k == "--ppm-dir") a.ppm_dir = val();
        else {
            std::fprintf(stderr, "unknown option %s\n", k.c_str());
            std::exit(2);
        }
    }
    if (a.in.empty()) {
        std::fprintf(stderr, "usage: vz_serve --in FILE [--budget-mb N] [--max-active N] [--realtime] "
                             "[--accum-dir DIR] [--ppm-dir DIR]\n");
        std::exit(2);
    }
    return a;
}

struct Fail {
    std::string msg;
};

void ck(vzStatus s, const char* what)
{
    if (s != VZ_SUCCESS) throw Fail{std::string(what) + ": " + vzStatusString(s) + ": " + vzGetLastError()};
}
void hck(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Fail{std::string(what) + ": " + hipGetErrorString(e)};
}

// One request from the stream, from submission until its result is written.
struct Job {
    vzRequestDesc desc;
    uint64_t id = 0;          // server request id (0: not submitted)
    vzStatus status = VZ_SUCCESS;
    std::string error;
    void* accum = nullptr;
    void* image = nullptr;
    vzRequestResult res{};
    bool done = false;
};

void write_outputs(const Args& a, Job& j)
{
    const vzRequestDesc& r = j.desc;
    const size_t npx = (size_t)r.width * r.height;
    const size_t abytes = npx * ((r.flags & VZ_ACCUM_FP64) ? 32 : 16);
    if (!a.accum_dir.empty()) {
        std::vector<unsigned char> h(abytes);
        hck(hipMemcpy(h.data(), j.accum, abytes, hipMemcpyDeviceToHost), "accum download");
        const std::string path = a.accum_dir + "/" + std::to_string(r.tag) + ".accum";
        FILE* f = std::fopen(path.c_str(), "wb");
        if (!f || std::fwrite(h.data(), 1, h.size(), f) != h.size()) throw Fail{"cannot write " + path};
        std::fclose(f);
    }
    if (!a.ppm_dir.empty()) {
        std::vector<unsigned char> px(npx * 4);
        hck(hipMemcpy(px.data(), j.image, px.size(), hipMemcpyDeviceToHost), "image download");
        const std::string path = a.ppm_dir + "/" + std::to_string(r.tag) + ".ppm";
        FILE* f = std::fopen(path.c_str(), "wb");
        if (!f) throw Fail{"cannot write " + path};
        std::fprintf(f, "P6\n%u %u\n255\n", r.width, r.height);
        for (int y = (int)r.height - 1; y >= 0; y--)
            for (uint32_t x = 0; x < r.width; x++) std::fwrite(&px[4 * ((size_t)y * r.width + x)], 1, 3, f);
        std::fclose(f);
    }
}

void release(Job& j)
{
    if (j.accum) (void)hipFree(j.accum);
    if (j.image) (void)hipFree(j.image);
    j.accum = j.image = nullptr;
}

// Polls the unfinished submitted jobs; writes and frees the finished ones.
void reap(const Args& a, vzServer srv, std::vector<Job>& jobs, size_t& first_open)
{
    for (size_t i = first_open; i < jobs.size(); i++) {
        Job& j = jobs[i];
        if (j.done || j.id == 0) continue;
        const vzStatus s = vzServerPoll(srv, j.id, &j.res);
        if (s == VZ_ERROR_NOT_READY) continue;
        if (s != VZ_SUCCESS) {
            j.status = s;
            j.error = vzGetLastError();
        } else if (j.res.status != VZ_SUCCESS) {
            j.status = (vzStatus)j.res.status;
            j.error = "request failed on the server";
        } else {
            write_outputs(a, j);
        }
        release(j);
        j.done = true;
    }
    while (first_open < jobs.size() && jobs[first_open].done) first_open++;
}

// Reads the next request record (header and inline scene) into rec; false
// at the end of the file.  Throws on a truncated record or an inline scene
// that runs past the end of the file.
bool read_record(FILE* in, uint64_t file_size, uint64_t& offset, uint64_t index, std::vector<unsigned char>& rec)
{
    rec.resize(VZ_WIRE_REQUEST_HEADER);
    const size_t got = std::fread(rec.data(), 1, VZ_WIRE_REQUEST_HEADER, in);
    if (got == 0) return false;
    const std::string where = "record " + std::to_string(index);
    if (got < VZ_WIRE_REQUEST_HEADER) throw Fail{where + ": truncated header"};
    uint32_t inline_bytes = 0;  // offset 88 of the request header
    std::memcpy(&inline_bytes, rec.data() + 88, 4);
    offset += VZ_WIRE_REQUEST_HEADER;
    if (inline_bytes > file_size - offset) throw Fail{where + ": inline scene runs past the end of the file"};
    offset += inline_bytes;
    rec.resize(VZ_WIRE_REQUEST_HEADER + (size_t)inline_bytes);
    if (inline_bytes && std::fread(rec.data() + VZ_WIRE_REQUEST_HEADER, 1, inline_bytes, in) != inline_bytes)
        throw Fail{where + ": truncated inline scene"};
    return true;
}

// First pass: every record decodes (sizes, checksums, values), before any
// GPU work.  Returns the number of records.
uint64_t validate_stream(FILE* in, uint64_t file_size)
{
    uint64_t offset = 0, index = 0;
    std::vector<unsigned char> rec;
    vzRequestDesc d;
    while (read_record(in, file_size, offset, index, rec)) {
        if (vzWireDecodeRequest(rec.data(), rec.size(), &d, nullptr, nullptr) != VZ_SUCCESS)
            throw Fail{"record " + std::to_string(index) + ": " + vzGetLastError()};
        index++;
    }
    if (index == 0) throw Fail{"the stream holds no request record"};
    return index;
}

double pct(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t k = (size_t)std::min<double>((double)v.size() - 1, std::max(0.0, q * (double)v.size() - 1e-9));
    return v[k];
}

}  // namespace

int main(int argc, char** argv)
{
    const Args a = parse(argc, argv);
    FILE* in = std::fopen(a.in.c_str(), "rb");
    if (!in) {
        std::fprintf(stderr, "vz_serve: cannot open %s\n", a.in.c_str());
        return 2;
    }
    vzContext ctx = nullptr;
    vzServer srv = nullptr;
    std::map<uint64_t, vzScene> scenes;
    std::vector<Job> jobs;
    size_t first_open = 0;
    int rc = 0;
    uint64_t inline_built = 0, inline_ignored = 0;
    double wall_s = 0.0;
    try {
        std::fseek(in, 0, SEEK_END);
        const long end = std::ftell(in);
        if (end < 0) throw Fail{a.in + " is not a seekable file"};
        const uint64_t file_size = (uint64_t)end;
        std::fseek(in, 0, SEEK_SET);
        validate_stream(in, file_size);
        std::fseek(in, 0, SEEK_SET);
        ck(vzCreate(0, &ctx), "vzCreate");
        vzServeConfig cfg{};
        cfg.hbm_budget_bytes = a.budget_mb << 20;
        cfg.max_active = a.max_active;
        // cfg.tuning stays zero: the serving defaults (12M-ray queue, 8x8 tiles, split 4)
        ck(vzServerCreate(ctx, &cfg, &srv), "vzServerCreate");
        const auto t0 = std::chrono::steady_clock::now();
        auto now_ms = [&] {
            return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        };
        uint64_t offset = 0;
        std::vector<unsigned char> rec;
        for (uint64_t index = 0; read_record(in, file_size, offset, index, rec); index++) {
            const std::string where = "record " + std::to_string(index);
            Job j;
            size_t sb = 0;
            if (vzWireDecodeRequest(rec.data(), rec.size(), &j.desc, &sb, nullptr) != VZ_SUCCESS)
                throw Fail{where + ": " + vzGetLastError()};
            const vzRequestDesc& r = j.desc;
            if (a.realtime)
                while (now_ms() < r.arrival_ms) {
                    reap(a, srv, jobs, first_open);
                    std::this_thread::sleep_for(std::chrono::microseconds(200));
                }
            vzScene sc = nullptr;
            auto it = scenes.find(r.scene_key);
            if (it != scenes.end()) {
                sc = it->second;
                if (sb) inline_ignored++;
            } else if (sb) {
                const unsigned char* sr = rec.data() + VZ_WIRE_REQUEST_HEADER;
                uint32_t n = 0;
                ck(vzWireDecodeScene(sr, sb, nullptr, nullptr, 0, &n, nullptr), "scene record");
                std::vector<vzSphere> sph(n);
                std::vector<vzMaterial> mat(n);
                ck(vzWireDecodeScene(sr, sb, sph.data(), mat.data(), n, &n, nullptr), "scene record");
                ck(vzSceneCreate(ctx, sph.data(), mat.data(), n, nullptr, &sc), "vzSceneCreate");
                scenes[r.scene_key] = sc;
                inline_built++;
            } else {
                j.status = VZ_ERROR_INVALID_VALUE;
                j.error = "scene key " + std::to_string(r.scene_key) + " is not resident and the record has no scene";
                j.done = true;
                jobs.push_back(j);
                continue;
            }
            const size_t npx = (size_t)r.width * r.height;
            const size_t abytes = npx * ((r.flags & VZ_ACCUM_FP64) ? 32 : 16);
            hck(hipMalloc(&j.accum, abytes), "accum allocation");
            hck(hipMalloc(&j.image, npx * 4), "image allocation");
            // the server renders on its own non-blocking stream: the clear must
            // be complete before the request is submitted
            hck(hipMemsetAsync(j.accum, 0, abytes, nullptr), "accum clear");
            hck(hipStreamSynchronize(nullptr), "accum clear");
            vzLaunchParams p;
            ck(vzRequestLaunchParams(&r, j.image, j.accum, &p), "vzRequestLaunchParams");
            const vzStatus sub = vzServerSubmit(srv, sc, &p, &j.id);
            if (sub != VZ_SUCCESS) {
                j.status = sub;
                j.error = vzGetLastError();
                j.id = 0;
                j.done = true;
                release(j);
            }
            jobs.push_back(j);
            reap(a, srv, jobs, first_open);
        }
        ck(vzServerDrain(srv), "vzServerDrain");
        reap(a, srv, jobs, first_open);
        wall_s = now_ms() / 1e3;
        std::vector<double> lat;
        uint64_t segments = 0, ok = 0;
        for (const Job& j : jobs) {
            const vzRequestDesc& r = j.desc;
            if (j.status == VZ_SUCCESS && j.done) {
                ok++;
                segments += j.res.segments;
                lat.push_back(j.res.finish_ms - j.res.submit_ms);
                std::printf("{\"tag\": %llu, \"scene_key\": %llu, \"status\": \"VZ_SUCCESS\", \"width\": %u, "
                            "\"height\": %u, \"spp\": %u, \"depth\": %u, \"fp64\": %s, \"segments\": %llu, "
                            "\"arrival_ms\": %.3f, \"submit_ms\": %.3f, \"admit_ms\": %.3f, \"finish_ms\": %.3f, "
                            "\"latency_ms\": %.3f}\n",
                            (unsigned long long)r.tag, (unsigned long long)r.scene_key, r.width, r.height,
                            r.samples_per_pixel, r.max_depth, (r.flags & VZ_ACCUM_FP64) ? "true" : "false",
                            (unsigned long long)j.res.segments, r.arrival_ms, j.res.submit_ms, j.res.admit_ms,
                            j.res.finish_ms, j.res.finish_ms - j.res.submit_ms);
            } else {
                rc = 1;
                std::string msg = j.error;
                for (char& c : msg)
                    if (c == '"' || c == '\\') c = '\'';
                std::printf("{\"tag\": %llu, \"scene_key\": %llu, \"status\": \"%s\", \"error\": \"%s\"}\n",
                            (unsigned long long)r.tag, (unsigned long long)r.scene_key, vzStatusString(j.status),
                            msg.c_str());
            }
        }
        std::printf("{\"requests\": %zu, \"ok\": %llu, \"segments\": %llu, \"wall_s\": %.6f, "
                    "\"mray_segments_per_s\": %.2f, \"p50_ms\": %.3f, \"p99_ms\": %.3f, \"scenes_resident\": %zu, "
                    "\"inline_scenes_built\": %llu, \"inline_scenes_ignored\": %llu, \"realtime\": %s}\n",
                    jobs.size(), (unsigned long long)ok, (unsigned long long)segments, wall_s,
                    wall_s > 0 ? segments / wall_s / 1e6 : 0.0, pct(lat, 0.50), pct(lat, 0.99), scenes.size(),
                    (unsigned long long)inline_built, (unsigned long long)inline_ignored,
                    a.realtime ? "true" : "false");
    } catch (const Fail& f) {
        std::fprintf(stderr, "vz_serve: %s\n", f.msg.c_str());
        rc = 1;
    }
    std::fclose(in);
    // teardown in dependency order: server (drains), scenes, context
    if (srv) (void)vzServerDestroy(srv);
    for (Job& j : jobs) release(j);
    for (auto& kv : scenes) (void)vzSceneDestroy(kv.second);
    if (ctx && vzDestroy(ctx) != VZ_SUCCESS) {
        std::fprintf(stderr, "vz_serve: vzDestroy: %s\n", vzGetLastError());
        rc = 1;
    }
    return rc;
}
