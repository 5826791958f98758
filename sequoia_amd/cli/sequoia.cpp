// sequoia — the reference's command-line harness on the MI355X library.
//
// Same command line, defaults and output files as main.cpp:339-538:
//   sequoia <scene.obj> [-builder recursive|horizontal|grid|grid_sse|grid_avx|morton]
//                       [-tracer basic | packet <1|2|4|8> | bundle <16|64>]
//                       [-threads <n>] [-res <w> <h>] [-benchmark]
// A run loads the OBJ, places run()'s camera (scene box centre, z pulled back
// by the box half-depth, main.cpp:311-318), builds, traces and writes
// render.bmp; -benchmark builds with all six builders, traces basic,
// packet<8> and bundle<64> at 1920x1080 on the recursive tree and writes
// benchmark1.bmp .. benchmark3.bmp (main.cpp:339-414). Builds and traces run
// on GPU 0 through the C ABI; times are wall-clock (steady_clock) around the
// synchronous library calls, scene upload included in the build time.
//
// Differences: the morton builder's tree carries refitted boxes (the
// reference writes none, so its morton renders are not meaningful); a
// missing option value, an unsupported width or a failing call prints an
// error and exits with status 1 where the reference asserts; -threads only
// sets the number reported (device work does not use host threads).
#include "sequoia.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Args {
    sq_builder builder = SQ_BUILDER_RECURSIVE;
    int tracer = 0;  // 0 basic, 1 packet, 2 bundle
    uint32_t packet_width = 8, bundle_width = 64;
    uint32_t w = 1920, h = 1080;
    uint32_t threads = 64;
    bool benchmark = false;
    std::string file;
};

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

[[noreturn]] void die(const char* what) {
    std::fprintf(stderr, "Error: %s: %s\n", what, sq_last_error());
    std::exit(1);
}

void check(int rc, const char* what) {
    if (rc != SQ_OK) die(what);
}

void usage() {
    std::printf(
        "Usage:\n"
        "    sequoia <scene.obj> [<args>]\n"
        "        args:\n"
        "            -builder <type> : Specify the builder to use. One of the following:\n"
        "                recursive      : binned SAH, one split per task (GPU: levels of tasks in parallel)\n"
        "                horizontal     : same tree as recursive\n"
        "                grid           : 4x4x4 centroid grid, SAH inside each cell\n"
        "                grid_sse       : grid with the stable partition and tight cell boxes\n"
        "                grid_avx       : same tree as grid_sse\n"
        "                morton         : 30-bit Morton codes, radix sort, top-down emission (boxes refitted)\n"
        "            -tracer  <type> : Specify the traverser to use. One of the following:\n"
        "                basic          : A single ray traverser\n"
        "                packet <n>     : A packet traverser of n rays (1, 2, 4, 8)\n"
        "                bundle <n>     : A bundle traverser of n rays (16, 64)\n"
        "            -threads <n>    : Maximum number of threads (reported only)\n"
        "            -res <x> <y>    : Set the output resolution\n"
        "            -benchmark      : run a benchmark mode testing all builders and traversers\n");
}

const char* need(int argc, char** argv, int& i, const char* opt) {
    if (++i >= argc) {
        std::fprintf(stderr, "Error: %s needs a value\n", opt);
        std::exit(1);
    }
    return argv[i];
}

Args parse(int argc, char** argv) {
    Args a;
    if (argc < 2) {
        std::printf("Error: not enough arguments\n");
        usage();
        std::exit(0);
    }
    a.file = argv[1];
    for (int i = 2; i < argc; i++) {
        const std::string opt = argv[i];
        if (opt == "-builder") {
            const std::string v = need(argc, argv, i, "-builder");
            if (v == "recursive") a.builder = SQ_BUILDER_RECURSIVE;
            else if (v == "horizontal") a.builder = SQ_BUILDER_HORIZONTAL;
            else if (v == "grid") a.builder = SQ_BUILDER_GRID;
            else if (v == "grid_sse") a.builder = SQ_BUILDER_GRID_SSE;
            else if (v == "grid_avx") a.builder = SQ_BUILDER_GRID_AVX;
            else if (v == "morton") a.builder = SQ_BUILDER_MORTON;
            else {
                std::printf("Error: invalid builder type\n");
                std::exit(0);
            }
        } else if (opt == "-tracer") {
            const std::string v = need(argc, argv, i, "-tracer");
            if (v == "basic") {
                a.tracer = 0;
            } else if (v == "packet") {
                a.tracer = 1;
                a.packet_width = (uint32_t)std::atoi(need(argc, argv, i, "-tracer packet"));
                if (a.packet_width != 1 && a.packet_width != 2 && a.packet_width != 4 && a.packet_width != 8) {
                    std::fprintf(stderr, "Error: invalid packet width %u (1, 2, 4 or 8)\n", a.packet_width);
                    std::exit(1);
                }
            } else if (v == "bundle") {
                a.tracer = 2;
                a.bundle_width = (uint32_t)std::atoi(need(argc, argv, i, "-tracer bundle"));
                if (a.bundle_width != 16 && a.bundle_width != 64) {
                    std::fprintf(stderr, "Error: invalid bundle size %u (16 or 64)\n", a.bundle_width);
                    std::exit(1);
                }
            } else {
                std::printf("Error: invalid tracer type\n");
                std::exit(0);
            }
        } else if (opt == "-threads") {
            a.threads = (uint32_t)std::atoi(need(argc, argv, i, "-threads"));
        } else if (opt == "-res") {
            a.w = (uint32_t)std::atoi(need(argc, argv, i, "-res"));
            a.h = (uint32_t)std::atoi(need(argc, argv, i, "-res"));
            if (a.w == 0 || a.h == 0) {
                std::fprintf(stderr, "Error: -res needs a positive width and height\n");
                std::exit(1);
            }
        } else if (opt == "-benchmark") {
            a.benchmark = true;
        }
        // other words are ignored, as by the reference
    }
    return a;
}

struct Scene {
    std::vector<sq_triangle> tris;
    uint32_t verts = 0;
};

Scene load(const std::string& file) {
    sq_triangle* t = nullptr;
    uint32_t n = 0, nv = 0;
    check(sq_load_obj(file.c_str(), &t, &n, &nv), "loading the OBJ file");
    Scene s;
    s.tris.assign(t, t + n);
    s.verts = nv;
    sq_host_free(t);
    std::printf("Geometry\n");
    std::printf("  faces:        %u\n", n);
    std::printf("  verts:        %u\n", nv);
    return s;
}

// Build wall time (upload + build, synchronous) in ms.
sq_bvh* build(sq_ctx* ctx, const Scene& s, sq_builder kind, sq_scene** scene_out, double* ms) {
    const double t0 = now_ms();
    sq_scene* sc = nullptr;
    check(sq_scene_upload(ctx, s.tris.data(), (uint32_t)s.tris.size(), &sc), "uploading the scene");
    sq_bvh* b = nullptr;
    check(sq_build(ctx, sc, kind, nullptr, &b), "building");
    check(sq_ctx_sync(ctx), "building");
    *ms = now_ms() - t0;
    *scene_out = sc;
    return b;
}

std::vector<uint8_t> trace(sq_ctx* ctx, const sq_bvh* b, const Scene& s, uint32_t w, uint32_t h, int tracer,
                           uint32_t width, double* ms) {
    sq_camera cam;
    check(sq_reference_camera(s.tris.data(), (uint32_t)s.tris.size(), w, h, &cam), "placing the camera");
    sq_trace_params p;
    std::memset(&p, 0, sizeof p);
    p.semantics = SQ_TRACE_EXHAUSTIVE;
    p.heatmap = tracer == 0 ? SQ_HEATMAP_RAY : (tracer == 1 ? SQ_HEATMAP_PACKET : SQ_HEATMAP_BUNDLE);
    p.width = tracer == 0 ? 0 : width;
    std::vector<uint8_t> rgb((size_t)w * h * 3);
    sq_trace_out out;
    std::memset(&out, 0, sizeof out);
    out.rgb = rgb.data();
    const double t0 = now_ms();
    check(sq_trace_camera(ctx, b, &cam, &p, &out, 0), "tracing");
    *ms = now_ms() - t0;
    return rgb;
}

void save(const char* name, const std::vector<uint8_t>& rgb, uint32_t w, uint32_t h) {
    check(sq_write_bmp(name, rgb.data(), w, h), "writing the image");
}

void run(sq_ctx* ctx, const Args& a) {
    double t0 = now_ms();
    Scene s = load(a.file);
    std::printf("Load time:     %.0fms\n", now_ms() - t0);
    if (s.tris.empty()) {
        std::fprintf(stderr, "Error: the scene has no triangles\n");
        std::exit(1);
    }
    double bms = 0, tms = 0;
    sq_scene* sc = nullptr;
    sq_bvh* b = build(ctx, s, a.builder, &sc, &bms);
    std::printf("Build time:    %.0fms\n", bms);
    const uint32_t width = a.tracer == 1 ? a.packet_width : a.bundle_width;
    std::vector<uint8_t> img = trace(ctx, b, s, a.w, a.h, a.tracer, width, &tms);
    std::printf("Traverse time: %.0fms\n", tms);
    save("render.bmp", img, a.w, a.h);
    sq_bvh_free(b);
    sq_scene_free(sc);
}

void benchmark(sq_ctx* ctx, const Args& a) {
    double t0 = now_ms();
    Scene s = load(a.file);
    const double load_s = (now_ms() - t0) / 1e3;
    if (s.tris.empty()) {
        std::fprintf(stderr, "Error: the scene has no triangles\n");
        std::exit(1);
    }
    const sq_builder kinds[6] = {SQ_BUILDER_RECURSIVE, SQ_BUILDER_HORIZONTAL, SQ_BUILDER_GRID,
                                 SQ_BUILDER_GRID_SSE, SQ_BUILDER_GRID_AVX, SQ_BUILDER_MORTON};
    double bms[6];
    sq_bvh* rec = nullptr;
    sq_scene* rec_scene = nullptr;
    for (int k = 0; k < 6; k++) {
        sq_scene* sc = nullptr;
        sq_bvh* b = build(ctx, s, kinds[k], &sc, &bms[k]);
        if (k == 0) {
            rec = b;
            rec_scene = sc;
        } else {
            sq_bvh_free(b);
            sq_scene_free(sc);
        }
    }
    const uint32_t w = 1920, h = 1080;
    double tms[3];
    std::vector<uint8_t> img1 = trace(ctx, rec, s, w, h, 0, 0, &tms[0]);
    std::vector<uint8_t> img2 = trace(ctx, rec, s, w, h, 1, 8, &tms[1]);
    std::vector<uint8_t> img3 = trace(ctx, rec, s, w, h, 2, 64, &tms[2]);
    const double n = (double)s.tris.size();
    std::printf("Load time: %.2fs\n", load_s);
    std::printf("\n");
    std::printf("             recursive horizontal       grid   grid_sse   grid_avx     morton\n");
    std::printf("build time: % 8.0fms % 8.0fms % 8.0fms % 8.0fms % 8.0fms % 8.0fms\n", bms[0], bms[1], bms[2],
                bms[3], bms[4], bms[5]);
    std::printf("tri/s     : % 9.0fm % 9.0fm % 9.0fm % 9.0fm % 9.0fm % 9.0fm\n", n / (1e3 * bms[0]),
                n / (1e3 * bms[1]), n / (1e3 * bms[2]), n / (1e3 * bms[3]), n / (1e3 * bms[4]), n / (1e3 * bms[5]));
    std::printf("\n");
    std::printf("             basic packet bundle\n");
    std::printf("trace time: % 4.0fms % 4.0fms % 4.0fms\n", tms[0], tms[1], tms[2]);
    const double mr = w * h / 1e6;
    std::printf("mrps      : % 6.1f % 6.1f % 6.1f\n", mr / (tms[0] / 1e3), mr / (tms[1] / 1e3), mr / (tms[2] / 1e3));
    save("benchmark1.bmp", img1, w, h);
    save("benchmark2.bmp", img2, w, h);
    save("benchmark3.bmp", img3, w, h);
    sq_bvh_free(rec);
    sq_scene_free(rec_scene);
}

}  // namespace

int main(int argc, char** argv) {
    const Args a = parse(argc, argv);
    sq_ctx* ctx = nullptr;
    check(sq_ctx_create(0, &ctx), "opening GPU 0");
    std::printf("Running on GPU 0 (%s)\n", sq_version());
    std::printf("  with %u threads\n\n", a.threads);
    if (a.benchmark) benchmark(ctx, a);
    else run(ctx, a);
    sq_ctx_destroy(ctx);
    return 0;
}
