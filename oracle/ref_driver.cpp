// ref_driver.cpp — C entry points over the reference's own builders and
// tracers, linked with the reference sources (builder.cpp, tracer.cpp,
// algorithms.cpp) into oracle/_ref/libseqref.so by oracle/Makefile.
//
// TEST INFRASTRUCTURE ONLY: used in this container to write golden vectors
// (tests/golden/make_golden.py) and, where the built library is present, as
// the timed CPU baseline. Nothing here restates reference arithmetic: every
// number returned is computed by the reference's functions.
#include "aabb.h"
#include "algorithms.h"
#include "builder.h"
#include "ray.h"
#include "tracer.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

// The builders allocate Node1[] with new[] and leave unused slots
// uninitialised (Node1() {}); zero-filling makes those slots deterministic so
// whole arrays can be compared. Bound inside this library with -Bsymbolic.
void* operator new[](std::size_t sz) {
    void* p = std::calloc(1, sz ? sz : 1);
    if (!p) throw std::bad_alloc();
    return p;
}
void operator delete[](void* p) noexcept { std::free(p); }
void operator delete[](void* p, std::size_t) noexcept { std::free(p); }

// Same layout and name as the key/value pairs inside builder.cpp:2196-2246 so
// that setup_morton/setup_morton2 link against this declaration.
typedef struct alignas(8) RadixInput32 {
    uint32_t code;
    unsigned index;
    __forceinline operator uint32_t() const { return code; }
    __forceinline bool operator<(const RadixInput32& m) const { return code < m.code; }
} RadixInput32;
typedef struct alignas(16) RadixInput64 {
    uint64_t code;
    unsigned index;
    __forceinline operator uint64_t() const { return code; }
    __forceinline bool operator<(const RadixInput64& m) const { return code < m.code; }
} RadixInput64;

// main.cpp:108 and the stb writer compiled into main.o (main.cpp:25-26);
// main.cpp's main() is compiled as seqref_cli_main (oracle/Makefile).
std::vector<Triangle> loadOBJFromFile(std::string filename);
int seqref_cli_main(int argc, char** argv);
extern "C" void stbi_flip_vertically_on_write(int flag);
extern "C" int stbi_write_bmp(char const* filename, int w, int h, int comp, const void* data);

void setup_morton(std::vector<Triangle>& primitives, std::vector<RadixInput32>& morton_codes, AABB1& scene_aabb);
void setup_morton2(std::vector<Triangle>& primitives, std::vector<RadixInput64>& morton_codes, AABB1& scene_aabb);

static std::vector<Triangle> to_tris(const float* t, uint32_t n) {
    std::vector<Triangle> v(n);
    std::memcpy(v.data(), t, sizeof(Triangle) * (size_t)n);
    return v;
}

extern "C" {

uint32_t ref_sizeof_node() { return sizeof(Node1); }
uint32_t ref_sizeof_triangle() { return sizeof(Triangle); }

uint32_t ref_interleave32(uint32_t x, uint32_t y, uint32_t z) { return bit_interleave_32(x, y, z); }
uint64_t ref_interleave64(uint64_t x, uint64_t y, uint64_t z) { return bit_interleave_64(x, y, z); }

// kind: 0 recursive, 1 horizontal, 2 grid, 3 grid_sse, 4 grid_avx, 5 morton, 6 morton2.
// nodes_out must hold 2n (+128 for the grid builders) entries of 32 B.
int ref_build(int kind, const float* tris, uint32_t n, void* nodes_out, uint32_t* root) {
    std::vector<Triangle> prims = to_tris(tris, n);
    Hierarchy h;
    size_t cap = (size_t)2 * n;
    switch (kind) {
        case 0: h = Builder::build_hierarchy(prims); break;
        case 1: h = Builder::build_hierarchy_horizontal(prims); break;
        case 2: h = Builder::build_hierarchy_grid(prims); cap += 128; break;
        case 3: h = Builder::build_hierarchy_grid_sse(prims); cap += 128; break;
        case 4: h = Builder::build_hierarchy_grid_m128(prims); cap += 128; break;
        case 5: h = Builder::build_hierarchy_morton(prims); break;
        case 6: h = Builder::build_hierarchy_morton2(prims); break;
        default: return -1;
    }
    std::memcpy(nodes_out, h.node_base, cap * sizeof(Node1));
    *root = h.root_index;
    delete[] h.node_base;
    return 0;
}

// Wall time (steady_clock, ms) of `reps` builds of one kind, each on a fresh
// copy of the triangles made outside the timed region; times[r] per build.
// The CPU baseline's build figures, measured the way main.cpp:376-392 brackets
// the builder call (with steady_clock in place of clock()).
int ref_time_build(int kind, const float* tris, uint32_t n, uint32_t reps, double* times) {
    for (uint32_t r = 0; r < reps; r++) {
        std::vector<Triangle> prims = to_tris(tris, n);
        Hierarchy h;
        auto t0 = std::chrono::steady_clock::now();
        switch (kind) {
            case 0: h = Builder::build_hierarchy(prims); break;
            case 1: h = Builder::build_hierarchy_horizontal(prims); break;
            case 2: h = Builder::build_hierarchy_grid(prims); break;
            case 3: h = Builder::build_hierarchy_grid_sse(prims); break;
            case 4: h = Builder::build_hierarchy_grid_m128(prims); break;
            case 5: h = Builder::build_hierarchy_morton(prims); break;
            case 6: h = Builder::build_hierarchy_morton2(prims); break;
            default: return -1;
        }
        auto t1 = std::chrono::steady_clock::now();
        times[r] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        delete[] h.node_base;
    }
    return 0;
}

// setup_morton + radix_sort<uint32_t> exactly as build_hierarchy_morton runs them.
void ref_morton32(const float* tris, uint32_t n, uint32_t* codes_unsorted, uint32_t* codes, uint32_t* idx) {
    std::vector<Triangle> prims = to_tris(tris, n);
    std::vector<RadixInput32> mc(n), scratch(n);
    AABB1 scene;
    setup_morton(prims, mc, scene);
    for (uint32_t i = 0; i < n; i++) codes_unsorted[i] = mc[i].code;
    radix_sort<uint32_t, RadixInput32>(mc.begin(), mc.end(), scratch.begin(), scratch.end(), 8);
    for (uint32_t i = 0; i < n; i++) { codes[i] = mc[i].code; idx[i] = mc[i].index; }
}

void ref_morton64(const float* tris, uint32_t n, uint64_t* codes_unsorted, uint64_t* codes, uint32_t* idx) {
    std::vector<Triangle> prims = to_tris(tris, n);
    std::vector<RadixInput64> mc(n), scratch(n);
    AABB1 scene;
    setup_morton2(prims, mc, scene);
    for (uint32_t i = 0; i < n; i++) codes_unsorted[i] = mc[i].code;
    radix_sort<uint64_t, RadixInput64>(mc.begin(), mc.end(), scratch.begin(), scratch.end(), 8);
    for (uint32_t i = 0; i < n; i++) { codes[i] = mc[i].code; idx[i] = mc[i].index; }
}

// radix_sort over caller keys (values = original positions).
void ref_radix_sort32(const uint32_t* keys, uint32_t n, uint32_t* out_keys, uint32_t* out_idx) {
    std::vector<RadixInput32> v(n), s(n);
    for (uint32_t i = 0; i < n; i++) { v[i].code = keys[i]; v[i].index = i; }
    radix_sort<uint32_t, RadixInput32>(v.begin(), v.end(), s.begin(), s.end(), 8);
    for (uint32_t i = 0; i < n; i++) { out_keys[i] = v[i].code; out_idx[i] = v[i].index; }
}

static Hierarchy make_h(void* nodes, uint32_t root, std::vector<Triangle>& prims) {
    Hierarchy h;
    h.node_base = (Node1*)nodes;
    h.node_count = 0;
    h.tri_base = prims.data();
    h.plucker_tri_base = nullptr;
    h.tri_count = (unsigned)prims.size();
    h.root_index = root;
    return h;
}

// mode 0: trace(); 1: trace_packet<K>; 2: trace_bundle<K>. rgb holds w*h*3.
int ref_trace(int mode, int K, void* nodes, uint32_t root, const float* tris, uint32_t n, uint32_t w,
              uint32_t hgt, const float* origin, uint8_t* rgb) {
    std::vector<Triangle> prims = to_tris(tris, n);
    Hierarchy h = make_h(nodes, root, prims);
    vec3f1 o(origin[0], origin[1], origin[2]);
    uint8_t* img = nullptr;
    if (mode == 0) img = trace(h, w, hgt, o);
    else if (mode == 1) {
        switch (K) {
            case 1: img = trace_packet<1>(h, w, hgt, o); break;
            case 2: img = trace_packet<2>(h, w, hgt, o); break;
            case 4: img = trace_packet<4>(h, w, hgt, o); break;
            case 8: img = trace_packet<8>(h, w, hgt, o); break;
            default: return -1;
        }
    } else if (mode == 2) {
        switch (K) {
            case 16: img = trace_bundle<16>(h, w, hgt, o); break;
            case 64: img = trace_bundle<64>(h, w, hgt, o); break;
            default: return -1;
        }
    } else return -1;
    std::memcpy(rgb, img, (size_t)w * hgt * 3);
    delete[] img;
    return 0;
}

// traverse() over caller rays (Ray<1> layout: o, tmin, d, tmax).
void ref_traverse_rays(void* nodes, uint32_t root, const float* tris, uint32_t n, const float* rays,
                       uint32_t nrays, uint32_t* counts, uint32_t* hit) {
    std::vector<Triangle> prims = to_tris(tris, n);
    Hierarchy h = make_h(nodes, root, prims);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)nrays; i++) {
        Ray<1> r;
        const float* q = rays + 8 * i;
        r.origin[0][0] = q[0]; r.origin[1][0] = q[1]; r.origin[2][0] = q[2];
        r.tmin = simdf<1>::broadcast(q[3]);
        r.direction[0][0] = q[4]; r.direction[1][0] = q[5]; r.direction[2][0] = q[6];
        r.tmax = simdf<1>::broadcast(q[7]);
        HitAttributes<1> a;
        traverse(h, root, 1, r, a);
        counts[i] = (uint32_t)a.num_tests[0];
        hit[i] = a.hit[0] ? 1u : 0u;
    }
}

// loadOBJFromFile (main.cpp:108-179): copies up to `cap` triangles, returns
// how many the file holds.
uint32_t ref_load_obj(const char* path, float* out, uint32_t cap) {
    std::vector<Triangle> t = loadOBJFromFile(std::string(path));
    const size_t n = t.size() < cap ? t.size() : cap;
    std::memcpy(out, t.data(), sizeof(Triangle) * n);
    return (uint32_t)t.size();
}

// The reference command line (main.cpp:339-538). It ends with exit(0), so
// callers run it in a child process; files go to the working directory.
int ref_cli_main(int argc, char** argv) { return seqref_cli_main(argc, argv); }

// The heatmap file as main.cpp:335-336 writes it.
int ref_write_bmp(const char* path, int w, int h, const uint8_t* rgb) {
    stbi_flip_vertically_on_write(1);
    return stbi_write_bmp(path, w, h, 3, rgb);
}

}  // extern "C"
