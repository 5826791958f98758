// compat.cpp — the reference's Builder / trace entry points (sequoia_compat.h)
// implemented over the C ABI. One process-wide context serves every call;
// hierarchies built here keep their device copy for later traces.
#include "sequoia_compat.h"

#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "sequoia.h"

namespace {

struct Entry {
    sq_scene* scene = nullptr;
    sq_bvh* bvh = nullptr;
};

std::mutex g_mu;
sq_ctx* g_ctx = nullptr;
int g_device = 0;
std::unordered_map<const Node1*, Entry> g_live;

void check(int rc, const char* what) {
    if (rc != SQ_OK) throw std::runtime_error(std::string(what) + ": " + sq_last_error());
}

sq_ctx* ctx() {
    if (!g_ctx) check(sq_ctx_create(g_device, &g_ctx), "sequoia context");
    return g_ctx;
}

Hierarchy build(std::vector<Triangle>& prims, sq_builder kind) {
    if (prims.empty()) throw std::runtime_error("Builder: empty primitive list");
    std::lock_guard<std::mutex> lock(g_mu);
    sq_ctx* c = ctx();
    sq_scene* s = nullptr;
    check(sq_scene_upload(c, reinterpret_cast<const sq_triangle*>(prims.data()), (uint32_t)prims.size(), &s),
          "scene upload");
    sq_build_params p;
    std::memset(&p, 0, sizeof(p));
    sq_bvh* b = nullptr;
    int rc = sq_build(c, s, kind, &p, &b);
    if (rc != SQ_OK) {
        sq_scene_free(s);
        check(rc, "build");
    }
    uint32_t root = 0, cap = 0, depth = 0;
    sq_bvh_info(b, &root, &cap, &depth);
    Node1* nodes = new Node1[cap];
    rc = sq_bvh_download(b, reinterpret_cast<sq_node*>(nodes), cap);
    if (rc != SQ_OK) {
        delete[] nodes;
        sq_bvh_free(b);
        sq_scene_free(s);
        check(rc, "download");
    }
    g_live[nodes] = Entry{s, b};
    Hierarchy h;
    h.node_base = nodes;
    h.node_count = cap;
    h.tri_base = prims.data();
    h.plucker_tri_base = nullptr;
    h.tri_count = (unsigned)prims.size();
    h.root_index = root;
    return h;
}

// Largest triangle index referenced by a leaf reachable from the root.
uint32_t tri_extent(const Hierarchy& h) {
    std::vector<uint32_t> st{h.root_index};
    uint32_t mx = 0;
    size_t guard = 0;
    while (!st.empty() && guard++ < (size_t)h.node_count * 4 + 64) {
        uint32_t i = st.back();
        st.pop_back();
        if (h.node_count && i >= h.node_count) throw std::runtime_error("trace: node index outside the hierarchy");
        const Node1& n = h.node_base[i];
        if (n.flags & AABB_FLAGS_LEAF) {
            if (n.count && n.child + n.count > mx) mx = n.child + n.count;
        } else {
            for (uint32_t j = 0; j < n.count; j++) st.push_back(n.child + j);
        }
    }
    return mx;
}

const sq_bvh* device_tree(Hierarchy& h) {
    auto it = g_live.find(h.node_base);
    if (it != g_live.end()) return it->second.bvh;
    // A hierarchy built elsewhere: upload its nodes and triangles once.
    if (!h.node_count) throw std::runtime_error("trace: hierarchy without node_count");
    uint32_t ntri = h.tri_count ? h.tri_count : tri_extent(h);
    sq_ctx* c = ctx();
    Entry e;
    check(sq_scene_upload(c, reinterpret_cast<const sq_triangle*>(h.tri_base), ntri, &e.scene), "scene upload");
    int rc = sq_bvh_from_host(c, e.scene, reinterpret_cast<const sq_node*>(h.node_base), h.node_count, h.root_index,
                              &e.bvh);
    if (rc != SQ_OK) {
        sq_scene_free(e.scene);
        check(rc, "hierarchy upload");
    }
    g_live[h.node_base] = e;
    return e.bvh;
}

uint8_t* trace_mode(Hierarchy& h, uint32_t w, uint32_t hh, vec3f1 origin, sq_heatmap mode, uint32_t K) {
    std::lock_guard<std::mutex> lock(g_mu);
    const sq_bvh* b = device_tree(h);
    sq_camera cam;
    std::memset(&cam, 0, sizeof(cam));
    cam.origin = {origin.x, origin.y, origin.z};
    cam.tan_half_fov = 1.0f;
    cam.width = w;
    cam.height = hh;
    cam.spp = 1;
    cam.tmin = 0.0f;
    cam.tmax = 10000.0f;
    sq_trace_params p;
    std::memset(&p, 0, sizeof(p));
    p.semantics = SQ_TRACE_EXHAUSTIVE;
    p.heatmap = mode;
    p.width = K;
    uint8_t* img = new uint8_t[(size_t)w * hh * 3];
    sq_trace_out out;
    std::memset(&out, 0, sizeof(out));
    out.rgb = img;
    int rc = sq_trace_camera(ctx(), b, &cam, &p, &out, 0);
    if (rc != SQ_OK) {
        delete[] img;
        check(rc, "trace");
    }
    return img;
}

}  // namespace

Hierarchy Builder::build_hierarchy(std::vector<Triangle>& p) { return build(p, SQ_BUILDER_RECURSIVE); }
Hierarchy Builder::build_hierarchy_morton(std::vector<Triangle>& p) { return build(p, SQ_BUILDER_MORTON); }
Hierarchy Builder::build_hierarchy_morton2(std::vector<Triangle>& p) { return build(p, SQ_BUILDER_MORTON2); }
Hierarchy Builder::build_hierarchy_grid(std::vector<Triangle>& p) { return build(p, SQ_BUILDER_GRID); }
Hierarchy Builder::build_hierarchy_grid_sse(std::vector<Triangle>& p) { return build(p, SQ_BUILDER_GRID_SSE); }
Hierarchy Builder::build_hierarchy_grid_m128(std::vector<Triangle>& p) { return build(p, SQ_BUILDER_GRID_AVX); }
Hierarchy Builder::build_hierarchy_horizontal(std::vector<Triangle>& p) { return build(p, SQ_BUILDER_HORIZONTAL); }

uint8_t* trace(Hierarchy& h, uint32_t w, uint32_t hh, vec3f1 origin) {
    return trace_mode(h, w, hh, origin, SQ_HEATMAP_RAY, 1);
}

template <size_t K>
uint8_t* trace_packet(Hierarchy& h, uint32_t w, uint32_t hh, vec3f1 origin) {
    return trace_mode(h, w, hh, origin, SQ_HEATMAP_PACKET, (uint32_t)K);
}
template <size_t K>
uint8_t* trace_bundle(Hierarchy& h, uint32_t w, uint32_t hh, vec3f1 origin) {
    return trace_mode(h, w, hh, origin, SQ_HEATMAP_BUNDLE, (uint32_t)K);
}

template uint8_t* trace_packet<1>(Hierarchy&, uint32_t, uint32_t, vec3f1);
template uint8_t* trace_packet<2>(Hierarchy&, uint32_t, uint32_t, vec3f1);
template uint8_t* trace_packet<4>(Hierarchy&, uint32_t, uint32_t, vec3f1);
template uint8_t* trace_packet<8>(Hierarchy&, uint32_t, uint32_t, vec3f1);
template uint8_t* trace_bundle<16>(Hierarchy&, uint32_t, uint32_t, vec3f1);
template uint8_t* trace_bundle<64>(Hierarchy&, uint32_t, uint32_t, vec3f1);

// Bit i of x lands at bit 3i, of y at 3i+1, of z at 3i+2 (the reference's
// pdep masks 0x49249249 / 0x92492492 / 0x24924924 and their 64-bit widening);
// bits that fall beyond the word are dropped, as pdep drops them.
template <typename U>
static U interleave3(U x, U y, U z) {
    const unsigned W = sizeof(U) * 8;
    U r = 0;
    for (unsigned i = 0; 3 * i < W; i++) {
        r |= ((x >> i) & U(1)) << (3 * i);
        if (3 * i + 1 < W) r |= ((y >> i) & U(1)) << (3 * i + 1);
        if (3 * i + 2 < W) r |= ((z >> i) & U(1)) << (3 * i + 2);
    }
    return r;
}

uint32_t bit_interleave_32(const uint32_t& x, const uint32_t& y, const uint32_t& z) { return interleave3(x, y, z); }
uint64_t bit_interleave_64(const uint64_t& x, const uint64_t& y, const uint64_t& z) { return interleave3(x, y, z); }

namespace sequoia_compat {

void sort_pairs(uint64_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, uint32_t digit_bits) {
    std::lock_guard<std::mutex> lock(g_mu);
    check(sq_radix_sort_pairs(ctx(), keys, vals, n, key_bits, digit_bits), "radix_sort");
}

void set_device(int device) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_device = device;
}

void release(Hierarchy& h) {
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_live.find(h.node_base);
    if (it != g_live.end()) {
        sq_bvh_free(it->second.bvh);
        sq_scene_free(it->second.scene);
        g_live.erase(it);
    }
    delete[] h.node_base;
    h.node_base = nullptr;
}

}  // namespace sequoia_compat
