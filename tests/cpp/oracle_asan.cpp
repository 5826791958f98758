// CPU run of every oracle entry point (oracle/oracle.cpp, node2.cpp,
// nodek.cpp, plucker.cpp) on small seeded scenes, built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_sanitizers.py: each Morton, SAH
// and grid variant is built, then traced exhaustively, closest-hit, per
// camera, per packet and bundle, through the quantised two-child nodes and
// the precomputed-triangle test. Buffers are sized exactly as the header
// states, so any read or write past them aborts. A few cheap agreements are
// checked on the way (closest == exhaustive hits, the Node2
// traversal reaches the same triangles).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "oracle.h"

static uint64_t rng = 0x2545F4914F6CDD1Dull;
static float frand() {
    rng = rng * 6364136223846793005ull + 1442695040888963407ull;
    return (float)((rng >> 40) & 0xFFFFFF) / 16777216.0f;
}

static int failures = 0;
#define CHECK(c, ...)                                  \
    do {                                               \
        if (!(c)) {                                    \
            std::printf("FAIL %s: ", #c);              \
            std::printf(__VA_ARGS__);                  \
            std::printf("\n");                         \
            failures++;                                \
        }                                              \
    } while (0)

// kind 0: soup of small triangles in the unit cube; 1: one triangle repeated
// (every centroid equal); 2: a flat sheet (zero extent along z); 3: long
// thin triangles spanning the scene.
static std::vector<sqo_tri> scene(uint32_t n, int kind) {
    std::vector<sqo_tri> t(n);
    for (uint32_t i = 0; i < n; i++) {
        float c[3] = {frand(), frand(), kind == 2 ? 0.5f : frand()};
        const float s = kind == 3 ? 0.9f : 0.05f;
        sqo_vec3* v[3] = {&t[i].v0, &t[i].v1, &t[i].v2};
        for (int k = 0; k < 3; k++) {
            v[k]->x = c[0] + s * (frand() - 0.5f);
            v[k]->y = c[1] + s * (frand() - 0.5f);
            v[k]->z = kind == 2 ? 0.5f : c[2] + s * (frand() - 0.5f);
        }
        if (kind == 1 && i) t[i] = t[0];
    }
    return t;
}

static std::vector<sqo_ray> rays(uint32_t n) {
    std::vector<sqo_ray> r(n);
    for (auto& x : r) {
        x.ox = frand() * 3 - 1;
        x.oy = frand() * 3 - 1;
        x.oz = -1.5f;
        x.dx = frand() - 0.5f;
        x.dy = frand() - 0.5f;
        x.dz = 1.0f;
        x.tmin = 0.0f;
        x.tmax = 1e30f;
    }
    return r;
}

static sqo_camera camera(uint32_t w, uint32_t h, uint32_t spp, uint32_t jitter) {
    sqo_camera c;
    c.origin[0] = 0.5f;
    c.origin[1] = 0.5f;
    c.origin[2] = -1.5f;
    c.tan_half_fov = 0.6f;
    c.width = w;
    c.height = h;
    c.spp = spp;
    c.jitter = jitter;
    c.seed = 7;
    c.tmin = 0.0f;
    c.tmax = 1e30f;
    return c;
}

static void trace_all(const std::vector<sqo_node>& nodes, uint32_t root, const std::vector<sqo_tri>& tris,
                      const char* tag) {
    const uint32_t n = (uint32_t)tris.size(), nr = 257;
    const auto r = rays(nr);
    std::vector<uint32_t> counts(nr), hid(nr), any(nr), hid2(nr), hid3(nr), cnt3(nr);
    std::vector<float> ht(nr), ht2(nr), ht3(nr);
    sqo_trace_rays(nodes.data(), root, tris.data(), r.data(), nr, counts.data(), hid.data(), ht.data(), any.data());
    sqo_trace_rays_closest(nodes.data(), root, tris.data(), r.data(), nr, hid2.data(), ht2.data());
    for (uint32_t i = 0; i < nr; i++) CHECK(hid[i] == hid2[i], "%s ray %u closest %u vs %u", tag, i, hid2[i], hid[i]);

    // camera, packets, bundles
    const sqo_camera cam = camera(13, 9, 2, 1);
    std::vector<uint32_t> pc(13 * 9), chid(13 * 9 * 2);
    std::vector<uint8_t> rgb(13 * 9 * 3);
    std::vector<float> cht(13 * 9 * 2);
    sqo_trace_camera(nodes.data(), root, tris.data(), &cam, pc.data(), rgb.data(), chid.data(), cht.data());
    const sqo_camera pcam = camera(17, 11, 1, 0);
    std::vector<uint8_t> prgb(17 * 11 * 3);
    for (uint32_t K : {1u, 2u, 4u, 8u}) {
        const uint32_t kx = (uint32_t)std::floor(std::sqrt((double)K)), ky = K / kx;
        std::vector<uint32_t> packets((size_t)((17 + kx - 1) / kx) * ((11 + ky - 1) / ky));
        sqo_trace_packet(nodes.data(), root, tris.data(), &pcam, K, packets.data(), prgb.data());
    }
    for (uint32_t K : {16u, 64u}) sqo_trace_bundle(nodes.data(), root, tris.data(), &pcam, K, prgb.data());

    // quantised two-child nodes: capacities grow until the conversion fits
    uint32_t cap = (uint32_t)nodes.size(), tcap = n;
    for (int attempt = 0; attempt < 12; attempt++) {
        std::vector<sqo_node2> n2(cap);
        std::vector<sqo_tri> tris2(tcap);
        std::vector<uint32_t> perm(tcap);
        uint32_t nn = 0, nt = 0;
        sqo_node2_root r2;
        const int rc = sqo_node2_convert(nodes.data(), (uint32_t)nodes.size(), root, tris.data(), n, 0, n2.data(), cap,
                                         &nn, tris2.data(), perm.data(), tcap, &nt, &r2);
        if (rc == -5) { cap *= 2; continue; }
        if (rc == -4) { tcap *= 2; continue; }
        if (rc != 0) break;  // an unsupported tree (e.g. a leaf of more than 7 triangles)
        for (uint32_t i = 0; i < nn; i++) {
            float a0[3], a1[3], b0[3], b1[3];
            sqo_node2_decode(&n2[i], a0, a1, b0, b1);
        }
        sqo_node2_trace(n2.data(), &r2, tris2.data(), perm.data(), r.data(), nr, 1, cnt3.data(), hid3.data(),
                        ht3.data());
        for (uint32_t i = 0; i < nr; i++) CHECK(hid3[i] == hid[i], "%s node2 ray %u %u vs %u", tag, i, hid3[i], hid[i]);
        sqo_node2_trace(n2.data(), &r2, tris2.data(), perm.data(), r.data(), nr, 0, cnt3.data(), hid3.data(),
                        ht3.data());
        break;
    }

    // wide nodes, K = 2, 4, 8, collapsed or not: capacity grows until it fits
    for (uint32_t K : {2u, 4u, 8u}) {
        for (uint32_t expand : {0u, 1u}) {
            uint32_t wcap = 1 + (uint32_t)nodes.size() / 4;
            for (int attempt = 0; attempt < 12; attempt++) {
                std::vector<uint32_t> wide((size_t)wcap * 8 * K);
                uint32_t nn = 0, levels = 0;
                sqo_nodek_root rk;
                const int rc = sqo_nodek_convert(nodes.data(), (uint32_t)nodes.size(), root, K, expand, wide.data(),
                                                 wcap, &nn, &levels, &rk);
                if (rc == -4) { wcap *= 2; continue; }
                CHECK(rc == 0, "%s nodek K=%u convert %d", tag, K, rc);
                sqo_nodek_trace(wide.data(), K, &rk, tris.data(), r.data(), nr, 1, cnt3.data(), hid3.data(),
                                ht3.data());
                sqo_nodek_trace(wide.data(), K, &rk, tris.data(), r.data(), nr, 0, cnt3.data(), hid3.data(),
                                ht3.data());
                break;
            }
        }
    }

    std::vector<sqo_plucker> pl(n);
    sqo_plucker_precompute(tris.data(), n, pl.data());
    sqo_trace_rays_plucker(nodes.data(), root, pl.data(), r.data(), nr, cnt3.data(), hid3.data(), ht3.data());
}

int main() {
    uint32_t built = 0;
    for (uint32_t n : {1u, 2u, 3u, 7u, 64u, 333u, 2000u}) {
        for (int kind = 0; kind < 4; kind++) {
            const auto tris = scene(n, kind);
            char tag[96];
            float lo[3], hi[3];
            sqo_scene_bounds(tris.data(), n, lo, hi);
            std::vector<uint64_t> codes(n);
            for (uint32_t bits : {10u, 20u, 21u}) sqo_morton_codes(tris.data(), n, bits, codes.data());

            for (uint32_t bits : {10u, 20u, 21u})
                for (uint32_t layout : {0u, 1u}) {
                    std::vector<sqo_node> nodes(2 * (size_t)n);
                    std::vector<uint64_t> sc(n);
                    std::vector<uint32_t> si(n);
                    const uint32_t root = sqo_build_morton(tris.data(), n, bits, layout, 1, nodes.data(), sc.data(),
                                                           si.data());
                    for (uint32_t i = 1; i < n; i++) CHECK(sc[i - 1] <= sc[i], "morton sort n=%u", n);
                    std::snprintf(tag, sizeof(tag), "morton%u/%u n=%u kind=%d", 3 * bits, layout, n, kind);
                    trace_all(nodes, root, tris, tag);
                    built++;
                }
            for (uint32_t bins : {8u, 16u, 32u})
                for (uint32_t part : {0u, 1u})
                    for (uint32_t sa : {0u, 1u, 2u})
                        for (uint32_t layout : {0u, 1u}) {
                            std::vector<sqo_node> nodes(sqo_sah_node_capacity(n));
                            const uint32_t root = sqo_build_sah(tris.data(), n, bins, part, sa, layout, nodes.data());
                            std::snprintf(tag, sizeof(tag), "sah%u/%u/%u/%u n=%u kind=%d", bins, part, sa, layout, n,
                                          kind);
                            if (sa == 0 && part == 0) trace_all(nodes, root, tris, tag);
                            built++;
                        }
            for (uint32_t dim : {2u, 4u})
                for (uint32_t cell : {0u, 1u})
                    for (uint32_t layout : {0u, 1u}) {
                        std::vector<sqo_node> nodes(sqo_grid_node_capacity(n, dim));
                        const uint32_t root = sqo_build_grid(tris.data(), n, dim, 8, 0, 0, cell, layout, nodes.data());
                        std::snprintf(tag, sizeof(tag), "grid%u/%u/%u n=%u kind=%d", dim, cell, layout, n, kind);
                        trace_all(nodes, root, tris, tag);
                        built++;
                    }
        }
    }
    // radix sorts on their own, with ragged sizes and every digit width used
    for (uint32_t n : {0u, 1u, 5u, 1000u}) {
        std::vector<uint32_t> k32(n), v(n);
        std::vector<uint64_t> k64(n);
        for (uint32_t i = 0; i < n; i++) {
            k32[i] = (uint32_t)(rng = rng * 6364136223846793005ull + 1);
            k64[i] = rng >> 1;
            v[i] = i;
        }
        for (uint32_t b : {4u, 8u, 11u, 16u}) {
            auto k = k32;
            auto w = v;
            sqo_radix_sort32(k.data(), w.data(), n, b);
            for (uint32_t i = 1; i < n; i++) CHECK(k[i - 1] <= k[i], "sort32 b=%u", b);
            auto k2 = k64;
            auto w2 = v;
            sqo_radix_sort64(k2.data(), w2.data(), n, b, 63);
            for (uint32_t i = 1; i < n; i++) CHECK(k2[i - 1] <= k2[i], "sort64 b=%u", b);
        }
    }
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("oracle sanitizer run: %u trees built and traced\n", built);
    return 0;
}
