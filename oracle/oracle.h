/*
 * oracle.h — CPU restatement of Sequoia's BVH build and ray traversal.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests, the smoke
 * entry and the bench's CPU-baseline leg may call it; the product library
 * (libsequoia_amd.so) never links, loads or calls it.
 *
 * Every function restates a reference function (file:line under the
 * reference's src/). Floating-point expressions mirror the fused
 * multiply-add pattern that g++ 11.4 -O3 (x86-64-v3) chose for the same
 * source expression; each site names the pattern it follows. The oracle
 * is compiled with -ffp-contract=off so that only the explicit fmaf calls
 * fuse.
 *
 * Parity pinning: see tests/test_oracle_*.py. Every function below is
 * checked against golden vectors written by the compiled reference
 * (tests/golden/), against closed forms or against brute force.
 */
#ifndef SEQUOIA_ORACLE_H
#define SEQUOIA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z; } sqo_vec3;
typedef struct { sqo_vec3 v0, v1, v2; } sqo_tri;                 /* 36 B, reference Triangle */
typedef struct {                                                  /* 32 B, reference Node1    */
    float min[3]; uint16_t count; uint16_t flags; float max[3]; uint32_t child;
} sqo_node;
typedef struct { float ox, oy, oz, tmin, dx, dy, dz, tmax; } sqo_ray; /* 32 B, Ray<1> layout */

enum { SQO_LAYOUT_REFERENCE = 0, SQO_LAYOUT_FIXED = 1 };
enum { SQO_PART_HOARE = 0, SQO_PART_STABLE = 1 };
enum { SQO_SA_SCALAR = 0, SQO_SA_SSE = 1, SQO_SA_M128 = 2 };
enum { SQO_CELLBOX_SCENE = 0, SQO_CELLBOX_TIGHT = 1 };

/* ---- Morton codes and sorting (builder.cpp:2204-2279, algorithms.*) ---- */
uint32_t sqo_interleave32(uint32_t x, uint32_t y, uint32_t z);
uint64_t sqo_interleave64(uint64_t x, uint64_t y, uint64_t z);
/* bits_per_axis: 10 (30-bit codes), 20 (60-bit), 21 (63-bit). */
void sqo_scene_bounds(const sqo_tri* tris, uint32_t n, float out_min[3], float out_max[3]);
void sqo_morton_codes(const sqo_tri* tris, uint32_t n, uint32_t bits_per_axis, uint64_t* codes);
void sqo_radix_sort32(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t digit_bits);
void sqo_radix_sort64(uint64_t* keys, uint32_t* vals, uint32_t n, uint32_t digit_bits, uint32_t key_bits);

/* ---- Morton / LBVH emission (builder.cpp:2281-2409) + refit (new) ---- */
uint32_t sqo_find_split(const uint64_t* codes, uint32_t count);
/* Writes nodes[0 .. 2n) (zeroed first); returns the root index (0). */
uint32_t sqo_emit_morton(const uint64_t* sorted_codes, const uint32_t* sorted_idx, uint32_t n,
                         uint32_t layout, sqo_node* nodes);
/* Bottom-up AABB refit of every node reachable from root (+ orphan leaves). */
void sqo_refit(sqo_node* nodes, uint32_t node_count, uint32_t root, const sqo_tri* tris);
/* Full Morton build: codes, stable sort, emission, optional refit. */
uint32_t sqo_build_morton(const sqo_tri* tris, uint32_t n, uint32_t bits_per_axis, uint32_t layout,
                          uint32_t refit, sqo_node* nodes, uint64_t* out_sorted_codes,
                          uint32_t* out_sorted_idx);

/* ---- Binned SAH (builder.cpp:114-147,334-628,1053-1156,1365-1663) ---- */
/* nodes must hold 2n entries; returns the root index (2n-1). */
uint32_t sqo_build_sah(const sqo_tri* tris, uint32_t n, uint32_t num_bins, uint32_t partition,
                       uint32_t sa_kind, uint32_t layout, sqo_node* nodes);
/* ---- Grid (builder.cpp:27-35,1665-2194); nodes hold 2n + 2*dim^3 ---- */
uint32_t sqo_build_grid(const sqo_tri* tris, uint32_t n, uint32_t grid_dim, uint32_t num_bins,
                        uint32_t partition, uint32_t sa_kind, uint32_t cell_box, uint32_t layout,
                        sqo_node* nodes);

/* ---- Rays (tracer.cpp:195-216 generalised) ---- */
typedef struct {
    float origin[3];
    float tan_half_fov;   /* 1.0 reproduces the reference camera exactly */
    uint32_t width, height, spp;
    uint32_t jitter;      /* 0: pixel centre (+0.5); 1: hashed sub-pixel offsets */
    uint32_t seed;
    float tmin, tmax;
} sqo_camera;
void sqo_camera_ray(const sqo_camera* cam, uint32_t px, uint32_t py, uint32_t sample, sqo_ray* out);
void sqo_camera_rays(const sqo_camera* cam, sqo_ray* out); /* ray r = (py*w + px)*spp + s */

/* ---- Traversal (tracer.h:13-35,83-116; tracer.cpp:47-65) ---- */
int sqo_slab(const sqo_node* node, const sqo_ray* ray);
/* Reference Möller–Trumbore mask; *t_out receives t. */
int sqo_tri_test(const sqo_tri* tri, const sqo_ray* ray, float* t_out);
/* Exhaustive (reference) traversal. counts: node tests per ray; hit_id/hit_t:
 * closest valid hit (0xFFFFFFFF / +inf if none); any_hit: reference hit mask. */
void sqo_trace_rays(const sqo_node* nodes, uint32_t root, const sqo_tri* tris, const sqo_ray* rays,
                    uint32_t nrays, uint32_t* counts, uint32_t* hit_id, float* hit_t, uint32_t* any_hit);
/* Closest-hit traversal with t culling (counts not comparable to the exhaustive mode). */
void sqo_trace_rays_closest(const sqo_node* nodes, uint32_t root, const sqo_tri* tris,
                            const sqo_ray* rays, uint32_t nrays, uint32_t* hit_id, float* hit_t);
/* Camera trace: per-pixel summed counts and RGB8 heatmap (tracer.cpp:195-230). */
void sqo_trace_camera(const sqo_node* nodes, uint32_t root, const sqo_tri* tris, const sqo_camera* cam,
                      uint32_t* pixel_counts, uint8_t* rgb, uint32_t* hit_id, float* hit_t);
/* Packet heatmap, K in {1,2,4,8} (tracer.cpp:88-125,232-281). */
void sqo_trace_packet(const sqo_node* nodes, uint32_t root, const sqo_tri* tris, const sqo_camera* cam,
                      uint32_t K, uint32_t* packet_counts, uint8_t* rgb);
/* Bundle heatmap, K in {16,64} (tracer.h:39-80, tracer.cpp:128-156,292-362). */
void sqo_trace_bundle(const sqo_node* nodes, uint32_t root, const sqo_tri* tris, const sqo_camera* cam,
                      uint32_t K, uint8_t* rgb);
uint8_t sqo_heat_byte(uint32_t count, uint32_t spp);

/* ---- Quantised two-child nodes (Node2, aabb.h:44-74; encoding defined in node2.cpp) ---- */
typedef struct { float c[3]; uint32_t w0, w1; uint8_t q[12]; } sqo_node2;          /* 32 B */
typedef struct { float min[3], max[3]; uint32_t kind, child, count, pad; } sqo_node2_root;
/* kind: 1 = the root is a leaf group (child = first triangle, count), 2 = Node2 `child`.
 * Returns 0, or < 0 for an unsupported tree (leaf of > 7 triangles, index out
 * of range, non-finite box, capacity). */
int sqo_node2_convert(const sqo_node* nodes, uint32_t node_count, uint32_t root, const sqo_tri* tris,
                      uint32_t ntris, uint32_t max_leaf, sqo_node2* out, uint32_t node_cap, uint32_t* out_nodes,
                      sqo_tri* tris2, uint32_t* perm, uint32_t tri_cap, uint32_t* out_tris,
                      sqo_node2_root* root_out);
void sqo_node2_decode(const sqo_node2* n, float a_min[3], float a_max[3], float b_min[3], float b_max[3]);
/* counts: box tests per ray (root + two per visited Node2); hits report original triangle ids. */
void sqo_node2_trace(const sqo_node2* nodes, const sqo_node2_root* root, const sqo_tri* tris2, const uint32_t* perm,
                     const sqo_ray* rays, uint32_t nrays, uint32_t closest, uint32_t* counts, uint32_t* hit_id,
                     float* hit_t);

/* ---- Wide nodes (Node<K>, aabb.h:26-32; meaning defined in nodek.cpp) ----
 * A Node<K> is 8 K 32-bit words: min x[K], min y[K], min z[K], flags[K],
 * max x[K], max y[K], max z[K], child[K]. K = 2, 4 or 8. */
typedef struct { float min[3], max[3]; uint32_t kind, child, count, pad; } sqo_nodek_root;
/* kind: 1 = the root is a leaf (child = first triangle, count), 2 = wide node `child` with `count` used slots,
 * 0 = an interior root with an empty list.
 * expand: 1 = pull grandchildren up (largest half surface area first) while they fit, 0 = one Node1 child
 * list per wide node (lists longer than K still become K runs). Returns 0, or < 0: -1 malformed tree,
 * -3 bad K, -4 node_cap too small. */
int sqo_nodek_convert(const sqo_node* nodes, uint32_t node_count, uint32_t root, uint32_t K, uint32_t expand,
                      uint32_t* out, uint32_t node_cap, uint32_t* out_nodes, uint32_t* out_levels,
                      sqo_nodek_root* root_out);
/* counts: box tests per ray (root + one per used slot of each visited node); hits use the Node1 tree's ids. */
void sqo_nodek_trace(const uint32_t* nodes, uint32_t K, const sqo_nodek_root* root, const sqo_tri* tris,
                     const sqo_ray* rays, uint32_t nrays, uint32_t closest, uint32_t* counts, uint32_t* hit_id,
                     float* hit_t);

/* ---- Precomputed-triangle (PluckerTriangle, aabb.h:182-195) test; plucker.cpp ---- */
typedef struct {
    float nu, nv, np, pu, pv;
    int32_t ci;                 /* dominant normal axis, 3 = never hit */
    float e0u, e0v, e1u, e1v;
    int32_t pad0, pad1;
} sqo_plucker;                  /* 48 B */
void sqo_plucker_precompute(const sqo_tri* tris, uint32_t n, sqo_plucker* out);
int sqo_plucker_test(const sqo_plucker* p, const sqo_ray* r, float* t_out);
void sqo_trace_rays_plucker(const sqo_node* nodes, uint32_t root, const sqo_plucker* pl, const sqo_ray* rays,
                            uint32_t nrays, uint32_t* counts, uint32_t* hit_id, float* hit_t);

/* Sizing helpers. */
uint32_t sqo_sah_node_capacity(uint32_t n);
uint32_t sqo_grid_node_capacity(uint32_t n, uint32_t grid_dim);

#ifdef __cplusplus
}
#endif
#endif
