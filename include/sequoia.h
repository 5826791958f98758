/*
 * sequoia.h — C ABI of the MI355X (gfx950) BVH build and ray-traversal library.
 *
 * Plain C types, host or device pointers and sizes; no framework types.
 * All functions return SQ_OK (0) or a negative SQ_ERR_* code; the message
 * of the last failure on the calling thread is available from
 * sq_last_error(). No function aborts the process.
 *
 * Reference interfaces replaced (paths relative to the reference's src/):
 *   Triangle / Node1 / Hierarchy ........ aabb.h:17-24, 34-42, 197-205
 *   Ray<1> / HitAttributes<1> ........... ray.h:8-14, 25-30
 *   Builder::build_hierarchy_* .......... builder.h:81-104 (called from main.cpp:258-277)
 *   trace / trace_packet / trace_bundle . tracer.h:129-135 (called from main.cpp:279-300)
 *   traverse ............................ tracer.h:118
 *   radix_sort / bit_interleave_* ....... algorithms.h:89-116
 * The C++ header sequoia_compat.h keeps the reference's own signatures on
 * top of this ABI.
 */
#ifndef SEQUOIA_H
#define SEQUOIA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define SQ_OK 0
#define SQ_ERR_INVALID (-1)     /* bad argument */
#define SQ_ERR_HIP (-2)         /* HIP runtime failure (no device, launch failure, ...) */
#define SQ_ERR_NOMEM (-3)       /* device allocation failed */
#define SQ_ERR_BUDGET (-4)      /* request exceeds the configured device-memory budget */
#define SQ_ERR_UNSUPPORTED (-5) /* option combination not implemented */
#define SQ_ERR_OVERFLOW (-6)    /* a traversal exceeded its stack bound; results invalid */
#define SQ_ERR_TIMEOUT (-7)     /* a served request expired in the queue, or a wait timed out */
#define SQ_ERR_CANCELLED (-8)   /* a served request was cancelled before it ran */

/* ---- data layouts (byte-identical to the reference) ---- */
typedef struct { float x, y, z; } sq_vec3;
/* Triangle (aabb.h:17-24): 36 B, AoS. */
typedef struct { sq_vec3 v0, v1, v2; } sq_triangle;
/* Node1 (aabb.h:34-42): 32 B. Leaf: flags=1, child = first triangle index,
 * count = triangles. Interior: flags=0, child = first of `count` consecutive
 * child nodes. */
typedef struct { sq_vec3 min; uint16_t count; uint16_t flags; sq_vec3 max; uint32_t child; } sq_node;
/* Ray<1> (ray.h:8-14): origin, tmin, direction, tmax; 32 B. */
typedef struct { sq_vec3 origin; float tmin; sq_vec3 dir; float tmax; } sq_ray;

#define SQ_NODE_LEAF 1u
#define SQ_MISS 0xFFFFFFFFu

typedef struct sq_ctx sq_ctx;     /* one device + one HIP stream + workspace */
typedef struct sq_scene sq_scene; /* device-resident triangles */
typedef struct sq_bvh sq_bvh;     /* device-resident Node1[] built over a scene */

/* ---- context ---- */
const char* sq_last_error(void);
const char* sq_version(void);
/* Creates a context on `device` with its own HIP stream (default priority). */
int sq_ctx_create(int device, sq_ctx** out);
void sq_ctx_destroy(sq_ctx* ctx);
int sq_ctx_sync(sq_ctx* ctx);
/* Opaque hipStream_t of the context (work of this library is ordered on it). */
void* sq_ctx_stream(sq_ctx* ctx);

/* Device times (ms) of the last build/trace on this context, measured with
 * HIP events recorded on the context stream. total_build_ms and trace_ms are
 * always measured; the build's phase times (bounds_ms .. grid_ms) only after
 * sq_ctx_set_phase_timing(ctx, 1), since every phase boundary's event leaves
 * the device idle for about 10 us. total_build_ms spans the build's device
 * work from its first kernel to the copy of its result record to the host
 * (the host's wait for that copy is not part of it). */
typedef struct {
    float bounds_ms, morton_ms, sort_ms, emit_ms, refit_ms, sah_ms, grid_ms, total_build_ms;
    float trace_ms;
    uint64_t node_tests, tri_tests; /* filled when the trace was asked to count work */
    uint32_t sort_passes, tree_depth, top_levels, small_tasks;
    uint32_t top_retries; /* always 0: no build phase waits on other workgroups or retries */
} sq_timings;
int sq_get_timings(sq_ctx* ctx, sq_timings* out);
/* 1: time each build phase separately from now on; 0 (the default): not. */
int sq_ctx_set_phase_timing(sq_ctx* ctx, int enable);

/* ---- device memory for callers without their own allocator ---- */
int sq_device_alloc(sq_ctx* ctx, size_t bytes, void** out);
void sq_device_free(sq_ctx* ctx, void* ptr);
/* Device-to-device copy by a plain streaming kernel (also the reference
 * kernel for calibrating hardware byte counters). */
int sq_device_copy(sq_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Copies device -> host on the context stream and waits. */
int sq_copy_to_host(sq_ctx* ctx, void* host_dst, const void* dev_src, size_t bytes);
/* Copies host -> device on the context stream and waits. */
int sq_copy_to_device(sq_ctx* ctx, void* dev_dst, const void* host_src, size_t bytes);

/* ---- scenes ---- */
int sq_scene_upload(sq_ctx* ctx, const sq_triangle* host_tris, uint32_t n, sq_scene** out);
/* Wraps caller-owned device memory (not freed by sq_scene_free). */
int sq_scene_wrap_device(sq_ctx* ctx, const sq_triangle* dev_tris, uint32_t n, sq_scene** out);
void sq_scene_free(sq_scene* s);
const sq_triangle* sq_scene_device_ptr(const sq_scene* s);
uint32_t sq_scene_count(const sq_scene* s);

/* ---- builders (builder.h:81-104) ---- */
typedef enum {
    SQ_BUILDER_RECURSIVE = 0,  /* build_hierarchy: binned SAH, Hoare partition, root 2n-1 */
    SQ_BUILDER_HORIZONTAL = 1, /* build_hierarchy_horizontal: same tree as RECURSIVE */
    SQ_BUILDER_GRID = 2,       /* build_hierarchy_grid: cell nodes carry the scene box */
    SQ_BUILDER_GRID_SSE = 3,   /* build_hierarchy_grid_sse: stable partition, tight cell boxes */
    SQ_BUILDER_GRID_AVX = 4,   /* build_hierarchy_grid_m128: same tree as GRID_SSE */
    SQ_BUILDER_MORTON = 5,     /* build_hierarchy_morton: 30-bit codes */
    SQ_BUILDER_MORTON2 = 6,    /* build_hierarchy_morton2: 60-bit codes */
    SQ_BUILDER_MORTON63 = 7    /* 63-bit codes (21 bits per axis) */
} sq_builder;

#define SQ_LAYOUT_REFERENCE 0u /* reproduce the reference slot allocator exactly */
#define SQ_LAYOUT_FIXED 1u     /* give a right singleton child its own slot */

typedef struct {
    uint32_t morton_bits_per_axis; /* 0 = builder default (10, 20 or 21) */
    uint32_t radix_digit_bits;     /* 0 = 8 */
    uint32_t num_bins;             /* 0 = 8 (reference NUM_BINS); 2..64 */
    uint32_t grid_dim;             /* 0 = 4 (reference GRID_DIM); power of two <= 8 */
    uint32_t layout;               /* SQ_LAYOUT_* */
    uint32_t no_refit;             /* Morton builders: 1 = leave node boxes zero like the reference */
    uint32_t subtree_threshold;    /* 0 = default (1024); largest range one workgroup finishes alone (8..2048) */
    uint32_t reserved;
} sq_build_params;

int sq_build(sq_ctx* ctx, const sq_scene* scene, sq_builder kind, const sq_build_params* params, sq_bvh** out);
/* Wraps a host-built Node1[] (uploaded) so it can be traced on the device. */
int sq_bvh_from_host(sq_ctx* ctx, const sq_scene* scene, const sq_node* nodes, uint32_t node_count, uint32_t root,
                     sq_bvh** out);
int sq_bvh_info(const sq_bvh* bvh, uint32_t* root, uint32_t* node_capacity, uint32_t* depth);
int sq_bvh_download(const sq_bvh* bvh, sq_node* host_nodes, uint32_t capacity);
const sq_node* sq_bvh_device_nodes(const sq_bvh* bvh);
void sq_bvh_free(sq_bvh* bvh);

/* ---- low-level kernels exported for tests and benchmarks ---- */
/* bit_interleave_32/64 (algorithms.cpp:8-22) + setup_morton quantisation
 * (builder.cpp:2204-2237): one code per triangle, written to host memory. */
int sq_morton_codes(sq_ctx* ctx, const sq_scene* scene, uint32_t bits_per_axis, uint64_t* host_codes);
/* Stable LSD radix sort of (key, value) pairs (algorithms.h:89-113) in host
 * memory; key_bits selects the number of passes. */
int sq_radix_sort_pairs(sq_ctx* ctx, uint64_t* host_keys, uint32_t* host_vals, uint32_t n, uint32_t key_bits,
                        uint32_t digit_bits);
/* Recomputes every node box of a device tree bottom-up. */
int sq_refit(sq_ctx* ctx, sq_bvh* bvh);

/* ---- tracing (tracer.h:129-135) ---- */
/* Pinhole camera. tan_half_fov = 1 reproduces the reference trace() camera
 * (tracer.cpp:202-216) bit-exactly; jitter = 1 replaces the pixel centre by
 * hashed sub-pixel offsets; ray (px, py, s) is ray index (py*width+px)*spp+s. */
typedef struct {
    sq_vec3 origin;
    float tan_half_fov;
    uint32_t width, height, spp, jitter, seed;
    float tmin, tmax;
} sq_camera;

typedef enum {
    SQ_TRACE_EXHAUSTIVE = 0, /* reference semantics: every hit box visited; counts comparable */
    SQ_TRACE_CLOSEST = 1     /* ordered traversal culling boxes beyond the closest hit */
} sq_semantics;
typedef enum {
    SQ_HEATMAP_RAY = 0,    /* trace(): per-ray node tests */
    SQ_HEATMAP_PACKET = 1, /* trace_packet<K>, K in {1,2,4,8} */
    SQ_HEATMAP_BUNDLE = 2  /* trace_bundle<K>, K in {16,64} */
} sq_heatmap;

#define SQ_KERNEL_AUTO 0u    /* per-lane rays (the measured faster one on primary rays) */
#define SQ_KERNEL_RAYS 1u    /* one ray per lane, per-lane stacks, dynamic refill */
#define SQ_KERNEL_PACKETS 2u /* one 64-ray packet per wave, union walk with lane masks */
typedef struct {
    uint32_t semantics;   /* sq_semantics */
    uint32_t heatmap;     /* sq_heatmap */
    uint32_t width;       /* packet/bundle K */
    uint32_t block_size;  /* 0 = default; multiple of 64 */
    uint32_t stack_slots; /* 0 = default; LDS stack entries per lane */
    uint32_t count_work;  /* 1 = accumulate total node/triangle tests into sq_timings */
    uint32_t kernel;      /* camera rays: SQ_KERNEL_* (packets: exhaustive semantics only) */
    uint32_t blocks_per_cu; /* per-ray kernel: 0 = default (one camera's rays: 12 waves per CU, 16 for
                               exhaustive Node2 steps; 16 otherwise; at most what fits); else 1..64 */
    uint32_t tri_test;    /* 0: reference Moller-Trumbore; 1: precomputed triangles (sq_bvh_plucker) */
    uint32_t lists_per_step; /* per-ray kernel, exhaustive semantics: child lists (Node1) or Node2s
                                tested per step, 0 = default (2), else 2 or 4; Node<K> trees (both
                                semantics): nodes per step, 0 = default, 1, or 2 for K = 4 */
} sq_trace_params;

/* Outputs (any may be NULL). hit_id/hit_t: one per ray (SQ_MISS / +inf on
 * miss). counts: one per pixel (camera) or per ray (ray list), node tests
 * summed over samples; with exhaustive semantics they are the reference's
 * counts, with closest semantics the box tests this traversal made (they
 * depend on the step: camera rays test two child lists per step, ray lists
 * one). rgb: width*height*3 heatmap bytes uint8(min(1, tests/200)*255)
 * (tracer.cpp:222-225). */
typedef struct {
    uint32_t* hit_id;
    float* hit_t;
    uint32_t* counts;
    uint8_t* rgb;
} sq_trace_out;

/* out_on_device = 0: out pointers are host memory; 1: device memory. */
int sq_trace_camera(sq_ctx* ctx, const sq_bvh* bvh, const sq_camera* cam, const sq_trace_params* params,
                    const sq_trace_out* out, int out_on_device);
/* rays: host (io_on_device = 0) or device (1) array of n rays. */
int sq_trace_rays(sq_ctx* ctx, const sq_bvh* bvh, const sq_ray* rays, uint32_t n, const sq_trace_params* params,
                  const sq_trace_out* out, int io_on_device);

/* ---- allocation-free tracing ----
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sq_trace_rays with device rays and device outputs
 * (io_on_device = 1) whose scratch (tile counter, fault bits, work totals,
 * per-lane stack spill areas) is one fixed array of SQ_NOALLOC_SCRATCH_BYTES
 * per device, set up by sq_ctx_create, instead of the context's workspace;
 * it calls no allocating routine. Node1 trees, ray heatmaps. The launch
 * takes as many workgroups as the spill area holds for the tree's depth (at
 * most sq_trace_rays' launch); a tree so deep that fewer than one workgroup
 * per CU fits fails with SQ_ERR_BUDGET. Every call on a device shares that
 * device's array, so calls that overlap on one device (from other contexts
 * or threads) race: the caller serialises them. Outputs equal
 * sq_trace_rays(..., io_on_device = 1). */
#define SQ_NOALLOC_SCRATCH_BYTES (64u << 20)
int sq_trace_rays_noalloc(sq_ctx* ctx, const sq_bvh* bvh, const sq_ray* dev_rays, uint32_t n,
                          const sq_trace_params* params, const sq_trace_out* dev_out);

/* ---- precomputed triangles (PluckerTriangle, aabb.h:182-241) ----
 * The reference declares the 48-byte record and leaves its conversion
 * commented out, with no intersection routine. Here it is the projection
 * method of that layout (Wald 2004, 7.1): the plane as P[w] + nu P[u] +
 * nv P[v] = np along the dominant normal axis w = ci, and the two edges
 * scaled by the projected determinant (e0u, e0v, e1u, e1v). A hit needs
 * beta, gamma >= 0, beta + gamma <= 1, t > 1e-9, tmin < t < tmax. It changes
 * hit arithmetic: t differs from Moller-Trumbore in the last bits and hits
 * within rounding of an edge may differ; box tests (counts) do not change. */
typedef struct {
    float nu, nv, np, pu, pv;
    int32_t ci;               /* dominant normal axis; 3 = degenerate, never hit */
    float e0u, e0v, e1u, e1v;
    int32_t pad0, pad1;
} sq_plucker;
/* Precomputes the records of the tree's scene on the device (kept with the
 * tree, as the reference's Hierarchy::plucker_tri_base; freed with it);
 * *out receives their device address (may be NULL). Traces with
 * sq_trace_params.tri_test = 1 then use them (per-ray kernel, Node1 trees). */
int sq_bvh_plucker(sq_ctx* ctx, sq_bvh* bvh, const sq_plucker** out);

/* ---- quantised two-child nodes (Node2, aabb.h:44-74) ----
 * The reference declares Node2 (a float centre, 24-bit child fields, an 8-bit
 * size, 3-bit counts, leaf bits, twelve 8-bit bounds) and never uses it; here
 * it is a traversal format converted from any Node1 tree. Each Node2 holds
 * the two children of one node: a leaf group of 1..7 triangles (contiguous in
 * a reordered triangle array) or another Node2. Bound q of a child decodes to
 * c + (q - 128) * 2^(size - 127) and rounds outwards, so the decoded boxes
 * contain the Node1 boxes: closest hits equal the Node1 tree's, while a child
 * pair costs one 32-byte fetch. Counts are box tests of this tree (the root
 * box, then two per visited Node2), not the reference's Node1 counts. */
typedef struct {
    float c[3];             /* centre */
    uint32_t a_child_size;  /* a_child (24) | size (8) */
    uint32_t b_child_bits;  /* b_child (24) | a_count (3) | a_leaf (1) | b_count (3) | b_leaf (1) */
    uint8_t q[12];          /* a_min_x a_max_x b_min_x b_max_x, then y, then z */
} sq_node2;
typedef struct {
    float min[3], max[3];   /* root box, tested unquantised */
    uint32_t kind;          /* 1: leaf group (child = first triangle, count), 2: Node2 `child` */
    uint32_t child, count, reserved;
} sq_node2_root;
typedef struct sq_bvh2 sq_bvh2;
/* Converts a device Node1 tree (needs leaves of at most 7 triangles, as every
 * builder here makes; breadth-first, deterministic). A list of several Node1
 * leaves becomes one leaf group when it holds at most max_leaf triangles
 * (1..7; 0 = 1, measured fastest), else a subtree of Node2s over them. */
int sq_bvh2_create(sq_ctx* ctx, const sq_bvh* bvh, uint32_t max_leaf, sq_bvh2** out);
int sq_bvh2_info(const sq_bvh2* t, uint32_t* nodes, uint32_t* tris, uint32_t* levels, sq_node2_root* root);
/* Host copies of the nodes, the reordered triangles and their original indices (any may be NULL). */
int sq_bvh2_download(const sq_bvh2* t, sq_node2* nodes, uint32_t node_cap, sq_triangle* tris, uint32_t* perm,
                     uint32_t tri_cap);
void sq_bvh2_free(sq_bvh2* t);
/* As sq_trace_camera / sq_trace_rays (ray heatmaps, per-ray kernel); hit ids
 * are original triangle indices. */
int sq_trace2_camera(sq_ctx* ctx, const sq_bvh2* t, const sq_camera* cam, const sq_trace_params* params,
                     const sq_trace_out* out, int out_on_device);
int sq_trace2_rays(sq_ctx* ctx, const sq_bvh2* t, const sq_ray* rays, uint32_t n, const sq_trace_params* params,
                   const sq_trace_out* out, int io_on_device);

/* ---- wide nodes (Node<K>, aabb.h:26-32) ----
 * The reference declares template <size_t K> struct Node (vec3f<K> min,
 * simdi<K> flags, vec3f<K> max, simdi<K> child: structure-of-arrays, 32 K
 * bytes) and never uses it; here it is a traversal format converted from any
 * Node1 tree, for K = 4 (128 B) and K = 8 (256 B). Slot j of a node is one
 * child: a Node1 node (its box as stored there) or a run of a longer Node1
 * list (the union box); flags[j] is in Node1's count | flags << 16 order: a
 * leaf slot copies the Node1 leaf (bit 16 set, count triangles from child[j],
 * the Node1 tree's triangle ids), an interior slot names wide node child[j]
 * and holds how many of its slots are used (used slots first; 0 with child
 * 0xFFFFFFFF for a Node1 node with an empty list). Unused slots: min +inf,
 * max -inf, flags 0, child 0xFFFFFFFF. Conversion pulls grandchildren into
 * free slots, largest half surface area first (SQ_BVHK_NO_COLLAPSE: one Node1
 * list per node, so a tree without lists longer than K is traced with exactly
 * the Node1 tree's box tests and counts); nodes are numbered breadth first.
 * Counts are box tests of this tree: the root box, then one per used slot of
 * each visited node. */
typedef struct {
    float min_x[4], min_y[4], min_z[4];
    uint32_t flags[4];
    float max_x[4], max_y[4], max_z[4];
    uint32_t child[4];
} sq_node4;
typedef struct {
    float min_x[8], min_y[8], min_z[8];
    uint32_t flags[8];
    float max_x[8], max_y[8], max_z[8];
    uint32_t child[8];
} sq_node8;
typedef struct {
    float min[3], max[3];   /* root box, tested as stored */
    uint32_t kind;          /* 0: empty list, 1: leaf (child = first triangle, count), 2: node `child`, count used slots */
    uint32_t child, count, reserved;
} sq_nodek_root;
typedef struct sq_bvhk sq_bvhk;
enum { SQ_BVHK_NO_COLLAPSE = 1 };
/* Converts a device Node1 tree into Node<k> (k = 4 or 8) on the device.
 * flags: 0 or SQ_BVHK_NO_COLLAPSE. The tree keeps using the Node1 tree's
 * scene (its triangles are not copied). */
int sq_bvhk_create(sq_ctx* ctx, const sq_bvh* bvh, uint32_t k, uint32_t flags, sq_bvhk** out);
int sq_bvhk_info(const sq_bvhk* t, uint32_t* k, uint32_t* nodes, uint32_t* levels, sq_nodek_root* root);
/* Host copy of the nodes (32 k bytes each; node_cap in nodes). */
int sq_bvhk_download(const sq_bvhk* t, void* nodes, uint32_t node_cap);
void sq_bvhk_free(sq_bvhk* t);
/* As sq_trace_camera / sq_trace_rays (ray heatmaps, per-ray kernel). */
int sq_tracek_camera(sq_ctx* ctx, const sq_bvhk* t, const sq_camera* cam, const sq_trace_params* params,
                     const sq_trace_out* out, int out_on_device);
int sq_tracek_rays(sq_ctx* ctx, const sq_bvhk* t, const sq_ray* rays, uint32_t n, const sq_trace_params* params,
                   const sq_trace_out* out, int io_on_device);

/* ---- batches: many scenes built and traced together ----
 * No counterpart in the reference, which builds and traces one Hierarchy per
 * call (Builder::build_hierarchy_morton, builder.cpp:2474-2508; trace,
 * tracer.cpp:195-230); every scene's tree and outputs equal what those calls
 * produce for it alone (and what sq_build / sq_trace_camera produce).
 *
 * Triangles of all scenes live in one device buffer (scene s = counts[s]
 * consecutive triangles). A build runs the Morton pipeline once over the
 * concatenation (per-scene bounds and codes, one radix sort by (scene, code),
 * per-scene emission and refit in shared launches). A trace runs one
 * persistent launch whose waves steal 8x8-pixel tiles of every scene. */
typedef struct sq_batch sq_batch;

int sq_batch_create(sq_ctx* ctx, const sq_triangle* tris, const uint32_t* counts, uint32_t nscenes,
                    int tris_on_device, sq_batch** out);
void sq_batch_free(sq_batch* b);
uint32_t sq_batch_count(const sq_batch* b);

/* Morton builders only (MORTON, MORTON2, MORTON63); at most 2048 scenes. */
int sq_batch_build(sq_ctx* ctx, sq_batch* b, sq_builder kind, const sq_build_params* params);
/* Scene s's tree, owned by the batch (sq_bvh_free ignores it); usable with
 * sq_bvh_info / sq_bvh_download / sq_trace_camera / sq_trace_rays. */
int sq_batch_bvh(sq_batch* b, uint32_t scene, const sq_bvh** out);

#define SQ_OUT_HIT_ID 1u
#define SQ_OUT_HIT_T 2u
#define SQ_OUT_COUNTS 4u
#define SQ_OUT_RGB 8u
/* One camera per scene; allocates batch-owned device outputs (mask of
 * SQ_OUT_*). params: semantics, block_size, stack_slots, count_work. */
int sq_batch_set_cameras(sq_ctx* ctx, sq_batch* b, const sq_camera* cams, const sq_trace_params* params,
                         uint32_t outputs);
/* Traces every scene's camera in one launch (needs a build and cameras). */
int sq_batch_trace(sq_ctx* ctx, sq_batch* b);
/* Build + trace as one HIP graph: captured on the first call (or after the
 * builder, parameters or cameras change), replayed afterwards. */
int sq_batch_run(sq_ctx* ctx, sq_batch* b, sq_builder kind, const sq_build_params* params);
/* Device pointers of scene s's outputs (NULL where not requested). */
int sq_batch_outputs(const sq_batch* b, uint32_t scene, sq_trace_out* device_out);
/* Copies scene s's outputs to host arrays (any may be NULL). */
int sq_batch_download(sq_ctx* ctx, const sq_batch* b, uint32_t scene, const sq_trace_out* host_out);
/* Device time (ms) of the last sq_batch_build / sq_batch_trace / sq_batch_run. */
int sq_batch_times(const sq_batch* b, float* build_ms, float* trace_ms, float* run_ms);

/* Autotuner: times candidate radix digit widths (4..8) and subtree
 * thresholds (512..2048) for a Morton sq_batch_build, or bin counts 8, 16
 * and 32 for a binned-SAH one (recursive / horizontal: the bin count shapes
 * the trees, so each candidate is built and traced with the default trace
 * and the fastest build + trace kept), and the traversal kernel
 * (per-lane rays with stack slots 8..32, launch shapes of 8 to 24 waves
 * per CU in workgroups of 64 or 128 and 2 or 4 child lists per step, or wave
 * packets for exhaustive semantics) for sq_batch_trace, on this batch (which must have cameras
 * set), keeps the fastest in the batch and reports them. Apart from the bin
 * count, none of them changes any result; with it, results equal the
 * single-scene builds with that bin count. */
typedef struct {
    uint32_t radix_digit_bits, subtree_threshold, stack_slots, block_size;
    float build_ms, trace_ms;           /* best times */
    float default_build_ms, default_trace_ms;
    uint32_t candidates_tried;
    uint32_t kernel;                    /* SQ_KERNEL_RAYS or SQ_KERNEL_PACKETS */
    uint32_t blocks_per_cu;             /* 0 = the default launch shape */
    uint32_t lists_per_step;            /* per-ray kernel: 2 or 4 */
    uint32_t num_bins;                  /* binned-SAH batches: the chosen bin count; 0 otherwise */
} sq_tune_result;
int sq_batch_autotune(sq_ctx* ctx, sq_batch* b, sq_builder kind, uint32_t repeats, sq_tune_result* out);

/* ---- serving: a continuous-batching request loop ----
 * No counterpart in the reference (its main.cpp runs one build and one trace
 * per process). A server owns one device, a scheduler thread and a pool of
 * device memory. Requests (build + trace of a new scene, or trace of a
 * resident scene) queue in arrival order; whenever one of the in-flight wave
 * slots is free, the scheduler admits every queued request that fits the
 * wave limits and the memory budget into the next wave: the wave's new
 * scenes are built together (the batched Morton pipeline) and all its traces
 * run in one persistent launch, on the slot's own stream, overlapping the
 * other slot's wave. Built trees stay resident in HBM under their key for
 * later trace-only requests; unreferenced scenes are evicted LRU when the
 * budget needs room. Every result equals the single-scene calls' on the same
 * input (morton63 tree, sq_trace_camera / sq_trace_rays). */
typedef struct sq_server sq_server;

typedef struct {
    uint64_t budget_bytes;      /* device bytes for resident scenes + in-flight waves; 0 = 200 GB */
    uint32_t max_wave_requests; /* 0 = 256 */
    uint32_t max_wave_rays;     /* 0 = 16 Mi (a single larger request still forms its own wave) */
    uint32_t max_wave_tris;     /* 0 = 8 Mi */
    uint32_t semantics;         /* sq_semantics of every trace */
    uint32_t inflight_waves;    /* wave slots on the device at once, 1 or 2; 0 = 2 */
    uint32_t reserved[3];       /* [0]: MiB of pinned input and output staging per slot, allocated at
                                   creation (0 = 256); larger waves grow it. [1]: GiB of device pool
                                   reserved at creation (0 = 16, at most the budget); it grows as needed */
} sq_server_config;

/* A request. Host buffers (tris, rays, out) must stay valid until it completes. */
typedef struct {
    uint64_t key;              /* scene key */
    const sq_triangle* tris;   /* non-NULL: build this scene under `key`; NULL: trace resident scene `key` */
    uint32_t ntris;
    uint32_t nrays;            /* > 0: trace the ray list `rays`; 0: trace `camera` */
    const sq_ray* rays;
    sq_camera camera;
    sq_trace_out out;          /* destinations, any may be NULL (counts: per pixel / per listed ray) */
    float timeout_ms;          /* > 0: fail with SQ_ERR_TIMEOUT if not admitted within this time */
    uint32_t flags;            /* SQ_REQ_DEVICE_OUT: `out` holds device pointers */
} sq_request;

/* The request's outputs are device memory of the server's device: the trace
 * writes them in place (no staging, no device-to-host copy); they hold the
 * results once the request is DONE. Requests in flight must not share them. */
#define SQ_REQ_DEVICE_OUT 1u

#define SQ_REQ_QUEUED 0
#define SQ_REQ_RUNNING 1
#define SQ_REQ_DONE 2
#define SQ_REQ_FAILED 3

typedef struct {
    int32_t state;      /* SQ_REQ_* */
    int32_t status;     /* SQ_OK or the failure (BUDGET: larger than the budget; TIMEOUT; CANCELLED;
                           INVALID: trace of a scene that is not resident) */
    float latency_ms;   /* submit -> outputs in host memory */
    float queue_ms;     /* submit -> admitted into a wave */
    uint64_t wave;
    uint32_t fault;     /* status OVERFLOW: 1 stack bound, 2 node index, 4 triangle index */
    uint32_t reserved;
} sq_request_status;

typedef struct {
    uint64_t submitted, completed, failed, waves, evictions;
    uint64_t resident_scenes, resident_bytes, inflight_bytes, pool_reserved_bytes;
    uint64_t rays_traced, tris_built;
    float p50_ms, p99_ms, mean_ms, max_ms; /* latency over completed requests */
    float span_s;                          /* first submit -> last completion */
    float mean_wave_requests;
    float stage_in_ms, stage_out_ms;       /* host time copying inputs to / outputs from pinned staging */
    float wave_ms;                         /* summed enqueue -> retire time of the waves */
} sq_server_stats;

int sq_server_create(int device, const sq_server_config* cfg, sq_server** out);
/* Finishes in-flight waves, fails queued requests (CANCELLED), frees everything. */
void sq_server_destroy(sq_server* s);
int sq_server_submit(sq_server* s, const sq_request* r, uint64_t* id);
/* Waits up to timeout_ms (< 0: forever) for request `id`; SQ_ERR_TIMEOUT if still running. */
int sq_server_wait(sq_server* s, uint64_t id, float timeout_ms, sq_request_status* st);
/* Cancels a queued request (SQ_ERR_INVALID once it is in a wave or finished). */
int sq_server_cancel(sq_server* s, uint64_t id);
/* Waits until every submitted request has finished. */
int sq_server_drain(sq_server* s);
/* Drops resident scene `key` (SQ_ERR_INVALID if unknown or in use). */
int sq_server_evict(sq_server* s, uint64_t key);
int sq_server_get_stats(sq_server* s, sq_server_stats* out);
int sq_server_reset_stats(sq_server* s);

/* ---- deterministic workload generators (host) ---- */
uint32_t sq_gen_icosphere_count(uint32_t target);
int sq_gen_icosphere(uint32_t target, uint32_t seed, float amplitude, sq_triangle* out);
int sq_gen_soup(uint32_t n, uint32_t seed, float sigma, sq_triangle* out);
int sq_gen_clusters(uint32_t n, uint32_t clusters, uint32_t seed, sq_triangle* out);
int sq_gen_secondary_rays(const sq_triangle* tris, uint32_t ntri, uint32_t nrays, uint32_t seed, float tmin,
                          sq_ray* out);
/* run()'s camera (main.cpp:311-318). */
int sq_reference_camera(const sq_triangle* tris, uint32_t n, uint32_t w, uint32_t h, sq_camera* out);
int sq_pinhole_camera(const sq_triangle* tris, uint32_t n, uint32_t w, uint32_t h, float fov_deg, float distance,
                      uint32_t spp, uint32_t jitter, uint32_t seed, sq_camera* out);

/* ---- scene files (host) ----
 * Replaces loadOBJFromFile (main.cpp:108-179): the same triangles for every
 * file that loader reads ('v' and 'f' records, quads fanned, 1-based or
 * negative indices, 255-byte line pieces, the reference's tokenizer);
 * inputs it cannot read get defined results (a last token without a trailing
 * delimiter is read, an empty fourth index is ignored) or SQ_ERR_INVALID (an
 * index that names no vertex, a record with too few fields, a missing file).
 * *out is allocated with malloc: release it with sq_host_free. nverts may be
 * NULL. Nothing is printed. */
int sq_load_obj(const char* path, sq_triangle** out, uint32_t* ntris, uint32_t* nverts);
void sq_host_free(void* p);
/* Replaces stbi_flip_vertically_on_write(true) + stbi_write_bmp(name, w, h,
 * 3, img) (main.cpp:335-336): a 24-bit BMP whose first stored row is memory
 * row 0 (the bottom of a heatmap), BGR, rows padded to 4 bytes; the same
 * bytes as the reference's file. rgb is h rows of w RGB8 pixels. */
size_t sq_bmp_size(uint32_t w, uint32_t h);
int sq_encode_bmp(const uint8_t* rgb, uint32_t w, uint32_t h, uint8_t* out, size_t cap);
int sq_write_bmp(const char* path, const uint8_t* rgb, uint32_t w, uint32_t h);

#ifdef __cplusplus
}
#endif
#endif
