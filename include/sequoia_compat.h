/*
 * sequoia_compat.h — the reference's C++ entry points, served by the MI355X
 * library. A caller written against the reference's headers (aabb.h,
 * builder.h, tracer.h) includes this header instead and relinks against
 * libsequoia_amd.so; the types keep the reference's layout and the functions
 * its signatures and mangled names, so objects compiled against the
 * reference's own headers (its main.cpp, for one) link against the library
 * unchanged:
 *
 *   Triangle, Node1, Hierarchy ............. aabb.h:17-24, 34-42, 197-205
 *   Builder::build_hierarchy{,_morton,...} . builder.h:81-104
 *   trace / trace_packet<K> / trace_bundle<K>  tracer.h:129-135
 *   radix_sort<Key, T>, bit_interleave_32/64 .. algorithms.h:89-116
 *
 * Differences a caller can observe:
 *   - Hierarchy::node_count and tri_count are filled in (the reference leaves
 *     them unset), and tri_base is set by every builder (the reference's
 *     horizontal, grid and grid_avx builders leave it dangling).
 *   - Morton hierarchies carry refitted node boxes (the reference writes
 *     none, so its Morton trees cannot be traced).
 *   - Failures throw std::runtime_error with the library's message (the
 *     reference asserts or exits).
 * node_base is allocated with new[]; release it with delete[] (or
 * sequoia_compat::release, which also drops the device copy).
 */
#ifndef SEQUOIA_COMPAT_H
#define SEQUOIA_COMPAT_H

#include <cstddef>
#include <cstdint>
#include <vector>

/* The reference's vec3f1 lives in namespace vec (vec3.h:10,130) and has a
 * user-provided copy constructor (vec3.h:143), which makes it non-trivial for
 * the purpose of calls: the Itanium C++ ABI passes such an argument as a
 * pointer to a caller-made copy, and mangles it as vec::vec3f1. This type
 * keeps both properties, so trace(..., vec3f1) and its templates link and
 * pass the origin exactly as in objects compiled against the reference's
 * tracer.h:129-135. */
namespace vec {
struct vec3f1 {
    float x, y, z;
    vec3f1() {}
    vec3f1(float a) : x(a), y(a), z(a) {}
    vec3f1(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    vec3f1(const vec3f1& a) : x(a.x), y(a.y), z(a.z) {}
    vec3f1& operator=(const vec3f1& o) {
        x = o.x;
        y = o.y;
        z = o.z;
        return *this;
    }
};
}  // namespace vec
using vec::vec3f1;

struct Triangle {
    vec3f1 v0, v1, v2;
    Triangle() {}
    Triangle(const vec3f1& a, const vec3f1& b, const vec3f1& c) : v0(a), v1(b), v2(c) {}
};

enum AABB_FLAGS { AABB_FLAGS_NONE = 0x00, AABB_FLAGS_LEAF = 0x01 };

struct Node1 {
    vec3f1 min;
    uint16_t count;
    uint16_t flags;
    vec3f1 max;
    uint32_t child;
    Node1() {}
};

struct PluckerTriangle;

struct Hierarchy {
    Node1* node_base;
    unsigned node_count;
    Triangle* tri_base;
    PluckerTriangle* plucker_tri_base;
    unsigned tri_count;
    unsigned root_index;
};

static_assert(sizeof(vec3f1) == 12, "vec3f1 layout");
static_assert(sizeof(Triangle) == 36, "Triangle layout");
static_assert(sizeof(Node1) == 32, "Node1 layout");

class Builder {
public:
    static Hierarchy build_hierarchy(std::vector<Triangle>& primitives);
    static Hierarchy build_hierarchy_morton(std::vector<Triangle>& primitives);
    static Hierarchy build_hierarchy_morton2(std::vector<Triangle>& primitives);
    static Hierarchy build_hierarchy_grid(std::vector<Triangle>& primitives);
    static Hierarchy build_hierarchy_grid_sse(std::vector<Triangle>& primitives);
    static Hierarchy build_hierarchy_grid_m128(std::vector<Triangle>& primitives);
    static Hierarchy build_hierarchy_horizontal(std::vector<Triangle>& primitives);
};

uint8_t* trace(Hierarchy& hierarchy, uint32_t w, uint32_t h, vec3f1 origin);
template <size_t K>
uint8_t* trace_packet(Hierarchy& hierarchy, uint32_t w, uint32_t h, vec3f1 origin);
template <size_t K>
uint8_t* trace_bundle(Hierarchy& hierarchy, uint32_t w, uint32_t h, vec3f1 origin);

/* algorithms.cpp:8-22: bit i of x, y, z placed at bits 3i, 3i+1, 3i+2 (bits
 * beyond the word dropped), as the reference's pdep masks place them. Host
 * functions. */
uint32_t bit_interleave_32(const uint32_t& x, const uint32_t& y, const uint32_t& z);
uint64_t bit_interleave_64(const uint64_t& x, const uint64_t& y, const uint64_t& z);

namespace sequoia_compat {
/* Device used by the shim (default 0); call before the first build. */
void set_device(int device);
/* Frees node_base and the device-resident copy of the hierarchy. */
void release(Hierarchy& h);
/* Stable LSD sort of (key, value) pairs in host memory on the GPU, by the
 * low key_bits bits in passes of digit_bits (1..8) bits. */
void sort_pairs(uint64_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, uint32_t digit_bits);
}  // namespace sequoia_compat

/* algorithms.h:89-113: stable LSD radix sort of [begin, end) by (Key)(*it),
 * four passes of b bits for a 32-bit Key and eight for a 64-bit one (so the
 * low 4b / 8b bits of the key order the elements), result in [begin, end).
 * Here the passes run on the GPU; the scratch range is not used and its
 * contents are left unchanged (the reference leaves intermediate passes
 * there). b must be 1..8 (the reference's counts[1<<b] table). */
template <typename Key, typename T>
static void radix_sort(typename std::vector<T>::iterator begin, typename std::vector<T>::iterator end,
                       typename std::vector<T>::iterator scratch, typename std::vector<T>::iterator scratch_end,
                       unsigned b) {
    (void)scratch;
    (void)scratch_end;
    static_assert(sizeof(Key) == 4 || sizeof(Key) == 8, "radix_sort: 32- or 64-bit keys");
    const size_t n = (size_t)(end - begin);
    if (n == 0) return;
    std::vector<uint64_t> keys(n);
    std::vector<uint32_t> order(n);
    std::vector<T> copy(begin, end);
    for (size_t i = 0; i < n; i++) {
        keys[i] = (uint64_t)(Key)copy[i];
        order[i] = (uint32_t)i;
    }
    const unsigned passes = sizeof(Key) == 4 ? 4u : 8u;
    sequoia_compat::sort_pairs(keys.data(), order.data(), (uint32_t)n, passes * b, b);
    for (size_t i = 0; i < n; i++) begin[i] = copy[order[i]];
}

#endif
