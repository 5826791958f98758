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

namespace sequoia_compat {
/* Device used by the shim (default 0); call before the first build. */
void set_device(int device);
/* Frees node_base and the device-resident copy of the hierarchy. */
void release(Hierarchy& h);
}  // namespace sequoia_compat

#endif
