// A caller written against the reference's API (Builder::build_hierarchy_*,
// trace, trace_packet<K>, trace_bundle<K>), compiled against
// sequoia_compat.h and linked with libsequoia_amd.so.
//
//   compat_demo <tris.bin> <w> <h> <out_prefix>
// reads N x 9 float32 triangles, builds with every builder, traces with the
// run() camera (main.cpp:311-318) and writes <prefix>_<builder>.nodes and
// <prefix>_<mode>.rgb for the recursive tree.
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "sequoia_compat.h"

static void write_file(const std::string& path, const void* data, size_t bytes) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { perror(path.c_str()); exit(1); }
    fwrite(data, 1, bytes, f);
    fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 5) { fprintf(stderr, "usage: compat_demo tris.bin w h prefix\n"); return 2; }
    FILE* f = fopen(argv[1], "rb");
    if (!f) { perror(argv[1]); return 1; }
    fseek(f, 0, SEEK_END);
    long bytes = ftell(f);
    fseek(f, 0, SEEK_SET);
    std::vector<Triangle> prims(bytes / sizeof(Triangle));
    if (fread(prims.data(), sizeof(Triangle), prims.size(), f) != prims.size()) return 1;
    fclose(f);
    uint32_t w = atoi(argv[2]), h = atoi(argv[3]);
    std::string pre = argv[4];

    struct B { const char* name; Hierarchy (*fn)(std::vector<Triangle>&); };
    B builders[] = {{"recursive", Builder::build_hierarchy}, {"horizontal", Builder::build_hierarchy_horizontal},
                    {"grid", Builder::build_hierarchy_grid}, {"grid_sse", Builder::build_hierarchy_grid_sse},
                    {"grid_avx", Builder::build_hierarchy_grid_m128}, {"morton", Builder::build_hierarchy_morton},
                    {"morton2", Builder::build_hierarchy_morton2}};
    Hierarchy rec{};
    for (auto& b : builders) {
        Hierarchy hh = b.fn(prims);
        write_file(pre + "_" + b.name + ".nodes", hh.node_base, sizeof(Node1) * hh.node_count);
        printf("%s root %u nodes %u\n", b.name, hh.root_index, hh.node_count);
        if (std::string(b.name) == "recursive") rec = hh;
        else sequoia_compat::release(hh);
    }
    // run()'s camera: scene centre in x,y; z = 2*min.z - centre.z
    vec3f1 mn(FLT_MAX), mx(-FLT_MAX);
    for (auto& t : prims)
        for (const vec3f1* v : {&t.v0, &t.v1, &t.v2}) {
            mn.x = v->x < mn.x ? v->x : mn.x; mn.y = v->y < mn.y ? v->y : mn.y; mn.z = v->z < mn.z ? v->z : mn.z;
            mx.x = mx.x < v->x ? v->x : mx.x; mx.y = mx.y < v->y ? v->y : mx.y; mx.z = mx.z < v->z ? v->z : mx.z;
        }
    vec3f1 c((mn.x + mx.x) * 0.5f, (mn.y + mx.y) * 0.5f, (mn.z + mx.z) * 0.5f);
    vec3f1 cam(c.x, c.y, 2.0f * mn.z - c.z);
    uint8_t* img = trace(rec, w, h, cam);
    write_file(pre + "_basic.rgb", img, (size_t)w * h * 3);
    delete[] img;
    img = trace_packet<8>(rec, w, h, cam);
    write_file(pre + "_packet8.rgb", img, (size_t)w * h * 3);
    delete[] img;
    img = trace_bundle<64>(rec, w, h, cam);
    write_file(pre + "_bundle64.rgb", img, (size_t)w * h * 3);
    delete[] img;
    sequoia_compat::release(rec);
    printf("ok\n");
    return 0;
}
