// A caller of the reference's algorithms.h entry points (radix_sort<Key, T>,
// bit_interleave_32/64) built against sequoia_compat.h.
//
//   compat_algorithms interleave x y z           -> prints bit_interleave_32 and _64 of (x, y, z)
//   compat_algorithms sort32 <in> <out> <b>      -> in: n uint32 keys; out: n sorted keys, then n indices
//   compat_algorithms sort64 <in> <out> <b>      -> in: n uint64 keys; out: n sorted keys, then n indices (u32)
#include "sequoia_compat.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

// the key/value pairs of builder.cpp:2196-2246: code first, converts to its key
struct alignas(8) Pair32 {
    uint32_t code;
    uint32_t index;
    operator uint32_t() const { return code; }
};
struct alignas(16) Pair64 {
    uint64_t code;
    uint32_t index;
    operator uint64_t() const { return code; }
};

template <typename Key, typename P>
static int sort_file(const char* in, const char* out, unsigned b) {
    FILE* f = std::fopen(in, "rb");
    if (!f) return 2;
    std::vector<Key> keys;
    Key k;
    while (std::fread(&k, sizeof(Key), 1, f) == 1) keys.push_back(k);
    std::fclose(f);
    std::vector<P> v(keys.size()), scratch(keys.size());
    for (size_t i = 0; i < keys.size(); i++) {
        v[i].code = keys[i];
        v[i].index = (uint32_t)i;
    }
    radix_sort<Key, P>(v.begin(), v.end(), scratch.begin(), scratch.end(), b);
    FILE* o = std::fopen(out, "wb");
    if (!o) return 2;
    for (auto& p : v) std::fwrite(&p.code, sizeof(Key), 1, o);
    for (auto& p : v) std::fwrite(&p.index, 4, 1, o);
    std::fclose(o);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) return 1;
    std::string cmd = argv[1];
    try {
        if (cmd == "interleave" && argc == 5) {
            uint64_t x = std::strtoull(argv[2], nullptr, 0), y = std::strtoull(argv[3], nullptr, 0),
                     z = std::strtoull(argv[4], nullptr, 0);
            std::printf("%u %llu\n", bit_interleave_32((uint32_t)x, (uint32_t)y, (uint32_t)z),
                        (unsigned long long)bit_interleave_64(x, y, z));
            return 0;
        }
        if (cmd == "sort32" && argc == 5) return sort_file<uint32_t, Pair32>(argv[2], argv[3], std::atoi(argv[4]));
        if (cmd == "sort64" && argc == 5) return sort_file<uint64_t, Pair64>(argv[2], argv[3], std::atoi(argv[4]));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 3;
    }
    return 1;
}
