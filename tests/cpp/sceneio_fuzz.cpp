// CPU fuzz of the scene-file code (sequoia_amd/csrc/sceneio.cpp): seeded
// mutations of an OBJ grammar file (byte flips, cuts, inserted tokens and
// record pieces, long lines past the 255-byte read, large and negative
// indices, empty and unterminated files) through sq_load_obj, and random
// image sizes and buffer capacities through sq_encode_bmp / sq_write_bmp.
// Built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitizers.py; any memory error or undefined behaviour aborts.
//
//   sceneio_fuzz <seed.obj> <work dir> <iterations>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <cstdarg>

#include "sequoia.h"

// The library's error slot lives in its context code; this driver links
// sceneio.cpp alone, so the message is formatted into a local buffer.
namespace sq {
static char last_error[512];
void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(last_error, sizeof(last_error), fmt, ap);
    va_end(ap);
}
}  // namespace sq

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint32_t rnd() {
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (uint32_t)(rng_state >> 11);
}
static uint32_t rnd(uint32_t n) { return n ? rnd() % n : 0; }

static const char* kPieces[] = {"v ", "f ", "vn ", "vt ", "# ", "\n", "\r\n", "\t", " ", "/", "//", "-",
                                "1", "-1", "2147483647", "-2147483648", "99999999999", "1e38", "-1e-38", "nan",
                                "inf", "0.5", "3/4/5", "f 1 2 3 4 5\n", "v 1 2 3\n", "f -1 -2 -3\n", "\n\n"};

static std::string mutate(const std::string& seed) {
    std::string s = seed;
    const uint32_t ops = 1 + rnd(8);
    for (uint32_t o = 0; o < ops; o++) {
        const uint32_t pos = s.empty() ? 0 : rnd((uint32_t)s.size() + 1);
        switch (rnd(7)) {
            case 0:  // flip a byte
                if (!s.empty()) s[rnd((uint32_t)s.size())] = (char)rnd(256);
                break;
            case 1:  // cut
                s.resize(pos);
                break;
            case 2:  // insert a grammar piece
                s.insert(pos, kPieces[rnd(sizeof(kPieces) / sizeof(kPieces[0]))]);
                break;
            case 3: {  // a line longer than one 255-byte read
                std::string line = rnd(2) ? "f" : "v";
                const uint32_t len = 200 + rnd(400);
                while (line.size() < len) line += " " + std::to_string((int)rnd(20) - 5);
                s.insert(pos, line + (rnd(2) ? "\n" : ""));
                break;
            }
            case 4:  // delete a range
                if (pos < s.size()) s.erase(pos, 1 + rnd(16));
                break;
            case 5:  // duplicate a range
                if (pos < s.size()) s.insert(pos, s.substr(pos, 1 + rnd(64)));
                break;
            default:  // drop the final newline
                if (!s.empty() && s.back() == '\n') s.pop_back();
                break;
        }
    }
    return s;
}

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::string seed;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) seed.append(buf, n);
    std::fclose(f);
    const std::string path = std::string(argv[2]) + "/fuzz.obj";
    const long iters = std::atol(argv[3]);
    long loaded = 0, rejected = 0;
    for (long it = 0; it < iters; it++) {
        std::string text = it == 0 ? std::string() : (it == 1 ? seed : mutate(seed));
        FILE* o = std::fopen(path.c_str(), "wb");
        if (!o) return 2;
        std::fwrite(text.data(), 1, text.size(), o);
        std::fclose(o);
        sq_triangle* tris = nullptr;
        uint32_t nt = 0, nv = 0;
        const int rc = sq_load_obj(path.c_str(), &tris, &nt, &nv);
        if (rc == SQ_OK) {
            loaded++;
            // touch every triangle the loader returned
            volatile float acc = 0.f;
            for (uint32_t i = 0; i < nt; i++) acc += tris[i].v0.x + tris[i].v2.z;
            (void)acc;
        } else {
            rejected++;
            if (tris) return 3;  // a failed load returns no array
        }
        sq_host_free(tris);
    }
    // BMP encoding: random sizes, exact and short capacities
    for (int it = 0; it < 300; it++) {
        const uint32_t w = 1 + rnd(67), h = 1 + rnd(45);
        std::vector<uint8_t> rgb((size_t)w * h * 3);
        for (auto& b : rgb) b = (uint8_t)rnd(256);
        const size_t need = sq_bmp_size(w, h);
        std::vector<uint8_t> out(need);
        if (sq_encode_bmp(rgb.data(), w, h, out.data(), need) != SQ_OK) return 4;
        if (need > 0 && sq_encode_bmp(rgb.data(), w, h, out.data(), need - 1 - rnd((uint32_t)need)) == SQ_OK) return 5;
        if (it % 50 == 0 && sq_write_bmp((std::string(argv[2]) + "/fuzz.bmp").c_str(), rgb.data(), w, h) != SQ_OK)
            return 6;
    }
    std::printf("sceneio fuzz: %ld files, %ld loaded, %ld rejected\n", iters, loaded, rejected);
    return 0;
}
