// sceneio.cpp — scene files on the host: Wavefront OBJ input and the heatmap
// BMP output, with the reference's reading and byte layout.
//
// OBJ (loadOBJFromFile, main.cpp:84-179):
//   * the file is read in pieces of at most 255 bytes, each ending at a
//     newline (a longer line continues as a new record);
//   * a token ends at the first delimiter, which is consumed; two delimiters
//     in a row give an empty token; the first token of a record ends at
//     space, tab, CR or LF, the later ones at space, tab or LF;
//   * only "v x y z" and "f i j k [l]" records count: coordinates as by atof
//     (C locale), indices as by atoi (the text before a '/'), 1-based or,
//     when negative, relative to the vertices read so far; a fourth index
//     fans the face into (i, j, k) and (i, k, l), further ones are ignored;
//   * triangles are assembled after the whole file is read, so a face may
//     name a vertex that comes later.
// Inputs the reference cannot read (it dereferences a null token or indexes
// past its vertex array) are given a defined result here: a required token
// not followed by a delimiter (last line without a newline, or cut at 255
// bytes) is the text up to the end of the piece; an empty fourth index is no
// index; an index that names no vertex fails the load with SQ_ERR_INVALID.
//
// BMP (stbi_write_bmp after stbi_flip_vertically_on_write(1), main.cpp:335-336
// with stb_image_write.h:428-475): 14-byte file header and 40-byte info header
// (24 bpp, no compression, zero size and resolution fields), then the rows in
// memory order (row 0 first), each pixel as B, G, R, each row zero-padded to
// a multiple of 4 bytes.
#include "sequoia.h"

#include <clocale>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <locale.h>
#include <string>
#include <vector>

namespace sq {
void set_error(const char* fmt, ...);  // context.hip: the message sq_last_error() returns
}

namespace {

int fail(int code, const char* msg) {
    sq::set_error("%s", msg);
    return code;
}

bool is_delim(char c, const char* delims) { return c != '\0' && std::strchr(delims, c) != nullptr; }

// Next token of the piece at `s` under the reference's rule. Returns false
// (and leaves `s`) when no delimiter follows; `rest_ok` then takes the text
// up to the end of the piece instead, if there is any.
bool take(const char*& s, const char* delims, std::string& tok, bool rest_ok) {
    const char* p = s;
    while (*p && !is_delim(*p, delims)) p++;
    if (*p) {
        tok.assign(s, (size_t)(p - s));
        s = p + 1;
        return true;
    }
    if (rest_ok && p != s) {
        tok.assign(s, (size_t)(p - s));
        s = p;
        return true;
    }
    return false;
}

locale_t c_locale() {
    static locale_t loc = newlocale(LC_NUMERIC_MASK, "C", (locale_t)0);
    return loc;
}

float parse_float(const std::string& t) { return (float)strtod_l(t.c_str(), nullptr, c_locale()); }

long parse_int(const std::string& t) {
    // atoi: optional sign and decimal digits, the rest ignored
    const long v = std::strtol(t.c_str(), nullptr, 10);
    return (long)(int)v;
}

}  // namespace

extern "C" {

void sq_host_free(void* p) { std::free(p); }

int sq_load_obj(const char* path, sq_triangle** out, uint32_t* ntris, uint32_t* nverts) {
    if (!path || !out || !ntris) return fail(SQ_ERR_INVALID, "sq_load_obj: null argument");
    *out = nullptr;
    *ntris = 0;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) {
        sq::set_error("cannot open OBJ file %s", path);
        return SQ_ERR_INVALID;
    }
    std::vector<float> verts;        // xyz
    std::vector<int64_t> idx;        // resolved 0-based indices, three per triangle
    std::vector<uint32_t> idx_line;  // record number of each triangle (for errors)
    char piece[256];
    const char* kFirst = " \t\n\r";
    const char* kRest = " \t\n";
    uint32_t record = 0;
    std::string tok;
    int rc = SQ_OK;
    while (std::fgets(piece, sizeof piece, fp)) {
        record++;
        const char* s = piece;
        if (!take(s, kFirst, tok, false)) continue;
        if (tok == "v") {
            float f[3];
            for (int k = 0; k < 3; k++) {
                if (!take(s, kRest, tok, true)) {
                    sq::set_error("%s: record %u: vertex with fewer than 3 coordinates", path,
                                  record);
                    rc = SQ_ERR_INVALID;
                    break;
                }
                f[k] = parse_float(tok);
            }
            if (rc) break;
            verts.insert(verts.end(), f, f + 3);
        } else if (tok == "f") {
            const int64_t nv = (int64_t)(verts.size() / 3);
            auto resolve = [&](const std::string& t) -> int64_t {
                const long i = parse_int(t);
                return i < 0 ? nv + i : (int64_t)i - 1;
            };
            int64_t v[3];
            for (int k = 0; k < 3; k++) {
                if (!take(s, kRest, tok, true)) {
                    sq::set_error("%s: record %u: face with fewer than 3 indices", path, record);
                    rc = SQ_ERR_INVALID;
                    break;
                }
                v[k] = resolve(tok);
            }
            if (rc) break;
            idx.insert(idx.end(), v, v + 3);
            idx_line.push_back(record);
            if (take(s, kRest, tok, false) && !tok.empty()) {
                const int64_t w = resolve(tok);
                const int64_t q[3] = {v[0], v[2], w};
                idx.insert(idx.end(), q, q + 3);
                idx_line.push_back(record);
            }
        }
    }
    std::fclose(fp);
    if (rc) return rc;
    const int64_t nv = (int64_t)(verts.size() / 3);
    const size_t nt = idx.size() / 3;
    for (size_t i = 0; i < idx.size(); i++)
        if (idx[i] < 0 || idx[i] >= nv) {
            sq::set_error("%s: record %u: face index names no vertex (%lld of %lld)", path,
                          idx_line[i / 3], (long long)idx[i] + 1, (long long)nv);
            return SQ_ERR_INVALID;
        }
    if (nt > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "sq_load_obj: too many triangles");
    sq_triangle* t = (sq_triangle*)std::malloc(nt ? nt * sizeof(sq_triangle) : 1);
    if (!t) return fail(SQ_ERR_INVALID, "sq_load_obj: out of host memory");
    float* d = (float*)t;
    for (size_t i = 0; i < idx.size(); i++) std::memcpy(d + 3 * i, &verts[3 * (size_t)idx[i]], 12);
    *out = t;
    *ntris = (uint32_t)nt;
    if (nverts) *nverts = (uint32_t)nv;
    return SQ_OK;
}

size_t sq_bmp_size(uint32_t w, uint32_t h) {
    const size_t row = (size_t)w * 3, pad = (4 - row % 4) % 4;
    return 54 + (row + pad) * h;
}

int sq_encode_bmp(const uint8_t* rgb, uint32_t w, uint32_t h, uint8_t* out, size_t cap) {
    if (!rgb || !out) return fail(SQ_ERR_INVALID, "sq_encode_bmp: null argument");
    if (w > 0x7FFFFFFFu || h > 0x7FFFFFFFu) return fail(SQ_ERR_INVALID, "sq_encode_bmp: image too large");
    const size_t total = sq_bmp_size(w, h);
    if (cap < total) return fail(SQ_ERR_INVALID, "sq_encode_bmp: output buffer too small");
    if (total > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "sq_encode_bmp: image too large for BMP");
    uint8_t* p = out;
    auto u16 = [&](uint32_t v) { *p++ = (uint8_t)v; *p++ = (uint8_t)(v >> 8); };
    auto u32 = [&](uint32_t v) { u16(v & 0xFFFF); u16(v >> 16); };
    *p++ = 'B';
    *p++ = 'M';
    u32((uint32_t)total);
    u16(0);
    u16(0);
    u32(54);
    u32(40);
    u32(w);
    u32(h);
    u16(1);
    u16(24);
    for (int i = 0; i < 6; i++) u32(0);
    const size_t row = (size_t)w * 3, pad = (4 - row % 4) % 4;
    for (uint32_t y = 0; y < h; y++) {
        const uint8_t* src = rgb + (size_t)y * row;
        for (uint32_t x = 0; x < w; x++) {
            *p++ = src[3 * x + 2];
            *p++ = src[3 * x + 1];
            *p++ = src[3 * x];
        }
        for (size_t k = 0; k < pad; k++) *p++ = 0;
    }
    return SQ_OK;
}

int sq_write_bmp(const char* path, const uint8_t* rgb, uint32_t w, uint32_t h) {
    if (!path) return fail(SQ_ERR_INVALID, "sq_write_bmp: null path");
    std::vector<uint8_t> buf(sq_bmp_size(w, h));
    int rc = sq_encode_bmp(rgb, w, h, buf.data(), buf.size());
    if (rc) return rc;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) {
        sq::set_error("cannot create %s", path);
        return SQ_ERR_INVALID;
    }
    const size_t wr = std::fwrite(buf.data(), 1, buf.size(), fp);
    const int cl = std::fclose(fp);
    if (wr != buf.size() || cl != 0) {
        sq::set_error("short write to %s", path);
        return SQ_ERR_INVALID;
    }
    return SQ_OK;
}

}  // extern "C"
