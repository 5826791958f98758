// scenegen.cpp — deterministic synthetic scenes and ray sets (host code).
//
// Counter-based random numbers (splitmix64 of seed, stream and index) make
// every generator independent of thread count and call order. Generation is
// done in double precision and rounded once to float; -0 is canonicalised to
// +0 so that box min/max reductions never depend on the order in which equal
// zeros of opposite sign meet.
#include "sequoia.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
struct Rng {
    uint64_t state;
    Rng(uint64_t seed, uint64_t stream, uint64_t index)
        : state(splitmix64(seed * 0x632BE59BD9B4E019ull ^ splitmix64(stream * 0x8CB92BA72F3D8DD7ull ^ index))) {}
    uint64_t next() { state += 0x9E3779B97F4A7C15ull; return splitmix64(state); }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double gauss() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline float canon(double v) { return (float)v + 0.0f; }

struct D3 { double x, y, z; };
inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 mul(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline D3 normalize(D3 a) { double l = std::sqrt(dot(a, a)); return l > 0 ? mul(a, 1.0 / l) : D3{0, 0, 1}; }

inline sq_vec3 tof(D3 a) { return {canon(a.x), canon(a.y), canon(a.z)}; }

// 3-D value noise: hashed lattice values in [-1,1], smoothstep-trilinear.
double lattice(int64_t i, int64_t j, int64_t k, uint32_t seed, int octave) {
    uint64_t h = splitmix64((uint64_t)i * 0x9E3779B1ull ^ splitmix64((uint64_t)j * 0x85EBCA77ull ^
                 splitmix64((uint64_t)k * 0xC2B2AE3Dull ^ ((uint64_t)seed << 8) ^ (uint64_t)octave)));
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
double value_noise(D3 p, uint32_t seed, int octave) {
    double fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    double tx = p.x - fx, ty = p.y - fy, tz = p.z - fz;
    tx = tx * tx * (3 - 2 * tx); ty = ty * ty * (3 - 2 * ty); tz = tz * tz * (3 - 2 * tz);
    double c[2][2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed, octave);
    double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx, x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx, x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}
// 3 octaves (frequencies 2, 4, 8; amplitudes 1, 1/2, 1/4), normalised to [-1,1].
double fbm3(D3 p, uint32_t seed) {
    double s = 0, amp = 1, freq = 2, norm = 0;
    for (int o = 0; o < 3; o++) { s += amp * value_noise(mul(p, freq), seed, o); norm += amp; amp *= 0.5; freq *= 2; }
    return s / norm;
}

const double kIcoV[12][3] = {
    {-1, 1.618033988749895, 0}, {1, 1.618033988749895, 0}, {-1, -1.618033988749895, 0}, {1, -1.618033988749895, 0},
    {0, -1, 1.618033988749895}, {0, 1, 1.618033988749895}, {0, -1, -1.618033988749895}, {0, 1, -1.618033988749895},
    {1.618033988749895, 0, -1}, {1.618033988749895, 0, 1}, {-1.618033988749895, 0, -1}, {-1.618033988749895, 0, 1}};
const int kIcoF[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};

}  // namespace

extern "C" {

// Frequency-f geodesic subdivision of the icosahedron has 20 f^2 faces; the
// largest f with 20 f^2 <= target is used and the first (target - 20 f^2)
// triangles are bisected through an edge midpoint (same surface), so the
// count is exactly `target` whenever target >= 20 and target <= 40 f^2.
uint32_t sq_gen_icosphere_count(uint32_t target) {
    if (target < 20) return 20;
    uint32_t f = (uint32_t)std::floor(std::sqrt((double)target / 20.0));
    while (20ull * (f + 1) * (f + 1) <= target) f++;
    while (f > 1 && 20ull * f * f > target) f--;
    uint64_t base = 20ull * f * f;
    uint64_t extra = target - base;
    if (extra > base) extra = base;
    return (uint32_t)(base + extra);
}

int sq_gen_icosphere(uint32_t target, uint32_t seed, float amplitude, sq_triangle* out) {
    if (!out) return SQ_ERR_INVALID;
    uint32_t total = sq_gen_icosphere_count(target);
    uint32_t f = 1;
    while (20ull * (f + 1) * (f + 1) <= (target < 20 ? 20 : target)) f++;
    uint64_t base = 20ull * f * f;
    uint64_t extra = total - base;
    auto vert = [&](int face, double a, double b) -> D3 {
        const double* A = kIcoV[kIcoF[face][0]];
        const double* B = kIcoV[kIcoF[face][1]];
        const double* C = kIcoV[kIcoF[face][2]];
        double w = 1.0 - a - b;
        D3 p{A[0] * w + B[0] * a + C[0] * b, A[1] * w + B[1] * a + C[1] * b, A[2] * w + B[2] * a + C[2] * b};
        D3 d = normalize(p);
        double r = 1.0 + (double)amplitude * fbm3(d, seed);
        return mul(d, r);
    };
    std::vector<sq_triangle> tris;
    tris.reserve(total);
    for (int face = 0; face < 20; face++) {
        for (uint32_t i = 0; i < f; i++) {
            for (uint32_t j = 0; j + i < f; j++) {
                double a0 = (double)i / f, b0 = (double)j / f, s = 1.0 / f;
                D3 p0 = vert(face, a0, b0), p1 = vert(face, a0 + s, b0), p2 = vert(face, a0, b0 + s);
                tris.push_back({tof(p0), tof(p1), tof(p2)});
                if (i + j + 1 < f) {
                    D3 q0 = vert(face, a0 + s, b0), q1 = vert(face, a0 + s, b0 + s), q2 = vert(face, a0, b0 + s);
                    tris.push_back({tof(q0), tof(q1), tof(q2)});
                }
            }
        }
    }
    // Bisect the first `extra` triangles: (v0,v1,v2) -> (v0,v1,m), (v0,m,v2), m = mid(v1,v2).
    std::vector<sq_triangle> result;
    result.reserve(total);
    for (uint64_t t = 0; t < tris.size(); t++) {
        if (t < extra) {
            const sq_triangle& T = tris[t];
            sq_vec3 m = {canon(((double)T.v1.x + T.v2.x) * 0.5), canon(((double)T.v1.y + T.v2.y) * 0.5),
                         canon(((double)T.v1.z + T.v2.z) * 0.5)};
            result.push_back({T.v0, T.v1, m});
            result.push_back({T.v0, m, T.v2});
        } else {
            result.push_back(tris[t]);
        }
    }
    std::memcpy(out, result.data(), sizeof(sq_triangle) * result.size());
    return (int)result.size();
}

// Triangle soup: centroid ~ U[0,1)^3, edge vectors a, b ~ N(0, sigma^2)^3,
// v0 = c - (a+b)/3, v1 = v0 + a, v2 = v0 + b (centroid exactly c in reals).
int sq_gen_soup(uint32_t n, uint32_t seed, float sigma, sq_triangle* out) {
    if (!out) return SQ_ERR_INVALID;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        Rng r(seed, 1, (uint64_t)i);
        D3 c{r.uniform(), r.uniform(), r.uniform()};
        D3 a{r.gauss() * sigma, r.gauss() * sigma, r.gauss() * sigma};
        D3 b{r.gauss() * sigma, r.gauss() * sigma, r.gauss() * sigma};
        D3 v0 = sub(c, mul(add(a, b), 1.0 / 3.0));
        out[i] = {tof(v0), tof(add(v0, a)), tof(add(v0, b))};
    }
    return (int)n;
}

// Clustered scene: `clusters` Gaussian clusters with centres ~ U[0,1)^3 and
// sigma ~ U[0.02, 0.2]; each triangle picks a cluster uniformly; its size is
// log-normal (mu -5, sigma 1) and its edges are that size times N(0,1)^3.
int sq_gen_clusters(uint32_t n, uint32_t clusters, uint32_t seed, sq_triangle* out) {
    if (!out || clusters == 0) return SQ_ERR_INVALID;
    std::vector<D3> centre(clusters);
    std::vector<double> sig(clusters);
    for (uint32_t k = 0; k < clusters; k++) {
        Rng r(seed, 2, k);
        centre[k] = {r.uniform(), r.uniform(), r.uniform()};
        sig[k] = 0.02 + 0.18 * r.uniform();
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        Rng r(seed, 3, (uint64_t)i);
        uint32_t k = (uint32_t)(r.uniform() * clusters);
        if (k >= clusters) k = clusters - 1;
        D3 c{centre[k].x + sig[k] * r.gauss(), centre[k].y + sig[k] * r.gauss(), centre[k].z + sig[k] * r.gauss()};
        double size = std::exp(-5.0 + 1.0 * r.gauss());
        D3 a{size * r.gauss(), size * r.gauss(), size * r.gauss()};
        D3 b{size * r.gauss(), size * r.gauss(), size * r.gauss()};
        D3 v0 = sub(c, mul(add(a, b), 1.0 / 3.0));
        out[i] = {tof(v0), tof(add(v0, a)), tof(add(v0, b))};
    }
    return (int)n;
}

// Secondary rays: origin uniform on a uniformly chosen triangle (barycentric
// square-root map), direction cosine-weighted about the geometric normal
// e1 x e2 (Malley's method), tmin as given, tmax = FLT_MAX.
int sq_gen_secondary_rays(const sq_triangle* tris, uint32_t ntri, uint32_t nrays, uint32_t seed, float tmin,
                          sq_ray* out) {
    if (!tris || !out || ntri == 0) return SQ_ERR_INVALID;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)nrays; i++) {
        Rng r(seed, 4, (uint64_t)i);
        uint32_t t = (uint32_t)(r.uniform() * ntri);
        if (t >= ntri) t = ntri - 1;
        const sq_triangle& T = tris[t];
        D3 A{T.v0.x, T.v0.y, T.v0.z}, B{T.v1.x, T.v1.y, T.v1.z}, C{T.v2.x, T.v2.y, T.v2.z};
        double r1 = r.uniform(), r2 = r.uniform(), sr = std::sqrt(r1);
        D3 p = add(add(mul(A, 1.0 - sr), mul(B, sr * (1.0 - r2))), mul(C, sr * r2));
        D3 nrm = normalize(cross(sub(B, A), sub(C, A)));
        D3 helper = std::fabs(nrm.x) > 0.9 ? D3{0, 1, 0} : D3{1, 0, 0};
        D3 tx = normalize(cross(helper, nrm));
        D3 ty = cross(nrm, tx);
        double u1 = r.uniform(), u2 = r.uniform();
        double rad = std::sqrt(u1), phi = 6.283185307179586 * u2;
        double lx = rad * std::cos(phi), ly = rad * std::sin(phi), lz = std::sqrt(std::fmax(0.0, 1.0 - u1));
        D3 d = normalize(add(add(mul(tx, lx), mul(ty, ly)), mul(nrm, lz)));
        sq_ray& o = out[i];
        o.origin = tof(p); o.tmin = tmin; o.dir = tof(d); o.tmax = 3.402823466e+38f;
    }
    return (int)nrays;
}

// run() camera (main.cpp:311-318): scene box centre in x,y; z = 2*min.z - centre.z;
// looks along +z with the 90-degree frustum (tan_half_fov = 1), tmin 0, tmax 1e4.
int sq_reference_camera(const sq_triangle* tris, uint32_t n, uint32_t w, uint32_t h, sq_camera* cam) {
    if (!tris || !cam || n == 0) return SQ_ERR_INVALID;
    float mn[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
    float mx[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
    for (uint32_t i = 0; i < n; i++) {
        const float* v = &tris[i].v0.x;
        for (int k = 0; k < 9; k++) {
            float x = v[k];
            int a = k % 3;
            mn[a] = (x < mn[a]) ? x : mn[a];
            mx[a] = (mx[a] < x) ? x : mx[a];
        }
    }
    float cx = (mn[0] + mx[0]) * 0.5f, cy = (mn[1] + mx[1]) * 0.5f, cz = (mn[2] + mx[2]) * 0.5f;
    std::memset(cam, 0, sizeof(*cam));
    cam->origin = {cx, cy, 2.0f * mn[2] - cz};
    cam->tan_half_fov = 1.0f;
    cam->width = w; cam->height = h; cam->spp = 1; cam->jitter = 0; cam->seed = 0;
    cam->tmin = 0.0f; cam->tmax = 10000.0f;
    return SQ_OK;
}

// Pinhole camera `distance` in front of the scene-box centre along -z looking
// +z, vertical field of view `fov_deg`.
int sq_pinhole_camera(const sq_triangle* tris, uint32_t n, uint32_t w, uint32_t h, float fov_deg, float distance,
                      uint32_t spp, uint32_t jitter, uint32_t seed, sq_camera* cam) {
    if (!tris || !cam || n == 0) return SQ_ERR_INVALID;
    sq_camera ref;
    sq_reference_camera(tris, n, w, h, &ref);
    float mn_z = 3.402823466e+38f, mx_z = -3.402823466e+38f;
    for (uint32_t i = 0; i < n; i++) {
        const float* v = &tris[i].v0.x;
        for (int k = 2; k < 9; k += 3) { mn_z = (v[k] < mn_z) ? v[k] : mn_z; mx_z = (mx_z < v[k]) ? v[k] : mx_z; }
    }
    float cz = (mn_z + mx_z) * 0.5f;
    *cam = ref;
    cam->origin.z = cz - distance;
    cam->tan_half_fov = (float)std::tan((double)fov_deg * 0.5 * 3.141592653589793 / 180.0);
    cam->spp = spp ? spp : 1;
    cam->jitter = jitter;
    cam->seed = seed;
    cam->tmax = 3.402823466e+38f;
    return SQ_OK;
}

}  // extern "C"
