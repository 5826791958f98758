// radix_sort.hip — stable LSD radix sort of (key, uint32 value) pairs.
//
// Same result as radix_sort / parallel_counting_sort (algorithms.h:41-113):
// every pass is a stable counting sort on one digit, so the output is the
// input ordered by (key, original position).
//
// Per pass (reduce-then-scan, three kernels):
//   upsweep   each 4096-key tile builds its digit histogram; each wave counts
//             its own 1024 keys with ballot-matched peer groups into a
//             wave-private LDS row (no atomics),
//   scan      one workgroup per digit scans that digit's tile counts, then
//             one workgroup scans the 2^b digit totals,
//   downsweep each tile ranks its keys stably (wave-private running counts +
//             per-digit prefix over waves), stages the tile in LDS in digit
//             order and writes contiguous runs per digit.
// Algorithmic traffic per pass: read keys (upsweep) + read keys/values +
// write keys/values (downsweep).
#include "internal.h"

namespace sq {

constexpr int kRsWaves = 4;
constexpr int kRsThreads = 64 * kRsWaves;
constexpr int kRsRounds = 16;                          // 64-key rounds per wave
constexpr int kRsWaveKeys = 64 * kRsRounds;            // 1024 keys per wave
constexpr int kRsTile = kRsWaveKeys * kRsWaves;        // 4096 keys per tile
constexpr int kRadixMax = 256;

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, uint32_t shift, uint32_t mask) {
    return (uint32_t)(key >> shift) & mask;
}

// Lanes of the wave holding the same digit as this lane (among `active`).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint32_t bits, uint64_t active) {
    uint64_t peers = active;
    for (uint32_t b = 0; b < bits; b++) {
        uint64_t bal = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    return peers;
}

template <typename K>
__global__ __launch_bounds__(kRsThreads) void k_rs_upsweep(const K* __restrict__ keys, uint32_t n, uint32_t shift,
                                                           uint32_t bits, uint32_t* __restrict__ hist,
                                                           uint32_t num_tiles) {
    __shared__ uint32_t s_cnt[kRsWaves][kRadixMax];
    const uint32_t radix = 1u << bits, mask = radix - 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t i = threadIdx.x; i < kRsWaves * kRadixMax; i += kRsThreads) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRsTile + (size_t)wave * kRsWaveKeys;
    for (int r = 0; r < kRsRounds; r++) {
        size_t idx = base + (size_t)r * 64 + lane;
        bool valid = idx < n;
        uint64_t active = __ballot(valid);
        if (active == 0) break;
        uint32_t d = valid ? digit_of(keys[idx], shift, mask) : 0u;
        uint64_t peers = match_digit(d, bits, active);
        if (valid) {
            uint64_t lower = peers & ((1ull << lane) - 1ull);
            if (lower == 0) s_cnt[wave][d] += (uint32_t)__popcll(peers);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < radix; d += kRsThreads) {
        uint32_t s = 0;
        for (int w = 0; w < kRsWaves; w++) s += s_cnt[w][d];
        hist[(size_t)d * num_tiles + blockIdx.x] = s;
    }
}

// Offsets of the digit-major [digit][tile] counts in two steps: each digit
// row is scanned over its tiles by its own workgroup (row prefix + row
// total), then one workgroup scans the row totals into digit bases.
constexpr int kScanThreads = 256;
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_rows(uint32_t* __restrict__ hist, uint32_t num_tiles,
                                                            uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_w[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t* row = hist + (size_t)blockIdx.x * num_tiles;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < num_tiles; base += kScanThreads) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < num_tiles ? row[i] : 0u;
        uint32_t incl = wave_incl_scan(v, lane);
        if (lane == 63) s_w[wave] = incl;
        __syncthreads();
        uint32_t wb = 0, tot = 0;
        for (int w = 0; w < kScanThreads / 64; w++) {
            uint32_t t = s_w[w];
            wb += w < wave ? t : 0u;
            tot += t;
        }
        if (i < num_tiles) row[i] = carry + wb + incl - v;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_totals(uint32_t* __restrict__ totals, uint32_t radix) {
    __shared__ uint32_t s_w[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t v = threadIdx.x < radix ? totals[threadIdx.x] : 0u;
    uint32_t incl = wave_incl_scan(v, lane);
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();
    uint32_t wb = 0;
    for (int w = 0; w < wave; w++) wb += s_w[w];
    if (threadIdx.x < radix) totals[threadIdx.x] = wb + incl - v;
}

template <typename K>
__global__ __launch_bounds__(kRsThreads) void k_rs_downsweep(const K* __restrict__ keys_in,
                                                             const uint32_t* __restrict__ vals_in,
                                                             K* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                             uint32_t n, uint32_t shift, uint32_t bits,
                                                             const uint32_t* __restrict__ offsets,
                                                             const uint32_t* __restrict__ digit_base, uint32_t num_tiles) {
    __shared__ uint32_t s_cnt[kRsWaves][kRadixMax];  // per-wave running counts, then per-wave prefix
    __shared__ uint32_t s_start[kRadixMax];          // tile-local start of each digit
    __shared__ uint32_t s_gbase[kRadixMax];          // global offset of each digit for this tile
    __shared__ K s_keys[kRsTile];
    __shared__ uint32_t s_vals[kRsTile];

    const uint32_t radix = 1u << bits, mask = radix - 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t i = threadIdx.x; i < kRsWaves * kRadixMax; i += kRsThreads) (&s_cnt[0][0])[i] = 0;
    for (uint32_t d = threadIdx.x; d < radix; d += kRsThreads) s_gbase[d] = digit_base[d] + offsets[(size_t)d * num_tiles + blockIdx.x];
    __syncthreads();

    const size_t tile_base = (size_t)blockIdx.x * kRsTile;
    const size_t base = tile_base + (size_t)wave * kRsWaveKeys;
    K key[kRsRounds];
    uint32_t val[kRsRounds], rank[kRsRounds];
#pragma unroll
    for (int r = 0; r < kRsRounds; r++) {
        size_t idx = base + (size_t)r * 64 + lane;
        bool valid = idx < n;
        key[r] = valid ? keys_in[idx] : (K)0;
        val[r] = valid ? vals_in[idx] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kRsRounds; r++) {
        size_t idx = base + (size_t)r * 64 + lane;
        bool valid = idx < n;
        uint64_t active = __ballot(valid);
        uint32_t d = digit_of(key[r], shift, mask);
        rank[r] = 0;
        if (active == 0) continue;
        uint64_t peers = match_digit(d, bits, active);
        uint64_t lower = peers & ((1ull << lane) - 1ull);
        uint32_t prior = valid ? s_cnt[wave][d] : 0u;  // all lanes read before the leader writes
        rank[r] = prior + (uint32_t)__popcll(lower);
        __builtin_amdgcn_wave_barrier();
        if (valid && lower == 0) s_cnt[wave][d] = prior + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // Per-digit exclusive prefix over waves (in place) and tile histogram.
    __shared__ uint32_t s_hist[kRadixMax];
    for (uint32_t d = threadIdx.x; d < radix; d += kRsThreads) {
        uint32_t run = 0;
        for (int w = 0; w < kRsWaves; w++) {
            uint32_t c = s_cnt[w][d];
            s_cnt[w][d] = run;
            run += c;
        }
        s_hist[d] = run;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t d = 0; d < radix; d++) {
            s_start[d] = run;
            run += s_hist[d];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRsRounds; r++) {
        size_t idx = base + (size_t)r * 64 + lane;
        if (idx < n) {
            uint32_t d = digit_of(key[r], shift, mask);
            uint32_t pos = s_start[d] + s_cnt[wave][d] + rank[r];
            if (pos < (uint32_t)kRsTile) {
                s_keys[pos] = key[r];
                s_vals[pos] = val[r];
            }
        }
    }
    __syncthreads();
    const uint32_t tile_count = (uint32_t)((n - tile_base) < (size_t)kRsTile ? (n - tile_base) : (size_t)kRsTile);
    for (uint32_t i = threadIdx.x; i < tile_count; i += kRsThreads) {
        K k = s_keys[i];
        uint32_t d = digit_of(k, shift, mask);
        size_t out = (size_t)s_gbase[d] + (i - s_start[d]);
        if (out < n) {
            keys_out[out] = k;
            vals_out[out] = s_vals[i];
        }
    }
}

template <typename K>
static int radix_sort_t(sq_ctx* c, uint32_t n, uint32_t key_bits, uint32_t bits, int in, int* out_slot) {
    const uint32_t num_tiles = (n + kRsTile - 1) / kRsTile;
    const uint32_t radix = 1u << bits;
    SQ_TRY(c->hist.ensure((size_t)radix * num_tiles * 4 + kRadixMax * 4));
    uint32_t* totals = c->hist.as<uint32_t>() + (size_t)radix * num_tiles;
    const uint32_t passes = (key_bits + bits - 1) / bits;
    int cur = in;
    for (uint32_t p = 0; p < passes; p++) {
        uint32_t shift = p * bits;
        const K* kin = c->keys[cur].as<K>();
        const uint32_t* vin = c->vals[cur].as<uint32_t>();
        K* kout = c->keys[cur ^ 1].as<K>();
        uint32_t* vout = c->vals[cur ^ 1].as<uint32_t>();
        hipLaunchKernelGGL(k_rs_upsweep<K>, dim3(num_tiles), dim3(kRsThreads), 0, c->stream, kin, n, shift, bits,
                           c->hist.as<uint32_t>(), num_tiles);
        hipLaunchKernelGGL(k_scan_rows, dim3(radix), dim3(kScanThreads), 0, c->stream, c->hist.as<uint32_t>(),
                           num_tiles, totals);
        hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kScanThreads), 0, c->stream, totals, radix);
        hipLaunchKernelGGL(k_rs_downsweep<K>, dim3(num_tiles), dim3(kRsThreads), 0, c->stream, kin, vin, kout, vout, n,
                           shift, bits, c->hist.as<uint32_t>(), totals, num_tiles);
        SQ_CHECK_HIP(hipGetLastError());
        cur ^= 1;
    }
    c->last.sort_passes = passes;
    *out_slot = cur;
    return SQ_OK;
}

int radix_sort(sq_ctx* c, uint32_t n, uint32_t key_bits, uint32_t digit_bits, bool key64, int in, int* out_slot) {
    if (n == 0) { *out_slot = in; return SQ_OK; }
    if (key64) return radix_sort_t<uint64_t>(c, n, key_bits, digit_bits, in, out_slot);
    return radix_sort_t<uint32_t>(c, n, key_bits, digit_bits, in, out_slot);
}

}  // namespace sq
