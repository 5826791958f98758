// radix_sort.hip — stable LSD radix sort of (key, uint32 value) pairs.
//
// Same result as radix_sort / parallel_counting_sort (algorithms.h:41-113):
// every pass is a stable counting sort on one digit, so the output is the
// input ordered by (key, original position).
//
// Plain builds and sq_radix_sort_pairs run each key pass as one kernel with
// decoupled look-back (k_rs_onesweep, below), after one read of the keys
// that counts the digits of every pass. Batched and served builds, and the
// segment passes, use reduce-then-scan, three kernels per pass:
//   upsweep   each 4096-key tile builds its digit histogram; each wave counts
//             its own 256 keys with ballot-matched peer groups into a
//             wave-private LDS row (no atomics),
//   scan      one workgroup per digit scans that digit's tile counts, then
//             one workgroup scans the 2^b digit totals,
//   downsweep each tile ranks its keys stably (wave-private running counts +
//             per-digit prefix over waves), stages the tile in LDS in digit
//             order and writes contiguous runs per digit.
// Algorithmic traffic per pass: read keys (upsweep) + read keys/values +
// write keys/values (downsweep).
//
// Segment passes (batched builds): the digit is taken from the segment that
// holds the pair's value (values are indices into a concatenation of
// segments, offsets in LDS, binary search). Run after the key passes they
// group the pairs by segment while keeping each segment's key order, i.e.
// sort by (segment, key, value) with the keys carried along.
#include "internal.h"

namespace sq {

// 16 waves x 4 rounds per 4096-key tile measured fastest (1M 63-bit keys: the
// 8-pass sort 0.19 ms, against 0.20 at 8 x 8 and 0.24 at 4 x 16)
constexpr int kRsWaves = 16;
constexpr int kRsThreads = 64 * kRsWaves;
constexpr int kRsRounds = 4;                           // 64-key rounds per wave
constexpr int kRsWaveKeys = 64 * kRsRounds;            // 256 keys per wave
constexpr int kRsTile = kRsWaveKeys * kRsWaves;        // 4096 keys per tile
constexpr int kRadixMax = 256;
constexpr uint32_t kMaxSegs = 2048;  // segments per segment pass (LDS table)

// Digit source of a pass: the key, or the segment of the value.
struct SegTable {
    const uint32_t* off;  // nseg + 1 ascending offsets, off[0] = 0
    uint32_t nseg;
};

__device__ __forceinline__ void load_seg_lds(uint32_t* s_off, const SegTable& t) {
    for (uint32_t i = threadIdx.x; i <= t.nseg; i += blockDim.x) s_off[i] = t.off[i];
}
// Largest s with off[s] <= v (v < off[nseg]).
__device__ __forceinline__ uint32_t seg_of(const uint32_t* s_off, uint32_t nseg, uint32_t v) {
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (s_off[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

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

template <typename K, bool SEG>
__global__ __launch_bounds__(kRsThreads) void k_rs_upsweep(const K* __restrict__ keys,
                                                           const uint32_t* __restrict__ vals, uint32_t n,
                                                           uint32_t shift, uint32_t bits, uint32_t* __restrict__ hist,
                                                           uint32_t num_tiles, SegTable seg) {
    __shared__ uint32_t s_cnt[kRsWaves][kRadixMax];
    __shared__ uint32_t s_off[SEG ? kMaxSegs + 1 : 1];
    const uint32_t radix = 1u << bits, mask = radix - 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t i = threadIdx.x; i < kRsWaves * kRadixMax; i += kRsThreads) (&s_cnt[0][0])[i] = 0;
    if (SEG) load_seg_lds(s_off, seg);
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRsTile + (size_t)wave * kRsWaveKeys;
    for (int r = 0; r < kRsRounds; r++) {
        size_t idx = base + (size_t)r * 64 + lane;
        bool valid = idx < n;
        uint64_t active = __ballot(valid);
        if (active == 0) break;
        uint32_t d = 0;
        if (valid) d = SEG ? (seg_of(s_off, seg.nseg, vals[idx]) >> shift) & mask : digit_of(keys[idx], shift, mask);
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

template <typename K, bool SEG>
__global__ __launch_bounds__(kRsThreads) void k_rs_downsweep(const K* __restrict__ keys_in,
                                                             const uint32_t* __restrict__ vals_in,
                                                             K* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                             uint32_t n, uint32_t shift, uint32_t bits,
                                                             const uint32_t* __restrict__ offsets,
                                                             const uint32_t* __restrict__ digit_base, uint32_t num_tiles,
                                                             SegTable seg) {
    __shared__ uint32_t s_cnt[kRsWaves][kRadixMax];  // per-wave running counts, then per-wave prefix
    __shared__ uint32_t s_start[kRadixMax];          // tile-local start of each digit
    __shared__ uint32_t s_gbase[kRadixMax];          // global offset of each digit for this tile
    __shared__ K s_keys[kRsTile];
    __shared__ uint32_t s_vals[kRsTile];
    __shared__ uint32_t s_off[SEG ? kMaxSegs + 1 : 1];

    const uint32_t radix = 1u << bits, mask = radix - 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t i = threadIdx.x; i < kRsWaves * kRadixMax; i += kRsThreads) (&s_cnt[0][0])[i] = 0;
    if (SEG) load_seg_lds(s_off, seg);
    for (uint32_t d = threadIdx.x; d < radix; d += kRsThreads) s_gbase[d] = digit_base[d] + offsets[(size_t)d * num_tiles + blockIdx.x];
    __syncthreads();

    const size_t tile_base = (size_t)blockIdx.x * kRsTile;
    const size_t base = tile_base + (size_t)wave * kRsWaveKeys;
    K key[kRsRounds];
    uint32_t val[kRsRounds], rank[kRsRounds], dig[kRsRounds];
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
        uint32_t d = SEG ? (valid ? (seg_of(s_off, seg.nseg, val[r]) >> shift) & mask : 0u)
                         : digit_of(key[r], shift, mask);
        dig[r] = d;
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
            uint32_t d = dig[r];
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
        uint32_t d = SEG ? (seg_of(s_off, seg.nseg, s_vals[i]) >> shift) & mask : digit_of(k, shift, mask);
        size_t out = (size_t)s_gbase[d] + (i - s_start[d]);
        if (out < n) {
            keys_out[out] = k;
            vals_out[out] = s_vals[i];
        }
    }
}

// ---- single-pass key passes (plain builds) ----
// Digit totals of every pass come from one read of the keys up front
// (k_rs_ghist). Each pass is then one kernel: a workgroup takes the next
// tile by ticket, ranks its keys as the downsweep does, publishes its digit
// counts, and finds the keys of earlier tiles with each digit by looking back
// over the earlier tiles' published words (decoupled look-back): a word
// holds either that tile's own count (flag 1) or the count of all tiles up to
// and including it (flag 2); the walk stops at the first flag-2 word. Tiles
// are taken in ticket order, and a tile publishes its own counts before it
// looks back, so every word it waits for belongs to a workgroup that is
// already running and never waits for it: the wait always ends. A wait that
// nevertheless spins kSpinLimit times sets the error word (pinned host
// memory) and the pass continues with a zero count (the host then reports
// the sort as failed).
constexpr uint32_t kFlagAgg = 1u << 30, kFlagPrefix = 2u << 30, kValueMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 20;  // about 30 ms of s_sleep(1) per word
constexpr uint32_t kMaxKeyPasses = 64;

// Layout of the single-pass area (words): digit totals [kMaxKeyPasses][256],
// tickets [kMaxKeyPasses], a spare word, then per pass num_tiles x radix
// status words.
__host__ __device__ constexpr size_t os_ghist() { return 0; }
__host__ __device__ constexpr size_t os_tickets() { return (size_t)kMaxKeyPasses * kRadixMax; }
__host__ __device__ constexpr size_t os_spare() { return os_tickets() + kMaxKeyPasses; }
__host__ __device__ constexpr size_t os_status() { return os_spare() + 64; }

// Digit totals of passes p0 .. kpasses-1 (ghist row p = pass p).
template <typename K>
__global__ __launch_bounds__(kRsThreads) void k_rs_ghist(const K* __restrict__ keys, uint32_t n, uint32_t bits,
                                                         uint32_t p0, uint32_t kpasses, uint32_t* __restrict__ ghist) {
    __shared__ uint32_t s_h[kMaxKeyPasses * 2 > 8 * kRadixMax ? kMaxKeyPasses * 2 : 8 * kRadixMax];
    const uint32_t radix = 1u << bits, mask = radix - 1, cells = (kpasses - p0) * radix;
    for (uint32_t i = threadIdx.x; i < cells; i += kRsThreads) s_h[i] = 0;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * kRsThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kRsThreads) {
        const K k = keys[i];
        for (uint32_t p = p0; p < kpasses; p++) atomicAdd(&s_h[(p - p0) * radix + digit_of(k, p * bits, mask)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < cells; i += kRsThreads) {
        const uint32_t v = s_h[i];
        if (v) atomicAdd(&ghist[(p0 + i / radix) * kRadixMax + (i % radix)], v);
    }
}

template <typename K>
__global__ __launch_bounds__(kRsThreads) void k_rs_onesweep(const K* __restrict__ keys_in,
                                                            const uint32_t* __restrict__ vals_in,
                                                            K* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                            uint32_t n, uint32_t shift, uint32_t bits,
                                                            const uint32_t* __restrict__ ghist_pass,
                                                            uint32_t* __restrict__ ticket,
                                                            uint32_t* __restrict__ status, uint32_t* __restrict__ err) {
    __shared__ uint32_t s_cnt[kRsWaves][kRadixMax];
    __shared__ uint32_t s_start[kRadixMax];
    __shared__ uint32_t s_gbase[kRadixMax];
    __shared__ uint32_t s_hist[kRadixMax];
    __shared__ uint32_t s_dbase[kRadixMax];
    __shared__ K s_keys[kRsTile];
    __shared__ uint32_t s_vals[kRsTile];
    __shared__ uint32_t s_tile;

    const uint32_t radix = 1u << bits, mask = radix - 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    for (uint32_t i = threadIdx.x; i < kRsWaves * kRadixMax; i += kRsThreads) (&s_cnt[0][0])[i] = 0;
    // digit bases: exclusive scan of this pass's digit totals (one wave)
    if (wave == 0) {
        uint32_t v[kRadixMax / 64], run = 0;
#pragma unroll
        for (int j = 0; j < kRadixMax / 64; j++) {
            const uint32_t d = (uint32_t)lane * (kRadixMax / 64) + j;
            v[j] = d < radix ? ghist_pass[d] : 0u;
            run += v[j];
        }
        const uint32_t incl = wave_incl_scan(run, lane);
        uint32_t ex = incl - run;
#pragma unroll
        for (int j = 0; j < kRadixMax / 64; j++) {
            const uint32_t d = (uint32_t)lane * (kRadixMax / 64) + j;
            s_dbase[d] = ex;
            ex += v[j];
        }
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    const size_t tile_base = (size_t)tile * kRsTile;
    const size_t base = tile_base + (size_t)wave * kRsWaveKeys;
    K key[kRsRounds];
    uint32_t val[kRsRounds], rank[kRsRounds], dig[kRsRounds];
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
        dig[r] = d;
        rank[r] = 0;
        if (active == 0) continue;
        uint64_t peers = match_digit(d, bits, active);
        uint64_t lower = peers & ((1ull << lane) - 1ull);
        uint32_t prior = valid ? s_cnt[wave][d] : 0u;
        rank[r] = prior + (uint32_t)__popcll(lower);
        __builtin_amdgcn_wave_barrier();
        if (valid && lower == 0) s_cnt[wave][d] = prior + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < radix; d += kRsThreads) {
        uint32_t run = 0;
        for (int w = 0; w < kRsWaves; w++) {
            uint32_t c = s_cnt[w][d];
            s_cnt[w][d] = run;
            run += c;
        }
        s_hist[d] = run;
        // publish this tile's own count (tile 0: its count is the prefix)
        __hip_atomic_store(&status[(size_t)tile * radix + d], (tile == 0 ? kFlagPrefix : kFlagAgg) | run,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // look back: keys with digit d in earlier tiles (a window of 16 words
    // read at once measured no faster)
    for (uint32_t d = threadIdx.x; d < radix; d += kRsThreads) {
        uint32_t excl = 0;
        for (uint32_t t = tile; t-- > 0;) {
            uint32_t w, spins = 0;
            while (true) {
                w = __hip_atomic_load(&status[(size_t)t * radix + d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (w & (kFlagAgg | kFlagPrefix)) break;
                if (++spins > kSpinLimit) {
                    *err = 1u;  // host-visible word, read after the call's final sync
                    w = kFlagPrefix;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            excl += w & kValueMask;
            if (w & kFlagPrefix) break;
        }
        if (tile > 0)
            __hip_atomic_store(&status[(size_t)tile * radix + d], kFlagPrefix | ((excl + s_hist[d]) & kValueMask),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_gbase[d] = s_dbase[d] + excl;
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
            uint32_t d = dig[r];
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

// ---- local finish (plain builds with 8-bit digits and three or more passes) ----
// The two most significant digits run as global single-pass passes, which
// leaves the pairs ordered by their top 16 key bits (stably). Each workgroup
// then takes a span of about kLocSpan positions whose ends are moved back to
// the start of their top-16-bit run, so spans hold whole runs and together
// cover the array once, and sorts it by every digit in which its keys differ
// (LSD, stable, in LDS). Every pass is a stable counting sort on one digit,
// as in the global passes, so the result is the same (key, original position)
// order. A span longer than kLocCap (one run of equal top bits longer than
// the LDS holds) is sorted by the same workgroup in chunks through global
// memory: per digit a count over the span, then the chunks in order, each
// ranked stably and scattered behind the earlier chunks' keys.
constexpr int kLocCap = 8192;                        // keys sorted in LDS by one workgroup
constexpr int kLocRounds = kLocCap / kRsThreads;     // 8 keys per thread
constexpr int kLocWaveKeys = kLocCap / kRsWaves;     // 512 keys per wave
constexpr uint32_t kLocSpan = kLocCap / 2;           // nominal span per workgroup

// Smallest q <= p whose key has the same prefix (key >> pshift) as position p
// (keys ordered by prefix): one wave, 64-ary search, first over the kLocSpan
// positions before p, then over all of [0, p] when the run reaches further.
template <typename K>
__device__ uint32_t run_start(const K* __restrict__ keys, uint32_t p, uint32_t pshift, int lane) {
    if (p == 0) return 0;
    const K target = keys[p] >> pshift;
    uint32_t lo = p > kLocSpan ? p - kLocSpan : 0, hi = p;
    bool widened = lo == 0;
    while (lo < hi) {  // the answer lies in [lo, hi]; keys[hi] has the prefix
        const uint32_t span = hi - lo;
        const uint32_t pos = lo + (uint32_t)(((uint64_t)span * (uint32_t)lane) >> 6);
        const uint64_t eq = __ballot((keys[pos] >> pshift) == target);
        if (eq & 1ull) {  // lane 0 probes lo itself
            if (widened) return lo;
            hi = lo;
            lo = 0;
            widened = true;
            continue;
        }
        const int k = eq ? __ffsll((long long)eq) - 1 : 64;
        const uint32_t below = __shfl(pos, k - 1, 64);
        const uint32_t at = __shfl(pos, k < 64 ? k : 63, 64);
        lo = below + 1;
        if (k < 64) hi = at;
    }
    return lo;
}

// Exclusive scan of s_hist[0..256) into s_start by wave 0 (4 digits per lane).
__device__ __forceinline__ void digit_starts(const uint32_t* s_hist, uint32_t* s_start, int wave, int lane) {
    if (wave != 0) return;
    uint32_t v[kRadixMax / 64], run = 0;
#pragma unroll
    for (int j = 0; j < kRadixMax / 64; j++) {
        v[j] = s_hist[lane * (kRadixMax / 64) + j];
        run += v[j];
    }
    uint32_t ex = wave_incl_scan(run, lane) - run;
#pragma unroll
    for (int j = 0; j < kRadixMax / 64; j++) {
        s_start[lane * (kRadixMax / 64) + j] = ex;
        ex += v[j];
    }
}

// Stable ranks of one workgroup's kLocCap keys (position w*512 + r*64 + lane,
// the first m valid) on digit `shift`: dig[r], rank[r] within the wave, then
// s_cnt[w][d] = keys of digit d in earlier waves, s_hist[d] = total of d.
template <typename K>
__device__ __forceinline__ void local_rank(const K (&key)[kLocRounds], uint32_t m, uint32_t shift,
                                           uint32_t (&dig)[kLocRounds], uint32_t (&rank)[kLocRounds],
                                           uint32_t (*s_cnt)[kRadixMax], uint32_t* s_hist, int wave, int lane) {
    for (uint32_t i = threadIdx.x; i < kRsWaves * kRadixMax; i += kRsThreads) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kLocRounds; r++) {
        const uint32_t i = (uint32_t)(wave * kLocWaveKeys + r * 64 + lane);
        const bool valid = i < m;
        const uint64_t active = __ballot(valid);
        const uint32_t d = digit_of(key[r], shift, 0xFFu);
        dig[r] = d;
        rank[r] = 0;
        if (active == 0) continue;
        const uint64_t peers = match_digit(d, 8, active);
        const uint64_t lower = peers & ((1ull << lane) - 1ull);
        const uint32_t prior = valid ? s_cnt[wave][d] : 0u;
        rank[r] = prior + (uint32_t)__popcll(lower);
        __builtin_amdgcn_wave_barrier();
        if (valid && lower == 0) s_cnt[wave][d] = prior + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < kRadixMax; d += kRsThreads) {
        uint32_t run = 0;
        for (int w = 0; w < kRsWaves; w++) {
            const uint32_t c = s_cnt[w][d];
            s_cnt[w][d] = run;
            run += c;
        }
        s_hist[d] = run;
    }
    __syncthreads();
}

// spans[j] = start of the top-bit run holding position j * kLocSpan (n past
// the end), j = 0 .. nspans: one wave per entry. Computed before any span is
// sorted, since a span longer than the LDS is sorted in place of its input.
template <typename K>
__global__ __launch_bounds__(256) void k_rs_spans(const K* __restrict__ keys, uint32_t n, uint32_t pshift,
                                                  uint32_t nspans, uint32_t* __restrict__ spans) {
    const uint32_t j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j > nspans) return;
    const uint64_t nom = (uint64_t)j * kLocSpan;
    const uint32_t v = nom >= n ? n : (nom == 0 ? 0u : run_start<K>(keys, (uint32_t)nom, pshift, lane));
    if (lane == 0) spans[j] = v;
}

// Sorts span [spans[b], spans[b+1]) of workgroup b (all of [0, n) when spans
// is null) by every digit in which its keys differ.
template <typename K>
__global__ __launch_bounds__(kRsThreads) void k_rs_local(const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
                                                         K* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                         uint32_t n, uint32_t kpasses, const uint32_t* __restrict__ spans) {
    __shared__ K s_k[kLocCap];
    __shared__ uint32_t s_v[kLocCap];
    __shared__ uint32_t s_cnt[kRsWaves][kRadixMax];
    __shared__ uint32_t s_hist[kRadixMax];
    __shared__ uint32_t s_start[kRadixMax];
    __shared__ uint32_t s_span[2];
    __shared__ uint32_t s_bits[4];  // OR low/high, AND low/high of the span's keys
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x < 2) s_span[threadIdx.x] = spans ? spans[blockIdx.x + threadIdx.x] : (threadIdx.x ? n : 0u);
    if (threadIdx.x >= 64 && threadIdx.x < 68) s_bits[threadIdx.x - 64] = threadIdx.x < 66 ? 0u : 0xFFFFFFFFu;
    __syncthreads();
    const uint32_t s = s_span[0], m = s_span[1] > s ? s_span[1] - s : 0u;
    if (m == 0) return;
    const K* kin = keys_in + s;
    const uint32_t* vin = vals_in + s;
    K* kout = keys_out + s;
    uint32_t* vout = vals_out + s;

    // digits in which the span's keys differ
    uint64_t any = 0, all = ~0ull;
    for (uint32_t i = threadIdx.x; i < m && m > (uint32_t)kLocCap; i += kRsThreads) {
        const uint64_t k = (uint64_t)kin[i];
        any |= k;
        all &= k;
    }
    K key[kLocRounds];
    uint32_t val[kLocRounds], dig[kLocRounds], rank[kLocRounds];
    if (m <= (uint32_t)kLocCap) {
#pragma unroll
        for (int r = 0; r < kLocRounds; r++) {
            const uint32_t i = (uint32_t)(wave * kLocWaveKeys + r * 64 + lane);
            key[r] = i < m ? kin[i] : (K)0;
            val[r] = i < m ? vin[i] : 0u;
            if (i < m) {
                any |= (uint64_t)key[r];
                all &= (uint64_t)key[r];
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        any |= __shfl_xor(any, off, 64);
        all &= __shfl_xor(all, off, 64);
    }
    if (lane == 0) {
        atomicOr(&s_bits[0], (uint32_t)any);
        atomicOr(&s_bits[1], (uint32_t)(any >> 32));
        atomicAnd(&s_bits[2], (uint32_t)all);
        atomicAnd(&s_bits[3], (uint32_t)(all >> 32));
    }
    __syncthreads();
    const uint64_t differ = (((uint64_t)s_bits[1] << 32) | s_bits[0]) ^ (((uint64_t)s_bits[3] << 32) | s_bits[2]);

    if (m <= (uint32_t)kLocCap) {
        for (uint32_t p = 0; p < kpasses; p++) {
            const uint32_t shift = p * 8;
            if (((differ >> shift) & 0xFFu) == 0) continue;  // one digit value: the pass keeps the order
            local_rank<K>(key, m, shift, dig, rank, s_cnt, s_hist, wave, lane);
            digit_starts(s_hist, s_start, wave, lane);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kLocRounds; r++) {
                const uint32_t i = (uint32_t)(wave * kLocWaveKeys + r * 64 + lane);
                if (i < m) {
                    const uint32_t pos = s_start[dig[r]] + s_cnt[wave][dig[r]] + rank[r];
                    if (pos < (uint32_t)kLocCap) {
                        s_k[pos] = key[r];
                        s_v[pos] = val[r];
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kLocRounds; r++) {
                const uint32_t i = (uint32_t)(wave * kLocWaveKeys + r * 64 + lane);
                if (i < m) {
                    key[r] = s_k[i];
                    val[r] = s_v[i];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kLocRounds; r++) {
            const uint32_t i = (uint32_t)(wave * kLocWaveKeys + r * 64 + lane);
            if (i < m) {
                kout[i] = key[r];
                vout[i] = val[r];
            }
        }
        return;
    }

    // A span longer than the LDS: chunked passes through global memory,
    // ping-ponging between the span's input and output ranges.
    K* ka[2] = {const_cast<K*>(kin), kout};
    uint32_t* va[2] = {const_cast<uint32_t*>(vin), vout};
    int src = 0;
    for (uint32_t p = 0; p < kpasses; p++) {
        const uint32_t shift = p * 8;
        if (((differ >> shift) & 0xFFu) == 0) continue;
        const K* ks = ka[src];
        const uint32_t* vs = va[src];
        K* kd = ka[src ^ 1];
        uint32_t* vd = va[src ^ 1];
        for (uint32_t d = threadIdx.x; d < kRadixMax; d += kRsThreads) s_hist[d] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < m; i += kRsThreads) atomicAdd(&s_hist[digit_of(ks[i], shift, 0xFFu)], 1u);
        __syncthreads();
        digit_starts(s_hist, s_start, wave, lane);
        __syncthreads();
        // s_start[d]: where the next chunk's keys of digit d go (span-relative)
        for (uint32_t c0 = 0; c0 < m; c0 += kLocCap) {
            const uint32_t cm = m - c0 < (uint32_t)kLocCap ? m - c0 : (uint32_t)kLocCap;
#pragma unroll
            for (int r = 0; r < kLocRounds; r++) {
                const uint32_t i = (uint32_t)(wave * kLocWaveKeys + r * 64 + lane);
                key[r] = i < cm ? ks[c0 + i] : (K)0;
                val[r] = i < cm ? vs[c0 + i] : 0u;
            }
            local_rank<K>(key, cm, shift, dig, rank, s_cnt, s_hist, wave, lane);
#pragma unroll
            for (int r = 0; r < kLocRounds; r++) {
                const uint32_t i = (uint32_t)(wave * kLocWaveKeys + r * 64 + lane);
                if (i < cm) {
                    const uint32_t pos = s_start[dig[r]] + s_cnt[wave][dig[r]] + rank[r];
                    if (pos < m) {
                        kd[pos] = key[r];
                        vd[pos] = val[r];
                    }
                }
            }
            __syncthreads();
            for (uint32_t d = threadIdx.x; d < kRadixMax; d += kRsThreads) s_start[d] += s_hist[d];
            __syncthreads();
        }
        src ^= 1;
    }
    if (src == 0) {  // an even number of passes ran: the sorted span is in the input range
        for (uint32_t i = threadIdx.x; i < m; i += kRsThreads) {
            kout[i] = kin[i];
            vout[i] = vin[i];
        }
    }
}

template <typename K>
static int sort_passes(const SortBufs& b, uint32_t n, uint32_t key_bits, uint32_t bits, uint32_t seg_passes,
                       SegTable seg, int in, int* out_slot, uint32_t* passes_out) {
    const uint32_t num_tiles = (n + kRsTile - 1) / kRsTile;
    const uint32_t radix = 1u << bits;
    uint32_t* hist = b.hist;
    uint32_t* totals = hist + (size_t)radix * num_tiles;
    const uint32_t kpasses = (key_bits + bits - 1) / bits;
    int cur = in;
    // the two top digits as global passes, then spans of equal top bits sorted locally
    const bool local = b.os && seg_passes == 0 && bits == 8 && kpasses >= 3;
    if (local && n <= (uint32_t)kLocCap) {
        hipLaunchKernelGGL((k_rs_local<K>), dim3(1), dim3(kRsThreads), 0, b.stream, (const K*)b.keys[cur],
                           b.vals[cur], (K*)b.keys[cur ^ 1], b.vals[cur ^ 1], n, kpasses, (const uint32_t*)nullptr);
        SQ_CHECK_HIP(hipGetLastError());
        if (passes_out) *passes_out = kpasses;
        *out_slot = cur ^ 1;
        return SQ_OK;
    }
    const uint32_t p_first = local ? kpasses - 2 : 0;
    if (b.os) {
        // single-pass key passes: zero the area, count the passes' digits once
        const size_t words = os_status() + (size_t)kpasses * num_tiles * radix;
        SQ_TRY(launch_fill_u32_on(b.stream, b.num_cus, b.os, 0u, words));
        uint32_t gb = (n + kRsThreads * 16 - 1) / (kRsThreads * 16);
        if (gb > (uint32_t)b.num_cus * 4) gb = (uint32_t)b.num_cus * 4;
        hipLaunchKernelGGL((k_rs_ghist<K>), dim3(gb), dim3(kRsThreads), 0, b.stream, (const K*)b.keys[cur], n, bits,
                           p_first, kpasses, b.os + os_ghist());
        SQ_CHECK_HIP(hipGetLastError());
    }
    for (uint32_t p = p_first; p < kpasses + seg_passes; p++) {
        const bool segp = p >= kpasses;
        if (b.os && !segp) {
            const K* kin = (const K*)b.keys[cur];
            hipLaunchKernelGGL((k_rs_onesweep<K>), dim3(num_tiles), dim3(kRsThreads), 0, b.stream, kin, b.vals[cur],
                               (K*)b.keys[cur ^ 1], b.vals[cur ^ 1], n, p * bits, bits,
                               b.os + os_ghist() + (size_t)p * kRadixMax, b.os + os_tickets() + p,
                               b.os + os_status() + (size_t)p * num_tiles * radix, b.os_err);
            SQ_CHECK_HIP(hipGetLastError());
            cur ^= 1;
            continue;
        }
        const uint32_t shift = (segp ? p - kpasses : p) * bits;
        const K* kin = (const K*)b.keys[cur];
        const uint32_t* vin = b.vals[cur];
        K* kout = (K*)b.keys[cur ^ 1];
        uint32_t* vout = b.vals[cur ^ 1];
        if (segp)
            hipLaunchKernelGGL((k_rs_upsweep<K, true>), dim3(num_tiles), dim3(kRsThreads), 0, b.stream, kin, vin, n,
                               shift, bits, hist, num_tiles, seg);
        else
            hipLaunchKernelGGL((k_rs_upsweep<K, false>), dim3(num_tiles), dim3(kRsThreads), 0, b.stream, kin, vin, n,
                               shift, bits, hist, num_tiles, seg);
        hipLaunchKernelGGL(k_scan_rows, dim3(radix), dim3(kScanThreads), 0, b.stream, hist, num_tiles, totals);
        hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kScanThreads), 0, b.stream, totals, radix);
        if (segp)
            hipLaunchKernelGGL((k_rs_downsweep<K, true>), dim3(num_tiles), dim3(kRsThreads), 0, b.stream, kin, vin,
                               kout, vout, n, shift, bits, hist, totals, num_tiles, seg);
        else
            hipLaunchKernelGGL((k_rs_downsweep<K, false>), dim3(num_tiles), dim3(kRsThreads), 0, b.stream, kin, vin,
                               kout, vout, n, shift, bits, hist, totals, num_tiles, seg);
        SQ_CHECK_HIP(hipGetLastError());
        cur ^= 1;
    }
    if (local) {
        // span table in the (otherwise unused) three-kernel histogram area
        const uint32_t grid = (uint32_t)(((uint64_t)n + kLocSpan - 1) / kLocSpan);
        hipLaunchKernelGGL((k_rs_spans<K>), dim3(grid / 4 + 1), dim3(256), 0, b.stream, (const K*)b.keys[cur], n,
                           p_first * bits, grid, hist);
        hipLaunchKernelGGL((k_rs_local<K>), dim3(grid), dim3(kRsThreads), 0, b.stream, (const K*)b.keys[cur],
                           b.vals[cur], (K*)b.keys[cur ^ 1], b.vals[cur ^ 1], n, kpasses, (const uint32_t*)hist);
        SQ_CHECK_HIP(hipGetLastError());
        cur ^= 1;
    }
    if (passes_out) *passes_out = kpasses + seg_passes;
    *out_slot = cur;
    return SQ_OK;
}

size_t sort_hist_bytes(uint32_t n, uint32_t digit_bits) {
    const size_t num_tiles = (n + kRsTile - 1) / kRsTile;
    return ((size_t)1 << digit_bits) * num_tiles * 4 + kRadixMax * 4;
}

size_t sort_onesweep_bytes(uint32_t n, uint32_t key_bits, uint32_t digit_bits) {
    const size_t num_tiles = (n + kRsTile - 1) / kRsTile;
    const size_t kpasses = (key_bits + digit_bits - 1) / digit_bits;
    return (os_status() + kpasses * num_tiles * ((size_t)1 << digit_bits)) * 4;
}

uint32_t sort_segment_passes(uint32_t nseg, uint32_t digit_bits) {
    uint32_t segbits = 0;
    while ((1u << segbits) < nseg) segbits++;
    return (segbits + digit_bits - 1) / digit_bits;
}

int radix_sort_bufs(const SortBufs& b, uint32_t n, uint32_t key_bits, uint32_t digit_bits, bool key64,
                    const uint32_t* d_seg_off, uint32_t nseg, int in, int* out_slot, uint32_t* passes) {
    if (n == 0) { *out_slot = in; if (passes) *passes = 0; return SQ_OK; }
    if (digit_bits < 1 || digit_bits > 8) SQ_RET(SQ_ERR_INVALID, "radix digit bits %u outside 1..8", digit_bits);
    if (nseg > kMaxSegs) SQ_RET(SQ_ERR_INVALID, "radix sort: %u segments exceed %u", nseg, kMaxSegs);
    SegTable seg{d_seg_off, nseg};
    const uint32_t sp = (d_seg_off && nseg > 1) ? sort_segment_passes(nseg, digit_bits) : 0;
    if (key64) return sort_passes<uint64_t>(b, n, key_bits, digit_bits, sp, seg, in, out_slot, passes);
    return sort_passes<uint32_t>(b, n, key_bits, digit_bits, sp, seg, in, out_slot, passes);
}

int radix_sort(sq_ctx* c, uint32_t n, uint32_t key_bits, uint32_t digit_bits, bool key64, int in, int* out_slot) {
    if (n == 0) { *out_slot = in; return SQ_OK; }
    if (key_bits > kMaxKeyPasses) SQ_RET(SQ_ERR_INVALID, "radix sort: %u key bits", key_bits);
    SQ_TRY(c->hist.ensure(sort_hist_bytes(n, digit_bits)));
    SortBufs b{c->stream, {c->keys[0].p, c->keys[1].p}, {c->vals[0].as<uint32_t>(), c->vals[1].as<uint32_t>()},
               c->hist.as<uint32_t>(), nullptr, nullptr, c->num_cus};
    // single-pass key passes while prefix counts fit the 30-bit status words
    const bool os = n < (1u << 30);
    if (os) {
        SQ_TRY(c->sort_os.ensure(sort_onesweep_bytes(n, key_bits, digit_bits)));
        b.os = c->sort_os.as<uint32_t>();
        SQ_TRY(sort_error_word(c, &b.os_err));  // checked after the call's final sync
    }
    uint32_t passes = 0;
    SQ_TRY(radix_sort_bufs(b, n, key_bits, digit_bits, key64, nullptr, 0, in, out_slot, &passes));
    c->last.sort_passes = passes;
    return SQ_OK;
}

}  // namespace sq
