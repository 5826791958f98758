// hostcopy.cpp — multi-threaded host memcpy for staging large inputs and
// outputs through pinned buffers (one core copies a few GB/s; the serving
// loop moves up to hundreds of MB per wave).
#include <cstddef>
#include <cstdint>
#include <cstring>

#include <omp.h>

namespace sq {

void par_memcpy(void* dst, const void* src, size_t bytes) {
    constexpr size_t kChunk = 4u << 20;
    if (bytes < 2 * kChunk) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const int64_t nchunks = (int64_t)((bytes + kChunk - 1) / kChunk);
    int threads = omp_get_max_threads();
    if (threads > 16) threads = 16;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int64_t c = 0; c < nchunks; c++) {
        const size_t off = (size_t)c * kChunk;
        const size_t n = bytes - off < kChunk ? bytes - off : kChunk;
        std::memcpy((char*)dst + off, (const char*)src + off, n);
    }
}

}  // namespace sq
