// noalloc.hip — allocation-free tracing of device ray lists.
//
// sq_trace_rays_noalloc is sq_trace_rays with device inputs and outputs
// whose scratch (tile counter, fault bits, work totals, per-lane stack spill
// areas) is one fixed array per device instead of the context's grow-only
// workspace, so a call allocates nothing: for code under HIP graph capture,
// tight serving loops and harnesses that forbid hipMalloc.
//
// The array is a file-scope __device__ variable. Its host-side name is only
// a stand-in for the symbol, never a device address: noalloc_setup (called
// by sq_ctx_create with the device current) asks the runtime for this
// device's address of it once and keeps that per device; the entry point
// passes only that address to the kernels. Every call on one device shares
// the array and leaves its contents behind; the counters the kernel needs
// at zero are cleared by a kernel at the start of each launch (the spill
// areas need no clearing: a lane reads back only what it pushed). Calls
// that overlap on one device race on the array.
#include <mutex>

#include "trace_common.h"

namespace sq {

constexpr size_t kScratchBytes = SQ_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch size");

static __device__ __align__(256) uint32_t scratch[kScratchBytes / 4];

constexpr int kMaxDevices = 64;
static std::mutex g_mu;
static uint32_t* g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) SQ_RET(SQ_ERR_INVALID, "device %d: at most %d devices", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch_ptr[device]) return SQ_OK;
    int cur = -1;
    SQ_CHECK_HIP(hipGetDevice(&cur));
    if (cur != device) SQ_RET(SQ_ERR_INVALID, "noalloc setup for device %d while device %d is current", device, cur);
    void* ptr = nullptr;
    SQ_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) SQ_RET(SQ_ERR_HIP, "hipGetSymbolAddress returned no address for the trace scratch");
    g_scratch_ptr[device] = (uint32_t*)ptr;
    return SQ_OK;
}

static uint32_t* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return g_scratch_ptr[device];
}

}  // namespace sq

using namespace sq;

extern "C" int sq_trace_rays_noalloc(sq_ctx* c, const sq_bvh* b, const sq_ray* rays, uint32_t n,
                                     const sq_trace_params* p, const sq_trace_out* out) {
    if (!c || !b || !rays || !out) SQ_RET(SQ_ERR_INVALID, "sq_trace_rays_noalloc: NULL argument");
    if (p && p->heatmap != SQ_HEATMAP_RAY)
        SQ_RET(SQ_ERR_UNSUPPORTED, "sq_trace_rays_noalloc: heatmap mode %u", p->heatmap);
    if (n == 0) return SQ_OK;
    uint32_t* base = scratch_ptr(c->device);
    if (!base) SQ_RET(SQ_ERR_INVALID, "sq_trace_rays_noalloc: no scratch address for device %d", c->device);
    const TraceScratch ext{base, kScratchBytes / 4};
    return trace_rays_device(c, b, nullptr, rays, n, p, out, &ext);
}
