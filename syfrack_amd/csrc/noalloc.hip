// The primary trace without any allocation: syf_trace_primary_noalloc.
//
// syf_trace_primary keeps its hit queue and counters in device memory that syf_job_create
// allocates.  Callers that cannot allocate (code under HIP graph capture, tight serving
// loops) pass their own device planes and stream instead, and the trace's scratch -- the
// counters and the hit queue -- lives in `scratch`, a file-scope device array of fixed size
// that every GPU running the library holds until the process ends.
//
// In host code `scratch` names a host-side stand-in, not device memory.  Its device address is
// taken once per device with hipGetSymbolAddress (noalloc_setup, called by syf_ctx_create with
// the device current) and only that address reaches a kernel.
//
// Every call on one device shares that device's array, and the array keeps the last call's
// contents: the counters are cleared by a fill kernel at each call; the hit queue is filled
// once per device (every field empty, every done flag 0) and each call's trailing pass empties
// the entries it used, with done flags that count only when they hold the call's epoch.  Calls
// that overlap on one device -- on two streams, or from two threads -- race on the array.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "api_internal.hpp"
#include "kernels.hpp"

namespace syf {
namespace {

// Entries of the fixed hit queue: SYF_NOALLOC_MAX_PIXELS rounded to 64, plus one 64-entry
// block for each wave of the widest primary grid (at most 4 workgroups of 4 waves per CU, 256
// CUs).
constexpr uint32_t NA_MAX_CU = 256;
constexpr size_t NA_QUEUE_CAP = size_t(SYF_NOALLOC_MAX_PIXELS) + size_t(NA_MAX_CU) * 4u * 4u * 64u;
constexpr size_t NA_COUNTER_WORDS = 256;  // COUNTER_BYTES (824) rounded up to 1 KiB
static_assert(COUNTER_BYTES <= NA_COUNTER_WORDS * sizeof(uint32_t), "counters do not fit");
static_assert(NA_QUEUE_CAP % 64 == 0, "queue capacity must be whole blocks");
constexpr size_t NA_WORDS = NA_COUNTER_WORDS + HITQ_FIELDS * NA_QUEUE_CAP + NA_QUEUE_CAP / 64;

}  // namespace
}  // namespace syf

// 31.5 MB per device: counters, then six queue fields of NA_QUEUE_CAP words, then done flags.
static __device__ __align__(256) uint32_t scratch[syf::NA_WORDS];

namespace syf {
namespace {

__global__ void k_fill_words(uint32_t* __restrict__ p, size_t n, uint32_t value) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = value;
}

hipError_t fill_words(uint32_t* p, size_t n, uint32_t value, hipStream_t s) {
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_words, dim3(uint32_t(blocks)), dim3(256), 0, s, p, n, value);
    return hipGetLastError();
}

constexpr int NA_MAX_DEVICES = 64;
struct DeviceScratch {
    uint32_t* ptr = nullptr;  // device address of `scratch` on this device
    uint32_t epoch = 0;       // done-flag value of the latest call (never 0 once a call ran)
    bool queue_empty = false; // every queue field empty and no stale flag can match the next epoch
};
std::mutex g_na_mutex;
DeviceScratch g_na[NA_MAX_DEVICES];

}  // namespace

// Looks up this device's address of `scratch`, once per device; the device must be current.
int noalloc_setup(int device, void** out) {
    if (device < 0 || device >= NA_MAX_DEVICES) return fail(SYF_E_NO_DEVICE, "device %d above the library's %d", device, NA_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_na_mutex);
    DeviceScratch& d = g_na[device];
    if (!d.ptr) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(SYF_E_DEVICE, "hipGetSymbolAddress(scratch) returned a null address");
        d.ptr = static_cast<uint32_t*>(p);
    }
    *out = d.ptr;
    return SYF_OK;
}

}  // namespace syf

using namespace syf;

extern "C" int syf_trace_primary_noalloc(syf_ctx* ctx, const syf_job_desc* desc, uint32_t sample, float* depth,
                                         float* aux4, double* depth64, uint32_t* steps, void* hip_stream) {
    clear_error();
    if (!ctx || !desc) return fail(SYF_E_INVALID, "ctx and desc must not be null");
    if (!depth || !aux4 || !depth64 || !steps)
        return fail(SYF_E_INVALID, "depth, aux4, depth64 and steps must be device pointers, not null");
    if ((reinterpret_cast<uintptr_t>(aux4) & 15u) || (reinterpret_cast<uintptr_t>(depth64) & 7u) ||
        (reinterpret_cast<uintptr_t>(depth) & 3u) || (reinterpret_cast<uintptr_t>(steps) & 3u))
        return fail(SYF_E_INVALID, "aux4 must be 16-byte, depth64 8-byte, depth and steps 4-byte aligned");
    JobP P;
    int rc = lower_job(*desc, P);
    if (rc != SYF_OK) return rc;
    const size_t pixels = size_t(desc->width) * desc->height;
    if (pixels > SYF_NOALLOC_MAX_PIXELS)
        return fail(SYF_E_INVALID, "width x height: %ux%u = %zu pixels, above the allocation-free trace's maximum of %u",
                    desc->width, desc->height, pixels, SYF_NOALLOC_MAX_PIXELS);
    const size_t cap_needed = hitq_capacity(desc->width, desc->height, ctx->cu_count);
    if (cap_needed > NA_QUEUE_CAP || ctx->cu_count > int(NA_MAX_CU))
        return fail(SYF_E_INVALID, "a %ux%u trace on %d CUs needs %zu hit-queue entries, above the %zu of the fixed scratch",
                    desc->width, desc->height, ctx->cu_count, cap_needed, NA_QUEUE_CAP);
    if (ctx->device < 0 || ctx->device >= NA_MAX_DEVICES) return fail(SYF_E_NO_DEVICE, "device %d out of range", ctx->device);
    if (hipError_t de = hipSetDevice(ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");

    std::lock_guard<std::mutex> lock(g_na_mutex);
    DeviceScratch& d = g_na[ctx->device];
    if (!d.ptr) return fail(SYF_E_DEVICE, "the scratch of device %d was not set up (syf_ctx_create)", ctx->device);
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    uint32_t* const counters = d.ptr;
    uint32_t* const q = d.ptr + NA_COUNTER_WORDS;
    const size_t cap = NA_QUEUE_CAP;
    TraceBuffers B{depth,
                   reinterpret_cast<float4*>(aux4),
                   depth64,
                   steps,
                   counters,
                   HitQueue{q, q + cap, q + 2 * cap, q + 3 * cap, q + 4 * cap, q + 5 * cap, counters + COUNTER_HITS,
                            counters + COUNTER_CURSOR, q + HITQ_FIELDS * cap, counters + COUNTER_QERR, uint32_t(cap), 0u},
                   reinterpret_cast<unsigned long long*>(counters + COUNTER_WORDS)};
    // counters start at zero; the queue is emptied on the first call on the device and when the
    // epoch wraps (otherwise each call's trailing pass has emptied what it used)
    hipError_t e = fill_words(counters, NA_COUNTER_WORDS, 0u, s);
    if (++d.epoch == 0u) {
        d.epoch = 1u;
        d.queue_empty = false;
    }
    if (!d.queue_empty) {
        if (e == hipSuccess) e = fill_words(q, HITQ_FIELDS * cap, HITQ_EMPTY, s);
        if (e == hipSuccess) e = fill_words(q + HITQ_FIELDS * cap, cap / 64, 0u, s);
    }
    B.q.epoch = d.epoch;
    d.queue_empty = false;
    if (e == hipSuccess) e = launch_trace_primary(P, sample, B, ctx->cu_count, s);
    if (e == hipSuccess) e = launch_hits_tail(P, sample, B, ctx->cu_count, s);
    if (e == hipSuccess) d.queue_empty = true;
    return e == hipSuccess ? SYF_OK : hip_fail(e, "syf_trace_primary_noalloc");
}
