// Host-side launchers of the library's kernels (each defined next to its kernels), the device
// structures they share with the kernels, and the family dispatch of a launch.
#pragma once

#include <hip/hip_runtime.h>

#include "params.hpp"

#ifndef __HIPCC_RTC__
#include <tuple>
#include <type_traits>
#include <utility>
#endif

namespace syf {

// Layout of TraceBuffers::counters: eight ray counters 16 words apart, then the hit queue's
// reserved count, its claim cursor and its error flag, 16 words apart.
constexpr uint32_t COUNTER_HITS = 128;
constexpr uint32_t COUNTER_CURSOR = 144;
constexpr uint32_t COUNTER_QERR = 160;
constexpr uint32_t COUNTER_WORDS = 176;
// One allocation holds the counters and, after them, the CTR_COUNT work counters.
constexpr size_t COUNTER_BYTES = COUNTER_WORDS * sizeof(uint32_t) + CTR_COUNT * sizeof(unsigned long long);

// Hit queue of a primary trace: HITQ_FIELDS fields per entry in separate arrays of `capacity`
// words (room for every pixel plus one 64-entry block of padding per wave), every word
// HITQ_EMPTY before the launch (the trailing pass empties the entries it covered); one done
// flag per 64-entry block, which counts only when it holds the launch's epoch.
constexpr uint32_t HITQ_EMPTY = 0xffffffffu;
constexpr uint32_t HITQ_SKIP = 0xfffffffeu;  // pixel of a padding entry
constexpr uint32_t HITQ_FIELDS = 6;
struct HitQueue {
    uint32_t* pix;
    uint32_t* t;       // march distance of the hit (float bits)
    uint32_t* t_prev;  // march position before the hit (float bits): lower end of the refine bracket
    uint32_t* steps;
    uint32_t* r_prev;  // the march's radius |d| * safety at t_prev (float bits)
    uint32_t* r_hit;   // the march's radius at t (float bits)
    uint32_t* count;   // entries reserved by the march
    uint32_t* cursor;  // blocks claimed by shading waves
    uint32_t* done;    // per block: `epoch` once a shading wave of the fused kernel has shaded it
    uint32_t* error;   // nonzero if a block did not fit (never at hitq_capacity())
    uint32_t capacity; // entries, a multiple of 64
    uint32_t epoch;    // nonzero, distinct from every stale done flag
};

// Device buffers of a primary trace.
struct TraceBuffers {
    float* depth;
    float4* aux;
    double* depth64;
    uint32_t* steps;
    uint32_t* counters;        // COUNTER_WORDS words: ray counters (one per XCD, 64 B apart), queue count, cursor
    HitQueue q;
    unsigned long long* ctr;   // work counters (CTR_*)
};

#ifndef __HIPCC_RTC__
// Entries of the hit queue a width x height primary trace needs on a device of cu_count CUs.
size_t hitq_capacity(uint32_t width, uint32_t height, int cu_count);
// The fused march + hit-shading kernel, then the trailing pass over the hits it left.
hipError_t launch_trace_primary(const JobP& P, uint32_t sample, const TraceBuffers& B, int cu_count, hipStream_t s);
hipError_t launch_hits_tail(const JobP& P, uint32_t sample, const TraceBuffers& B, int cu_count, hipStream_t s);
// Work item of the batched serving kernel: a tile of 64 pixels of request `slot`, 2^tws wide
// (8 x 8, 16 x 4 or 32 x 2), samples [s0, s0 + ns).
struct BatchItem {
    uint32_t slot, x0, y0, s0, ns, tws;
};
// Device planes of one request.
struct BatchPlanes {
    float4* accum;
    float* depth;
    float4* aux;
    double* depth64;
};
// One persistent launch of the flattened path tracer over work items that all belong to
// kernel family `fam` (kernel_family of their jobs).  claim: 8 counters, 16 words apart; ctr:
// CTR_COUNT work counters; exit_count: one word.  All three are zero before the launch and are
// left zero by it: the launch's last wave copies the work counters to host_out (host memory)
// and clears them.  Items starting at sample 0 begin the pixels' accumulation and leave the
// refine bracket in their DEPTH64 plane; launch_refine_first_hits over the same items replaces
// it by the FP64 hit.
constexpr uint32_t CLAIM_WORDS = 8 * 16;
// `steps`: field evaluations per lane per trip of the kernel's loop (4, 8 or 16; other values
// round up to the next of these).
hipError_t launch_paths(uint32_t fam, const JobP* jobs, const BatchPlanes* planes, const BatchItem* items,
                        uint32_t n_items, uint32_t* claim, unsigned long long* ctr, uint32_t* exit_count,
                        unsigned long long* host_out, uint32_t grid_blocks, uint32_t steps, hipStream_t s);
// Writes slot `slot` of the batched launches' job and plane tables (one small kernel: it runs on
// the stream of the launch that first reads the slot).
hipError_t launch_admit(JobP* jobs, BatchPlanes* planes, uint32_t slot, const JobP& P, const BatchPlanes& pl,
                        hipStream_t s);
hipError_t launch_refine_first_hits(uint32_t fam, const JobP* jobs, const BatchPlanes* planes,
                                    const BatchItem* items, uint32_t n_items, unsigned long long* ctr,
                                    hipStream_t s);
// P: the job's host copy (its family picks the kernel); dP: the same constants in device memory.
// vol: the session planes (VOLUME, MOMENTS, SURFACE) of a job with fog, null otherwise.
hipError_t launch_render_tile(const JobP& P, const JobP* dP, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                              uint32_t s0, uint32_t ns, float4* accum, float* depth, float4* aux, double* depth64,
                              float4* vol, unsigned long long* ctr, hipStream_t s);

hipError_t launch_de_probe(const FractalP& F, const float4* pts, uint32_t n, float4* out, float4* traps,
                           hipStream_t s);

// Kernels of a specialised graph topology (FAM_JIT), compiled at run time by graph.hip: one
// entry per family-templated kernel, in this order.
enum : uint32_t {
    JIT_k_primary,
    JIT_k_primary_exact,
    JIT_k_hits_tail,
    JIT_k_de_probe,
    JIT_k_render_tile,
    JIT_k_wf_intersect,
    JIT_k_wf_shade,
    JIT_k_wf_commit,
    JIT_k_render_tile_full,
    JIT_KERNELS
};
// Launches kernel `which` of the compiled topology `jit` (FractalP::gjit) with its arguments'
// addresses, like hipModuleLaunchKernel.
hipError_t jit_launch_args(const void* jit, uint32_t which, dim3 grid, dim3 block, size_t shmem, hipStream_t s,
                           void** args);
// The arguments converted to the parameter types of `kernel` (the library's instantiation of the
// same template, whose signature the run-time one shares), then launched.
template <class... Params, class... Args>
hipError_t jit_launch_as(void (*)(Params...), const void* jit, uint32_t which, dim3 grid, dim3 block, size_t shmem,
                         hipStream_t s, Args&&... a) {
    static_assert(sizeof...(Params) == sizeof...(Args), "argument count of a run-time kernel launch");
    std::tuple<std::decay_t<Params>...> v(std::forward<Args>(a)...);
    return std::apply(
        [&](auto&... x) {
            void* args[] = {static_cast<void*>(&x)...};
            return jit_launch_args(jit, which, grid, block, shmem, s, args);
        },
        v);
}

// Launches KERNEL<family of F> with hipLaunchKernelGGL's arguments (grid, block, shared memory,
// stream, kernel arguments); a specialised graph topology launches its run-time compiled
// instance.  Sets `err` to the launch status.
#define SYF_DISPATCH(err, F, KERNEL, grid, block, shmem, stream, ...)                                       \
    switch (kernel_family(F)) {                                                                           \
    case SYF_FAMILY_MANDELBULB: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_MANDELBULB>, grid, block, shmem, stream, __VA_ARGS__); err = hipGetLastError(); break; \
    case SYF_FAMILY_MANDELBOX: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_MANDELBOX>, grid, block, shmem, stream, __VA_ARGS__); err = hipGetLastError(); break;   \
    case SYF_FAMILY_MENGER: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_MENGER>, grid, block, shmem, stream, __VA_ARGS__); err = hipGetLastError(); break;         \
    case FAM_BULB_TRIG: hipLaunchKernelGGL(KERNEL<FAM_BULB_TRIG>, grid, block, shmem, stream, __VA_ARGS__); err = hipGetLastError(); break;                 \
    case FAM_BULB8: hipLaunchKernelGGL(KERNEL<FAM_BULB8>, grid, block, shmem, stream, __VA_ARGS__); err = hipGetLastError(); break;                         \
    case FAM_GRAPH: hipLaunchKernelGGL(KERNEL<FAM_GRAPH>, grid, block, shmem, stream, __VA_ARGS__); err = hipGetLastError(); break;                         \
    case FAM_JIT: err = jit_launch_as(&KERNEL<FAM_GRAPH>, (F).gjit, JIT_##KERNEL, grid, block, shmem, stream, __VA_ARGS__); break;                         \
    default: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_SPHERE>, grid, block, shmem, stream, __VA_ARGS__); err = hipGetLastError(); break;                        \
    }
#endif  // __HIPCC_RTC__

}  // namespace syf
