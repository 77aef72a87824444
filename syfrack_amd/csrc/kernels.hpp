// Host-side launchers of the library's kernels (each defined next to its kernels).
#pragma once

#include <hip/hip_runtime.h>

#include "params.hpp"

namespace syf {

// Layout of TraceBuffers::counters: eight ray counters 16 words apart, then the hit count.
constexpr uint32_t COUNTER_HITS = 128;
constexpr uint32_t COUNTER_WORDS = 144;

// Device buffers of a primary trace.
struct TraceBuffers {
    float* depth;
    float4* aux;
    double* depth64;
    uint32_t* steps;
    float* t_prev;             // per pixel: march position before the hit (refine bracket)
    uint32_t* hit_list;        // per pixel capacity: pixels that hit, for the FP64 refine
    uint32_t* counters;        // COUNTER_WORDS words: ray counters (one per XCD, 64 B apart), hit_list length
    unsigned long long* ctr;   // work counters (CTR_*)
};

hipError_t launch_trace_primary(const JobP& P, uint32_t sample, const TraceBuffers& B, int cu_count, hipStream_t s);
hipError_t launch_refine64(const JobP& P, uint32_t sample, const TraceBuffers& B, int cu_count, hipStream_t s);
// Work item of the batched serving kernel: an 8x8 tile of request `slot`, samples
// [s0, s0 + ns).
struct BatchItem {
    uint32_t slot, x0, y0, s0, ns;
};
// Device planes of one request.
struct BatchPlanes {
    float4* accum;
    float* depth;
    float4* aux;
    double* depth64;
};
// One persistent launch of the flattened path tracer over work items that all belong to
// kernel family `fam` (kernel_family of their jobs).  claim: 8 counters, 16 words apart
// (zeroed before the launch).  Items starting at sample 0 leave the refine bracket in their
// DEPTH64 plane; launch_refine_first_hits over the same items replaces it by the FP64 hit.
constexpr uint32_t CLAIM_WORDS = 8 * 16;
hipError_t launch_paths(uint32_t fam, const JobP* jobs, const BatchPlanes* planes, const BatchItem* items,
                        uint32_t n_items, uint32_t* claim, unsigned long long* ctr, uint32_t grid_blocks,
                        hipStream_t s);
hipError_t launch_refine_first_hits(uint32_t fam, const JobP* jobs, const BatchPlanes* planes,
                                    const BatchItem* items, uint32_t n_items, unsigned long long* ctr,
                                    hipStream_t s);
hipError_t launch_render_tile(const JobP& P, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t s0,
                              uint32_t ns, float4* accum, float* depth, float4* aux, double* depth64,
                              unsigned long long* ctr, hipStream_t s);

hipError_t launch_de_probe(const FractalP& F, const float4* pts, uint32_t n, float4* out, float4* traps,
                           hipStream_t s);

}  // namespace syf
