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
hipError_t launch_de_probe(const FractalP& F, const float4* pts, uint32_t n, float4* out, float4* traps,
                           hipStream_t s);

}  // namespace syf
