// Host-side launchers of the library's kernels (each defined next to its kernels).
#pragma once

#include <hip/hip_runtime.h>

#include "params.hpp"

namespace syf {

hipError_t launch_trace_primary(const JobP& P, uint32_t sample, float* depth, float4* aux, uint32_t* steps,
                                float* t_prev, unsigned long long* ctr, hipStream_t s);
hipError_t launch_refine64(const JobP& P, uint32_t sample, const float* depth, const float* t_prev, double* depth64,
                           unsigned long long* ctr, hipStream_t s);
hipError_t launch_de_probe(const FractalP& F, const float4* pts, uint32_t n, float4* out, float4* traps,
                           hipStream_t s);

}  // namespace syf
