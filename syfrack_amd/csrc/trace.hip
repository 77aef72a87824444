// Primary-ray sphere trace and distance probe kernels.
//
// The primary trace runs in one persistent kernel and a short trailing pass:
//   k_primary    FP32 march of every camera ray (march_segment, pathtrace.wgsl:78-145) as a
//                lane-persistent state machine; hits go to a hit queue.  The job's march_math
//                picks the FP32 arithmetic: SYF_MATH_FAST (DeStepperFast: hardware reciprocal,
//                reciprocal square root and logarithm, FMA; the default) or SYF_MATH_EXACT
//                (k_primary_exact: the IEEE operations of the oracle's FP32 twin).  A wave whose lanes
//                have no rays left lowers its issue priority and takes queued hits: the
//                tetrahedral normal (common.wgsl:376-387, FP32) and the FP64 refine of the hit
//                distance.  The hit work thus fills the SIMDs the march's tail leaves idle
//                instead of running after it.
//   k_hits_tail  the hits no wave took (blocks claimed before they were complete, or queued
//                after the last consumer left), then empties every queue entry the trace used.
// The march reads nothing from HBM: the job's constants sit in the kernel argument segment and
// each ray is derived from (pixel, sample, seed) by the PCG stream; HBM traffic is the output
// planes (32 bytes per pixel) and the hit queue (24 bytes per hit, written and read back).
#include <algorithm>
#include <cstdlib>

#include "trace_kernels.hpp"

namespace syf {


// Workgroups of a primary-trace launch.  Two workgroups (eight waves) per CU: a frame's time is
// set by its longest rays (up to max_steps evaluations in sequence on one lane); more resident
// waves share each SIMD's issue slots with them and stretch that critical path (measured at
// 1080p, march alone: 2 per CU 0.80 ms, 1 per CU 1.00 ms, 3 per CU 0.88 ms, 8 per CU 1.65 ms).
// At least one workgroup per ray lane, or that lane's tiles would never be traced.
// Workgroups per CU of a primary-trace launch: 2, or SYF_PRIMARY_WG_PER_CU (1 to 4) from the
// environment, read once (a tuning override; the planes do not depend on it).
static uint32_t primary_wg_per_cu() {
    static const uint32_t v = [] {
        const char* e = std::getenv("SYF_PRIMARY_WG_PER_CU");
        const long n = e ? std::strtol(e, nullptr, 10) : 0;
        return uint32_t(n >= 1 && n <= 4 ? n : 2);
    }();
    return v;
}

static uint32_t primary_blocks(uint32_t width, uint32_t height, int cu_count) {
    const uint32_t n_slots = ((width + 7u) / 8u) * ((height + 7u) / 8u) * 64u;
    const uint32_t fill = uint32_t(cu_count > 0 ? cu_count : 256) * primary_wg_per_cu();
    return std::max<uint32_t>(std::min<uint32_t>((n_slots + 255u) / 256u, fill), RAY_LANES);
}

size_t hitq_capacity(uint32_t width, uint32_t height, int cu_count) {
    // every pixel, plus one padded block per wave (4 waves per workgroup)
    const size_t pixels = size_t(width) * height;
    return (pixels + HIT_BATCH - 1) / HIT_BATCH * HIT_BATCH + size_t(primary_blocks(width, height, cu_count)) * 4u * HIT_BATCH;
}

hipError_t launch_trace_primary(const JobP& P, uint32_t sample, const TraceBuffers& B, int cu_count, hipStream_t s) {
    const uint32_t tiles_x = (P.width + 7u) / 8u, tiles_y = (P.height + 7u) / 8u;
    const uint32_t n_slots = tiles_x * tiles_y * 64u;
    const dim3 grid(primary_blocks(P.width, P.height, cu_count)), block(256);
    hipError_t e = hipSuccess;
    if (P.march_math == SYF_MATH_EXACT) {
        SYF_DISPATCH(e, P.f, k_primary_exact, grid, block, 0, s, P, sample, tiles_x, n_slots, B.counters, B.depth,
                     B.aux, B.depth64, B.steps, B.q, B.ctr);
    } else {
        SYF_DISPATCH(e, P.f, k_primary, grid, block, 0, s, P, sample, tiles_x, n_slots, B.counters, B.depth, B.aux,
                     B.depth64, B.steps, B.q, B.ctr);
    }
    return e;
}

hipError_t launch_hits_tail(const JobP& P, uint32_t sample, const TraceBuffers& B, int cu_count, hipStream_t s) {
    // two workgroups per CU, each lane striding over several hits: measured 2.3x faster than
    // one hit per lane (eight workgroups per CU) for a whole 1080p frame's hits
    const uint32_t n = P.width * P.height;
    const uint32_t fill = uint32_t(cu_count > 0 ? cu_count : 256) * 2u;
    const dim3 grid(std::min<uint32_t>((n + 255u) / 256u, fill)), block(256);
    hipError_t e = hipSuccess;
    SYF_DISPATCH(e, P.f, k_hits_tail, grid, block, 0, s, P, sample, B.q, B.aux, B.depth64, B.ctr);
    return e;
}

hipError_t launch_de_probe(const FractalP& F, const float4* pts, uint32_t n, float4* out, float4* traps,
                           hipStream_t s) {
    const dim3 grid((n + 255u) / 256u), block(256);
    hipError_t e = hipSuccess;
    SYF_DISPATCH(e, F, k_de_probe, grid, block, 0, s, F, pts, n, out, traps);
    return e;
}

}  // namespace syf
