// Primary-ray sphere trace and distance probe kernels.
//
// The primary trace runs in two kernels:
//   k_march  FP32 march of every camera ray (march_segment, pathtrace.wgsl:78-145) as a
//            lane-persistent state machine; hits are queued;
//   k_hits   for every queued hit, the tetrahedral normal (common.wgsl:376-387, FP32) and
//            the FP64 refine of the hit distance.
// The march reads nothing from HBM: the job's constants sit in the kernel argument segment and
// each ray is derived from (pixel, sample, seed) by the PCG stream; HBM traffic is the output
// planes and the hit queue (about 32 bytes per ray).
#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"

namespace syf {

// Pixel of a ray index in tile-major order (8x8 tiles, tiles row-major): a wave's batch of
// consecutive indices covers neighbouring pixels.
__device__ __forceinline__ void ray_pixel(uint32_t idx, uint32_t tiles_x, uint32_t& px, uint32_t& py) {
    const uint32_t tile = idx >> 6, l = idx & 63u;
    px = (tile % tiles_x) * 8u + (l & 7u);
    py = (tile / tiles_x) * 8u + (l >> 3);
}

// Rays a wave claims at a time: two 8x8 tiles, one atomic per claim.  The tiles are dealt to
// eight counters, one per XCD (a workgroup uses counter blockIdx % 8, the XCD the dispatcher
// places it on), so the claims of the chip's 8 L2 slices never contend for one address.
constexpr uint32_t RAY_CHUNK = 128;
constexpr uint32_t RAY_LANES = 8;
// Hits are staged per wave in LDS and published to the hit queue 64 at a time.
constexpr uint32_t HIT_BATCH = 64;
// A wave generates new camera rays when at least this many of its lanes are idle (or none is
// marching): ray generation is a branch every lane of the wave pays for, so it is batched.
constexpr uint32_t REFILL_BATCH = 8;

__device__ __forceinline__ uint64_t lanes_below() {
    const uint32_t lane = threadIdx.x & 63u;
    return lane == 0u ? 0ull : (~0ull >> (64u - lane));
}

// FP32 march of all primary rays.  Every trip of the loop runs at most one Iterate body on
// each lane; a lane whose field evaluation completes in the trip applies the fence and
// advances its march (the arithmetic of march(), in the same order), queueing a hit or writing
// a miss when the ray ends.  Idle lanes take new rays from a wave-owned chunk of the launch's
// ray counter.  Deep and shallow evaluations, long and short rays share a wave without waiting
// for each other.
// Global ray index of slot j of ray lane `xl` (tiles dealt round-robin to the lanes).
__device__ __forceinline__ uint32_t lane_ray(uint32_t j, uint32_t xl) {
    return (((j >> 6) * RAY_LANES + xl) << 6) | (j & 63u);
}

template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_march(const JobP P, uint32_t sample, uint32_t tiles_x, uint32_t n_slots,
                                               uint32_t* __restrict__ next, float* __restrict__ depth,
                                               float4* __restrict__ aux, uint32_t* __restrict__ steps_out,
                                               float* __restrict__ t_prev_out, uint32_t* __restrict__ hit_list,
                                               uint32_t* __restrict__ hit_count, unsigned long long* __restrict__ ctr) {
    enum : uint32_t { IDLE = 0, MARCH = 1, DONE = 2 };
    __shared__ uint32_t hit_stage[4][2 * HIT_BATCH];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t xl = blockIdx.x % RAY_LANES;
    const uint32_t lane_slots = ((n_slots >> 6) + RAY_LANES - 1u) / RAY_LANES * 64u;  // slots per ray lane
    uint32_t staged = 0;  // wave-uniform count of staged hits
    const FractalP& F = P.f;
    const float safety = F.safety;
    const float hit_cone = P.pixel_cone * P.pixel_tolerance;
    uint32_t mode = IDLE;
    V3<float> ro = mk(0.0f, 0.0f, 0.0f), rd = ro;
    float t = 0.0f, omega = 1.2f, prev_radius = 0.0f, step_len = 0.0f, cand_t = 0.0f, cand_err = 1e30f, t_last = 0.0f;
    uint32_t i = 0, pix = 0;
    DeStepper<FAM> de;
    bool de_done = false;
    Work w;
    uint32_t hits = 0, rays = 0;
    // the wave's chunk of ray indices [chunk_next, chunk_end) (wave-uniform)
    uint32_t chunk_next = 0, chunk_end = 0;
    bool exhausted = false;
    for (;;) {
        const bool need = mode == IDLE;
        const uint64_t mask = __ballot(need);
        const uint32_t idle = uint32_t(__popcll(mask));
        const uint64_t marching = __ballot(mode == MARCH);
        if (idle && (idle >= REFILL_BATCH || !marching)) {
            if (chunk_next == chunk_end && !exhausted) {
                uint32_t base = 0;
                if (lane == 0u) base = atomicAdd(&next[xl * 16u], RAY_CHUNK);
                base = __shfl(base, 0, 64);
                if (base >= lane_slots) {
                    exhausted = true;
                } else {
                    chunk_next = base;
                    chunk_end = min(base + RAY_CHUNK, lane_slots);
                }
            }
            const uint32_t avail = chunk_end - chunk_next;
            if (need) {
                const uint32_t rank = uint32_t(__popcll(mask & lanes_below()));
                if (rank < avail) {
                    uint32_t px, py;
                    const uint32_t idx = lane_ray(chunk_next + rank, xl);
                    ray_pixel(idx, tiles_x, px, py);
                    if (idx < n_slots && px < P.width && py < P.height) {
                        pix = py * P.width + px;
                        uint32_t rng = rng_init(px, py, sample, P.seed);
                        camera_ray(P.cam, P.width, P.height, px, py, rng, ro, rd);
                        t = 0.0f;
                        omega = 1.2f;
                        prev_radius = 0.0f;
                        step_len = 0.0f;
                        cand_t = 0.0f;
                        cand_err = 1e30f;
                        t_last = 0.0f;
                        i = 0;
                        mode = MARCH;
                        ++rays;
                        de_done = de.start(F, ro);
                    }
                } else if (exhausted) {
                    mode = DONE;
                }
            }
            chunk_next += idle < avail ? idle : avail;
        }
        if (__ballot(mode != DONE) == 0ull) break;
        const bool active = mode == MARCH;
        if (active && !de_done) de_done = de.step(F);
        bool hit_now = false;
        if (active && de_done) {
            const De<float> s = de.finish(F);
            w.de_evals++;
            w.de_iters += s.runs;
            w.steps++;
            const uint32_t nsteps = i + 1u;  // march() reports i + 1 for the step just evaluated
            const float radius = fabsf(s.d) * safety;
            const bool sor_fail = omega > 1.0f && (radius + prev_radius) < step_len * 1.0001f;
            if (sor_fail) {
                step_len = step_len - omega * step_len;
                omega = 1.0f;
            } else {
                step_len = radius * omega;
            }
            prev_radius = radius;
            const float err = radius / fmaxf(t, 1e-6f);
            if (!sor_fail && err < cand_err) {
                cand_t = t;
                cand_err = err;
            }
            uint32_t outcome = 0;  // 0 continue, 1 hit, 2 miss
            float t_hit = t, tp_hit = t_last;
            if (!sor_fail && err < hit_cone) {
                outcome = 1;
            } else {
                if (!sor_fail && t + step_len >= P.t_end) {
                    if (radius >= P.t_end - t) outcome = 2;
                    omega = 1.0f;
                    step_len = radius;
                }
                if (outcome == 0u) {
                    t_last = t;
                    t = t + step_len;
                    ++i;
                    if (i >= P.max_steps) {
                        if (cand_err < hit_cone * 4.0f) {
                            outcome = 1;
                            t_hit = cand_t;
                            tp_hit = cand_t;  // best-candidate hit: no refine bracket
                        } else {
                            outcome = 2;
                        }
                    } else if (t >= P.t_end) {
                        outcome = 2;
                    }
                }
            }
            if (outcome == 0u) {
                de_done = de.start(F, add(ro, scl(rd, t)));
            } else {
                steps_out[pix] = nsteps;
                if (outcome == 1u) {
                    depth[pix] = t_hit;
                    t_prev_out[pix] = tp_hit;
                    hit_now = true;
                    ++hits;
                } else {
                    depth[pix] = -1.0f;
                    aux[pix] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
                }
                mode = IDLE;
            }
        }
        // stage this trip's hits in LDS; publish them 64 at a time (one atomic per batch)
        const uint64_t hm = __ballot(hit_now);
        if (hm) {
            if (hit_now) hit_stage[wave][staged + uint32_t(__popcll(hm & lanes_below()))] = pix;
            staged += uint32_t(__popcll(hm));
            if (staged >= HIT_BATCH) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                uint32_t base = 0;
                if (lane == 0u) base = atomicAdd(hit_count, HIT_BATCH);
                base = __shfl(base, 0, 64);
                hit_list[base + lane] = hit_stage[wave][lane];
                const uint32_t rest = staged - HIT_BATCH;
                const uint32_t carry = lane < rest ? hit_stage[wave][HIT_BATCH + lane] : 0u;
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if (lane < rest) hit_stage[wave][lane] = carry;
                staged = rest;
            }
        }
    }
    if (staged) {  // the wave's last hits
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        uint32_t base = 0;
        if (lane == 0u) base = atomicAdd(hit_count, staged);
        base = __shfl(base, 0, 64);
        if (lane < staged) hit_list[base + lane] = hit_stage[wave][lane];
    }
    flush_work(w, hits, rays, ctr);
    const uint32_t pr = wave_sum(rays);
    if (lane == 0u) atomicAdd(&ctr[CTR_PRIMARY], (unsigned long long)pr);
}

// Normal and FP64 refine of every queued hit.  The camera ray is regenerated from its PCG
// stream (the FP32 ray the march followed); the normal takes four FP32 taps at the hit point
// (common.wgsl:376-387, h = max(t * cone / 2, 1e-4) * pixel_tolerance); the refine locates the
// hit rule's crossing in FP64 between the march's last two positions.  Every lane holds a hit,
// and all four taps of a lane sit at the same depth of the field.
template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_hits(const JobP P, uint32_t sample, const uint32_t* __restrict__ hit_list,
                                              const uint32_t* __restrict__ hit_count, const float* __restrict__ depth,
                                              const float* __restrict__ t_prev, const uint32_t* __restrict__ steps,
                                              float4* __restrict__ aux, double* __restrict__ depth64,
                                              unsigned long long* __restrict__ ctr) {
    const uint32_t n = *hit_count;
    Work w, wn;  // refine; normal taps (counted apart from the march's evaluations)
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t pix = hit_list[k];
        const uint32_t px = pix % P.width, py = pix / P.width;
        uint32_t rng = rng_init(px, py, sample, P.seed);
        V3<float> ro, rd;
        camera_ray(P.cam, P.width, P.height, px, py, rng, ro, rd);
        const float t = depth[pix];
        const V3<float> p = add(ro, scl(rd, t));
        const V3<float> nv = normal_at<float, FAM>(P.f, p, fmaxf(t * P.pixel_cone * 0.5f, 1e-4f) * P.pixel_tolerance, wn);
        aux[pix] = make_float4(nv.x, nv.y, nv.z, float(steps[pix]));
        depth64[pix] = refine64<FAM>(P.f, cvt<double>(ro), cvt<double>(rd), double(t_prev[pix]), double(t),
                                     double(P.pixel_cone) * double(P.pixel_tolerance), w);
    }
    flush_work(w, 0u, 0u, ctr);
    const uint32_t a = wave_sum(wn.de_evals), b = wave_sum(wn.de_iters);
    if ((threadIdx.x & 63u) == 0u) {
        atomicAdd(&ctr[CTR_HIT_EVALS], (unsigned long long)a);
        atomicAdd(&ctr[CTR_HIT_ITERS], (unsigned long long)b);
    }
}

// Misses of the FP64 plane: -1 wherever the march found no hit.
__global__ __launch_bounds__(256) void k_depth64_init(const float* __restrict__ depth, double* __restrict__ depth64,
                                                      uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (depth[i] < 0.0f) depth64[i] = -1.0;
}

template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_de_probe(const FractalP F, const float4* __restrict__ pts, uint32_t n,
                                                  float4* __restrict__ out, float4* __restrict__ traps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 q = pts[i];
    De<float> s = de_eval<float, FAM>(F, mk(q.x, q.y, q.z));
    out[i] = make_float4(s.d, s.trap0, float(s.iter), 0.0f);
    traps[i] = make_float4(s.trap0, 1.0f, 1.0f, 1.0f);
}

#define SYF_DISPATCH_FAMILY(fam, KERNEL, ...)                                                      \
    switch (fam) {                                                                                 \
    case SYF_FAMILY_MANDELBULB: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_MANDELBULB>, __VA_ARGS__); break; \
    case SYF_FAMILY_MANDELBOX: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_MANDELBOX>, __VA_ARGS__); break;   \
    case SYF_FAMILY_MENGER: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_MENGER>, __VA_ARGS__); break;         \
    case FAM_BULB_TRIG: hipLaunchKernelGGL(KERNEL<FAM_BULB_TRIG>, __VA_ARGS__); break;                 \
    case FAM_BULB8: hipLaunchKernelGGL(KERNEL<FAM_BULB8>, __VA_ARGS__); break;                         \
    default: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_SPHERE>, __VA_ARGS__); break;                        \
    }

hipError_t launch_trace_primary(const JobP& P, uint32_t sample, const TraceBuffers& B, int cu_count, hipStream_t s) {
    const uint32_t tiles_x = (P.width + 7u) / 8u, tiles_y = (P.height + 7u) / 8u;
    const uint32_t n_slots = tiles_x * tiles_y * 64u;
    // Two workgroups (eight waves) per CU.  A frame's time is set by its longest rays (up to
    // max_steps evaluations in sequence on one lane); more resident waves share each SIMD's
    // issue slots with them and stretch that critical path: measured at 1080p, 2 per CU
    // 0.80 ms, 1 per CU 1.00 ms, 3 per CU 0.88 ms, 8 per CU 1.65 ms.
    const uint32_t fill = uint32_t(cu_count > 0 ? cu_count : 256) * 2u;
    // at least one workgroup per ray lane, or that lane's tiles would never be traced
    const dim3 grid(std::max<uint32_t>(std::min<uint32_t>((n_slots + 255u) / 256u, fill), RAY_LANES)), block(256);
    SYF_DISPATCH_FAMILY(kernel_family(P.f), k_march, grid, block, 0, s, P, sample, tiles_x, n_slots, B.counters,
                        B.depth, B.aux, B.steps, B.t_prev, B.hit_list, B.counters + COUNTER_HITS, B.ctr);
    return hipGetLastError();
}

hipError_t launch_refine64(const JobP& P, uint32_t sample, const TraceBuffers& B, int cu_count, hipStream_t s) {
    const uint32_t n = P.width * P.height;
    // two workgroups per CU, each lane striding over several hits: measured 2.3x faster than
    // one hit per lane (eight workgroups per CU) at 1080p; a lane's successive hits lie far
    // apart, so the expensive refines of silhouette regions spread over all lanes
    const uint32_t fill = uint32_t(cu_count > 0 ? cu_count : 256) * 2u;
    const dim3 grid(std::min<uint32_t>((n + 255u) / 256u, fill)), block(256);
    hipLaunchKernelGGL(k_depth64_init, grid, block, 0, s, B.depth, B.depth64, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    SYF_DISPATCH_FAMILY(kernel_family(P.f), k_hits, grid, block, 0, s, P, sample, B.hit_list, B.counters + COUNTER_HITS, B.depth,
                        B.t_prev, B.steps, B.aux, B.depth64, B.ctr);
    return hipGetLastError();
}

hipError_t launch_de_probe(const FractalP& F, const float4* pts, uint32_t n, float4* out, float4* traps,
                           hipStream_t s) {
    const dim3 grid((n + 255u) / 256u), block(256);
    SYF_DISPATCH_FAMILY(kernel_family(F), k_de_probe, grid, block, 0, s, F, pts, n, out, traps);
    return hipGetLastError();
}

}  // namespace syf
