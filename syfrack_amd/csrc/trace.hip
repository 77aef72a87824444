// Primary-ray sphere trace and distance probe kernels.
//
// Launch shape: 256-thread workgroups = four wave64s; each wave owns one 8x8 pixel tile, so
// the rays of a wave are spatial neighbours and take similar step counts (the wave's
// divergence is the spread of its 64 rays).  The march reads nothing from HBM: the job's
// constants sit in the kernel argument segment and every ray is derived from
// (pixel, sample, seed) by the PCG stream, so the only HBM traffic is the output planes.
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

__device__ __forceinline__ uint64_t lanes_below() {
    const uint32_t lane = threadIdx.x & 63u;
    return lane == 0u ? 0ull : (~0ull >> (64u - lane));
}

// Primary-ray trace as a lane-persistent state machine.  Every trip of the loop runs at most
// one Iterate body on each lane; a lane whose field evaluation completes in the trip applies
// the fence and then either advances its march (march_segment, pathtrace.wgsl:78-145), takes
// the next tap of its normal (common.wgsl:376-387), or writes its pixel.  A lane without a
// ray takes the next one from the launch's ray counter (one atomic per wave).  Deep and
// shallow evaluations, long and short rays therefore share a wave without waiting for each
// other, and one field-evaluation site serves march steps and normal taps.  The arithmetic of
// each evaluation and step is march()'s and normal_at()'s, in the same order.
template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_trace_primary(const JobP P, uint32_t sample, uint32_t tiles_x,
                                                       uint32_t n_slots, uint32_t* __restrict__ next,
                                                       float* __restrict__ depth, float4* __restrict__ aux,
                                                       uint32_t* __restrict__ steps_out, float* __restrict__ t_prev_out,
                                                       uint32_t* __restrict__ hit_list, uint32_t* __restrict__ hit_count,
                                                       unsigned long long* __restrict__ ctr) {
    enum : uint32_t { IDLE = 0, MARCH = 1, NORMAL = 2, DONE = 3 };
    const uint32_t lane = threadIdx.x & 63u;
    const FractalP& F = P.f;
    const float safety = F.safety;
    const float hit_cone = P.pixel_cone * P.pixel_tolerance;
    uint32_t mode = IDLE;
    V3<float> ro = mk(0.0f, 0.0f, 0.0f), rd = ro, g = ro;
    float t = 0.0f, omega = 1.2f, prev_radius = 0.0f, step_len = 0.0f, cand_t = 0.0f, cand_err = 1e30f, t_last = 0.0f;
    float h = 0.0f, t_hit = 0.0f, tp_hit = 0.0f;
    uint32_t i = 0, nsteps = 0, tap = 0, pix = 0;
    DeStepper<FAM> de;
    bool de_done = false;
    Work w;
    uint32_t hits = 0, rays = 0;
    for (;;) {
        // refill idle lanes from the ray counter
        const bool need = mode == IDLE;
        const uint64_t mask = __ballot(need);
        if (mask) {
            uint32_t base = 0;
            if (lane == 0u) base = atomicAdd(next, uint32_t(__popcll(mask)));
            base = __shfl(base, 0, 64);
            if (need) {
                const uint32_t idx = base + uint32_t(__popcll(mask & lanes_below()));
                if (idx >= n_slots) {
                    mode = DONE;
                } else {
                    uint32_t px, py;
                    ray_pixel(idx, tiles_x, px, py);
                    if (px < P.width && py < P.height) {
                        pix = py * P.width + px;
                        uint32_t rng = rng_init(px, py, sample, P.seed);
                        camera_ray(P.cam, P.width, P.height, px, py, rng, ro, rd);
                        t = 0.0f; omega = 1.2f; prev_radius = 0.0f; step_len = 0.0f;
                        cand_t = 0.0f; cand_err = 1e30f; t_last = 0.0f;
                        i = 0;
                        nsteps = 0;
                        mode = MARCH;
                        ++rays;
                        de_done = de.start(F, ro);
                    }
                }
            }
        }
        if (__ballot(mode != DONE) == 0ull) break;
        const bool active = mode == MARCH || mode == NORMAL;
        if (active && !de_done) de_done = de.step(F);
        bool produced_hit = false;
        if (active && de_done) {
            const De<float> s = de.finish(F);
            w.de_evals++;
            w.de_iters += s.runs;
            if (mode == MARCH) {
                w.steps++;
                nsteps = i + 1u;
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
                // outcome: 0 continue, 1 hit, 2 miss
                uint32_t outcome = 0;
                if (!sor_fail && err < hit_cone) {
                    outcome = 1;
                    t_hit = t;
                    tp_hit = t_last;
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
                } else if (outcome == 1u) {
                    mode = NORMAL;
                    tap = 0;
                    g = mk(0.0f, 0.0f, 0.0f);
                    h = fmaxf(t_hit * P.pixel_cone * 0.5f, 1e-4f) * P.pixel_tolerance;
                    const V3<float> p = add(ro, scl(rd, t_hit));
                    de_done = de.start(F, add(p, scl(mk(1.0f, -1.0f, -1.0f), h)));
                } else {
                    depth[pix] = -1.0f;
                    aux[pix] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
                    steps_out[pix] = nsteps;
                    mode = IDLE;
                }
            } else {
                const float kx = (tap == 0u || tap == 3u) ? 1.0f : -1.0f;
                const float ky = (tap >= 2u) ? 1.0f : -1.0f;
                const float kz = (tap == 1u || tap == 3u) ? 1.0f : -1.0f;
                g = add(g, scl(mk(kx, ky, kz), s.d));
                ++tap;
                if (tap == 4u) {
                    const V3<float> n = nrm(g);
                    depth[pix] = t_hit;
                    t_prev_out[pix] = tp_hit;
                    aux[pix] = make_float4(n.x, n.y, n.z, float(nsteps));
                    steps_out[pix] = nsteps;
                    ++hits;
                    produced_hit = true;
                    mode = IDLE;
                } else {
                    const float nx = (tap == 3u) ? 1.0f : -1.0f;
                    const float ny = (tap >= 2u) ? 1.0f : -1.0f;
                    const float nz = (tap == 1u || tap == 3u) ? 1.0f : -1.0f;
                    const V3<float> p = add(ro, scl(rd, t_hit));
                    de_done = de.start(F, add(p, scl(mk(nx, ny, nz), h)));
                }
            }
        }
        // queue the hit for the FP64 refine (one atomic per wave)
        const uint64_t hm = __ballot(produced_hit);
        if (hm) {
            uint32_t base = 0;
            if (lane == 0u) base = atomicAdd(hit_count, uint32_t(__popcll(hm)));
            base = __shfl(base, 0, 64);
            if (produced_hit) hit_list[base + uint32_t(__popcll(hm & lanes_below()))] = pix;
        }
    }
    flush_work(w, hits, rays, ctr);
    const uint32_t pr = wave_sum(rays);
    if (lane == 0u) atomicAdd(&ctr[CTR_PRIMARY], (unsigned long long)pr);
}

// FP64 refine of the queued first hits: the camera ray is regenerated from its PCG stream
// (the same FP32 ray the march followed) and the hit rule's crossing is located in FP64
// between the march's last two positions.  Every lane holds a hit, so no lane idles on a
// miss; kept out of the march kernel so the FP32 march keeps its small register budget.
template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_refine64(const JobP P, uint32_t sample, const uint32_t* __restrict__ hit_list,
                                                  const uint32_t* __restrict__ hit_count,
                                                  const float* __restrict__ depth, const float* __restrict__ t_prev,
                                                  double* __restrict__ depth64, unsigned long long* __restrict__ ctr) {
    const uint32_t n = *hit_count;
    Work w;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t pix = hit_list[k];
        const uint32_t px = pix % P.width, py = pix / P.width;
        uint32_t rng = rng_init(px, py, sample, P.seed);
        V3<float> ro, rd;
        camera_ray(P.cam, P.width, P.height, px, py, rng, ro, rd);
        depth64[pix] = refine64<FAM>(P.f, cvt<double>(ro), cvt<double>(rd), double(t_prev[pix]), double(depth[pix]),
                                     double(P.pixel_cone) * double(P.pixel_tolerance), w);
    }
    flush_work(w, 0u, 0u, ctr);
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
    default: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_SPHERE>, __VA_ARGS__); break;                        \
    }

hipError_t launch_trace_primary(const JobP& P, uint32_t sample, const TraceBuffers& B, int cu_count, hipStream_t s) {
    const uint32_t tiles_x = (P.width + 7u) / 8u, tiles_y = (P.height + 7u) / 8u;
    const uint32_t n_slots = tiles_x * tiles_y * 64u;
    const uint32_t fill = uint32_t(cu_count > 0 ? cu_count : 256) * 8u;
    const dim3 grid(std::min<uint32_t>((n_slots + 255u) / 256u, fill)), block(256);
    hipError_t e = hipMemsetAsync(B.counters, 0, 2 * sizeof(uint32_t), s);
    if (e != hipSuccess) return e;
    SYF_DISPATCH_FAMILY(kernel_family(P.f), k_trace_primary, grid, block, 0, s, P, sample, tiles_x, n_slots,
                        B.counters, B.depth, B.aux, B.steps, B.t_prev, B.hit_list, B.counters + 1, B.ctr);
    return hipGetLastError();
}

hipError_t launch_refine64(const JobP& P, uint32_t sample, const TraceBuffers& B, int cu_count, hipStream_t s) {
    const uint32_t n = P.width * P.height;
    const uint32_t fill = uint32_t(cu_count > 0 ? cu_count : 256) * 8u;
    const dim3 grid(std::min<uint32_t>((n + 255u) / 256u, fill)), block(256);
    hipLaunchKernelGGL(k_depth64_init, grid, block, 0, s, B.depth, B.depth64, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    SYF_DISPATCH_FAMILY(kernel_family(P.f), k_refine64, grid, block, 0, s, P, sample, B.hit_list, B.counters + 1,
                        B.depth, B.t_prev, B.depth64, B.ctr);
    return hipGetLastError();
}

hipError_t launch_de_probe(const FractalP& F, const float4* pts, uint32_t n, float4* out, float4* traps,
                           hipStream_t s) {
    const dim3 grid((n + 255u) / 256u), block(256);
    SYF_DISPATCH_FAMILY(kernel_family(F), k_de_probe, grid, block, 0, s, F, pts, n, out, traps);
    return hipGetLastError();
}

}  // namespace syf
