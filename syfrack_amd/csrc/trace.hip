// Primary-ray sphere trace and distance probe kernels.
//
// Launch shape: 256-thread workgroups = four wave64s; each wave owns one 8x8 pixel tile, so
// the rays of a wave are spatial neighbours and take similar step counts (the wave's
// divergence is the spread of its 64 rays).  The march reads nothing from HBM: the job's
// constants sit in the kernel argument segment and every ray is derived from
// (pixel, sample, seed) by the PCG stream, so the only HBM traffic is the output planes.
#include "device.hpp"
#include "kernels.hpp"

namespace syf {

template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_trace_primary(const JobP P, uint32_t sample, uint32_t tiles_x,
                                                       uint32_t tiles, float* __restrict__ depth,
                                                       float4* __restrict__ aux, uint32_t* __restrict__ steps,
                                                       float* __restrict__ t_prev,
                                                       unsigned long long* __restrict__ ctr) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t tile = blockIdx.x * 4u + wave;
    if (tile >= tiles) return;  // whole wave leaves together
    const uint32_t px = (tile % tiles_x) * 8u + (lane & 7u);
    const uint32_t py = (tile / tiles_x) * 8u + (lane >> 3);
    Work w;
    uint32_t hit_count = 0, rays = 0;
    if (px < P.width && py < P.height) {
        rays = 1;
        uint32_t rng = rng_init(px, py, sample, P.seed);
        V3<float> ro, rd;
        camera_ray(P.cam, P.width, P.height, px, py, rng, ro, rd);
        Hit<float> h = march<float, FAM>(P.f, ro, rd, P.pixel_cone, P.pixel_tolerance, P.max_steps, P.t_end, w);
        const size_t i = size_t(py) * P.width + px;
        steps[i] = h.steps;
        if (h.hit) {
            hit_count = 1;
            V3<float> p = add(ro, scl(rd, h.t));
            V3<float> n = normal_at<float, FAM>(P.f, p, fmaxf(h.t * P.pixel_cone * 0.5f, 1e-4f) * P.pixel_tolerance, w);
            depth[i] = h.t;
            t_prev[i] = h.t_prev;
            aux[i] = make_float4(n.x, n.y, n.z, float(h.steps));
        } else {
            depth[i] = -1.0f;
            aux[i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        }
    }
    flush_work(w, hit_count, rays, ctr);
}

// FP64 refine of every first hit: the camera ray is regenerated from its PCG stream (the
// same FP32 ray the march followed) and walked in FP64 from the FP32 hit distance.  Kept
// out of the march kernel so the FP32 march keeps its small register budget.
template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_refine64(const JobP P, uint32_t sample, uint32_t tiles_x, uint32_t tiles,
                                                  const float* __restrict__ depth, const float* __restrict__ t_prev,
                                                  double* __restrict__ depth64, unsigned long long* __restrict__ ctr) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t tile = blockIdx.x * 4u + wave;
    if (tile >= tiles) return;
    const uint32_t px = (tile % tiles_x) * 8u + (lane & 7u);
    const uint32_t py = (tile / tiles_x) * 8u + (lane >> 3);
    Work w;
    if (px < P.width && py < P.height) {
        const size_t i = size_t(py) * P.width + px;
        const float t = depth[i];
        double out = -1.0;
        if (t >= 0.0f) {
            uint32_t rng = rng_init(px, py, sample, P.seed);
            V3<float> ro, rd;
            camera_ray(P.cam, P.width, P.height, px, py, rng, ro, rd);
            out = refine64<FAM>(P.f, cvt<double>(ro), cvt<double>(rd), double(t_prev[i]), double(t),
                                double(P.pixel_cone) * double(P.pixel_tolerance), w);
        }
        depth64[i] = out;
    }
    flush_work(w, 0u, 0u, ctr);
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

hipError_t launch_trace_primary(const JobP& P, uint32_t sample, float* depth, float4* aux, uint32_t* steps,
                                float* t_prev, unsigned long long* ctr, hipStream_t s) {
    const uint32_t tiles_x = (P.width + 7u) / 8u, tiles_y = (P.height + 7u) / 8u;
    const uint32_t tiles = tiles_x * tiles_y;
    const dim3 grid((tiles + 3u) / 4u), block(256);
    SYF_DISPATCH_FAMILY(kernel_family(P.f), k_trace_primary, grid, block, 0, s, P, sample, tiles_x, tiles, depth,
                        aux, steps, t_prev, ctr);
    return hipGetLastError();
}

hipError_t launch_refine64(const JobP& P, uint32_t sample, const float* depth, const float* t_prev, double* depth64,
                           unsigned long long* ctr, hipStream_t s) {
    const uint32_t tiles_x = (P.width + 7u) / 8u, tiles_y = (P.height + 7u) / 8u;
    const uint32_t tiles = tiles_x * tiles_y;
    const dim3 grid((tiles + 3u) / 4u), block(256);
    SYF_DISPATCH_FAMILY(kernel_family(P.f), k_refine64, grid, block, 0, s, P, sample, tiles_x, tiles, depth, t_prev,
                        depth64, ctr);
    return hipGetLastError();
}

hipError_t launch_de_probe(const FractalP& F, const float4* pts, uint32_t n, float4* out, float4* traps,
                           hipStream_t s) {
    const dim3 grid((n + 255u) / 256u), block(256);
    SYF_DISPATCH_FAMILY(kernel_family(F), k_de_probe, grid, block, 0, s, F, pts, n, out, traps);
    return hipGetLastError();
}

}  // namespace syf
