// Path-traced forms of the distance-estimator path.
//
// Megakernel (syf_render_tile): one lane per pixel runs its samples in ascending order, every
// bounce and shadow ray inline, like pathtrace.wgsl main (:1203-1276).  The one-launch form.
//
// Wavefront (syf_render_wavefront): per sample, generate -> (intersect -> shade/extend) x
// rounds -> commit, the stages of wavefront.wgsl / wavefront.rs:310-419 re-laid for CDNA:
//   * path state lives in structure-of-arrays planes in HBM (one float or u32 plane per field,
//     lane i touches element i: every load and store is a coalesced 256-byte wave access);
//   * survivors are compacted inside the shade kernel with a wave64 ballot and a popcount
//     prefix (one atomic per wave, no LDS scan and no separate compaction pass);
//   * stage kernels use a grid that fills the chip once and stride over the live queue, so
//     shrinking queues cost no empty launches and no host round trip;
//   * the per-sample chain is captured once in a hipGraph and replayed per sample; the sample
//     index is a device counter bumped by the commit stage;
//   * queue invariants are status bits checked once per call (capacity overflow = 1,
//     compaction overflow = 2, path alive at commit = 4; wavefront.wgsl:183,220,236).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "api_internal.hpp"
#include "kernels.hpp"
#include "shade.hpp"

namespace syf {

// Trace one segment: nearest emitter, then the march up to it (trace_geometry_lights).
template <uint32_t FAM>
__device__ __forceinline__ SegHit trace_segment(const JobP& P, F3 o, F3 d, Work& w) {
    uint32_t light;
    const float nearest = nearest_emitter(P, o, d, P.t_end, light);
    Hit<float> hm = march<float, FAM>(P.f, o, d, P.pixel_cone, P.pixel_tolerance, P.max_steps, nearest, w);
    SegHit sh{hm.hit, hm.exhausted, hm.t, hm.steps, NO_LIGHT, hm.s.trap0, hm.t_prev};
    if (!hm.hit && !hm.exhausted && light != NO_LIGHT) {
        sh.hit = true;
        sh.t = nearest;
        sh.light = light;
    }
    return sh;
}

template <uint32_t FAM> struct Tracer {
    const JobP& P;
    Work& w;
    uint32_t& shadow_rays;
    __device__ F3 normal(F3 p, float h) const { return normal_at<float, FAM>(P.f, p, h, w); }
    __device__ bool visible(F3 o, F3 d, float t_end) const {
        ++shadow_rays;
        uint32_t light;
        const float nearest = nearest_emitter(P, o, d, t_end, light);
        Hit<float> hm = march<float, FAM>(P.f, o, d, P.pixel_cone, P.pixel_tolerance, P.max_steps, nearest, w);
        return !hm.hit && !hm.exhausted && light == NO_LIGHT;
    }
};

// ---------------------------------------------------------------- megakernel

// All samples s0 .. s0+ns-1 of one pixel, in ascending order, accumulated as
// accum += (r if finite, 1) (pathtrace.wgsl:1203-1276).  A call that starts at sample 0 also
// writes the first-hit planes: depth, normal + steps, and the first surface hit refined in FP64
// between the march's last two positions (refine64, as the primary trace), -1 otherwise.
template <uint32_t FAM>
__device__ __forceinline__ void render_pixel(const JobP& P, uint32_t px, uint32_t py, uint32_t s0, uint32_t ns,
                                             float4* __restrict__ accum, float* __restrict__ depth,
                                             float4* __restrict__ aux, double* __restrict__ depth64, Work& work,
                                             uint32_t& segs, uint32_t& shadow, uint32_t& prim, uint32_t& hits) {
    const size_t idx = size_t(py) * P.width + px;
    Tracer<FAM> tr{P, work, shadow};
    float4 acc = accum[idx];
    float fd0 = -1.0f, tp0 = -1.0f;
    float4 fa0 = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    for (uint32_t s = 0; s < ns; ++s) {
        PathState st;
        path_start(P, px, py, s0 + s, st);
        float fd = -1.0f;
        float4 fa = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        ++prim;
        while (st.alive) {
            SegHit sh = trace_segment<FAM>(P, st.o, st.d, work);
            ++segs;
            if (st.bounce == 0u && sh.hit && sh.light == NO_LIGHT) {
                ++hits;
                if (s == 0u) tp0 = sh.t_prev;
            }
            path_bounce(P, st, sh, fd, fa, [&](F3 p, float hh) { return tr.normal(p, hh); },
                        [&](F3 o, F3 d, float te) { return tr.visible(o, d, te); });
        }
        const F3 r = st.L;
        const bool ok = r.x == r.x && r.y == r.y && r.z == r.z && fabsf(r.x) < 1e30f && fabsf(r.y) < 1e30f &&
                        fabsf(r.z) < 1e30f;
        acc.x += ok ? r.x : 0.0f;
        acc.y += ok ? r.y : 0.0f;
        acc.z += ok ? r.z : 0.0f;
        acc.w += 1.0f;
        if (s == 0u) {
            fd0 = fd;
            fa0 = fa;
        }
    }
    accum[idx] = acc;
    if (s0 == 0u && ns > 0u) {
        depth[idx] = fd0;
        aux[idx] = fa0;
        if (depth64) {
            double d64 = -1.0;
            if (fd0 >= 0.0f && tp0 >= 0.0f) {
                uint32_t rng = rng_init(px, py, 0u, P.seed);
                F3 ro, rd;
                camera_ray(P.cam, P.width, P.height, px, py, rng, ro, rd);
                d64 = refine64_of_march<FAM>(P.f, ro, rd, tp0, fd0, double(P.pixel_cone) * double(P.pixel_tolerance),
                                             work);
            }
            depth64[idx] = d64;
        }
    }
}

template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_render_tile(const JobP P, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                                                     uint32_t s0, uint32_t ns, uint32_t tiles_x, uint32_t tiles,
                                                     float4* __restrict__ accum, float* __restrict__ depth,
                                                     float4* __restrict__ aux, double* __restrict__ depth64,
                                                     unsigned long long* __restrict__ ctr) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t tile = blockIdx.x * 4u + wave;
    if (tile >= tiles) return;
    const uint32_t lx = (tile % tiles_x) * 8u + (lane & 7u), ly = (tile / tiles_x) * 8u + (lane >> 3);
    Work work;
    uint32_t segs = 0, shadow = 0, prim = 0, hits = 0;
    if (lx < w && ly < h)
        render_pixel<FAM>(P, x0 + lx, y0 + ly, s0, ns, accum, depth, aux, depth64, work, segs, shadow, prim, hits);
    flush_work(work, hits, segs + shadow, ctr);
    uint32_t p = wave_sum(prim);
    if (lane == 0u) atomicAdd(&ctr[CTR_PRIMARY], (unsigned long long)p);
}

// ---------------------------------------------------------------- wavefront

enum : uint32_t {
    // float planes
    WF_OX, WF_OY, WF_OZ, WF_DX, WF_DY, WF_DZ, WF_TX, WF_TY, WF_TZ, WF_LX, WF_LY, WF_LZ, WF_NX, WF_NY, WF_NZ,
    WF_PDF, WF_TOTAL, WF_FD, WF_FA0, WF_FA1, WF_FA2, WF_FA3, WF_HT, WF_HTRAP,
    WF_TPREV,  // first segment: march position before a surface hit (FP64 refine bracket), else -1
    // u32 planes
    WF_RNG, WF_BOUNCE, WF_HSTEPS, WF_HLIGHT,
    WF_PLANES
};
// WF_BOUNCE packs bounce (bits 0-15), alive (16), prev_from_bsdf (17), hit (18), exhausted (19).

struct WfCtl {
    uint32_t count[2];  // queue lengths (ping-pong)
    uint32_t sample;    // samples committed by this call
    uint32_t error;     // status bits
    uint32_t rays, shadow, hits, primary;
};

struct WfP {
    uint32_t x0, y0, w, h, capacity, sample_start, tiles_x;
};

struct WavefrontArena {
    uint32_t capacity = 0;
    void* planes = nullptr;     // WF_PLANES x capacity x 4 B
    uint32_t* queue = nullptr;  // 2 x capacity
    WfCtl* ctl = nullptr;
    hipGraphExec_t graph = nullptr;
    uint32_t graph_key[8] = {0};
};

struct Soa {
    float* f;
    uint32_t* u;
    uint32_t cap;
    __device__ float& F(uint32_t k, uint32_t i) const { return f[size_t(k) * cap + i]; }
    __device__ uint32_t& U(uint32_t k, uint32_t i) const { return u[size_t(k) * cap + i]; }
};

__device__ __forceinline__ void wf_save(const Soa& S, uint32_t id, const PathState& st) {
    S.F(WF_OX, id) = st.o.x; S.F(WF_OY, id) = st.o.y; S.F(WF_OZ, id) = st.o.z;
    S.F(WF_DX, id) = st.d.x; S.F(WF_DY, id) = st.d.y; S.F(WF_DZ, id) = st.d.z;
    S.F(WF_TX, id) = st.throughput.x; S.F(WF_TY, id) = st.throughput.y; S.F(WF_TZ, id) = st.throughput.z;
    S.F(WF_LX, id) = st.L.x; S.F(WF_LY, id) = st.L.y; S.F(WF_LZ, id) = st.L.z;
    S.F(WF_NX, id) = st.prev_n.x; S.F(WF_NY, id) = st.prev_n.y; S.F(WF_NZ, id) = st.prev_n.z;
    S.F(WF_PDF, id) = st.prev_pdf; S.F(WF_TOTAL, id) = st.prev_total;
    S.U(WF_RNG, id) = st.rng;
    uint32_t b = S.U(WF_BOUNCE, id) & ~0x3ffffu;
    S.U(WF_BOUNCE, id) = b | (st.bounce & 0xffffu) | (st.alive ? 0x10000u : 0u) | (st.prev_from_bsdf ? 0x20000u : 0u);
}
__device__ __forceinline__ PathState wf_load(const Soa& S, uint32_t id) {
    PathState st;
    st.o = mk(S.F(WF_OX, id), S.F(WF_OY, id), S.F(WF_OZ, id));
    st.d = mk(S.F(WF_DX, id), S.F(WF_DY, id), S.F(WF_DZ, id));
    st.throughput = mk(S.F(WF_TX, id), S.F(WF_TY, id), S.F(WF_TZ, id));
    st.L = mk(S.F(WF_LX, id), S.F(WF_LY, id), S.F(WF_LZ, id));
    st.prev_n = mk(S.F(WF_NX, id), S.F(WF_NY, id), S.F(WF_NZ, id));
    st.prev_pdf = S.F(WF_PDF, id);
    st.prev_total = S.F(WF_TOTAL, id);
    st.rng = S.U(WF_RNG, id);
    const uint32_t b = S.U(WF_BOUNCE, id);
    st.bounce = b & 0xffffu;
    st.alive = (b & 0x10000u) != 0u;
    st.prev_from_bsdf = (b & 0x20000u) != 0u;
    return st;
}

__global__ void k_wf_reset(WfCtl* ctl) {
    if (threadIdx.x == 0) {
        ctl->count[0] = ctl->count[1] = 0;
        ctl->sample = 0;
        ctl->error = 0;
        ctl->rays = ctl->shadow = ctl->hits = ctl->primary = 0;
    }
}

__global__ __launch_bounds__(256) void k_wf_generate(const JobP P, const WfP W, Soa S, uint32_t* __restrict__ queue,
                                                     WfCtl* __restrict__ ctl, unsigned long long* __restrict__ ctr) {
    const uint32_t n = W.w * W.h;
    const uint32_t sample = W.sample_start + ctl->sample;
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < n; id += gridDim.x * blockDim.x) {
        const uint32_t px = W.x0 + id % W.w, py = W.y0 + id / W.w;
        PathState st;
        path_start(P, px, py, sample, st);
        S.U(WF_BOUNCE, id) = 0u;
        wf_save(S, id, st);
        S.F(WF_FD, id) = -1.0f;
        S.F(WF_FA0, id) = 0.0f; S.F(WF_FA1, id) = 0.0f; S.F(WF_FA2, id) = 0.0f; S.F(WF_FA3, id) = -1.0f;
        queue[id] = id;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->count[0] = n;
        ctl->count[1] = 0;
        atomicAdd(&ctl->primary, n);
        atomicAdd(&ctr[CTR_PRIMARY], (unsigned long long)n);
    }
}

// Intersect every path of queue `in`: the extension segment of its current bounce, as
// trace_segment does it (the nearest emitter caps the march; a miss in front of an emitter is
// a hit on it).  Lane-persistent: each wave owns a contiguous range of the queue, and a lane
// whose march has ended takes the range's next path, so lanes do not wait for one another's
// marches.  Every trip of the loop runs one field evaluation and march()'s step on each busy
// lane (march_advance: the same arithmetic in the same order, so the segments equal
// trace_segment's bit for bit).  Measured on the 1080p Mandelbox frame (wavefront form, Mrays/s):
// a thread per path marching to the end 77.2, this 80.3; one Iterate body per trip 68.0, and
// four evaluations per trip 79.6.
template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_wf_intersect(const JobP P, const WfP W, Soa S, const uint32_t* __restrict__ queue,
                                                      uint32_t in, WfCtl* __restrict__ ctl,
                                                      unsigned long long* __restrict__ ctr) {
    const uint32_t count = ctl->count[in];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t per = (count + waves - 1u) / waves;
    const uint32_t r1 = uint32_t(min(uint64_t(count), uint64_t(wave + 1u) * per));
    uint32_t next = uint32_t(min(uint64_t(count), uint64_t(wave) * per));  // wave-uniform
    const uint64_t below = lane == 0u ? 0ull : (~0ull >> (64u - lane));
    const FractalP& F = P.f;
    const float hit_cone = P.pixel_cone * P.pixel_tolerance;
    Work work;
    uint32_t segs = 0;
    bool busy = false, de_done = false;
    uint32_t id = 0, light = NO_LIGHT;
    MarchState m;
    DeStepper<FAM> de;
    for (;;) {
        // idle lanes take the next paths of the wave's range
        const uint64_t idle = __ballot(!busy);
        if (idle != 0ull && next < r1) {
            if (!busy) {
                const uint32_t at = next + uint32_t(__popcll(idle & below));
                if (at < r1) {
                    id = queue[size_t(in) * W.capacity + at];
                    const F3 o = mk(S.F(WF_OX, id), S.F(WF_OY, id), S.F(WF_OZ, id));
                    const F3 d = mk(S.F(WF_DX, id), S.F(WF_DY, id), S.F(WF_DZ, id));
                    const float nearest = nearest_emitter(P, o, d, P.t_end, light);
                    // nearest > 0 (the ray extent is >= 200, emitter distances exceed 1e-6), so
                    // march() evaluates its first position
                    march_reset(m, o, d, nearest);
                    de_done = de.start(F, add(o, scl(d, 0.0f)));
                    busy = true;
                    ++segs;
                }
            }
            next = min(r1, next + uint32_t(__popcll(idle)));
        }
        if (__ballot(busy) == 0ull) break;  // the range is used up and every march has ended
        if (busy) {
            while (!de_done) de_done = de.step(F);
            const De<float> s = de.finish(F);
            work.de_evals++;
            work.de_iters += s.runs;
            work.steps++;
            MarchEnd e;
            if (!march_advance(m, s, F.safety, hit_cone, P.max_steps, e)) {
                de_done = de.start(F, add(m.ro, scl(m.rd, m.t)));
            } else {
                SegHit sh{e.hit, e.exhausted, e.t, e.steps, NO_LIGHT, e.trap0, e.t_prev};
                if (!e.hit && !e.exhausted && light != NO_LIGHT) {
                    sh.hit = true;
                    sh.t = m.t_end;
                    sh.light = light;
                }
                S.F(WF_HT, id) = sh.t;
                S.F(WF_HTRAP, id) = sh.trap0;
                S.U(WF_HSTEPS, id) = sh.steps;
                S.U(WF_HLIGHT, id) = sh.light;
                uint32_t b = S.U(WF_BOUNCE, id) & ~(0x40000u | 0x80000u);
                if ((b & 0xffffu) == 0u) S.F(WF_TPREV, id) = sh.hit && sh.light == NO_LIGHT ? sh.t_prev : -1.0f;
                S.U(WF_BOUNCE, id) = b | (sh.hit ? 0x40000u : 0u) | (sh.exhausted ? 0x80000u : 0u);
                busy = false;
            }
        }
    }
    flush_work(work, 0u, segs, ctr);
}

// Shade every path of queue `in` (normal, surface, next-event estimation with its shadow
// march, continuation, roulette) and append the survivors to queue `in ^ 1`.
template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_wf_shade(const JobP P, const WfP W, Soa S, uint32_t* __restrict__ queue,
                                                  uint32_t in, WfCtl* __restrict__ ctl,
                                                  unsigned long long* __restrict__ ctr) {
    const uint32_t count = ctl->count[in];
    const uint32_t lane = threadIdx.x & 63u;
    Work work;
    uint32_t shadow = 0, hits = 0;
    Tracer<FAM> tr{P, work, shadow};
    // every lane of a wave runs the same number of strides, so the ballot below is wave-wide
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t base0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    for (uint32_t b0 = base0; b0 < count; b0 += stride) {
        const uint32_t i = b0 + lane;
        bool alive = false;
        uint32_t id = 0;
        if (i < count) {
            id = queue[size_t(in) * W.capacity + i];
            PathState st = wf_load(S, id);
            const uint32_t bw = S.U(WF_BOUNCE, id);
            SegHit sh{(bw & 0x40000u) != 0u, (bw & 0x80000u) != 0u, S.F(WF_HT, id), S.U(WF_HSTEPS, id),
                      S.U(WF_HLIGHT, id), S.F(WF_HTRAP, id)};
            if (st.bounce == 0u && sh.hit && sh.light == NO_LIGHT) ++hits;
            float fd = S.F(WF_FD, id);
            float4 fa = make_float4(S.F(WF_FA0, id), S.F(WF_FA1, id), S.F(WF_FA2, id), S.F(WF_FA3, id));
            const bool first = st.bounce == 0u;
            path_bounce(P, st, sh, fd, fa, [&](F3 p, float hh) { return tr.normal(p, hh); },
                        [&](F3 o, F3 d, float te) { return tr.visible(o, d, te); });
            wf_save(S, id, st);
            if (first) {
                S.F(WF_FD, id) = fd;
                S.F(WF_FA0, id) = fa.x; S.F(WF_FA1, id) = fa.y; S.F(WF_FA2, id) = fa.z; S.F(WF_FA3, id) = fa.w;
            }
            alive = st.alive;
        }
        // wave64 compaction: ballot, popcount prefix, one atomic per wave
        const unsigned long long mask = __ballot(alive);
        const uint32_t total = uint32_t(__popcll(mask));
        uint32_t base = 0;
        if (lane == 0u && total) base = atomicAdd(&ctl->count[in ^ 1u], total);
        base = __shfl(base, 0, 64);
        if (alive) {
            const uint32_t slot = base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
            if (slot < W.capacity) queue[size_t(in ^ 1u) * W.capacity + slot] = id;
            else atomicOr(&ctl->error, 2u);
        }
    }
    flush_work(work, hits, shadow, ctr);
}

__global__ void k_wf_next_round(uint32_t in, uint32_t capacity, WfCtl* ctl) {
    if (threadIdx.x == 0) {
        if (ctl->count[in ^ 1u] > capacity) {
            atomicOr(&ctl->error, 1u);
            ctl->count[in ^ 1u] = 0;
        }
        ctl->count[in] = 0;
    }
}

// Accumulate each path of the tile into the film (wf_commit, wavefront.wgsl:232-260).
template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_wf_commit(const JobP P, const WfP W, Soa S, float4* __restrict__ accum,
                                                   float* __restrict__ depth, float4* __restrict__ aux,
                                                   double* __restrict__ depth64, WfCtl* __restrict__ ctl,
                                                   unsigned long long* __restrict__ ctr) {
    const uint32_t n = W.w * W.h;
    const uint32_t sample = W.sample_start + ctl->sample;
    Work work;
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < n; id += gridDim.x * blockDim.x) {
        const uint32_t px = W.x0 + id % W.w, py = W.y0 + id / W.w;
        const size_t idx = size_t(py) * P.width + px;
        if (S.U(WF_BOUNCE, id) & 0x10000u) {
            atomicOr(&ctl->error, 4u);
            continue;
        }
        const F3 r = mk(S.F(WF_LX, id), S.F(WF_LY, id), S.F(WF_LZ, id));
        const bool ok = r.x == r.x && r.y == r.y && r.z == r.z && fabsf(r.x) < 1e30f && fabsf(r.y) < 1e30f &&
                        fabsf(r.z) < 1e30f;
        float4 a = accum[idx];
        a.x += ok ? r.x : 0.0f;
        a.y += ok ? r.y : 0.0f;
        a.z += ok ? r.z : 0.0f;
        a.w += 1.0f;
        accum[idx] = a;
        if (sample == 0u) {
            const float fd = S.F(WF_FD, id), tp = S.F(WF_TPREV, id);
            depth[idx] = fd;
            aux[idx] = make_float4(S.F(WF_FA0, id), S.F(WF_FA1, id), S.F(WF_FA2, id), S.F(WF_FA3, id));
            double d64 = -1.0;
            if (fd >= 0.0f && tp >= 0.0f) {  // the camera ray of sample 0, regenerated from its PCG stream
                uint32_t rng = rng_init(px, py, 0u, P.seed);
                F3 ro, rd;
                camera_ray(P.cam, P.width, P.height, px, py, rng, ro, rd);
                d64 = refine64_of_march<FAM>(P.f, ro, rd, tp, fd, double(P.pixel_cone) * double(P.pixel_tolerance),
                                             work);
            }
            depth64[idx] = d64;
        }
    }
    flush_work(work, 0u, 0u, ctr);
}

__global__ void k_wf_end_sample(WfCtl* ctl) {
    if (threadIdx.x == 0) ctl->sample += 1u;
}

#define SYF_DISPATCH_FAMILY(fam, KERNEL, ...)                                                      \
    switch (fam) {                                                                                 \
    case SYF_FAMILY_MANDELBULB: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_MANDELBULB>, __VA_ARGS__); break; \
    case SYF_FAMILY_MANDELBOX: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_MANDELBOX>, __VA_ARGS__); break;   \
    case SYF_FAMILY_MENGER: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_MENGER>, __VA_ARGS__); break;         \
    case FAM_BULB_TRIG: hipLaunchKernelGGL(KERNEL<FAM_BULB_TRIG>, __VA_ARGS__); break;                 \
    case FAM_BULB8: hipLaunchKernelGGL(KERNEL<FAM_BULB8>, __VA_ARGS__); break;                         \
    case FAM_GRAPH: hipLaunchKernelGGL(KERNEL<FAM_GRAPH>, __VA_ARGS__); break;                         \
    default: hipLaunchKernelGGL(KERNEL<SYF_FAMILY_SPHERE>, __VA_ARGS__); break;                        \
    }

// ---------------------------------------------------------------- batched serving form

// One wave takes one work item: an 8x8 tile of one request and a range of its samples.  The
// request's parameters are read from the job table (wave-uniform, scalar loads) and its
// family is a wave-uniform switch, so requests of every family and size share one launch.
// Per pixel the arithmetic is the one-launch form's: a request renders the same alone and in
// any mix.  Items are claimed from eight counters (one per XCD, blockIdx % 8) to keep the
// claims of different L2 slices off one address.
hipError_t launch_render_tile(const JobP& P, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t s0,
                              uint32_t ns, float4* accum, float* depth, float4* aux, double* depth64,
                              unsigned long long* ctr, hipStream_t s) {
    const uint32_t tiles_x = (w + 7u) / 8u, tiles = tiles_x * ((h + 7u) / 8u);
    const dim3 grid((tiles + 3u) / 4u), block(256);
    SYF_DISPATCH_FAMILY(kernel_family(P.f), k_render_tile, grid, block, 0, s, P, x0, y0, w, h, s0, ns, tiles_x,
                        tiles, accum, depth, aux, depth64, ctr);
    return hipGetLastError();
}

void release_wavefront(syf_job* j) {
    if (!j || !j->wf) return;
    WavefrontArena* a = j->wf;
    if (a->graph) hipGraphExecDestroy(a->graph);
    hipFree(a->planes);
    hipFree(a->queue);
    hipFree(a->ctl);
    delete a;
    j->wf = nullptr;
}

}  // namespace syf

using namespace syf;

namespace {

// Stage chain of one sample, enqueued on `s` (captured into a graph by the caller).
hipError_t enqueue_wavefront_sample(syf_job* j, const WfP& W, int cu_count) {
    WavefrontArena* a = j->wf;
    const JobP& P = j->P;
    Soa S{(float*)a->planes, (uint32_t*)a->planes, a->capacity};
    const uint32_t n = W.w * W.h;
    const uint32_t fill = uint32_t(cu_count > 0 ? cu_count : 256) * 8u;  // workgroups that fill the chip
    const dim3 block(256);
    const dim3 gen_grid(std::min<uint32_t>((n + 255u) / 256u, fill));
    hipStream_t s = j->stream;
    hipLaunchKernelGGL(k_wf_generate, gen_grid, block, 0, s, P, W, S, a->queue, a->ctl, j->ctr);
    const uint32_t fam = kernel_family(P.f);
    for (uint32_t r = 0; r <= P.max_bounces; ++r) {
        const uint32_t in = r & 1u;
        const dim3 grid(std::min<uint32_t>((n + 255u) / 256u, fill));
        SYF_DISPATCH_FAMILY(fam, k_wf_intersect, grid, block, 0, s, P, W, S, a->queue, in, a->ctl, j->ctr);
        SYF_DISPATCH_FAMILY(fam, k_wf_shade, grid, block, 0, s, P, W, S, a->queue, in, a->ctl, j->ctr);
        hipLaunchKernelGGL(k_wf_next_round, dim3(1), dim3(64), 0, s, in, a->capacity, a->ctl);
    }
    SYF_DISPATCH_FAMILY(fam, k_wf_commit, gen_grid, block, 0, s, P, W, S, (float4*)j->plane[SYF_PLANE_ACCUM],
                        (float*)j->plane[SYF_PLANE_DEPTH], (float4*)j->plane[SYF_PLANE_AUX],
                        (double*)j->plane[SYF_PLANE_DEPTH64], a->ctl, j->ctr);
    hipLaunchKernelGGL(k_wf_end_sample, dim3(1), dim3(64), 0, s, a->ctl);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int syf_render_tile(syf_job* j, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sample_start,
                    uint32_t sample_count) {
    clear_error();
    if (!j) return fail(SYF_E_INVALID, "job must not be null");
    if (w == 0 || h == 0 || x >= j->P.width || y >= j->P.height || w > j->P.width - x || h > j->P.height - y)
        return fail(SYF_E_INVALID, "tile (%u,%u %ux%u) must lie inside the %ux%u image", x, y, w, h, j->P.width,
                    j->P.height);
    if (hipError_t de = hipSetDevice(j->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    hipError_t e = hipMemsetAsync(j->ctr, 0, sizeof(unsigned long long) * CTR_COUNT, j->stream);
    if (e == hipSuccess) e = hipEventRecord(j->ev0, j->stream);
    if (e == hipSuccess)
        e = launch_render_tile(j->P, x, y, w, h, sample_start, sample_count, (float4*)j->plane[SYF_PLANE_ACCUM],
                               (float*)j->plane[SYF_PLANE_DEPTH], (float4*)j->plane[SYF_PLANE_AUX],
                               (double*)j->plane[SYF_PLANE_DEPTH64], j->ctr, j->stream);
    if (e == hipSuccess) e = hipEventRecord(j->ev1, j->stream);
    if (e == hipSuccess) e = hipEventRecord(j->ev2, j->stream);
    j->timed = true;
    j->extra_ms = 0.0;
    return e == hipSuccess ? SYF_OK : hip_fail(e, "syf_render_tile");
}

int syf_render_wavefront(syf_job* j, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sample_start,
                         uint32_t sample_count) {
    clear_error();
    if (!j) return fail(SYF_E_INVALID, "job must not be null");
    if (w == 0 || h == 0 || x >= j->P.width || y >= j->P.height || w > j->P.width - x || h > j->P.height - y)
        return fail(SYF_E_INVALID, "tile (%u,%u %ux%u) must lie inside the %ux%u image", x, y, w, h, j->P.width,
                    j->P.height);
    if (hipError_t de = hipSetDevice(j->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    const uint32_t n = w * h;
    hipError_t e = hipSuccess;
    if (!j->wf || j->wf->capacity < n) {
        release_wavefront(j);
        WavefrontArena* a = new WavefrontArena();
        a->capacity = n;
        if ((e = hipMalloc(&a->planes, size_t(WF_PLANES) * n * 4)) != hipSuccess ||
            (e = hipMalloc(&a->queue, size_t(2) * n * 4)) != hipSuccess ||
            (e = hipMalloc(&a->ctl, sizeof(WfCtl))) != hipSuccess) {
            hipFree(a->planes);
            hipFree(a->queue);
            delete a;
            return hip_fail(e, "wavefront arena");
        }
        j->wf = a;
    }
    WavefrontArena* a = j->wf;
    WfP W{x, y, w, h, a->capacity, sample_start, 0};
    const uint32_t key[8] = {x, y, w, h, sample_start, a->capacity, 0, 0};
    e = hipMemsetAsync(j->ctr, 0, sizeof(unsigned long long) * CTR_COUNT, j->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, j->stream, a->ctl);
        e = hipGetLastError();
    }
    if (e == hipSuccess && (!a->graph || std::memcmp(key, a->graph_key, sizeof key) != 0)) {
        if (a->graph) {
            hipGraphExecDestroy(a->graph);
            a->graph = nullptr;
        }
        hipGraph_t g = nullptr;
        e = hipStreamBeginCapture(j->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t e2 = enqueue_wavefront_sample(j, W, j->ctx->cu_count);
            e = hipStreamEndCapture(j->stream, &g);
            if (e == hipSuccess) e = e2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&a->graph, g, nullptr, nullptr, 0);
        if (g) hipGraphDestroy(g);
        if (e == hipSuccess) std::memcpy(a->graph_key, key, sizeof key);
    }
    if (e == hipSuccess) e = hipEventRecord(j->ev0, j->stream);
    for (uint32_t s = 0; s < sample_count && e == hipSuccess; ++s) e = hipGraphLaunch(a->graph, j->stream);
    if (e == hipSuccess) e = hipEventRecord(j->ev1, j->stream);
    if (e == hipSuccess) e = hipEventRecord(j->ev2, j->stream);
    WfCtl ctl{};
    if (e == hipSuccess) e = hipMemcpyAsync(&ctl, a->ctl, sizeof ctl, hipMemcpyDeviceToHost, j->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(j->stream);
    j->timed = true;
    j->extra_ms = 0.0;
    if (e != hipSuccess) return hip_fail(e, "syf_render_wavefront");
    if (ctl.error)
        return fail(SYF_E_QUEUE, "wavefront queue invariant broken (status bits 0x%x: 1 = active above capacity, "
                                 "2 = compaction overflow, 4 = path alive at commit)", ctl.error);
    return SYF_OK;
}

}  // extern "C"
