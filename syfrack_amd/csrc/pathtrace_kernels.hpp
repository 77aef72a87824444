// Device code of the path-traced forms: the one-launch megakernel (k_render_tile) and the
// wavefront stages, templated on the kernel family.  Included by pathtrace.hip, which
// instantiates and launches the library's families, and by the run-time compiled units of
// specialised graph topologies (graph.hip), which instantiate the FAM_JIT family.
#pragma once

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
    __device__ __forceinline__ F3 normal(F3 p, float h) const { return normal_at<float, FAM>(P.f, p, h, w); }
    // The field sample at a surface hit in full (the march's own evaluation at that point gives
    // the same values): every normalised trap, the clamped iteration count, the material.
    __device__ __forceinline__ ShadeSample sample(F3 p) const {
        ShadeSample s;
        w.de_evals++;
        if constexpr (FAM == FAM_GRAPH || FAM == FAM_JIT) {
            GRaw<float> g;
            if constexpr (FAM == FAM_JIT) g = graph_jit<float>(P.f.gprog, p);
            else g = graph_run<float>(P.f.gprog, p);
            w.de_iters += g.runs;
            auto norm = [&](float t, uint32_t k) { return isfinite(t) ? fminf(t / P.f.gtrap_scale[k], 1.0f) : 1.0f; };
            s.t0 = norm(g.r.t0, 0u);
            s.t1 = norm(g.r.t1, 1u);
            s.t2 = norm(g.r.t2, 2u);
            s.t3 = norm(g.r.t3, 3u);
            s.iter = g.r.iter < P.f.iteration_product ? g.r.iter : P.f.iteration_product;
            s.material = g.r.mat;
        } else {
            const De<float> d = de_eval<float, FAM>(P.f, p);
            w.de_iters += d.runs;
            s.t0 = d.trap0;
            s.t1 = s.t2 = s.t3 = 1.0f;
            s.iter = d.iter;
            s.material = 0u;
        }
        return s;
    }
    // A shadow ray's march, out of line: the visibility tests of a bounce share one copy of the
    // march, and the registers of the shading live across the call instead of beside an inlined
    // march (the megakernel then runs three waves per SIMD instead of two).  The job constants
    // are in device memory (job_ref), so the call passes their address without a copy.
    __device__ __attribute__((noinline)) bool visible(F3 o, F3 d, float t_end) const {
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
// FOG: the job has fog (the session planes vol are accumulated too).
template <uint32_t FAM, bool FOG>
__device__ __forceinline__ void render_pixel(const JobP& P, uint32_t px, uint32_t py, uint32_t s0, uint32_t ns,
                                             float4* __restrict__ accum, float* __restrict__ depth,
                                             float4* __restrict__ aux, double* __restrict__ depth64,
                                             float4* __restrict__ vol, Work& work, uint32_t& segs, uint32_t& shadow,
                                             uint32_t& prim, uint32_t& hits) {
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
            path_bounce<FOG>(P, st, sh, fd, fa,
                             [&](F3 p, float hh) __attribute__((always_inline)) { return tr.normal(p, hh); },
                             [&](F3 o, F3 d, float te) __attribute__((always_inline)) { return tr.visible(o, d, te); },
                             [&](F3 p) __attribute__((always_inline)) { return tr.sample(p); });
        }
        const F3 r = st.L;
        const bool ok = r.x == r.x && r.y == r.y && r.z == r.z && fabsf(r.x) < 1e30f && fabsf(r.y) < 1e30f &&
                        fabsf(r.z) < 1e30f;
        acc.x += ok ? r.x : 0.0f;
        acc.y += ok ? r.y : 0.0f;
        acc.z += ok ? r.z : 0.0f;
        acc.w += 1.0f;
        if (FOG && vol) accumulate_session(vol, size_t(P.width) * P.height, idx, r, st, ok);
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

template <uint32_t FAM, bool FOG>
__device__ __forceinline__ void render_tile(const JobP* __restrict__ job, uint32_t x0, uint32_t y0, uint32_t w,
                                            uint32_t h, uint32_t s0, uint32_t ns, uint32_t tiles_x, uint32_t tiles,
                                            float4* __restrict__ accum, float* __restrict__ depth,
                                            float4* __restrict__ aux, double* __restrict__ depth64,
                                            float4* __restrict__ vol, unsigned long long* __restrict__ ctr) {
    const JobP& P = job_ref(job);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t tile = blockIdx.x * 4u + wave;
    if (tile >= tiles) return;
    const uint32_t lx = (tile % tiles_x) * 8u + (lane & 7u), ly = (tile / tiles_x) * 8u + (lane >> 3);
    Work work;
    uint32_t segs = 0, shadow = 0, prim = 0, hits = 0;
    if (lx < w && ly < h)
        render_pixel<FAM, FOG>(P, x0 + lx, y0 + ly, s0, ns, accum, depth, aux, depth64, vol, work, segs, shadow, prim,
                               hits);
    flush_work(work, hits, segs + shadow, ctr);
    uint32_t p = wave_sum(prim);
    if (lane == 0u) atomicAdd(&ctr[CTR_PRIMARY], (unsigned long long)p);
}
// The megakernel over a tile of a job without fog (vol unused).
template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_render_tile(const JobP* __restrict__ job, uint32_t x0, uint32_t y0, uint32_t w,
                                                     uint32_t h, uint32_t s0, uint32_t ns, uint32_t tiles_x,
                                                     uint32_t tiles, float4* __restrict__ accum,
                                                     float* __restrict__ depth, float4* __restrict__ aux,
                                                     double* __restrict__ depth64, float4* __restrict__ vol,
                                                     unsigned long long* __restrict__ ctr) {
    render_tile<FAM, false>(job, x0, y0, w, h, s0, ns, tiles_x, tiles, accum, depth, aux, depth64, vol, ctr);
}
// The megakernel of a job with fog or finish detail patterns (both compiled in): 30-40 more
// registers (two waves per SIMD instead of three).  vol: the session planes of a job with fog,
// else null.
template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_render_tile_full(const JobP* __restrict__ job, uint32_t x0, uint32_t y0,
                                                         uint32_t w, uint32_t h, uint32_t s0, uint32_t ns,
                                                         uint32_t tiles_x, uint32_t tiles, float4* __restrict__ accum,
                                                         float* __restrict__ depth, float4* __restrict__ aux,
                                                         double* __restrict__ depth64, float4* __restrict__ vol,
                                                         unsigned long long* __restrict__ ctr) {
    render_tile<FAM, true>(job, x0, y0, w, h, s0, ns, tiles_x, tiles, accum, depth, aux, depth64, vol, ctr);
}

// ---------------------------------------------------------------- wavefront

enum : uint32_t {
    // float planes
    WF_OX, WF_OY, WF_OZ, WF_DX, WF_DY, WF_DZ, WF_TX, WF_TY, WF_TZ, WF_LX, WF_LY, WF_LZ, WF_NX, WF_NY, WF_NZ,
    WF_PDF, WF_TOTAL, WF_FD, WF_FA0, WF_FA1, WF_FA2, WF_FA3, WF_HT, WF_HTRAP,
    WF_TPREV,  // first segment: march position before a surface hit (FP64 refine bracket), else -1
    WF_VX, WF_VY, WF_VZ,  // fog: the camera segment's in-scattered radiance
    WF_TW, WF_TS,         // the ray footprint for finish detail: width, spread
    // u32 planes
    WF_RNG, WF_BOUNCE, WF_HSTEPS, WF_HLIGHT,
    WF_PLANES
};
// WF_BOUNCE packs bounce (bits 0-15), alive (16), prev_from_bsdf (17), hit (18), exhausted (19),
// prev_from_volume (20), prev_volume_nee (21), volume_path (22).

struct WfCtl {
    uint32_t count[2];  // queue lengths (ping-pong)
    uint32_t sample;    // samples committed by this call
    uint32_t error;     // status bits
    uint32_t rays, shadow, hits, primary;
};

struct WfP {
    uint32_t x0, y0, w, h, capacity, sample_start, tiles_x;
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
    S.F(WF_VX, id) = st.primary_volume.x; S.F(WF_VY, id) = st.primary_volume.y; S.F(WF_VZ, id) = st.primary_volume.z;
    S.F(WF_TW, id) = st.tex_width; S.F(WF_TS, id) = st.tex_spread;
    S.U(WF_RNG, id) = st.rng;
    uint32_t b = S.U(WF_BOUNCE, id) & 0xc0000u;  // hit and exhausted belong to the intersect stage
    S.U(WF_BOUNCE, id) = b | (st.bounce & 0xffffu) | (st.alive ? 0x10000u : 0u) | (st.prev_from_bsdf ? 0x20000u : 0u) |
                         (st.prev_from_volume ? 0x100000u : 0u) | (st.prev_volume_nee ? 0x200000u : 0u) |
                         (st.volume_path ? 0x400000u : 0u);
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
    st.prev_from_volume = (b & 0x100000u) != 0u;
    st.prev_volume_nee = (b & 0x200000u) != 0u;
    st.volume_path = (b & 0x400000u) != 0u;
    st.primary_volume = mk(S.F(WF_VX, id), S.F(WF_VY, id), S.F(WF_VZ, id));
    st.tex_width = S.F(WF_TW, id);
    st.tex_spread = S.F(WF_TS, id);
    st.sample = 0u;  // set by the stage that shades (it knows the pixel and the sample)
    st.point_stream = 0u;
    return st;
}

#ifndef __HIPCC_RTC__  // family-independent stages: the library's own build only
__global__ void k_wf_reset(WfCtl* ctl) {
    if (threadIdx.x == 0) {
        ctl->count[0] = ctl->count[1] = 0;
        ctl->sample = 0;
        ctl->error = 0;
        ctl->rays = ctl->shadow = ctl->hits = ctl->primary = 0;
    }
}
#endif


#ifndef __HIPCC_RTC__  // family-independent stages: the library's own build only
__global__ __launch_bounds__(256) void k_wf_generate(const JobP* __restrict__ job, const WfP W, Soa S, uint32_t* __restrict__ queue,
                                                     WfCtl* __restrict__ ctl, unsigned long long* __restrict__ ctr) {
    const JobP& P = job_ref(job);
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
#endif


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
__global__ __launch_bounds__(256) void k_wf_intersect(const JobP* __restrict__ job, const WfP W, Soa S, const uint32_t* __restrict__ queue,
                                                      uint32_t in, WfCtl* __restrict__ ctl,
                                                      unsigned long long* __restrict__ ctr) {
    const JobP& P = job_ref(job);
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
__global__ __launch_bounds__(256) void k_wf_shade(const JobP* __restrict__ job, const WfP W, Soa S, uint32_t* __restrict__ queue,
                                                  uint32_t in, WfCtl* __restrict__ ctl,
                                                  unsigned long long* __restrict__ ctr) {
    const JobP& P = job_ref(job);
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
            st.sample = W.sample_start + ctl->sample;
            st.point_stream = pcg(pcg((W.x0 + id % W.w) ^ P.seed) ^ pcg((W.y0 + id / W.w) ^ 0x27d4eb2du));
            const uint32_t bw = S.U(WF_BOUNCE, id);
            SegHit sh{(bw & 0x40000u) != 0u, (bw & 0x80000u) != 0u, S.F(WF_HT, id), S.U(WF_HSTEPS, id),
                      S.U(WF_HLIGHT, id), S.F(WF_HTRAP, id)};
            if (st.bounce == 0u && sh.hit && sh.light == NO_LIGHT) ++hits;
            float fd = S.F(WF_FD, id);
            float4 fa = make_float4(S.F(WF_FA0, id), S.F(WF_FA1, id), S.F(WF_FA2, id), S.F(WF_FA3, id));
            const bool first = st.bounce == 0u;
            path_bounce<true>(P, st, sh, fd, fa, [&](F3 p, float hh) __attribute__((always_inline)) { return tr.normal(p, hh); },
                        [&](F3 o, F3 d, float te) __attribute__((always_inline)) { return tr.visible(o, d, te); },
                        [&](F3 p) __attribute__((always_inline)) { return tr.sample(p); });
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

#ifndef __HIPCC_RTC__  // family-independent stages: the library's own build only
__global__ void k_wf_next_round(uint32_t in, uint32_t capacity, WfCtl* ctl) {
    if (threadIdx.x == 0) {
        if (ctl->count[in ^ 1u] > capacity) {
            atomicOr(&ctl->error, 1u);
            ctl->count[in ^ 1u] = 0;
        }
        ctl->count[in] = 0;
    }
}
#endif


// Accumulate each path of the tile into the film (wf_commit, wavefront.wgsl:232-260).
template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_wf_commit(const JobP* __restrict__ job, const WfP W, Soa S, float4* __restrict__ accum,
                                                   float* __restrict__ depth, float4* __restrict__ aux,
                                                   double* __restrict__ depth64, float4* __restrict__ vol,
                                                   WfCtl* __restrict__ ctl, unsigned long long* __restrict__ ctr) {
    const JobP& P = job_ref(job);
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
        if (vol) accumulate_session(vol, size_t(P.width) * P.height, idx, r, wf_load(S, id), ok);
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

#ifndef __HIPCC_RTC__  // family-independent stages: the library's own build only
__global__ void k_wf_end_sample(WfCtl* ctl) {
    if (threadIdx.x == 0) ctl->sample += 1u;
}
#endif


}  // namespace syf
