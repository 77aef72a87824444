// Flattened path tracer of the batched server: one lane-persistent state machine per lane.
//
// The megakernel (k_render_tile) runs a pixel's paths as nested loops -- samples, bounces,
// march steps -- so a wave waits at every level for its slowest lane: the longest march of a
// segment, the longest path of a sample, the pixel with the longest paths.  Here every trip of
// the kernel's loop runs a few field evaluations on each evaluating lane, each followed by
// that lane's own march step or normal tap; a lane whose march or normal ended runs the
// shading code up to its next field evaluation.  Lanes therefore sit at different depths of
// different paths and never wait for one another's marches; a lane whose pixel has no samples
// left takes the next pixel of the wave's item.  (Running one Iterate body per trip instead,
// as the primary march does, costs more than it saves here: the loop's bookkeeping for the
// much larger path state exceeds a body.)
//
// A path's arithmetic and random draws are those of render_pixel (pathtrace.hip), in the
// same order: the march is march() (pathtrace.wgsl:78-145) with the evaluation split into
// DeStepper start/step/finish, the normal is normal_at's four taps, and the shading is
// path_bounce / direct_light (shade.hpp) cut at each visibility test.  A pixel's samples run
// in ascending order on one lane and are added to its accumulation in that order, so the
// planes are bit-identical to the megakernel's.
//
// Work items are 8x8 tiles of one request (its job and planes come from per-slot tables).
// All lanes of a wave work for one request at a time, so the job's constants are read with
// wave-uniform scalar loads; a wave changes request only when all its lanes are idle.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "shade.hpp"

namespace syf {

namespace {

// Idle lanes that trigger a refill while others still work.  Measured on single served
// 1024^2 16-spp requests (Mrays/s, Mandelbulb / Mandelbox / Menger): 1 281 / 257 / 1064,
// 4 285 / 259 / 1047, 8 280 / 257 / 1036, 16 268 / 248 / 1000.
constexpr uint32_t PATH_REFILL = 4;
// Field evaluations (with their march steps) a lane runs per trip of the loop, STEPS (4, 8 or
// 16, chosen per kernel family by the server's autotuner): the loop's own bookkeeping is paid
// once per STEPS evaluations, at the cost of lanes whose march ends early in the trip idling
// for the rest of it (measured on single requests: 8 beats 1 by 1.4x; 4 and 16 are each best
// on some families).  The planes do not depend on it.

// The per-slot job table is read through the constant address space (job_ref): its loads are
// then known to be invariant, so the wave-uniform constants of the request being served are
// loaded with scalar loads the compiler may hoist and keep, as it does for kernel arguments.
__device__ __forceinline__ const JobP& job_of(const JobP* jobs, uint32_t slot) { return job_ref(jobs + slot); }

__device__ __forceinline__ uint64_t below_mask() {
    const uint32_t lane = threadIdx.x & 63u;
    return lane == 0u ? 0ull : (~0ull >> (64u - lane));
}


// Lane modes: evaluating (a march step or a normal tap in flight) or due for shading with one
// of four events.
enum : uint32_t {
    L_IDLE = 0,
    L_EXT,        // extension march
    L_SHADOW,     // shadow march
    L_NORMAL,     // normal taps
    L_EV_START,   // a sample is to start (path_start)
    L_EV_SEGMENT, // the extension march ended
    L_EV_NORMAL,  // the normal is known: shade the surface
    L_EV_SHADOW,  // a shadow march ended: continue next-event estimation
};

template <uint32_t FAM> struct Lane {
    uint32_t mode;
    // unit: pixel (py << 16 | px) of the wave's item, samples [s, s_end)
    uint32_t pix, s, s_end, s_first;
    float4 acc;
    float fd0, tp0;  // first-hit planes of the unit's first sample
    float4 fa0;
    PathState st;
    // the extension segment being shaded (how its march ended)
    float seg_t, seg_trap0;
    uint32_t seg_steps;
    bool seg_hit, seg_exhausted;
    uint32_t ext_light;  // emitter in front of the extension march (NO_LIGHT if none)
    // evaluation in flight; the march's origin is st.o (extension) or p_off (shadow)
    DeStepper<FAM> de;
    bool de_done;
    MarchState m;
    // normal taps; tap_p, g and tap_h share registers with p_off, dl and p_spec (the taps
    // end before the surface is shaded)
    uint32_t tap;
    // surface and next-event estimation; light_i holds the hit sample's iteration count from
    // the evaluation after the normal taps until the surface is shaded (next-event estimation
    // starts after that)
    Surf surf;
    F3 p_off, dl, pend;
    float p_spec, area_total, sel_w;
    uint32_t sel, light_i;
    bool last;
};

// Tap directions of normal_at (common.wgsl:376-387).
__device__ __forceinline__ F3 tap_dir(uint32_t i) {
    return mk((i == 0u || i == 3u) ? 1.0f : -1.0f, (i >= 2u) ? 1.0f : -1.0f, (i == 1u || i == 3u) ? 1.0f : -1.0f);
}

template <uint32_t FAM>
__device__ __forceinline__ void start_march(const FractalP& F, Lane<FAM>& L, F3 o, F3 d, float t_end, uint32_t mode) {
    // t_end > 0 on every call (the ray extent is >= 200, emitter distances exceed 1e-6 and
    // area-light shadow ends exceed 9e-5), so march() always evaluates its first position
    march_reset(L.m, o, d, t_end);
    L.mode = mode;
    L.de_done = L.de.start(F, add(o, scl(d, 0.0f)));
}

// The rest of direct_light (shade.hpp) from light L.light_i on, up to the next shadow ray:
// the contribution a shadow ray carries is computed before its march (no random draw follows
// the visibility test) and added when the march finds the light unoccluded.  Returns true
// when a shadow march was started; false when next-event estimation is done.
template <uint32_t FAM>
__device__ __forceinline__ bool continue_direct(const JobP& P, Lane<FAM>& L, uint32_t& shadow) {
    const F3 v = neg3(L.st.d);
    const Surf& s = L.surf;
    for (; L.light_i < P.light_count; ++L.light_i) {
        const LightP& l = P.lights[L.light_i];
        F3 dir;
        float t_end;
        if (l.kind == SYF_LIGHT_ENV) {
            const float mx = fmaxf(fmaxf(fmaxf(l.color[0], l.color[1]), l.color[2]),
                                   fmaxf(fmaxf(fmaxf(l.horizon[0], l.horizon[1]), l.horizon[2]),
                                         fmaxf(fmaxf(l.ground[0], l.ground[1]), l.ground[2])));
            if (mx <= 0.0f) continue;
            const float prob = environment_probability(l, L.area_total);
            if (rand01(L.st.rng) >= prob) continue;
            const float r1 = rand01(L.st.rng);
            const float r2 = rand01(L.st.rng);
            dir = cosine_sample(s.n, r1, r2);
            const float nl = fmaxf(dot3(s.n, dir), 0.0f);
            if (dot3(s.n, dir) <= 0.0f || nl <= 0.0f) continue;
            const float pl = prob * nl / PI_F;
            const float pb = L.last ? 0.0f : bsdf_pdf(s, v, dir, L.p_spec);
            const float w = L.last ? 1.0f : mis_power(pl, pb);
            const F3 f = brdf_eval(s, v, dir);
            const F3 e = environment_value(l, dir);
            L.pend = mk(f.x * nl * e.x * 1.0f * w / pl, f.y * nl * e.y * 1.0f * w / pl, f.z * nl * e.z * 1.0f * w / pl);
            t_end = P.t_end;
        } else if (l.kind == SYF_LIGHT_SUN) {
            const float r1 = rand01(L.st.rng);
            const float r2 = rand01(L.st.rng);
            dir = cone_sample(ld3(l.a), l.cos_max, r1, r2);
            const float nl = dot3(s.n, dir);
            if (nl <= 0.0f || dot3(s.n, dir) <= 0.0f) continue;
            const float pl = sun_pdf(l.cos_max);
            const float pb = L.last ? 0.0f : bsdf_pdf(s, v, dir, L.p_spec);
            const float w = L.last ? 1.0f : mis_power(pl, pb);
            const F3 f = brdf_eval(s, v, dir);
            L.pend = mk(f.x * nl * l.color[0] * w, f.y * nl * l.color[1] * w, f.z * nl * l.color[2] * w);
            t_end = P.t_end;
        } else if (l.kind == SYF_LIGHT_AREA) {
            if (L.light_i != L.sel) continue;
            const float u1 = rand01(L.st.rng);
            const float u2 = rand01(L.st.rng);
            const F3 q = add(add(ld3(l.a), scl(ld3(l.b), u1 - 0.5f)), scl(ld3(l.c), u2 - 0.5f));
            const F3 to_q = sub(q, L.p_off);
            const float dist2 = dot3(to_q, to_q);
            if (dist2 < 1e-8f) continue;
            const float dist = sqrtf(dist2);
            dir = mk(to_q.x / dist, to_q.y / dist, to_q.z / dist);
            const float nl = dot3(s.n, dir);
            const float cos_l = dot3(ld3(l.n), neg3(dir));
            if (nl <= 0.0f || cos_l <= 0.0f || dot3(s.n, dir) <= 0.0f) continue;
            t_end = fmaxf(0.0f, dist - fmaxf(1e-5f, dist * 1e-6f));
            const float pl = (L.sel_w / L.area_total) * (dist2 / fmaxf(l.area * cos_l, 1e-20f));
            const float pb = L.last ? 0.0f : bsdf_pdf(s, v, dir, L.p_spec);
            const float w = L.last ? 1.0f : mis_power(pl, pb);
            const F3 f = brdf_eval(s, v, dir);
            L.pend = mk(f.x * nl * l.color[0] * 1.0f * w / pl, f.y * nl * l.color[1] * 1.0f * w / pl,
                        f.z * nl * l.color[2] * 1.0f * w / pl);
        } else {
            continue;
        }
        // visible() (Tracer): an emitter in front of the shadow ray decides it without a march
        ++shadow;
        uint32_t light;
        const float nearest = nearest_emitter(P, L.p_off, dir, t_end, light);
        if (light != NO_LIGHT) continue;
        ++L.light_i;
        start_march(P.f, L, L.p_off, dir, nearest, L_SHADOW);
        return true;
    }
    return false;
}

// After next-event estimation: the continuation of path_bounce.  Returns true when the path
// goes on (st.o, st.d hold its next ray).
template <uint32_t FAM> __device__ __forceinline__ bool continue_path(const JobP& P, Lane<FAM>& L) {
    PathState& st = L.st;
    st.L = add(st.L, mul3(st.throughput, L.dl));
    if (L.last) return false;
    const Surf& s = L.surf;
    const F3 v = neg3(st.d);
    const float r0 = rand01(st.rng), r1 = rand01(st.rng), r2 = rand01(st.rng);
    F3 l;
    if (r0 < L.p_spec) l = reflect3(neg3(v), sample_ggx_h(s.n, v, s.alpha, r1, r2));
    else l = cosine_sample(s.n, r1, r2);
    const float nl = dot3(s.n, l);
    if (nl <= 0.0f || dot3(s.n, l) <= 0.0f) return false;
    const float pdf = bsdf_pdf(s, v, l, L.p_spec);
    if (pdf <= 1e-6f) return false;
    const F3 f = brdf_eval(s, v, l);
    st.throughput = mk(st.throughput.x * (f.x * nl / pdf), st.throughput.y * (f.y * nl / pdf),
                       st.throughput.z * (f.z * nl / pdf));
    st.prev_pdf = pdf;
    st.prev_from_bsdf = true;
    st.prev_n = s.n;
    st.prev_total = L.area_total;
    if (st.bounce >= 2u) {
        const float q = clampT(fmaxf(st.throughput.x, fmaxf(st.throughput.y, st.throughput.z)), 0.05f, 0.95f);
        if (rand01(st.rng) > q) return false;
        st.throughput = mk(st.throughput.x / q, st.throughput.y / q, st.throughput.z / q);
    }
    st.o = L.p_off;
    st.d = l;
    st.bounce++;
    return true;
}

// The surface part of path_bounce once the normal is known, up to next-event estimation (a
// fractal surface, or the ground plane).  The served families are classic graphs: a hit's
// sample is trap slot 0, unwritten slots (1) and material 0, with the iteration count the
// extra evaluation recorded when the shading reads it.
template <uint32_t FAM> __device__ __forceinline__ void shade_surface(const JobP& P, Lane<FAM>& L, F3 n, bool ground) {
    PathState& st = L.st;
    if (st.bounce == 0u) {
        if (L.s == L.s_first) {  // first-hit planes come from the unit's first sample
            L.fd0 = L.seg_t;
            L.fa0 = make_float4(n.x, n.y, n.z, float(L.seg_steps));
        }
        if (!ground) {
            const float t = clampT(float(L.seg_steps) / 96.0f, 0.0f, 1.0f);
            st.throughput = scl(st.throughput, 1.0f - P.cavity_darkness * (t * t * (3.0f - 2.0f * t)));
        }
    }
    const F3 p = add(st.o, scl(st.d, L.seg_t));
    const F3 v = neg3(st.d);
    ShadeSample smp = ground ? ground_sample(P) : carried_sample(P, L.seg_trap0);
    if (!ground && P.sh.full_sample) smp.iter = L.light_i;
    F3 emission;
    L.surf = make_surface<false>(P, p, n, smp, 1e-4f, emission);  // the server takes no finish detail patterns
    st.L = add(st.L, mul3(st.throughput, emission));
    const float offset = ground ? 1e-3f : fmaxf(L.seg_t * P.pixel_cone * 2.0f, 1e-3f) * P.pixel_tolerance;
    L.p_off = add(p, scl(n, offset));
    const float nv = fmaxf(dot3(L.surf.n, v), 1e-4f);
    L.p_spec = spec_probability(L.surf, nv);
    L.area_total = area_total_weight(P, L.p_off, L.surf.n);
    L.last = st.bounce == P.max_bounces;
    // direct_light: choose one area light by weight
    L.sel = NO_LIGHT;
    L.sel_w = 0.0f;
    if (L.area_total > 0.0f) {
        const float threshold = rand01(st.rng) * L.area_total;
        float sum = 0.0f;
        LIGHT_LOOP(i) {
            if (P.lights[i].kind != SYF_LIGHT_AREA) continue;
            const float w = area_weight(L.p_off, L.surf.n, P.lights[i]);
            sum += w;
            if (w > 0.0f) {
                L.sel = i;
                L.sel_w = w;
                if (threshold < sum) break;
            }
        }
    }
    L.dl = mk(0.0f, 0.0f, 0.0f);
    L.light_i = 0;
}

}  // namespace

// Waves per SIMD the register allocation of k_paths is held to (3: 168 VGPRs, no vector
// spills except one in the trigonometric Mandelbulb; measured 1.2x faster than the
// unconstrained 2-wave allocation).
constexpr int PATHS_WAVES = 3;

template <uint32_t FAM, uint32_t STEPS>
__global__ __launch_bounds__(256, PATHS_WAVES) void k_paths(const JobP* __restrict__ jobs, const BatchPlanes* __restrict__ planes,
                                               const BatchItem* __restrict__ items, uint32_t n_items,
                                               uint32_t* __restrict__ claim, unsigned long long* __restrict__ ctr,
                                               uint32_t* __restrict__ exit_count,
                                               unsigned long long* __restrict__ host_out) {
    const uint32_t lane = threadIdx.x & 63u;
    // The items are cut into eight contiguous ranges, one per XCD (a workgroup starts on the
    // range of blockIdx % 8, the XCD the dispatcher places it on): consecutive claims of a
    // wave stay within one request for long runs, and each range's counter sits in its own
    // 64-byte line.  A wave whose range is used up moves on to the next ones.
    uint32_t xl = blockIdx.x % 8u, ranges_left = 8;
    // wave-uniform: the request served, the item whose pixels are being handed out, and an
    // item of another request held until the wave drains
    uint32_t slot = ~0u;
    BatchItem pool{0, 0, 0, 0, 0, 3}, pending{0, 0, 0, 0, 0, 3};
    uint32_t pool_next = 64, have_pending = 0, exhausted = 0;
    Lane<FAM> L;
    L.mode = L_IDLE;
    L.de_done = false;
    Work work;
    uint32_t segs = 0, shadow = 0, prim = 0, hits = 0;
    for (;;) {
        // ---- refill: idle lanes take pixels of the wave's item
        const uint64_t idle_mask = __ballot(L.mode == L_IDLE);
        const uint32_t idle = uint32_t(__popcll(idle_mask));
        const bool any_busy = idle < 64u;
        if (idle && (idle >= PATH_REFILL || !any_busy)) {
            while (pool_next >= 64u && !have_pending && !exhausted) {
                const uint32_t r0 = uint32_t(uint64_t(n_items) * xl / 8u), r1 = uint32_t(uint64_t(n_items) * (xl + 1u) / 8u);
                uint32_t k = 0;
                if (lane == 0u) k = atomicAdd(&claim[xl * 16u], 1u);
                k = __builtin_amdgcn_readfirstlane(k);
                if (r0 + k >= r1) {
                    xl = (xl + 1u) % 8u;
                    if (--ranges_left == 0u) exhausted = 1;
                } else {
                    const BatchItem it = items[r0 + k];
                    if (__builtin_amdgcn_readfirstlane(it.slot) == slot) {
                        pool = it;
                        pool_next = 0;
                    } else {
                        pending = it;
                        have_pending = 1;
                    }
                }
            }
            if (pool_next >= 64u && have_pending && !any_busy) {  // drained: switch request
                pool = pending;
                slot = __builtin_amdgcn_readfirstlane(pending.slot);
                pool_next = 0;
                have_pending = 0;
            }
            if (pool_next < 64u) {
                const JobP& P = job_of(jobs, slot);
                const uint32_t rank = uint32_t(__popcll(idle_mask & below_mask()));
                if (L.mode == L_IDLE && pool_next + rank < 64u) {
                    const uint32_t u = pool_next + rank;
                    const uint32_t px = pool.x0 + (u & ((1u << pool.tws) - 1u)), py = pool.y0 + (u >> pool.tws);
                    if (px < P.width && py < P.height && pool.ns > 0u) {
                        L.pix = (py << 16) | px;
                        L.s = pool.s0;
                        L.s_first = pool.s0;
                        L.s_end = pool.s0 + pool.ns;
                        // the first pass starts the accumulation (the planes are not cleared)
                        L.acc = pool.s0 == 0u ? make_float4(0.0f, 0.0f, 0.0f, 0.0f)
                                              : planes[slot].accum[size_t(py) * P.width + px];
                        L.fd0 = -1.0f;
                        L.tp0 = -1.0f;
                        L.fa0 = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
                        L.mode = L_EV_START;
                    }
                }
                pool_next = min(64u, pool_next + idle);
            }
        }
        if (__ballot(L.mode != L_IDLE) == 0ull && pool_next >= 64u && !have_pending && exhausted) break;
        if (L.mode == L_IDLE) continue;
        // ---- evaluation: up to STEPS field evaluations per evaluating lane, each
        // followed by its march step or normal tap; a lane stops early when its march or
        // normal ends
        if (L.mode <= L_NORMAL) {
            const JobP& P = job_of(jobs, slot);
            const FractalP& F = P.f;
            for (uint32_t k = 0; k < STEPS && L.mode <= L_NORMAL; ++k) {
            while (!L.de_done) L.de_done = L.de.step(F);
            {
                const De<float> s = L.de.finish(F);
                work.de_evals++;
                work.de_iters += s.runs;
                if (L.mode == L_NORMAL) {
                    // normal_at: g += k * d over the four taps (tap_p = p_off, g = dl, h = p_spec);
                    // then, if the shading reads it, the hit's own sample for its iteration count
                    if (L.tap < 4u) {
                        L.dl = add(L.dl, scl(tap_dir(L.tap), s.d));
                        if (++L.tap < 4u) L.de_done = L.de.start(F, add(L.p_off, scl(tap_dir(L.tap), L.p_spec)));
                        else if (P.sh.full_sample) L.de_done = L.de.start(F, L.p_off);
                        else L.mode = L_EV_NORMAL;
                    } else {
                        L.light_i = s.iter;
                        L.mode = L_EV_NORMAL;
                    }
                } else {
                    work.steps++;
                    MarchEnd e;
                    if (!march_advance(L.m, s, F.safety, P.pixel_cone * P.pixel_tolerance, P.max_steps, e)) {
                        const F3 ro = L.mode == L_SHADOW ? L.p_off : L.st.o;
                        L.de_done = L.de.start(F, add(ro, scl(L.m.rd, L.m.t)));
                    } else if (L.mode == L_EXT) {
                        L.seg_hit = e.hit;
                        L.seg_exhausted = e.exhausted;
                        L.seg_t = e.t;
                        L.seg_trap0 = e.trap0;
                        L.seg_steps = e.steps;
                        if (e.hit && L.st.bounce == 0u) {
                            ++hits;
                            if (L.s == L.s_first) L.tp0 = e.t_prev;
                        }
                        L.mode = L_EV_SEGMENT;
                    } else {
                        if (!e.hit && !e.exhausted) L.dl = add(L.dl, L.pend);
                        L.mode = L_EV_SHADOW;
                    }
                }
            }
            }
        }
        // ---- shading: a lane whose march or normal ended takes its path up to its next field
        // evaluation.  (Holding such lanes back until several have one, so that the wave runs
        // the shading code once for all of them, measured slower: the lanes idle meanwhile.)
        if (L.mode < L_EV_START) continue;
        const JobP& P = job_of(jobs, slot);
        PathState& st = L.st;
        bool nee = false, new_sample = false, path_end = false, new_segment = false;
        if (L.mode == L_EV_START) {
            new_sample = true;
        } else if (L.mode == L_EV_SEGMENT) {
            // path_bounce up to the normal (trace_segment turns a miss in front of an
            // emitter into a hit on it)
            if (!L.seg_hit && !L.seg_exhausted && L.ext_light == GROUND_HIT) {
                shade_surface(P, L, ground_normal(st.d), true);
                nee = true;
            } else if (!L.seg_hit && !L.seg_exhausted && L.ext_light != NO_LIGHT) {
                const F3 gn = light_normal(P, L.ext_light);
                if (st.bounce == 0u && L.s == L.s_first) {
                    L.fd0 = L.seg_t;
                    L.fa0 = make_float4(gn.x, gn.y, gn.z, float(L.seg_steps));
                }
                st.L = add(st.L, mul3(st.throughput, area_emission(P, st.o, st.d, L.seg_t, st.prev_pdf,
                                                                   st.prev_from_bsdf, st.prev_n, st.prev_total)));
                path_end = true;
            } else if (!L.seg_hit) {
                if (!L.seg_exhausted) {
                    st.L = add(st.L, mul3(st.throughput, environment_emission(P, st.d, st.prev_pdf, st.prev_from_bsdf,
                                                                              st.prev_n, st.prev_total)));
                    st.L = add(st.L, mul3(st.throughput, sun_radiance(P, st.d, st.prev_pdf, st.prev_from_bsdf)));
                }
                path_end = true;
            } else {  // normal_at: four taps around the hit point
                L.p_off = add(st.o, scl(st.d, L.seg_t));
                L.p_spec = fmaxf(L.seg_t * P.pixel_cone * 0.5f, 1e-4f) * P.pixel_tolerance;
                L.dl = mk(0.0f, 0.0f, 0.0f);
                L.tap = 0;
                L.mode = L_NORMAL;
                L.de_done = L.de.start(P.f, add(L.p_off, scl(tap_dir(0u), L.p_spec)));
            }
        } else if (L.mode == L_EV_NORMAL) {
            shade_surface(P, L, nrm(L.dl), false);
            nee = true;
        } else {  // L_EV_SHADOW
            nee = true;
        }
        if (nee && !continue_direct(P, L, shadow)) {
            if (continue_path(P, L)) new_segment = true;
            else path_end = true;
        }
        if (path_end) {  // accumulate the sample: accum += (rgb if finite, 1)
            const F3 r = st.L;
            const bool ok = r.x == r.x && r.y == r.y && r.z == r.z && fabsf(r.x) < 1e30f && fabsf(r.y) < 1e30f &&
                            fabsf(r.z) < 1e30f;
            L.acc.x += ok ? r.x : 0.0f;
            L.acc.y += ok ? r.y : 0.0f;
            L.acc.z += ok ? r.z : 0.0f;
            L.acc.w += 1.0f;
            if (++L.s < L.s_end) {
                new_sample = true;
            } else {  // the pixel's samples are done: store its planes
                const BatchPlanes pl = planes[slot];
                const size_t idx = size_t(L.pix >> 16) * P.width + (L.pix & 0xffffu);
                pl.accum[idx] = L.acc;
                if (L.s_first == 0u) {
                    pl.depth[idx] = L.fd0;
                    pl.aux[idx] = L.fa0;
                    // the refine bracket, for k_refine_first_hits (-1: no surface hit to refine)
                    pl.depth64[idx] = L.fd0 >= 0.0f && L.tp0 >= 0.0f ? double(L.tp0) : -1.0;
                }
                L.mode = L_IDLE;
            }
        }
        if (new_sample) {
            path_start(P, L.pix & 0xffffu, L.pix >> 16, L.s, st);
            ++prim;
            new_segment = true;
        }
        if (new_segment) {  // trace_segment: the nearest emitter caps the extension march
            ++segs;
            const float nearest = nearest_emitter(P, st.o, st.d, P.t_end, L.ext_light);
            start_march(P.f, L, st.o, st.d, nearest, L_EXT);
        }
    }
    flush_work(work, hits, segs + shadow, ctr);
    const uint32_t p = wave_sum(prim);
    if (lane == 0u) {
        atomicAdd(&ctr[CTR_PRIMARY], (unsigned long long)p);
        // The launch's last wave (every wave of the grid counts itself out once its counts are
        // in) hands the counters to the host and clears them, the claim counters and the exit
        // count for the launch that reuses them: no clearing kernel runs between launches.
        const uint32_t waves = gridDim.x * (blockDim.x >> 6);
        if (__hip_atomic_fetch_add(exit_count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == waves - 1u) {
            for (uint32_t c = 0; c < CTR_COUNT; ++c)
                host_out[c] = __hip_atomic_exchange(&ctr[c], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (uint32_t x = 0; x < 8u; ++x) __hip_atomic_store(&claim[x * 16u], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(exit_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");  // system scope: the host reads host_out
        }
    }
}

// FP64 refine of the first surface hits of the items that start at sample 0: k_paths left the
// march's last position before the hit (the lower end of the bracket) in the FP64 plane, or -1
// where sample 0 hit no surface.  One wave per item, one lane per pixel; the camera ray of
// sample 0 is regenerated from its PCG stream (as the primary trace does for its hits).
template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_refine_first_hits(const JobP* __restrict__ jobs,
                                                           const BatchPlanes* __restrict__ planes,
                                                           const BatchItem* __restrict__ items, uint32_t n_items,
                                                           unsigned long long* __restrict__ ctr) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    Work w;
    for (uint32_t k = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); k < n_items; k += waves) {
        const BatchItem it = items[k];
        if (it.s0 != 0u || it.ns == 0u) continue;
        const uint32_t slot = __builtin_amdgcn_readfirstlane(it.slot);
        const JobP& P = job_of(jobs, slot);
        const BatchPlanes pl = planes[slot];
        const uint32_t px = it.x0 + (lane & ((1u << it.tws) - 1u)), py = it.y0 + (lane >> it.tws);
        if (px >= P.width || py >= P.height) continue;
        const size_t idx = size_t(py) * P.width + px;
        const double tp = pl.depth64[idx];
        if (!(tp >= 0.0)) continue;
        uint32_t rng = rng_init(px, py, 0u, P.seed);
        F3 ro, rd;
        camera_ray(P.cam, P.width, P.height, px, py, rng, ro, rd);
        // tp is the float t_prev the flattened march stored (exact in a double)
        pl.depth64[idx] = refine64_of_march<FAM>(P.f, ro, rd, float(tp), pl.depth[idx],
                                                 double(P.pixel_cone) * double(P.pixel_tolerance), w);
    }
    flush_work(w, 0u, 0u, ctr);
}

hipError_t launch_refine_first_hits(uint32_t fam, const JobP* jobs, const BatchPlanes* planes,
                                    const BatchItem* items, uint32_t n_items, unsigned long long* ctr,
                                    hipStream_t s) {
    if (n_items == 0) return hipSuccess;
    if (fam >= FAM_COUNT) return hipErrorInvalidValue;
    const dim3 grid(std::min<uint32_t>((n_items + 3u) / 4u, 4096u)), block(256);
    switch (fam) {
    case SYF_FAMILY_MANDELBULB: hipLaunchKernelGGL(k_refine_first_hits<SYF_FAMILY_MANDELBULB>, grid, block, 0, s, jobs, planes, items, n_items, ctr); break;
    case SYF_FAMILY_MANDELBOX: hipLaunchKernelGGL(k_refine_first_hits<SYF_FAMILY_MANDELBOX>, grid, block, 0, s, jobs, planes, items, n_items, ctr); break;
    case SYF_FAMILY_MENGER: hipLaunchKernelGGL(k_refine_first_hits<SYF_FAMILY_MENGER>, grid, block, 0, s, jobs, planes, items, n_items, ctr); break;
    case FAM_BULB_TRIG: hipLaunchKernelGGL(k_refine_first_hits<FAM_BULB_TRIG>, grid, block, 0, s, jobs, planes, items, n_items, ctr); break;
    case FAM_BULB8: hipLaunchKernelGGL(k_refine_first_hits<FAM_BULB8>, grid, block, 0, s, jobs, planes, items, n_items, ctr); break;
    default: hipLaunchKernelGGL(k_refine_first_hits<SYF_FAMILY_SPHERE>, grid, block, 0, s, jobs, planes, items, n_items, ctr); break;
    }
    return hipGetLastError();
}

// Writes a request's entries of the slot tables; the entries travel as kernel arguments.
__global__ __launch_bounds__(64) void k_admit(JobP* __restrict__ jobs, BatchPlanes* __restrict__ planes, uint32_t slot,
                                              const JobP P, const BatchPlanes pl) {
    constexpr uint32_t words = sizeof(JobP) / 4u;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&P);
    uint32_t* dst = reinterpret_cast<uint32_t*>(jobs + slot);
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
    if (threadIdx.x == 0u) planes[slot] = pl;
}

hipError_t launch_admit(JobP* jobs, BatchPlanes* planes, uint32_t slot, const JobP& P, const BatchPlanes& pl,
                        hipStream_t s) {
    static_assert(sizeof(JobP) % 4u == 0u, "JobP is copied in words");
    hipLaunchKernelGGL(k_admit, dim3(1), dim3(64), 0, s, jobs, planes, slot, P, pl);
    return hipGetLastError();
}

template <uint32_t STEPS>
static void launch_paths_steps(uint32_t fam, dim3 grid, const JobP* jobs, const BatchPlanes* planes,
                               const BatchItem* items, uint32_t n_items, uint32_t* claim, unsigned long long* ctr,
                               uint32_t* exit_count, unsigned long long* host_out, hipStream_t s) {
    const dim3 block(256);
    switch (fam) {
    case SYF_FAMILY_MANDELBULB: hipLaunchKernelGGL((k_paths<SYF_FAMILY_MANDELBULB, STEPS>), grid, block, 0, s, jobs, planes, items, n_items, claim, ctr, exit_count, host_out); break;
    case SYF_FAMILY_MANDELBOX: hipLaunchKernelGGL((k_paths<SYF_FAMILY_MANDELBOX, STEPS>), grid, block, 0, s, jobs, planes, items, n_items, claim, ctr, exit_count, host_out); break;
    case SYF_FAMILY_MENGER: hipLaunchKernelGGL((k_paths<SYF_FAMILY_MENGER, STEPS>), grid, block, 0, s, jobs, planes, items, n_items, claim, ctr, exit_count, host_out); break;
    case FAM_BULB_TRIG: hipLaunchKernelGGL((k_paths<FAM_BULB_TRIG, STEPS>), grid, block, 0, s, jobs, planes, items, n_items, claim, ctr, exit_count, host_out); break;
    case FAM_BULB8: hipLaunchKernelGGL((k_paths<FAM_BULB8, STEPS>), grid, block, 0, s, jobs, planes, items, n_items, claim, ctr, exit_count, host_out); break;
    default: hipLaunchKernelGGL((k_paths<SYF_FAMILY_SPHERE, STEPS>), grid, block, 0, s, jobs, planes, items, n_items, claim, ctr, exit_count, host_out); break;
    }
}

hipError_t launch_paths(uint32_t fam, const JobP* jobs, const BatchPlanes* planes, const BatchItem* items,
                        uint32_t n_items, uint32_t* claim, unsigned long long* ctr, uint32_t* exit_count,
                        unsigned long long* host_out, uint32_t grid_blocks, uint32_t steps, hipStream_t s) {
    if (n_items == 0) return hipSuccess;
    if (fam >= FAM_COUNT) return hipErrorInvalidValue;
    const uint32_t waves = (n_items + 7u) / 8u * 8u;
    const dim3 grid(std::max<uint32_t>(8u, std::min<uint32_t>(grid_blocks, (waves + 3u) / 4u)));
    if (steps <= 4u)
        launch_paths_steps<4>(fam, grid, jobs, planes, items, n_items, claim, ctr, exit_count, host_out, s);
    else if (steps <= 8u)
        launch_paths_steps<8>(fam, grid, jobs, planes, items, n_items, claim, ctr, exit_count, host_out, s);
    else
        launch_paths_steps<16>(fam, grid, jobs, planes, items, n_items, claim, ctr, exit_count, host_out, s);
    return hipGetLastError();
}

}  // namespace syf
