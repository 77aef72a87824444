// Device code of the primary-ray trace and the distance probe (k_primary, k_primary_exact,
// k_hits_tail, k_de_probe), templated on the kernel family.  Included by trace.hip, which
// instantiates and launches the library's families, and by the run-time compiled units of
// specialised graph topologies (graph.hip), which instantiate the FAM_JIT family.
#pragma once

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
// Measured at 1080p (ms per frame): 32 rays 1.00, 64 0.907, 128 0.882, 256 0.94 (256 also
// leaves 650-700 hits to the trailing pass).
constexpr uint32_t RAY_CHUNK = 128;
constexpr uint32_t RAY_LANES = 8;
// Hits are staged per wave in LDS and published to the hit queue 64 at a time.
constexpr uint32_t HIT_BATCH = 64;
// A wave generates new camera rays when at least this many of its lanes are idle (or none is
// marching): ray generation is a branch every lane of the wave pays for, so it is batched.
// Measured at 1080p (ms per frame): 4 0.910, 8 0.890, 16 0.877, 24 0.884, 32 0.898.
constexpr uint32_t REFILL_BATCH = 16;
// Issue priorities: marching waves are served first; hit consumers take the slots left over.
constexpr short PRIO_MARCH = 1, PRIO_HITS = 0;
// Queue blocks a shading wave claims at once.  Measured at 1080p (ms per frame, frames back to
// back): 2 0.973, 3 0.911, 4 0.950 (with 4, claims run past the reserved blocks at the end
// and leave 45-57 hits to the trailing pass; at 8 that took the frame to 1.47 ms).
constexpr uint32_t SHADE_CLAIM = 3;

__device__ __forceinline__ uint64_t lanes_below() {
    const uint32_t lane = threadIdx.x & 63u;
    return lane == 0u ? 0ull : (~0ull >> (64u - lane));
}

// The hit queue's entries cross XCDs, whose L2 caches are not coherent with each other:
// fields are written and read with agent-scope atomic accesses (coherent at the device), and
// every field starts as HITQ_EMPTY (filled once, then emptied by each trace's trailing pass).
// A field is thus either empty or final, and an entry whose fields all differ from HITQ_EMPTY
// is complete; no fence is needed.  No field can be HITQ_EMPTY when written: pixel < 2^32 - 2,
// distances are finite floats, radii have a clear sign bit, steps <= max_steps.  Entries are
// reserved in blocks of 64; a wave's last, partial block is padded with HITQ_SKIP entries.
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_agent(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Normal and FP64 refine of one hit.  The camera ray is regenerated from its PCG stream (the
// FP32 ray the march followed); the normal takes four FP32 taps at the hit point
// (common.wgsl:376-387, h = max(t * cone / 2, 1e-4) * pixel_tolerance); the refine locates the
// hit rule's crossing in FP64 between the march's last two positions, whose radii the march
// queued with the hit.
template <uint32_t FAM, bool FAST>
__device__ __forceinline__ void shade_hit(const JobP& P, uint32_t sample, uint32_t pix, float t, float t_prev,
                                          float r_prev, float r_hit, uint32_t steps, float4* __restrict__ aux,
                                          double* __restrict__ depth64, Work& w64, Work& wn) {
    const uint32_t px = pix % P.width, py = pix / P.width;
    uint32_t rng = rng_init(px, py, sample, P.seed);
    V3<float> ro, rd;
    camera_ray(P.cam, P.width, P.height, px, py, rng, ro, rd);
    const V3<float> p = add(ro, scl(rd, t));
    const V3<float> nv =
        normal_at<float, FAM, FAST>(P.f, p, fmaxf(t * P.pixel_cone * 0.5f, 1e-4f) * P.pixel_tolerance, wn);
    aux[pix] = make_float4(nv.x, nv.y, nv.z, float(steps));
    const double cone64 = double(P.pixel_cone) * double(P.pixel_tolerance);
    depth64[pix] = refine64<FAM>(P.f, cvt<double>(ro), cvt<double>(rd), double(t_prev), double(t),
                                 refine_end_g(r_prev, t_prev, cone64), refine_end_g(r_hit, t, cone64), cone64, w64);
}

// Global ray index of slot j of ray lane `xl` (tiles dealt round-robin to the lanes).
__device__ __forceinline__ uint32_t lane_ray(uint32_t j, uint32_t xl) {
    return (((j >> 6) * RAY_LANES + xl) << 6) | (j & 63u);
}

// Publishes the first `n` staged hits of a wave (n <= 64) as one block of the hit queue: one
// atomic reserves the block, each lane writes one entry's fields (lanes >= n a skip entry).
__device__ __forceinline__ void publish_hits(const uint32_t (*stage)[2 * HIT_BATCH], uint32_t n, const HitQueue& Q,
                                             uint32_t lane) {
    uint32_t base = 0;
    if (lane == 0u) base = atomicAdd(Q.count, HIT_BATCH);
    base = __shfl(base, 0, 64);
    if (base + HIT_BATCH > Q.capacity) {  // cannot happen at the capacity hitq_capacity() gives
        if (lane == 0u) atomicAdd(Q.error, 1u);
        return;
    }
    const uint32_t k = base + lane;
    const bool real = lane < n;
    st_agent(Q.t + k, real ? stage[1][lane] : 0u);
    st_agent(Q.t_prev + k, real ? stage[2][lane] : 0u);
    st_agent(Q.steps + k, real ? stage[3][lane] : 0u);
    st_agent(Q.r_prev + k, real ? stage[4][lane] : 0u);
    st_agent(Q.r_hit + k, real ? stage[5][lane] : 0u);
    st_agent(Q.pix + k, real ? stage[0][lane] : HITQ_SKIP);
}

// FP32 march of all primary rays, then hit shading.  March: every trip of the loop runs at
// most two Iterate bodies on each lane; a lane whose field evaluation completes in the trip
// applies the fence and advances its march (the arithmetic of march(), in the same order),
// queueing a hit or writing a miss when the ray ends.  Idle lanes take new rays from a
// wave-owned chunk of the launch's ray counter.  Deep and shallow evaluations, long and short
// rays share a wave without waiting for each other.  Hits: once no ray is left for any of its
// lanes, the wave claims runs of complete queue entries (one per lane) at the queue's cursor
// and shades them, until it finds none complete; it never waits for one.
// (The unroll request on the claim loop below is declined for the larger fields; that is
// expected, so the compiler's note about it is not shown.)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wpass-failed"
// FAST selects the FP32 arithmetic (SYF_MATH_FAST: DeStepperFast for the march and the normal
// taps, a hardware reciprocal in the hit rule's ratio; otherwise the exact arithmetic of the
// oracle's FP32 twin).  The FP64 refine is the same in both: with reciprocal and reciprocal
// square root estimates refined by Newton steps instead of the correctly rounded FP64 division
// and square root it took the 1080p frame to 0.766 against 0.727 ms (measured).
template <uint32_t FAM, bool FAST>
__device__ __forceinline__ void primary_trace(const JobP& P, uint32_t sample, uint32_t tiles_x, uint32_t n_slots,
                                              uint32_t* __restrict__ next, float* __restrict__ depth,
                                              float4* __restrict__ aux, double* __restrict__ depth64,
                                              uint32_t* __restrict__ steps_out, const HitQueue& Q,
                                              unsigned long long* __restrict__ ctr) {
    enum : uint32_t { IDLE = 0, MARCH = 1, DONE = 2 };
    // per wave: staged hits' pixel, t, t before the hit, steps and the radii at both
    __shared__ uint32_t hit_stage[4][HITQ_FIELDS][2 * HIT_BATCH];
    __builtin_amdgcn_s_setprio(PRIO_MARCH);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t (*stage)[2 * HIT_BATCH] = hit_stage[wave];
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
    cond_t<FAST, DeStepperFast<float, FAM>, DeStepper<FAM>> de;
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
        // up to two Iterate bodies per trip: the march step below then runs for half as many
        // trips of a deep evaluation (1080p frame 0.693 against 0.710 ms with one)
        if (active && !de_done) {
            de_done = de.step(F);
            if (!de_done) de_done = de.step(F);
        }
        bool hit_now = false;
        uint32_t hit_steps = 0;
        float t_hit = 0.0f, tp_hit = 0.0f, r_hit = 0.0f, rp_hit = 0.0f;
        if (active && de_done) {
            // the march step of march() in straight-line form: every decision is a select, so
            // the lanes that finished an evaluation run one branch-free block (the reference's
            // order of tests: back-step, best candidate, hit, t_end guard, step budget)
            const De<float> s = de.finish(F);
            w.de_evals++;
            w.de_iters += s.runs;
            w.steps++;
            const uint32_t nsteps = i + 1u;  // march() reports i + 1 for the step just evaluated
            const float radius = fabsf(s.d) * safety;
            const bool sor_fail = omega > 1.0f && (radius + prev_radius) < step_len * 1.0001f;
            step_len = sor_fail ? step_len - omega * step_len : radius * omega;
            omega = sor_fail ? 1.0f : omega;
            rp_hit = prev_radius;  // the radius at t_last
            r_hit = radius;
            prev_radius = radius;
            const float err = FAST ? radius * rcp_hw(fmaxf(t, 1e-6f)) : radius / fmaxf(t, 1e-6f);
            const bool better = !sor_fail && err < cand_err;
            cand_t = better ? t : cand_t;
            cand_err = better ? err : cand_err;
            const bool is_hit = !sor_fail && err < hit_cone;
            const bool guard = !sor_fail && !is_hit && t + step_len >= P.t_end;
            const bool guard_miss = guard && radius >= P.t_end - t;
            omega = guard ? 1.0f : omega;
            step_len = guard ? radius : step_len;
            const bool advance = !is_hit && !guard_miss;
            const float t_next = t + step_len;
            const uint32_t i_next = i + 1u;
            const bool out_of_steps = advance && i_next >= P.max_steps;
            const bool cand_hit = out_of_steps && cand_err < hit_cone * 4.0f;
            const bool past_end = advance && !out_of_steps && t_next >= P.t_end;
            const uint32_t outcome =  // 0 continue, 1 hit, 2 miss
                (is_hit || cand_hit) ? 1u : ((guard_miss || out_of_steps || past_end) ? 2u : 0u);
            t_hit = cand_hit ? cand_t : t;
            tp_hit = cand_hit ? cand_t : t_last;  // best-candidate hit: no refine bracket
            t_last = advance ? t : t_last;
            t = advance ? t_next : t;
            i = advance ? i_next : i;
            const bool next_done = de.start(F, add(ro, scl(rd, t)));  // unused once the ray has ended
            de_done = outcome == 0u ? next_done : de_done;
            if (outcome != 0u) {
                steps_out[pix] = nsteps;
                if (outcome == 1u) {
                    depth[pix] = t_hit;
                    hit_now = true;
                    hit_steps = nsteps;
                    ++hits;
                } else {
                    depth[pix] = -1.0f;
                    depth64[pix] = -1.0;
                    aux[pix] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
                }
                mode = IDLE;
            }
        }
        // stage this trip's hits in LDS; publish them 64 at a time (one atomic per batch)
        const uint64_t hm = __ballot(hit_now);
        if (hm) {
            if (hit_now) {
                const uint32_t at = staged + uint32_t(__popcll(hm & lanes_below()));
                stage[0][at] = pix;
                stage[1][at] = __float_as_uint(t_hit);
                stage[2][at] = __float_as_uint(tp_hit);
                stage[3][at] = hit_steps;
                stage[4][at] = __float_as_uint(rp_hit);
                stage[5][at] = __float_as_uint(r_hit);
            }
            staged += uint32_t(__popcll(hm));
            if (staged >= HIT_BATCH) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                publish_hits(stage, HIT_BATCH, Q, lane);
                const uint32_t rest = staged - HIT_BATCH;
                uint32_t carry[HITQ_FIELDS];
#pragma unroll
                for (int f = 0; f < int(HITQ_FIELDS); ++f) carry[f] = lane < rest ? stage[f][HIT_BATCH + lane] : 0u;
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if (lane < rest) {
#pragma unroll
                    for (int f = 0; f < int(HITQ_FIELDS); ++f) stage[f][lane] = carry[f];
                }
                staged = rest;
            }
        }
    }
    if (staged) {  // the wave's last hits
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        publish_hits(stage, staged, Q, lane);
    }
    // Hit shading, at the lower priority.  A claim takes the block at the cursor with one
    // fetch-and-add (never retried); a wave stops when the cursor has caught up with the blocks
    // reserved.  A claimed block that is not reserved yet, or whose entries are not all
    // complete, is left without its done flag: the trailing pass shades it.
    __builtin_amdgcn_s_setprio(PRIO_HITS);
    Work wn;  // normal taps (counted apart from the march's evaluations)
    const uint32_t cap_blocks = Q.capacity / HIT_BATCH, npix = P.width * P.height;
    const double cone64 = double(P.pixel_cone) * double(P.pixel_tolerance);
    for (uint32_t guard = 0; guard < cap_blocks; ++guard) {  // every trip claims distinct blocks
        // SHADE_CLAIM blocks per claim, one hit of each per lane.  All lanes first take the
        // normals of their hits (the same work on every lane); then each lane runs the FP64
        // refines of its hits back to back, one field evaluation per trip, so a wave takes the
        // longest lane's sum of refines rather than the sum of each block's longest refine
        // (a refine takes 0 to 40 FP64 evaluations).
        uint32_t b0 = 0, reserved = 0;
        if (lane == 0u) {
            reserved = ld_agent(Q.count) / HIT_BATCH;
            b0 = ld_agent(Q.cursor) < reserved ? atomicAdd(Q.cursor, SHADE_CLAIM) : HITQ_EMPTY;
            if (b0 != HITQ_EMPTY && b0 + SHADE_CLAIM > reserved) reserved = ld_agent(Q.count) / HIT_BATCH;
        }
        b0 = __shfl(b0, 0, 64);
        reserved = __shfl(reserved, 0, 64);
        if (b0 >= reserved || b0 >= cap_blocks) break;
        uint32_t h_pix[SHADE_CLAIM];
        float h_t[SHADE_CLAIM], h_tp[SHADE_CLAIM], h_rp[SHADE_CLAIM], h_rh[SHADE_CLAIM];
        // unrolled where the field is small enough (the power-8 Mandelbulb, measured)
#pragma unroll
        for (uint32_t j = 0; j < SHADE_CLAIM; ++j) {
            h_pix[j] = HITQ_SKIP;
            const uint32_t b = b0 + j;
            // a block not reserved yet, or still being written, is left to the trailing pass
            if (b >= reserved || b >= cap_blocks) continue;
            const uint32_t k = b * HIT_BATCH + lane;
            const uint32_t e_pix = ld_agent(Q.pix + k), e_t = ld_agent(Q.t + k), e_tp = ld_agent(Q.t_prev + k),
                           e_st = ld_agent(Q.steps + k), e_rp = ld_agent(Q.r_prev + k), e_rh = ld_agent(Q.r_hit + k);
            const bool ready = e_pix != HITQ_EMPTY && e_t != HITQ_EMPTY && e_tp != HITQ_EMPTY && e_st != HITQ_EMPTY &&
                               e_rp != HITQ_EMPTY && e_rh != HITQ_EMPTY;
            if (__ballot(ready) != ~0ull) continue;
            if (lane == 0u) Q.done[b] = Q.epoch;
            if (e_pix >= npix) continue;  // a padding entry
            h_pix[j] = e_pix;
            h_t[j] = __uint_as_float(e_t);
            h_tp[j] = __uint_as_float(e_tp);
            h_rp[j] = __uint_as_float(e_rp);
            h_rh[j] = __uint_as_float(e_rh);
            const uint32_t px = e_pix % P.width, py = e_pix / P.width;
            uint32_t rng = rng_init(px, py, sample, P.seed);
            V3<float> ro, rd;
            camera_ray(P.cam, P.width, P.height, px, py, rng, ro, rd);
            const V3<float> nv = normal_at<float, FAM, FAST>(P.f, add(ro, scl(rd, h_t[j])),
                                                       fmaxf(h_t[j] * P.pixel_cone * 0.5f, 1e-4f) * P.pixel_tolerance, wn);
            aux[e_pix] = make_float4(nv.x, nv.y, nv.z, float(e_st));
        }
        // the refines, one evaluation per trip (Refine64: refine64's steps)
        uint32_t j = 0;
        bool refining = false;
        Refine64 rf;
        V3<double> ro64 = mk(0.0, 0.0, 0.0), rd64 = ro64;
        for (uint32_t trip = 0; trip <= SHADE_CLAIM * uint32_t(REFINE_ITERS); ++trip) {
            // a lane between hits starts its next one
            while (!refining && j < SHADE_CLAIM) {
                const uint32_t pix = h_pix[j];
                double out = 0.0;
                if (pix != HITQ_SKIP &&
                    !rf.begin(double(h_tp[j]), double(h_t[j]), refine_end_g(h_rp[j], h_tp[j], cone64),
                              refine_end_g(h_rh[j], h_t[j], cone64), out)) {
                    const uint32_t px = pix % P.width, py = pix / P.width;
                    uint32_t rng = rng_init(px, py, sample, P.seed);
                    V3<float> ro, rd;
                    camera_ray(P.cam, P.width, P.height, px, py, rng, ro, rd);
                    ro64 = cvt<double>(ro);
                    rd64 = cvt<double>(rd);
                    refining = true;
                } else {
                    if (pix != HITQ_SKIP) depth64[pix] = out;  // no evaluation needed (no bracket: the march t)
                    ++j;
                }
            }
            if (__ballot(refining) == 0ull) break;
            if (!refining) continue;
            double out;
            if (rf.update(refine_g<FAM>(F, ro64, rd64, rf.x, cone64, w), out)) {
                depth64[h_pix[j]] = out;
                ++j;
                refining = false;
            }
        }
    }
    flush_work(w, hits, rays, ctr);
    const uint32_t pr = wave_sum(rays), a = wave_sum(wn.de_evals), b = wave_sum(wn.de_iters);
    if (lane == 0u) {
        atomicAdd(&ctr[CTR_PRIMARY], (unsigned long long)pr);
        atomicAdd(&ctr[CTR_HIT_EVALS], (unsigned long long)a);
        atomicAdd(&ctr[CTR_HIT_ITERS], (unsigned long long)b);
    }
}

template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_primary(const JobP P, uint32_t sample, uint32_t tiles_x, uint32_t n_slots,
                                                 uint32_t* __restrict__ next, float* __restrict__ depth,
                                                 float4* __restrict__ aux, double* __restrict__ depth64,
                                                 uint32_t* __restrict__ steps_out, const HitQueue Q,
                                                 unsigned long long* __restrict__ ctr) {
    primary_trace<FAM, true>(P, sample, tiles_x, n_slots, next, depth, aux, depth64, steps_out, Q, ctr);
}

template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_primary_exact(const JobP P, uint32_t sample, uint32_t tiles_x,
                                                       uint32_t n_slots, uint32_t* __restrict__ next,
                                                       float* __restrict__ depth, float4* __restrict__ aux,
                                                       double* __restrict__ depth64, uint32_t* __restrict__ steps_out,
                                                       const HitQueue Q, unsigned long long* __restrict__ ctr) {
    primary_trace<FAM, false>(P, sample, tiles_x, n_slots, next, depth, aux, depth64, steps_out, Q, ctr);
}

#pragma clang diagnostic pop

// The hits no wave of k_primary shaded: the entries of every reserved block without this
// launch's done flag, one per lane, each lane striding over several (a lane's successive hits
// lie far apart, so the expensive refines of silhouette regions spread over all lanes).  The
// queue is complete and coherent at this launch's start.  Every reserved entry is then emptied
// for the next trace, so the queue needs no clearing between traces.  Counted apart
// (CTR_TAIL_*) so that the fused kernel's own work is known.
template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_hits_tail(const JobP P, uint32_t sample, const HitQueue Q,
                                                   float4* __restrict__ aux, double* __restrict__ depth64,
                                                   unsigned long long* __restrict__ ctr) {
    const uint32_t n = min(*Q.count, Q.capacity);
    Work w, wn;
    uint32_t tail = 0;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t pix = Q.pix[k];
        // not shaded by k_primary, and not a padding entry
        if (Q.done[k / HIT_BATCH] != Q.epoch && pix < P.width * P.height) {
            // the job's arithmetic, as k_primary's shading waves (the planes must not depend on
            // which kernel shaded a hit)
            if (P.march_math == SYF_MATH_EXACT)
                shade_hit<FAM, false>(P, sample, pix, __uint_as_float(Q.t[k]), __uint_as_float(Q.t_prev[k]),
                                      __uint_as_float(Q.r_prev[k]), __uint_as_float(Q.r_hit[k]), Q.steps[k], aux,
                                      depth64, w, wn);
            else
                shade_hit<FAM, true>(P, sample, pix, __uint_as_float(Q.t[k]), __uint_as_float(Q.t_prev[k]),
                                     __uint_as_float(Q.r_prev[k]), __uint_as_float(Q.r_hit[k]), Q.steps[k], aux,
                                     depth64, w, wn);
            ++tail;
        }
        Q.pix[k] = HITQ_EMPTY;
        Q.t[k] = HITQ_EMPTY;
        Q.t_prev[k] = HITQ_EMPTY;
        Q.steps[k] = HITQ_EMPTY;
        Q.r_prev[k] = HITQ_EMPTY;
        Q.r_hit[k] = HITQ_EMPTY;
    }
    const uint32_t a = wave_sum(wn.de_evals), b = wave_sum(wn.de_iters), c = wave_sum(w.de64_evals),
                   d = wave_sum(w.de64_iters), h = wave_sum(tail);
    if ((threadIdx.x & 63u) == 0u && h) {
        atomicAdd(&ctr[CTR_HIT_EVALS], (unsigned long long)a);
        atomicAdd(&ctr[CTR_HIT_ITERS], (unsigned long long)b);
        atomicAdd(&ctr[CTR_DE64_EVALS], (unsigned long long)c);
        atomicAdd(&ctr[CTR_DE64_ITERS], (unsigned long long)d);
        atomicAdd(&ctr[CTR_TAIL_HITS], (unsigned long long)h);
        atomicAdd(&ctr[CTR_TAIL_EVALS], (unsigned long long)a);
        atomicAdd(&ctr[CTR_TAIL_ITERS], (unsigned long long)b);
        atomicAdd(&ctr[CTR_TAIL_DE64_EVALS], (unsigned long long)c);
        atomicAdd(&ctr[CTR_TAIL_DE64_ITERS], (unsigned long long)d);
    }
}

template <uint32_t FAM>
__global__ __launch_bounds__(256) void k_de_probe(const FractalP F, const float4* __restrict__ pts, uint32_t n,
                                                  float4* __restrict__ out, float4* __restrict__ traps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 q = pts[i];
    if constexpr (FAM == FAM_GRAPH || FAM == FAM_JIT) {
        // every trap slot and the material index (PreparedGraph::evaluate, eval.rs:98-114)
        GRaw<float> g;
        if constexpr (FAM == FAM_JIT) g = graph_jit<float>(F.gprog, mk(q.x, q.y, q.z));
        else g = graph_run<float>(F.gprog, mk(q.x, q.y, q.z));
        auto norm = [&](float t, int k) { return isfinite(t) ? fminf(t / F.gtrap_scale[k], 1.0f) : 1.0f; };
        const float t0 = norm(g.r.t0, 0);
        const uint32_t it = g.r.iter < F.iteration_product ? g.r.iter : F.iteration_product;
        out[i] = make_float4(g.r.d, t0, float(it), float(g.r.mat));
        traps[i] = make_float4(t0, norm(g.r.t1, 1), norm(g.r.t2, 2), norm(g.r.t3, 3));
    } else {
        De<float> s = de_eval<float, FAM>(F, mk(q.x, q.y, q.z));
        out[i] = make_float4(s.d, s.trap0, float(s.iter), 0.0f);
        traps[i] = make_float4(s.trap0, 1.0f, 1.0f, 1.0f);
    }
}

}  // namespace syf
