// The control machine: interactive geometry of a node graph without per-edit compilation
// (ControlSession, crates/syfrack-render/src/control/mod.rs:145-515; shaders/control.wgsl).
//
// The reference runs a fixed per-lane instruction interpreter whose registers live in a 64 MiB
// storage buffer, advancing every pixel by 128 graph instructions per dispatch.  Here a pixel
// is a small resumable machine (36 bytes of state: stage, march position, relaxation factor,
// previous radius, step length, step count, normal accumulator) and a dispatch advances the
// image by a bounded number of whole field evaluations:
//   * every field evaluation runs the library's graph interpreter (graph_run): the program is
//     wave-uniform, each lane evaluates its own point, so lanes at different stages of their
//     pixels (marching, normal taps) share each evaluation;
//   * the unfinished pixels sit in a ring of pixel indices; a wave claims pixels from it as
//     lanes fall idle (a wave-level claim once 16 lanes are idle), runs at most `quantum`
//     evaluations, then saves the state of the pixels it still holds and returns them to the
//     ring.  A dispatch therefore costs at most quantum evaluations per wave, whatever the rays;
//   * per pixel the state machine is control.wgsl's: the over-relaxed march of march_segment
//     with pixel tolerance 1 and no step limit (control.wgsl:113-136: a ray continues until a
//     hit or a proven exit), the four tetrahedral taps of the normal (:54-61, :79-84) and the
//     headlight shade (:85-97).  A NaN or -inf distance stops the pixel and records the fence
//     that produced it (+inf is the library's empty distance, eval.rs:224-228).
// Each pixel's arithmetic does not depend on the schedule, so an image is the same for every
// quantum, grid and claim order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "api_internal.hpp"
#include "device.hpp"

namespace syf {

enum : uint32_t {
    CTL_FRESH = 0u,  // not started
    CTL_MARCH = 1u,
    CTL_TAP0 = 2u,   // normal taps 0..3: stages 2..5
    CTL_DONE = 6u,
    CTL_ERROR = 8u,
};
constexpr uint32_t CTL_NONE = 0xffffffffu;
constexpr uint32_t CTL_REFILL = 16u;  // idle lanes of a wave before it claims more pixels

struct CtlPlanes {  // structure of arrays, one element per pixel
    uint32_t* stage;
    float* t;       // march position; from the hit on, the hit distance
    float* omega;   // over-relaxation factor
    float* prevr;   // previous step's radius
    float* step;    // step length
    uint32_t* steps;
    float* nx;      // normal accumulator (sum of tap direction x tap distance)
    float* ny;
    float* nz;
    float4* accum;
    float* depth;
    float4* aux;
    uint32_t* ring;  // unfinished pixels, indexed by counter % pixels
};

struct CtlCounters {
    unsigned long long claim;  // next ring position to claim
    unsigned long long tail;   // next ring position to fill
    uint32_t errors;           // pixels stopped by an undefined distance
    uint32_t error_node;       // the fence of the last such pixel
};

// The pixel-centre ray of control_init (control.wgsl:271-274).
__device__ __forceinline__ V3<float> control_ray(const JobP& P, uint32_t pix) {
    const uint32_t px = pix % P.width, py = pix / P.width;
    const float ndcx = ((float(px) + 0.5f) / float(P.width)) * 2.0f - 1.0f;
    const float ndcy = ((float(py) + 0.5f) / float(P.height)) * 2.0f - 1.0f;
    const V3<float> f = ld3(P.cam.forward), r = ld3(P.cam.right), u = ld3(P.cam.up);
    return nrm(add(add(f, scl(r, ndcx * P.cam.tan_half)), scl(u, -ndcy * P.cam.tan_half / P.cam.aspect)));
}

// stencil(i) of control.wgsl:50-53 (normal_at's taps, common.wgsl:376-387)
__device__ __forceinline__ V3<float> control_tap(uint32_t i) {
    return mk((i == 0u || i == 3u) ? 1.0f : -1.0f, i >= 2u ? 1.0f : -1.0f, (i == 1u || i == 3u) ? 1.0f : -1.0f);
}

// safe_normalize (common.wgsl:242-247)
__device__ __forceinline__ V3<float> control_safe_normalize(V3<float> v) {
    const float l = len3(v);
    const float rcp = 1.0f / l;
    if (!(l > 0.0f) || !(rcp > 0.0f) || !(rcp <= 3.4028235e38f)) return mk(0.0f, 1.0f, 0.0f);
    return scl(v, rcp);
}

__global__ __launch_bounds__(256) void k_control_init(CtlPlanes S, uint32_t n, CtlCounters* ctr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0u) {
        ctr->claim = 0ull;
        ctr->tail = n;
        ctr->errors = 0u;
        ctr->error_node = 0u;
    }
    if (i >= n) return;
    S.stage[i] = CTL_FRESH;
    S.accum[i] = make_float4(0.025f, 0.03f, 0.04f, 1.0f);
    S.depth[i] = -1.0f;
    S.aux[i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    S.ring[i] = i;
}

// Sets the claim position of the next quantum (the host's view of the ring's head).
__global__ void k_control_arm(CtlCounters* ctr, unsigned long long head) { ctr->claim = head; }

// One quantum: waves claim pixels from ring positions [claim, end) and run at most `trips`
// field evaluations each; the pixels still unfinished go back to the ring after `end`.
__global__ __launch_bounds__(256) void k_control_quantum(const JobP P, CtlPlanes S, uint32_t cap,
                                                         const uint32_t* __restrict__ node_ids,
                                                         CtlCounters* ctr, unsigned long long end, uint32_t trips) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    const FractalP& F = P.f;
    const float cone = P.pixel_cone;  // pixel tolerance 1 (control/mod.rs:258)
    const V3<float> ro = ld3(P.cam.origin);
    const V3<float> light = nrm(mk(0.4f, 0.7f, 0.6f));
    uint32_t pix = CTL_NONE, stage = CTL_DONE, steps = 0u;
    float t = 0.0f, omega = 1.2f, prevr = 0.0f, step = 0.0f;
    V3<float> nacc = mk(0.0f, 0.0f, 0.0f);
    bool exhausted = false;
    for (uint32_t trip = 0;; ++trip) {
        const uint64_t idle = __ballot(pix == CTL_NONE);
        const uint32_t n_idle = uint32_t(__popcll(idle));
        if (!exhausted && trip < trips && (n_idle >= CTL_REFILL || n_idle == 64u)) {
            unsigned long long base = 0ull;
            if (lane == 0u) base = atomicAdd(&ctr->claim, (unsigned long long)n_idle);
            const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(base));
            const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(base >> 32));
            base = (unsigned long long)hi << 32 | lo;
            if (base + n_idle >= end) exhausted = true;
            if (pix == CTL_NONE) {
                const unsigned long long pos = base + uint32_t(__popcll(idle & below));
                if (pos < end) {
                    pix = S.ring[pos % cap];
                    stage = S.stage[pix];
                    if (stage == CTL_FRESH) {  // control_init's machine (control.wgsl:276-278)
                        stage = CTL_MARCH;
                        t = 0.0f;
                        omega = 1.2f;
                        prevr = 0.0f;
                        step = 0.0f;
                        steps = 1u;
                        nacc = mk(0.0f, 0.0f, 0.0f);
                    } else {
                        t = S.t[pix];
                        omega = S.omega[pix];
                        prevr = S.prevr[pix];
                        step = S.step[pix];
                        steps = S.steps[pix];
                        nacc = mk(S.nx[pix], S.ny[pix], S.nz[pix]);
                    }
                }
            }
        }
        const bool busy = pix != CTL_NONE;
        if (__ballot(busy) == 0ull || trip >= trips) break;
        // the point this lane evaluates: along its ray, or a normal tap around its hit
        V3<float> p = ro;
        V3<float> rd = mk(0.0f, 0.0f, 1.0f);
        float h = 0.0f;
        if (busy) {
            rd = control_ray(P, pix);
            p = add(ro, scl(rd, t));
            if (stage >= CTL_TAP0) {
                h = fmaxf(t * cone * 0.5f, 1e-4f);  // complete_hit (control.wgsl:54-61)
                p = add(p, scl(control_tap(stage - CTL_TAP0), h));
            }
        }
        const GRaw<float> g = graph_run<float>(F.gprog, p);
        const float d = g.r.d;
        if (!busy) continue;
        if (d != d || d == -INFINITY) {  // an undefined distance (control.wgsl:75-78)
            uint32_t node = 0u;
            (void)graph_run<float, true>(F.gprog, p, node_ids, &node);
            atomicAdd(&ctr->errors, 1u);
            atomicExch(&ctr->error_node, node);
            S.stage[pix] = CTL_ERROR;
            pix = CTL_NONE;
            continue;
        }
        if (stage >= CTL_TAP0) {  // finish_sample's normal taps and shade (control.wgsl:79-97)
            const uint32_t i = stage - CTL_TAP0;
            nacc = add(nacc, scl(control_tap(i), d));
            if (i < 3u) {
                ++stage;
            } else {
                const V3<float> n = control_safe_normalize(nacc);
                const float lit = 0.20f + 0.80f * fmaxf(dot3(n, light), 0.0f);
                S.accum[pix] = make_float4(0.68f * lit, 0.70f * lit, 0.73f * lit, 1.0f);
                S.depth[pix] = t;
                S.aux[pix] = make_float4(n.x, n.y, n.z, float(steps));
                S.stage[pix] = CTL_DONE;
                pix = CTL_NONE;
            }
            continue;
        }
        // the march (control.wgsl:113-136: march_segment's rules without a step limit)
        const float radius = fabsf(d) * F.safety;
        const bool failed = omega > 1.0f && radius + prevr < step * 1.0001f;
        if (failed) {
            step = step - omega * step;
            omega = 1.0f;
        } else {
            step = radius * omega;
        }
        prevr = radius;
        const float err = radius / fmaxf(t, 1e-6f);
        bool miss = false;
        if (!failed && err < cone) {  // complete_hit: t stays, the taps follow
            stage = CTL_TAP0;
            nacc = mk(0.0f, 0.0f, 0.0f);
            continue;
        }
        if (!failed && t + step >= P.t_end) {
            if (radius >= P.t_end - t) {
                miss = true;
            } else {
                omega = 1.0f;
                step = radius;
            }
        }
        if (!miss) {
            t = t + step;
            miss = t >= P.t_end;
        }
        if (miss) {  // complete_miss (control.wgsl:62-65), no medium
            S.accum[pix] = make_float4(0.025f, 0.035f, 0.05f, 1.0f);
            S.stage[pix] = CTL_DONE;
            pix = CTL_NONE;
            continue;
        }
        steps = min(steps, 0xfffffffeu) + 1u;
    }
    // the pixels still held go back to the ring
    const bool keep = pix != CTL_NONE;
    if (keep) {
        S.stage[pix] = stage;
        S.t[pix] = t;
        S.omega[pix] = omega;
        S.prevr[pix] = prevr;
        S.step[pix] = step;
        S.steps[pix] = steps;
        S.nx[pix] = nacc.x;
        S.ny[pix] = nacc.y;
        S.nz[pix] = nacc.z;
    }
    const uint64_t km = __ballot(keep);
    if (km) {
        unsigned long long base = 0ull;
        if (lane == 0u) base = atomicAdd(&ctr->tail, (unsigned long long)__popcll(km));
        const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(base));
        const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(base >> 32));
        base = (unsigned long long)hi << 32 | lo;
        if (keep) S.ring[(base + uint32_t(__popcll(km & below))) % cap] = pix;
    }
}

// Renderer::control_distances: the interpreter at points, in syf_de_probe's layout.
__global__ __launch_bounds__(256) void k_control_probe(const FractalP F, const float4* __restrict__ pts, uint32_t n,
                                                       float4* __restrict__ out, float4* __restrict__ traps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const float4 q = pts[min(i, n - 1u)];
    const GRaw<float> g = graph_run<float>(F.gprog, mk(q.x, q.y, q.z));
    if (i >= n) return;
    auto norm = [&](float t, int k) { return isfinite(t) ? fminf(t / F.gtrap_scale[k], 1.0f) : 1.0f; };
    const float t0 = norm(g.r.t0, 0);
    const uint32_t it = g.r.iter < F.iteration_product ? g.r.iter : F.iteration_product;
    out[i] = make_float4(g.r.d, t0, float(it), float(g.r.mat));
    traps[i] = make_float4(t0, norm(g.r.t1, 1), norm(g.r.t2, 2), norm(g.r.t3, 3));
}

}  // namespace syf

using namespace syf;

struct syf_control {
    syf_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    JobP P{};
    uint32_t width = 0, height = 0, quantum = SYF_CONTROL_QUANTUM;
    size_t pixels = 0;
    void* block = nullptr;  // the machine planes, image planes and ring
    CtlPlanes S{};
    CtlCounters* ctr = nullptr;
    GOp* prog = nullptr;        // the program and the node of each instruction
    uint32_t* node_ids = nullptr;
    size_t prog_cap = 0;        // instructions the program buffers hold
    CtlCounters* status = nullptr;  // pinned host copy of the counters after a quantum
    hipEvent_t done = nullptr;
    bool pending = false;
    bool initialize = true;
    unsigned long long head = 0, tail = 0, end = 0;
    uint32_t dispatches = 0;
    std::string failure;
};

namespace {

constexpr size_t CONTROL_MAX_PIXELS = size_t(1) << 26;
constexpr uint32_t CONTROL_MAX_QUANTUM = 4096;

void control_free_planes(syf_control* c) {
    (void)hipFree(c->block);
    c->block = nullptr;
    c->S = CtlPlanes{};
    c->ctr = nullptr;
}

void control_free(syf_control* c) {
    if (!c) return;
    // teardown: every release is attempted, whatever an earlier one returned
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    control_free_planes(c);
    (void)hipFree(c->prog);
    (void)hipFree(c->node_ids);
    if (c->status) (void)hipHostFree(c->status);
    if (c->done) (void)hipEventDestroy(c->done);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

hipError_t control_alloc_planes(syf_control* c, uint32_t width, uint32_t height) {
    const size_t n = size_t(width) * height;
    // nine 4-byte state planes; then ACCUM and AUX (16 bytes each), DEPTH and the ring (4 bytes
    // each), then the counters on a 256-byte boundary
    const size_t a16 = (9 * n * 4 + 255) & ~size_t(255);
    char* base = nullptr;
    hipError_t e = hipMalloc(&base, a16 + n * 40 + 512);
    if (e != hipSuccess) return e;
    c->block = base;
    uint32_t* w = reinterpret_cast<uint32_t*>(base);
    c->S.stage = w;
    c->S.t = reinterpret_cast<float*>(w + n);
    c->S.omega = reinterpret_cast<float*>(w + 2 * n);
    c->S.prevr = reinterpret_cast<float*>(w + 3 * n);
    c->S.step = reinterpret_cast<float*>(w + 4 * n);
    c->S.steps = w + 5 * n;
    c->S.nx = reinterpret_cast<float*>(w + 6 * n);
    c->S.ny = reinterpret_cast<float*>(w + 7 * n);
    c->S.nz = reinterpret_cast<float*>(w + 8 * n);
    char* img = base + a16;
    c->S.accum = reinterpret_cast<float4*>(img);
    c->S.aux = reinterpret_cast<float4*>(img + n * 16);
    c->S.depth = reinterpret_cast<float*>(img + n * 32);
    c->S.ring = reinterpret_cast<uint32_t*>(img + n * 36);
    c->ctr = reinterpret_cast<CtlCounters*>((reinterpret_cast<uintptr_t>(img + n * 40) + 255) & ~uintptr_t(255));
    c->width = width;
    c->height = height;
    c->pixels = n;
    return hipSuccess;
}

// Lowers and compiles (host only) a scene for the session: the job constants with pixel
// tolerance 1 (control/mod.rs:258) and the interpreter program.
int control_lower(const syf_job_desc& desc, const syf_graph& graph, JobP& P, std::vector<GOp>& prog,
                  std::vector<uint32_t>& node_of) {
    if (size_t(desc.width) * desc.height > CONTROL_MAX_PIXELS)
        return fail(SYF_E_INVALID, "width x height (%ux%u) exceeds the control machine's %zu pixels", desc.width,
                    desc.height, CONTROL_MAX_PIXELS);
    syf_job_desc d = desc;
    d.pixel_tolerance = 1.0f;
    if (d.max_steps < 1) d.max_steps = 1;  // unused: the machine has no step limit
    if (int rc = graph_lower_interpreted(d, graph, P, prog, node_of); rc != SYF_OK) return rc;
    return SYF_OK;
}

// Uploads the program (after the work in flight on the stream has finished).
int control_upload(syf_control* c, const std::vector<GOp>& prog, const std::vector<uint32_t>& node_of) {
    hipError_t e = hipSuccess;
    if (prog.size() > c->prog_cap) {
        (void)hipFree(c->prog);
        (void)hipFree(c->node_ids);
        c->prog = nullptr;
        c->node_ids = nullptr;
        c->prog_cap = 0;
        if ((e = hipMalloc(&c->prog, prog.size() * sizeof(GOp))) != hipSuccess ||
            (e = hipMalloc(&c->node_ids, prog.size() * sizeof(uint32_t))) != hipSuccess)
            return hip_fail(e, "control program allocation");
        c->prog_cap = prog.size();
    }
    e = hipMemcpy(c->prog, prog.data(), prog.size() * sizeof(GOp), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->node_ids, node_of.data(), node_of.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "control program upload");
    c->P.f.gprog = c->prog;
    return SYF_OK;
}

void control_restart(syf_control* c) {
    c->pending = false;
    c->initialize = true;
    c->head = 0;
    c->tail = c->pixels;
    c->end = 0;
    c->dispatches = 0;
    c->failure.clear();
}

int control_failed(const syf_control* c) { return fail(SYF_E_INVALID, "%s", c->failure.c_str()); }

}  // namespace

extern "C" {

int syf_control_create(syf_ctx* ctx, const syf_job_desc* desc, const syf_graph* graph, uint32_t quantum,
                       syf_control** out) {
    clear_error();
    if (!ctx || !desc || !graph || !out) return fail(SYF_E_INVALID, "ctx, desc, graph and out must not be null");
    if (quantum > CONTROL_MAX_QUANTUM)
        return fail(SYF_E_INVALID, "quantum must be at most %u field evaluations, got %u", CONTROL_MAX_QUANTUM, quantum);
    JobP P;
    std::vector<GOp> prog;
    std::vector<uint32_t> node_of;
    if (int rc = control_lower(*desc, *graph, P, prog, node_of); rc != SYF_OK) return rc;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    syf_control* c = new syf_control();
    c->ctx = ctx;
    c->P = P;
    c->quantum = quantum ? quantum : SYF_CONTROL_QUANTUM;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->done, hipEventDisableTiming)) != hipSuccess ||
        (e = hipHostMalloc(reinterpret_cast<void**>(&c->status), sizeof(CtlCounters))) != hipSuccess ||
        (e = control_alloc_planes(c, desc->width, desc->height)) != hipSuccess) {
        control_free(c);
        return hip_fail(e, "syf_control_create allocation");
    }
    if (int rc = control_upload(c, prog, node_of); rc != SYF_OK) {
        control_free(c);
        return rc;
    }
    control_restart(c);
    *out = c;
    return SYF_OK;
}

int syf_control_destroy(syf_control* c) {
    clear_error();
    control_free(c);
    return SYF_OK;
}

int syf_control_set_scene(syf_control* c, const syf_job_desc* desc, const syf_graph* graph) {
    clear_error();
    if (!c || !desc || !graph) return fail(SYF_E_INVALID, "control, desc and graph must not be null");
    JobP P;
    std::vector<GOp> prog;
    std::vector<uint32_t> node_of;
    // a new topology is committed only after validation (control/mod.rs:352-353)
    if (int rc = control_lower(*desc, *graph, P, prog, node_of); rc != SYF_OK) return rc;
    if (hipError_t de = hipSetDevice(c->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    // the quantum in flight reads the program and planes: let it finish, and drop its status
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "syf_control_set_scene");
    if (desc->width != c->width || desc->height != c->height) {
        control_free_planes(c);
        if ((e = control_alloc_planes(c, desc->width, desc->height)) != hipSuccess)
            return hip_fail(e, "syf_control_set_scene allocation");
    }
    const GOp* keep = c->prog;
    c->P = P;
    c->P.f.gprog = keep;
    if (int rc = control_upload(c, prog, node_of); rc != SYF_OK) return rc;
    control_restart(c);
    return SYF_OK;
}

int syf_control_tick(syf_control* c) {
    clear_error();
    if (!c) return fail(SYF_E_INVALID, "control must not be null");
    if (!c->failure.empty()) return control_failed(c);
    if (hipError_t de = hipSetDevice(c->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    if (c->pending) {
        const hipError_t q = hipEventQuery(c->done);
        if (q == hipErrorNotReady) return SYF_OK;
        if (q != hipSuccess) return hip_fail(q, "syf_control_tick");
        c->pending = false;
        const CtlCounters st = *c->status;
        if (st.errors != 0u) {
            char buf[160];
            std::snprintf(buf, sizeof buf,
                          "control graph produced a non-finite distance or invalid execution near node %u",
                          st.error_node);
            c->failure = buf;
            return control_failed(c);
        }
        c->head = std::min<unsigned long long>(st.claim, c->end);
        c->tail = st.tail;
        ++c->dispatches;
    }
    if (c->tail == c->head) return SYF_OK;  // complete
    hipError_t e = hipSuccess;
    const dim3 block(256);
    if (c->initialize) {
        hipLaunchKernelGGL(k_control_init, dim3(uint32_t((c->pixels + 255) / 256)), block, 0, c->stream, c->S,
                           uint32_t(c->pixels), c->ctr);
        e = hipGetLastError();
        c->initialize = false;
        c->head = 0;
        c->tail = c->pixels;
    } else {
        hipLaunchKernelGGL(k_control_arm, dim3(1), dim3(1), 0, c->stream, c->ctr, c->head);
        e = hipGetLastError();
    }
    // waves enough for the pending pixels, at most eight per CU
    const unsigned long long pending = c->tail - c->head;
    const uint32_t cus = uint32_t(c->ctx->cu_count > 0 ? c->ctx->cu_count : 256);
    const uint32_t blocks = uint32_t(std::min<unsigned long long>((pending + 255) / 256, cus * 2u));
    c->end = c->tail;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_control_quantum, dim3(blocks), block, 0, c->stream, c->P, c->S, uint32_t(c->pixels),
                           c->node_ids, c->ctr, c->end, c->quantum);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->status, c->ctr, sizeof(CtlCounters), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->done, c->stream);
    if (e != hipSuccess) return hip_fail(e, "syf_control_tick");
    c->pending = true;
    return SYF_OK;
}

int syf_control_wait(syf_control* c, double timeout_s) {
    clear_error();
    if (!c) return fail(SYF_E_INVALID, "control must not be null");
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    for (;;) {
        if (int rc = syf_control_tick(c); rc != SYF_OK) return rc;
        if (!c->pending && c->tail == c->head) return SYF_OK;
        if (std::chrono::steady_clock::now() > deadline)
            return fail(SYF_E_TIMEOUT, "control evaluation timed out after %g s", timeout_s);
        if (c->pending) {
            const hipError_t e = hipEventSynchronize(c->done);
            if (e != hipSuccess) return hip_fail(e, "syf_control_wait");
        }
    }
}

int syf_control_progress(syf_control* c, uint32_t* completed_pixels, uint32_t* dispatches, uint32_t* complete,
                         uint32_t* width, uint32_t* height) {
    clear_error();
    if (!c) return fail(SYF_E_INVALID, "control must not be null");
    const unsigned long long remaining = c->tail - c->head;
    if (completed_pixels) *completed_pixels = uint32_t(c->pixels - remaining);
    if (dispatches) *dispatches = c->dispatches;
    if (complete) *complete = (!c->pending && remaining == 0ull) ? 1u : 0u;
    if (width) *width = c->width;
    if (height) *height = c->height;
    return SYF_OK;
}

int syf_control_readback(syf_control* c, uint32_t plane, void* host, size_t bytes) {
    clear_error();
    if (!c || !host) return fail(SYF_E_INVALID, "control and host must not be null");
    const void* src = nullptr;
    size_t need = 0;
    switch (plane) {
    case SYF_PLANE_ACCUM: src = c->S.accum; need = c->pixels * 16; break;
    case SYF_PLANE_DEPTH: src = c->S.depth; need = c->pixels * 4; break;
    case SYF_PLANE_AUX: src = c->S.aux; need = c->pixels * 16; break;
    default: return fail(SYF_E_INVALID, "plane: the control image has ACCUM, DEPTH and AUX, not plane %u", plane);
    }
    if (bytes != need) return fail(SYF_E_INVALID, "bytes must be %zu, got %zu", need, bytes);
    if (hipError_t de = hipSetDevice(c->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    hipError_t e = hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? SYF_OK : hip_fail(e, "syf_control_readback");
}

int syf_control_pick(syf_control* c, float u, float v, uint32_t* hit, float* depth, float* aux4) {
    clear_error();
    if (!c || !hit || !depth || !aux4) return fail(SYF_E_INVALID, "control and outputs must not be null");
    const float cu = std::fmin(std::fmax(u, 0.0f), 0.9999f), cv = std::fmin(std::fmax(v, 0.0f), 0.9999f);
    const size_t x = size_t(cu * float(c->width)), y = size_t(cv * float(c->height));
    const size_t i = y * c->width + x;
    if (hipError_t de = hipSetDevice(c->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    float d = -1.0f;
    float a[4] = {0, 0, 0, 0};
    hipError_t e = hipMemcpyAsync(&d, c->S.depth + i, 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(a, c->S.aux + i, 16, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "syf_control_pick");
    *hit = d > 0.0f ? 1u : 0u;
    *depth = d;
    std::memcpy(aux4, a, sizeof a);
    return SYF_OK;
}

int syf_control_probe(syf_ctx* ctx, const syf_graph* graph, const float* xyz4, size_t n, float* out4,
                      float* traps4) {
    clear_error();
    if (!ctx || !graph || (n && (!xyz4 || !out4 || !traps4)))
        return fail(SYF_E_INVALID, "ctx, graph and buffers must not be null");
    if (n > (size_t(1) << 30)) return fail(SYF_E_INVALID, "n must be below 2^30, got %zu", n);
    // a probe needs no camera: a minimal valid description
    syf_job_desc d;
    std::memset(&d, 0, sizeof d);
    d.width = d.height = 1;
    d.max_steps = 1;
    d.pixel_tolerance = 1.0f;
    d.camera.focal_mm = 35.0f;
    d.camera.target[2] = -1.0f;
    JobP P;
    std::vector<GOp> prog;
    std::vector<uint32_t> node_of;
    if (int rc = graph_lower_interpreted(d, *graph, P, prog, node_of); rc != SYF_OK) return rc;
    if (n == 0) return SYF_OK;
    if (hipError_t de = hipSetDevice(ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    GOp* dprog = nullptr;
    float4 *dp = nullptr, *dout = nullptr, *dtr = nullptr;
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&dprog, prog.size() * sizeof(GOp));
    if (e == hipSuccess) e = hipMalloc(&dp, n * 16);
    if (e == hipSuccess) e = hipMalloc(&dout, n * 16);
    if (e == hipSuccess) e = hipMalloc(&dtr, n * 16);
    if (e == hipSuccess) e = hipMemcpyAsync(dprog, prog.data(), prog.size() * sizeof(GOp), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dp, xyz4, n * 16, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        P.f.gprog = dprog;
        hipLaunchKernelGGL(k_control_probe, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, P.f, dp, uint32_t(n),
                           dout, dtr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out4, dout, n * 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(traps4, dtr, n * 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dprog);
    (void)hipFree(dp);
    (void)hipFree(dout);
    (void)hipFree(dtr);
    if (s) (void)hipStreamDestroy(s);
    return e == hipSuccess ? SYF_OK : hip_fail(e, "syf_control_probe");
}

}  // extern "C"
