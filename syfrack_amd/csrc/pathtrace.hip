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
#include "pathtrace_kernels.hpp"

namespace syf {

struct WavefrontArena {
    uint32_t capacity = 0;
    void* planes = nullptr;     // WF_PLANES x capacity x 4 B
    uint32_t* queue = nullptr;  // 2 x capacity
    WfCtl* ctl = nullptr;
    hipGraphExec_t graph = nullptr;
    uint32_t graph_key[8] = {0};
};

// ---------------------------------------------------------------- launchers

// The one-launch form over a tile: four 8x8 pixel tiles per workgroup (one per wave).
hipError_t launch_render_tile(const JobP& P, const JobP* dP, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                              uint32_t s0, uint32_t ns, float4* accum, float* depth, float4* aux, double* depth64,
                              float4* vol, unsigned long long* ctr, hipStream_t s) {
    const uint32_t tiles_x = (w + 7u) / 8u, tiles = tiles_x * ((h + 7u) / 8u);
    const dim3 grid((tiles + 3u) / 4u), block(256);
    hipError_t e = hipSuccess;
    if (P.sh.fog_density > 0.0f || P.sh.detail) {
        SYF_DISPATCH(e, P.f, k_render_tile_full, grid, block, 0, s, dP, x0, y0, w, h, s0, ns, tiles_x, tiles, accum,
                     depth, aux, depth64, vol, ctr);
    } else {
        SYF_DISPATCH(e, P.f, k_render_tile, grid, block, 0, s, dP, x0, y0, w, h, s0, ns, tiles_x, tiles, accum, depth,
                     aux, depth64, vol, ctr);
    }
    return e;
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
    const JobP* dP = j->dP;
    hipLaunchKernelGGL(k_wf_generate, gen_grid, block, 0, s, dP, W, S, a->queue, a->ctl, j->ctr);
    hipError_t e = hipGetLastError();
    for (uint32_t r = 0; r <= P.max_bounces && e == hipSuccess; ++r) {
        const uint32_t in = r & 1u;
        const dim3 grid(std::min<uint32_t>((n + 255u) / 256u, fill));
        SYF_DISPATCH(e, P.f, k_wf_intersect, grid, block, 0, s, dP, W, S, a->queue, in, a->ctl, j->ctr);
        if (e != hipSuccess) break;
        SYF_DISPATCH(e, P.f, k_wf_shade, grid, block, 0, s, dP, W, S, a->queue, in, a->ctl, j->ctr);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(k_wf_next_round, dim3(1), dim3(64), 0, s, in, a->capacity, a->ctl);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return e;
    SYF_DISPATCH(e, P.f, k_wf_commit, gen_grid, block, 0, s, dP, W, S, (float4*)j->plane[SYF_PLANE_ACCUM],
                 (float*)j->plane[SYF_PLANE_DEPTH], (float4*)j->plane[SYF_PLANE_AUX],
                 (double*)j->plane[SYF_PLANE_DEPTH64], (float4*)j->vplanes, a->ctl, j->ctr);
    if (e != hipSuccess) return e;
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
        e = launch_render_tile(j->P, j->dP, x, y, w, h, sample_start, sample_count, (float4*)j->plane[SYF_PLANE_ACCUM],
                               (float*)j->plane[SYF_PLANE_DEPTH], (float4*)j->plane[SYF_PLANE_AUX],
                               (double*)j->plane[SYF_PLANE_DEPTH64], (float4*)j->vplanes, j->ctr, j->stream);
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
