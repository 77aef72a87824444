// Render server: continuous batching of uneven requests on one device.
//
// A scheduler thread owns the device side.  Each cycle it
//   1. admits queued requests in arrival order while their footprint (accumulation, depth,
//      aux and FP64 depth planes, 44 bytes per pixel) fits the HBM budget and a slot is free;
//   2. plans one launch: work items (8x8 tile, sample range) of the active requests, shortest
//      remaining work first, until the launch holds enough items to fill the device and its
//      estimated device time reaches the target (a running estimate of ns per
//      pixel-sample-segment, refreshed from every launch that filled the device).  The target
//      is short while requests arrive and long once they have stopped (quiet mode);
//   3. enqueues it (one persistent kernel per kernel family present, the families on streams
//      of their own so that they run concurrently) and keeps up to three launches in flight, so
//      that the next launch's workgroups take the CUs the running launch's tail frees.  Nothing else is enqueued
//      between launches: the kernels read their work items from pinned host memory, and each
//      launch's last wave hands its counters to the host and clears them.  (A copy or clear
//      enqueued between two launches runs as a kernel of its own, which waits for the running
//      launch's workgroups to leave the CUs and so kept consecutive launches from overlapping.)
//   4. retires the oldest launch when its end event completes: adds its samples to the
//      requests' progress and stamps the requests it finished.
// Accumulation of a pixel is owned by one work item per sample range, and a tile's next range
// is planned only after the launch holding its previous range retired, so samples land in
// ascending order with no float atomics; the other tiles of the request go on meanwhile.
// The plain mode runs one launch per request (all samples), in arrival order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "api_internal.hpp"
#include "kernels.hpp"

namespace syf {

struct Request {
    uint64_t id = 0;
    JobP P{};
    uint32_t spp = 0;
    uint32_t state = SYF_REQ_QUEUED;
    uint32_t slot = ~0u;
    uint32_t tiles_x = 0, tiles = 0;
    uint32_t tws = 3;        // its tiles are 2^tws x 2^(6 - tws) pixels (its family's choice)
    uint32_t quantum_long = 0;  // samples per pass in quiet mode (its family's choice)
    uint64_t total = 0;      // tile-samples of the whole request
    uint64_t planned = 0;    // tile-samples launched
    uint64_t completed = 0;  // tile-samples whose launches finished
    // Passes: the request's samples go in quanta of q_len samples starting at q_start; a pass
    // plans the tiles in ascending order (q_cursor = next tile).  Tile t of a pass may be
    // planned only once the launch holding tile t of the previous pass has retired: ret_prev
    // and ret_cur are the retired tile prefixes of passes `pass - 1` and `pass`.
    uint32_t q_start = 0, q_len = 0, q_cursor = 0;
    uint32_t pass = 0, ret_prev = 0, ret_cur = 0;
    bool pass_open = false;
    size_t pixels = 0;
    uint64_t footprint = 0;
    void* accum = nullptr;
    void* depth = nullptr;
    void* aux = nullptr;
    void* depth64 = nullptr;
    BatchPlanes planes{};  // source of the request's entry in the device plane table
    // Batched mode: the slot-table entries are written by the first launch that holds items of
    // the request; until that launch retires, no other launch takes its items.
    bool uploaded = false, upload_in_flight = false;
    double submit_ms = 0, start_ms = -1, done_ms = -1;
    bool cancelled = false;
    bool released = false;  // released by the caller; deleted once no launch refers to it
    uint32_t refs = 0;      // launches in flight (or a plain render) that refer to the request
};

// A launch's share of one request: tile-samples, and the pass and end of the tile range it
// holds (tiles of one pass, ascending).
struct Share {
    Request* r;
    uint64_t samples;
    uint32_t pass, tile_end;
};

struct Launch {
    hipEvent_t start = nullptr, end = nullptr;
    std::vector<Share> progress;
    std::vector<Request*> uploads;  // requests whose slot-table entries this launch writes
    uint32_t buf = 0;
    uint32_t families = 0;  // bit f: a kernel of family f ran
    double units = 0;
    bool full = false;  // enough items to fill the device: its time measures throughput
};

}  // namespace syf

using namespace syf;

// Batched launches in flight at once; each has its own work-item and counter buffers, streams
// and events.
constexpr int NBUF = 3;
// Quiet mode (see plan_launch): after LATENCY_HOLD_MS without a submission, launches of
// QUIET_FACTOR x the latency-mode target with QUIET_QUANTUM samples per pass.  Measured on
// the 256-request uneven batch (Mrays/s; the plain form 142): 4x / 4 samples 336, 8x / 4
// 359, 16x / 8 404, 32x / 16 427, everything in one launch 526; the Poisson stream's rays/s
// and preview latency did not change (its requests arrive within 0.3 s, in latency mode).
constexpr double QUIET_FACTOR = 32.0;
constexpr uint32_t QUIET_QUANTUM = 16;
// Requests count as still arriving for max(LATENCY_HOLD_MS, 10 x the mean gap between recent
// submissions), at most LATENCY_HOLD_MAX_MS: a batch submitted at once goes quiet after 200 ms,
// a Poisson stream of 20/s (mean gap 50 ms) only after 500 ms without a submission, which its
// gaps exceed with probability e^-10 (with a fixed 200 ms, 2 % of its gaps switched the server
// to quiet launches of about a second, and previews arriving then waited for them: p99 768 ms).
constexpr double LATENCY_HOLD_MS = 200.0;
constexpr double LATENCY_HOLD_MAX_MS = 2000.0;

struct syf_server {
    syf_ctx* ctx = nullptr;
    syf_server_config cfg{};
    hipStream_t stream = nullptr;  // admissions, plain-mode launches, calibration
    // Batched launches: launch b (b = launch index % NBUF) runs all its kernels on stream ls[b].
    // Consecutive launches run concurrently (the planner keeps their pixels disjoint), and no
    // launch waits on another stream: a wait on a stream that shares a hardware queue with a
    // running launch would hold it back until that launch ends.
    hipStream_t ls[NBUF] = {};
    std::mutex mu;
    std::condition_variable cv;
    std::thread worker;
    bool stop = false;
    std::map<uint64_t, Request*> reqs;
    std::deque<Request*> queue;
    std::vector<Request*> active;
    std::deque<Launch> inflight;
    uint64_t next_id = 1;
    uint32_t max_slots = 0;
    std::vector<uint32_t> free_slots;
    JobP* d_jobs = nullptr;
    JobP* d_plain_job = nullptr;  // plain mode: the job constants of the request being rendered
    BatchPlanes* d_planes = nullptr;
    BatchItem* h_items[NBUF] = {};  // pinned, coherent: read by the kernels in place
    size_t items_cap[NBUF] = {};
    // per launch buffer and kernel family: claim counters, work counters and exit count (zero
    // between launches), and the host copy of the work counters
    uint32_t* d_claim[NBUF] = {};
    unsigned long long* d_ctr[NBUF] = {};
    uint32_t* d_exit[NBUF] = {};
    unsigned long long* h_out[NBUF] = {};
    unsigned long long* d_rctr = nullptr;  // counters of the refine kernels (not reported)
    unsigned long long* d_pctr = nullptr;  // counters of plain-mode launches
    hipEvent_t ev_pool[2 * NBUF] = {};
    uint32_t quantum = 0, quantum_long = 0, target_us = 0, quiet_us = 0, grid = 0;
    // per kernel family: persistent workgroups, field evaluations per loop trip of k_paths,
    // tile width (log2) and quiet-mode samples per pass (the autotuner's choices, or the
    // configuration's)
    uint32_t fam_grid[FAM_COUNT] = {0}, fam_steps[FAM_COUNT] = {0}, fam_tws[FAM_COUNT] = {0},
             fam_quantum[FAM_COUNT] = {0};
    double autotune_ms = 0.0;
    double last_submit_ms = -1e30;  // host clock of the latest submission
    double gap_ms = 0.0;            // running mean of the gaps between submissions (below 5 s)
    uint32_t fill_items = 0;  // work items (one wave each) a launch needs to fill the device
    double ns_per_unit = 0.0;
    uint64_t hbm_in_use = 0, hbm_peak = 0, budget = 0;
    uint64_t submitted = 0, completed = 0, cancelled = 0, launches = 0, rays = 0, primary = 0;
    double device_ms = 0.0;
    std::chrono::steady_clock::time_point t0;
    int error = SYF_OK;
    std::string error_text;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

constexpr uint32_t TILE_PIXELS = 64;  // pixels of a work item: one per lane of a wave
constexpr size_t BYTES_PER_PIXEL = 16 + 4 + 16 + 8;  // accum, depth, aux, depth64

// Device time estimate of a work item: pixel-samples x segments per path.
inline double item_units(const Request* r, uint32_t ns) {
    return double(TILE_PIXELS) * double(ns) * double(r->P.max_bounces + 1u);
}
// Work item t of a pass of request r
inline BatchItem tile_item(const Request* r, uint32_t t) {
    const uint32_t tw = 1u << r->tws, th = TILE_PIXELS >> r->tws;
    return BatchItem{r->slot, (t % r->tiles_x) * tw, (t / r->tiles_x) * th, r->q_start, r->q_len, r->tws};
}

void set_server_error(syf_server* s, int code, const char* what, hipError_t e) {
    s->error = code;
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    s->error_text = buf;
}

// Allocates a request's planes and its slot (caller holds the lock).
bool admit(syf_server* s, Request* r) {
    if (s->free_slots.empty()) return false;
    if (s->hbm_in_use + r->footprint > s->budget) return false;
    hipError_t e = hipMalloc(&r->accum, r->pixels * 16);
    if (e == hipSuccess) e = hipMalloc(&r->depth, r->pixels * 4);
    if (e == hipSuccess) e = hipMalloc(&r->aux, r->pixels * 16);
    if (e == hipSuccess) e = hipMalloc(&r->depth64, r->pixels * 8);
    // batched launches write every pixel of every plane in the request's first pass; the
    // megakernel of the plain mode adds to the accumulation
    if (s->cfg.mode == SYF_SERVE_PLAIN) {
        if (e == hipSuccess) e = hipMemsetAsync(r->accum, 0, r->pixels * 16, s->stream);
        if (e == hipSuccess) e = hipMemsetAsync(r->depth, 0, r->pixels * 4, s->stream);
        if (e == hipSuccess) e = hipMemsetAsync(r->aux, 0, r->pixels * 16, s->stream);
        if (e == hipSuccess) e = hipMemsetAsync(r->depth64, 0, r->pixels * 8, s->stream);
    }
    if (e != hipSuccess) {
        hipFree(r->accum);
        hipFree(r->depth);
        hipFree(r->aux);
        hipFree(r->depth64);
        r->accum = r->depth = r->aux = r->depth64 = nullptr;
        return false;
    }
    r->slot = s->free_slots.back();
    s->free_slots.pop_back();
    r->planes = BatchPlanes{(float4*)r->accum, (float*)r->depth, (float4*)r->aux, (double*)r->depth64};
    // plain mode: the megakernel takes the job by value; batched mode: the first launch that
    // holds the request's items writes its slot-table entries
    r->uploaded = false;
    r->upload_in_flight = false;
    if (e != hipSuccess) {
        set_server_error(s, SYF_E_DEVICE, "request upload", e);
        return false;
    }
    s->hbm_in_use += r->footprint;
    s->hbm_peak = std::max(s->hbm_peak, s->hbm_in_use);
    r->state = SYF_REQ_ACTIVE;
    s->active.push_back(r);
    return true;
}

void free_request_planes(syf_server* s, Request* r) {
    if (r->accum || r->depth || r->aux || r->depth64) {
        hipFree(r->accum);
        hipFree(r->depth);
        hipFree(r->aux);
        hipFree(r->depth64);
        r->accum = r->depth = r->aux = r->depth64 = nullptr;
        s->hbm_in_use -= r->footprint;
    }
    if (r->slot != ~0u) {
        s->free_slots.push_back(r->slot);
        r->slot = ~0u;
    }
}

bool ensure_items(syf_server* s, uint32_t buf, size_t n) {
    if (n <= s->items_cap[buf]) return true;
    size_t cap = std::max<size_t>(n, s->items_cap[buf] * 2);
    // the launches of this buffer have retired: nothing reads the old items
    hipHostFree(s->h_items[buf]);
    s->h_items[buf] = nullptr;
    s->items_cap[buf] = 0;
    if (hipHostMalloc(&s->h_items[buf], cap * sizeof(BatchItem), hipHostMallocCoherent) != hipSuccess) return false;
    s->items_cap[buf] = cap;
    return true;
}

// Plans and enqueues one batched launch (caller holds the lock).  Returns false when there
// is nothing to launch.
bool plan_launch(syf_server* s) {
    const uint32_t buf = uint32_t(s->launches % NBUF);
    std::vector<Request*> order;
    for (Request* r : s->active)
        if (!r->cancelled && r->planned < r->total && !r->upload_in_flight) order.push_back(r);
    if (order.empty()) return false;
    std::stable_sort(order.begin(), order.end(), [](const Request* a, const Request* b) {
        return (a->total - a->planned) * (a->P.max_bounces + 1u) < (b->total - b->planned) * (b->P.max_bounces + 1u);
    });
    // Latency mode (launches of target_us, s->quantum samples per pass) while requests arrive
    // (a submission within the last LATENCY_HOLD_MS) or one waits for admission.  Otherwise
    // nothing new has come in for a while: launches of quiet_us with s->quantum_long samples
    // per pass, for fewer launch tails and less per-sample bookkeeping; a request arriving then
    // waits for those longer launches.  Measured on one 1080p Mandelbox request at 16 spp: 141
    // Mrays/s at (30 ms, 1 sample), 205 at (100 ms, 4 samples), 215 for the whole request in
    // one launch.
    const double hold_ms = std::min(std::max(LATENCY_HOLD_MS, 10.0 * s->gap_ms), LATENCY_HOLD_MAX_MS);
    const bool latency = !s->queue.empty() || s->now_ms() - s->last_submit_ms < hold_ms;
    const double ns = std::max(s->ns_per_unit, 1e-6);
    const double budget_units = latency ? double(s->target_us) * 1000.0 / ns
                                : s->quiet_us == SYF_QUIET_UNBOUNDED ? 1e300 : double(s->quiet_us) * 1000.0 / ns;
    std::vector<BatchItem> by_family[FAM_COUNT];
    size_t n_items = 0;
    Launch L;
    double units = 0.0;
    for (Request* r : order) {
        if (units >= budget_units && n_items >= s->fill_items) break;
        // A launch may hold the end of one pass of a request and the start of the next: the
        // frontier keeps every pixel in at most one of its items.
        uint64_t seg = 0;  // tile-samples of the current pass in this launch
        std::vector<BatchItem>& items = by_family[kernel_family(r->P.f)];
        auto close_segment = [&](uint32_t tile_end) {
            L.progress.push_back(Share{r, seg, r->pass, tile_end});
            r->refs++;
            seg = 0;
        };
        while (r->planned < r->total && (units < budget_units || n_items < s->fill_items)) {
            if (!r->pass_open) {  // open the next pass
                r->q_start = uint32_t(r->planned / r->tiles);
                r->q_len = std::min(latency ? s->quantum : r->quantum_long, r->spp - r->q_start);
                // a request whose remaining work fits in a launch goes in one piece: it then
                // needs one launch, not one per quantum
                if (double(r->total - r->planned) * double(TILE_PIXELS) * double(r->P.max_bounces + 1u) <=
                    budget_units)
                    r->q_len = r->spp - r->q_start;
                if (r->q_start == 0u) {
                    r->pass = 0;
                    r->ret_prev = r->tiles;  // no previous pass
                } else {
                    r->pass++;
                    r->ret_prev = r->ret_cur;
                }
                r->ret_cur = 0;
                r->q_cursor = 0;
                r->pass_open = true;
            }
            const uint32_t t = r->q_cursor;
            // the previous pass's item of this tile is still in flight: its samples must land first
            if (t >= r->ret_prev) break;
            items.push_back(tile_item(r, t));
            n_items++;
            units += item_units(r, r->q_len);
            r->planned += r->q_len;
            seg += r->q_len;
            if (++r->q_cursor == r->tiles) {
                r->pass_open = false;
                close_segment(r->tiles);
            }
        }
        if (seg) close_segment(r->q_cursor);
        if (r->refs && r->start_ms < 0) r->start_ms = s->now_ms();
        if (r->refs && !r->uploaded && !r->upload_in_flight) {
            r->upload_in_flight = true;
            L.uploads.push_back(r);
        }
    }
    if (n_items == 0) return false;
    if (!ensure_items(s, buf, n_items)) {
        s->error = SYF_E_DEVICE;
        s->error_text = "work-item buffers: out of memory";
        return false;
    }
    // one persistent kernel per kernel family present, back to back on the stream; the items
    // of a family stay in shortest-remaining-first order
    uint32_t first[FAM_COUNT], count[FAM_COUNT];
    bool first_quantum[FAM_COUNT];
    size_t at = 0;
    for (uint32_t f = 0; f < FAM_COUNT; ++f) {
        first[f] = uint32_t(at);
        count[f] = uint32_t(by_family[f].size());
        first_quantum[f] = std::any_of(by_family[f].begin(), by_family[f].end(),
                                       [](const BatchItem& it) { return it.s0 == 0u; });
        if (count[f]) std::memcpy(s->h_items[buf] + at, by_family[f].data(), count[f] * sizeof(BatchItem));
        at += count[f];
    }
    L.buf = buf;
    L.units = units;
    L.full = n_items >= s->fill_items;
    L.start = s->ev_pool[2 * buf];
    L.end = s->ev_pool[2 * buf + 1];
    hipStream_t ls = s->ls[buf];
    hipError_t e = hipEventRecord(L.start, ls);
    for (Request* r : L.uploads)
        if (e == hipSuccess) e = launch_admit(s->d_jobs, s->d_planes, r->slot, r->P, r->planes, ls);
    for (uint32_t f = 0; f < FAM_COUNT && e == hipSuccess; ++f)
        if (count[f]) {
            e = launch_paths(f, s->d_jobs, s->d_planes, s->h_items[buf] + first[f], count[f],
                             s->d_claim[buf] + f * CLAIM_WORDS, s->d_ctr[buf] + f * CTR_COUNT, s->d_exit[buf] + f * 16u,
                             s->h_out[buf] + f * CTR_COUNT, latency ? s->fam_grid[f] : s->grid, s->fam_steps[f],
                             ls);
            if (e == hipSuccess && first_quantum[f])
                e = launch_refine_first_hits(f, s->d_jobs, s->d_planes, s->h_items[buf] + first[f], count[f],
                                             s->d_rctr, ls);
            L.families |= 1u << f;
        }
    if (e == hipSuccess) e = hipEventRecord(L.end, ls);
    if (e != hipSuccess) {
        set_server_error(s, SYF_E_DEVICE, "batch launch", e);
        return false;
    }
    s->launches++;
    s->inflight.push_back(std::move(L));
    return true;
}

// Marks an active request done (caller holds the lock).
void finish(syf_server* s, Request* r, double t) {
    r->state = SYF_REQ_DONE;
    r->done_ms = t;
    s->completed++;
    auto it = std::find(s->active.begin(), s->active.end(), r);
    if (it != s->active.end()) s->active.erase(it);
}

// Retires the oldest launch after its end event (entered with the lock held; waits without it).
void retire(syf_server* s, std::unique_lock<std::mutex>& lk) {
    Launch L = std::move(s->inflight.front());
    s->inflight.pop_front();
    lk.unlock();
    hipError_t e = hipEventSynchronize(L.end);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, L.start, L.end);
    // the counters each family's kernel left in host memory
    unsigned long long c[CTR_COUNT] = {0};
    for (uint32_t f = 0; f < FAM_COUNT; ++f)
        if (L.families >> f & 1u)
            for (uint32_t k = 0; k < CTR_COUNT; ++k)
                c[k] += __atomic_load_n(&s->h_out[L.buf][f * CTR_COUNT + k], __ATOMIC_ACQUIRE);
    const double t = s->now_ms();
    lk.lock();
    if (e != hipSuccess) {
        set_server_error(s, SYF_E_DEVICE, "batch completion", e);
        return;
    }
    s->device_ms += ms;
    s->rays += c[CTR_RAYS];
    s->primary += c[CTR_PRIMARY];
    // a launch too small to fill the device is latency-bound: its time per unit says nothing
    // about throughput and would shrink the next launches
    if (L.full && L.units > 0 && ms > 0) {
        const double meas = double(ms) * 1e6 / L.units;
        s->ns_per_unit = s->ns_per_unit > 0 ? 0.7 * s->ns_per_unit + 0.3 * meas : meas;
    }
    for (Request* r : L.uploads) {
        r->uploaded = true;
        r->upload_in_flight = false;
    }
    for (const Share& sh : L.progress) {
        Request* r = sh.r;
        r->completed += sh.samples;
        // launches retire in order, so a pass's retired tiles form a growing prefix
        if (sh.pass == r->pass) r->ret_cur = sh.tile_end;
        else if (sh.pass + 1u == r->pass) r->ret_prev = sh.tile_end;
        r->refs--;
        if (r->completed >= r->total && r->state == SYF_REQ_ACTIVE) finish(s, r, t);
        if (r->released && r->refs == 0) delete r;
    }
    s->cv.notify_all();
}

bool run_plain(syf_server* s, std::unique_lock<std::mutex>& lk) {
    Request* r = nullptr;
    while (!s->queue.empty()) {
        r = s->queue.front();
        if (r->cancelled) {
            s->queue.pop_front();
            r = nullptr;
            continue;
        }
        if (!admit(s, r)) {
            if (s->hbm_in_use > 0 || !s->active.empty()) return false;  // wait for releases
            r->state = SYF_REQ_CANCELLED;
            r->cancelled = true;
            s->cancelled++;
            s->queue.pop_front();
            s->cv.notify_all();
            r = nullptr;
            continue;
        }
        s->queue.pop_front();
        break;
    }
    if (!r) return false;
    r->start_ms = s->now_ms();
    r->planned = r->total;
    r->refs++;
    lk.unlock();
    hipError_t e = hipMemsetAsync(s->d_pctr, 0, CTR_COUNT * sizeof(unsigned long long), s->stream);
    // the job constants, outside the timed span (the host copy stays alive: the launch is
    // waited for below)
    if (e == hipSuccess) e = hipMemcpyAsync(s->d_plain_job, &r->P, sizeof(JobP), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipEventRecord(s->ev_pool[0], s->stream);
    if (e == hipSuccess)
        e = launch_render_tile(r->P, s->d_plain_job, 0, 0, r->P.width, r->P.height, 0, r->spp, (float4*)r->accum,
                               (float*)r->depth, (float4*)r->aux, (double*)r->depth64, nullptr, s->d_pctr, s->stream);
    if (e == hipSuccess) e = hipEventRecord(s->ev_pool[1], s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(s->ev_pool[1]);
    float ms = 0.0f;
    unsigned long long c[CTR_COUNT] = {0};
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, s->ev_pool[0], s->ev_pool[1]);
    if (e == hipSuccess) e = hipMemcpy(c, s->d_pctr, sizeof c, hipMemcpyDeviceToHost);
    const double t = s->now_ms();
    lk.lock();
    r->refs--;
    if (e != hipSuccess) {
        set_server_error(s, SYF_E_DEVICE, "plain launch", e);
        return false;
    }
    s->launches++;
    s->device_ms += ms;
    s->rays += c[CTR_RAYS];
    s->primary += c[CTR_PRIMARY];
    r->completed = r->total;
    if (r->state == SYF_REQ_ACTIVE) finish(s, r, t);  // a cancellation during the render wins
    if (r->released && r->refs == 0) delete r;
    s->cv.notify_all();
    return true;
}

void worker_main(syf_server* s) {
    (void)hipSetDevice(s->ctx->device);  // a failure shows in the calls that follow
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        if (s->error != SYF_OK) {  // stays stopped: waiters see the error, destroy cleans up
            s->cv.notify_all();
            s->cv.wait(lk, [&] { return s->stop; });
            break;
        }
        if (s->cfg.mode == SYF_SERVE_PLAIN) {
            s->cv.wait(lk, [&] { return s->stop || !s->queue.empty(); });
            if (s->stop) break;
            if (!run_plain(s, lk) && !s->stop) s->cv.wait_for(lk, std::chrono::milliseconds(5));  // memory held
            continue;
        }
        // admissions in arrival order
        while (!s->queue.empty()) {
            Request* r = s->queue.front();
            if (r->cancelled) {
                s->queue.pop_front();
                continue;
            }
            if (s->active.size() >= s->cfg.max_active) break;
            if (!admit(s, r)) {
                if (s->error != SYF_OK || s->hbm_in_use > 0 || !s->inflight.empty()) break;  // wait for releases
                // nothing else holds device memory and the allocation still failed: refuse it
                r->state = SYF_REQ_CANCELLED;
                r->cancelled = true;
                s->cancelled++;
                s->queue.pop_front();
                s->cv.notify_all();
                continue;
            }
            s->queue.pop_front();
        }
        bool planned = false;
        while (!s->stop && s->inflight.size() < NBUF && plan_launch(s)) planned = true;
        if (!s->inflight.empty()) {
            retire(s, lk);
            continue;
        }
        if (s->stop) break;
        // idle: woken by a submission, a release or a cancellation
        if (!planned) s->cv.wait_for(lk, std::chrono::milliseconds(5));
    }
}

// The calibration scene of a kernel family: 512x512, 4 samples, 2 bounces under a white sky, the
// family's bundled graph (Mandelbulb powers 5, 8 and 6.5 for the three Mandelbulb kernels) seen
// as the serving mix's random scenes see it.
syf_job_desc calibration_desc(uint32_t kfam) {
    syf_job_desc d;
    std::memset(&d, 0, sizeof d);
    syf_fractal& f = d.fractal;
    f.plane_normal[1] = 1.0f;
    float dist = 3.2f;
    if (kfam == SYF_FAMILY_MANDELBOX) {
        f.family = SYF_FAMILY_MANDELBOX;
        f.iterations = 12;
        f.bailout = 1000.0f;
        f.scale = 2.0f;
        f.min_radius = 0.5f;
        f.fixed_radius = 1.0f;
        f.fold_limit = 1.0f;
        f.trap_scale = 2.0f;
        f.safety = 0.6f;
        f.bounding_radius_hint = 10.392304f;
        dist = 11.0f;
    } else if (kfam == SYF_FAMILY_MENGER) {
        f.family = SYF_FAMILY_MENGER;
        f.iterations = 5;
        f.bailout = 1e6f;
        f.scale = 3.0f;
        for (int k = 0; k < 3; ++k) f.offset[k] = f.half[k] = 1.0f;
        f.trap_scale = 4.0f;
        f.safety = 0.9f;
        f.bounding_radius_hint = 4.464102f;
        dist = 4.2f;
    } else if (kfam == SYF_FAMILY_SPHERE) {
        f.family = SYF_FAMILY_SPHERE;
        f.radius = 1.0f;
        f.plane_offset = -1.0f;
        f.trap_scale = 1.0f;
        f.safety = 1.0f;
        f.bounding_radius_hint = 100.0f;
        dist = 3.0f;
    } else {
        f.family = SYF_FAMILY_MANDELBULB;
        f.iterations = 10;
        f.bailout = 4.0f;
        f.power = kfam == FAM_BULB8 ? 8.0f : kfam == FAM_BULB_TRIG ? 6.5f : 5.0f;
        f.trap_scale = 4.0f;
        f.safety = 0.7f;
        f.bounding_radius_hint = 2.5f;
    }
    d.camera.mode = SYF_CAMERA_ORBIT;
    d.camera.distance = dist;
    d.camera.azimuth = 30.0f;
    d.camera.elevation = 20.0f;
    d.camera.focal_mm = 35.0f;
    d.material.albedo[0] = d.material.albedo[1] = d.material.albedo[2] = 0.8f;
    d.material.roughness = 0.4f;
    d.material.specular = 0.04f;
    d.light_count = 1;
    d.lights[0].kind = SYF_LIGHT_ENV;
    for (int i = 0; i < 3; ++i) d.lights[0].zenith[i] = d.lights[0].horizon[i] = d.lights[0].ground[i] = 1.0f;
    d.lights[0].intensity = 1.0f;
    d.width = d.height = 512;
    d.max_bounces = 2;
    d.max_steps = 256;
    d.pixel_tolerance = 1.0f;
    d.seed = 1;
    return d;
}

// Autotuner.  For each kernel family it times its calibration scene, in one persistent launch
// over 64-pixel items holding all 4 samples (4096 items, the least a launch holds on a 256-CU
// device; the better of two launches), choosing in turn:
//  * the field evaluations per loop trip, 4, 8 or 16, on the default grid of 3 workgroups per
//    CU with 8 x 8 tiles;
//  * the grid: 2 workgroups per CU against 3.  The grid serves latency-mode launches, which are
//    of that size; quiet-mode launches, many times longer, keep 3 workgroups per CU (with the
//    tuned grid they took the uneven batch from 535 to 495 Mrays/s, while the 200/s serving
//    loop gained 350 -> 363);
//  * the tile shape, 16 x 4 or 32 x 2 pixels against 8 x 8 (rows of a tile share more of their
//    paths' first bounces, squares more of their march lengths);
//  * the quiet-mode samples per pass, 8 against 16, from the time per sample of launches whose
//    items hold 8 and 16 samples (one launch each).  Longer items amortise a lane refill and a
//    claim over more paths and lengthen the launch's tail.
// A choice leaves the default only when it is at least 3 % faster: a first version on a 384^2,
// 2-sample scene (fewer items than the grid has waves) picked 1 workgroup per CU for some
// families and took the uneven 256-request batch from 535 to 392 Mrays/s.  A choice the
// configuration fixes (grid_blocks, path_steps, tile_width, quantum_samples) is not tuned.  The
// cost model (ns per pixel-sample-segment) is seeded with the power-8 Mandelbulb's best rate.
// None of the choices changes a served image: they only schedule the same per-pixel arithmetic.
void autotune(syf_server* s) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t cu = uint32_t(s->ctx->cu_count > 0 ? s->ctx->cu_count : 256);
    const uint32_t W = 512, H = 512, SPP = 4;
    const size_t pixels = size_t(W) * H;
    const uint32_t tiles = uint32_t(pixels / TILE_PIXELS);
    float4 *accum = nullptr, *aux = nullptr;
    float* depth = nullptr;
    double* depth64 = nullptr;
    JobP* djob = nullptr;
    BatchPlanes* dpl = nullptr;
    BatchItem* ditems = nullptr;
    uint32_t* claim = nullptr;
    uint32_t* exit_count = nullptr;
    unsigned long long* ctr = nullptr;
    unsigned long long* host_out = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    bool ok = hipMalloc(&accum, pixels * 16) == hipSuccess && hipMalloc(&aux, pixels * 16) == hipSuccess &&
              hipMalloc(&depth, pixels * 4) == hipSuccess && hipMalloc(&depth64, pixels * 8) == hipSuccess &&
              hipMalloc(&djob, sizeof(JobP)) == hipSuccess && hipMalloc(&dpl, sizeof(BatchPlanes)) == hipSuccess &&
              hipMalloc(&ditems, size_t(tiles) * sizeof(BatchItem)) == hipSuccess &&
              hipMalloc(&claim, CLAIM_WORDS * 4) == hipSuccess && hipMalloc(&ctr, CTR_COUNT * 8) == hipSuccess &&
              hipMalloc(&exit_count, 16 * 4) == hipSuccess &&
              hipHostMalloc(&host_out, CTR_COUNT * 8, hipHostMallocCoherent) == hipSuccess &&
              hipMemset(claim, 0, CLAIM_WORDS * 4) == hipSuccess && hipMemset(ctr, 0, CTR_COUNT * 8) == hipSuccess &&
              hipMemset(exit_count, 0, 16 * 4) == hipSuccess && hipEventCreate(&a) == hipSuccess &&
              hipEventCreate(&b) == hipSuccess;
    BatchPlanes pl{accum, depth, aux, depth64};
    ok = ok && hipMemcpy(dpl, &pl, sizeof pl, hipMemcpyHostToDevice) == hipSuccess;
    // the calibration's items: every 64-pixel tile of the image, 2^tws pixels wide, samples [0, ns)
    uint32_t items_tws = ~0u, items_ns = 0;
    auto set_items = [&](uint32_t tws, uint32_t ns) {
        if (!ok || (tws == items_tws && ns == items_ns)) return;
        const uint32_t tw = 1u << tws, th = TILE_PIXELS >> tws, tiles_x = W / tw;
        std::vector<BatchItem> items;
        for (uint32_t t = 0; t < tiles; ++t) items.push_back(BatchItem{0, (t % tiles_x) * tw, (t / tiles_x) * th, 0, ns, tws});
        ok = hipMemcpy(ditems, items.data(), size_t(tiles) * sizeof(BatchItem), hipMemcpyHostToDevice) == hipSuccess;
        items_tws = tws;
        items_ns = ns;
    };
    // device time of one launch of family f's calibration (the better of `reps`; the claim and
    // work counters and the exit count are left zero by each launch)
    auto time_launch = [&](uint32_t f, uint32_t grid, uint32_t steps, uint32_t tws, uint32_t ns, int reps) -> double {
        set_items(tws, ns);
        float best = 1e30f;
        for (int rep = 0; rep < reps && ok; ++rep) {
            float ms = 0.0f;
            ok = hipEventRecord(a, s->stream) == hipSuccess &&
                 launch_paths(f, djob, dpl, ditems, tiles, claim, ctr, exit_count, host_out, grid, steps, s->stream) ==
                     hipSuccess &&
                 hipEventRecord(b, s->stream) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
                 hipEventElapsedTime(&ms, a, b) == hipSuccess;
            best = std::min(best, ms);
        }
        return ok ? double(best) : 1e30;
    };
    constexpr double MARGIN = 0.97;  // a challenger must be 3 % faster than the incumbent
    for (uint32_t f = 0; ok && f < FAM_COUNT; ++f) {
        JobP P;
        const syf_job_desc d = calibration_desc(f);
        if (lower_job(d, P) != SYF_OK || kernel_family(P.f) != f) continue;
        ok = hipMemcpy(djob, &P, sizeof P, hipMemcpyHostToDevice) == hipSuccess;
        uint32_t steps = s->fam_steps[f], grid = s->fam_grid[f], tws = s->fam_tws[f];
        // the incumbent: the defaults or the fixed values
        double best = time_launch(f, grid, steps, tws, SPP, 2);
        if (s->cfg.path_steps == 0)
            for (uint32_t k : {4u, 16u}) {
                const double ms = time_launch(f, grid, k, tws, SPP, 2);
                if (ms < best * MARGIN) best = ms, steps = k;
            }
        if (s->cfg.grid_blocks == 0) {
            const double ms = time_launch(f, cu * 2, steps, tws, SPP, 2);
            if (ms < best * MARGIN) best = ms, grid = cu * 2;
        }
        const double best_4 = best;
        if (s->cfg.tile_width == 0)
            for (uint32_t w : {4u, 5u}) {
                const double ms = time_launch(f, grid, steps, w, SPP, 2);
                if (ms < best * MARGIN) best = ms, tws = w;
            }
        if (s->cfg.quantum_samples == 0) {
            // quiet-mode launches run on the default grid; ms per sample at 16 and 8 samples per item
            const uint32_t qgrid = s->grid;
            const double per16 = time_launch(f, qgrid, steps, tws, 16, 1) / 16.0;
            const double per8 = time_launch(f, qgrid, steps, tws, 8, 1) / 8.0;
            s->fam_quantum[f] = per8 < per16 * MARGIN ? 8u : 16u;
        }
        s->fam_steps[f] = steps;
        s->fam_grid[f] = grid;
        s->fam_tws[f] = tws;
        if (f == FAM_BULB8 && best_4 < 1e29)
            s->ns_per_unit = best * 1e6 / (double(pixels) * SPP * (d.max_bounces + 1));
    }
    hipFree(accum);
    hipFree(aux);
    hipFree(depth);
    hipFree(depth64);
    hipFree(djob);
    hipFree(dpl);
    hipFree(ditems);
    hipFree(claim);
    hipFree(ctr);
    hipFree(exit_count);
    hipHostFree(host_out);
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    s->autotune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

Request* find(syf_server* s, uint64_t id) {
    auto it = s->reqs.find(id);
    return it == s->reqs.end() ? nullptr : it->second;
}

}  // namespace

extern "C" {

int syf_server_create(syf_ctx* ctx, const syf_server_config* cfg, syf_server** out) {
    clear_error();
    if (!ctx || !cfg || !out) return fail(SYF_E_INVALID, "ctx, cfg and out must not be null");
    if (cfg->mode > SYF_SERVE_PLAIN) return fail(SYF_E_INVALID, "mode: unknown serving mode %u", cfg->mode);
    if (cfg->quantum_samples > 4096) return fail(SYF_E_INVALID, "quantum_samples must be at most 4096");
    if (cfg->path_steps != 0 && cfg->path_steps != 4 && cfg->path_steps != 8 && cfg->path_steps != 16)
        return fail(SYF_E_INVALID, "path_steps must be 0 (autotuned), 4, 8 or 16, got %u", cfg->path_steps);
    if (cfg->tile_width != 0 && cfg->tile_width != 8 && cfg->tile_width != 16 && cfg->tile_width != 32)
        return fail(SYF_E_INVALID, "tile_width must be 0 (autotuned), 8, 16 or 32, got %u", cfg->tile_width);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    syf_server* s = new syf_server();
    s->ctx = ctx;
    s->cfg = *cfg;
    if (s->cfg.max_active == 0) s->cfg.max_active = 4096;
    s->t0 = std::chrono::steady_clock::now();
    size_t free_b = 0, total_b = 0;
    hipMemGetInfo(&free_b, &total_b);
    s->budget = cfg->hbm_budget_bytes ? cfg->hbm_budget_bytes : uint64_t(double(free_b) * 0.9);
    s->max_slots = s->cfg.max_active;
    for (uint32_t i = 0; i < s->max_slots; ++i) s->free_slots.push_back(s->max_slots - 1u - i);
    const uint32_t cu = uint32_t(ctx->cu_count > 0 ? ctx->cu_count : 256);
    // one sample per item of a long request.  Two gained 8-15 % on single 1024^2 Mandelbulb and
    // Mandelbox requests, but not on the 256-request uneven batch (233 against 244-259 Mrays/s),
    // and it doubled the preview latency of the Poisson stream (p50 142 against 74 ms).
    s->quantum = cfg->quantum_samples ? cfg->quantum_samples : 1u;
    // a quantum the caller chose holds in both modes
    s->quantum_long = cfg->quantum_samples ? cfg->quantum_samples : QUIET_QUANTUM;
    s->grid = cfg->grid_blocks ? cfg->grid_blocks : cu * 3u;
    for (uint32_t f = 0; f < FAM_COUNT; ++f) {
        s->fam_grid[f] = s->grid;
        s->fam_steps[f] = cfg->path_steps ? cfg->path_steps : 8u;
        s->fam_tws[f] = cfg->tile_width == 32 ? 5u : cfg->tile_width == 16 ? 4u : 3u;
        s->fam_quantum[f] = s->quantum_long;
    }
    // each launch ends with a tail as long as its slowest pixel's quantum: launches of a few
    // tens of ms keep that below a few per cent, and a preview waits at most two of them
    s->target_us = cfg->batch_target_us ? cfg->batch_target_us : 30000u;
    s->quiet_us = cfg->quiet_target_us ? cfg->quiet_target_us
                                        : uint32_t(std::min(double(s->target_us) * QUIET_FACTOR, 4.0e9));
    s->fill_items = cu * 16u;  // two workgroups of four waves per CU, twice over
    s->ns_per_unit = 0.05;
    // The launch streams first: the runtime gives a new stream a hardware queue of its own while
    // the process holds fewer than GPU_MAX_HW_QUEUES (4 by default), and otherwise shares one.
    // Created before the server's other stream, the three launch streams take queues of their
    // own whenever the process has at most one other stream (its null stream), and launches
    // in flight then never wait for each other in a shared queue.
    for (int b = 0; b < NBUF; ++b)
        if ((e = hipStreamCreateWithFlags(&s->ls[b], hipStreamNonBlocking)) != hipSuccess) {
            syf_server_destroy(s);
            return hip_fail(e, "server streams");
        }
    if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&s->d_jobs, size_t(s->max_slots) * sizeof(JobP))) != hipSuccess ||
        (e = hipMalloc(&s->d_plain_job, sizeof(JobP))) != hipSuccess ||
        (e = hipMalloc(&s->d_planes, size_t(s->max_slots) * sizeof(BatchPlanes))) != hipSuccess) {
        syf_server_destroy(s);
        return hip_fail(e, "server allocation");
    }
    for (int b = 0; b < NBUF; ++b) {
        if ((e = hipMalloc(&s->d_claim[b], FAM_COUNT * CLAIM_WORDS * sizeof(uint32_t))) != hipSuccess ||
            (e = hipMalloc(&s->d_ctr[b], FAM_COUNT * CTR_COUNT * sizeof(unsigned long long))) != hipSuccess ||
            (e = hipMalloc(&s->d_exit[b], FAM_COUNT * 16 * sizeof(uint32_t))) != hipSuccess ||
            (e = hipHostMalloc(&s->h_out[b], FAM_COUNT * CTR_COUNT * sizeof(unsigned long long), hipHostMallocCoherent)) !=
                hipSuccess ||
            (e = hipMemset(s->d_claim[b], 0, FAM_COUNT * CLAIM_WORDS * sizeof(uint32_t))) != hipSuccess ||
            (e = hipMemset(s->d_ctr[b], 0, FAM_COUNT * CTR_COUNT * sizeof(unsigned long long))) != hipSuccess ||
            (e = hipMemset(s->d_exit[b], 0, FAM_COUNT * 16 * sizeof(uint32_t))) != hipSuccess) {
            syf_server_destroy(s);
            return hip_fail(e, "server allocation");
        }
    }
    if ((e = hipMalloc(&s->d_rctr, CTR_COUNT * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMalloc(&s->d_pctr, CTR_COUNT * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMemset(s->d_rctr, 0, CTR_COUNT * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipDeviceSynchronize()) != hipSuccess) {
        syf_server_destroy(s);
        return hip_fail(e, "server allocation");
    }
    for (auto& ev : s->ev_pool) {
        if ((e = hipEventCreate(&ev)) != hipSuccess) {
            syf_server_destroy(s);
            return hip_fail(e, "server events");
        }
    }
    if (s->cfg.mode == SYF_SERVE_BATCHED &&
        (cfg->grid_blocks == 0 || cfg->path_steps == 0 || cfg->tile_width == 0 || cfg->quantum_samples == 0))
        autotune(s);
    s->worker = std::thread(worker_main, s);
    *out = s;
    return SYF_OK;
}

int syf_server_destroy(syf_server* s) {
    clear_error();
    if (!s) return SYF_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->ctx->device);  // a failure shows in the calls that follow
    if (s->stream) hipStreamSynchronize(s->stream);
    // released requests still referred to by launches left behind by an error
    std::vector<Request*> orphans;
    for (auto& L : s->inflight)
        for (const Share& sh : L.progress)
            if (sh.r->released && std::find(orphans.begin(), orphans.end(), sh.r) == orphans.end())
                orphans.push_back(sh.r);
    for (Request* r : orphans) delete r;
    for (auto& kv : s->reqs) {
        free_request_planes(s, kv.second);
        delete kv.second;
    }
    hipFree(s->d_jobs);
    hipFree(s->d_plain_job);
    hipFree(s->d_planes);
    for (int b = 0; b < NBUF; ++b) {
        hipHostFree(s->h_items[b]);
        hipFree(s->d_claim[b]);
        hipFree(s->d_ctr[b]);
        hipFree(s->d_exit[b]);
        hipHostFree(s->h_out[b]);
    }
    hipFree(s->d_rctr);
    hipFree(s->d_pctr);
    for (auto& ev : s->ev_pool)
        if (ev) hipEventDestroy(ev);
    for (int b = 0; b < NBUF; ++b)
        if (s->ls[b]) {
            hipStreamSynchronize(s->ls[b]);
            hipStreamDestroy(s->ls[b]);
        }
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
    return SYF_OK;
}

int syf_server_submit(syf_server* s, const syf_job_desc* desc, uint32_t spp, uint64_t* id) {
    clear_error();
    if (!s || !desc || !id) return fail(SYF_E_INVALID, "server, desc and id must not be null");
    if (spp < 1 || spp > 65536) return fail(SYF_E_INVALID, "spp must be between 1 and 65536, got %u", spp);
    Request* r = new Request();
    int rc = lower_job(*desc, r->P);
    if (rc != SYF_OK) {
        delete r;
        return rc;
    }
    if (r->P.sh.fog_density > 0.0f) {
        delete r;
        return fail(SYF_E_INVALID, "fog.density: the server traces no participating medium; render a job with fog "
                                   "with syf_render_tile or syf_render_wavefront");
    }
    if (r->P.sh.detail) {
        delete r;
        return fail(SYF_E_INVALID, "finishes: the server renders no finish detail patterns; render a job with them "
                                   "with syf_render_tile or syf_render_wavefront");
    }
    r->spp = spp;
    r->pixels = size_t(desc->width) * desc->height;
    r->footprint = uint64_t(r->pixels) * BYTES_PER_PIXEL;
    const uint32_t fam = kernel_family(r->P.f);
    r->tws = s->fam_tws[fam];
    r->quantum_long = s->fam_quantum[fam];
    const uint32_t tw = 1u << r->tws, th = TILE_PIXELS >> r->tws;
    r->tiles_x = (desc->width + tw - 1) / tw;
    r->tiles = r->tiles_x * ((desc->height + th - 1) / th);
    r->total = uint64_t(r->tiles) * spp;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->error != SYF_OK) {
        delete r;
        return fail(s->error, "server stopped after an error: %s", s->error_text.c_str());
    }
    if (r->footprint > s->budget) {
        const uint64_t fp = r->footprint;
        delete r;
        return fail(SYF_E_ADMISSION, "request needs %llu device bytes, above the server's HBM budget of %llu",
                    (unsigned long long)fp, (unsigned long long)s->budget);
    }
    r->id = s->next_id++;
    r->submit_ms = s->now_ms();
    const double gap = r->submit_ms - s->last_submit_ms;
    if (gap < 5000.0) s->gap_ms = 0.9 * s->gap_ms + 0.1 * gap;
    s->last_submit_ms = r->submit_ms;
    s->reqs[r->id] = r;
    s->queue.push_back(r);
    s->submitted++;
    *id = r->id;
    s->cv.notify_all();
    return SYF_OK;
}

int syf_server_poll(syf_server* s, uint64_t id, syf_request_info* out) {
    clear_error();
    if (!s || !out) return fail(SYF_E_INVALID, "server and out must not be null");
    std::lock_guard<std::mutex> lk(s->mu);
    Request* r = find(s, id);
    if (!r) return fail(SYF_E_NOT_FOUND, "unknown request id %llu", (unsigned long long)id);
    std::memset(out, 0, sizeof *out);
    out->state = r->state;
    out->samples_done = r->tiles ? uint32_t(r->completed / r->tiles) : 0;
    out->spp = r->spp;
    out->footprint_bytes = r->footprint;
    out->submit_ms = r->submit_ms;
    out->start_ms = r->start_ms;
    out->done_ms = r->done_ms;
    return SYF_OK;
}

int syf_server_wait(syf_server* s, uint64_t id, uint32_t timeout_ms) {
    clear_error();
    if (!s) return fail(SYF_E_INVALID, "server must not be null");
    std::unique_lock<std::mutex> lk(s->mu);
    Request* r = find(s, id);
    if (!r) return fail(SYF_E_NOT_FOUND, "unknown request id %llu", (unsigned long long)id);
    const bool ok = s->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] {
        return r->state == SYF_REQ_DONE || r->state == SYF_REQ_CANCELLED || s->error != SYF_OK;
    });
    if (s->error != SYF_OK) return fail(s->error, "server error: %s", s->error_text.c_str());
    if (!ok) return fail(SYF_E_TIMEOUT, "request %llu not finished after %u ms", (unsigned long long)id, timeout_ms);
    return SYF_OK;
}

int syf_server_wait_all(syf_server* s, uint32_t timeout_ms) {
    clear_error();
    if (!s) return fail(SYF_E_INVALID, "server must not be null");
    std::unique_lock<std::mutex> lk(s->mu);
    const bool ok = s->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] {
        return s->error != SYF_OK || s->completed + s->cancelled >= s->submitted;
    });
    if (s->error != SYF_OK) return fail(s->error, "server error: %s", s->error_text.c_str());
    if (!ok) return fail(SYF_E_TIMEOUT, "requests still running after %u ms", timeout_ms);
    return SYF_OK;
}

int syf_server_result(syf_server* s, uint64_t id, uint32_t plane, void* host, size_t bytes) {
    clear_error();
    if (!s || !host) return fail(SYF_E_INVALID, "server and host must not be null");
    void* src = nullptr;
    size_t need = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        Request* r = find(s, id);
        if (!r) return fail(SYF_E_NOT_FOUND, "unknown request id %llu", (unsigned long long)id);
        if (r->state != SYF_REQ_DONE) return fail(SYF_E_INVALID, "request %llu is not finished", (unsigned long long)id);
        switch (plane) {
        case SYF_PLANE_ACCUM: src = r->accum; need = r->pixels * 16; break;
        case SYF_PLANE_DEPTH: src = r->depth; need = r->pixels * 4; break;
        case SYF_PLANE_AUX: src = r->aux; need = r->pixels * 16; break;
        case SYF_PLANE_DEPTH64: src = r->depth64; need = r->pixels * 8; break;
        default: return fail(SYF_E_INVALID, "plane: a request has ACCUM, DEPTH, AUX and DEPTH64 planes, got %u", plane);
        }
        if (!src) return fail(SYF_E_INVALID, "request %llu was released", (unsigned long long)id);
    }
    if (bytes != need) return fail(SYF_E_INVALID, "bytes must be %zu, got %zu", need, bytes);
    if (hipError_t de = hipSetDevice(s->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    hipError_t e = hipMemcpy(host, src, bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? SYF_OK : hip_fail(e, "syf_server_result");
}

int syf_server_release(syf_server* s, uint64_t id) {
    clear_error();
    if (!s) return fail(SYF_E_INVALID, "server must not be null");
    std::lock_guard<std::mutex> lk(s->mu);
    Request* r = find(s, id);
    if (!r) return fail(SYF_E_NOT_FOUND, "unknown request id %llu", (unsigned long long)id);
    if (r->state != SYF_REQ_DONE && r->state != SYF_REQ_CANCELLED)
        return fail(SYF_E_INVALID, "request %llu is still running", (unsigned long long)id);
    if (hipError_t de = hipSetDevice(s->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    // launches of a cancelled request may still write its planes
    hipStreamSynchronize(s->stream);
    for (int b = 0; b < NBUF; ++b)
        if (s->ls[b]) hipStreamSynchronize(s->ls[b]);
    free_request_planes(s, r);
    s->reqs.erase(r->id);
    r->released = true;
    if (r->refs == 0) delete r;  // otherwise the last launch that refers to it deletes it
    s->cv.notify_all();
    return SYF_OK;
}

int syf_server_cancel(syf_server* s, uint64_t id) {
    clear_error();
    if (!s) return fail(SYF_E_INVALID, "server must not be null");
    std::lock_guard<std::mutex> lk(s->mu);
    Request* r = find(s, id);
    if (!r) return fail(SYF_E_NOT_FOUND, "unknown request id %llu", (unsigned long long)id);
    if (r->state == SYF_REQ_DONE || r->state == SYF_REQ_CANCELLED) return SYF_OK;
    r->cancelled = true;
    if (r->state == SYF_REQ_QUEUED) {
        auto q = std::find(s->queue.begin(), s->queue.end(), r);
        if (q != s->queue.end()) s->queue.erase(q);
    } else {
        auto it = std::find(s->active.begin(), s->active.end(), r);
        if (it != s->active.end()) s->active.erase(it);
        // launches already holding its items still write its planes; syf_server_release
        // waits for them before freeing the planes
    }
    r->state = SYF_REQ_CANCELLED;
    s->cancelled++;
    s->cv.notify_all();
    return SYF_OK;
}

int syf_server_stats(syf_server* s, syf_server_counters* out) {
    clear_error();
    if (!s || !out) return fail(SYF_E_INVALID, "server and out must not be null");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof *out);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->cancelled = s->cancelled;
    out->launches = s->launches;
    out->rays = s->rays;
    out->primary_rays = s->primary;
    out->hbm_in_use = s->hbm_in_use;
    out->hbm_peak = s->hbm_peak;
    out->device_ms = s->device_ms;
    out->wall_ms = s->now_ms();
    out->quantum_samples = s->quantum;
    out->batch_target_us = s->target_us;
    out->grid_blocks = s->fam_grid[FAM_BULB8];
    for (uint32_t f = 0; f < FAM_COUNT && f < SYF_KERNEL_FAMILIES; ++f) {
        out->family_grid_blocks[f] = s->fam_grid[f];
        out->family_path_steps[f] = s->fam_steps[f];
        out->family_tile_width[f] = 1u << s->fam_tws[f];
        out->family_quiet_quantum[f] = s->fam_quantum[f];
    }
    out->autotune_ms = s->autotune_ms;
    return SYF_OK;
}

}  // extern "C"
