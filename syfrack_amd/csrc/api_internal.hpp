// Internal definitions of the C ABI's opaque handles.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "params.hpp"

struct syf_ctx {
    int device = 0;
    int cu_count = 0;
    char arch[64] = {0};
    void* scratch = nullptr;  // this device's address of the allocation-free trace's scratch array
};

namespace syf {
struct WavefrontArena;
}

struct syf_job {
    syf_ctx* ctx = nullptr;
    syf_job_desc desc{};
    syf::JobP P{};
    size_t pixels = 0;
    hipStream_t stream = nullptr;
    void* plane[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    void* vplanes = nullptr;  // a job with fog: VOLUME, MOMENTS, SURFACE (float4 each), one allocation
    unsigned long long* ctr = nullptr;  // work counters: CTR_COUNT words inside the counters allocation
    uint32_t* hitq = nullptr;      // hit queue of the primary trace: HITQ_FIELDS fields x hitq_cap entries, then done flags
    size_t hitq_cap = 0;
    uint32_t hitq_epoch = 0;       // done-flag value of the latest primary trace (never 0)
    bool hitq_empty = false;       // every queue field is HITQ_EMPTY once the stream reaches the next trace
    uint32_t* counters = nullptr;  // ray counters, hit-queue count and cursor of the primary trace, then ctr
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;  // start, after dominant kernel, end
    bool timed = false;
    bool traced = false;  // a primary trace ran: its hit-queue error flag is checked by syf_job_stats
    double extra_ms = 0.0;
    syf::WavefrontArena* wf = nullptr;
    syf::JobP* dP = nullptr;         // the job constants in device memory (the path-tracing kernels read them there)
    syf::GOp* gprog = nullptr;       // a graph job's compiled program (device memory)
    std::vector<uint8_t> gprog_host;  // its bytes, for the job fingerprint
};

namespace syf {
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
void clear_error();
int lower_fractal(const syf_fractal& f, FractalP& F);
int lower_job(const syf_job_desc& d, JobP& P, const FractalP* graph_fractal = nullptr);
// Allocates a job's planes, counters and hit queue for a lowered description (syf_job_create).
int job_alloc(syf_ctx* ctx, const syf_job_desc& desc, const JobP& P, syf_job** out);
size_t plane_bytes(uint32_t plane, size_t pixels);
void release_wavefront(syf_job* j);
int noalloc_setup(int device, void** out);
// A graph in the interpreter's form, for callers that keep their own program buffer (the
// control machine): the job lowered from desc with the graph's fractal parameters, the compiled
// program and the node of each instruction.  Validates like syf_job_create_graph.
int graph_lower_interpreted(const syf_job_desc& desc, const syf_graph& graph, JobP& P, std::vector<GOp>& prog,
                            std::vector<uint32_t>& node_of);
}  // namespace syf
