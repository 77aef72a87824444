// Entry points whose kernels are not built yet: they fail loudly.
#include <hip/hip_runtime.h>

#include "api_internal.hpp"

namespace syf {
void release_wavefront(syf_job*) {}
}  // namespace syf

using namespace syf;

extern "C" {
int syf_render_tile(syf_job*, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t) {
    return fail(SYF_E_INVALID, "syf_render_tile is not available in this build");
}
int syf_render_wavefront(syf_job*, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t) {
    return fail(SYF_E_INVALID, "syf_render_wavefront is not available in this build");
}
int syf_planar_create(syf_ctx*, const syf_planar_desc*, syf_planar**) {
    return fail(SYF_E_INVALID, "syf_planar_create is not available in this build");
}
int syf_planar_destroy(syf_planar*) { return SYF_OK; }
int syf_planar_run(syf_planar*) { return fail(SYF_E_INVALID, "not available"); }
int syf_planar_sync(syf_planar*) { return fail(SYF_E_INVALID, "not available"); }
int syf_planar_readback(syf_planar*, uint32_t, void*, size_t) { return fail(SYF_E_INVALID, "not available"); }
int syf_planar_stats(syf_planar*, uint64_t*, uint64_t*, double*) { return fail(SYF_E_INVALID, "not available"); }
}
