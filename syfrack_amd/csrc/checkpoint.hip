// Checkpoints and export payloads of a render job (host code).
//
// RenderSession::save / resume (session.rs:1199-1330) keep everything but the device planes in
// the caller's state.ron and the planes as raw little-endian files beside it; the export
// payload (export.rs:1226-1324) is one file of planes with its byte count and an FNV-1a-64
// checksum.  The Rust host keeps writing state.ron (its serde of Scene and RenderConfig); the
// library writes and reads the planes, replaces a checkpoint directory atomically, and gives
// the transport identity the host stores beside the scene.
#include <hip/hip_runtime.h>

#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <system_error>
#include <vector>

#include "api_internal.hpp"

namespace fs = std::filesystem;
using namespace syf;

namespace {

// The reference's state files (session.rs:33-45), plus this path's FP64 depth and step planes.
const char* const STATE_FILES[] = {"state.ron",  "accum.bin",          "depth.bin",   "normals.bin",
                                   "guiding.bin", "volume.bin",         "volume-moments.bin",
                                   "surface.bin", "volume-components.bin", "export.bin",
                                   "depth64.bin", "steps.bin"};

// Transport version of this library: changes whenever a kernel change alters any sample.
const char TRANSPORT_TAG[] = "syfrack_hip path tracer, classic graphs, transport 3\n";

uint64_t fnv1a(const void* data, size_t n, uint64_t hash) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) hash = (hash ^ uint64_t(p[i])) * 0x100000001b3ull;
    return hash;
}

int check_replaceable(const fs::path& dir) {
    if (dir.filename().empty()) return fail(SYF_E_INVALID, "checkpoint path %s has no final component", dir.c_str());
    std::error_code ec;
    if (!fs::is_directory(dir, ec)) return SYF_OK;
    for (const auto& e : fs::directory_iterator(dir, ec)) {
        const std::string name = e.path().filename().string();
        bool known = false;
        for (const char* f : STATE_FILES) known = known || name == f;
        if (!known)
            return fail(SYF_E_INVALID, "%s is not a SyFrack state directory (contains \"%s\"); refusing to replace it",
                        dir.c_str(), name.c_str());
    }
    if (ec) return fail(SYF_E_INVALID, "reading %s: %s", dir.c_str(), ec.message().c_str());
    return SYF_OK;
}

int write_file(const fs::path& p, const void* data, size_t n) {
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) return fail(SYF_E_INVALID, "creating %s: %s", p.c_str(), std::strerror(errno));
    const bool ok = (n == 0 || std::fwrite(data, 1, n, f) == n) && std::fflush(f) == 0 && fsync(fileno(f)) == 0;
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) return fail(SYF_E_INVALID, "writing %s: %s", p.c_str(), std::strerror(errno));
    return SYF_OK;
}

// Reads a whole file that must hold exactly `n` bytes.
int read_file(const fs::path& p, void* data, size_t n) {
    std::error_code ec;
    const auto size = fs::file_size(p, ec);
    if (ec) return fail(SYF_E_INVALID, "reading %s: %s", p.c_str(), ec.message().c_str());
    if (size != n) return fail(SYF_E_INVALID, "%s holds %llu bytes, expected %zu", p.c_str(), (unsigned long long)size, n);
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) return fail(SYF_E_INVALID, "opening %s: %s", p.c_str(), std::strerror(errno));
    const bool ok = n == 0 || std::fread(data, 1, n, f) == n;
    std::fclose(f);
    if (!ok) return fail(SYF_E_INVALID, "reading %s", p.c_str());
    return SYF_OK;
}

// The mean colour of the session's readback (session.rs decode_readback): sum / max(n, 1), alpha 1.
void mean_color(const std::vector<float>& accum, std::vector<float>& out) {
    out.resize(accum.size());
    for (size_t i = 0; i < accum.size(); i += 4) {
        const float n = std::max(accum[i + 3], 1.0f);
        out[i] = accum[i] / n;
        out[i + 1] = accum[i + 1] / n;
        out[i + 2] = accum[i + 2] / n;
        out[i + 3] = 1.0f;
    }
}

template <class T> int read_plane(syf_job* j, uint32_t plane, std::vector<T>& v) {
    v.resize(plane_bytes(plane, j->pixels) / sizeof(T));
    return syf_readback(j, plane, v.data(), v.size() * sizeof(T));
}

}  // namespace

extern "C" {

uint64_t syf_fnv1a64(const void* data, size_t n, uint64_t hash) {
    return data || n == 0 ? fnv1a(data, n, hash) : hash;
}

int syf_job_fingerprint(syf_job* j, uint64_t* out) {
    clear_error();
    if (!j || !out) return fail(SYF_E_INVALID, "job and out must not be null");
    uint64_t h = fnv1a(TRANSPORT_TAG, sizeof TRANSPORT_TAG - 1, 0xcbf29ce484222325ull);
    // the job parameters without the device address of a graph program, then the program itself
    JobP P = j->P;
    P.f.gprog = nullptr;
    h = fnv1a(&P, sizeof P, h);
    *out = j->gprog_host.empty() ? h : fnv1a(j->gprog_host.data(), j->gprog_host.size(), h);
    return SYF_OK;
}

int syf_upload(syf_job* j, uint32_t plane, const void* host, size_t bytes) {
    clear_error();
    if (!j || !host) return fail(SYF_E_INVALID, "job and host must not be null");
    if (plane > SYF_PLANE_SURFACE || (plane >= SYF_PLANE_VOLUME && !j->vplanes))
        return fail(SYF_E_INVALID, "plane: %u cannot be uploaded (without fog the session planes follow from ACCUM)",
                    plane);
    const size_t need = plane_bytes(plane, j->pixels);
    if (bytes != need) return fail(SYF_E_INVALID, "bytes must be %zu for plane %u, got %zu", need, plane, bytes);
    if (hipError_t de = hipSetDevice(j->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    void* dst = plane < SYF_PLANE_VOLUME ? j->plane[plane]
                                         : static_cast<char*>(j->vplanes) + (plane - SYF_PLANE_VOLUME) * j->pixels * 16;
    hipError_t e = hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, j->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(j->stream);
    return e == hipSuccess ? SYF_OK : hip_fail(e, "syf_upload");
}

int syf_checkpoint_save(syf_job* j, const char* dir_c, const char* state_text) {
    clear_error();
    if (!j || !dir_c || !state_text) return fail(SYF_E_INVALID, "job, dir and state_text must not be null");
    const fs::path dir(dir_c);
    if (int rc = check_replaceable(dir); rc != SYF_OK) return rc;
    const std::string pid = std::to_string(getpid());
    const fs::path tmp = dir.parent_path() / (dir.filename().string() + ".tmp-" + pid);
    const fs::path aside = dir.parent_path() / (dir.filename().string() + ".old-" + pid);
    std::error_code ec;
    fs::remove_all(tmp, ec);
    if (!fs::create_directories(tmp, ec) || ec)
        return fail(SYF_E_INVALID, "creating %s: %s", tmp.c_str(), ec ? ec.message().c_str() : "exists");
    int rc = SYF_OK;
    {
        std::vector<float> accum, depth, aux, vol, mom, surf;
        std::vector<double> depth64;
        std::vector<uint32_t> steps;
        const uint32_t planes[] = {SYF_PLANE_ACCUM, SYF_PLANE_DEPTH, SYF_PLANE_AUX, SYF_PLANE_VOLUME, SYF_PLANE_MOMENTS,
                                   SYF_PLANE_SURFACE};
        std::vector<float>* bufs[] = {&accum, &depth, &aux, &vol, &mom, &surf};
        const char* names[] = {"accum.bin", "depth.bin", "normals.bin", "volume.bin", "volume-moments.bin", "surface.bin"};
        for (int k = 0; k < 6 && rc == SYF_OK; ++k) {
            rc = read_plane(j, planes[k], *bufs[k]);
            if (rc == SYF_OK) rc = write_file(tmp / names[k], bufs[k]->data(), bufs[k]->size() * sizeof(float));
        }
        // the volume components are not enabled on this path: the reference keeps a 48-byte buffer
        const unsigned char components[48] = {0};
        if (rc == SYF_OK) rc = write_file(tmp / "volume-components.bin", components, sizeof components);
        if (rc == SYF_OK) rc = read_plane(j, SYF_PLANE_DEPTH64, depth64);
        if (rc == SYF_OK) rc = write_file(tmp / "depth64.bin", depth64.data(), depth64.size() * sizeof(double));
        if (rc == SYF_OK) rc = read_plane(j, SYF_PLANE_STEPS, steps);
        if (rc == SYF_OK) rc = write_file(tmp / "steps.bin", steps.data(), steps.size() * sizeof(uint32_t));
        if (rc == SYF_OK) rc = write_file(tmp / "state.ron", state_text, std::strlen(state_text));
    }
    if (rc != SYF_OK) {
        fs::remove_all(tmp, ec);
        return rc;
    }
    // one complete checkpoint always exists: the old one moves aside and goes only once the new
    // one is in place (session.rs:1277-1306)
    fs::remove_all(aside, ec);
    if (ec) {
        fs::remove_all(tmp, ec);
        return fail(SYF_E_INVALID, "clearing %s", aside.c_str());
    }
    const bool had_old = fs::exists(dir, ec);
    if (had_old) {
        fs::rename(dir, aside, ec);
        if (ec) {
            const std::string m = ec.message();
            fs::remove_all(tmp, ec);
            return fail(SYF_E_INVALID, "moving %s aside: %s", dir.c_str(), m.c_str());
        }
    }
    fs::rename(tmp, dir, ec);
    if (ec) {
        const std::string m = ec.message();
        if (had_old) fs::rename(aside, dir, ec);
        fs::remove_all(tmp, ec);
        return fail(SYF_E_INVALID, "renaming into %s: %s", dir.c_str(), m.c_str());
    }
    if (had_old) fs::remove_all(aside, ec);
    return SYF_OK;
}

int syf_checkpoint_load(syf_job* j, const char* dir_c) {
    clear_error();
    if (!j || !dir_c) return fail(SYF_E_INVALID, "job and dir must not be null");
    const fs::path dir(dir_c);
    const size_t px = j->pixels;
    std::vector<float> accum(4 * px), depth(px), aux(4 * px);
    int rc = read_file(dir / "accum.bin", accum.data(), accum.size() * 4);
    if (rc == SYF_OK) rc = read_file(dir / "depth.bin", depth.data(), depth.size() * 4);
    if (rc == SYF_OK) rc = read_file(dir / "normals.bin", aux.data(), aux.size() * 4);
    if (rc != SYF_OK) return rc;
    if ((rc = syf_upload(j, SYF_PLANE_ACCUM, accum.data(), accum.size() * 4)) != SYF_OK) return rc;
    if ((rc = syf_upload(j, SYF_PLANE_DEPTH, depth.data(), depth.size() * 4)) != SYF_OK) return rc;
    if ((rc = syf_upload(j, SYF_PLANE_AUX, aux.data(), aux.size() * 4)) != SYF_OK) return rc;
    // a job with fog resumes its session planes (volume.bin, volume-moments.bin, surface.bin)
    if (j->vplanes) {
        const char* names[] = {"volume.bin", "volume-moments.bin", "surface.bin"};
        std::vector<float> v(4 * px);
        for (uint32_t k = 0; k < 3; ++k) {
            if ((rc = read_file(dir / names[k], v.data(), v.size() * 4)) != SYF_OK) return rc;
            if ((rc = syf_upload(j, SYF_PLANE_VOLUME + k, v.data(), v.size() * 4)) != SYF_OK) return rc;
        }
    }
    // this path's own planes, when the checkpoint has them (a reference checkpoint does not)
    std::error_code ec;
    if (fs::exists(dir / "depth64.bin", ec)) {
        std::vector<double> d64(px);
        if ((rc = read_file(dir / "depth64.bin", d64.data(), px * 8)) != SYF_OK) return rc;
        if ((rc = syf_upload(j, SYF_PLANE_DEPTH64, d64.data(), px * 8)) != SYF_OK) return rc;
    }
    if (fs::exists(dir / "steps.bin", ec)) {
        std::vector<uint32_t> st(px);
        if ((rc = read_file(dir / "steps.bin", st.data(), px * 4)) != SYF_OK) return rc;
        if ((rc = syf_upload(j, SYF_PLANE_STEPS, st.data(), px * 4)) != SYF_OK) return rc;
    }
    return SYF_OK;
}

int syf_export_write(syf_job* j, const char* path, uint64_t* bytes, uint64_t* hash) {
    clear_error();
    if (!j || !path || !bytes || !hash) return fail(SYF_E_INVALID, "job, path, bytes and hash must not be null");
    std::vector<float> accum, depth, aux, color;
    int rc = read_plane(j, SYF_PLANE_ACCUM, accum);
    if (rc == SYF_OK) rc = read_plane(j, SYF_PLANE_DEPTH, depth);
    if (rc == SYF_OK) rc = read_plane(j, SYF_PLANE_AUX, aux);
    if (rc != SYF_OK) return rc;
    mean_color(accum, color);
    // the payload order of export.rs planes(): colour, depth, normals (no layers on this path)
    std::vector<unsigned char> payload;
    payload.reserve((color.size() + depth.size() + aux.size()) * 4);
    for (const std::vector<float>* v : {&color, &depth, &aux}) {
        const unsigned char* p = reinterpret_cast<const unsigned char*>(v->data());
        payload.insert(payload.end(), p, p + v->size() * 4);
    }
    if ((rc = write_file(path, payload.data(), payload.size())) != SYF_OK) return rc;
    *bytes = payload.size();
    *hash = fnv1a(payload.data(), payload.size(), 0xcbf29ce484222325ull);
    return SYF_OK;
}

int syf_export_read(const char* path, uint64_t bytes, uint64_t hash, uint32_t width, uint32_t height, float* color4,
                    float* depth, float* normals4) {
    clear_error();
    if (!path) return fail(SYF_E_INVALID, "path must not be null");
    const size_t px = size_t(width) * height;
    const size_t expected = px * (16 + 4 + 16);
    if (bytes != expected) return fail(SYF_E_INVALID, "export.bin has an invalid byte count");
    std::vector<unsigned char> payload(expected);
    if (int rc = read_file(path, payload.data(), expected); rc != SYF_OK) return rc;
    if (fnv1a(payload.data(), expected, 0xcbf29ce484222325ull) != hash)
        return fail(SYF_E_INVALID, "export.bin checksum mismatch");
    if (color4) std::memcpy(color4, payload.data(), px * 16);
    if (depth) std::memcpy(depth, payload.data() + px * 16, px * 4);
    if (normals4) std::memcpy(normals4, payload.data() + px * 20, px * 16);
    return SYF_OK;
}

}  // extern "C"
