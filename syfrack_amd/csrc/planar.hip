// 2D escape-time kernels: PlanarFractal::sample (planar.rs:77-194, planar.wgsl:16-91) for
// the parameter-plane families, in four modes:
//   FP32     direct iteration in f32, the reference's arithmetic (radius = |z| by sqrt);
//   FP64     direct iteration in f64 with the squared escape test |z|^2 > bailout^2, so
//            every operation is an exactly rounded IEEE add or multiply and the iteration
//            counts are bit-exact against any IEEE FP64 evaluation of the same loop;
//   PERTURB  deep zoom: a reference orbit Z_n (host, double-double, stored rounded to f32)
//            and per-pixel FP32 deltas
//            delta <- (2 Z + delta) delta + delta_c (Burning Ship: with diffabs).  A pixel
//            whose orbit loses precision against the reference (Pauldelbrot's criterion
//            |Z + delta| < 1e-3 |Z|) or outlives it is a glitch: it is queued (wave64 ballot,
//            one atomic per wave) and re-run against a new reference, the glitched pixel that
//            got furthest; after the reference budget (8 by default) the few left take the
//            PERTURB64 loop below, so no pixel is ever iterated from a rounded c;
//   PERTURB64 FP64 deltas against the view centre's orbit (double-double, stored rounded to
//            FP64) with rebasing (Zhuoran 2021):
//            when |Z_m + delta| < |delta|, or the reference has escaped, the delta takes the
//            whole z and the orbit index restarts at Z_0 = 0.  One reference serves every
//            pixel and c = centre + offset keeps the offset's full FP64 precision (about
//            1e-28 at span 1e-12), where the direct FP64 loop's c = fl(centre + offset) is
//            off by up to half an ulp of the centre: half the pixel spacing of a 4096-pixel
//            view at span 1e-12, and several pixels at smaller spans.
// Each wave owns an 8x8 pixel tile (neighbours escape at similar counts).  The pixel
// offsets are host-computed FP64 tables (one per column, one per row), so the device and
// any checker derive the same c from the same doubles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "api_internal.hpp"
#include "device.hpp"

namespace syf {

struct PlanarP {
    uint32_t family, power, iterations, width, height, tiles_x, tiles;
    uint32_t orbit_len;  // PERTURB: index of the last orbit entry
    double cx, cy;       // centre
    double cr, sr;       // rotation cos, sin
    double jx, jy;       // Julia constant
    float bailout;
    float b2f;           // bailout^2 in f32 (PERTURB escape test)
    double b2;           // bailout^2 in f64
};

__device__ __forceinline__ void pixel_of(const PlanarP& P, uint32_t& px, uint32_t& py, bool& valid) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t tile = blockIdx.x * 4u + wave;
    px = (tile % P.tiles_x) * 8u + (lane & 7u);
    py = (tile / P.tiles_x) * 8u + (lane >> 3);
    valid = tile < P.tiles && px < P.width && py < P.height;
}

__device__ __forceinline__ void offset_of(const PlanarP& P, const double* fx, const double* fy, uint32_t px,
                                          uint32_t py, double& dx, double& dy) {
    const double a = fx[px], b = fy[py];
    dx = P.cr * a - P.sr * b;
    dy = P.sr * a + P.cr * b;
}

__device__ __forceinline__ void flush_iters(uint32_t it, uint32_t rebases, unsigned long long* ctr) {
    uint32_t a = wave_sum(it), b = wave_sum(rebases);
    if ((threadIdx.x & 63u) == 0u) {
        atomicAdd(&ctr[0], (unsigned long long)a);
        atomicAdd(&ctr[1], (unsigned long long)b);
    }
}

// FP32 direct (planar.rs:107-164 restricted to the polynomial parameter-plane/Julia path,
// twist 0, no inversion).
template <uint32_t FAMILY>
__global__ __launch_bounds__(256) void k_planar_fp32(const PlanarP P, const double* __restrict__ fx,
                                                     const double* __restrict__ fy, float* __restrict__ smooth,
                                                     uint32_t* __restrict__ iters, unsigned long long* ctr) {
    uint32_t px, py;
    bool valid;
    pixel_of(P, px, py, valid);
    uint32_t count = 0;
    if (valid) {
        double dx, dy;
        offset_of(P, fx, fy, px, py, dx, dy);
        const float ptx = float(P.cx + dx), pty = float(P.cy + dy);
        const bool pp = FAMILY != SYF_PLANAR_JULIA;
        const float cx = pp ? ptx : float(P.jx), cy = pp ? pty : float(P.jy);
        float zx = pp ? 0.0f : ptx, zy = pp ? 0.0f : pty;
        bool escaped = false;
        for (uint32_t i = 0; i < P.iterations; ++i) {
            float radius = sqrtf(zx * zx + zy * zy);
            if (radius > P.bailout) {
                escaped = true;
                break;
            }
            float bx = zx, by = zy;
            if (FAMILY == SYF_PLANAR_BURNING_SHIP) {
                bx = fabsf(bx);
                by = fabsf(by);
            }
            float wx = 1.0f, wy = 0.0f;  // complex_power: repeated multiplication from 1
            for (uint32_t k = 0; k < P.power; ++k) {
                float nx = wx * bx - wy * by, ny = wx * by + wy * bx;
                wx = nx;
                wy = ny;
            }
            zx = wx + cx;
            zy = wy + cy;
            ++count;
        }
        float radius = fminf(sqrtf(zx * zx + zy * zy), 1e20f);
        escaped = escaped || radius > P.bailout;
        float s = float(P.iterations);
        if (escaped) s = float(count) + 1.0f - logf(logf(fmaxf(radius, 1.0001f))) / logf(float(P.power));
        const size_t i = size_t(py) * P.width + px;
        smooth[i] = s;
        iters[i] = count;
    }
    flush_iters(count, 0u, ctr);
}

// FP64 direct: the squared escape test, exact IEEE operations only.
template <uint32_t FAMILY>
__global__ __launch_bounds__(256) void k_planar_fp64(const PlanarP P, const double* __restrict__ fx,
                                                     const double* __restrict__ fy, float* __restrict__ smooth,
                                                     uint32_t* __restrict__ iters, unsigned long long* ctr) {
    uint32_t px, py;
    bool valid;
    pixel_of(P, px, py, valid);
    uint32_t count = 0;
    if (valid) {
        double dx, dy;
        offset_of(P, fx, fy, px, py, dx, dy);
        const double ptx = P.cx + dx, pty = P.cy + dy;
        const bool pp = FAMILY != SYF_PLANAR_JULIA;
        const double cx = pp ? ptx : P.jx, cy = pp ? pty : P.jy;
        double zx = pp ? 0.0 : ptx, zy = pp ? 0.0 : pty;
        bool escaped = false;
        for (uint32_t i = 0; i < P.iterations; ++i) {
            double r2 = zx * zx + zy * zy;
            if (r2 > P.b2) {
                escaped = true;
                break;
            }
            double x = zx, y = zy;
            if (FAMILY == SYF_PLANAR_BURNING_SHIP) {
                x = x < 0.0 ? -x : x;
                y = y < 0.0 ? -y : y;
            }
            double nx = x * x - y * y + cx;
            double ny = 2.0 * x * y + cy;
            zx = nx;
            zy = ny;
            ++count;
        }
        double r2 = zx * zx + zy * zy;
        escaped = escaped || r2 > P.b2;
        double r = fmin(sqrt(r2), 1e20);
        float s = float(P.iterations);
        if (escaped) s = float(double(count) + 1.0 - log(log(fmax(r, 1.0001))) / log(2.0));
        const size_t i = size_t(py) * P.width + px;
        smooth[i] = s;
        iters[i] = count;
    }
    flush_iters(count, 0u, ctr);
}

// |c + d| - |c| without cancellation (the Burning Ship perturbation of |x|).
__device__ __forceinline__ float diffabs(float c, float d) {
    const float cd = c + d;
    if (c >= 0.0f) return cd >= 0.0f ? d : -(2.0f * c + d);
    return cd > 0.0f ? 2.0f * c + d : -d;
}

// Glitch queue and reference choice of one perturbation pass.
struct PerturbCtl {
    uint32_t glitches;        // length of the output queue
    uint32_t pad;
    unsigned long long best;  // max over glitches of (iterations << 32 | ~pixel)
    uint32_t claim[8 * 16];   // per-XCD claim counters of 64-item groups, 64 bytes apart
};

// Reference-orbit entries a perturbation pass loads at once (one scalar load of 8 float2 each
// 8): the loop otherwise waits for a load at every iteration.
constexpr uint32_t ORBIT_CHUNK = 16;

// Perturbation pass against one reference orbit (rounded to f32 pairs), over the pixels of
// `list` (all pixels when list is null).  ref_dx/ref_dy: the reference's offset from the view
// centre, so a pixel's delta_c = its offset - the reference offset (FP64, then rounded).
template <uint32_t FAMILY>
__global__ __launch_bounds__(256) void k_planar_perturb(const PlanarP P, const double* __restrict__ fx,
                                                        const double* __restrict__ fy,
                                                        const float2* __restrict__ orbit, uint32_t orbit_len,
                                                        double ref_dx, double ref_dy,
                                                        const uint32_t* __restrict__ list, uint32_t n_items,
                                                        uint32_t* __restrict__ glitch_out, PerturbCtl* __restrict__ pc,
                                                        float* __restrict__ smooth, uint32_t* __restrict__ iters,
                                                        unsigned long long* ctr) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t done_iters = 0, glitched_total = 0;
    // Waves claim 64-item groups one at a time: a group's cost ranges from a few iterations to
    // all of them, and a fixed grid-stride assignment left long tails (measured 13 % slower on
    // the 4096^2 views).  Group g belongs to counter g % 8, the counter of the XCD that
    // workgroup blockIdx % 8 runs on; a wave whose counter is used up moves on to the next
    // ones, so every group is taken whatever the grid size.
    uint32_t xl = blockIdx.x % 8u;
    const uint32_t groups = (n_items + 63u) / 64u;
    for (uint32_t ranges_left = 8u; ranges_left > 0u;) {
        const uint32_t mine = groups > xl ? (groups - xl + 7u) / 8u : 0u;
        uint32_t k = 0;
        if (lane == 0u) k = atomicAdd(&pc->claim[xl * 16u], 1u);
        k = __builtin_amdgcn_readfirstlane(k);
        if (k >= mine) {
            xl = (xl + 1u) % 8u;
            --ranges_left;
            continue;
        }
        const uint32_t b0 = (k * 8u + xl) * 64u;
        const uint32_t item = b0 + lane;
        bool glitch = false;
        uint32_t pix = 0, count = 0;
        if (item < n_items) {
            pix = list ? list[item] : item;
            const uint32_t px = pix % P.width, py = pix / P.width;
            double ddx, ddy;
            offset_of(P, fx, fy, px, py, ddx, ddy);
            const float dcx = float(ddx - ref_dx), dcy = float(ddy - ref_dy);
            float ex = 0.0f, ey = 0.0f;
            bool escaped = false;
            float zx = 0.0f, zy = 0.0f;
            // The orbit index is the same on every lane still iterating, so the reference values
            // are wave-uniform loads; they are fetched ORBIT_CHUNK at a time.  The iteration is
            // branch-free: every lane computes it, and a lane that has stopped (escape or
            // glitch) keeps its state through selects (measured 1.5x faster than branching out
            // of the loop, whose mask updates took most of the issue slots).
            bool active = true;
            for (uint32_t n0 = 0; n0 < P.iterations; n0 += ORBIT_CHUNK) {
                float2 Zc[ORBIT_CHUNK];
#pragma unroll
                for (uint32_t k = 0; k < ORBIT_CHUNK; ++k) Zc[k] = orbit[min(n0 + k, P.iterations)];
#pragma unroll
                for (uint32_t k = 0; k < ORBIT_CHUNK; ++k) {
                    const uint32_t n = n0 + k;
                    const float2 Z = Zc[k];
                    const float zx1 = Z.x + ex, zy1 = Z.y + ey;
                    const float r2 = zx1 * zx1 + zy1 * zy1;
                    const bool live = active && n < P.iterations;
                    const bool esc = r2 > P.b2f;
                    // precision lost against the reference, or the reference has escaped
                    const bool gl = r2 < 1e-6f * (Z.x * Z.x + Z.y * Z.y) || n == orbit_len;
                    float nx, ny;
                    if (FAMILY == SYF_PLANAR_BURNING_SHIP) {
                        nx = (2.0f * Z.x + ex) * ex - (2.0f * Z.y + ey) * ey + dcx;
                        const float XY = Z.x * Z.y;
                        const float dXY = (Z.x + ex) * ey + ex * Z.y;  // (X+ex)(Y+ey) - XY
                        ny = 2.0f * diffabs(XY, dXY) + dcy;
                    } else {
                        const float tx = 2.0f * Z.x + ex, ty = 2.0f * Z.y + ey;
                        nx = tx * ex - ty * ey + dcx;
                        ny = tx * ey + ty * ex + dcy;
                    }
                    const bool step = live && !esc && !gl;
                    zx = live ? zx1 : zx;
                    zy = live ? zy1 : zy;
                    escaped = escaped || (live && esc);
                    glitch = glitch || (live && !esc && gl);
                    active = step;
                    ex = step ? nx : ex;
                    ey = step ? ny : ey;
                    count += step ? 1u : 0u;
                }
                if (__ballot(active) == 0ull) break;
            }
            if (!glitch && !escaped && count == P.iterations && P.iterations <= orbit_len) {
                const float2 Z = orbit[count];
                zx = Z.x + ex;
                zy = Z.y + ey;
                escaped = zx * zx + zy * zy > P.b2f;
            }
            if (!glitch) {
                float sm = float(P.iterations);
                if (escaped) {
                    float r = fminf(sqrtf(zx * zx + zy * zy), 1e20f);
                    sm = float(count) + 1.0f - logf(logf(fmaxf(r, 1.0001f))) / logf(2.0f);
                }
                smooth[pix] = sm;
                iters[pix] = count;
            }
            done_iters += count;
        }
        const unsigned long long mask = __ballot(glitch);
        const uint32_t total = uint32_t(__popcll(mask));
        uint32_t base = 0;
        if (lane == 0u && total) base = atomicAdd(&pc->glitches, total);
        base = __shfl(base, 0, 64);
        if (glitch) {
            glitch_out[base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)))] = pix;
            atomicMax(&pc->best, ((unsigned long long)count << 32) | (unsigned long long)(0xffffffffu - pix));
            ++glitched_total;
        }
    }
    flush_iters(done_iters, glitched_total, ctr);
}

// |c + d| - |c| without cancellation, in FP64.
__device__ __forceinline__ double diffabs64(double c, double d) {
    const double cd = c + d;
    if (c >= 0.0) return cd >= 0.0 ? d : -(2.0 * c + d);
    return cd > 0.0 ? 2.0 * c + d : -d;
}

// FP64 perturbation with rebasing over a pixel list (all pixels when list is null), against
// one FP64 orbit of C = centre + (ref_dx, ref_dy) whose last entry is orbit[orbit_len].
// A lane's orbit index m is its own (rebasing restarts it), so the orbit is read by per-lane
// gathers; it is 16 bytes an entry and stays in L2.  Every operation is an IEEE add or
// multiply rounded as written, so the counts are bit-exact against any FP64 restatement of
// this loop.
template <uint32_t FAMILY>
__global__ __launch_bounds__(256) void k_planar_rebase(const PlanarP P, const double* __restrict__ fx,
                                                       const double* __restrict__ fy,
                                                       const double2* __restrict__ orbit, uint32_t orbit_len,
                                                       double ref_dx, double ref_dy,
                                                       const uint32_t* __restrict__ list, uint32_t n_items,
                                                       float* __restrict__ smooth, uint32_t* __restrict__ iters,
                                                       unsigned long long* ctr) {
    uint32_t total = 0, rebased = 0;
    for (uint32_t item = blockIdx.x * blockDim.x + threadIdx.x; item < n_items; item += gridDim.x * blockDim.x) {
        const uint32_t pix = list ? list[item] : item;
        const uint32_t px = pix % P.width, py = pix / P.width;
        double ddx, ddy;
        offset_of(P, fx, fy, px, py, ddx, ddy);
        const double dcx = ddx - ref_dx, dcy = ddy - ref_dy;
        double ex = 0.0, ey = 0.0, zx = 0.0, zy = 0.0;
        uint32_t m = 0, count = 0;
        bool escaped = false;
        // Z = orbit[m] always; the next two entries are loaded two steps ahead, so a gather's
        // latency overlaps the arithmetic of the steps before it (entries past the last are
        // never used: the step at m == orbit_len rebases first)
        double2 Z = make_double2(0.0, 0.0);
        double2 Z1 = orbit[min(1u, orbit_len)], Z2 = orbit[min(2u, orbit_len)];
        for (uint32_t i = 0; i < P.iterations; ++i) {
            zx = Z.x + ex;
            zy = Z.y + ey;
            const double r2 = zx * zx + zy * zy;
            if (r2 > P.b2) {
                escaped = true;
                break;
            }
            if (m == orbit_len || r2 < ex * ex + ey * ey) {  // rebase onto Z_0 = 0
                ex = zx;
                ey = zy;
                m = 0;
                Z = make_double2(0.0, 0.0);
                Z1 = orbit[min(1u, orbit_len)];
                Z2 = orbit[min(2u, orbit_len)];
                ++rebased;
            }
            double nx, ny;
            if (FAMILY == SYF_PLANAR_BURNING_SHIP) {
                nx = (2.0 * Z.x + ex) * ex - (2.0 * Z.y + ey) * ey + dcx;
                const double XY = Z.x * Z.y;
                const double dXY = (Z.x + ex) * ey + ex * Z.y;
                ny = 2.0 * diffabs64(XY, dXY) + dcy;
            } else {
                const double tx = 2.0 * Z.x + ex, ty = 2.0 * Z.y + ey;
                nx = tx * ex - ty * ey + dcx;
                ny = tx * ey + ty * ex + dcy;
            }
            ex = nx;
            ey = ny;
            ++m;
            ++count;
            Z = Z1;
            Z1 = Z2;
            Z2 = orbit[min(m + 2u, orbit_len)];
        }
        if (!escaped) {
            zx = Z.x + ex;
            zy = Z.y + ey;
        }
        const double r2 = zx * zx + zy * zy;
        escaped = escaped || r2 > P.b2;
        float sm = float(P.iterations);
        if (escaped) sm = float(double(count) + 1.0 - log(log(fmax(fmin(sqrt(r2), 1e20), 1.0001))) / log(2.0));
        smooth[pix] = sm;
        iters[pix] = count;
        total += count;
    }
    const uint32_t a = wave_sum(total), b = wave_sum(rebased);
    if ((threadIdx.x & 63u) == 0u) {
        atomicAdd(&ctr[0], (unsigned long long)a);
        atomicAdd(&ctr[2], (unsigned long long)b);
    }
}

}  // namespace syf

using namespace syf;

struct syf_planar {
    syf_ctx* ctx = nullptr;
    syf_planar_desc desc{};
    PlanarP P{};
    hipStream_t stream = nullptr;
    double* fx = nullptr;
    double* fy = nullptr;
    float* smooth = nullptr;
    uint32_t* iters = nullptr;
    float2* orbit = nullptr;
    double2* orbit64 = nullptr;             // PERTURB64 / fallback: the centre's orbit in FP64
    size_t orbit_cap = 0;
    uint32_t* list[2] = {nullptr, nullptr};  // PERTURB glitch queues (ping-pong)
    PerturbCtl* pc = nullptr;
    std::vector<double> hfx, hfy;           // host copies of the offset tables
    uint32_t max_references = 8;            // PERTURB: FP32-delta references before the fallback
    uint32_t references = 0;                // references used by the last PERTURB run
    uint32_t fallback = 0;                  // pixels left to the FP64 rebasing loop by the last run
    unsigned long long* ctr = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
};

namespace {

void planar_free(syf_planar* p) {
    if (!p) return;
    // teardown: every release is attempted, whatever an earlier one returned
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    (void)hipFree(p->fx);
    (void)hipFree(p->fy);
    (void)hipFree(p->smooth);
    (void)hipFree(p->iters);
    (void)hipFree(p->orbit);
    (void)hipFree(p->orbit64);
    (void)hipFree(p->list[0]);
    (void)hipFree(p->list[1]);
    (void)hipFree(p->pc);
    (void)hipFree(p->ctr);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

// Double-double arithmetic (Dekker 1971; Knuth's TwoSum): hi + lo with |lo| <= ulp(hi)/2,
// about 106 bits, each step an IEEE operation rounded as written (fma gives the exact product
// error), so the oracle's restatement of these steps produces the same bits.
struct DD {
    double hi, lo;
};
DD dd_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}
DD dd_norm(double s, double e) {  // |s| >= |e|
    const double h = s + e;
    return DD{h, e - (h - s)};
}
DD dd_add(DD a, DD b) {
    const DD s = dd_two_sum(a.hi, b.hi);
    return dd_norm(s.hi, s.lo + (a.lo + b.lo));
}
DD dd_mul(DD a, DD b) {
    const double p = a.hi * b.hi;
    return dd_norm(p, std::fma(a.hi, b.hi, -p) + (a.hi * b.lo + a.lo * b.hi));
}
DD dd_neg(DD a) { return DD{-a.hi, -a.lo}; }
DD dd_abs(DD a) { return a.hi < 0.0 ? dd_neg(a) : a; }

// Reference orbit of C = centre + (ox, oy), held exactly as a double-double: Z_0 = 0,
// Z_{n+1} = f(Z_n) + C iterated in double-double (Burning Ship on |Re|, |Im|) and stored
// rounded to FP64 (and, when f32 is not null, to f32), up to the first entry with
// |Z|^2 > bailout^2 or the iteration budget; returns the index of its last entry.  An orbit
// iterated in FP64 itself drifts from C's true orbit in chaotic views (its rounding errors
// grow with the dynamics), and every pixel perturbed against it inherits the drift; this
// one is the true orbit to about 1e-32 per step.
uint32_t reference_orbit(const syf_planar_desc& d, double ox, double oy, std::vector<double2>& out,
                         std::vector<float2>* f32) {
    const double b2 = double(d.bailout) * double(d.bailout);
    const bool ship = d.family == SYF_PLANAR_BURNING_SHIP;
    const DD cx = dd_two_sum(d.center[0], ox), cy = dd_two_sum(d.center[1], oy);
    DD zx{0.0, 0.0}, zy{0.0, 0.0};
    out.clear();
    out.reserve(size_t(d.iterations) + 1);
    out.push_back(make_double2(0.0, 0.0));
    for (uint32_t n = 0; n < d.iterations; ++n) {
        const DD x = ship ? dd_abs(zx) : zx, y = ship ? dd_abs(zy) : zy;
        const DD xx = dd_mul(x, x), yy = dd_mul(y, y), xy = dd_mul(x, y);
        zx = dd_add(dd_add(xx, dd_neg(yy)), cx);
        zy = dd_add(DD{2.0 * xy.hi, 2.0 * xy.lo}, cy);
        out.push_back(make_double2(zx.hi, zy.hi));
        if (zx.hi * zx.hi + zy.hi * zy.hi > b2) break;
    }
    if (f32) {
        f32->resize(out.size());
        for (size_t i = 0; i < out.size(); ++i) (*f32)[i] = make_float2(float(out[i].x), float(out[i].y));
    }
    return uint32_t(out.size() - 1);
}

// PERTURB64 over `list` (all pixels when null): the view centre's FP64 orbit, FP64 deltas
// with rebasing.
hipError_t run_rebase(syf_planar* p, const uint32_t* list, uint32_t n_items) {
    const syf_planar_desc& d = p->desc;
    std::vector<double2> orbit;
    const uint32_t orbit_len = reference_orbit(d, 0.0, 0.0, orbit, nullptr);
    hipError_t e = hipMemcpyAsync(p->orbit64, orbit.data(), orbit.size() * sizeof(double2), hipMemcpyHostToDevice,
                                  p->stream);
    // the copy reads the host vector: it must finish before the vector is freed
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return e;
    const int cus = p->ctx->cu_count > 0 ? p->ctx->cu_count : 256;
    const dim3 block(256), grid(std::min<uint32_t>((n_items + 255u) / 256u, uint32_t(cus) * 8u));
    if (d.family == SYF_PLANAR_MANDELBROT)
        hipLaunchKernelGGL(k_planar_rebase<SYF_PLANAR_MANDELBROT>, grid, block, 0, p->stream, p->P, p->fx, p->fy,
                           p->orbit64, orbit_len, 0.0, 0.0, list, n_items, p->smooth, p->iters, p->ctr);
    else
        hipLaunchKernelGGL(k_planar_rebase<SYF_PLANAR_BURNING_SHIP>, grid, block, 0, p->stream, p->P, p->fx, p->fy,
                           p->orbit64, orbit_len, 0.0, 0.0, list, n_items, p->smooth, p->iters, p->ctr);
    return hipGetLastError();
}

// Perturbation passes: reference 0 is the view centre; each later reference is the glitched
// pixel of the previous pass that got furthest (ties: lowest pixel index), and only that
// pass's glitches are re-run.  Glitches left after max_references take the FP64 rebasing
// loop (PERTURB64) against the centre.
hipError_t run_perturbation(syf_planar* p) {
    const syf_planar_desc& d = p->desc;
    const PlanarP& P = p->P;
    const uint32_t n = d.width * d.height;
    const int cus = p->ctx->cu_count > 0 ? p->ctx->cu_count : 256;
    const dim3 block(256);
    double ref_dx = 0.0, ref_dy = 0.0;
    const uint32_t* in = nullptr;  // all pixels
    uint32_t n_in = n, which = 0;
    std::vector<double2> orbit64;
    std::vector<float2> orbit;
    hipError_t e = hipSuccess;
    p->references = 0;
    p->fallback = 0;
    for (uint32_t pass = 0; pass < p->max_references && n_in > 0; ++pass) {
        const uint32_t orbit_len = reference_orbit(d, ref_dx, ref_dy, orbit64, &orbit);
        e = hipMemcpyAsync(p->orbit, orbit.data(), orbit.size() * sizeof(float2), hipMemcpyHostToDevice, p->stream);
        if (e == hipSuccess) e = hipMemsetAsync(p->pc, 0, sizeof(PerturbCtl), p->stream);
        if (e != hipSuccess) return e;
        const dim3 grid(std::min<uint32_t>((n_in + 255u) / 256u, uint32_t(cus) * 8u));
        uint32_t* out = p->list[which];
        if (d.family == SYF_PLANAR_MANDELBROT)
            hipLaunchKernelGGL(k_planar_perturb<SYF_PLANAR_MANDELBROT>, grid, block, 0, p->stream, P, p->fx, p->fy,
                               p->orbit, orbit_len, ref_dx, ref_dy, in, n_in, out, p->pc, p->smooth, p->iters, p->ctr);
        else
            hipLaunchKernelGGL(k_planar_perturb<SYF_PLANAR_BURNING_SHIP>, grid, block, 0, p->stream, P, p->fx, p->fy,
                               p->orbit, orbit_len, ref_dx, ref_dy, in, n_in, out, p->pc, p->smooth, p->iters, p->ctr);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        PerturbCtl pc{};
        e = hipMemcpyAsync(&pc, p->pc, sizeof pc, hipMemcpyDeviceToHost, p->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
        if (e != hipSuccess) return e;
        p->references = pass + 1;
        n_in = pc.glitches;
        in = out;
        which ^= 1u;
        if (n_in == 0) break;
        const uint32_t pix = 0xffffffffu - uint32_t(pc.best & 0xffffffffull);
        const uint32_t px = pix % d.width, py = pix / d.width;
        ref_dx = P.cr * p->hfx[px] - P.sr * p->hfy[py];
        ref_dy = P.sr * p->hfx[px] + P.cr * p->hfy[py];
    }
    if (n_in > 0) {
        p->fallback = n_in;
        e = run_rebase(p, in, n_in);
    }
    return e;
}

}  // namespace

extern "C" {

int syf_planar_create(syf_ctx* ctx, const syf_planar_desc* d, syf_planar** out) {
    clear_error();
    if (!ctx || !d || !out) return fail(SYF_E_INVALID, "ctx, desc and out must not be null");
    if (d->family != SYF_PLANAR_MANDELBROT && d->family != SYF_PLANAR_BURNING_SHIP && d->family != SYF_PLANAR_JULIA)
        return fail(SYF_E_INVALID, "family: unsupported planar family %u (Julia 1, Mandelbrot 2, Burning Ship 3)",
                    d->family);
    if (d->mode > SYF_PLANAR_PERTURB64) return fail(SYF_E_INVALID, "mode: unknown mode %u", d->mode);
    const bool perturb = d->mode == SYF_PLANAR_PERTURB || d->mode == SYF_PLANAR_PERTURB64;
    if (perturb && d->family == SYF_PLANAR_JULIA)
        return fail(SYF_E_INVALID, "mode: perturbation is for the parameter-plane families");
    if (d->max_references > 64)
        return fail(SYF_E_INVALID, "max_references must be at most 64 (0 selects 8), got %u", d->max_references);
    if (d->width < 1 || d->height < 1 || d->width > 32768 || d->height > 32768)
        return fail(SYF_E_INVALID, "width x height (%ux%u) must be within 1..32768 per side", d->width, d->height);
    if (d->iterations < 1 || d->iterations > (1u << 24))
        return fail(SYF_E_INVALID, "iterations must be between 1 and 2^24, got %u", d->iterations);
    if (!(d->bailout >= 2.0f && d->bailout <= 1e4f))
        return fail(SYF_E_INVALID, "bailout must be between 2 and 1e4, got %g", double(d->bailout));
    if (!(d->span > 0.0) || !std::isfinite(d->span)) return fail(SYF_E_INVALID, "span must be positive and finite");
    if (!std::isfinite(d->center[0]) || !std::isfinite(d->center[1]))
        return fail(SYF_E_INVALID, "center must be finite");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    syf_planar* p = new syf_planar();
    p->ctx = ctx;
    p->desc = *d;
    p->max_references = d->max_references ? d->max_references : 8u;
    PlanarP& P = p->P;
    P.family = d->family;
    P.power = 2;
    P.iterations = d->iterations;
    P.width = d->width;
    P.height = d->height;
    P.tiles_x = (d->width + 7u) / 8u;
    P.tiles = P.tiles_x * ((d->height + 7u) / 8u);
    P.cx = d->center[0];
    P.cy = d->center[1];
    const double rot = double(d->rotation) * (3.14159265358979323846 / 180.0);
    P.cr = std::cos(rot);
    P.sr = std::sin(rot);
    P.jx = d->julia[0];
    P.jy = d->julia[1];
    P.bailout = d->bailout;
    P.b2f = d->bailout * d->bailout;
    P.b2 = double(d->bailout) * double(d->bailout);
    const size_t n = size_t(d->width) * d->height;
    std::vector<double> fx(d->width), fy(d->height);
    for (uint32_t i = 0; i < d->width; ++i) fx[i] = ((double(i) + 0.5) / double(d->width) - 0.5) * d->span;
    const double vspan = d->span * double(d->height) / double(d->width);
    for (uint32_t j = 0; j < d->height; ++j) fy[j] = (0.5 - (double(j) + 0.5) / double(d->height)) * vspan;
    if ((e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&p->fx, fx.size() * 8)) != hipSuccess || (e = hipMalloc(&p->fy, fy.size() * 8)) != hipSuccess ||
        (e = hipMalloc(&p->smooth, n * 4)) != hipSuccess || (e = hipMalloc(&p->iters, n * 4)) != hipSuccess ||
        (e = hipMalloc(&p->ctr, 3 * sizeof(unsigned long long))) != hipSuccess || (e = hipEventCreate(&p->ev0)) != hipSuccess ||
        (e = hipEventCreate(&p->ev1)) != hipSuccess) {
        planar_free(p);
        return hip_fail(e, "syf_planar_create allocation");
    }
    if (perturb) {
        p->orbit_cap = size_t(d->iterations) + 1;
        if ((e = hipMalloc(&p->orbit64, p->orbit_cap * sizeof(double2))) != hipSuccess) {
            planar_free(p);
            return hip_fail(e, "syf_planar_create perturbation buffers");
        }
    }
    if (d->mode == SYF_PLANAR_PERTURB) {
        if ((e = hipMalloc(&p->orbit, p->orbit_cap * sizeof(float2))) != hipSuccess ||
            (e = hipMalloc(&p->list[0], n * 4)) != hipSuccess || (e = hipMalloc(&p->list[1], n * 4)) != hipSuccess ||
            (e = hipMalloc(&p->pc, sizeof(PerturbCtl))) != hipSuccess) {
            planar_free(p);
            return hip_fail(e, "syf_planar_create perturbation buffers");
        }
    }
    p->hfx = fx;
    p->hfy = fy;
    e = hipMemcpy(p->fx, fx.data(), fx.size() * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->fy, fy.data(), fy.size() * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        planar_free(p);
        return hip_fail(e, "syf_planar_create upload");
    }
    *out = p;
    return SYF_OK;
}

int syf_planar_destroy(syf_planar* p) {
    clear_error();
    planar_free(p);
    return SYF_OK;
}

int syf_planar_run(syf_planar* p) {
    clear_error();
    if (!p) return fail(SYF_E_INVALID, "planar must not be null");
    if (hipError_t de = hipSetDevice(p->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    const syf_planar_desc& d = p->desc;
    hipError_t e = hipMemsetAsync(p->ctr, 0, 3 * sizeof(unsigned long long), p->stream);
    if (e == hipSuccess) e = hipEventRecord(p->ev0, p->stream);
    const dim3 grid((p->P.tiles + 3u) / 4u), block(256);
    const PlanarP& P = p->P;
    if (e == hipSuccess) {
        switch (d.mode) {
        case SYF_PLANAR_FP32:
            if (d.family == SYF_PLANAR_MANDELBROT)
                hipLaunchKernelGGL(k_planar_fp32<SYF_PLANAR_MANDELBROT>, grid, block, 0, p->stream, P, p->fx, p->fy, p->smooth, p->iters, p->ctr);
            else if (d.family == SYF_PLANAR_BURNING_SHIP)
                hipLaunchKernelGGL(k_planar_fp32<SYF_PLANAR_BURNING_SHIP>, grid, block, 0, p->stream, P, p->fx, p->fy, p->smooth, p->iters, p->ctr);
            else
                hipLaunchKernelGGL(k_planar_fp32<SYF_PLANAR_JULIA>, grid, block, 0, p->stream, P, p->fx, p->fy, p->smooth, p->iters, p->ctr);
            break;
        case SYF_PLANAR_FP64:
            if (d.family == SYF_PLANAR_MANDELBROT)
                hipLaunchKernelGGL(k_planar_fp64<SYF_PLANAR_MANDELBROT>, grid, block, 0, p->stream, P, p->fx, p->fy, p->smooth, p->iters, p->ctr);
            else if (d.family == SYF_PLANAR_BURNING_SHIP)
                hipLaunchKernelGGL(k_planar_fp64<SYF_PLANAR_BURNING_SHIP>, grid, block, 0, p->stream, P, p->fx, p->fy, p->smooth, p->iters, p->ctr);
            else
                hipLaunchKernelGGL(k_planar_fp64<SYF_PLANAR_JULIA>, grid, block, 0, p->stream, P, p->fx, p->fy, p->smooth, p->iters, p->ctr);
            break;
        case SYF_PLANAR_PERTURB:
            e = run_perturbation(p);
            break;
        default:
            p->references = 1;
            p->fallback = 0;
            e = run_rebase(p, nullptr, d.width * d.height);
            break;
        }
        if (e == hipSuccess) e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(p->ev1, p->stream);
    p->timed = true;
    return e == hipSuccess ? SYF_OK : hip_fail(e, "syf_planar_run");
}

int syf_planar_sync(syf_planar* p) {
    clear_error();
    if (!p) return fail(SYF_E_INVALID, "planar must not be null");
    hipError_t e = hipStreamSynchronize(p->stream);
    return e == hipSuccess ? SYF_OK : hip_fail(e, "syf_planar_sync");
}

int syf_planar_readback(syf_planar* p, uint32_t plane, void* host, size_t bytes) {
    clear_error();
    if (!p || !host) return fail(SYF_E_INVALID, "planar and host must not be null");
    if (plane > SYF_PLANAR_ITER) return fail(SYF_E_INVALID, "plane: unknown planar plane %u", plane);
    const size_t need = size_t(p->desc.width) * p->desc.height * 4;
    if (bytes != need) return fail(SYF_E_INVALID, "bytes must be %zu, got %zu", need, bytes);
    if (hipError_t de = hipSetDevice(p->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    hipError_t e = hipMemcpyAsync(host, plane == SYF_PLANAR_SMOOTH ? (void*)p->smooth : (void*)p->iters, bytes,
                                  hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    return e == hipSuccess ? SYF_OK : hip_fail(e, "syf_planar_readback");
}

int syf_planar_stats(syf_planar* p, uint64_t* iterations, uint64_t* rebases, double* kernel_ms) {
    clear_error();
    if (!p) return fail(SYF_E_INVALID, "planar must not be null");
    unsigned long long c[3] = {0, 0, 0};
    hipError_t e = hipMemcpyAsync(c, p->ctr, sizeof c, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return hip_fail(e, "syf_planar_stats");
    float ms = 0.0f;
    if (p->timed && hipEventElapsedTime(&ms, p->ev0, p->ev1) != hipSuccess) ms = 0.0f;
    if (iterations) *iterations = c[0];
    if (rebases) *rebases = c[1];
    if (kernel_ms) *kernel_ms = ms;
    return SYF_OK;
}

int syf_planar_glitches(syf_planar* p, uint32_t* references, uint32_t* fallback_pixels, uint64_t* rebasings) {
    clear_error();
    if (!p) return fail(SYF_E_INVALID, "planar must not be null");
    unsigned long long c[3] = {0, 0, 0};
    hipError_t e = hipMemcpyAsync(c, p->ctr, sizeof c, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return hip_fail(e, "syf_planar_glitches");
    if (references) *references = p->references;
    if (fallback_pixels) *fallback_pixels = p->fallback;
    if (rebasings) *rebasings = c[2];
    return SYF_OK;
}

}  // extern "C"
