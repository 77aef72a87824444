// Device-side building blocks of the gfx950 distance-estimator path: vector algebra, the
// classic distance fields, the PCG stream, camera rays, the sphere-tracing march, the
// tetrahedral normal and the FP64 final-hit refine.
//
// Everything is written for the wave64 FP32 VALU (no MFMA: per-ray branchy arithmetic).
// Functions are templated on the arithmetic type: float for the march, double for the
// final-hit refine.  The library is compiled with -ffp-contract=off, so each expression
// rounds as written, in the order of the reference formulas cited below (paths relative to
// the SyFrack source tree).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "params.hpp"

#ifndef INFINITY  // run-time compiled units have no <math.h>
#define INFINITY __builtin_inff()
#endif

namespace syf {

template <bool B, class X, class Y> struct cond_type { using type = X; };
template <class X, class Y> struct cond_type<false, X, Y> { using type = Y; };
template <bool B, class X, class Y> using cond_t = typename cond_type<B, X, Y>::type;

template <class T> struct V3 {
    T x, y, z;
};
template <class T> __device__ __forceinline__ V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> __device__ __forceinline__ V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> __device__ __forceinline__ V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> __device__ __forceinline__ V3<T> scl(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> __device__ __forceinline__ T dot3(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __device__ __forceinline__ T len3(V3<T> a) { return sqrt(dot3(a, a)); }
template <class T> __device__ __forceinline__ V3<T> nrm(V3<T> a) { return scl(a, T(1) / len3(a)); }
template <class T> __device__ __forceinline__ T clampT(T v, T lo, T hi) { return fmin(fmax(v, lo), hi); }
template <class T> __device__ __forceinline__ V3<T> cvt(V3<float> a) { return {T(a.x), T(a.y), T(a.z)}; }
template <class T> __device__ __forceinline__ V3<T> cross3(V3<T> a, V3<T> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ V3<float> ld3(const float* p) { return {p[0], p[1], p[2]}; }

// Result of one field evaluation: distance, normalised trap slot 0 (the classics write
// only slot 0; slots 1-3 are 1), iteration count clamped to the iteration product, and the
// number of Iterate-body executions (work accounting).
template <class T> struct De {
    T d;
    T trap0;
    uint32_t iter;
    uint32_t runs;
};

// integral_spherical_power (eval.rs:129-178).
template <class T>
__device__ __forceinline__ void spherical_power(V3<T>& p, T r, uint32_t n, T& deriv) {
    if (n == 1u) {
        deriv = T(1);
        return;
    }
    T rho = sqrt(p.x * p.x + p.y * p.y);
    T tx = clampT(p.z / r, T(-1), T(1)), ty = rho / r;
    T fx = T(1), fy = T(0);
    if (rho > T(0)) {
        fx = p.x / rho;
        fy = p.y / rho;
    }
    if (n == 8u || n == 6u) {
        T t2x = tx * tx - ty * ty, t2y = T(2) * tx * ty;
        T f2x = fx * fx - fy * fy, f2y = T(2) * fx * fy;
        T t4x = t2x * t2x - t2y * t2y, t4y = T(2) * t2x * t2y;
        T f4x = f2x * f2x - f2y * f2y, f4y = T(2) * f2x * f2y;
        T r2 = r * r;
        T r4 = r2 * r2;
        T ax, ay, bx, by, prev;
        if (n == 8u) {
            ax = t4x * t4x - t4y * t4y;
            ay = T(2) * t4x * t4y;
            bx = f4x * f4x - f4y * f4y;
            by = T(2) * f4x * f4y;
            prev = r4 * r2 * r;
        } else {
            ax = t2x * t4x - t2y * t4y;
            ay = t2x * t4y + t2y * t4x;
            bx = f2x * f4x - f2y * f4y;
            by = f2x * f4y + f2y * f4x;
            prev = r4 * r;
        }
        T s = prev * r;
        p = mk((ay * bx) * s, (ay * by) * s, ax * s);
        deriv = T(n) * prev;
        return;
    }
    T ax = T(1), ay = T(0), bx = T(1), by = T(0);
    uint32_t e = n;
    while (e != 0u) {
        if (e & 1u) {
            T nx = ax * tx - ay * ty, ny = ax * ty + ay * tx;
            ax = nx;
            ay = ny;
            nx = bx * fx - by * fy;
            ny = bx * fy + by * fx;
            bx = nx;
            by = ny;
        }
        e >>= 1;
        if (e != 0u) {
            T nx = tx * tx - ty * ty, ny = tx * ty + ty * tx;
            tx = nx;
            ty = ny;
            nx = fx * fx - fy * fy;
            ny = fx * fy + fy * fx;
            fx = nx;
            fy = ny;
        }
    }
    T prev = T(1);
    for (uint32_t i = 1; i < n; ++i) prev = prev * r;
    T s = prev * r;
    p = mk((ay * bx) * s, (ay * by) * s, ax * s);
    deriv = T(n) * prev;
}

// Trigonometric Power (eval.rs:462-475) for non-integer powers or tiny radii.
template <class T> __device__ __forceinline__ void trig_power(V3<T>& p, T radius, T power, T& dr) {
    T r = fmax(radius, T(1e-12));
    T theta = radius >= T(1e-12) ? atan2(sqrt(p.x * p.x + p.y * p.y), p.z) : acos(clampT(p.z / r, T(-1), T(1)));
    T phi = atan2(p.y, p.x);
    dr = dr * (pow(r, power - T(1)) * power);
    T zr = pow(r, power);
    T st = sin(theta * power), ct = cos(theta * power);
    T sp = sin(phi * power), cp = cos(phi * power);
    p = mk((st * cp) * zr, (st * sp) * zr, ct * zr);
}

// Integer power of an orbit point closer than 1e-12 to the origin.  The reference takes its
// trigonometric form there with the radius clamped to 1e-12 (eval.rs:462-475): the new point
// has length 1e-12^n and dr is multiplied by n * 1e-12^(n-1).  For n >= 4 the length
// underflows to zero in f32 and this is exact; for n = 2, 3 the point is set to the origin
// instead of a vector of length 1e-24 or 1e-36, which leaves the distance unchanged and moves
// a trap by less than 1e-23.  Keeping trigonometry out of the integer-power kernels keeps
// their code small.
template <class T> __device__ __forceinline__ void tiny_power(V3<T>& p, uint32_t n, T& dr) {
    T prev = T(1);
    for (uint32_t i = 1; i < n; ++i) prev = prev * T(1e-12);
    dr = dr * (prev * T(n));
    p = mk(T(0), T(0), T(0));
}

}  // namespace syf

#include "graph_device.hpp"

namespace syf {

// A node graph's field evaluation as the march and the shading use it: the distance, the
// normalised trap of the shading's slot, the iteration count clamped to the iteration product
// (eval.rs:98-114), and the Iterate-body runs.
template <class T, bool JIT> __device__ __forceinline__ De<T> graph_de(const FractalP& F, V3<T> c) {
    GRaw<T> g;
    if constexpr (JIT) g = graph_jit<T>(F.gprog, c);
    else g = graph_run<T>(F.gprog, c);
    const uint32_t slot = F.gshade_slot;
    const T t = slot == 0u ? g.r.t0 : slot == 1u ? g.r.t1 : slot == 2u ? g.r.t2 : g.r.t3;
    De<T> o;
    o.d = g.r.d;
    o.trap0 = isfinite(t) ? fmin(t / T(F.trap_scale), T(1)) : T(1);
    o.iter = g.r.iter < F.iteration_product ? g.r.iter : F.iteration_product;
    o.runs = g.runs;
    return o;
}

// Graph::evaluate for one classic family (eval.rs:98-114, :346-556, :558-599, :601-621).
template <class T, uint32_t FAM> __device__ __forceinline__ De<T> de_eval(const FractalP& F, V3<T> c) {
    if constexpr (FAM == FAM_GRAPH || FAM == FAM_JIT) return graph_de<T, FAM == FAM_JIT>(F, c);
    V3<T> z = c;
    T dr = T(1);
    T trap = T(INFINITY);
    uint32_t iter = 0;
    T d;
    const T bail = T(F.bailout);
    if constexpr (is_bulb(FAM)) {
        for (uint32_t k = 0; k < F.iterations; ++k) {
            T r = len3(z);
            if (r > bail) break;
            trap = fmin(trap, r);
            if constexpr (FAM != FAM_BULB_TRIG) {
                const uint32_t n = FAM == FAM_BULB8 ? 8u : F.ipower;
                if (r >= T(1e-12)) {
                    T deriv;
                    spherical_power(z, r, n, deriv);
                    dr = dr * deriv;
                } else {
                    tiny_power(z, n, dr);
                }
            } else {
                trig_power(z, r, T(F.power), dr);
            }
            z = add(z, c);
            dr = dr + T(1);
            ++iter;
        }
        T r = fmax(len3(z), T(1e-12));
        d = T(0.5) * log(r) * r / dr;
    } else if constexpr (FAM == SYF_FAMILY_MANDELBOX) {
        const T L = T(F.fold_limit), mr2 = T(F.mr2), fr2 = T(F.fr2), k_in = T(F.fr2_over_mr2);
        const T s = T(F.scale), as = T(F.abs_scale);
        for (uint32_t k = 0; k < F.iterations; ++k) {
            if (len3(z) > bail) break;
            z = mk(clampT(z.x, -L, L) * T(2) - z.x, clampT(z.y, -L, L) * T(2) - z.y, clampT(z.z, -L, L) * T(2) - z.z);
            T r2 = dot3(z, z);
            trap = fmin(trap, sqrt(r2));
            T kf = r2 < mr2 ? k_in : (r2 < fr2 ? fr2 / r2 : T(1));
            z = scl(z, kf);
            dr = dr * kf;
            z = scl(z, s);
            dr = dr * as;
            z = add(z, c);
            dr = dr + T(1);
            ++iter;
        }
        d = len3(z) / fabs(dr);
    } else if constexpr (FAM == SYF_FAMILY_MENGER) {
        const T s = T(F.scale), ox = T(F.off_s1[0]), oy = T(F.off_s1[1]), oz = T(F.off_s1[2]);
        const T thr = T(F.menger_shift_threshold);
        for (uint32_t k = 0; k < F.iterations; ++k) {
            if (len3(z) > bail) break;
            T wx = fabs(z.x), wy = fabs(z.y), wz = fabs(z.z), tmp;
            if (wx < wy) { tmp = wx; wx = wy; wy = tmp; }
            if (wx < wz) { tmp = wx; wx = wz; wz = tmp; }
            if (wy < wz) { tmp = wy; wy = wz; wz = tmp; }
            wx = wx * s - ox;
            wy = wy * s - oy;
            wz = wz * s - oz;
            if (wz < thr) wz = wz + oz;
            z = mk(wx, wy, wz);
            dr = dr * s;
            trap = fmin(trap, len3(z));
            ++iter;
        }
        T qx = fabs(z.x) - T(F.half[0]), qy = fabs(z.y) - T(F.half[1]), qz = fabs(z.z) - T(F.half[2]);
        V3<T> qm = mk(fmax(qx, T(0)), fmax(qy, T(0)), fmax(qz, T(0)));
        d = (len3(qm) + fmin(fmax(qx, fmax(qy, qz)), T(0))) / dr;
    } else {  // sphere graph: Trap -> Sphere U Plane
        T r = len3(z);
        trap = r;
        T a = (r - T(F.radius)) / dr;
        T b = (dot3(z, mk(T(F.plane_n[0]), T(F.plane_n[1]), T(F.plane_n[2]))) - T(F.plane_offset)) / dr;
        d = fmin(a, b);
    }
    De<T> o;
    o.d = d;
    o.trap0 = isfinite(trap) ? fmin(trap / T(F.trap_scale), T(1)) : T(1);
    o.iter = iter < F.iteration_product ? iter : F.iteration_product;
    o.runs = iter;
    return o;
}

// One field evaluation split into resumable body executions (FP32), for kernels that
// interleave evaluations of different depths on the lanes of a wave: start() seeds the point
// and reports whether the orbit is already done (|p| > bailout or no iterations); step() runs
// one Iterate body plus the next bailout test and reports completion; finish() applies the
// fence and the trap normalisation.  The arithmetic is de_eval's, in the same order, so the
// results are identical.
template <uint32_t FAM> struct DeStepper {
    V3<float> c, z;
    float dr, trap, r;  // r = |z| of the current orbit point
    uint32_t k;

    __device__ __forceinline__ bool start(const FractalP& F, V3<float> p) {
        c = p;
        z = p;
        dr = 1.0f;
        trap = INFINITY;
        k = 0;
        if constexpr (FAM == SYF_FAMILY_SPHERE) {
            r = len3(z);
            return true;
        } else {
            r = len3(z);
            return F.iterations == 0u || r > F.bailout;
        }
    }

    // One Iterate body (the caller guarantees the orbit is not done).
    __device__ __forceinline__ bool step(const FractalP& F) {
        if constexpr (is_bulb(FAM)) {
            trap = fminf(trap, r);
            if constexpr (FAM != FAM_BULB_TRIG) {
                const uint32_t n = FAM == FAM_BULB8 ? 8u : F.ipower;
                if (r >= 1e-12f) {
                    float deriv;
                    spherical_power(z, r, n, deriv);
                    dr = dr * deriv;
                } else {
                    tiny_power(z, n, dr);
                }
            } else {
                trig_power(z, r, F.power, dr);
            }
            z = add(z, c);
            dr = dr + 1.0f;
        } else if constexpr (FAM == SYF_FAMILY_MANDELBOX) {
            const float L = F.fold_limit;
            z = mk(clampT(z.x, -L, L) * 2.0f - z.x, clampT(z.y, -L, L) * 2.0f - z.y, clampT(z.z, -L, L) * 2.0f - z.z);
            const float r2 = dot3(z, z);
            trap = fminf(trap, sqrtf(r2));
            const float kf = r2 < F.mr2 ? F.fr2_over_mr2 : (r2 < F.fr2 ? F.fr2 / r2 : 1.0f);
            z = scl(z, kf);
            dr = dr * kf;
            z = scl(z, F.scale);
            dr = dr * F.abs_scale;
            z = add(z, c);
            dr = dr + 1.0f;
        } else if constexpr (FAM == SYF_FAMILY_MENGER) {
            float wx = fabsf(z.x), wy = fabsf(z.y), wz = fabsf(z.z), tmp;
            if (wx < wy) { tmp = wx; wx = wy; wy = tmp; }
            if (wx < wz) { tmp = wx; wx = wz; wz = tmp; }
            if (wy < wz) { tmp = wy; wy = wz; wz = tmp; }
            wx = wx * F.scale - F.off_s1[0];
            wy = wy * F.scale - F.off_s1[1];
            wz = wz * F.scale - F.off_s1[2];
            if (wz < F.menger_shift_threshold) wz = wz + F.off_s1[2];
            z = mk(wx, wy, wz);
            dr = dr * F.scale;
            trap = fminf(trap, len3(z));
        }
        ++k;
        r = len3(z);
        return k >= F.iterations || r > F.bailout;
    }

    __device__ __forceinline__ De<float> finish(const FractalP& F) const {
        float d;
        float tr = trap;
        if constexpr (is_bulb(FAM)) {
            const float rr = fmaxf(r, 1e-12f);
            d = 0.5f * logf(rr) * rr / dr;
        } else if constexpr (FAM == SYF_FAMILY_MANDELBOX) {
            d = r / fabsf(dr);
        } else if constexpr (FAM == SYF_FAMILY_MENGER) {
            const float qx = fabsf(z.x) - F.half[0], qy = fabsf(z.y) - F.half[1], qz = fabsf(z.z) - F.half[2];
            const V3<float> qm = mk(fmaxf(qx, 0.0f), fmaxf(qy, 0.0f), fmaxf(qz, 0.0f));
            d = (len3(qm) + fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f)) / dr;
        } else {
            tr = r;
            const float a = (r - F.radius) / dr;
            const float b = (dot3(z, mk(F.plane_n[0], F.plane_n[1], F.plane_n[2])) - F.plane_offset) / dr;
            d = fminf(a, b);
        }
        De<float> o;
        o.d = d;
        o.trap0 = isfinite(tr) ? fminf(tr / F.trap_scale, 1.0f) : 1.0f;
        o.iter = k < F.iteration_product ? k : F.iteration_product;
        o.runs = k;
        return o;
    }
};

// Hardware reciprocal, reciprocal square root and base-2 logarithm (v_rcp_f32, v_rsq_f32,
// v_log_f32: about 1 ulp, no range scaling) for the fast FP32 march.
__device__ __forceinline__ float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq_hw(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float ln_hw(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
// sqrt(x) for x >= 0 as x * rsq(x), 0 at 0
__device__ __forceinline__ float sqrt_hw(float x) { return x > 0.0f ? x * rsq_hw(x) : 0.0f; }

// The operations of the fast FP32 distance field: the hardware estimates above and FMA.
template <class T> struct FastOps;
template <> struct FastOps<float> {
    static constexpr float tiny = 1.17549435e-38f;  // least normal
    __device__ static __forceinline__ float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    __device__ static __forceinline__ float rcp(float x) { return rcp_hw(x); }
    __device__ static __forceinline__ float rsq(float x) { return rsq_hw(x); }
    __device__ static __forceinline__ float ln(float x) { return ln_hw(x); }
};

// The DeStepper interface with the fast arithmetic of the primary trace (SYF_MATH_FAST), for
// the FP32 march and normal taps: the same formulas in the same order of
// operations, but every quotient becomes a product with a reciprocal estimate, every square
// root comes from a reciprocal square root estimate, and products and sums contract into
// FMAs (FastOps).  The orbit keeps |z|^2 and tests it against bailout^2; the power body takes
// 1/r and 1/rho from one reciprocal square root each (no division); minimum traps keep the
// least |z|^2 and take its root once.  FP32 results differ from the exact stepper by a few ulps
// per operation, which the orbit amplifies: the march is held to the FP64 oracle's tolerance,
// not to the FP32 restatement step for step.
template <class T, uint32_t FAM> struct DeStepperFast {
    using O = FastOps<T>;
    V3<T> c, z;
    T dr, trap2, r2;  // r2 = |z|^2 of the current orbit point, trap2 = least r2 seen
    uint32_t k;

    __device__ __forceinline__ static T norm2(V3<T> v) { return O::fma(v.x, v.x, O::fma(v.y, v.y, v.z * v.z)); }
    __device__ __forceinline__ static T root(T x) { return x > T(0) ? x * O::rsq(x) : T(0); }

    __device__ __forceinline__ bool start(const FractalP& F, V3<T> p) {
        c = p;
        z = p;
        dr = T(1);
        trap2 = T(INFINITY);
        k = 0;
        r2 = norm2(p);
        if constexpr (FAM == SYF_FAMILY_SPHERE) return true;
        return F.iterations == 0u || r2 > T(F.bailout) * T(F.bailout);
    }

    // z^n + c for an integer power (integral_spherical_power, eval.rs:129-178), r >= 1e-12
    __device__ __forceinline__ void power_add(uint32_t n, T inv_r, T r) {
        const T rho2 = O::fma(z.x, z.x, z.y * z.y);
        const T inv_rho = O::rsq(fmax(rho2, O::tiny));
        const T rho = rho2 * inv_rho;
        T tx = fmin(fmax(z.z * inv_r, T(-1)), T(1)), ty = rho * inv_r;
        T fx = rho2 > T(0) ? z.x * inv_rho : T(1), fy = rho2 > T(0) ? z.y * inv_rho : T(0);
        T ax, ay, bx, by, prev;
        if (n == 8u || n == 6u) {
            const T t2x = O::fma(tx, tx, -(ty * ty)), t2y = (tx + tx) * ty;
            const T f2x = O::fma(fx, fx, -(fy * fy)), f2y = (fx + fx) * fy;
            const T t4x = O::fma(t2x, t2x, -(t2y * t2y)), t4y = (t2x + t2x) * t2y;
            const T f4x = O::fma(f2x, f2x, -(f2y * f2y)), f4y = (f2x + f2x) * f2y;
            const T rr2 = r * r, r4 = rr2 * rr2;
            if (n == 8u) {
                ax = O::fma(t4x, t4x, -(t4y * t4y));
                ay = (t4x + t4x) * t4y;
                bx = O::fma(f4x, f4x, -(f4y * f4y));
                by = (f4x + f4x) * f4y;
                prev = r4 * rr2 * r;
            } else {
                ax = O::fma(t2x, t4x, -(t2y * t4y));
                ay = O::fma(t2x, t4y, t2y * t4x);
                bx = O::fma(f2x, f4x, -(f2y * f4y));
                by = O::fma(f2x, f4y, f2y * f4x);
                prev = r4 * r;
            }
        } else {
            ax = T(1); ay = T(0); bx = T(1); by = T(0);
            uint32_t e = n;
            while (e != 0u) {
                if (e & 1u) {
                    T nx = O::fma(ax, tx, -(ay * ty)), ny = O::fma(ax, ty, ay * tx);
                    ax = nx; ay = ny;
                    nx = O::fma(bx, fx, -(by * fy));
                    ny = O::fma(bx, fy, by * fx);
                    bx = nx; by = ny;
                }
                e >>= 1;
                if (e != 0u) {
                    T nx = O::fma(tx, tx, -(ty * ty)), ny = (tx + tx) * ty;
                    tx = nx; ty = ny;
                    nx = O::fma(fx, fx, -(fy * fy));
                    ny = (fx + fx) * fy;
                    fx = nx; fy = ny;
                }
            }
            prev = T(1);
            for (uint32_t i = 1; i < n; ++i) prev = prev * r;
        }
        const T sc = prev * r;
        const T ays = ay * sc;
        z = mk(O::fma(ays, bx, c.x), O::fma(ays, by, c.y), O::fma(ax, sc, c.z));
        dr = O::fma(dr, T(n) * prev, T(1));
    }

    // One Iterate body (the caller guarantees the orbit is not done).
    __device__ __forceinline__ bool step(const FractalP& F) {
        if constexpr (is_bulb(FAM)) {
            trap2 = fmin(trap2, r2);
            const T inv_r = O::rsq(r2), r = r2 * inv_r;
            if constexpr (FAM != FAM_BULB_TRIG) {
                const uint32_t n = FAM == FAM_BULB8 ? 8u : F.ipower;
                if (r >= T(1e-12)) {
                    power_add(n, inv_r, r);
                } else {
                    tiny_power(z, n, dr);
                    z = add(z, c);
                    dr = dr + T(1);
                }
            } else {
                trig_power(z, r, T(F.power), dr);
                z = add(z, c);
                dr = dr + T(1);
            }
        } else if constexpr (FAM == SYF_FAMILY_MANDELBOX) {
            const T L = T(F.fold_limit);
            z = mk(O::fma(clampT(z.x, -L, L), T(2), -z.x), O::fma(clampT(z.y, -L, L), T(2), -z.y),
                   O::fma(clampT(z.z, -L, L), T(2), -z.z));
            const T q2 = norm2(z);
            trap2 = fmin(trap2, q2);
            const T kf = q2 < T(F.mr2) ? T(F.fr2_over_mr2) : (q2 < T(F.fr2) ? T(F.fr2) * O::rcp(q2) : T(1));
            const T ks = kf * T(F.scale);
            z = mk(O::fma(z.x, ks, c.x), O::fma(z.y, ks, c.y), O::fma(z.z, ks, c.z));
            dr = O::fma(dr * kf, T(F.abs_scale), T(1));
        } else if constexpr (FAM == SYF_FAMILY_MENGER) {
            T wx = fabs(z.x), wy = fabs(z.y), wz = fabs(z.z), tmp;
            if (wx < wy) { tmp = wx; wx = wy; wy = tmp; }
            if (wx < wz) { tmp = wx; wx = wz; wz = tmp; }
            if (wy < wz) { tmp = wy; wy = wz; wz = tmp; }
            wx = O::fma(wx, T(F.scale), -T(F.off_s1[0]));
            wy = O::fma(wy, T(F.scale), -T(F.off_s1[1]));
            wz = O::fma(wz, T(F.scale), -T(F.off_s1[2]));
            if (wz < T(F.menger_shift_threshold)) wz = wz + T(F.off_s1[2]);
            z = mk(wx, wy, wz);
            dr = dr * T(F.scale);
        }
        ++k;
        r2 = norm2(z);
        if constexpr (FAM == SYF_FAMILY_MENGER) trap2 = fmin(trap2, r2);
        return k >= F.iterations || r2 > T(F.bailout) * T(F.bailout);
    }

    __device__ __forceinline__ De<T> finish(const FractalP& F) const {
        T d;
        T tr = trap2 < T(INFINITY) ? root(trap2) : T(INFINITY);
        if constexpr (is_bulb(FAM)) {
            const T rr = fmax(root(r2), T(1e-12));
            d = T(0.5) * O::ln(rr) * rr * O::rcp(dr);
        } else if constexpr (FAM == SYF_FAMILY_MANDELBOX) {
            d = root(r2) * O::rcp(fabs(dr));
        } else if constexpr (FAM == SYF_FAMILY_MENGER) {
            const T qx = fabs(z.x) - T(F.half[0]), qy = fabs(z.y) - T(F.half[1]), qz = fabs(z.z) - T(F.half[2]);
            const V3<T> qm = mk(fmax(qx, T(0)), fmax(qy, T(0)), fmax(qz, T(0)));
            d = (root(norm2(qm)) + fmin(fmax(qx, fmax(qy, qz)), T(0))) * O::rcp(dr);
        } else {
            const T r = root(r2);
            tr = r;
            const T inv = O::rcp(dr);
            const T a = (r - T(F.radius)) * inv;
            const T b = (dot3(z, mk(T(F.plane_n[0]), T(F.plane_n[1]), T(F.plane_n[2]))) - T(F.plane_offset)) * inv;
            d = fmin(a, b);
        }
        De<T> o;
        o.d = d;
        o.trap0 = isfinite(tr) ? fmin(tr / T(F.trap_scale), T(1)) : T(1);
        o.iter = k < F.iteration_product ? k : F.iteration_product;
        o.runs = k;
        return o;
    }
};

// A node graph is evaluated whole by start() (the interpreter runs its loops itself); the
// fast arithmetic does not apply to it.
template <class T, bool JIT> struct GraphStepper {
    De<T> res;
    __device__ __forceinline__ bool start(const FractalP& F, V3<T> p) {
        res = graph_de<T, JIT>(F, p);
        return true;
    }
    __device__ __forceinline__ bool step(const FractalP&) { return true; }
    __device__ __forceinline__ De<T> finish(const FractalP&) const { return res; }
};
template <> struct DeStepper<FAM_GRAPH> : GraphStepper<float, false> {};
template <> struct DeStepper<FAM_JIT> : GraphStepper<float, true> {};
template <class T> struct DeStepperFast<T, FAM_GRAPH> : GraphStepper<T, false> {};
template <class T> struct DeStepperFast<T, FAM_JIT> : GraphStepper<T, true> {};

// One whole field evaluation with the fast arithmetic.
template <class T, uint32_t FAM> __device__ __forceinline__ De<T> de_eval_fast(const FractalP& F, V3<T> p) {
    DeStepperFast<T, FAM> s;
    if (!s.start(F, p)) {
        for (uint32_t k = 0; k < F.iterations; ++k)
            if (s.step(F)) break;
    }
    return s.finish(F);
}

// Family chosen at run time (uniform per job or per tile): used where several families
// share one launch.
template <class T> __device__ __forceinline__ De<T> de_eval_any(const FractalP& F, V3<T> p) {
    switch (F.family) {
    case SYF_FAMILY_MANDELBULB:
        if (F.ipower == 8u) return de_eval<T, FAM_BULB8>(F, p);
        return F.ipower != 0u ? de_eval<T, SYF_FAMILY_MANDELBULB>(F, p) : de_eval<T, FAM_BULB_TRIG>(F, p);
    case SYF_FAMILY_MANDELBOX: return de_eval<T, SYF_FAMILY_MANDELBOX>(F, p);
    case SYF_FAMILY_MENGER: return de_eval<T, SYF_FAMILY_MENGER>(F, p);
    case SYF_FAMILY_GRAPH: return graph_de<T, false>(F, p);
    default: return de_eval<T, SYF_FAMILY_SPHERE>(F, p);
    }
}

// ---------------------------------------------------------------- RNG (pathtrace.wgsl:20-37)

__device__ __host__ __forceinline__ uint32_t pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __host__ __forceinline__ uint32_t rng_init(uint32_t px, uint32_t py, uint32_t sample, uint32_t seed) {
    uint32_t h = pcg(seed);
    h = pcg(h ^ px);
    h = pcg(h ^ py);
    return pcg(h ^ sample);
}
__device__ __forceinline__ float rand01(uint32_t& s) {
    s = pcg(s);
    return float(s >> 8u) * (1.0f / 16777216.0f);
}

// camera_ray (pathtrace.wgsl:781-800).
__device__ __forceinline__ void camera_ray(const CameraP& C, uint32_t w, uint32_t h, uint32_t px, uint32_t py,
                                           uint32_t& rng, V3<float>& ro, V3<float>& rd) {
    float jx = rand01(rng);
    float jy = rand01(rng);
    float ndcx = ((float(px) + jx) / float(w)) * 2.0f - 1.0f;
    float ndcy = ((float(py) + jy) / float(h)) * 2.0f - 1.0f;
    V3<float> f = ld3(C.forward), r = ld3(C.right), u = ld3(C.up);
    V3<float> dir = nrm(add(add(f, scl(r, ndcx * C.tan_half)), scl(u, -ndcy * C.tan_half / C.aspect)));
    V3<float> origin = ld3(C.origin);
    if (C.lens_radius > 0.0f) {
        V3<float> focus = add(origin, scl(dir, C.focus / fmaxf(dot3(dir, f), 1e-4f)));
        float rr = C.lens_radius * sqrtf(rand01(rng));
        float phi = 2.0f * 3.14159265358979f * rand01(rng);
        origin = add(add(origin, scl(r, rr * cosf(phi))), scl(u, rr * sinf(phi)));
        dir = nrm(sub(focus, origin));
    }
    ro = origin;
    rd = dir;
}

// ---------------------------------------------------------------- march (pathtrace.wgsl:78-145)

struct Work {
    uint32_t de_evals = 0, de_iters = 0, steps = 0, de64_evals = 0, de64_iters = 0;
};

template <class T> struct Hit {
    bool hit, exhausted;
    T t;
    T t_prev;  // march position evaluated before the hit: lower end of the refine bracket
    T r_prev, r_hit;  // |d| * safety at t_prev and at t (the ends of the refine bracket)
    uint32_t steps;
    De<T> s;
};

// Sphere tracing with over-relaxation and fallback (Keinert et al. 2014); hit when the
// distance, relative to t, falls below the pixel cone times pixel_tolerance.
template <class T, uint32_t FAM>
__device__ __forceinline__ Hit<T> march(const FractalP& F, V3<T> ro, V3<T> rd, T pixel_cone, T pixel_tol,
                                        uint32_t max_steps, T t_end, Work& w) {
    const T safety = T(F.safety);
    const T hit_cone = pixel_cone * pixel_tol;
    T t = T(0), omega = T(1.2), prev_radius = T(0), step_len = T(0), cand_t = T(0), cand_err = T(1e30);
    T t_last = T(0);
    De<T> cand{};
    Hit<T> out;
    out.hit = false;
    out.exhausted = false;
    out.t = t_end;
    out.t_prev = T(0);
    out.r_prev = T(0);
    out.r_hit = T(0);
    out.steps = 0;
    out.s = De<T>{T(200), T(1), 0u, 0u};
    for (uint32_t i = 0; i < max_steps; ++i) {
        if (t >= t_end) return out;
        out.steps = i + 1u;
        De<T> s;
        if constexpr (FAM == 0xffffffffu) s = de_eval_any<T>(F, add(ro, scl(rd, t)));
        else s = de_eval<T, FAM>(F, add(ro, scl(rd, t)));
        w.de_evals++;
        w.de_iters += s.runs;
        w.steps++;
        T radius = fabs(s.d) * safety;
        bool sor_fail = omega > T(1) && (radius + prev_radius) < step_len * T(1.0001);
        if (sor_fail) {
            step_len = step_len - omega * step_len;
            omega = T(1);
        } else {
            step_len = radius * omega;
        }
        const T r_last = prev_radius;  // the radius at t_last
        prev_radius = radius;
        T err = radius / fmax(t, T(1e-6));
        if (!sor_fail && err < cand_err) {
            cand_t = t;
            cand = s;
            cand_err = err;
        }
        if (!sor_fail && err < hit_cone) {
            out.hit = true;
            out.t = t;
            out.t_prev = t_last;
            out.r_prev = r_last;
            out.r_hit = radius;
            out.s = s;
            return out;
        }
        if (!sor_fail && t + step_len >= t_end) {
            if (radius >= t_end - t) return out;
            omega = T(1);
            step_len = radius;
        }
        t_last = t;
        t = t + step_len;
    }
    if (cand_err < hit_cone * T(4)) {
        out.hit = true;
        out.t = cand_t;
        out.t_prev = cand_t;  // best-candidate hit: no bracket, the refine keeps t
        out.s = cand;
    } else {
        out.exhausted = true;
    }
    return out;
}

// normal_at (common.wgsl:376-387).
template <class T, uint32_t FAM, bool FAST = false>
__device__ __forceinline__ V3<T> normal_at(const FractalP& F, V3<T> p, T h, Work& w) {
    V3<T> g = mk(T(0), T(0), T(0));
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const T kx = (i == 0 || i == 3) ? T(1) : T(-1);
        const T ky = (i >= 2) ? T(1) : T(-1);
        const T kz = (i == 1 || i == 3) ? T(1) : T(-1);
        V3<T> k = mk(kx, ky, kz);
        De<T> s;
        if constexpr (FAM == 0xffffffffu) s = de_eval_any<T>(F, add(p, scl(k, h)));
        else if constexpr (FAST) s = de_eval_fast<T, FAM>(F, add(p, scl(k, h)));
        else s = de_eval<T, FAM>(F, add(p, scl(k, h)));
        w.de_evals++;
        w.de_iters += s.runs;
        g = add(g, scl(k, s.d));
    }
    if constexpr (FAST) return scl(g, FastOps<T>::rsq(dot3(g, g)));
    return nrm(g);
}

// Final-hit refine in FP64: the crossing of g(t) = |d(t)| * safety - hit_cone * t (the
// march's own pixel-footprint hit rule) between the last position before the hit (g > 0)
// and the hit (g <= 0), by Illinois regula falsi to a bracket of REFINE_REL_TOL * t (at most
// REFINE_ITERS evaluations); the bracket midpoint is returned.  g at the two ends is taken
// from the FP32 march, which evaluated the field there: g = radius - hit_cone * t with the
// march's radius |d| * safety.  Only the ends' signs and the first interpolation point depend
// on them; every interior point is evaluated in FP64.  Measured on a 480x270 Mandelbulb, this
// takes 6.2 FP64 evaluations per hit against 8.2 when both ends are evaluated again in FP64,
// and the results agree with a 1e-14 refine to the same 5e-10.  When the ends do not bracket
// a crossing (a preceding over-relaxation back-step, or a best-candidate hit after the step
// budget) the march t is kept.  The tolerance, 1e-9, is 60 times below an FP32 ulp and 1e5
// times below the 1e-4 the refined distance is held to; near the surface the FP64 field
// itself carries relative noise of about 1e-7 (ten orbit iterations amplify its rounding),
// and a tighter bracket (1e-12 measured: 9.5 evaluations per hit against 8.0, a 1080p frame
// 1.07 against 0.96 ms) chases that noise.
//
// Refine64 holds the refine between two evaluations of g, so that kernels can run it one
// evaluation at a time (begin, then g at x and update until it reports the result);
// refine64() runs it in one loop.  Both take the same steps in the same order.
constexpr int REFINE_ITERS = 40;
constexpr double REFINE_REL_TOL = 1e-9;

// g at a bracket end from the FP32 march's radius there.
__device__ __forceinline__ double refine_end_g(float radius, float t, double hit_cone) {
    return double(radius) - hit_cone * double(t);
}

struct Refine64 {
    double a, b, ga, gb, x;  // bracket, g at its ends, next point to evaluate
    int side;
    uint32_t it;             // interior evaluations done

    // Returns true when the refine needs no evaluation, with its result in out: no bracket
    // (t_prev >= t_hit, or g does not change sign between the ends: the march t), or a
    // bracket already within the tolerance.  Otherwise x is the first point to evaluate.
    __device__ __forceinline__ bool begin(double t_prev, double t_hit, double g_prev, double g_hit, double& out) {
        out = t_hit;
        if (!(t_prev < t_hit) || !(g_prev > 0.0 && g_hit <= 0.0)) return true;
        a = t_prev;
        b = t_hit;
        ga = g_prev;
        gb = g_hit;
        side = 0;
        it = 0;
        return next(out);
    }

    // Takes g(x); returns true with the refined t in out once the refine has ended.
    __device__ __forceinline__ bool update(double gx, double& out) {
        if (gx > 0.0) {
            a = x;
            ga = gx;
            if (side == 1) gb *= 0.5;
            side = 1;
        } else {
            b = x;
            gb = gx;
            if (side == -1) ga *= 0.5;
            side = -1;
        }
        ++it;
        return next(out);
    }

    __device__ __forceinline__ bool next(double& out) {
        if (it >= uint32_t(REFINE_ITERS) || b - a <= REFINE_REL_TOL * b) {
            out = 0.5 * (a + b);
            return true;
        }
        double c = b - gb * (b - a) / (gb - ga);
        if (!(c > a && c < b)) c = 0.5 * (a + b);
        x = c;
        return false;
    }
};

// g(t) of the refine at a point of the ray.
template <uint32_t FAM>
__device__ __forceinline__ double refine_g(const FractalP& F, V3<double> ro, V3<double> rd, double t, double hit_cone,
                                           Work& w) {
    De<double> s;
    if constexpr (FAM == 0xffffffffu) s = de_eval_any<double>(F, add(ro, scl(rd, t)));
    else s = de_eval<double, FAM>(F, add(ro, scl(rd, t)));
    w.de64_evals++;
    w.de64_iters += s.runs;
    return fabs(s.d) * double(F.safety) - hit_cone * t;
}

template <uint32_t FAM>
__device__ __forceinline__ double refine64(const FractalP& F, V3<double> ro, V3<double> rd, double t_prev,
                                           double t_hit, double g_prev, double g_hit, double hit_cone, Work& w) {
    Refine64 r;
    double out;
    if (r.begin(t_prev, t_hit, g_prev, g_hit, out)) return out;
    // at most REFINE_ITERS evaluations: update() ends the refine by then
    for (int e = 0; e < REFINE_ITERS; ++e)
        if (r.update(refine_g<FAM>(F, ro, rd, r.x, hit_cone, w), out)) break;
    return out;
}

// The march's radius |d| * safety at t along an FP32 ray: the value march() computed there
// (the same FP32 position and evaluation), for callers that kept only the march's t values.
template <uint32_t FAM>
__device__ __forceinline__ float march_radius(const FractalP& F, V3<float> ro, V3<float> rd, float t) {
    De<float> s;
    if constexpr (FAM == 0xffffffffu) s = de_eval_any<float>(F, add(ro, scl(rd, t)));
    else s = de_eval<float, FAM>(F, add(ro, scl(rd, t)));
    return fabsf(s.d) * F.safety;
}

// refine64 of an FP32 march's hit known by its t values only (t_prev < t_hit): the radii at
// the ends are evaluated again in FP32.
template <uint32_t FAM>
__device__ __forceinline__ double refine64_of_march(const FractalP& F, V3<float> ro, V3<float> rd, float t_prev,
                                                    float t_hit, double hit_cone, Work& w) {
    if (!(t_prev < t_hit)) return double(t_hit);
    const double g_prev = refine_end_g(march_radius<FAM>(F, ro, rd, t_prev), t_prev, hit_cone);
    const double g_hit = refine_end_g(march_radius<FAM>(F, ro, rd, t_hit), t_hit, hit_cone);
    return refine64<FAM>(F, cvt<double>(ro), cvt<double>(rd), double(t_prev), double(t_hit), g_prev, g_hit,
                         hit_cone, w);
}

// ---------------------------------------------------------------- wave-level helpers

// Sum over the 64 lanes of a wave (butterfly on the cross-lane permute network).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Adds a wave's work counters to the launch totals: one atomic per counter per wave.
__device__ __forceinline__ void flush_work(const Work& w, uint32_t hits, uint32_t rays, unsigned long long* ctr) {
    uint32_t a = wave_sum(w.de_evals), b = wave_sum(w.de_iters), c = wave_sum(w.steps);
    uint32_t d = wave_sum(w.de64_evals), e = wave_sum(w.de64_iters), f = wave_sum(hits), g = wave_sum(rays);
    if ((threadIdx.x & 63u) == 0u) {
        atomicAdd(&ctr[CTR_DE_EVALS], (unsigned long long)a);
        atomicAdd(&ctr[CTR_DE_ITERS], (unsigned long long)b);
        atomicAdd(&ctr[CTR_STEPS], (unsigned long long)c);
        atomicAdd(&ctr[CTR_DE64_EVALS], (unsigned long long)d);
        atomicAdd(&ctr[CTR_DE64_ITERS], (unsigned long long)e);
        atomicAdd(&ctr[CTR_HITS], (unsigned long long)f);
        atomicAdd(&ctr[CTR_RAYS], (unsigned long long)g);
    }
}

}  // namespace syf
