// Oracle: CPU restatement of SyFrack's sphere-traced distance-estimator path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg load this library, and only as the checker: the product path never links,
// loads or calls it.
//
// Every function below restates one piece of the reference (paths relative to the SyFrack
// source tree) in plain C++17, templated on the arithmetic type so that one body gives the
// FP32 restatement (the reference's precision: Rust f32 without contraction, glam vector
// algebra) and its FP64 twin.  Build with -ffp-contract=off so that no multiply-add is fused.
//
// Parity pins (what fixes each function independently of this file):
//   de_eval (Mandelbulb/Mandelbox/Menger/sphere)  -> the reference's 64-point fixtures
//       crates/syfrack-core/tests/fixtures/parity/*.ron (distance 1e-5, traps 1e-5,
//       iterations exact), the f64 spherical-power identity (eval.rs:180-222), the
//       Mandelbulb 2pi/7 z-rotation symmetry (fractal.rs:164-178), sphere/plane closed forms.
//   camera_basis                                  -> closed-form look-at basis checks.
//   pcg / rng_init / rand                         -> the integer hash written out by hand
//       in the tests on a few inputs (known answers) and its bijectivity on 2^20 inputs.
//   march / normal / refine                       -> analytic unit-sphere scenes (centre
//       depth, hit radius error shrinking with pixel_tolerance, unit normals; the KATs of
//       gpu_render.rs:113-145,479-490,696-726) and brute-force fine stepping in FP64.
//   planar_sample_f32 / escape_f64 / escape_f128  -> planar.rs:275-345 known answers,
//       c = 0 never escapes, |c| > 2 escapes at once, conjugate symmetry, agreement of
//       the FP64 and quad-precision loops where both are exact.
//   radiance (white furnace, Lambert under an area light)   -> closed forms.
// A function without such a pin would be marked "parity unpinned" here; none is.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <algorithm>
#include <limits>
#include <quadmath.h>

#include "../include/syfrack_hip.h"

namespace {

template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> neg(V3<T> a) { return {-a.x, -a.y, -a.z}; }
// glam Vec3::dot: (x*x') + (y*y') + (z*z'), left to right.
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
// glam Vec3::cross.
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
// glam normalize: v * (1 / |v|).
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / length(a)); }
template <class T> inline V3<T> vabs(V3<T> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
// Rust f32::min / max (NaN-ignoring; the traps start at +inf).
template <class T> inline T fmin_(T a, T b) { return std::fmin(a, b); }
template <class T> inline T fmax_(T a, T b) { return std::fmax(a, b); }
template <class T> inline T clampv(T v, T lo, T hi) { return fmin_(fmax_(v, lo), hi); }

template <class T> struct Vec2 {
    T x, y;
};

// ---------------------------------------------------------------- distance field

template <class T> struct DeOut {
    T d;
    T traps[4];
    uint32_t iter;
    uint32_t material;
    uint32_t body_runs;  // Iterate-body executions (work count, not a result)
};

// integral_spherical_power (eval.rs:129-178): z^n in spherical coordinates without
// inverse trigonometry; returns the power and the radial derivative n * r^(n-1).
template <class T> void spherical_power(V3<T> p, T r, uint32_t n, V3<T>* out, T* deriv) {
    if (n == 1) {
        *out = p;
        *deriv = T(1);
        return;
    }
    auto mul = [](Vec2<T> a, Vec2<T> b) { return Vec2<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; };
    T rho = std::sqrt(p.x * p.x + p.y * p.y);
    Vec2<T> theta{clampv(p.z / r, T(-1), T(1)), rho / r};
    Vec2<T> phi = rho > T(0) ? Vec2<T>{p.x / rho, p.y / rho} : Vec2<T>{T(1), T(0)};
    if (n == 6 || n == 8) {
        auto square = [](Vec2<T> a) { return Vec2<T>{a.x * a.x - a.y * a.y, T(2) * a.x * a.y}; };
        Vec2<T> t2 = square(theta), f2 = square(phi);
        Vec2<T> t4 = square(t2), f4 = square(f2);
        T r2 = r * r;
        T r4 = r2 * r2;
        Vec2<T> t, f;
        T previous;
        if (n == 6) {
            t = mul(t2, t4);
            f = mul(f2, f4);
            previous = r4 * r;
        } else {
            t = square(t4);
            f = square(f4);
            previous = r4 * r2 * r;
        }
        *out = v3(t.y * f.x, t.y * f.y, t.x) * (previous * r);
        *deriv = T(n) * previous;
        return;
    }
    Vec2<T> t{T(1), T(0)}, f{T(1), T(0)};
    uint32_t e = n;
    while (e != 0) {
        if (e & 1u) {
            t = mul(t, theta);
            f = mul(f, phi);
        }
        e >>= 1;
        if (e != 0) {
            theta = mul(theta, theta);
            phi = mul(phi, phi);
        }
    }
    T radial_previous = T(1);
    for (uint32_t i = 1; i < n; ++i) radial_previous = radial_previous * r;
    *out = v3(t.y * f.x, t.y * f.y, t.x) * (radial_previous * r);
    *deriv = T(n) * radial_previous;
}

template <class T> T tpow(T a, T b) { return std::pow(a, b); }

// Power node (eval.rs:456-476): integer powers 1..16 take the exact spherical product,
// any other power the trigonometric form.
template <class T> void power_node(V3<T>* p, T* dr, T power) {
    T radius = length(*p);
    if (radius >= T(1e-12) && power >= T(1) && power <= T(16) && power == std::floor(power)) {
        V3<T> q;
        T d;
        spherical_power(*p, radius, uint32_t(power), &q, &d);
        *p = q;
        *dr = *dr * d;
    } else {
        T r = fmax_(radius, T(1e-12));
        T theta = radius >= T(1e-12) ? std::atan2(std::sqrt(p->x * p->x + p->y * p->y), p->z)
                                      : std::acos(clampv(p->z / r, T(-1), T(1)));
        T phi = std::atan2(p->y, p->x);
        *dr = *dr * (tpow(r, power - T(1)) * power);
        T zr = tpow(r, power);
        T st = std::sin(theta * power), ct = std::cos(theta * power);
        T sp = std::sin(phi * power), cp = std::cos(phi * power);
        *p = v3(st * cp, st * sp, ct) * zr;
    }
}

template <class T> struct Fractal {
    uint32_t family, iterations;
    T bailout, power, scale, min_radius, fixed_radius, fold_limit;
    V3<T> offset, half, plane_n;
    T radius, plane_offset, trap_scale, safety, bound_hint;
    uint32_t iteration_product;
};

template <class T> Fractal<T> lower_fractal(const syf_fractal& f) {
    Fractal<T> r;
    r.family = f.family;
    r.iterations = f.iterations;
    r.bailout = T(f.bailout);
    r.power = T(f.power);
    r.scale = T(f.scale);
    r.min_radius = T(f.min_radius);
    r.fixed_radius = T(f.fixed_radius);
    r.fold_limit = T(f.fold_limit);
    r.offset = v3(T(f.offset[0]), T(f.offset[1]), T(f.offset[2]));
    r.half = v3(T(f.half[0]), T(f.half[1]), T(f.half[2]));
    // Plane normal: try_normalize().unwrap_or(Y), done in f32 as the reference.
    V3<float> n{f.plane_normal[0], f.plane_normal[1], f.plane_normal[2]};
    float ln = length(n), rcp = 1.0f / ln;
    V3<float> nn = (ln > 0.0f && rcp > 0.0f && std::isfinite(rcp)) ? n * rcp : V3<float>{0.0f, 1.0f, 0.0f};
    r.plane_n = v3(T(nn.x), T(nn.y), T(nn.z));
    r.radius = T(f.radius);
    r.plane_offset = T(f.plane_offset);
    r.trap_scale = T(f.trap_scale);
    r.safety = T(f.safety);  // step_scale = safety / 1 (no Twist/Bend: eval.rs:828-849)
    r.bound_hint = T(f.bounding_radius_hint);
    // iteration_product (validate.rs:112-137): the Iterate count; 1 without an Iterate.
    r.iteration_product = f.family == SYF_FAMILY_SPHERE ? 1u : f.iterations;
    return r;
}

// Graph::evaluate for the classic graphs (eval.rs:98-114 normalisation, :346-556 point
// ops in the bundled body order, :558-599 fences, :601-621 union).
template <class T> DeOut<T> de_eval(const Fractal<T>& F, V3<T> p) {
    const T INF = std::numeric_limits<T>::infinity();
    DeOut<T> o;
    T trap0 = INF;
    V3<T> z = p;
    T dr = T(1);
    uint32_t iter = 0;
    T d;
    switch (F.family) {
    case SYF_FAMILY_MANDELBULB: {
        for (uint32_t k = 0; k < F.iterations; ++k) {
            if (length(z) > F.bailout) break;
            trap0 = fmin_(trap0, length(z));          // Trap(Norm, Minimum)
            power_node(&z, &dr, F.power);             // Power
            z = z + p;                                // AddOrigin
            dr = dr + T(1);
            ++iter;
        }
        T r = fmax_(length(z), T(1e-12));              // BulbFence
        d = T(0.5) * std::log(r) * r / dr;
        break;
    }
    case SYF_FAMILY_MANDELBOX: {
        for (uint32_t k = 0; k < F.iterations; ++k) {
            if (length(z) > F.bailout) break;
            T L = F.fold_limit;                       // BoxFold: clamp(p)*2 - p
            z = v3(clampv(z.x, -L, L) * T(2) - z.x, clampv(z.y, -L, L) * T(2) - z.y,
                   clampv(z.z, -L, L) * T(2) - z.z);
            trap0 = fmin_(trap0, length(z));          // Trap
            T r2 = dot(z, z);                         // SphereFold
            T mr2 = F.min_radius * F.min_radius;
            T fr2 = F.fixed_radius * F.fixed_radius;
            T kf = r2 < mr2 ? fr2 / mr2 : (r2 < fr2 ? fr2 / r2 : T(1));
            z = z * kf;
            dr = dr * kf;
            z = z * F.scale;                          // Scale
            dr = dr * std::fabs(F.scale);
            z = z + p;                                // AddOrigin
            dr = dr + T(1);
            ++iter;
        }
        d = (length(z) - T(0)) / std::fabs(dr);       // NormFence(inset 0)
        break;
    }
    case SYF_FAMILY_MENGER: {
        for (uint32_t k = 0; k < F.iterations; ++k) {
            if (length(z) > F.bailout) break;
            V3<T> w = vabs(z);                        // MengerFold
            if (w.x < w.y) w = v3(w.y, w.x, w.z);
            if (w.x < w.z) w = v3(w.z, w.y, w.x);
            if (w.y < w.z) w = v3(w.x, w.z, w.y);
            T s = F.scale, s1 = F.scale - T(1);
            w = v3(w.x * s - F.offset.x * s1, w.y * s - F.offset.y * s1, w.z * s - F.offset.z * s1);
            if (w.z < T(-0.5) * F.offset.z * s1) w.z = w.z + F.offset.z * s1;
            z = w;
            dr = dr * s;
            trap0 = fmin_(trap0, length(z));          // Trap
            ++iter;
        }
        V3<T> q = vabs(z) - F.half;                   // Box fence
        V3<T> qm = v3(fmax_(q.x, T(0)), fmax_(q.y, T(0)), fmax_(q.z, T(0)));
        d = (length(qm) + fmin_(fmax_(q.x, fmax_(q.y, q.z)), T(0))) / dr;
        break;
    }
    default: {  // sphere graph: Trap -> Sphere  U  Plane
        trap0 = fmin_(trap0, length(z));
        T a = (length(z) - F.radius) / dr;
        T b = (dot(z, F.plane_n) - F.plane_offset) / dr;
        d = fmin_(a, b);
        break;
    }
    }
    o.d = d;
    o.traps[0] = std::isfinite(trap0) ? fmin_(trap0 / F.trap_scale, T(1)) : T(1);
    o.traps[1] = o.traps[2] = o.traps[3] = T(1);
    o.iter = std::min(iter, F.iteration_product);
    o.material = 0;
    o.body_runs = iter;
    return o;
}

// ---------------------------------------------------------------- RNG (pathtrace.wgsl:20-37)

inline uint32_t pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
inline uint32_t rng_init(uint32_t px, uint32_t py, uint32_t sample, uint32_t seed) {
    uint32_t h = pcg(seed);
    h = pcg(h ^ px);
    h = pcg(h ^ py);
    h = pcg(h ^ sample);
    return h;
}
inline float rand01(uint32_t* s) {
    *s = pcg(*s);
    return float(*s >> 8u) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------- camera (camera.rs:58-104)

struct Basis {
    V3<float> origin, forward, right, up;
    float tan_half, aspect, lens_radius, focus;
};

inline float to_rad(float deg) { return deg * (3.14159265358979323846f / 180.0f); }

Basis camera_basis(const syf_camera& c, uint32_t width, uint32_t height) {
    V3<float> origin, target;
    if (c.mode == SYF_CAMERA_ORBIT) {
        float az = to_rad(c.azimuth), el = to_rad(c.elevation);
        V3<float> off = v3(std::cos(el) * std::sin(az), std::sin(el), std::cos(el) * std::cos(az)) * c.distance;
        target = v3(c.point[0], c.point[1], c.point[2]);
        origin = target + off;
    } else {
        origin = v3(c.position[0], c.position[1], c.position[2]);
        target = v3(c.target[0], c.target[1], c.target[2]);
    }
    V3<float> f = target - origin;
    float rcp = 1.0f / length(f);  // normalize_or(NEG_Z)
    V3<float> forward = (std::isfinite(rcp) && rcp > 0.0f) ? f * rcp : v3(0.0f, 0.0f, -1.0f);
    V3<float> world_up = std::fabs(forward.y) > 0.9999f ? v3(0.0f, 0.0f, 1.0f) : v3(0.0f, 1.0f, 0.0f);
    V3<float> right0 = normalize(cross(forward, world_up));
    V3<float> up0 = cross(right0, forward);
    float rr = to_rad(c.roll);
    float s = std::sin(rr), co = std::cos(rr);
    Basis b;
    b.origin = origin;
    b.forward = forward;
    b.right = right0 * co + up0 * s;
    b.up = up0 * co - right0 * s;
    b.tan_half = 18.0f / c.focal_mm;
    b.aspect = float(width) / float(height);
    b.lens_radius = c.f_number > 0.0f ? (c.focal_mm / 1000.0f) / (2.0f * c.f_number) : 0.0f;
    b.focus = c.focus_distance;
    return b;
}

// camera_ray (pathtrace.wgsl:781-800), FP32.
void camera_ray(const Basis& b, uint32_t w, uint32_t h, uint32_t px, uint32_t py, uint32_t* rng,
                V3<float>* ro, V3<float>* rd) {
    float jx = rand01(rng);
    float jy = rand01(rng);
    float ndcx = ((float(px) + jx) / float(w)) * 2.0f - 1.0f;
    float ndcy = ((float(py) + jy) / float(h)) * 2.0f - 1.0f;
    V3<float> dir = normalize(b.forward + b.right * (ndcx * b.tan_half) + b.up * (-ndcy * b.tan_half / b.aspect));
    V3<float> origin = b.origin;
    if (b.lens_radius > 0.0f) {
        V3<float> focus = origin + dir * (b.focus / std::fmax(dot(dir, b.forward), 1e-4f));
        float r = b.lens_radius * std::sqrt(rand01(rng));
        float phi = 2.0f * 3.14159265358979f * rand01(rng);
        origin = origin + b.right * (r * std::cos(phi)) + b.up * (r * std::sin(phi));
        dir = normalize(focus - origin);
    }
    *ro = origin;
    *rd = dir;
}

// ---------------------------------------------------------------- march (pathtrace.wgsl:78-145)

template <class T> struct Hit {
    bool hit, exhausted, candidate;
    T t;
    T t_prev;  // march position evaluated before the hit (bracket of the FP64 refine)
    uint32_t steps;
    DeOut<T> sample;
};

struct Count {
    uint64_t de_evals = 0, de_iters = 0, steps = 0, de64_evals = 0, de64_iters = 0;
};

template <class T>
Hit<T> march(const Fractal<T>& F, V3<T> ro, V3<T> rd, T pixel_cone, T pixel_tol, uint32_t max_steps,
             T t_end, Count* cnt) {
    const T safety = F.safety;
    const T hit_cone = pixel_cone * pixel_tol;
    T t = 0, omega = T(1.2), prev_radius = 0, step_len = 0, cand_t = 0, cand_err = T(1e30);
    T t_last = 0;
    DeOut<T> cand{};
    Hit<T> out{};
    out.hit = false;
    out.exhausted = false;
    out.candidate = false;
    out.t = t_end;
    out.steps = 0;
    for (uint32_t i = 0; i < max_steps; ++i) {
        if (t >= t_end) return out;
        out.steps = i + 1;
        DeOut<T> s = de_eval(F, ro + rd * t);
        cnt->de_evals++;
        cnt->de_iters += s.body_runs;
        cnt->steps++;
        T radius = std::fabs(s.d) * safety;
        bool sor_fail = omega > T(1) && (radius + prev_radius) < step_len * T(1.0001);
        if (sor_fail) {
            step_len = step_len - omega * step_len;
            omega = T(1);
        } else {
            step_len = radius * omega;
        }
        prev_radius = radius;
        T err = radius / fmax_(t, T(1e-6));
        if (!sor_fail && err < cand_err) {
            cand_t = t;
            cand = s;
            cand_err = err;
        }
        if (!sor_fail && err < hit_cone) {
            out.hit = true;
            out.t = t;
            out.t_prev = t_last;
            out.sample = s;
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
        out.candidate = true;
        out.t = cand_t;
        out.t_prev = cand_t;
        out.sample = cand;
    } else {
        out.exhausted = true;
    }
    return out;
}

// normal_at (common.wgsl:376-387): tetrahedral gradient, four taps.
template <class T> V3<T> normal_at(const Fractal<T>& F, V3<T> p, T h, Count* cnt) {
    const V3<T> k[4] = {v3(T(1), T(-1), T(-1)), v3(T(-1), T(-1), T(1)), v3(T(-1), T(1), T(-1)),
                        v3(T(1), T(1), T(1))};
    V3<T> g = v3(T(0), T(0), T(0));
    for (int i = 0; i < 4; ++i) {
        DeOut<T> s = de_eval(F, p + k[i] * h);
        cnt->de_evals++;
        cnt->de_iters += s.body_runs;
        g = g + k[i] * s.d;
    }
    return normalize(g);
}

// Final-hit refine in FP64 (no reference counterpart: the reference keeps the t of the
// step that met its pixel-footprint hit rule, which depends on where the steps landed).
// The refine finds, in FP64, where along the ray that rule first holds: the crossing of
//     g(t) = |d(t)| * safety - hit_cone * t
// between the last march position before the hit (g > 0) and the hit (g <= 0), by Illinois
// regula falsi until the bracket is <= 1e-12 * t wide (at most 40 evaluations after the two
// end points); the result is the bracket midpoint.  When the end points do not bracket a
// crossing in FP64 (a preceding over-relaxation back-step, or a best-candidate hit after the
// step budget), the march t is returned unchanged.
constexpr int REFINE_ITERS = 40;
double refine64(const Fractal<double>& F, V3<double> ro, V3<double> rd, double t_prev, double t_hit,
                double hit_cone, Count* cnt) {
    auto g = [&](double t) {
        DeOut<double> s = de_eval(F, ro + rd * t);
        cnt->de64_evals++;
        cnt->de64_iters += s.body_runs;
        return std::fabs(s.d) * F.safety - hit_cone * t;
    };
    if (!(t_prev < t_hit)) return t_hit;
    double a = t_prev, b = t_hit;
    double ga = g(a), gb = g(b);
    if (!(ga > 0.0 && gb <= 0.0)) return t_hit;
    int side = 0;
    for (int i = 0; i < REFINE_ITERS; ++i) {
        if (b - a <= 1e-12 * b) break;
        double c = b - gb * (b - a) / (gb - ga);
        if (!(c > a && c < b)) c = 0.5 * (a + b);
        double gc = g(c);
        if (gc > 0.0) {
            a = c;
            ga = gc;
            if (side == 1) gb *= 0.5;
            side = 1;
        } else {
            b = c;
            gb = gc;
            if (side == -1) ga *= 0.5;
            side = -1;
        }
    }
    return 0.5 * (a + b);
}

float ray_extent(const syf_job_desc& d, const Basis& b) {
    // uniforms.rs:781-782 and common.wgsl:206-208
    float p0z = std::fmax(2.0f * d.fractal.bounding_radius_hint + length(b.origin), 200.0f);
    return std::fmax(200.0f, p0z);
}

}  // namespace

namespace {
template <class T> struct Escape {
    uint32_t count;
    bool escaped;
    T radius2;
};
// Textbook escape-time loop (z <- z^2 + c, Burning Ship on |Re|, |Im|): escape when
// |z|^2 > bailout^2, the squared form of planar.rs:111-115 so that FP64 is exact.
template <class T> Escape<T> escape_loop(uint32_t family, T cx, T cy, T zx, T zy, uint32_t iterations, T b2) {
    uint32_t count = 0;
    bool escaped = false;
    for (uint32_t i = 0; i < iterations; ++i) {
        T r2 = zx * zx + zy * zy;
        if (r2 > b2) {
            escaped = true;
            break;
        }
        T x = zx, y = zy;
        if (family == SYF_PLANAR_BURNING_SHIP) {
            x = x < T(0) ? -x : x;
            y = y < T(0) ? -y : y;
        }
        T nx = x * x - y * y + cx;
        T ny = T(2) * x * y + cy;
        zx = nx;
        zy = ny;
        ++count;
    }
    T r2 = zx * zx + zy * zy;
    escaped = escaped || r2 > b2;
    return Escape<T>{count, escaped, r2};
}
}  // namespace

// ================================================================ C entry points (ctypes)

extern "C" {

uint32_t orc_pcg(uint32_t v) { return pcg(v); }
uint32_t orc_rng_init(uint32_t px, uint32_t py, uint32_t sample, uint32_t seed) {
    return rng_init(px, py, sample, seed);
}
uint32_t orc_fold_seed(uint64_t seed) { return uint32_t(seed) ^ uint32_t(seed >> 32); }

// Basis as 16 floats: origin, forward, right, up (xyz each), tan_half, aspect, lens, focus.
void orc_camera_basis(const syf_job_desc* d, float* out16) {
    Basis b = camera_basis(d->camera, d->width, d->height);
    const V3<float> v[4] = {b.origin, b.forward, b.right, b.up};
    for (int i = 0; i < 4; ++i) {
        out16[3 * i] = v[i].x;
        out16[3 * i + 1] = v[i].y;
        out16[3 * i + 2] = v[i].z;
    }
    out16[12] = b.tan_half;
    out16[13] = b.aspect;
    out16[14] = b.lens_radius;
    out16[15] = b.focus;
}

// Camera rays of one sample for every pixel: ro3, rd3 (FP32) per pixel, row-major.
void orc_camera_rays(const syf_job_desc* d, uint32_t sample, float* ro3, float* rd3) {
    Basis b = camera_basis(d->camera, d->width, d->height);
    uint32_t seed = orc_fold_seed(d->seed);
    for (uint32_t y = 0; y < d->height; ++y)
        for (uint32_t x = 0; x < d->width; ++x) {
            uint32_t rng = rng_init(x, y, sample, seed);
            V3<float> ro, rd;
            camera_ray(b, d->width, d->height, x, y, &rng, &ro, &rd);
            size_t i = size_t(y) * d->width + x;
            ro3[3 * i] = ro.x; ro3[3 * i + 1] = ro.y; ro3[3 * i + 2] = ro.z;
            rd3[3 * i] = rd.x; rd3[3 * i + 1] = rd.y; rd3[3 * i + 2] = rd.z;
        }
}

// DE at n points (xyz4 layout).  precision 32 or 64.  Outputs per point: distance,
// traps[4], iterations, and the body executions.
int orc_de_probe(const syf_fractal* f, int precision, const float* xyz4, size_t n, double* dist,
                 double* traps4, uint32_t* iters, uint32_t* body_runs) {
    if (precision == 32) {
        Fractal<float> F = lower_fractal<float>(*f);
        for (size_t i = 0; i < n; ++i) {
            DeOut<float> o = de_eval(F, v3(xyz4[4 * i], xyz4[4 * i + 1], xyz4[4 * i + 2]));
            dist[i] = o.d;
            for (int k = 0; k < 4; ++k) traps4[4 * i + k] = o.traps[k];
            iters[i] = o.iter;
            if (body_runs) body_runs[i] = o.body_runs;
        }
    } else if (precision == 64) {
        Fractal<double> F = lower_fractal<double>(*f);
        for (size_t i = 0; i < n; ++i) {
            DeOut<double> o = de_eval(F, v3(double(xyz4[4 * i]), double(xyz4[4 * i + 1]), double(xyz4[4 * i + 2])));
            dist[i] = o.d;
            for (int k = 0; k < 4; ++k) traps4[4 * i + k] = o.traps[k];
            iters[i] = o.iter;
            if (body_runs) body_runs[i] = o.body_runs;
        }
    } else {
        return -1;
    }
    return 0;
}

// Integer spherical power at n points, for the f64 identity check (eval.rs:180-222).
void orc_spherical_power(const float* xyz4, size_t n, uint32_t power, float* out4) {
    for (size_t i = 0; i < n; ++i) {
        V3<float> p = v3(xyz4[4 * i], xyz4[4 * i + 1], xyz4[4 * i + 2]);
        V3<float> q;
        float d;
        spherical_power(p, length(p), power, &q, &d);
        out4[4 * i] = q.x; out4[4 * i + 1] = q.y; out4[4 * i + 2] = q.z; out4[4 * i + 3] = d;
    }
}

// Primary-ray trace of sample `sample` for the pixel rectangle [x0,x0+w) x [y0,y0+h)
// (global pixel coordinates), precision 32 (FP32 march + FP64 refine, the product's
// contract) or 64 (FP64 march of the same FP32 camera rays + FP64 refine: the oracle of
// the hit distance).  Outputs per rectangle pixel, row-major: depth (march t or -1),
// depth64 (refined t or -1), aux4 (normal, steps | 0,0,0,-1), steps.
// counts[5]: de_evals, de_iterations, steps, de64_evals, de64_iterations.
int orc_trace_primary(const syf_job_desc* d, uint32_t sample, int precision, uint32_t x0, uint32_t y0,
                      uint32_t w, uint32_t h, double* depth, double* depth64, float* aux4,
                      uint32_t* steps, uint64_t* counts) {
    if (precision != 32 && precision != 64) return -1;
    Basis b = camera_basis(d->camera, d->width, d->height);
    uint32_t seed = orc_fold_seed(d->seed);
    float pixel_cone = b.tan_half * 2.0f / float(d->width);
    float t_end = ray_extent(*d, b);
    Fractal<float> F32 = lower_fractal<float>(d->fractal);
    Fractal<double> F64 = lower_fractal<double>(d->fractal);
    Count total;
#pragma omp parallel
    {
        Count c;
#pragma omp for schedule(dynamic, 1)
        for (long yy = 0; yy < long(h); ++yy) {
            for (uint32_t xx = 0; xx < w; ++xx) {
                uint32_t px = x0 + xx, py = y0 + uint32_t(yy);
                uint32_t rng = rng_init(px, py, sample, seed);
                V3<float> ro, rd;
                camera_ray(b, d->width, d->height, px, py, &rng, &ro, &rd);
                size_t i = size_t(yy) * w + xx;
                V3<double> ro64 = v3(double(ro.x), double(ro.y), double(ro.z));
                V3<double> rd64 = v3(double(rd.x), double(rd.y), double(rd.z));
                bool hit;
                double t, t_prev = 0.0;
                uint32_t st;
                V3<double> n;
                if (precision == 32) {
                    Hit<float> hh = march(F32, ro, rd, pixel_cone, d->pixel_tolerance, d->max_steps, t_end, &c);
                    hit = hh.hit;
                    t = hh.t;
                    t_prev = hh.t_prev;
                    st = hh.steps;
                    if (hit) {
                        V3<float> p = ro + rd * hh.t;
                        V3<float> nf = normal_at(F32, p, std::fmax(hh.t * pixel_cone * 0.5f, 1e-4f) * d->pixel_tolerance, &c);
                        n = v3(double(nf.x), double(nf.y), double(nf.z));
                    }
                } else {
                    Hit<double> hh = march(F64, ro64, rd64, double(pixel_cone), double(d->pixel_tolerance),
                                           d->max_steps, double(t_end), &c);
                    hit = hh.hit;
                    t = hh.t;
                    t_prev = hh.t_prev;
                    st = hh.steps;
                    if (hit) {
                        V3<double> p = ro64 + rd64 * hh.t;
                        n = normal_at(F64, p, std::fmax(hh.t * double(pixel_cone) * 0.5, 1e-4) * double(d->pixel_tolerance), &c);
                    }
                }
                steps[i] = st;
                if (hit) {
                    depth[i] = t;
                    depth64[i] = refine64(F64, ro64, rd64, t_prev, t, double(pixel_cone) * double(d->pixel_tolerance), &c);
                    aux4[4 * i] = float(n.x);
                    aux4[4 * i + 1] = float(n.y);
                    aux4[4 * i + 2] = float(n.z);
                    aux4[4 * i + 3] = float(st);
                } else {
                    depth[i] = -1.0;
                    depth64[i] = -1.0;
                    aux4[4 * i] = aux4[4 * i + 1] = aux4[4 * i + 2] = 0.0f;
                    aux4[4 * i + 3] = -1.0f;
                }
            }
        }
#pragma omp critical
        {
            total.de_evals += c.de_evals;
            total.de_iters += c.de_iters;
            total.steps += c.steps;
            total.de64_evals += c.de64_evals;
            total.de64_iters += c.de64_iters;
        }
    }
    if (counts) {
        counts[0] = total.de_evals;
        counts[1] = total.de_iters;
        counts[2] = total.steps;
        counts[3] = total.de64_evals;
        counts[4] = total.de64_iters;
    }
    return 0;
}

// Dense FP64 stepping along one ray (step `dt`) to the first sign change of the field,
// bisected to 1e-12: the brute-force pin of the march + refine on small images.
double orc_first_root_bruteforce(const syf_fractal* f, const double* ro3, const double* rd3, double t_max,
                                 double dt) {
    Fractal<double> F = lower_fractal<double>(*f);
    V3<double> ro = v3(ro3[0], ro3[1], ro3[2]), rd = v3(rd3[0], rd3[1], rd3[2]);
    double prev_t = 0.0;
    double prev_d = de_eval(F, ro).d;
    if (prev_d <= 0.0) return 0.0;
    for (double t = dt; t <= t_max; t += dt) {
        double dd = de_eval(F, ro + rd * t).d;
        if (dd <= 0.0) {
            double lo = prev_t, hi = t;
            for (int i = 0; i < 80 && hi - lo > 1e-13; ++i) {
                double mid = 0.5 * (lo + hi);
                if (de_eval(F, ro + rd * mid).d > 0.0) lo = mid; else hi = mid;
            }
            return 0.5 * (lo + hi);
        }
        prev_t = t;
        prev_d = dd;
    }
    (void)prev_d;
    return -1.0;
}


// ---------------------------------------------------------------- 2D escape time

// PlanarFractal::sample (planar.rs:77-194) in f32 for the polynomial families (Julia,
// Mandelbrot, Burning Ship; twist 0, no inversion), input already mapped to the plane.
// out: distance, traps[4]; returns the iteration count.
uint32_t orc_planar_sample_f32(uint32_t family, uint32_t power, uint32_t iterations, const float* julia2,
                               float bailout, float trap_radius, float stripes, float in_x, float in_y,
                               float* out5) {
    auto cmul = [](Vec2<float> a, Vec2<float> b) { return Vec2<float>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; };
    auto cpow = [&](Vec2<float> z, uint32_t p) {
        Vec2<float> r{1.0f, 0.0f};
        for (uint32_t i = 0; i < p; ++i) r = cmul(r, z);
        return r;
    };
    bool parameter_plane = family == SYF_PLANAR_MANDELBROT || family == SYF_PLANAR_BURNING_SHIP;
    Vec2<float> point{in_x, in_y};
    Vec2<float> c = parameter_plane ? point : Vec2<float>{julia2[0], julia2[1]};
    Vec2<float> z = parameter_plane ? Vec2<float>{0.0f, 0.0f} : point;
    Vec2<float> rotation{1.0f, 0.0f};  // twist 0: (cos 0, sin 0)
    float ring = 1e20f, stripe = 0.0f;
    uint32_t count = 0, measured = 0;
    float smooth = float(iterations);
    bool escaped = false;
    for (uint32_t i = 0; i < iterations; ++i) {
        float radius = std::sqrt(z.x * z.x + z.y * z.y);
        if (radius > bailout) {
            escaped = true;
            break;
        }
        ring = std::fmin(ring, std::fabs(radius - trap_radius));
        float angle = radius == 0.0f ? 0.0f : std::atan2(z.y, z.x);
        stripe += 0.5f + 0.5f * std::sin(stripes * angle);
        measured += 1;
        Vec2<float> base = family == SYF_PLANAR_BURNING_SHIP ? Vec2<float>{std::fabs(z.x), std::fabs(z.y)} : z;
        Vec2<float> pw = cmul(rotation, cpow(base, power));
        z = Vec2<float>{pw.x + c.x, pw.y + c.y};
        count += 1;
    }
    float radius = std::fmin(std::sqrt(z.x * z.x + z.y * z.y), 1e20f);
    escaped = escaped || radius > bailout;
    float lp = std::log(float(power));
    if (escaped) smooth = float(count) + 1.0f - std::log(std::log(std::fmax(radius, 1.0001f))) / lp;
    float final_angle = radius == 0.0f ? 0.5f : std::atan2(z.y, z.x) / 6.28318530717958647692f + 0.5f;
    float bands = stripe / float(std::max(measured, 1u));
    if (escaped && measured > 0) {
        float angle = radius == 0.0f ? 0.0f : std::atan2(z.y, z.x);
        float last = 0.5f + 0.5f * std::sin(stripes * angle);
        float full = (stripe + last) / float(measured + 1);
        float t = 1.0f + (std::log(std::log(bailout)) - std::log(std::log(std::fmax(radius, 1.0001f)))) / lp;
        t = std::fmin(std::fmax(t, 0.0f), 1.0f);
        bands += (full - bands) * t;
    }
    out5[0] = radius / (1.0f + radius);
    out5[1] = std::fmin(std::fmax(smooth / float(iterations), 0.0f), 1.0f);
    out5[2] = ring / (1.0f + ring);
    out5[3] = bands;
    out5[4] = final_angle;
    return count;
}

// Pixel offsets of the 2D view, in double: fx[i] for columns, fy[j] for rows (the
// product computes the same tables on its host side).
void orc_planar_offsets(const syf_planar_desc* d, double* fx, double* fy) {
    for (uint32_t i = 0; i < d->width; ++i)
        fx[i] = ((double(i) + 0.5) / double(d->width) - 0.5) * d->span;
    double vspan = d->span * double(d->height) / double(d->width);
    for (uint32_t j = 0; j < d->height; ++j)
        fy[j] = (0.5 - (double(j) + 0.5) / double(d->height)) * vspan;
}


// FP64 direct evaluation of the pixel rectangle [x0,x0+w) x [y0,y0+h): iteration counts
// and smooth counts (count + 1 - ln ln r / ln 2, planar.rs:162-164).
void orc_escape_f64(const syf_planar_desc* d, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t* iters,
                    float* smooth) {
    double* fx = new double[d->width];
    double* fy = new double[d->height];
    orc_planar_offsets(d, fx, fy);
    double rot = double(d->rotation) * (3.14159265358979323846 / 180.0);
    double cr = std::cos(rot), sr = std::sin(rot);
    double b2 = double(d->bailout) * double(d->bailout);
#pragma omp parallel for schedule(dynamic, 1)
    for (long jj = 0; jj < long(h); ++jj)
        for (uint32_t ii = 0; ii < w; ++ii) {
            uint32_t i = x0 + ii, j = y0 + uint32_t(jj);
            double dx = cr * fx[i] - sr * fy[j];
            double dy = sr * fx[i] + cr * fy[j];
            double px = d->center[0] + dx, py = d->center[1] + dy;
            bool pp = d->family != SYF_PLANAR_JULIA;
            Escape<double> e = escape_loop<double>(d->family, pp ? px : d->julia[0], pp ? py : d->julia[1],
                                                   pp ? 0.0 : px, pp ? 0.0 : py, d->iterations, b2);
            size_t k = size_t(jj) * w + ii;
            iters[k] = e.count;
            double r = std::fmin(std::sqrt(e.radius2), 1e20);
            smooth[k] = e.escaped ? float(double(e.count) + 1.0 - std::log(std::log(std::fmax(r, 1.0001))) / std::log(2.0))
                                  : float(d->iterations);
        }
    delete[] fx;
    delete[] fy;
}

// Quad-precision (binary128) direct evaluation at listed pixels, the deep-zoom pin of the
// perturbation mode: the pixel's c is centre + (float) offset, exactly the point whose
// orbit the FP32 deltas follow.
void orc_escape_f128(const syf_planar_desc* d, const uint32_t* px, const uint32_t* py, size_t n, uint32_t* iters,
                     float* smooth) {
    double* fx = new double[d->width];
    double* fy = new double[d->height];
    orc_planar_offsets(d, fx, fy);
    double rot = double(d->rotation) * (3.14159265358979323846 / 180.0);
    double cr = std::cos(rot), sr = std::sin(rot);
    __float128 b2 = (__float128)d->bailout * (__float128)d->bailout;
#pragma omp parallel for schedule(dynamic, 1)
    for (long k = 0; k < long(n); ++k) {
        double dx = cr * fx[px[k]] - sr * fy[py[k]];
        double dy = sr * fx[px[k]] + cr * fy[py[k]];
        __float128 cx = (__float128)d->center[0] + (__float128)(float)dx;
        __float128 cy = (__float128)d->center[1] + (__float128)(float)dy;
        Escape<__float128> e = escape_loop<__float128>(d->family, cx, cy, (__float128)0, (__float128)0, d->iterations, b2);
        iters[k] = e.count;
        double r = std::fmin(std::sqrt((double)e.radius2), 1e20);
        smooth[k] = e.escaped ? float(double(e.count) + 1.0 - std::log(std::log(std::fmax(r, 1.0001))) / std::log(2.0))
                              : float(d->iterations);
    }
    delete[] fx;
    delete[] fy;
}

}  // extern "C"
