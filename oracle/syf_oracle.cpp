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
//   reference_orbit_dd / perturb64 (deep zoom; no reference counterpart)
//       -> the double-double orbit equals the binary128 orbit rounded to FP64 where a plain
//       FP64 orbit drifts within two steps; PERTURB64 within 1 of binary128 at the exact
//       pixel points of 1e-12 views (Burning Ship's chaotic bands included) on every point
//       whose count survives a 64-ulp move of the offset, where the FP64 direct loop fails;
//       within 1 of FP64 direct on a whole-set view (tests/test_planar.py).
//   trig_power (non-integer Mandelbulb powers)    -> equals the integer spherical product
//       at n = 2..12 in FP32 and FP64, |z^n| = r^n, and continuity in the power
//       (tests/test_oracle_radiance.py).
//   radiance (NEE, BSDF continuation, MIS, light selection, VNDF, environment, emitters)
//       -> tests/test_oracle_radiance.py: the white and metallic furnaces and the rim
//       (gpu_render.rs:456-475,572-608); a Lambert plane under a gradient environment against
//       rho (H + 0.8 (Z - H)); a Lambert plane under a small emitter against rho/pi * E, with E
//       by quadrature over the emitter, NEE alone and with MIS; two identical emitters against
//       one of double intensity (gpu_render.rs:397-421); NEE-only against MIS + VNDF on a rough
//       metal plane, seen along and off the normal; surface-variation endpoints
//       (gpu_render.rs:282-306); cavity darkness monotone (gpu_render.rs:840-868).  Planted
//       mistakes (a dropped cos_l, an unweighted emitter hit, a VNDF without the alpha stretch,
//       a pdf without G1(v), a linear environment blend, a pdf without 1/pi, a missing
//       selection probability, swapped sin/cos) each fail at least one of them.
//   suns, flat ground, colouring ramp, finishes (cone_sample, sun_pdf, sun_radiance,
//       ground_intersection, ramp_sample, make_surface, Oren-Nayar in brdf_eval)
//       -> tests/test_oracle_shading.py: the ramp against palette.rs:225-260's known answers;
//       a Lambert ground under a sun against rho/pi * E * cos(theta) * (1 + cos r) / 2, NEE alone
//       and with MIS on disks up to 90 degrees; emission added once; every colouring slot on the
//       ground, whose sample is known; the orbit mix against OrbitColoring::value
//       (material.rs:39-53); the Oren-Nayar lobe against the qualitative model in its paper form
//       (sin alpha tan beta); the ground occluding an emitter below it.  Planted mistakes (sun
//       radiance without its pdf, max -> min in the Oren-Nayar denominator, a non-uniform cone
//       sample, a wrong normal slot value) each fail at least one of them.
//   height fog (fog_sigma, fog_segment_tau, fog_optical_depth, fog_invert_segment, fog_free_flight,
//       volume_coordinate, hg_phase, hg_sample, fog_segment_direct, fog_direct_light, the volume
//       branches of radiance and the session planes) -> tests/test_oracle_fog.py: the rotated
//       radical inverse; optical depth against quadrature of the height profile; free flight against
//       1 - exp(-tau) (Kolmogorov-Smirnov); the phase's normalisation, mean cosine and histogram; an
//       absorbing fog against exp(-tau) of the sky; the volume white furnace (a non-absorbing fog in a
//       uniform sky keeps its radiance); single scattering of an emitter against a quadrature; an
//       emitter counted once after a scattering event (radiance polynomial in the albedo); VOLUME +
//       SURFACE = ACCUM.  Planted mistakes (the distance pdf without the equiangular average, a
//       surface proposal for phase-sampled escapes, a sign in the phase sample, swapped Halton bases,
//       a homogeneous inversion of a height fog, a phase hit of an emitter counted again) each fail
//       at least one of them.
//   finish detail patterns (lattice, lattice_noise, lattice_fbm, finish_relief, the detail block
//       of make_surface, the ray footprint of radiance) -> tests/test_oracle_finish.py: the slope
//       against a finite difference of the height for patterns 1-7, 9 and 10; every band's fade to
//       its closed-form mean at a huge footprint; sand's ripple slope, weathered concrete's rain
//       slope and wet concrete's pool ripples in footprint ranges where their parts separate; the
//       lattice's range, mean and seeding.  Planted mistakes (the warp's Jacobian factor of dust, the
//       rosette's angular term, the pore slope's frequency, the rain slope's weight, an octave slope
//       without its frequency) each fail at least one of them.  Parity unpinned: the grain weights of
//       sand (2.5) and wet concrete (16), which add a shading slope no height follows.
// Every other function has such a pin.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <algorithm>
#include <limits>
#include <vector>
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

// The trigonometric branch of the Power node (eval.rs:462-475): polar angle from
// atan2(|xy|, z) (acos of the clamped cosine at the origin), azimuth atan2(y, x), both
// multiplied by the power; radius r^n and radial derivative n * r^(n-1).
template <class T> void trig_power(V3<T>* p, T* dr, T power) {
    T radius = length(*p);
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
        trig_power(p, dr, power);
    }
}

#include "graph_eval.hpp"

// The node graph that a job of family SYF_FAMILY_GRAPH evaluates (set by orc_graph_use).
OGraph g_graph;

template <class T> struct Fractal {
    const OGraph* graph;  // non-null for SYF_FAMILY_GRAPH
    uint32_t family, iterations;
    T bailout, power, scale, min_radius, fixed_radius, fold_limit;
    V3<T> offset, half, plane_n;
    T radius, plane_offset, trap_scale, safety, bound_hint;
    uint32_t iteration_product;
};

template <class T> Fractal<T> lower_fractal(const syf_fractal& f) {
    Fractal<T> r;
    r.graph = nullptr;
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
    if (f.family == SYF_FAMILY_GRAPH) {
        // GpuFractal of a graph: step scale, bounding hint and iteration product of the graph
        r.graph = &g_graph;
        r.safety = T(step_scale(g_graph));
        r.bound_hint = T(g_graph.hint);
        r.iteration_product = g_graph.iteration_product;
    }
    return r;
}

// Graph.bounding_radius_hint of the job's field (uniforms.rs:781-782)
inline float bound_hint_of(const syf_fractal& f) {
    return f.family == SYF_FAMILY_GRAPH ? g_graph.hint : f.bounding_radius_hint;
}

// Graph::evaluate for the classic graphs (eval.rs:98-114 normalisation, :346-556 point
// ops in the bundled body order, :558-599 fences, :601-621 union).
template <class T> DeOut<T> de_eval(const Fractal<T>& F, V3<T> p) {
    if (F.graph) return graph_evaluate(*F.graph, p);
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
    T r_prev, r_hit;  // |d| * safety at t_prev and at t: g at the bracket's ends
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
        const T r_last = prev_radius;  // the radius at t_last
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
            out.r_prev = r_last;
            out.r_hit = radius;
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
// regula falsi until the bracket is <= 1e-9 * t wide (at most 40 evaluations); the result is
// the bracket midpoint.  g at the two ends comes from the march, which evaluated the field
// there (g_prev, g_hit: the march's radius minus hit_cone * t); every interior point is
// evaluated in FP64.  When the ends do not bracket a crossing (a preceding over-relaxation
// back-step, or a best-candidate hit after the step budget), the march t is returned
// unchanged.
constexpr int REFINE_ITERS = 40;
double refine64(const Fractal<double>& F, V3<double> ro, V3<double> rd, double t_prev, double t_hit,
                double g_prev, double g_hit, double hit_cone, Count* cnt) {
    auto g = [&](double t) {
        DeOut<double> s = de_eval(F, ro + rd * t);
        cnt->de64_evals++;
        cnt->de64_iters += s.body_runs;
        return std::fabs(s.d) * F.safety - hit_cone * t;
    };
    if (!(t_prev < t_hit)) return t_hit;
    double a = t_prev, b = t_hit;
    double ga = g_prev, gb = g_hit;
    if (!(ga > 0.0 && gb <= 0.0)) return t_hit;
    int side = 0;
    for (int i = 0; i < REFINE_ITERS; ++i) {
        if (b - a <= 1e-9 * b) break;
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

// g at a bracket end from a march's radius there.
template <class T> double end_g(T radius, T t, double hit_cone) { return double(radius) - hit_cone * double(t); }

// refine64 of an FP32 march's hit known by its t values only: the radii at the ends are
// evaluated again in FP32 (the same positions and evaluations as the march's; not counted).
double refine64_of_march(const Fractal<float>& F32, const Fractal<double>& F64, V3<float> ro, V3<float> rd,
                         float t_prev, float t_hit, double hit_cone, Count* cnt) {
    if (!(t_prev < t_hit)) return double(t_hit);
    const float r_prev = std::fabs(de_eval(F32, ro + rd * t_prev).d) * F32.safety;
    const float r_hit = std::fabs(de_eval(F32, ro + rd * t_hit).d) * F32.safety;
    return refine64(F64, v3(double(ro.x), double(ro.y), double(ro.z)), v3(double(rd.x), double(rd.y), double(rd.z)),
                    double(t_prev), double(t_hit), end_g(r_prev, t_prev, hit_cone), end_g(r_hit, t_hit, hit_cone),
                    hit_cone, cnt);
}

// ---------------------------------------------------------------- path tracing (FP32)
//
// The radiance estimator of pathtrace.wgsl:802-1201 restricted to the scene subset of this
// path: area and environment lights, one base material with an optional trap-driven surface
// variation (Lambert + GGX metal), no fog, sun, HDRI, portal, ocean, optics, texture,
// guiding or phenomena (each of those branches is inert when its feature is absent).

constexpr float PI_F = 3.14159265358979f;
constexpr uint32_t NO_LIGHT = 0xffffffffu;
typedef V3<float> F3;

struct LightL {
    uint32_t kind;
    F3 a, b, c, color, horizon, ground, n;
    float area;
    float cos_max;  // sun: cosine of the disk's angular radius
};

struct SceneL {
    Fractal<float> F;
    Basis cam;
    uint32_t width, height, seed, max_bounces, max_steps;
    float pixel_tol, pixel_cone, t_end, cavity;
    LightL lights[SYF_MAX_LIGHTS];
    uint32_t light_count;
    syf_material mat;
    syf_coloring coloring;  // stop positions resolved (Ramp::position)
    uint32_t color_slot;    // the coloring's slot, SYF_COLOR_NONE without stops
    uint32_t finish_count;
    syf_finish finishes[SYF_MAX_FINISHES];
    syf_ground ground;
    syf_fog fog;  // density 0: none
};

inline F3 f3(const float* p) { return v3(p[0], p[1], p[2]); }
inline float luminance(F3 c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }
// WGSL mix(a, b, t) = a * (1 - t) + b * t
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline F3 mix3(F3 a, F3 b, float t) { return v3(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)); }
inline F3 mul3(F3 a, F3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline F3 reflect3(F3 e1, F3 e2) { return e1 - e2 * (2.0f * dot(e2, e1)); }
// common.wgsl:242-247
inline F3 safe_normalize(F3 v) {
    float l = length(v);
    float rcp = 1.0f / l;
    if (!(l > 0.0f) || !(rcp > 0.0f) || !(rcp <= 3.4028235e38f)) return v3(0.0f, 1.0f, 0.0f);
    return v * rcp;
}
inline float smoothstep01(float x) {
    float t = clampv((x - 0.0f) / (1.0f - 0.0f), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

SceneL lower_scene(const syf_job_desc& d) {
    SceneL S;
    S.F = lower_fractal<float>(d.fractal);
    S.cam = camera_basis(d.camera, d.width, d.height);
    S.width = d.width;
    S.height = d.height;
    S.seed = uint32_t(d.seed) ^ uint32_t(d.seed >> 32);
    S.max_bounces = d.max_bounces;
    S.max_steps = d.max_steps;
    S.pixel_tol = d.pixel_tolerance;
    S.pixel_cone = S.cam.tan_half * 2.0f / float(d.width);
    S.cavity = d.cavity_darkness;
    float p0z = std::fmax(2.0f * bound_hint_of(d.fractal) + length(S.cam.origin), 200.0f);
    S.t_end = std::fmax(200.0f, p0z);
    S.light_count = d.light_count;
    for (uint32_t i = 0; i < d.light_count; ++i) {
        const syf_light& l = d.lights[i];
        LightL& L = S.lights[i];
        L = LightL{};
        L.kind = l.kind;
        if (l.kind == SYF_LIGHT_AREA) {  // GpuLight::from_light (uniforms.rs:676-716)
            F3 n0 = f3(l.normal);
            float rcp = 1.0f / length(n0);
            F3 n = (std::isfinite(rcp) && rcp > 0.0f) ? n0 * rcp : v3(0.0f, 1.0f, 0.0f);
            F3 helper = std::fabs(n.y) > 0.999f ? v3(0.0f, 0.0f, 1.0f) : v3(0.0f, 1.0f, 0.0f);
            F3 u = normalize(cross(helper, n));
            F3 vv = cross(n, u);
            L.a = f3(l.position);
            L.b = u * l.size[0];
            L.c = vv * l.size[1];
            L.n = safe_normalize(cross(L.b, L.c));  // area_normal (lights.wgsl:164-166)
            L.area = length(cross(L.b, L.c));       // emitter_area (lights.wgsl:69-74)
            L.color = f3(l.color) * l.intensity;
        } else if (l.kind == SYF_LIGHT_SUN) {  // Light::Sun (uniforms.rs:656-668)
            F3 n0 = f3(l.normal);
            float rcp = 1.0f / length(n0);
            L.a = (std::isfinite(rcp) && rcp > 0.0f) ? n0 * rcp : v3(0.0f, 1.0f, 0.0f);
            L.cos_max = std::cos((l.size[0] * 0.5f) * (3.14159265358979323846f / 180.0f));
            L.color = f3(l.color) * l.intensity;
        } else {
            L.color = f3(l.zenith) * l.intensity;  // env: a = zenith, b = horizon, c = ground
            L.horizon = f3(l.horizon) * l.intensity;
            L.ground = f3(l.ground) * l.intensity;
        }
    }
    S.mat = d.material;
    // GpuColoring / GpuRamp (uniforms.rs:31-55, 194-214): positions from Ramp::position
    S.coloring = d.coloring;
    const uint32_t ns = d.coloring.stop_count;
    for (uint32_t i = 0; i < ns; ++i)
        if (!d.coloring.positioned)
            S.coloring.stops[i][3] = float(i) / float(d.coloring.cyclic ? ns : (ns > 1u ? ns - 1u : 1u));
    S.color_slot = ns > 0u ? d.coloring.slot : SYF_COLOR_NONE;
    S.finish_count = d.finish_count;
    for (uint32_t i = 0; i < SYF_MAX_FINISHES; ++i) S.finishes[i] = d.finishes[i];
    S.ground = d.ground;
    S.fog = d.fog;
    if (!(S.fog.density > 0.0f)) S.fog = syf_fog{};
    return S;
}

struct Surf {
    F3 n, diffuse, f0;
    float alpha;
    F3 emission;
    float sigma;  // diffuse_sigma: Oren-Nayar roughness, 0 = Lambert
};

// Branchless orthonormal basis, columns t, bt, n (pathtrace.wgsl:194-201).
inline void tangent_frame(F3 n, F3* t, F3* bt) {
    float s = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n.z);
    float b = n.x * n.y * a;
    *t = v3(1.0f + s * n.x * n.x * a, s * b, -s * n.x);
    *bt = v3(b, s + n.y * n.y * a, -n.y);
}
inline F3 frame_mul(F3 t, F3 bt, F3 n, F3 x) { return t * x.x + bt * x.y + n * x.z; }
// pathtrace.wgsl:203-208
F3 cosine_sample(F3 n, float r1, float r2) {
    float phi = 2.0f * PI_F * r1;
    float r = std::sqrt(r2);
    F3 t, bt;
    tangent_frame(n, &t, &bt);
    return frame_mul(t, bt, n, v3(r * std::cos(phi), r * std::sin(phi), std::sqrt(std::fmax(0.0f, 1.0f - r2))));
}
inline float ggx_d(float nh, float a2) {
    float d = nh * nh * (a2 - 1.0f) + 1.0f;
    return a2 / (PI_F * d * d);
}
inline float smith_g1(float nv, float a2) { return 2.0f * nv / (nv + std::sqrt(a2 + (1.0f - a2) * nv * nv)); }
inline F3 fresnel(F3 f0, float vh) {
    float k = std::pow(1.0f - vh, 5.0f);
    return v3(f0.x + (1.0f - f0.x) * k, f0.y + (1.0f - f0.y) * k, f0.z + (1.0f - f0.z) * k);
}
// Heitz 2018 VNDF sampling (pathtrace.wgsl:234-249).
F3 sample_ggx_h(F3 n, F3 v, float alpha, float r1, float r2) {
    F3 t, bt;
    tangent_frame(n, &t, &bt);
    F3 lv = v3(dot(t, v), dot(bt, v), dot(n, v));
    F3 vh = safe_normalize(v3(alpha * lv.x, alpha * lv.y, std::fmax(lv.z, 1e-5f)));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    F3 t1 = v3(1.0f, 0.0f, 0.0f);
    if (lensq > 0.0f) {
        float sl = std::sqrt(lensq);
        t1 = v3(-vh.y / sl, vh.x / sl, 0.0f);
    }
    F3 t2 = cross(vh, t1);
    float r = std::sqrt(r1);
    float phi = 2.0f * PI_F * r2;
    float x = r * std::cos(phi);
    float blend = 0.5f * (1.0f + vh.z);
    float y = mixf(std::sqrt(std::fmax(0.0f, 1.0f - x * x)), r * std::sin(phi), blend);
    F3 nh = t1 * x + t2 * y + vh * std::sqrt(std::fmax(0.0f, 1.0f - x * x - y * y));
    return frame_mul(t, bt, n, safe_normalize(v3(alpha * nh.x, alpha * nh.y, std::fmax(nh.z, 0.0f))));
}
// palette_mix / palette_sample (common.wgsl:331-362)
F3 palette_mix(const syf_coloring& r, uint32_t a, uint32_t b, float amount) {
    float f = clampv(amount, 0.0f, 1.0f);
    if (r.smooth) f = f * f * (3.0f - 2.0f * f);
    return mix3(f3(r.stops[a]), f3(r.stops[b]), f);
}
F3 ramp_sample(const syf_coloring& r, float value) {
    const uint32_t n = r.stop_count;
    if (n == 0u) return v3(1.0f, 1.0f, 1.0f);
    if (n == 1u) return f3(r.stops[0]);
    float t = value * r.frequency + r.phase;
    if (r.cyclic) t = t - std::floor(t);
    else t = clampv(t, 0.0f, 1.0f);
    if (r.gamma != 1.0f) t = std::pow(t, r.gamma);
    if (!r.positioned && !r.cyclic) {
        const float x = t * float(n - 1u);
        const uint32_t i = std::min(uint32_t(std::floor(x)), n - 2u);
        return palette_mix(r, i, i + 1u, x - float(i));
    }
    const float first = r.stops[0][3], last = r.stops[n - 1u][3];
    if (r.cyclic && (t < first || t >= last)) {
        const float x = t < first ? t + 1.0f : t;
        return palette_mix(r, n - 1u, 0u, (x - last) / std::fmax(1.0f + first - last, 1e-8f));
    }
    if (t <= first) return f3(r.stops[0]);
    if (t >= last) return f3(r.stops[n - 1u]);
    for (uint32_t i = 0; i < 15u; ++i)
        if (i + 1u < n && t < r.stops[i + 1u][3])
            return palette_mix(r, i, i + 1u, (t - r.stops[i][3]) / (r.stops[i + 1u][3] - r.stops[i][3]));
    return f3(r.stops[n - 1u]);
}
// surface_value (pathtrace.wgsl:265-269)
float surface_value(const SceneL& S, uint32_t slot, F3 n, const DeOut<float>& smp) {
    if (slot < 4u) return smp.traps[slot];
    if (slot == SYF_COLOR_NORMAL) return n.y * 0.5f + 0.5f;
    return float(smp.iter) / std::fmax(float(S.F.iteration_product), 1.0f);
}
// ---------------------------------------------------------------- finish detail patterns
// finish_detail.wgsl: per pattern, a height h in [0, 1] and its slope (the analytic gradient of
// h in world space for patterns 1-7, 9 and 10; a shading slope for 8, 11 and 12).
struct Relief {
    F3 g;
    float h;
};
inline float sstep(float a, float b, float x) {
    const float t = clampv((x - a) / (b - a), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
inline float sstep_d(float a, float b, float x) {  // finish_smooth_derivative
    const float t = clampv((x - a) / (b - a), 0.0f, 1.0f);
    return 6.0f * t * (1.0f - t) / (b - a);
}
inline float sgn(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
inline F3 hadamard(F3 a, float x, float y, float z) { return v3(a.x * x, a.y * y, a.z * z); }

// finish_hash (finish_detail.wgsl:3-7)
float lattice(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    const uint32_t h = pcg(uint32_t(x) ^ pcg(uint32_t(y) + 0x9e3779b9u) ^ pcg(uint32_t(z) + seed));
    return float(h >> 8) / 16777216.0f;
}
// finish_noise (finish_detail.wgsl:8-20): smoothstep-weighted trilinear lattice noise
Relief lattice_noise(F3 p, uint32_t seed) {
    const float px = std::floor(p.x), py = std::floor(p.y), pz = std::floor(p.z);
    const int32_t i = int32_t(px), j = int32_t(py), k = int32_t(pz);
    const float fx = p.x - px, fy = p.y - py, fz = p.z - pz;
    const float wx = fx * fx * (3.0f - 2.0f * fx), wy = fy * fy * (3.0f - 2.0f * fy), wz = fz * fz * (3.0f - 2.0f * fz);
    float c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(i + a, j + b, k + d, seed);
    const float y00 = mixf(c[0][0][0], c[1][0][0], wx), y10 = mixf(c[0][1][0], c[1][1][0], wx);
    const float y01 = mixf(c[0][0][1], c[1][0][1], wx), y11 = mixf(c[0][1][1], c[1][1][1], wx);
    const float z0 = mixf(y00, y10, wy), z1 = mixf(y01, y11, wy);
    const float ddx = mixf(mixf(c[1][0][0] - c[0][0][0], c[1][1][0] - c[0][1][0], wy),
                           mixf(c[1][0][1] - c[0][0][1], c[1][1][1] - c[0][1][1], wy), wz);
    const float ddy = mixf(y10 - y00, y11 - y01, wz);
    const float ddz = z1 - z0;
    Relief r;
    r.g = v3(ddx * 6.0f * fx * (1.0f - fx), ddy * 6.0f * fy * (1.0f - fy), ddz * 6.0f * fz * (1.0f - fz));
    r.h = mixf(z0, z1, wz);
    return r;
}
// finish_fbm (finish_detail.wgsl:21-33)
Relief lattice_fbm(F3 p, float width, uint32_t seed) {
    Relief out{v3(0.0f, 0.0f, 0.0f), 0.0f};
    float freq = 1.0f, amp = 0.5f;
    for (uint32_t o = 0; o < 3; ++o) {
        const float vis = 1.0f - sstep(0.35f, 0.85f, width * freq);
        if (vis > 0.0f) {
            const Relief n = lattice_noise(p * freq, seed + o * 97u);
            out.g = out.g + n.g * (amp * freq * vis);
            out.h += amp * mixf(0.5f, n.h, vis);
        } else {
            out.h += amp * 0.5f;
        }
        freq *= 2.13f;
        amp *= 0.5f;
    }
    out.g = v3(out.g.x / 0.875f, out.g.y / 0.875f, out.g.z / 0.875f);
    out.h /= 0.875f;
    return out;
}
// finish_eval (finish_detail.wgsl:38-151) without dunes (wind 0) or ocean
Relief finish_relief(uint32_t kind, uint32_t seed, float scale, F3 p, float footprint) {
    const F3 q = p * scale;
    const float width = footprint * scale;
    switch (kind) {
    case 12u: {
        const Relief basin = lattice_fbm(hadamard(q, 0.31f, 0.31f, 0.22f), width * 0.31f, seed + 911u);
        const float wet = sstep(0.47f, 0.59f, basin.h);
        const Relief grit = lattice_fbm(q * 16.0f, width * 16.0f, seed + 113u);
        const float vis = 1.0f - sstep(0.15f, 0.6f, width * 3.0f);
        const float ph = q.x * 2.8f + q.z * 4.1f;
        const F3 ripple = v3(2.8f, 0.0f, 4.1f) * (std::cos(ph) * 0.018f * vis);
        return Relief{mix3(grit.g * 16.0f, ripple, wet) * scale, wet};
    }
    case 10u:
    case 11u: {
        const Relief stains = lattice_fbm(q * 0.14f, width * 0.14f, seed + 17u);
        const Relief aggregate = lattice_fbm(q * 7.0f, width * 7.0f, seed + 113u);
        const Relief pore = lattice_noise(q * 23.0f, seed + 307u);
        const float vis = 1.0f - sstep(0.25f, 0.75f, width * 23.0f);
        const float pit = 1.0f - sstep(0.24f, 0.43f, pore.h);
        const F3 pit_g = pore.g * (-sstep_d(0.24f, 0.43f, pore.h) * 23.0f * vis);
        const float ph = (q.y / 1.25f + 0.5f) - std::floor(q.y / 1.25f + 0.5f) - 0.5f;
        const float board = 1.0f - sstep(0.03f, 0.15f, width);
        const float joint = (1.0f - sstep(0.006f, 0.027f, std::fabs(ph))) * board;
        const float joint_g = -sstep_d(0.006f, 0.027f, std::fabs(ph)) * sgn(ph) / 1.25f * board;
        const float height = 0.18f * stains.h + 0.65f * aggregate.h - 0.14f * mixf(0.2f, pit, vis) - 0.09f * joint;
        const F3 g = (stains.g * (0.18f * 0.14f) + aggregate.g * (0.65f * 7.0f) - pit_g * 0.14f -
                      v3(0.0f, 0.09f * joint_g, 0.0f)) * scale;
        if (kind == 11u) {
            const Relief rain = lattice_fbm(hadamard(q, 1.9f, 0.055f, 1.9f), width * 1.9f, seed + 601u);
            const float deposit = sstep(0.33f, 0.68f, rain.h);
            const F3 rain_g = hadamard(rain.g, 1.9f, 0.055f, 1.9f) * (scale * sstep_d(0.33f, 0.68f, rain.h));
            return Relief{g + rain_g * 0.12f, clampv(0.72f * deposit + 0.28f * stains.h, 0.0f, 1.0f)};
        }
        return Relief{g, clampv(height + 0.12f, 0.0f, 1.0f)};
    }
    case 9u: {
        const Relief wa = lattice_fbm(q * 0.17f, width * 0.17f, seed + 17u);
        const Relief wb = lattice_fbm(q * 0.17f, width * 0.17f, seed + 53u);
        const F3 warped = q + v3(2.8f * (wa.h - 0.5f), 0.0f, 2.8f * (wb.h - 0.5f));
        const Relief cloud = lattice_fbm(warped * 0.16f, width * 0.48f, seed + 211u);
        const Relief streak = lattice_fbm(hadamard(warped, 0.65f, 0.18f, 0.09f), width * 1.95f, seed + 397u);
        const float density = 0.78f * cloud.h + 0.22f * streak.h;
        const F3 g = cloud.g * (0.78f * 0.16f) + hadamard(streak.g, 0.65f, 0.18f, 0.09f) * 0.22f;
        const F3 dg = (g + (wa.g * g.x + wb.g * g.z) * (2.8f * 0.17f)) * scale;
        const float deposit = sstep(0.36f, 0.64f, density);
        const Relief grain = lattice_fbm(p * 48.0f, footprint * 48.0f, seed + 739u);
        const float powder = 0.85f + 0.15f * grain.h;
        return Relief{dg * (sstep_d(0.36f, 0.64f, density) * powder) + grain.g * (deposit * 0.15f * 48.0f),
                      deposit * powder};
    }
    case 8u: {
        const float x = p.x * scale, z = p.z * scale;
        const Relief warp = lattice_noise(v3(x * 0.045f, 0.0f, z * 0.045f), seed);
        const float ph = 2.0f * PI_F * (x + 1.7f * warp.h + 0.12f * std::sin(z * 0.22f));
        const float vis = std::exp(-((width * 3.0f) * (width * 3.0f)));
        const float ripple = 0.5f + 0.36f * std::sin(ph) + 0.10f * std::sin(2.0f * ph);
        const float a = 2.0f * PI_F * (0.36f * std::cos(ph) + 0.20f * std::cos(2.0f * ph)) * scale * vis;
        const Relief grain = lattice_fbm(p * 90.0f, footprint * 90.0f, seed + 139u);
        return Relief{v3(a * (1.0f + 0.0765f * warp.g.x), 0.0f, a * (0.0264f * std::cos(z * 0.22f) + 0.0765f * warp.g.z)) +
                          grain.g * 2.5f,
                      0.70f * mixf(0.5f, ripple, vis) + 0.30f * grain.h};
    }
    case 6u: {
        const float u = (q.x + 0.5f) - std::floor(q.x + 0.5f) - 0.5f, w = (q.z + 0.5f) - std::floor(q.z + 0.5f) - 0.5f;
        const float r = std::fmax(std::sqrt(u * u + w * w), 1e-5f);
        const float ang = std::atan2(w, u);
        const float flower = r - 0.25f - 0.07f * std::cos(ang * 8.0f);
        const float ring = r - 0.45f;
        float gu = sgn(flower) * (u / r + 0.56f * std::sin(ang * 8.0f) * (-w / (r * r)));
        float gw = sgn(flower) * (w / r + 0.56f * std::sin(ang * 8.0f) * (u / (r * r)));
        float dist = std::fabs(flower);
        if (std::fabs(ring) < dist) {
            dist = std::fabs(ring);
            gu = sgn(ring) * u / r;
            gw = sgn(ring) * w / r;
        }
        const float vis = 1.0f - sstep(0.10f, 0.40f, width);
        const float carving = 1.0f - sstep(0.012f, 0.045f, dist);
        const float k = -sstep_d(0.012f, 0.045f, dist) * vis * scale;
        return Relief{v3(gu * k, 0.0f, gw * k), mixf(0.2f, carving, vis)};
    }
    case 3u: {
        const Relief n = lattice_fbm(hadamard(q, 18.0f, 0.45f, 3.0f), width * 18.0f, seed);
        return Relief{hadamard(n.g, 18.0f, 0.45f, 3.0f) * scale, n.h};
    }
    default:
        break;
    }
    const Relief n = lattice_fbm(q, width, seed);
    Relief r = n;
    if (kind == 1u) {
        const F3 axis = v3(0.16f, 0.11f, 0.08f);
        const float ph = (dot(q, axis) + n.h * 2.8f) * PI_F * 2.0f;
        const float wave = std::sin(ph);
        const float veins = 1.0f - sstep(0.035f, 0.19f, std::fabs(wave));
        const F3 vg = (axis + n.g * 2.8f) * (-sstep_d(0.035f, 0.19f, std::fabs(wave)) * sgn(wave) * std::cos(ph) * PI_F * 2.0f);
        const float vis = 1.0f - sstep(0.18f, 0.45f, width);
        r.g = mix3(n.g, n.g * 0.65f + vg * 0.35f, vis);
        r.h = mixf(n.h, 0.65f * n.h + 0.35f * veins, vis);
    } else if (kind == 2u) {
        r.g = n.g * sstep_d(0.35f, 0.67f, n.h);
        r.h = sstep(0.35f, 0.67f, n.h);
    } else if (kind == 4u || kind == 7u) {
        const Relief grain = lattice_fbm(q * 3.7f, width * 3.7f, seed + 53u);
        r.g = n.g * 0.8f + grain.g * (0.2f * 3.7f);
        r.h = 0.8f * n.h + 0.2f * grain.h;
    } else if (kind == 5u) {
        const float a = q.x + 0.75f * n.h, b = q.z * 1.37f + q.x * 0.4f;
        const float vis = 1.0f - sstep(0.35f, 1.0f, width);
        r.g = ((v3(1.0f, 0.0f, 0.0f) + n.g * 0.75f) * (0.32f * std::cos(a)) + v3(0.4f, 0.0f, 1.37f) * (0.18f * std::cos(b))) * vis;
        r.h = 0.5f + (0.32f * std::sin(a) + 0.18f * std::sin(b)) * vis;
    }
    return Relief{r.g * scale, r.h};
}

// make_surface (pathtrace.wgsl:271-386) for the subset: colouring ramp, named finishes
// (emission; Oren-Nayar for the concrete patterns, sigma = 0.45 roughness), SurfaceVariation.
Surf make_surface(const SceneL& S, F3 p, F3 n, const DeOut<float>& smp, float footprint) {
    const syf_material& m = S.mat;
    F3 albedo = f3(m.albedo);
    const uint32_t slot = S.color_slot;
    const bool assigned = smp.material > 0u && smp.material <= S.finish_count;
    const float iters = float(smp.iter) / std::fmax(float(S.F.iteration_product), 1.0f);
    if (!assigned && slot == SYF_COLOR_ORBIT) {
        const syf_coloring& c = S.coloring;
        float value = c.weights[0] * smp.traps[0] + c.weights[1] * smp.traps[1] + c.weights[2] * smp.traps[2] +
                      c.weights[3] * smp.traps[3] + c.iteration_weight * iters + c.normal_weight * (n.y * 0.5f + 0.5f);
        const float gain = c.logarithmic;
        if (gain > 0.0f) {
            const float sg = value > 0.0f ? 1.0f : (value < 0.0f ? -1.0f : 0.0f);
            value = sg * std::log(1.0f + gain * std::fabs(value)) / std::log(1.0f + gain);
        }
        albedo = mul3(albedo, ramp_sample(c, value));
    } else if (!assigned && slot < 4u) {
        albedo = mul3(albedo, ramp_sample(S.coloring, smp.traps[slot]));
    } else if (!assigned && slot == SYF_COLOR_NORMAL) {
        albedo = mul3(albedo, ramp_sample(S.coloring, n.y * 0.5f + 0.5f));
    } else if (!assigned && slot == SYF_COLOR_ITER) {
        albedo = mul3(albedo, ramp_sample(S.coloring, iters));
    }
    float rough = m.roughness, metallic = m.metallic, spec = m.specular, sigma_scale = 0.0f;
    F3 emission = v3(0.0f, 0.0f, 0.0f);
    if (assigned) {
        const syf_finish& f = S.finishes[smp.material - 1u];
        albedo = f3(f.albedo);
        rough = f.roughness;
        metallic = f.metallic;
        spec = f.specular;
        emission = f3(f.emission);
        sigma_scale = (f.oren_nayar || f.pattern >= 10u) ? 0.45f : 0.0f;
    }
    if (!assigned && m.surface_enabled) {
        const float t = clampv(surface_value(S, m.surface_slot, n, smp), 0.0f, 1.0f);
        rough = m.surface_roughness[0] + (m.surface_roughness[1] - m.surface_roughness[0]) * t;
        metallic = m.surface_metallic[0] + (m.surface_metallic[1] - m.surface_metallic[0]) * t;
    }
    F3 shading_n = n;
    if (assigned && S.finishes[smp.material - 1u].pattern != 0u) {  // pathtrace.wgsl:306-329
        const syf_finish& f = S.finishes[smp.material - 1u];
        Relief d = finish_relief(f.pattern, f.seed, f.scale, p, footprint);
        if (f.pattern == 9u) {
            const float up = sstep(0.10f, 0.80f, n.y);
            d.g = d.g * up;
            d.h *= up;
        }
        if (f.pattern == 12u) d.h *= sstep(0.45f, 0.95f, n.y);
        const float amount = f.strength * d.h;
        albedo = mix3(albedo, f3(f.tint), amount);
        if (f.pattern == 9u) {
            rough += std::fmax(f.roughness_delta, 0.0f) * amount;
            metallic *= 1.0f - amount;
            spec = mixf(spec, 0.025f, amount);
        } else if (f.pattern == 12u) {
            rough = mixf(rough, clampv(rough + f.roughness_delta, 0.025f, 1.0f), amount);
            spec = mixf(spec, 0.0204f, amount);
            metallic *= 1.0f - amount;
        } else {
            rough += f.roughness_delta * (2.0f * d.h - 1.0f);
        }
        if (f.pattern == 2u) metallic *= 1.0f - 0.75f * amount;
        if (std::fabs(f.bump) > 0.0f) {  // finish_normal
            const F3 slope = d.g - n * dot(d.g, n);
            const float lim = 0.75f / std::fmax(std::sqrt(dot(slope, slope)), 1e-5f);
            shading_n = safe_normalize(n - slope * clampv(f.bump, -lim, lim));
        }
    }
    rough = clampv(rough, 0.02f, 1.0f);
    Surf s;
    s.n = shading_n;
    s.diffuse = albedo * (1.0f - metallic);
    s.f0 = mix3(v3(spec, spec, spec), albedo, metallic);
    s.alpha = rough * rough;
    s.emission = emission;
    s.sigma = rough * sigma_scale;
    return s;
}
// brdf_eval (pathtrace.wgsl:394-420): Lambert or Oren-Nayar diffuse + GGX.
F3 brdf_eval(const Surf& s, F3 v, F3 l) {
    float nl = dot(s.n, l), nv = dot(s.n, v);
    if (nl <= 0.0f || nv <= 0.0f) return v3(0.0f, 0.0f, 0.0f);
    F3 h = normalize(v + l);
    float nh = std::fmax(dot(s.n, h), 0.0f);
    float vh = std::fmax(dot(v, h), 0.0f);
    float a2 = s.alpha * s.alpha;
    F3 f = fresnel(s.f0, vh);
    float k = ggx_d(nh, a2) * smith_g1(nv, a2) * smith_g1(nl, a2);
    float den = std::fmax(4.0f * nv * nl, 1e-4f);
    F3 spec = v3(k * f.x / den, k * f.y / den, k * f.z / den);
    float df = 1.0f;
    if (s.sigma > 0.0f) {
        const float sigma2 = s.sigma * s.sigma;
        const float a = 1.0f - sigma2 / (2.0f * (sigma2 + 0.33f));
        const float b = 0.45f * sigma2 / (sigma2 + 0.09f);
        const float sin_l = std::sqrt(std::fmax(0.0f, 1.0f - nl * nl));
        const float sin_v = std::sqrt(std::fmax(0.0f, 1.0f - nv * nv));
        const float cos_phi = std::fmax(0.0f, dot(l - s.n * nl, v - s.n * nv) / std::fmax(sin_l * sin_v, 1e-8f));
        df = a + b * cos_phi * sin_l * sin_v / std::fmax(std::fmax(nl, nv), 1e-5f);
    }
    F3 one_f = v3(1.0f - f.x, 1.0f - f.y, 1.0f - f.z);
    F3 dif = mul3(s.diffuse, one_f) * df;
    return v3(dif.x / PI_F + spec.x, dif.y / PI_F + spec.y, dif.z / PI_F + spec.z);
}
inline float mis_power(float a, float b) {
    float scale = std::fmax(a, b);
    if (scale <= 0.0f) return 0.0f;
    float pa = a / scale, pb = b / scale;
    return pa * pa / (pa * pa + pb * pb);
}
inline float spec_probability(const Surf& s, float nv) {
    F3 f = fresnel(s.f0, nv);
    return clampv(luminance(f) / std::fmax(luminance(f) + luminance(s.diffuse), 1e-4f), 0.05f, 0.95f);
}
// bsdf_pdf (pathtrace.wgsl:443-452); continuation_pdf without guiding or transmission.
float bsdf_pdf(const Surf& s, F3 v, F3 l, float p_spec) {
    float nl = dot(s.n, l);
    if (nl <= 0.0f) return 0.0f;
    F3 hh = normalize(v + l);
    float nh = std::fmax(dot(s.n, hh), 0.0f);
    float a2 = s.alpha * s.alpha;
    float nv = std::fmax(dot(s.n, v), 1e-4f);
    float pdf_spec = ggx_d(nh, a2) * smith_g1(nv, a2) / (4.0f * nv);
    return p_spec * pdf_spec + (1.0f - p_spec) * nl / PI_F;
}
// environment_value / environment_probability (pathtrace.wgsl:164-175)
F3 environment_value(const LightL& l, F3 dir) {
    if (dir.y >= 0.0f) return mix3(l.horizon, l.color, std::sqrt(dir.y));
    return mix3(l.horizon, l.ground, std::sqrt(-dir.y));
}
float environment_probability(const LightL& l, float area_total) {
    float energy = std::fmax(std::fmax(luminance(l.color), luminance(l.horizon)), luminance(l.ground));
    if (energy <= 0.0f) return 0.0f;
    return clampv(4.0f * PI_F * energy / (4.0f * PI_F * energy + area_total), 0.05f, 1.0f);
}
// area_intersection (lights.wgsl:170-181)
float area_intersection(F3 ro, F3 rd, const LightL& l) {
    float den = dot(l.n, rd);
    if (std::fabs(den) < 1e-7f) return 1e30f;
    float t = dot(l.a - ro, l.n) / den;
    if (t <= 1e-6f) return 1e30f;
    F3 q = ro + rd * t - l.a;
    float x = dot(q, l.b) / dot(l.b, l.b);
    float y = dot(q, l.c) / dot(l.c, l.c);
    if (std::fabs(x) > 0.5f || std::fabs(y) > 0.5f) return 1e30f;
    return t;
}
// area_weight / area_total_weight (lights.wgsl:288-314)
float area_weight(F3 p, F3 n, const LightL& l) {
    F3 to = l.a - p;
    float dist2 = std::fmax(dot(to, to), 1e-6f);
    F3 dir = v3(to.x / std::sqrt(dist2), to.y / std::sqrt(dist2), to.z / std::sqrt(dist2));
    float facing = std::fmax(dot(l.n, neg(dir)), 0.0f);
    return luminance(l.color) * l.area * facing * std::fmax(dot(n, dir), 0.15f) / dist2;
}
float area_total_weight(const SceneL& S, F3 p, F3 n) {
    float total = 0.0f;
    for (uint32_t i = 0; i < S.light_count; ++i)
        if (S.lights[i].kind == SYF_LIGHT_AREA) total += area_weight(p, n, S.lights[i]);
    return total;
}
struct HitL {
    Hit<float> h;
    uint32_t light;
};
constexpr uint32_t GROUND_HIT = 0xfffffffeu;  // HitL.light of a segment that ended on the ground
// ground_intersection / ground_normal of a flat ground (landscape.wgsl:17-23, 71-78)
float ground_intersection(const SceneL& S, F3 ro, F3 rd, float t_end) {
    if (std::fabs(rd.y) < 1e-7f) return t_end;
    const float t = (S.ground.height - ro.y) / rd.y;
    return (t > 1e-5f && t < t_end) ? t : t_end;
}
inline F3 ground_normal(F3 rd) { return rd.y > 0.0f ? v3(0.0f, -1.0f, 0.0f) : v3(0.0f, 1.0f, 0.0f); }
// trace_geometry_lights (lights.wgsl:183-209): the ground, the area lights, then the march up to
// the nearest of them.
HitL trace_scene(const SceneL& S, F3 ro, F3 rd, float t_end, Count* c) {
    float nearest = t_end;
    uint32_t light = NO_LIGHT;
    if (S.ground.enabled) {
        const float t = ground_intersection(S, ro, rd, nearest);
        if (t > 1e-5f && t < nearest) {
            nearest = t;
            light = GROUND_HIT;
        }
    }
    for (uint32_t i = 0; i < S.light_count; ++i) {
        if (S.lights[i].kind != SYF_LIGHT_AREA) continue;
        float t = area_intersection(ro, rd, S.lights[i]);
        if (t < nearest) {
            nearest = t;
            light = i;
        }
    }
    HitL r;
    r.h = march(S.F, ro, rd, S.pixel_cone, S.pixel_tol, S.max_steps, nearest, c);
    r.light = NO_LIGHT;
    if (!r.h.hit && !r.h.exhausted && light != NO_LIGHT) {
        r.h.hit = true;
        r.h.t = nearest;
        r.light = light;
    }
    return r;
}
inline bool visible(const SceneL& S, F3 ro, F3 rd, float t_end, Count* c) {
    HitL h = trace_scene(S, ro, rd, t_end, c);
    return !h.h.hit && !h.h.exhausted;
}
inline float volume_light_probability(float weight, float total) {  // lights.wgsl:319-322
    return total <= 0.0f ? 0.0f : std::fmin(1.0f, 3.0f * weight / total);
}
// area_emission (lights.wgsl:330-357)
F3 area_emission(const SceneL& S, F3 ro, F3 rd, float distance, float prev_pdf, bool use_mis, F3 prev_n,
                 float prev_total, bool from_volume) {
    F3 L = v3(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < S.light_count; ++i) {
        const LightL& l = S.lights[i];
        if (l.kind != SYF_LIGHT_AREA) continue;
        float cos_l = dot(l.n, neg(rd));
        if (cos_l <= 0.0f) continue;
        float t = area_intersection(ro, rd, l);
        if (std::fabs(t - distance) > std::fmax(1e-6f, distance * 1e-6f)) continue;
        float w = 1.0f;
        if (use_mis) {
            float pl = distance * distance / std::fmax(l.area * cos_l, 1e-20f);
            float sel = 0.0f;
            if (prev_total > 0.0f)  // from_volume: a phase sample's light selection (lights.wgsl:346-350)
                sel = from_volume ? volume_light_probability(area_weight(ro, v3(0.0f, 0.0f, 0.0f), l), prev_total)
                                  : area_weight(ro, prev_n, l) / prev_total;
            pl *= sel;
            w = mis_power(prev_pdf, pl);
        }
        L = L + l.color * w;
    }
    return L;
}
// environment_emission (pathtrace.wgsl:177-189)
F3 environment_emission(const SceneL& S, F3 dir, float pdf, bool use_mis, F3 n, float area_total, bool from_volume) {
    F3 L = v3(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < S.light_count; ++i) {
        const LightL& l = S.lights[i];
        if (l.kind != SYF_LIGHT_ENV) continue;
        float w = 1.0f;
        if (use_mis) {
            float proposal = environment_probability(l, area_total) *
                             (from_volume ? 1.0f / (4.0f * PI_F) : std::fmax(dot(n, dir), 0.0f) / PI_F);
            w = mis_power(pdf, proposal);
        }
        L = L + environment_value(l, dir) * w;
    }
    return L;
}
// cone_sample / sun_pdf (pathtrace.wgsl:210-215, 432-434)
F3 cone_sample(F3 dir, float cos_max, float r1, float r2) {
    const float cos_t = 1.0f - r1 * (1.0f - cos_max);
    const float sin_t = std::sqrt(std::fmax(0.0f, 1.0f - cos_t * cos_t));
    const float phi = 2.0f * PI_F * r2;
    F3 t, bt;
    tangent_frame(dir, &t, &bt);
    return frame_mul(t, bt, dir, v3(sin_t * std::cos(phi), sin_t * std::sin(phi), cos_t));
}
inline float sun_pdf(float cos_max) { return 1.0f / (2.0f * PI_F * std::fmax(1.0f - cos_max, 1e-7f)); }
// sun_radiance (pathtrace.wgsl:454-467)
F3 sun_radiance(const SceneL& S, F3 dir, float pdf_bsdf, bool use_mis) {
    F3 L = v3(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < S.light_count; ++i) {
        const LightL& l = S.lights[i];
        if (l.kind != SYF_LIGHT_SUN || !(dot(dir, l.a) >= l.cos_max)) continue;
        const float pl = sun_pdf(l.cos_max);
        const float w = use_mis ? mis_power(pdf_bsdf, pl) : 1.0f;
        L = L + v3(l.color.x * pl * w, l.color.y * pl * w, l.color.z * pl * w);
    }
    return L;
}
// ---------------------------------------------------------------- height fog (pathtrace.wgsl:560-779)
float fog_sigma(const SceneL& S, F3 p) {  // :562-565
    return S.fog.density * std::exp(std::fmin(0.0f, -(p.y - S.fog.height) * S.fog.falloff));
}
float fog_segment_tau(const SceneL& S, F3 ro, F3 rd, float begin, float end) {  // :569-578
    const float distance = std::fmax(end - begin, 0.0f);
    if (distance == 0.0f) return 0.0f;
    const float sigma = fog_sigma(S, ro + rd * begin);
    const float mid_height = ro.y + rd.y * ((begin + end) * 0.5f);
    if (mid_height <= S.fog.height || S.fog.falloff == 0.0f) return S.fog.density * distance;
    const float k = rd.y * S.fog.falloff;
    const float x = k * distance;
    if (std::fabs(x) < 0.001f) return sigma * distance * (1.0f - x * 0.5f + x * x / 6.0f);
    return std::fmax(0.0f, (sigma - fog_sigma(S, ro + rd * end)) / k);
}
float fog_split(const SceneL& S, F3 ro, F3 rd, float distance) {  // :580-583
    if (std::fabs(rd.y) < 1e-8f) return distance;
    return clampv((S.fog.height - ro.y) / rd.y, 0.0f, distance);
}
float fog_optical_depth(const SceneL& S, F3 ro, F3 rd, float distance) {  // :585-588
    const float split = fog_split(S, ro, rd, distance);
    return fog_segment_tau(S, ro, rd, 0.0f, split) + fog_segment_tau(S, ro, rd, split, distance);
}
inline float fog_transmittance(const SceneL& S, F3 ro, F3 rd, float dist) {  // :590-592
    return std::exp(-fog_optical_depth(S, ro, rd, dist));
}
float fog_invert_segment(const SceneL& S, F3 ro, F3 rd, float begin, float end, float tau) {  // :594-604
    const float sigma = fog_sigma(S, ro + rd * begin);
    const float k = rd.y * S.fog.falloff;
    if (ro.y + rd.y * ((begin + end) * 0.5f) <= S.fog.height || std::fabs(k * (end - begin)) < 0.0001f)
        return std::fmin(end, begin + tau / std::fmax(sigma, 1e-30f));
    if (sigma > 1e-20f) return clampv(begin - std::log(std::fmax(1e-7f, 1.0f - tau * k / sigma)) / k, begin, end);
    const float remaining = std::fmax(0.0f, fog_segment_tau(S, ro, rd, begin, end) - tau);
    const float end_sigma = fog_sigma(S, ro + rd * end);
    return clampv(end + std::log(std::fmax(1e-7f, 1.0f + remaining * k / std::fmax(end_sigma, 1e-30f))) / (-k), begin,
                  end);
}
float fog_free_flight(const SceneL& S, F3 ro, F3 rd, float t_end, uint32_t* rng) {  // :606-615
    if (S.fog.density <= 0.0f) return t_end;
    const float tau = -std::log(std::fmax(1.0f - rand01(rng), 1e-7f));
    const float split = fog_split(S, ro, rd, t_end);
    const float first = fog_segment_tau(S, ro, rd, 0.0f, split);
    const float total = first + fog_segment_tau(S, ro, rd, split, t_end);
    if (tau >= total) return t_end;
    if (tau < first) return fog_invert_segment(S, ro, rd, 0.0f, split, tau);
    return fog_invert_segment(S, ro, rd, split, t_end, tau - first);
}
float volume_coordinate(uint32_t sample_index, uint32_t dimension, uint32_t stream) {  // :43-58
    static const uint32_t bases[6] = {2u, 3u, 5u, 7u, 11u, 13u};
    const uint32_t base = bases[dimension];
    uint32_t index = sample_index + 1u;
    float inverse = 1.0f / float(base);
    float value = 0.0f;
    while (index != 0u) {
        value += float(index % base) * inverse;
        index /= base;
        inverse /= float(base);
    }
    const float shift = float(pcg(stream ^ (dimension * 0x9e3779b9u)) >> 8u) / 16777216.0f;
    const float x = value + shift;
    return x - std::floor(x);
}
float hg_phase(float cos_theta, float g) {  // :717-720
    const float d = 1.0f + g * g - 2.0f * g * cos_theta;
    return (1.0f - g * g) / (4.0f * PI_F * d * std::sqrt(d));
}
F3 hg_sample(F3 dir, float g, float r1, float r2) {  // :722-731
    float cos_t = 1.0f - 2.0f * r1;
    if (std::fabs(g) > 1e-3f) {
        const float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * r1);
        cos_t = (1.0f + g * g - sq * sq) / (2.0f * g);
    }
    const float sin_t = std::sqrt(std::fmax(0.0f, 1.0f - cos_t * cos_t));
    const float phi = 2.0f * PI_F * r2;
    F3 t, bt;
    tangent_frame(dir, &t, &bt);
    return frame_mul(t, bt, dir, v3(sin_t * std::cos(phi), sin_t * std::sin(phi), cos_t));
}
float fog_segment_weight(const LightL& l, F3 ro, F3 rd, float end) {  // :621-633, area lights
    const float along = clampv(dot(l.a - ro, rd), 0.0f, end);
    const F3 delta = l.a - (ro + rd * along);
    return std::fmax(std::fmax(l.color.x, l.color.y), l.color.z) * l.area / (dot(delta, delta) + l.area);
}
// fog_segment_direct (:635-715): the emitters' light scattered toward [0, end) of the ray.
F3 fog_segment_direct(const SceneL& S, F3 ro, F3 rd, float end, uint32_t* rng, uint32_t point_index,
                      uint32_t point_stream, bool primary, Count* c) {
    if (end <= 0.0f || std::fmax(std::fmax(S.fog.color[0], S.fog.color[1]), S.fog.color[2]) <= 0.0f)
        return v3(0.0f, 0.0f, 0.0f);
    const float tau_end = fog_optical_depth(S, ro, rd, end);
    if (tau_end <= 0.0f) return v3(0.0f, 0.0f, 0.0f);
    float total = 0.0f;
    for (uint32_t i = 0; i < S.light_count; ++i)
        if (S.lights[i].kind == SYF_LIGHT_AREA) total += fog_segment_weight(S.lights[i], ro, rd, end);
    F3 result = v3(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < S.light_count; ++i) {
        const LightL& l = S.lights[i];
        if (l.kind != SYF_LIGHT_AREA) continue;
        const float probability = volume_light_probability(fog_segment_weight(l, ro, rd, end), total);
        const uint32_t stream = pcg(point_stream ^ (i * 0x85ebca6bu));
        float selection = rand01(rng);
        if (primary) selection = volume_coordinate(point_index, 5u, stream);
        if (probability <= 0.0f || selection >= probability) continue;
        float u1 = rand01(rng);
        float u2 = rand01(rng);
        if (primary) {
            u1 = volume_coordinate(point_index, 0u, stream);
            u2 = volume_coordinate(point_index, 1u, stream);
        }
        const F3 q = l.a + l.b * (u1 - 0.5f) + l.c * (u2 - 0.5f);  // sample_emitter_point
        const float tau_begin = fog_optical_depth(S, ro, rd, 0.0f);
        const float tau_finish = fog_optical_depth(S, ro, rd, end);
        const float tau_span = std::fmax(0.0f, tau_finish - tau_begin);
        if (tau_span <= 0.0f) continue;
        float conditional_mass = 1.0f - std::exp(-tau_span);
        if (tau_span < 0.001f) conditional_mass = tau_span * (1.0f - tau_span * 0.5f + tau_span * tau_span / 6.0f);
        const float interval_mass = std::exp(-tau_begin) * conditional_mass;
        if (interval_mass < 1e-30f) continue;
        const F3 delta = q - ro;
        const float projection = dot(delta, rd);
        const float perpendicular = length(delta - rd * projection);
        const float theta_a = std::atan2(0.0f - projection, std::fmax(perpendicular, 1e-6f));
        const float theta_b = std::atan2(end - projection, std::fmax(perpendicular, 1e-6f));
        const float span = theta_b - theta_a;
        const bool regular = perpendicular > 1e-4f && span > 1e-5f;
        float xi = rand01(rng);
        float technique = rand01(rng);
        if (primary) {
            xi = volume_coordinate(point_index, 3u, stream);
            technique = volume_coordinate(point_index, 4u, stream);
        }
        float t = mixf(0.0f, end, xi);
        if (technique < 0.5f) {  // equiangular
            if (regular) t = clampv(projection + perpendicular * std::tan(mixf(theta_a, theta_b, xi)), 0.0f, end);
        } else {  // distance, by the inverse optical depth
            const float tau = tau_begin - std::log(std::fmax(1.0f - xi * conditional_mass, 1e-7f));
            const float split = fog_split(S, ro, rd, end);
            const float first = fog_segment_tau(S, ro, rd, 0.0f, split);
            t = tau < first ? fog_invert_segment(S, ro, rd, 0.0f, split, tau)
                            : fog_invert_segment(S, ro, rd, split, end, tau - first);
        }
        const F3 p = ro + rd * t;
        const float transmittance = std::exp(-fog_optical_depth(S, ro, rd, t));
        const float sigma = fog_sigma(S, p);
        float angular_pdf = 1.0f / (end - 0.0f);
        if (regular)
            angular_pdf = perpendicular / (span * (perpendicular * perpendicular + (t - projection) * (t - projection)));
        const float distance_pdf = 0.5f * (angular_pdf + sigma * transmittance / interval_mass);
        const F3 to_light = q - p;
        const float distance2 = dot(to_light, to_light);
        if (distance2 < 1e-8f) continue;
        const float distance = std::sqrt(distance2);
        const F3 wi = v3(to_light.x / distance, to_light.y / distance, to_light.z / distance);
        const float cosine = dot(l.n, neg(wi));
        if (cosine <= 0.0f) continue;
        if (!visible(S, p, wi, std::fmax(0.0f, distance - std::fmax(1e-5f, distance * 1e-6f)), c)) continue;
        const float tr = transmittance * std::exp(-fog_optical_depth(S, p, wi, distance));
        const float area_pdf = distance2 / std::fmax(l.area * cosine, 1e-20f);  // area_direction_pdf
        const float k = sigma * tr * hg_phase(dot(rd, wi), S.fog.anisotropy) * 1.0f * 1.0f /
                        (probability * distance_pdf * area_pdf);
        result = result + v3(l.color.x * k, l.color.y * k, l.color.z * k);
    }
    return v3(result.x * S.fog.color[0], result.y * S.fog.color[1], result.z * S.fog.color[2]);
}
// fog_direct_light (:733-778): environment and suns at a scattering event, phase-weighted.
F3 fog_direct_light(const SceneL& S, F3 p, F3 rd, float area_total, bool last, uint32_t* rng, Count* c) {
    const float g = S.fog.anisotropy;
    F3 L = v3(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < S.light_count; ++i) {
        const LightL& l = S.lights[i];
        if (l.kind == SYF_LIGHT_ENV) {
            const float mx = std::fmax(std::fmax(std::fmax(l.color.x, l.color.y), l.color.z),
                                       std::fmax(std::fmax(std::fmax(l.horizon.x, l.horizon.y), l.horizon.z),
                                                 std::fmax(std::fmax(l.ground.x, l.ground.y), l.ground.z)));
            if (mx <= 0.0f) continue;
            const float probability = environment_probability(l, area_total);
            if (rand01(rng) >= probability) continue;
            const float z = 1.0f - 2.0f * rand01(rng);
            const float phi = 2.0f * PI_F * rand01(rng);
            const float radius = std::sqrt(std::fmax(0.0f, 1.0f - z * z));
            const F3 dir = v3(radius * std::cos(phi), z, radius * std::sin(phi));
            if (!visible(S, p, dir, S.t_end, c)) continue;
            const float tr = fog_transmittance(S, p, dir, S.t_end);
            const float ph = hg_phase(dot(rd, dir), g);
            const float pl = probability / (4.0f * PI_F);
            const float w = last ? 1.0f : mis_power(pl, ph);
            const F3 e = environment_value(l, dir);
            L = L + v3(e.x * ph * tr * w / pl, e.y * ph * tr * w / pl, e.z * ph * tr * w / pl);
        } else if (l.kind == SYF_LIGHT_SUN) {
            const float r1 = rand01(rng);
            const float r2 = rand01(rng);
            const F3 dir = cone_sample(l.a, l.cos_max, r1, r2);
            if (!visible(S, p, dir, S.t_end, c)) continue;
            const float tr = fog_transmittance(S, p, dir, S.t_end);
            const float ph = hg_phase(dot(rd, dir), g);
            const float pl = sun_pdf(l.cos_max);
            const float w = last ? 1.0f : mis_power(pl, ph);
            L = L + v3(l.color.x * ph * tr * w, l.color.y * ph * tr * w, l.color.z * ph * tr * w);
        }
    }
    return L;
}
// direct_light (pathtrace.wgsl:471-558): environment, sun and area branches.
F3 direct_light(const SceneL& S, F3 p, const Surf& s, F3 gn, F3 v, float p_spec, float area_total, bool last,
                bool in_fog, uint32_t* rng, Count* c) {
    F3 L = v3(0.0f, 0.0f, 0.0f);
    uint32_t selected = NO_LIGHT;
    float selected_weight = 0.0f;
    if (area_total > 0.0f) {
        float threshold = rand01(rng) * area_total;
        float sum = 0.0f;
        for (uint32_t i = 0; i < S.light_count; ++i) {
            if (S.lights[i].kind != SYF_LIGHT_AREA) continue;
            float w = area_weight(p, s.n, S.lights[i]);
            sum += w;
            if (w > 0.0f) {
                selected = i;
                selected_weight = w;
                if (threshold < sum) break;
            }
        }
    }
    for (uint32_t i = 0; i < S.light_count; ++i) {
        const LightL& l = S.lights[i];
        if (l.kind == SYF_LIGHT_ENV) {
            float mx = std::fmax(std::fmax(std::fmax(l.color.x, l.color.y), l.color.z),
                                 std::fmax(std::fmax(std::fmax(l.horizon.x, l.horizon.y), l.horizon.z),
                                           std::fmax(std::fmax(l.ground.x, l.ground.y), l.ground.z)));
            if (mx <= 0.0f) continue;
            float prob = environment_probability(l, area_total);
            if (rand01(rng) >= prob) continue;
            float r1 = rand01(rng);
            float r2 = rand01(rng);
            F3 dir = cosine_sample(s.n, r1, r2);
            float nl = std::fmax(dot(s.n, dir), 0.0f);
            if (dot(gn, dir) <= 0.0f || nl <= 0.0f) continue;
            if (!visible(S, p, dir, S.t_end, c)) continue;
            const float tr = in_fog ? fog_transmittance(S, p, dir, S.t_end) : 1.0f;
            float pl = prob * nl / PI_F;
            float pb = last ? 0.0f : bsdf_pdf(s, v, dir, p_spec);
            float w = last ? 1.0f : mis_power(pl, pb);
            F3 f = brdf_eval(s, v, dir);
            F3 e = environment_value(l, dir);
            L = L + v3(f.x * nl * e.x * tr * w / pl, f.y * nl * e.y * tr * w / pl, f.z * nl * e.z * tr * w / pl);
        } else if (l.kind == SYF_LIGHT_SUN) {
            // colour = irradiance: the disk's radiance is colour * sun_pdf
            const float r1 = rand01(rng);
            const float r2 = rand01(rng);
            const F3 dir = cone_sample(l.a, l.cos_max, r1, r2);
            const float nl = dot(s.n, dir);
            if (nl <= 0.0f || dot(gn, dir) <= 0.0f) continue;
            if (!visible(S, p, dir, S.t_end, c)) continue;
            const float tr = in_fog ? fog_transmittance(S, p, dir, S.t_end) : 1.0f;
            const float pl = sun_pdf(l.cos_max);
            const float pb = last ? 0.0f : bsdf_pdf(s, v, dir, p_spec);
            const float w = last ? 1.0f : mis_power(pl, pb);
            const F3 f = brdf_eval(s, v, dir);
            L = L + v3(f.x * nl * l.color.x * w * tr, f.y * nl * l.color.y * w * tr, f.z * nl * l.color.z * w * tr);
        } else if (l.kind == SYF_LIGHT_AREA) {
            if (i != selected) continue;
            float u1 = rand01(rng);
            float u2 = rand01(rng);
            F3 q = l.a + l.b * (u1 - 0.5f) + l.c * (u2 - 0.5f);  // sample_emitter_point (lights.wgsl:146-150)
            F3 to_q = q - p;
            float dist2 = dot(to_q, to_q);
            if (dist2 < 1e-8f) continue;
            float dist = std::sqrt(dist2);
            F3 dir = v3(to_q.x / dist, to_q.y / dist, to_q.z / dist);
            float nl = dot(s.n, dir);
            float cos_l = dot(l.n, neg(dir));
            if (nl <= 0.0f || cos_l <= 0.0f || dot(gn, dir) <= 0.0f) continue;
            float end = std::fmax(0.0f, dist - std::fmax(1e-5f, dist * 1e-6f));  // shadow_endpoint
            if (!visible(S, p, dir, end, c)) continue;
            const float tr = in_fog ? fog_transmittance(S, p, dir, dist) : 1.0f;
            float pl = (selected_weight / area_total) * (dist2 / std::fmax(l.area * cos_l, 1e-20f));
            float pb = last ? 0.0f : bsdf_pdf(s, v, dir, p_spec);
            float w = last ? 1.0f : mis_power(pl, pb);
            F3 f = brdf_eval(s, v, dir);
            L = L + v3(f.x * nl * l.color.x * tr * w / pl, f.y * nl * l.color.y * tr * w / pl,
                       f.z * nl * l.color.z * tr * w / pl);
        }
    }
    return L;
}

// radiance (pathtrace.wgsl:802-1201) for the subset; returns L and the first-hit planes.
// volume_out: the sample's volume radiance (the camera segment's in-scattering, or all of L when
// the camera ray scattered in the fog; pathtrace.wgsl:1197-1198).
F3 radiance(const SceneL& S, uint32_t px, uint32_t py, uint32_t sample, uint32_t* rng, float* first_depth,
            float* first_aux, float* first_tprev, F3* volume_out, Count* c) {
    F3 ro, rd;
    camera_ray(S.cam, S.width, S.height, px, py, rng, &ro, &rd);
    F3 throughput = v3(1.0f, 1.0f, 1.0f), L = v3(0.0f, 0.0f, 0.0f);
    float prev_pdf = 0.0f, prev_total = 0.0f;
    float tex_width = 0.0f, tex_spread = S.pixel_cone * 0.5f;  // pathtrace.wgsl:811-814
    bool prev_from_bsdf = false, prev_from_volume = false, prev_volume_nee = false, volume_path = false;
    F3 prev_n = v3(0.0f, 0.0f, 0.0f), primary_volume = v3(0.0f, 0.0f, 0.0f);
    const uint32_t point_stream = pcg(pcg(px ^ S.seed) ^ pcg(py ^ 0x27d4eb2du));  // :822
    const bool in_fog = S.fog.density > 0.0f;
    for (uint32_t bounce = 0; bounce <= S.max_bounces;) {
        HitL hl = trace_scene(S, ro, rd, S.t_end, c);
        const Hit<float>& h = hl.h;
        F3 gn = v3(0.0f, 0.0f, 0.0f);
        const bool ground = hl.light == GROUND_HIT;
        if (h.hit) {
            if (ground) {
                gn = ground_normal(rd);
            } else if (hl.light != NO_LIGHT) {
                gn = S.lights[hl.light].n;
            } else {
                gn = normal_at(S.F, ro + rd * h.t, std::fmax(h.t * S.pixel_cone * 0.5f, 1e-4f) * S.pixel_tol, c);
            }
            if (bounce == 0) {
                *first_depth = h.t;
                *first_tprev = hl.light == NO_LIGHT ? h.t_prev : -1.0f;  // refine bracket of a fractal hit
                first_aux[0] = gn.x; first_aux[1] = gn.y; first_aux[2] = gn.z; first_aux[3] = float(h.steps);
            }
        }
        if (in_fog) {  // :980-1010
            const float t_end = h.hit ? h.t : S.t_end;
            const F3 dv = mul3(throughput, fog_segment_direct(S, ro, rd, t_end, rng, sample, point_stream, bounce == 0, c));
            L = L + dv;
            if (bounce == 0) primary_volume = primary_volume + dv;
            const float t_s = fog_free_flight(S, ro, rd, t_end, rng);
            if (t_s < t_end) {
                if (bounce == 0) volume_path = true;
                const F3 p_s = ro + rd * t_s;
                tex_width += t_s * tex_spread;
                tex_spread = std::fmax(tex_spread, 0.25f);
                if (bounce == 0 && !h.hit) *first_depth = t_s;
                const float vat = area_total_weight(S, p_s, v3(0.0f, 0.0f, 0.0f));
                const bool last = bounce == S.max_bounces;
                const F3 fc = f3(S.fog.color);
                L = L + mul3(mul3(throughput, fc), fog_direct_light(S, p_s, rd, vat, last, rng, c));
                throughput = mul3(throughput, fc);
                if (last) break;
                const float r1 = rand01(rng);
                const float r2 = rand01(rng);
                const F3 nd = hg_sample(rd, S.fog.anisotropy, r1, r2);
                prev_pdf = hg_phase(dot(rd, nd), S.fog.anisotropy);
                prev_from_bsdf = true;
                prev_from_volume = true;
                prev_volume_nee = true;
                prev_n = v3(0.0f, 0.0f, 0.0f);
                prev_total = vat;
                ro = p_s;
                rd = nd;
                ++bounce;
                continue;
            }
        }
        if (!h.hit) {
            if (!h.exhausted) {
                F3 e = environment_emission(S, rd, prev_pdf, prev_from_bsdf, prev_n, prev_total, prev_from_volume);
                L = L + mul3(throughput, e);
                L = L + mul3(throughput, sun_radiance(S, rd, prev_pdf, prev_from_bsdf));
            }
            break;
        }
        F3 p = ro + rd * h.t;
        if (!ground && hl.light != NO_LIGHT) {
            if (!prev_volume_nee) {  // a fog segment already counted the emitters by its line integral
                F3 e = area_emission(S, ro, rd, h.t, prev_pdf, prev_from_bsdf, prev_n, prev_total, prev_from_volume);
                L = L + mul3(throughput, e);
            }
            break;
        }
        F3 n = gn;
        if (bounce == 0 && !ground) {
            float k = 1.0f - S.cavity * smoothstep01(float(h.steps) / 96.0f);
            throughput = throughput * k;
        }
        F3 v = neg(rd);
        // the ground's sample: the march's initial DeSample (pathtrace.wgsl:86) with the ground
        // finish as material (lights.wgsl:201)
        DeOut<float> smp = h.sample;
        if (ground) {
            smp.traps[0] = smp.traps[1] = smp.traps[2] = smp.traps[3] = 1.0f;
            smp.iter = 0u;
            smp.material = S.ground.finish;
        }
        tex_width += h.t * tex_spread;
        Surf s = make_surface(S, p, n, smp, std::fmax(tex_width, 1e-4f));
        L = L + mul3(throughput, s.emission);
        float offset = ground ? 1e-3f : std::fmax(h.t * S.pixel_cone * 2.0f, 1e-3f) * S.pixel_tol;
        F3 p_off = p + n * offset;
        float nv = std::fmax(dot(s.n, v), 1e-4f);
        float p_spec = spec_probability(s, nv);
        float area_total = area_total_weight(S, p_off, s.n);
        bool last = bounce == S.max_bounces;
        F3 dl = direct_light(S, p_off, s, n, v, p_spec, area_total, last, in_fog, rng, c);
        L = L + mul3(throughput, dl);
        if (last) break;
        float r0 = rand01(rng), r1 = rand01(rng), r2 = rand01(rng);
        F3 l;
        if (r0 < p_spec) {
            F3 hh = sample_ggx_h(s.n, v, s.alpha, r1, r2);
            l = reflect3(neg(v), hh);
        } else {
            l = cosine_sample(s.n, r1, r2);
        }
        float nl = dot(s.n, l);
        if (nl <= 0.0f || dot(n, l) <= 0.0f) break;
        float pdf = bsdf_pdf(s, v, l, p_spec);
        if (pdf <= 1e-6f) break;
        F3 f = brdf_eval(s, v, l);
        throughput = v3(throughput.x * (f.x * nl / pdf), throughput.y * (f.y * nl / pdf), throughput.z * (f.z * nl / pdf));
        prev_pdf = pdf;
        prev_from_bsdf = true;
        prev_n = s.n;
        prev_total = area_total;
        prev_from_volume = false;
        prev_volume_nee = false;
        if (bounce >= 2) {
            float q = clampv(std::fmax(throughput.x, std::fmax(throughput.y, throughput.z)), 0.05f, 0.95f);
            if (rand01(rng) > q) break;
            throughput = v3(throughput.x / q, throughput.y / q, throughput.z / q);
        }
        ro = p_off;
        rd = l;
        tex_spread = std::fmax(tex_spread, s.alpha * 0.25f);
        ++bounce;
    }
    *volume_out = volume_path ? L : primary_volume;
    return L;
}

float ray_extent(const syf_job_desc& d, const Basis& b) {
    // uniforms.rs:781-782 and common.wgsl:206-208
    float p0z = std::fmax(2.0f * bound_hint_of(d.fractal) + length(b.origin), 200.0f);
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

// Loads the node graph that jobs and probes of family SYF_FAMILY_GRAPH evaluate from then on.
// Returns 0, or -1 when an id does not resolve or the graph has a cycle.
int orc_graph_use(const syf_graph* g) { return load_graph(*g, &g_graph) ? 0 : -1; }

// Derived properties of the loaded graph: step scale (eval.rs:879-881), iteration product
// (validate.rs:112-137), trap scales (eval.rs:853-877), Lipschitz factor at `radius`
// (eval.rs:828-849), and the plan (node ids in evaluation order, at most plan_cap).
int orc_graph_props(float radius, float* step, uint32_t* product, float* scales4, float* lipschitz_at,
                    uint32_t* plan, uint32_t plan_cap, uint32_t* plan_len) {
    if (!g_graph.ok) return -1;
    *step = step_scale(g_graph);
    *product = g_graph.iteration_product;
    for (int i = 0; i < 4; ++i) scales4[i] = g_graph.trap_scales[i];
    *lipschitz_at = lipschitz(g_graph, radius);
    *plan_len = uint32_t(g_graph.plan.size());
    for (uint32_t i = 0; i < g_graph.plan.size() && i < plan_cap; ++i) plan[i] = g_graph.plan[i];
    return 0;
}

// One point node of the loaded graph applied to the start state at n points, as the
// reference's unit tests do (Graph::apply, eval.rs:341-344): out4 = (p', dr).
int orc_graph_point(uint32_t node_id, int precision, const float* xyz4, size_t n, double* out4) {
    if (!g_graph.node(node_id)) return -1;
    for (size_t i = 0; i < n; ++i) {
        if (precision == 64) {
            V3<double> p = v3(double(xyz4[4 * i]), double(xyz4[4 * i + 1]), double(xyz4[4 * i + 2]));
            PState<double> s{p, 1.0, 0u, {INFINITY, INFINITY, INFINITY, INFINITY}};
            s = apply_seeded(g_graph, node_id, s, p, p, 1.0);
            out4[4 * i] = s.p.x; out4[4 * i + 1] = s.p.y; out4[4 * i + 2] = s.p.z; out4[4 * i + 3] = s.dr;
        } else {
            V3<float> p = v3(xyz4[4 * i], xyz4[4 * i + 1], xyz4[4 * i + 2]);
            PState<float> s{p, 1.0f, 0u, {INFINITY, INFINITY, INFINITY, INFINITY}};
            s = apply_seeded(g_graph, node_id, s, p, p, 1.0f);
            out4[4 * i] = s.p.x; out4[4 * i + 1] = s.p.y; out4[4 * i + 2] = s.p.z; out4[4 * i + 3] = s.dr;
        }
    }
    return 0;
}

// The loaded graph at n points (xyz4), precision 32 or 64: distance, traps[4], iterations,
// material index, and the Iterate-body executions.
int orc_graph_probe(int precision, const float* xyz4, size_t n, double* dist, double* traps4, uint32_t* iters,
                    uint32_t* materials, uint32_t* body_runs) {
    if (!g_graph.ok) return -1;
    for (size_t i = 0; i < n; ++i) {
        DeOut<double> o;
        if (precision == 32) {
            DeOut<float> f = graph_evaluate(g_graph, v3(xyz4[4 * i], xyz4[4 * i + 1], xyz4[4 * i + 2]));
            o.d = f.d;
            for (int k = 0; k < 4; ++k) o.traps[k] = f.traps[k];
            o.iter = f.iter;
            o.material = f.material;
            o.body_runs = f.body_runs;
        } else if (precision == 64) {
            o = graph_evaluate(g_graph, v3(double(xyz4[4 * i]), double(xyz4[4 * i + 1]), double(xyz4[4 * i + 2])));
        } else {
            return -1;
        }
        dist[i] = o.d;
        for (int k = 0; k < 4; ++k) traps4[4 * i + k] = o.traps[k];
        iters[i] = o.iter;
        if (materials) materials[i] = o.material;
        if (body_runs) body_runs[i] = o.body_runs;
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

// The trigonometric Power branch forced at n points, any power (integer ones included), in
// precision 32 or 64: out4 = (z^n, n * r^(n-1)).  Pinned against the integer spherical
// product at integer powers.
void orc_trig_power(const float* xyz4, size_t n, float power, int precision, double* out4) {
    for (size_t i = 0; i < n; ++i) {
        if (precision == 64) {
            V3<double> p = v3(double(xyz4[4 * i]), double(xyz4[4 * i + 1]), double(xyz4[4 * i + 2]));
            double dr = 1.0;
            trig_power(&p, &dr, double(power));
            out4[4 * i] = p.x; out4[4 * i + 1] = p.y; out4[4 * i + 2] = p.z; out4[4 * i + 3] = dr;
        } else {
            V3<float> p = v3(xyz4[4 * i], xyz4[4 * i + 1], xyz4[4 * i + 2]);
            float dr = 1.0f;
            trig_power(&p, &dr, power);
            out4[4 * i] = p.x; out4[4 * i + 1] = p.y; out4[4 * i + 2] = p.z; out4[4 * i + 3] = dr;
        }
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
                double t, t_prev = 0.0, g_prev = 0.0, g_hit = 0.0;
                const double cone64 = double(pixel_cone) * double(d->pixel_tolerance);
                uint32_t st;
                V3<double> n;
                if (precision == 32) {
                    Hit<float> hh = march(F32, ro, rd, pixel_cone, d->pixel_tolerance, d->max_steps, t_end, &c);
                    hit = hh.hit;
                    t = hh.t;
                    t_prev = hh.t_prev;
                    g_prev = end_g(hh.r_prev, hh.t_prev, cone64);
                    g_hit = end_g(hh.r_hit, hh.t, cone64);
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
                    g_prev = end_g(hh.r_prev, hh.t_prev, cone64);
                    g_hit = end_g(hh.r_hit, hh.t, cone64);
                    st = hh.steps;
                    if (hit) {
                        V3<double> p = ro64 + rd64 * hh.t;
                        n = normal_at(F64, p, std::fmax(hh.t * double(pixel_cone) * 0.5, 1e-4) * double(d->pixel_tolerance), &c);
                    }
                }
                steps[i] = st;
                if (hit) {
                    depth[i] = t;
                    depth64[i] = refine64(F64, ro64, rd64, t_prev, t, g_prev, g_hit, cone64, &c);
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

// The control machine's image (ControlSession, control.wgsl:54-136,267-291) of a pixel
// rectangle for the job's field and camera: the pixel-centre ray (control.wgsl:271-274), the
// over-relaxed march of march_segment with pixel tolerance 1 and no step limit
// (control.wgsl:113-136; here at most max_evals evaluations, status 2 beyond), on a hit the
// four tetrahedral taps at h = max(t * cone / 2, 1e-4) summed as sum(k_i d_i) and
// safe-normalised, and the headlight shade (0.68, 0.70, 0.73) * (0.2 + 0.8 max(n.l, 0)) with
// l = normalize(0.4, 0.7, 0.6) (control.wgsl:79-97).  A miss shades (0.025, 0.035, 0.05).
// A NaN or -inf distance stops the pixel with status 1 (+inf is empty space).  Outputs are
// rectangle-local: accum4 (colour, 1), depth (-1 on a miss), aux4 (normal, steps; (0,0,0,-1)),
// status (0 done).
int orc_control_trace(const syf_job_desc* d, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t max_evals,
                      float* accum4, float* depth, float* aux4, uint32_t* status) {
    const Basis b = camera_basis(d->camera, d->width, d->height);
    const float cone = b.tan_half * 2.0f / float(d->width);
    const float t_end = ray_extent(*d, b);
    const Fractal<float> F = lower_fractal<float>(d->fractal);
    const V3<float> light = [] {
        const V3<float> l = v3(0.4f, 0.7f, 0.6f);
        return l * (1.0f / length(l));
    }();
    const V3<float> taps[4] = {v3(1.0f, -1.0f, -1.0f), v3(-1.0f, -1.0f, 1.0f), v3(-1.0f, 1.0f, -1.0f),
                               v3(1.0f, 1.0f, 1.0f)};
#pragma omp parallel for schedule(dynamic, 1)
    for (long yy = 0; yy < long(h); ++yy)
        for (uint32_t xx = 0; xx < w; ++xx) {
            const uint32_t px = x0 + xx, py = y0 + uint32_t(yy);
            const size_t i = size_t(yy) * w + xx;
            const float ndcx = ((float(px) + 0.5f) / float(d->width)) * 2.0f - 1.0f;
            const float ndcy = ((float(py) + 0.5f) / float(d->height)) * 2.0f - 1.0f;
            V3<float> rd = b.forward + b.right * (ndcx * b.tan_half) + b.up * (-ndcy * b.tan_half / b.aspect);
            rd = rd * (1.0f / length(rd));
            const V3<float> ro = b.origin;
            float* acc = accum4 + 4 * i;
            float* ax = aux4 + 4 * i;
            acc[0] = 0.025f, acc[1] = 0.03f, acc[2] = 0.04f, acc[3] = 1.0f;
            depth[i] = -1.0f;
            ax[0] = ax[1] = ax[2] = 0.0f, ax[3] = -1.0f;
            status[i] = 2u;
            float t = 0.0f, omega = 1.2f, prevr = 0.0f, step = 0.0f;
            uint32_t steps = 1u;
            bool hit = false, undefined = false;
            for (uint32_t e = 0; e < max_evals; ++e) {
                const float dist = de_eval(F, ro + rd * t).d;
                if (dist != dist || dist == -INFINITY) {
                    undefined = true;
                    break;
                }
                const float radius = std::fabs(dist) * F.safety;
                const bool failed = omega > 1.0f && radius + prevr < step * 1.0001f;
                if (failed) {
                    step = step - omega * step;
                    omega = 1.0f;
                } else {
                    step = radius * omega;
                }
                prevr = radius;
                const float err = radius / std::fmax(t, 1e-6f);
                if (!failed && err < cone) {
                    hit = true;
                    break;
                }
                bool miss = false;
                if (!failed && t + step >= t_end) {
                    if (radius >= t_end - t) {
                        miss = true;
                    } else {
                        omega = 1.0f;
                        step = radius;
                    }
                }
                if (!miss) {
                    t = t + step;
                    miss = t >= t_end;
                }
                if (miss) {
                    acc[0] = 0.025f, acc[1] = 0.035f, acc[2] = 0.05f;
                    status[i] = 0u;
                    break;
                }
                steps = std::min(steps, 0xfffffffeu) + 1u;
            }
            if (undefined) {
                status[i] = 1u;
                continue;
            }
            if (!hit) continue;
            const V3<float> hp = ro + rd * t;
            const float hh = std::fmax(t * cone * 0.5f, 1e-4f);
            V3<float> nacc = v3(0.0f, 0.0f, 0.0f);
            for (int k = 0; k < 4; ++k) {
                const float dist = de_eval(F, hp + taps[k] * hh).d;
                if (dist != dist || dist == -INFINITY) {
                    undefined = true;
                    break;
                }
                nacc = nacc + taps[k] * dist;
            }
            if (undefined) {
                status[i] = 1u;
                continue;
            }
            const float l = length(nacc), rcp = 1.0f / l;
            const V3<float> n = (l > 0.0f && rcp > 0.0f && rcp <= 3.4028235e38f) ? nacc * rcp : v3(0.0f, 1.0f, 0.0f);
            const float lit = 0.20f + 0.80f * std::fmax(dot(n, light), 0.0f);
            acc[0] = 0.68f * lit, acc[1] = 0.70f * lit, acc[2] = 0.73f * lit;
            depth[i] = t;
            ax[0] = n.x, ax[1] = n.y, ax[2] = n.z, ax[3] = float(steps);
            status[i] = 0u;
        }
    return 0;
}

// The fog's pieces for tests: optical depths along rays (ro3, rd3: n x 3) to dist[i]; free-flight
// distances along one ray with n independent PCG streams; the rotated Halton coordinate; the
// Henyey-Greenstein phase and its samples.
void orc_fog_optical_depth(const syf_job_desc* d, const float* ro3, const float* rd3, const float* dist, size_t n,
                           float* out) {
    const SceneL S = lower_scene(*d);
    for (size_t i = 0; i < n; ++i)
        out[i] = fog_optical_depth(S, v3(ro3[3 * i], ro3[3 * i + 1], ro3[3 * i + 2]),
                                   v3(rd3[3 * i], rd3[3 * i + 1], rd3[3 * i + 2]), dist[i]);
}
void orc_fog_free_flight(const syf_job_desc* d, const float* ro, const float* rd, float t_end, size_t n, float* out) {
    const SceneL S = lower_scene(*d);
    for (size_t i = 0; i < n; ++i) {
        uint32_t rng = pcg(uint32_t(i) * 2654435761u + 12345u);
        out[i] = fog_free_flight(S, v3(ro[0], ro[1], ro[2]), v3(rd[0], rd[1], rd[2]), t_end, &rng);
    }
}
// finish_relief at n points (xyz rows): out4 = slope xyz, height
void orc_finish_relief(uint32_t kind, uint32_t seed, float scale, const float* p, int n, float footprint,
                       float* out4) {
    for (int i = 0; i < n; ++i) {
        const Relief r = finish_relief(kind, seed, scale, v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), footprint);
        out4[4 * i] = r.g.x;
        out4[4 * i + 1] = r.g.y;
        out4[4 * i + 2] = r.g.z;
        out4[4 * i + 3] = r.h;
    }
}
float orc_volume_coordinate(uint32_t sample_index, uint32_t dimension, uint32_t stream) {
    return dimension < 6u ? volume_coordinate(sample_index, dimension, stream) : -1.0f;
}
float orc_hg_phase(float cos_theta, float g) { return hg_phase(cos_theta, g); }
void orc_hg_sample(float g, const float* dir, const float* r12, size_t n, float* out3) {
    for (size_t i = 0; i < n; ++i) {
        const F3 v = hg_sample(v3(dir[0], dir[1], dir[2]), g, r12[2 * i], r12[2 * i + 1]);
        out3[3 * i] = v.x;
        out3[3 * i + 1] = v.y;
        out3[3 * i + 2] = v.z;
    }
}

// palette_sample (common.wgsl:336-366) of a coloring's ramp at n values, rgb out; stop
// positions resolved as lower_scene does (Ramp::position, palette.rs:106-115).
void orc_ramp_sample(const syf_coloring* c, const float* values, size_t n, float* rgb) {
    syf_coloring r = *c;
    const uint32_t ns = std::min<uint32_t>(r.stop_count, SYF_MAX_STOPS);
    r.stop_count = ns;
    for (uint32_t i = 0; i < ns; ++i)
        if (!r.positioned) r.stops[i][3] = float(i) / float(r.cyclic ? ns : (ns > 1u ? ns - 1u : 1u));
    for (size_t i = 0; i < n; ++i) {
        const F3 v = ramp_sample(r, values[i]);
        rgb[3 * i] = v.x;
        rgb[3 * i + 1] = v.y;
        rgb[3 * i + 2] = v.z;
    }
}

// Megakernel accumulation (pathtrace.wgsl:1203-1276) of a pixel rectangle: samples
// sample_start .. sample_start+sample_count-1 in ascending order, accum += (r if finite, 1);
// depth/aux from the first sample of the call (written by the caller when sample_start == 0),
// and depth64: that sample's first surface hit refined in FP64 (refine64 between the march's
// last two positions, as the primary trace does), -1 on a miss or an emitter hit.
// Outputs are rectangle-local, row-major: accum4, depth, aux4, depth64.
// session4 (optional): the VOLUME, MOMENTS and SURFACE planes (3 x w x h float4, pathtrace.wgsl:
// 1243-1266), accumulated from their current values.
int orc_render(const syf_job_desc* d, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t sample_start,
               uint32_t sample_count, float* accum4, float* depth, float* aux4, double* depth64, float* session4) {
    SceneL S = lower_scene(*d);
    const Fractal<double> F64 = lower_fractal<double>(d->fractal);
    const double hit_cone = double(S.pixel_cone) * double(S.pixel_tol);
#pragma omp parallel for schedule(dynamic, 1)
    for (long yy = 0; yy < long(h); ++yy) {
        Count c;
        for (uint32_t xx = 0; xx < w; ++xx) {
            uint32_t px = x0 + xx, py = y0 + uint32_t(yy);
            size_t i = size_t(yy) * w + xx;
            float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            float fd = -1.0f, fa[4] = {0.0f, 0.0f, 0.0f, -1.0f};
            double f64 = -1.0;
            for (uint32_t s = 0; s < sample_count; ++s) {
                uint32_t rng = rng_init(px, py, sample_start + s, S.seed);
                float dd = -1.0f, aa[4] = {0.0f, 0.0f, 0.0f, -1.0f}, tp = -1.0f;
                F3 vol;
                F3 r = radiance(S, px, py, sample_start + s, &rng, &dd, aa, &tp, &vol, &c);
                bool ok = r.x == r.x && r.y == r.y && r.z == r.z && std::fabs(r.x) < 1e30f && std::fabs(r.y) < 1e30f &&
                          std::fabs(r.z) < 1e30f;
                acc[0] += ok ? r.x : 0.0f;
                acc[1] += ok ? r.y : 0.0f;
                acc[2] += ok ? r.z : 0.0f;
                acc[3] += 1.0f;
                if (session4) {  // Welford update of the volume, its moments, and the surface
                    float* V = session4 + 4 * i;
                    float* M = session4 + 4 * (size_t(w) * h + i);
                    float* Sf = session4 + 4 * (2 * size_t(w) * h + i);
                    const F3 v = ok ? vol : v3(0.0f, 0.0f, 0.0f);
                    const F3 surface = ok ? r - vol : v3(0.0f, 0.0f, 0.0f);
                    const float n = std::fmax(V[3], 1.0f);
                    const F3 old_mean = v3(V[0] / n, V[1] / n, V[2] / n);
                    const F3 delta = v - old_mean;
                    const float n1 = V[3] + 1.0f;
                    const F3 new_mean = old_mean + v3(delta.x / n1, delta.y / n1, delta.z / n1);
                    M[0] += std::fmax(0.0f, delta.x * (v.x - new_mean.x));
                    M[1] += std::fmax(0.0f, delta.y * (v.y - new_mean.y));
                    M[2] += std::fmax(0.0f, delta.z * (v.z - new_mean.z));
                    M[3] += 1.0f;
                    V[0] += v.x; V[1] += v.y; V[2] += v.z; V[3] += 1.0f;
                    Sf[0] += surface.x; Sf[1] += surface.y; Sf[2] += surface.z; Sf[3] += 1.0f;
                }
                if (s == 0) {
                    fd = dd;
                    for (int k = 0; k < 4; ++k) fa[k] = aa[k];
                    if (dd >= 0.0f && tp >= 0.0f) {
                        uint32_t rng0 = rng_init(px, py, sample_start, S.seed);
                        F3 ro, rd;
                        camera_ray(S.cam, S.width, S.height, px, py, &rng0, &ro, &rd);
                        f64 = refine64_of_march(S.F, F64, ro, rd, tp, dd, hit_cone, &c);
                    }
                }
            }
            for (int k = 0; k < 4; ++k) accum4[4 * i + k] = acc[k];
            depth[i] = fd;
            if (depth64) depth64[i] = f64;
            for (int k = 0; k < 4; ++k) aux4[4 * i + k] = fa[k];
        }
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

// FP32 direct evaluation of a pixel rectangle with the reference's f32 loop (planar.rs:
// 107-164: |z| by sqrt against the bailout, complex_power by repeated multiplication from
// 1, smooth count with ln of the power); c is the view point rounded to f32.
void orc_planar_f32(const syf_planar_desc* d, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t* iters,
                    float* smooth) {
    double* fx = new double[d->width];
    double* fy = new double[d->height];
    orc_planar_offsets(d, fx, fy);
    double rot = double(d->rotation) * (3.14159265358979323846 / 180.0);
    double cr = std::cos(rot), sr = std::sin(rot);
    const uint32_t power = 2;
#pragma omp parallel for schedule(dynamic, 1)
    for (long jj = 0; jj < long(h); ++jj)
        for (uint32_t ii = 0; ii < w; ++ii) {
            uint32_t i = x0 + ii, j = y0 + uint32_t(jj);
            double dx = cr * fx[i] - sr * fy[j];
            double dy = sr * fx[i] + cr * fy[j];
            float ptx = float(d->center[0] + dx), pty = float(d->center[1] + dy);
            bool pp = d->family != SYF_PLANAR_JULIA;
            float cx = pp ? ptx : float(d->julia[0]), cy = pp ? pty : float(d->julia[1]);
            float zx = pp ? 0.0f : ptx, zy = pp ? 0.0f : pty;
            uint32_t count = 0;
            bool escaped = false;
            for (uint32_t k = 0; k < d->iterations; ++k) {
                float radius = std::sqrt(zx * zx + zy * zy);
                if (radius > d->bailout) {
                    escaped = true;
                    break;
                }
                float bx = zx, by = zy;
                if (d->family == SYF_PLANAR_BURNING_SHIP) {
                    bx = std::fabs(bx);
                    by = std::fabs(by);
                }
                float wx = 1.0f, wy = 0.0f;
                for (uint32_t q = 0; q < power; ++q) {
                    float nx = wx * bx - wy * by, ny = wx * by + wy * bx;
                    wx = nx;
                    wy = ny;
                }
                zx = wx + cx;
                zy = wy + cy;
                ++count;
            }
            float radius = std::fmin(std::sqrt(zx * zx + zy * zy), 1e20f);
            escaped = escaped || radius > d->bailout;
            float s = float(d->iterations);
            if (escaped) s = float(count) + 1.0f - std::log(std::log(std::fmax(radius, 1.0001f))) / std::log(float(power));
            size_t k = size_t(jj) * w + ii;
            iters[k] = count;
            smooth[k] = s;
        }
    delete[] fx;
    delete[] fy;
}

// Quad-precision (binary128) direct evaluation at listed pixels, the deep-zoom pin of the
// perturbation mode: the pixel's c is centre + (float) offset, exactly the point whose
// orbit the FP32 deltas follow.
void orc_escape_f128(const syf_planar_desc* d, const uint32_t* px, const uint32_t* py, size_t n, double nudge,
                     uint32_t* iters, float* smooth) {
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
        // the f32 offset, optionally moved by `nudge` of its own ulps (a conditioning probe)
        const float fxo = float(dx), fyo = float(dy);
        const double ux = double(std::nextafter(std::fabs(fxo), INFINITY)) - std::fabs(double(fxo));
        const double uy = double(std::nextafter(std::fabs(fyo), INFINITY)) - std::fabs(double(fyo));
        __float128 cx = (__float128)d->center[0] + (__float128)fxo + (__float128)(nudge * ux);
        __float128 cy = (__float128)d->center[1] + (__float128)fyo + (__float128)(nudge * uy);
        Escape<__float128> e = escape_loop<__float128>(d->family, cx, cy, (__float128)0, (__float128)0, d->iterations, b2);
        iters[k] = e.count;
        double r = std::fmin(std::sqrt((double)e.radius2), 1e20);
        smooth[k] = e.escaped ? float(double(e.count) + 1.0 - std::log(std::log(std::fmax(r, 1.0001))) / std::log(2.0))
                              : float(d->iterations);
    }
    delete[] fx;
    delete[] fy;
}

// The same points as orc_escape_f128 (c = centre + the f32-rounded offset) iterated in FP64:
// how far a plain FP64 loop resolves them, the bar a perturbation method is held to.
void orc_escape_f64_points(const syf_planar_desc* d, const uint32_t* px, const uint32_t* py, size_t n,
                           uint32_t* iters) {
    double* fx = new double[d->width];
    double* fy = new double[d->height];
    orc_planar_offsets(d, fx, fy);
    double rot = double(d->rotation) * (3.14159265358979323846 / 180.0);
    double cr = std::cos(rot), sr = std::sin(rot);
    double b2 = double(d->bailout) * double(d->bailout);
#pragma omp parallel for schedule(dynamic, 1)
    for (long k = 0; k < long(n); ++k) {
        double dx = cr * fx[px[k]] - sr * fy[py[k]];
        double dy = sr * fx[px[k]] + cr * fy[py[k]];
        double cx = d->center[0] + double(float(dx));
        double cy = d->center[1] + double(float(dy));
        iters[k] = escape_loop<double>(d->family, cx, cy, 0.0, 0.0, d->iterations, b2).count;
    }
    delete[] fx;
    delete[] fy;
}

// Binary128 direct evaluation at c = centre + the pixel's FP64 offset itself (moved by `nudge`
// FP64 ulps of the offset, a conditioning probe): the point PERTURB64 follows, and the one a
// plain FP64 loop cannot represent once the pixel spacing nears an ulp of the centre.
void orc_escape_f128_exact(const syf_planar_desc* d, const uint32_t* px, const uint32_t* py, size_t n, double nudge,
                           uint32_t* iters, float* smooth) {
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
        const double ux = std::nextafter(std::fabs(dx), INFINITY) - std::fabs(dx);
        const double uy = std::nextafter(std::fabs(dy), INFINITY) - std::fabs(dy);
        __float128 cx = (__float128)d->center[0] + (__float128)dx + (__float128)(nudge * ux);
        __float128 cy = (__float128)d->center[1] + (__float128)dy + (__float128)(nudge * uy);
        Escape<__float128> e = escape_loop<__float128>(d->family, cx, cy, (__float128)0, (__float128)0, d->iterations, b2);
        iters[k] = e.count;
        double r = std::fmin(std::sqrt((double)e.radius2), 1e20);
        smooth[k] = e.escaped ? float(double(e.count) + 1.0 - std::log(std::log(std::fmax(r, 1.0001))) / std::log(2.0))
                              : float(d->iterations);
    }
    delete[] fx;
    delete[] fy;
}

// |c + d| - |c| without cancellation (the perturbation of |x|).
static double diffabs_f64(double c, double d) {
    const double cd = c + d;
    if (c >= 0.0) return cd >= 0.0 ? d : -(2.0 * c + d);
    return cd > 0.0 ? 2.0 * c + d : -d;
}

// Double-double arithmetic (Dekker 1971; Knuth's TwoSum): a value is hi + lo with |lo| <=
// ulp(hi)/2, about 106 bits.  Every step is an IEEE operation rounded as written (fma gives
// the exact product error), so any restatement of these steps gets the same bits.
struct DD {
    double hi, lo;
};
static DD dd_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return DD{s, e};
}
static DD dd_norm(double s, double e) {  // |s| >= |e|
    const double h = s + e;
    return DD{h, e - (h - s)};
}
static DD dd_add(DD a, DD b) {
    DD s = dd_two_sum(a.hi, b.hi);
    return dd_norm(s.hi, s.lo + (a.lo + b.lo));
}
static DD dd_mul(DD a, DD b) {
    const double p = a.hi * b.hi;
    const double e = std::fma(a.hi, b.hi, -p);
    return dd_norm(p, e + (a.hi * b.lo + a.lo * b.hi));
}
static DD dd_neg(DD a) { return DD{-a.hi, -a.lo}; }
static DD dd_abs(DD a) { return a.hi < 0.0 ? dd_neg(a) : a; }

// The reference orbit of the deep-zoom modes: C = centre + (ox, oy) held exactly as a
// double-double, Z_0 = 0, Z_{n+1} = f(Z_n) + C iterated in double-double (Burning Ship on
// |Re|, |Im|), each entry stored rounded to FP64, up to the first entry with
// |Z|^2 > bailout^2 (FP64 test on the rounded entry) or the iteration budget.
void orc_reference_orbit_dd(const syf_planar_desc* d, double ox, double oy, std::vector<double>& ZX,
                            std::vector<double>& ZY) {
    const double b2 = double(d->bailout) * double(d->bailout);
    const bool ship = d->family == SYF_PLANAR_BURNING_SHIP;
    const DD cx = dd_two_sum(d->center[0], ox), cy = dd_two_sum(d->center[1], oy);
    DD zx{0.0, 0.0}, zy{0.0, 0.0};
    ZX.assign(1, 0.0);
    ZY.assign(1, 0.0);
    for (uint32_t n = 0; n < d->iterations; ++n) {
        const DD x = ship ? dd_abs(zx) : zx, y = ship ? dd_abs(zy) : zy;
        const DD xx = dd_mul(x, x), yy = dd_mul(y, y), xy = dd_mul(x, y);
        zx = dd_add(dd_add(xx, dd_neg(yy)), cx);
        zy = dd_add(DD{2.0 * xy.hi, 2.0 * xy.lo}, cy);
        ZX.push_back(zx.hi);
        ZY.push_back(zy.hi);
        if (zx.hi * zx.hi + zy.hi * zy.hi > b2) break;
    }
}

// C-ABI views of the reference orbit for tests: the double-double orbit's FP64 entries, and
// the same recurrence in binary128 rounded to FP64 (its pin).  Both return the index of the
// last entry written (at most cap - 1).
uint32_t orc_reference_orbit(const syf_planar_desc* d, double ox, double oy, double* zx, double* zy, uint32_t cap) {
    std::vector<double> ZX, ZY;
    orc_reference_orbit_dd(d, ox, oy, ZX, ZY);
    const uint32_t n = uint32_t(std::min<size_t>(ZX.size(), cap));
    for (uint32_t i = 0; i < n; ++i) {
        zx[i] = ZX[i];
        zy[i] = ZY[i];
    }
    return n - 1;
}
uint32_t orc_reference_orbit_f128(const syf_planar_desc* d, double ox, double oy, double* zx, double* zy,
                                  uint32_t cap) {
    const __float128 b2 = (__float128)d->bailout * (__float128)d->bailout;
    const __float128 cx = (__float128)d->center[0] + (__float128)ox, cy = (__float128)d->center[1] + (__float128)oy;
    __float128 x = 0, y = 0;
    uint32_t n = 0;
    zx[0] = zy[0] = 0.0;
    while (n + 1 < cap && n < d->iterations) {
        __float128 a = x, b = y;
        if (d->family == SYF_PLANAR_BURNING_SHIP) {
            a = a < 0 ? -a : a;
            b = b < 0 ? -b : b;
        }
        const __float128 nx = a * a - b * b + cx, ny = 2 * a * b + cy;
        x = nx;
        y = ny;
        ++n;
        zx[n] = (double)x;
        zy[n] = (double)y;
        if (x * x + y * y > b2) break;
    }
    return n;
}

// The PERTURB64 loop at listed pixels: the view centre's reference orbit Z
// (orc_reference_orbit_dd), FP64 deltas delta <- (2 Z + delta) delta + delta_c with
// delta_c = the pixel's FP64 offset (Burning Ship: the |.| products through diffabs), and
// rebasing (Zhuoran 2021): when |Z_m + delta| < |delta| or Z has escaped, delta <- Z_m + delta
// and m <- 0.  Each operation rounds as written.  The deep-zoom mode has no counterpart in the
// reference (planar.rs:77-194 iterates f32 directly); this loop is pinned by binary128 at the
// exact pixel points (orc_escape_f128_exact) and by the FP64 direct loop where that resolves.
void orc_perturb64(const syf_planar_desc* d, const uint32_t* px, const uint32_t* py, size_t n, uint32_t* iters,
                   float* smooth) {
    double* fx = new double[d->width];
    double* fy = new double[d->height];
    orc_planar_offsets(d, fx, fy);
    double rot = double(d->rotation) * (3.14159265358979323846 / 180.0);
    double cr = std::cos(rot), sr = std::sin(rot);
    const double b2 = double(d->bailout) * double(d->bailout);
    const bool ship = d->family == SYF_PLANAR_BURNING_SHIP;
    std::vector<double> ZX, ZY;
    orc_reference_orbit_dd(d, 0.0, 0.0, ZX, ZY);
    const uint32_t last = uint32_t(ZX.size() - 1);
#pragma omp parallel for schedule(dynamic, 1)
    for (long k = 0; k < long(n); ++k) {
        const double dcx = cr * fx[px[k]] - sr * fy[py[k]];
        const double dcy = sr * fx[px[k]] + cr * fy[py[k]];
        double ex = 0.0, ey = 0.0, zx = 0.0, zy = 0.0;
        uint32_t m = 0, count = 0;
        bool escaped = false;
        for (uint32_t i = 0; i < d->iterations; ++i) {
            double X = ZX[m], Y = ZY[m];
            zx = X + ex;
            zy = Y + ey;
            const double r2 = zx * zx + zy * zy;
            if (r2 > b2) {
                escaped = true;
                break;
            }
            if (m == last || r2 < ex * ex + ey * ey) {
                ex = zx;
                ey = zy;
                m = 0;
                X = 0.0;
                Y = 0.0;
            }
            double nx, ny;
            if (ship) {
                nx = (2.0 * X + ex) * ex - (2.0 * Y + ey) * ey + dcx;
                const double XY = X * Y;
                const double dXY = (X + ex) * ey + ex * Y;
                ny = 2.0 * diffabs_f64(XY, dXY) + dcy;
            } else {
                const double tx = 2.0 * X + ex, ty = 2.0 * Y + ey;
                nx = tx * ex - ty * ey + dcx;
                ny = tx * ey + ty * ex + dcy;
            }
            ex = nx;
            ey = ny;
            ++m;
            ++count;
        }
        if (!escaped) {
            zx = ZX[m] + ex;
            zy = ZY[m] + ey;
        }
        const double r2 = zx * zx + zy * zy;
        escaped = escaped || r2 > b2;
        iters[k] = count;
        smooth[k] = escaped ? float(double(count) + 1.0 -
                                    std::log(std::log(std::fmax(std::fmin(std::sqrt(r2), 1e20), 1.0001))) / std::log(2.0))
                            : float(d->iterations);
    }
    delete[] fx;
    delete[] fy;
}

}  // extern "C"
