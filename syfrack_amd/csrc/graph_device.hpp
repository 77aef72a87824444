// Device interpreter of a compiled node graph (kernel family FAM_GRAPH).
//
// The reference generates one WGSL function per graph topology (graph_wgsl.rs:683-949); this
// library compiles a graph once, on the host, into a short register program (graph.hip) and
// every field evaluation of a graph job runs it here.  The program is wave-uniform: each lane
// evaluates its own point, every lane of a wave walks the same instructions, and the
// instruction stream is read with scalar loads.  Semantics follow eval.rs (Graph::evaluate):
//
//   * one current point state `cur` (PointState: p, dr, iter, traps) and one current distance
//     state `dc` (DistState: d, material, iter, traps); values that several nodes consume are
//     kept in a few point and distance slots;
//   * an Iterate is a structured loop (ITER_BEGIN ... ITER_END) over its body's point
//     operations applied to `cur` in body order; a lane whose orbit has passed the bailout, or
//     has run its iterations, stops executing the body (it keeps `cur`) while the wave goes on
//     until every lane has stopped (eval.rs:517-532);
//   * AddSeed adds the point that entered the innermost enclosing Iterate (the node's own input
//     outside one), AddOrigin and AddConstant the evaluated point c (eval.rs:477-491);
//   * distances, combinations and the trap normalisation as eval.rs:558-657, 98-114.
//
// Constants that the reference recomputes per evaluation from node parameters alone (squared
// radii, their quotient, Menger offsets times (scale - 1), the Kaleido sector, radians, the
// rotation quaternion, normalised normals) are folded on the host with the same f32
// operations and stored beside the raw parameters: the FP32 interpreter reads the folded
// values, which are the ones the reference computes, and the FP64 interpreter (final-hit
// refine) recomputes the products in double from the raw f32 parameters.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace syf {

enum GOpCode : uint32_t {
    GOP_END = 0,          // the result is dc
    GOP_P_START,          // cur = (c, dr 1, iter 0, traps +inf)
    GOP_P_SAVE,           // point slot a = cur
    GOP_P_LOAD,           // cur = point slot a
    GOP_ITER_BEGIN,       // a = level, b = iterations, c = index of the matching ITER_END; v0 = bailout
    GOP_ITER_END,         // a = level, b = iterations, c = index of the first body instruction; v0 = bailout
    GOP_TRANSLATE,        // v0..2 offset
    GOP_ROTATE,           // v0..3 quaternion (x, y, z, w)
    GOP_SCALE,            // v0 factor, v1 |factor|
    GOP_SCALE_AXES,       // v0..2 factors, v3 their maximum
    GOP_MIRROR,           // v0..2 unit normal, v3 offset
    GOP_BOX_FOLD,         // v0 limit
    GOP_BOX_FOLD_AXES,    // v0..2 limits, a = axis mask
    GOP_SPHERE_INVERSION, // v0 radius, v1 min_radius; f32 folds: v2 radius^2, v3 min_radius^2
    GOP_SPHERE_FOLD,      // v0 min_radius, v1 fixed_radius; f32 folds: v2 mr2, v3 fr2, v4 fr2 / mr2
    GOP_ABS,              // a = axis mask
    GOP_MENGER_FOLD,      // v0 scale, v1..3 offset; f32 folds: v4..6 offset * (scale - 1), v7 shift threshold
    GOP_KALEIDO,          // v0 sector (f32 TAU / planes), v1 phase in radians (f32)
    GOP_POWER,            // v0 power, a = integer power 1..16 or 0 for the trigonometric form
    GOP_ADD_ORIGIN,
    GOP_ADD_SEED,         // v0..2 multiplier, v3 max |multiplier|, a = seed level + 1 (0: own input)
    GOP_ADD_CONSTANT,     // v0..2 multiplier, v3..5 julia, v6 max |multiplier|
    GOP_TWIST,            // v0 rate
    GOP_BEND,             // v0 rate
    GOP_TILE,             // v0..2 period
    GOP_REPEAT_FINITE,    // v0..2 period, a, b, c counts
    GOP_TRAP,             // a = slot (0..3), b = measure, c = mode
    GOP_FENCE,            // a = SYF_NODE_* fence kind, or SYF_NODE_KINDS for a disabled fence; v = parameters
    GOP_D_SAVE,           // distance slot a = dc
    GOP_D_LOAD,           // dc = distance slot a
    GOP_MATTAG,           // dc.material = a
    GOP_BOUNDED,          // dc.d = max(dc.d, box distance of cur, half v0..2)
    GOP_COMBINE,          // a = SYF_NODE_* combination, b / c = slot of operand A / B or GRAPH_ACC (dc); v0 k
    GOP_COUNT,
};

struct GOp {
    uint32_t op, a, b, c;
    float v[8];
};

constexpr uint32_t GRAPH_PSLOTS = 4;  // point values kept for later consumers
constexpr uint32_t GRAPH_DSLOTS = 4;  // distance values kept for later consumers
constexpr uint32_t GRAPH_LEVELS = 3;  // nested Iterates (MAX_DEPTH, scene/graph/mod.rs:42)
constexpr uint32_t GRAPH_ACC = 0xffu;
constexpr uint32_t GRAPH_MAX_OPS = 4096;
// Instructions one evaluation can execute: every instruction once per run of its enclosing
// loops, at most the iteration product (256) plus the final check per loop.  A bound for the
// interpreter's loop, never reached by a verified program.
constexpr uint32_t GRAPH_MAX_STEPS = GRAPH_MAX_OPS * 260u;

template <class T> struct GPoint {
    V3<T> p;
    T dr;
    uint32_t iter;
    T t0, t1, t2, t3;
};
template <class T> struct GDist {
    T d;
    uint32_t mat, iter;
    T t0, t1, t2, t3;
};
// Raw result of a graph evaluation (before the trap normalisation) and the Iterate-body runs.
template <class T> struct GRaw {
    GDist<T> r;
    uint32_t runs;
};

template <class T> __device__ __forceinline__ T g_inf() { return T(INFINITY); }
// Rust clamp(x, 0, 1): NaN stays NaN
template <class T> __device__ __forceinline__ T clamp01(T x) {
    T h = x < T(0) ? T(0) : x;
    return h > T(1) ? T(1) : h;
}
template <class T> __device__ __forceinline__ T len2d(T x, T y) { return sqrt(x * x + y * y); }
// Rust rem_euclid for floats
template <class T> __device__ __forceinline__ T rem_euclid(T a, T b) {
    T r = fmod(a, b);
    return r < T(0) ? r + fabs(b) : r;
}

// measure (eval.rs:309-331)
template <class T> __device__ __forceinline__ T g_measure(uint32_t m, V3<T> p) {
    const T tau = T(6.283185307179586);
    switch (m) {
    case SYF_MEASURE_NORM: return len3(p);
    case SYF_MEASURE_AXIS_X: return len2d(p.y, p.z);
    case SYF_MEASURE_AXIS_Y: return len2d(p.x, p.z);
    case SYF_MEASURE_AXIS_Z: return len2d(p.x, p.y);
    case SYF_MEASURE_PLANE_Y: return fabs(p.y);
    case SYF_MEASURE_ANGLE_XY: return (p.x == T(0) && p.y == T(0)) ? T(0.5) : atan2(p.y, p.x) / tau + T(0.5);
    case SYF_MEASURE_ANGLE_XZ: return (p.x == T(0) && p.z == T(0)) ? T(0.5) : atan2(p.z, p.x) / tau + T(0.5);
    case SYF_MEASURE_ANGLE_YZ: return (p.y == T(0) && p.z == T(0)) ? T(0.5) : atan2(p.z, p.y) / tau + T(0.5);
    default: return log1p(len3(p));
    }
}

// terrain_hash / terrain_noise / rock_distance (eval.rs:235-273)
__device__ __forceinline__ uint32_t g_terrain_hash(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    uint32_t h = (uint32_t(x) * 0x8da6b343u) ^ (uint32_t(y) * 0xd8163841u) ^ (uint32_t(z) * 0xcb1ab31fu) ^ seed;
    h = (h ^ (h >> 16)) * 0x7feb352du;
    h = (h ^ (h >> 15)) * 0x846ca68bu;
    return h ^ (h >> 16);
}
// Rust `as i32` from a float: saturating, NaN -> 0
template <class T> __device__ __forceinline__ int32_t g_sat_i32(T v) {
    if (!(v == v)) return 0;
    if (v <= T(-2147483648.0)) return INT32_MIN;
    if (v >= T(2147483647.0)) return INT32_MAX;
    return int32_t(v);
}
template <class T> __device__ __forceinline__ T g_noise(V3<T> p, uint32_t seed) {
    const V3<T> base = mk(floor(p.x), floor(p.y), floor(p.z));
    const V3<T> f = sub(p, base);
    const T tx = f.x * f.x * (T(3) - T(2) * f.x), ty = f.y * f.y * (T(3) - T(2) * f.y),
            tz = f.z * f.z * (T(3) - T(2) * f.z);
    const uint32_t cx = uint32_t(g_sat_i32(base.x)), cy = uint32_t(g_sat_i32(base.y)), cz = uint32_t(g_sat_i32(base.z));
    auto h = [&](uint32_t x, uint32_t y, uint32_t z) {
        return T(g_terrain_hash(int32_t(cx + x), int32_t(cy + y), int32_t(cz + z), seed) >> 8) / T(16777216.0f);
    };
    auto mix = [](T a, T b, T w) { return a + (b - a) * w; };
    const T a = mix(h(0, 0, 0), h(1, 0, 0), tx);
    const T b = mix(h(0, 1, 0), h(1, 1, 0), tx);
    const T d = mix(h(0, 0, 1), h(1, 0, 1), tx);
    const T e = mix(h(0, 1, 1), h(1, 1, 1), tx);
    return mix(mix(a, b, ty), mix(d, e, ty), tz);
}
template <class T> __device__ __forceinline__ T g_rock(V3<T> p, V3<T> half, T erosion, uint32_t seed) {
    const T h = fmin(half.x, fmin(half.y, half.z));
    const V3<T> q = scl(p, T(1.15f) / h);
    const T relief = T(0.57f) * g_noise(q, seed) + T(0.29f) * g_noise(scl(q, T(2.03f)), seed + 97u) +
                     T(0.14f) * g_noise(scl(q, T(4.17f)), seed + 193u) - T(0.5f);
    const T l = len3(mk(p.x / half.x, p.y / half.y, p.z / half.z));
    return (l - T(1) - erosion * relief) * h / (T(1) + T(5.22f) * erosion);
}

template <class T> __device__ __forceinline__ T g_box(V3<T> p, T hx, T hy, T hz) {
    const T qx = fabs(p.x) - hx, qy = fabs(p.y) - hy, qz = fabs(p.z) - hz;
    return len3(mk(fmax(qx, T(0)), fmax(qy, T(0)), fmax(qz, T(0)))) + fmin(fmax(qx, fmax(qy, qz)), T(0));
}

// fence (eval.rs:558-599) of kind `kind` (SYF_NODE_KINDS: a disabled fence, +inf)
template <class T> __device__ __forceinline__ T g_fence(uint32_t kind, const GOp& o, const GPoint<T>& s) {
    const V3<T> p = s.p;
    switch (kind) {
    case SYF_NODE_SPHERE: return (len3(p) - T(o.v[0])) / s.dr;
    case SYF_NODE_TORUS: return (len2d(len2d(p.x, p.z) - T(o.v[0]), p.y) - T(o.v[1])) / s.dr;
    case SYF_NODE_CYLINDER: {
        const T qx = len2d(p.x, p.z) - T(o.v[0]), qy = fabs(p.y) - T(o.v[1]);
        return (len2d(fmax(qx, T(0)), fmax(qy, T(0))) + fmin(fmax(qx, qy), T(0))) / s.dr;
    }
    case SYF_NODE_BOX: return g_box(p, T(o.v[0]), T(o.v[1]), T(o.v[2])) / s.dr;
    case SYF_NODE_ROUND_BOX: {
        const T hx = T(o.v[0]), hy = T(o.v[1]), hz = T(o.v[2]);
        const T r = fmin(T(o.v[3]), fmin(hx, fmin(hy, hz)));
        const T qx = fabs(p.x) - (hx - r), qy = fabs(p.y) - (hy - r), qz = fabs(p.z) - (hz - r);
        return (len3(mk(fmax(qx, T(0)), fmax(qy, T(0)), fmax(qz, T(0)))) + fmin(fmax(qx, fmax(qy, qz)), T(0)) - r) /
               s.dr;
    }
    case SYF_NODE_ROCK: return g_rock(p, mk(T(o.v[0]), T(o.v[1]), T(o.v[2])), T(o.v[3]), o.b) / s.dr;
    case SYF_NODE_PLANE: return (dot3(p, mk(T(o.v[0]), T(o.v[1]), T(o.v[2]))) - T(o.v[3])) / s.dr;
    case SYF_NODE_NORM_FENCE: return (len3(p) - T(o.v[0])) / fabs(s.dr);
    case SYF_NODE_BULB_FENCE: {
        const T r = fmax(len3(p), T(1e-12));
        return T(0.5) * log(r) * r / s.dr;
    }
    default: return g_inf<T>();
    }
}

template <class T> __device__ __forceinline__ T g_trap_update(T t, T m, uint32_t mode) {
    if (mode == SYF_TRAP_MINIMUM) return fmin(t, m);
    if (mode == SYF_TRAP_MAXIMUM) return t >= T(1e30) ? m : fmax(t, m);
    return m;
}

// combine (eval.rs:601-657)
template <class T> __device__ __forceinline__ T g_blend_trap(T a, T b, T h) {
    const bool fa = isfinite(a), fb = isfinite(b);
    return (fa && fb) ? b + (a - b) * h : fa ? a : fb ? b : g_inf<T>();
}
template <class T> __device__ __forceinline__ GDist<T> g_combine(uint32_t kind, T k, const GDist<T>& A, const GDist<T>& B) {
    switch (kind) {
    case SYF_NODE_UNION: {
        GDist<T> r = A.d <= B.d ? A : B;
        r.d = fmin(A.d, B.d);
        return r;
    }
    case SYF_NODE_INTERSECTION: {
        GDist<T> r = A.d >= B.d ? A : B;
        r.d = fmax(A.d, B.d);
        return r;
    }
    case SYF_NODE_SUBTRACTION: {
        GDist<T> r = A;
        r.d = fmax(A.d, -B.d);
        return r;
    }
    default: {  // smooth union / intersection
        const bool uni = kind == SYF_NODE_SMOOTH_UNION;
        if (A.d >= T(1e30) || B.d >= T(1e30)) {
            GDist<T> r = A.d >= T(1e30) ? B : A;
            if (!uni) r.d = g_inf<T>();
            return r;
        }
        const T h = uni ? clamp01(T(0.5) + T(0.5) * (B.d - A.d) / k) : clamp01(T(0.5) - T(0.5) * (B.d - A.d) / k);
        const T base = B.d + (A.d - B.d) * h;
        const T bump = k * h * (T(1) - h);
        GDist<T> r;
        r.d = uni ? base - bump : base + bump;
        r.mat = h >= T(0.5) ? A.mat : B.mat;
        r.iter = h >= T(0.5) ? A.iter : B.iter;
        r.t0 = g_blend_trap(A.t0, B.t0, h);
        r.t1 = g_blend_trap(A.t1, B.t1, h);
        r.t2 = g_blend_trap(A.t2, B.t2, h);
        r.t3 = g_blend_trap(A.t3, B.t3, h);
        return r;
    }
    }
}

// One point operation (eval.rs:346-556) on the current point.  `op` and the integer fields
// a, b, cc are passed separately from the instruction so that specialised code passes them
// as constants (the branches below then fold away); sp, sdr are the AddSeed source.
template <class T>
__device__ __forceinline__ void g_point(uint32_t op, uint32_t a, uint32_t b, uint32_t cc, const GOp& o,
                                        GPoint<T>& cur, V3<T> c, V3<T> sp, T sdr) {
    switch (op) {
    case GOP_TRANSLATE: cur.p = add(cur.p, mk(T(o.v[0]), T(o.v[1]), T(o.v[2]))); break;
    case GOP_ROTATE: {
        const V3<T> bq = mk(T(o.v[0]), T(o.v[1]), T(o.v[2]));
        const T w = T(o.v[3]);
        const V3<T> v = cur.p;
        cur.p = add(add(scl(v, w * w - dot3(bq, bq)), scl(bq, dot3(v, bq) * T(2))), scl(cross3(bq, v), w * T(2)));
        break;
    }
    case GOP_SCALE:
        cur.p = scl(cur.p, T(o.v[0]));
        cur.dr = cur.dr * T(o.v[1]);
        break;
    case GOP_SCALE_AXES:
        cur.p = mk(cur.p.x * T(o.v[0]), cur.p.y * T(o.v[1]), cur.p.z * T(o.v[2]));
        cur.dr = cur.dr * T(o.v[3]);
        break;
    case GOP_MIRROR: {
        const V3<T> n = mk(T(o.v[0]), T(o.v[1]), T(o.v[2]));
        const T h = dot3(cur.p, n) - T(o.v[3]);
        if (h < T(0)) cur.p = sub(cur.p, scl(n, T(2) * h));
        break;
    }
    case GOP_BOX_FOLD: {
        const T l = T(o.v[0]);
        cur.p = mk(clampT(cur.p.x, -l, l) * T(2) - cur.p.x, clampT(cur.p.y, -l, l) * T(2) - cur.p.y,
                   clampT(cur.p.z, -l, l) * T(2) - cur.p.z);
        break;
    }
    case GOP_BOX_FOLD_AXES: {
        const T lx = T(o.v[0]), ly = T(o.v[1]), lz = T(o.v[2]);
        if (a & 1u) cur.p.x = clampT(cur.p.x, -lx, lx) * T(2) - cur.p.x;
        if (a & 2u) cur.p.y = clampT(cur.p.y, -ly, ly) * T(2) - cur.p.y;
        if (a & 4u) cur.p.z = clampT(cur.p.z, -lz, lz) * T(2) - cur.p.z;
        break;
    }
    case GOP_SPHERE_INVERSION: {
        T r2, m2;
        if constexpr (sizeof(T) == 4) {
            r2 = o.v[2];
            m2 = o.v[3];
        } else {
            r2 = T(o.v[0]) * T(o.v[0]);
            m2 = T(o.v[1]) * T(o.v[1]);
        }
        const T k = r2 / fmax(dot3(cur.p, cur.p), m2);
        cur.p = scl(cur.p, k);
        cur.dr = cur.dr * k;
        break;
    }
    case GOP_SPHERE_FOLD: {
        T mr2, fr2, q;
        if constexpr (sizeof(T) == 4) {
            mr2 = o.v[2];
            fr2 = o.v[3];
            q = o.v[4];
        } else {
            mr2 = T(o.v[0]) * T(o.v[0]);
            fr2 = T(o.v[1]) * T(o.v[1]);
            q = fr2 / mr2;
        }
        const T r2 = dot3(cur.p, cur.p);
        const T k = r2 < mr2 ? q : (r2 < fr2 ? fr2 / r2 : T(1));
        cur.p = scl(cur.p, k);
        cur.dr = cur.dr * k;
        break;
    }
    case GOP_ABS:
        if (a & 1u) cur.p.x = fabs(cur.p.x);
        if (a & 2u) cur.p.y = fabs(cur.p.y);
        if (a & 4u) cur.p.z = fabs(cur.p.z);
        break;
    case GOP_MENGER_FOLD: {
        const T s = T(o.v[0]);
        T ox, oy, oz, thr;
        if constexpr (sizeof(T) == 4) {
            ox = o.v[4];
            oy = o.v[5];
            oz = o.v[6];
            thr = o.v[7];
        } else {
            const T s1 = s - T(1);
            ox = T(o.v[1]) * s1;
            oy = T(o.v[2]) * s1;
            oz = T(o.v[3]) * s1;
            thr = T(-0.5) * T(o.v[3]) * s1;
        }
        T wx = fabs(cur.p.x), wy = fabs(cur.p.y), wz = fabs(cur.p.z), tmp;
        if (wx < wy) { tmp = wx; wx = wy; wy = tmp; }
        if (wx < wz) { tmp = wx; wx = wz; wz = tmp; }
        if (wy < wz) { tmp = wy; wy = wz; wz = tmp; }
        wx = wx * s - ox;
        wy = wy * s - oy;
        wz = wz * s - oz;
        if (wz < thr) wz = wz + oz;
        cur.p = mk(wx, wy, wz);
        cur.dr = cur.dr * s;
        break;
    }
    case GOP_KALEIDO: {
        const T sector = T(o.v[0]);
        T ang = atan2(cur.p.z, cur.p.x) + T(o.v[1]);
        ang = fabs(rem_euclid(ang + sector * T(0.5), sector) - sector * T(0.5));
        const T r = len2d(cur.p.x, cur.p.z);
        cur.p.x = r * cos(ang);
        cur.p.z = r * sin(ang);
        break;
    }
    case GOP_POWER: {
        const T radius = len3(cur.p);
        if (a != 0u && radius >= T(1e-12)) {
            T deriv;
            spherical_power(cur.p, radius, a, deriv);
            cur.dr = cur.dr * deriv;
        } else {
            trig_power(cur.p, radius, T(o.v[0]), cur.dr);
        }
        break;
    }
    case GOP_ADD_ORIGIN:
        cur.p = add(cur.p, c);
        cur.dr = cur.dr + T(1);
        break;
    case GOP_ADD_SEED:
        cur.p = add(cur.p, mk(sp.x * T(o.v[0]), sp.y * T(o.v[1]), sp.z * T(o.v[2])));
        cur.dr = cur.dr + sdr * T(o.v[3]);
        break;
    case GOP_ADD_CONSTANT:
        cur.p = add(cur.p, mk(c.x * T(o.v[0]) + T(o.v[3]), c.y * T(o.v[1]) + T(o.v[4]), c.z * T(o.v[2]) + T(o.v[5])));
        cur.dr = cur.dr + T(o.v[6]);
        break;
    case GOP_TWIST: {
        const T ang = T(o.v[0]) * cur.p.y;
        const T sn = sin(ang), cs = cos(ang);
        const T x = cur.p.x, z = cur.p.z;
        cur.p.x = cs * x - sn * z;
        cur.p.z = sn * x + cs * z;
        break;
    }
    case GOP_BEND: {
        const T ang = T(o.v[0]) * cur.p.x;
        const T sn = sin(ang), cs = cos(ang);
        const T x = cur.p.x, y = cur.p.y;
        cur.p.x = cs * x - sn * y;
        cur.p.y = sn * x + cs * y;
        break;
    }
    case GOP_TILE: {
        const T px = T(o.v[0]), py = T(o.v[1]), pz = T(o.v[2]);
        const T hx = px * T(0.5), hy = py * T(0.5), hz = pz * T(0.5);
        cur.p = mk(rem_euclid(cur.p.x + hx, px) - hx, rem_euclid(cur.p.y + hy, py) - hy, rem_euclid(cur.p.z + hz, pz) - hz);
        break;
    }
    case GOP_REPEAT_FINITE: {
        const T px = T(o.v[0]), py = T(o.v[1]), pz = T(o.v[2]);
        const T lx = T(float(o.a)) - T(1), ly = T(float(o.b)) - T(1), lz = T(float(o.c)) - T(1);
        const T hx = lx * T(0.5), hy = ly * T(0.5), hz = lz * T(0.5);
        const T cx = clampT(floor(cur.p.x / px + hx + T(0.5)), T(0), lx) - hx;
        const T cy = clampT(floor(cur.p.y / py + hy + T(0.5)), T(0), ly) - hy;
        const T cz = clampT(floor(cur.p.z / pz + hz + T(0.5)), T(0), lz) - hz;
        cur.p = sub(cur.p, mk(cx * px, cy * py, cz * pz));
        break;
    }
    case GOP_TRAP: {
        const T m = g_measure(b, cur.p);
        if (a == 0u) cur.t0 = g_trap_update(cur.t0, m, cc);
        else if (a == 1u) cur.t1 = g_trap_update(cur.t1, m, cc);
        else if (a == 2u) cur.t2 = g_trap_update(cur.t2, m, cc);
        else cur.t3 = g_trap_update(cur.t3, m, cc);
        break;
    }
    default: break;
    }
}

// The distance state of a fence of kind `kind` at the current point.
template <class T> __device__ __forceinline__ GDist<T> g_fence_state(uint32_t kind, const GOp& o, const GPoint<T>& s) {
    return GDist<T>{g_fence(kind, o, s), 0u, s.iter, s.t0, s.t1, s.t2, s.t3};
}

// Bounded: the branch distance clipped to the box of the point (eval.rs:722-731).
template <class T> __device__ __forceinline__ void g_bounded(const GOp& o, const GPoint<T>& s, GDist<T>& d) {
    d.d = fmax(d.d, g_box(s.p, T(o.v[0]), T(o.v[1]), T(o.v[2])) / s.dr);
}

// A 64-bit pointer made wave-uniform (its instructions are then read with scalar loads).
__device__ __forceinline__ const GOp* g_uniform(const GOp* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v)), hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return reinterpret_cast<const GOp*>((uint64_t(hi) << 32) | lo);
}

// Slot access by a wave-uniform index: one branch per slot.
template <class S, uint32_t N> __device__ __forceinline__ void g_store(S (&slots)[N], uint32_t i, const S& v) {
#pragma unroll
    for (uint32_t k = 0; k < N; ++k)
        if (i == k) slots[k] = v;
}
template <class S, uint32_t N> __device__ __forceinline__ S g_load(const S (&slots)[N], uint32_t i) {
    S v = slots[0];
#pragma unroll
    for (uint32_t k = 1; k < N; ++k)
        if (i == k) v = slots[k];
    return v;
}

#define SYF_GRAPH_POINT_OPS(X) \
    X(GOP_TRANSLATE) X(GOP_ROTATE) X(GOP_SCALE) X(GOP_SCALE_AXES) X(GOP_MIRROR) X(GOP_BOX_FOLD) X(GOP_BOX_FOLD_AXES) \
    X(GOP_SPHERE_INVERSION) X(GOP_SPHERE_FOLD) X(GOP_ABS) X(GOP_MENGER_FOLD) X(GOP_KALEIDO) X(GOP_POWER)          \
    X(GOP_ADD_ORIGIN) X(GOP_ADD_CONSTANT) X(GOP_TWIST) X(GOP_BEND) X(GOP_TILE) X(GOP_REPEAT_FINITE) X(GOP_TRAP)

// Whether a combination's result carries operand A's payload (material, iterations, and the
// fence it came from): the choices g_combine makes.
template <class T> __device__ __forceinline__ bool g_combine_takes_a(uint32_t kind, T k, const GDist<T>& A,
                                                                     const GDist<T>& B) {
    switch (kind) {
    case SYF_NODE_UNION: return A.d <= B.d;
    case SYF_NODE_INTERSECTION: return A.d >= B.d;
    case SYF_NODE_SUBTRACTION: return true;
    default: {
        if (A.d >= T(1e30) || B.d >= T(1e30)) return !(A.d >= T(1e30));
        const bool uni = kind == SYF_NODE_SMOOTH_UNION;
        const T h = uni ? clamp01(T(0.5) + T(0.5) * (B.d - A.d) / k) : clamp01(T(0.5) - T(0.5) * (B.d - A.d) / k);
        return h >= T(0.5);
    }
    }
}

// Interprets a graph program at c.  The program is valid by construction and checked on the
// host (graph.hip): every ITER_BEGIN/ITER_END pair is properly nested within GRAPH_LEVELS,
// every slot index is below the slot count, and the program ends with GOP_END.  With TRACK,
// it also follows which fence produced the result's distance (node_ids: the node of each
// instruction) and writes that node to *node_out: the diagnosis of an undefined distance.
template <class T, bool TRACK = false>
__device__ __noinline__ GRaw<T> graph_run(const GOp* prog_in, V3<T> c, const uint32_t* node_ids = nullptr,
                                          uint32_t* node_out = nullptr) {
    const GOp* prog = g_uniform(prog_in);
    const T INF = g_inf<T>();
    GPoint<T> cur{c, T(1), 0u, INF, INF, INF, INF};
    GDist<T> dc{INF, 0u, 0u, INF, INF, INF, INF};
    GPoint<T> ps[GRAPH_PSLOTS];
    GDist<T> ds[GRAPH_DSLOTS];
    V3<T> seed_p[GRAPH_LEVELS];
    T seed_dr[GRAPH_LEVELS];
    uint32_t kc[GRAPH_LEVELS];
#pragma unroll
    for (uint32_t i = 0; i < GRAPH_LEVELS; ++i) {
        seed_p[i] = c;
        seed_dr[i] = T(1);
        kc[i] = 0u;
    }
    uint32_t dead = 0u;  // bit L: this lane has left the loop at level L
    uint32_t runs = 0u;
    uint32_t pc = 0u;
    uint32_t dnode = 0u;  // TRACK: the fence of dc, and of each distance slot
    uint32_t dsnode[GRAPH_DSLOTS] = {};
    for (uint32_t guard = 0; guard < GRAPH_MAX_STEPS; ++guard) {
        const GOp& o = prog[pc];
        const uint32_t op = __builtin_amdgcn_readfirstlane(o.op);
        const bool live = dead == 0u;
        switch (op) {
        case GOP_END:
            if constexpr (TRACK) *node_out = dnode;
            return GRaw<T>{dc, runs};
        case GOP_P_START: cur = GPoint<T>{c, T(1), 0u, INF, INF, INF, INF}; break;
        case GOP_P_SAVE: g_store(ps, o.a, cur); break;
        case GOP_P_LOAD: cur = g_load(ps, o.a); break;
        case GOP_ITER_BEGIN: {
            const uint32_t lv = o.a;
            if (live) {
                g_store(seed_p, lv, cur.p);
                g_store(seed_dr, lv, cur.dr);
                g_store(kc, lv, 0u);
                if (o.b == 0u || len3(cur.p) > T(o.v[0])) dead |= 1u << lv;
            }
            if (!__any(dead == 0u)) {
                dead &= ~(1u << lv);
                pc = o.c + 1u;
                continue;
            }
            break;
        }
        case GOP_ITER_END: {
            const uint32_t lv = o.a;
            if (live) {
                const uint32_t k = g_load(kc, lv) + 1u;
                g_store(kc, lv, k);
                cur.iter += 1u;
                runs += 1u;
                if (k >= o.b || len3(cur.p) > T(o.v[0])) dead |= 1u << lv;
            }
            if (__any(dead == 0u)) {
                pc = o.c;
                continue;
            }
            dead &= ~(1u << lv);
            break;
        }
#define SYF_G_CASE(OP) \
        case OP: if (live) g_point<T>(OP, o.a, o.b, o.c, o, cur, c, cur.p, cur.dr); break;
        SYF_GRAPH_POINT_OPS(SYF_G_CASE)
#undef SYF_G_CASE
        case GOP_ADD_SEED:
            if (live) {
                const V3<T> sp = o.a == 0u ? cur.p : g_load(seed_p, o.a - 1u);
                const T sdr = o.a == 0u ? cur.dr : g_load(seed_dr, o.a - 1u);
                g_point<T>(GOP_ADD_SEED, o.a, o.b, o.c, o, cur, c, sp, sdr);
            }
            break;
        case GOP_FENCE:
            dc = g_fence_state(o.a, o, cur);
            if constexpr (TRACK) dnode = node_ids[pc];
            break;
        case GOP_D_SAVE:
            g_store(ds, o.a, dc);
            if constexpr (TRACK) g_store(dsnode, o.a, dnode);
            break;
        case GOP_D_LOAD:
            dc = g_load(ds, o.a);
            if constexpr (TRACK) dnode = g_load(dsnode, o.a);
            break;
        case GOP_MATTAG: dc.mat = o.a; break;
        case GOP_BOUNDED: g_bounded(o, cur, dc); break;
        case GOP_COMBINE: {
            const GDist<T> A = o.b == GRAPH_ACC ? dc : g_load(ds, o.b);
            const GDist<T> B = o.c == GRAPH_ACC ? dc : g_load(ds, o.c);
            if constexpr (TRACK) {
                const uint32_t na = o.b == GRAPH_ACC ? dnode : g_load(dsnode, o.b);
                const uint32_t nb = o.c == GRAPH_ACC ? dnode : g_load(dsnode, o.c);
                dnode = g_combine_takes_a(o.a, T(o.v[0]), A, B) ? na : nb;
            }
            dc = g_combine(o.a, T(o.v[0]), A, B);
            break;
        }
        default:
            if constexpr (TRACK) *node_out = dnode;
            return GRaw<T>{dc, runs};
        }
        ++pc;
    }
    if constexpr (TRACK) *node_out = dnode;
    return GRaw<T>{dc, runs};
}

// The specialised evaluator of one graph topology: generated from the compiled program on the
// host (graph.hip) and compiled at run time into the kernels of that topology (kernel family
// FAM_JIT); the library's own build never instantiates it.
template <class T> __device__ GRaw<T> graph_jit(const GOp* prog, V3<T> c);

}  // namespace syf
