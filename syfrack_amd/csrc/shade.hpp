// Device shading for the path-traced forms: the radiance estimator of pathtrace.wgsl:802-1201
// restricted to this path's scene subset: sun, area and environment lights; a flat ground
// plane; the base material with its colouring ramp and SurfaceVariation, and named finishes
// (Lambert or Oren-Nayar diffuse + GGX, emission); next-event estimation with MIS, Russian
// roulette; height fog and finish detail patterns.  HDRI, portals, ocean, dunes, optics,
// textures, guiding and phenomena are absent, so their branches are inert.  The random draws happen in
// the reference's order so that a path is a pure function of (pixel, sample, seed), whichever
// kernel form runs it.
#pragma once

#include "device.hpp"

namespace syf {

typedef V3<float> F3;

// The job constants of a path-tracing kernel, read from device memory through the constant
// address space: the loads are invariant (scalar where the address is wave-uniform), and a
// per-lane index into a table (palette stops, finishes) is an ordinary load.  (A job passed by
// value as a kernel argument is copied to scratch memory, over a kilobyte per lane, as soon as
// one access indexes it per lane or its address reaches a call.)
typedef const __attribute__((address_space(4))) JobP ConstJob;
__device__ __forceinline__ const JobP& job_ref(const JobP* p) { return *(const JobP*)((const ConstJob*)p); }
constexpr float PI_F = 3.14159265358979f;
constexpr uint32_t NO_LIGHT = 0xffffffffu;
constexpr uint32_t GROUND_HIT = 0xfffffffeu;  // a segment that ended on the ground plane

__device__ __forceinline__ F3 mul3(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 neg3(F3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float luminance(F3 c) { return dot3(c, mk(0.2126f, 0.7152f, 0.0722f)); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ F3 mix3(F3 a, F3 b, float t) { return {mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)}; }
__device__ __forceinline__ F3 safe_normalize(F3 v) {  // common.wgsl:242-247
    float l = len3(v);
    float rcp = 1.0f / l;
    if (!(l > 0.0f) || !(rcp > 0.0f) || !(rcp <= 3.4028235e38f)) return mk(0.0f, 1.0f, 0.0f);
    return scl(v, rcp);
}

struct Surf {
    F3 n, diffuse, f0;
    float alpha;
    float sigma;  // Oren-Nayar roughness (diffuse_sigma); 0: Lambert
};

// The field sample the shading reads at a surface hit (DeSample): normalised traps, the
// iteration count clamped to the iteration product, the material index.
struct ShadeSample {
    float t0, t1, t2, t3;  // traps of slots 0..3
    uint32_t iter, material;
};

// A surface hit's sample when the shading reads only the trap the march returns (trap0: the
// graph's shading slot, slot 0 of a classic family); the other slots are not read.
__device__ __forceinline__ ShadeSample carried_sample(const JobP& P, float trap0) {
    const uint32_t k = P.f.gshade_slot;
    return ShadeSample{k == 0u ? trap0 : 1.0f, k == 1u ? trap0 : 1.0f, k == 2u ? trap0 : 1.0f, k == 3u ? trap0 : 1.0f,
                       0u, 0u};
}
// The ground's sample: the march's initial DeSample (pathtrace.wgsl:86) with the ground finish
// as material (lights.wgsl:201).
__device__ __forceinline__ ShadeSample ground_sample(const JobP& P) {
    return ShadeSample{1.0f, 1.0f, 1.0f, 1.0f, 0u, P.sh.ground_finish};
}
// The trap of slot k (0..3).  A masked sum, exact for the finite traps of a sample (normalised
// to [0, 1]): a select chain over adjacent fields is folded into an indexed load, which puts
// the sample in scratch memory.
__device__ __forceinline__ float trap_at(const ShadeSample& s, uint32_t k) {
    return (k == 0u ? 1.0f : 0.0f) * s.t0 + (k == 1u ? 1.0f : 0.0f) * s.t1 + (k == 2u ? 1.0f : 0.0f) * s.t2 +
           (k == 3u ? 1.0f : 0.0f) * s.t3;
}

__device__ __forceinline__ float4 stop_at(const ShadeP& S, uint32_t i) {
    return make_float4(S.stops[i][0], S.stops[i][1], S.stops[i][2], S.stops[i][3]);
}
// palette_mix (common.wgsl:331-335)
__device__ __forceinline__ F3 palette_mix(const ShadeP& S, uint32_t a, uint32_t b, float amount) {
    float f = clampT(amount, 0.0f, 1.0f);
    if (S.smooth) f = f * f * (3.0f - 2.0f * f);
    const float4 A = stop_at(S, a), B = stop_at(S, b);
    return mk(A.x * (1.0f - f) + B.x * f, A.y * (1.0f - f) + B.y * f, A.z * (1.0f - f) + B.z * f);
}
// palette_sample / ramp_sample (common.wgsl:336-366)
__device__ __forceinline__ F3 ramp_sample(const ShadeP& S, float value) {
    const uint32_t n = S.stop_count;
    if (n == 0u) return mk(1.0f, 1.0f, 1.0f);
    if (n == 1u) return mk(S.stops[0][0], S.stops[0][1], S.stops[0][2]);
    float t = value * S.frequency + S.phase;
    if (S.cyclic) t = t - floorf(t);
    else t = clampT(t, 0.0f, 1.0f);
    if (S.gamma != 1.0f) t = powf(t, S.gamma);
    if (!S.positioned && !S.cyclic) {
        const float x = t * float(n - 1u);
        const uint32_t i = min(uint32_t(floorf(x)), n - 2u);
        return palette_mix(S, i, i + 1u, x - float(i));
    }
    const float first = S.stops[0][3];
    const float last = stop_at(S, n - 1u).w;
    if (S.cyclic && (t < first || t >= last)) {
        const float x = t < first ? t + 1.0f : t;
        return palette_mix(S, n - 1u, 0u, (x - last) / fmaxf(1.0f + first - last, 1e-8f));
    }
    if (t <= first) return mk(S.stops[0][0], S.stops[0][1], S.stops[0][2]);
    const float4 lastc = stop_at(S, n - 1u);
    if (t >= last) return mk(lastc.x, lastc.y, lastc.z);
    for (uint32_t i = 0; i + 1u < n; ++i) {
        const float4 a = stop_at(S, i), b = stop_at(S, i + 1u);
        if (t < b.w) return palette_mix(S, i, i + 1u, (t - a.w) / (b.w - a.w));
    }
    return mk(lastc.x, lastc.y, lastc.z);
}
__device__ __forceinline__ FinishP finish_at(const ShadeP& S, uint32_t i) { return S.finishes[i]; }

// Duff et al. 2017 orthonormal basis (pathtrace.wgsl:194-201); columns t, bt, n.
__device__ __forceinline__ void tangent_frame(F3 n, F3& t, F3& bt) {
    float s = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n.z);
    float b = n.x * n.y * a;
    t = mk(1.0f + s * n.x * n.x * a, s * b, -s * n.x);
    bt = mk(b, s + n.y * n.y * a, -n.y);
}
__device__ __forceinline__ F3 frame_mul(F3 t, F3 bt, F3 n, F3 x) { return add(add(scl(t, x.x), scl(bt, x.y)), scl(n, x.z)); }

__device__ __forceinline__ F3 cosine_sample(F3 n, float r1, float r2) {  // pathtrace.wgsl:203-208
    float phi = 2.0f * PI_F * r1;
    float r = sqrtf(r2);
    F3 t, bt;
    tangent_frame(n, t, bt);
    return frame_mul(t, bt, n, mk(r * cosf(phi), r * sinf(phi), sqrtf(fmaxf(0.0f, 1.0f - r2))));
}
__device__ __forceinline__ float ggx_d(float nh, float a2) {
    float d = nh * nh * (a2 - 1.0f) + 1.0f;
    return a2 / (PI_F * d * d);
}
__device__ __forceinline__ float smith_g1(float nv, float a2) { return 2.0f * nv / (nv + sqrtf(a2 + (1.0f - a2) * nv * nv)); }
__device__ __forceinline__ F3 fresnel(F3 f0, float vh) {
    float k = powf(1.0f - vh, 5.0f);
    return mk(f0.x + (1.0f - f0.x) * k, f0.y + (1.0f - f0.y) * k, f0.z + (1.0f - f0.z) * k);
}
// Heitz 2018 GGX visible-normal sampling (pathtrace.wgsl:234-249).
__device__ __forceinline__ F3 sample_ggx_h(F3 n, F3 v, float alpha, float r1, float r2) {
    F3 t, bt;
    tangent_frame(n, t, bt);
    F3 lv = mk(dot3(t, v), dot3(bt, v), dot3(n, v));
    F3 vh = safe_normalize(mk(alpha * lv.x, alpha * lv.y, fmaxf(lv.z, 1e-5f)));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    F3 t1 = mk(1.0f, 0.0f, 0.0f);
    if (lensq > 0.0f) {
        float sl = sqrtf(lensq);
        t1 = mk(-vh.y / sl, vh.x / sl, 0.0f);
    }
    F3 t2 = cross3(vh, t1);
    float r = sqrtf(r1);
    float phi = 2.0f * PI_F * r2;
    float x = r * cosf(phi);
    float blend = 0.5f * (1.0f + vh.z);
    float y = mixf(sqrtf(fmaxf(0.0f, 1.0f - x * x)), r * sinf(phi), blend);
    F3 nh = add(add(scl(t1, x), scl(t2, y)), scl(vh, sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y))));
    return frame_mul(t, bt, n, safe_normalize(mk(alpha * nh.x, alpha * nh.y, fmaxf(nh.z, 0.0f))));
}
__device__ __forceinline__ F3 reflect3(F3 e1, F3 e2) { return sub(e1, scl(e2, 2.0f * dot3(e2, e1))); }

// surface_value (pathtrace.wgsl:265-269)
__device__ __forceinline__ float surface_value(const JobP& P, uint32_t slot, F3 n, const ShadeSample& smp) {
    if (slot < 4u) return trap_at(smp, slot);
    if (slot == SYF_COLOR_NORMAL) return n.y * 0.5f + 0.5f;
    return float(smp.iter) / fmaxf(float(P.f.iteration_product), 1.0f);
}
}  // namespace syf
#include "finish_detail.hpp"
namespace syf {

// make_surface (pathtrace.wgsl:271-386) for this path's subset: the colouring ramp, a named
// finish (albedo, roughness, metallic, specular, emission; Oren-Nayar sigma for the concrete
// patterns), the SurfaceVariation.  The emission (Surface.emission) is returned apart: the
// caller adds it at once and the surface record stays small for the bounce's remainder.
template <bool DETAIL>
__device__ __forceinline__ Surf make_surface(const JobP& P, F3 p, F3 n, const ShadeSample& smp, float footprint,
                                             F3& emission) {
    const MaterialP& m = P.mat;
    const ShadeP& S = P.sh;
    F3 albedo = ld3(m.albedo);
    const uint32_t slot = S.color_slot;
    const bool assigned = smp.material > 0u && smp.material <= S.finish_count;
    const float iters = float(smp.iter) / fmaxf(float(P.f.iteration_product), 1.0f);
    if (!assigned && slot == SYF_COLOR_ORBIT) {
        float value = S.weights[0] * smp.t0 + S.weights[1] * smp.t1 + S.weights[2] * smp.t2 + S.weights[3] * smp.t3 +
                      S.iteration_weight * iters + S.normal_weight * (n.y * 0.5f + 0.5f);
        const float gain = S.logarithmic;
        if (gain > 0.0f) {
            const float sg = value > 0.0f ? 1.0f : (value < 0.0f ? -1.0f : 0.0f);
            value = sg * logf(1.0f + gain * fabsf(value)) / logf(1.0f + gain);
        }
        albedo = mul3(albedo, ramp_sample(S, value));
    } else if (!assigned && slot < 4u) {
        albedo = mul3(albedo, ramp_sample(S, trap_at(smp, slot)));
    } else if (!assigned && slot == SYF_COLOR_NORMAL) {
        albedo = mul3(albedo, ramp_sample(S, n.y * 0.5f + 0.5f));
    } else if (!assigned && slot == SYF_COLOR_ITER) {
        albedo = mul3(albedo, ramp_sample(S, iters));
    }
    float rough = m.roughness, metallic = m.metallic, spec = m.specular, sigma_scale = 0.0f;
    emission = mk(0.0f, 0.0f, 0.0f);
    if (assigned) {
        const FinishP f = finish_at(S, smp.material - 1u);
        albedo = ld3(f.albedo);
        rough = f.roughness;
        metallic = f.metallic;
        spec = f.specular;
        emission = ld3(f.emission);
        sigma_scale = f.sigma_scale;
    }
    if (!assigned && m.surface_enabled) {
        const float t = clampT(surface_value(P, m.surface_slot, n, smp), 0.0f, 1.0f);
        rough = m.surface[0] + (m.surface[1] - m.surface[0]) * t;
        metallic = m.surface[2] + (m.surface[3] - m.surface[2]) * t;
    }
    F3 shading_n = n;
    if (DETAIL && assigned && S.detail) {
        const FinishP& f = S.finishes[smp.material - 1u];
        if (f.pattern != 0u) apply_finish_detail(f, p, n, footprint, albedo, rough, metallic, spec, shading_n);
    }
    rough = clampT(rough, 0.02f, 1.0f);
    Surf s;
    s.n = shading_n;
    s.diffuse = scl(albedo, 1.0f - metallic);
    s.f0 = mix3(mk(spec, spec, spec), albedo, metallic);
    s.alpha = rough * rough;
    s.sigma = rough * sigma_scale;
    return s;
}
// brdf_eval (pathtrace.wgsl:394-420): Lambert or Oren-Nayar diffuse + GGX.
__device__ __forceinline__ F3 brdf_eval(const Surf& s, F3 v, F3 l) {
    float nl = dot3(s.n, l), nv = dot3(s.n, v);
    if (nl <= 0.0f || nv <= 0.0f) return mk(0.0f, 0.0f, 0.0f);
    F3 h = nrm(add(v, l));
    float nh = fmaxf(dot3(s.n, h), 0.0f);
    float vh = fmaxf(dot3(v, h), 0.0f);
    float a2 = s.alpha * s.alpha;
    F3 f = fresnel(s.f0, vh);
    float k = ggx_d(nh, a2) * smith_g1(nv, a2) * smith_g1(nl, a2);
    float den = fmaxf(4.0f * nv * nl, 1e-4f);
    float df = 1.0f;
    if (s.sigma > 0.0f) {  // Oren-Nayar (pathtrace.wgsl:407-418)
        const float sigma2 = s.sigma * s.sigma;
        const float a = 1.0f - sigma2 / (2.0f * (sigma2 + 0.33f));
        const float b = 0.45f * sigma2 / (sigma2 + 0.09f);
        const float sin_l = sqrtf(fmaxf(0.0f, 1.0f - nl * nl));
        const float sin_v = sqrtf(fmaxf(0.0f, 1.0f - nv * nv));
        const float cos_phi = fmaxf(0.0f, dot3(sub(l, scl(s.n, nl)), sub(v, scl(s.n, nv))) / fmaxf(sin_l * sin_v, 1e-8f));
        df = a + b * cos_phi * sin_l * sin_v / fmaxf(fmaxf(nl, nv), 1e-5f);
    }
    return mk(s.diffuse.x * (1.0f - f.x) * df / PI_F + k * f.x / den,
              s.diffuse.y * (1.0f - f.y) * df / PI_F + k * f.y / den,
              s.diffuse.z * (1.0f - f.z) * df / PI_F + k * f.z / den);
}
__device__ __forceinline__ float mis_power(float a, float b) {
    float scale = fmaxf(a, b);
    if (scale <= 0.0f) return 0.0f;
    float pa = a / scale, pb = b / scale;
    return pa * pa / (pa * pa + pb * pb);
}
__device__ __forceinline__ float spec_probability(const Surf& s, float nv) {
    F3 f = fresnel(s.f0, nv);
    return clampT(luminance(f) / fmaxf(luminance(f) + luminance(s.diffuse), 1e-4f), 0.05f, 0.95f);
}
// bsdf_pdf (pathtrace.wgsl:443-452) = continuation_pdf without guiding or transmission.
__device__ __forceinline__ float bsdf_pdf(const Surf& s, F3 v, F3 l, float p_spec) {
    float nl = dot3(s.n, l);
    if (nl <= 0.0f) return 0.0f;
    F3 hh = nrm(add(v, l));
    float nh = fmaxf(dot3(s.n, hh), 0.0f);
    float a2 = s.alpha * s.alpha;
    float nv = fmaxf(dot3(s.n, v), 1e-4f);
    float pdf_spec = ggx_d(nh, a2) * smith_g1(nv, a2) / (4.0f * nv);
    return p_spec * pdf_spec + (1.0f - p_spec) * nl / PI_F;
}
__device__ __forceinline__ F3 environment_value(const LightP& l, F3 dir) {  // pathtrace.wgsl:164-167
    if (dir.y >= 0.0f) return mix3(ld3(l.horizon), ld3(l.color), sqrtf(dir.y));
    return mix3(ld3(l.horizon), ld3(l.ground), sqrtf(-dir.y));
}
__device__ __forceinline__ float environment_probability(const LightP& l, float area_total) {
    float energy = fmaxf(fmaxf(luminance(ld3(l.color)), luminance(ld3(l.horizon))), luminance(ld3(l.ground)));
    if (energy <= 0.0f) return 0.0f;
    return clampT(4.0f * PI_F * energy / (4.0f * PI_F * energy + area_total), 0.05f, 1.0f);
}
__device__ __forceinline__ float area_intersection(F3 ro, F3 rd, const LightP& l) {  // lights.wgsl:170-181
    F3 n = ld3(l.n);
    float den = dot3(n, rd);
    if (fabsf(den) < 1e-7f) return 1e30f;
    float t = dot3(sub(ld3(l.a), ro), n) / den;
    if (t <= 1e-6f) return 1e30f;
    F3 q = sub(add(ro, scl(rd, t)), ld3(l.a));
    float x = dot3(q, ld3(l.b)) / dot3(ld3(l.b), ld3(l.b));
    float y = dot3(q, ld3(l.c)) / dot3(ld3(l.c), ld3(l.c));
    if (fabsf(x) > 0.5f || fabsf(y) > 0.5f) return 1e30f;
    return t;
}
__device__ __forceinline__ float area_weight(F3 p, F3 n, const LightP& l) {  // lights.wgsl:288-304
    F3 to = sub(ld3(l.a), p);
    float dist2 = fmaxf(dot3(to, to), 1e-6f);
    float sq = sqrtf(dist2);
    F3 dir = mk(to.x / sq, to.y / sq, to.z / sq);
    float facing = fmaxf(dot3(ld3(l.n), neg3(dir)), 0.0f);
    return luminance(ld3(l.color)) * l.area * facing * fmaxf(dot3(n, dir), 0.15f) / dist2;
}
// Loops over the job's lights run over the fixed capacity, unrolled, and stop at light_count:
// with constant indices the light table stays in registers/constant loads instead of being
// copied to scratch for a dynamic index.
#define LIGHT_LOOP(i) _Pragma("unroll") for (uint32_t i = 0; i < SYF_MAX_LIGHTS; ++i) if (i < P.light_count)
__device__ __forceinline__ F3 light_normal(const JobP& P, uint32_t k) {
    F3 n = mk(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (uint32_t i = 0; i < SYF_MAX_LIGHTS; ++i)
        if (i == k) n = ld3(P.lights[i].n);
    return n;
}
__device__ __forceinline__ float area_total_weight(const JobP& P, F3 p, F3 n) {
    float total = 0.0f;
    LIGHT_LOOP(i)
        if (P.lights[i].kind == SYF_LIGHT_AREA) total += area_weight(p, n, P.lights[i]);
    return total;
}
// ground_intersection for a flat ground (landscape.wgsl:17-23)
__device__ __forceinline__ float ground_intersection(const ShadeP& S, F3 ro, F3 rd, float t_end) {
    if (fabsf(rd.y) < 1e-7f) return t_end;
    const float t = (S.ground_height - ro.y) / rd.y;
    return (t > 1e-5f && t < t_end) ? t : t_end;
}
// ground_normal of a flat ground, facing the ray's origin (landscape.wgsl:71-78)
__device__ __forceinline__ F3 ground_normal(F3 rd) { return rd.y > 0.0f ? mk(0.0f, -1.0f, 0.0f) : mk(0.0f, 1.0f, 0.0f); }
// Nearest ground or emitter before t_end (trace_geometry_lights, lights.wgsl:183-209): light is
// the emitter's index, GROUND_HIT, or NO_LIGHT.
__device__ __forceinline__ float nearest_emitter(const JobP& P, F3 ro, F3 rd, float t_end, uint32_t& light) {
    float nearest = t_end;
    light = NO_LIGHT;
    if (P.sh.ground_enabled) {
        const float t = ground_intersection(P.sh, ro, rd, nearest);
        if (t > 1e-5f && t < nearest) {
            nearest = t;
            light = GROUND_HIT;
        }
    }
    LIGHT_LOOP(i) {
        if (P.lights[i].kind != SYF_LIGHT_AREA) continue;
        float t = area_intersection(ro, rd, P.lights[i]);
        if (t < nearest) {
            nearest = t;
            light = i;
        }
    }
    return nearest;
}
// volume_light_probability (lights.wgsl:319-322)
__device__ __forceinline__ float volume_light_probability(float weight, float total) {
    return total <= 0.0f ? 0.0f : fminf(1.0f, 3.0f * weight / total);
}
// area_emission (lights.wgsl:330-357); from_volume: the segment started at a scattering event in
// the fog, whose light selection is volume_light_probability.
__device__ __forceinline__ F3 area_emission(const JobP& P, F3 ro, F3 rd, float distance, float prev_pdf, bool use_mis,
                                            F3 prev_n, float prev_total, bool from_volume = false) {
    F3 L = mk(0.0f, 0.0f, 0.0f);
    LIGHT_LOOP(i) {
        const LightP& l = P.lights[i];
        if (l.kind != SYF_LIGHT_AREA) continue;
        float cos_l = dot3(ld3(l.n), neg3(rd));
        if (cos_l <= 0.0f) continue;
        float t = area_intersection(ro, rd, l);
        if (fabsf(t - distance) > fmaxf(1e-6f, distance * 1e-6f)) continue;
        float w = 1.0f;
        if (use_mis) {
            float pl = distance * distance / fmaxf(l.area * cos_l, 1e-20f);
            float sel = 0.0f;
            if (prev_total > 0.0f)
                sel = from_volume ? volume_light_probability(area_weight(ro, mk(0.0f, 0.0f, 0.0f), l), prev_total)
                                  : area_weight(ro, prev_n, l) / prev_total;
            pl *= sel;
            w = mis_power(prev_pdf, pl);
        }
        L = add(L, scl(ld3(l.color), w));
    }
    return L;
}
// environment_emission (pathtrace.wgsl:177-189): from_volume, the direction came from a phase
// sample in the fog, whose NEE proposal is uniform over the sphere.
__device__ __forceinline__ F3 environment_emission(const JobP& P, F3 dir, float pdf, bool use_mis, F3 n,
                                                   float area_total, bool from_volume = false) {
    F3 L = mk(0.0f, 0.0f, 0.0f);
    LIGHT_LOOP(i) {
        const LightP& l = P.lights[i];
        if (l.kind != SYF_LIGHT_ENV) continue;
        float w = 1.0f;
        if (use_mis)
            w = mis_power(pdf, environment_probability(l, area_total) *
                                   (from_volume ? 1.0f / (4.0f * PI_F) : fmaxf(dot3(n, dir), 0.0f) / PI_F));
        L = add(L, scl(environment_value(l, dir), w));
    }
    return L;
}

// cone_sample / sun_pdf (pathtrace.wgsl:210-215, 432-434)
__device__ __forceinline__ F3 cone_sample(F3 dir, float cos_max, float r1, float r2) {
    const float cos_t = 1.0f - r1 * (1.0f - cos_max);
    const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    const float phi = 2.0f * PI_F * r2;
    F3 t, bt;
    tangent_frame(dir, t, bt);
    return frame_mul(t, bt, dir, mk(sin_t * cosf(phi), sin_t * sinf(phi), cos_t));
}
__device__ __forceinline__ float sun_pdf(float cos_max) { return 1.0f / (2.0f * PI_F * fmaxf(1.0f - cos_max, 1e-7f)); }
// sun_radiance (pathtrace.wgsl:454-467): every sun disk containing dir, MIS-weighted against
// the BSDF sample that found it.
__device__ __forceinline__ F3 sun_radiance(const JobP& P, F3 dir, float pdf_bsdf, bool use_mis) {
    F3 L = mk(0.0f, 0.0f, 0.0f);
    LIGHT_LOOP(i) {
        const LightP& l = P.lights[i];
        if (l.kind != SYF_LIGHT_SUN || !(dot3(dir, ld3(l.a)) >= l.cos_max)) continue;
        const float pl = sun_pdf(l.cos_max);
        const float w = use_mis ? mis_power(pdf_bsdf, pl) : 1.0f;
        L = add(L, mk(l.color[0] * pl * w, l.color[1] * pl * w, l.color[2] * pl * w));
    }
    return L;
}

// ---------------------------------------------------------------- height fog
// The fog of pathtrace.wgsl:560-779: extinction sigma(p) = density exp(min(0, -(y - height)
// falloff)), exact Beer-Lambert optical depths split at the fog height, free-flight sampling by
// inverting the optical depth, the emitters' line integral along a segment (equiangular and
// distance sampling combined), Henyey-Greenstein scattering with next-event estimation of the
// environment and suns.
__device__ __forceinline__ float fog_sigma(const ShadeP& S, F3 p) {
    return S.fog_density * expf(fminf(0.0f, -(p.y - S.fog_height) * S.fog_falloff));
}
__device__ __forceinline__ float fog_segment_tau(const ShadeP& S, F3 ro, F3 rd, float begin, float end) {
    const float distance = fmaxf(end - begin, 0.0f);
    if (distance == 0.0f) return 0.0f;
    const float sigma = fog_sigma(S, add(ro, scl(rd, begin)));
    const float mid_height = ro.y + rd.y * ((begin + end) * 0.5f);
    if (mid_height <= S.fog_height || S.fog_falloff == 0.0f) return S.fog_density * distance;
    const float k = rd.y * S.fog_falloff;
    const float x = k * distance;
    if (fabsf(x) < 0.001f) return sigma * distance * (1.0f - x * 0.5f + x * x / 6.0f);
    return fmaxf(0.0f, (sigma - fog_sigma(S, add(ro, scl(rd, end)))) / k);
}
__device__ __forceinline__ float fog_split(const ShadeP& S, F3 ro, F3 rd, float distance) {
    if (fabsf(rd.y) < 1e-8f) return distance;
    return clampT((S.fog_height - ro.y) / rd.y, 0.0f, distance);
}
__device__ __forceinline__ float fog_optical_depth(const ShadeP& S, F3 ro, F3 rd, float distance) {
    const float split = fog_split(S, ro, rd, distance);
    return fog_segment_tau(S, ro, rd, 0.0f, split) + fog_segment_tau(S, ro, rd, split, distance);
}
__device__ __forceinline__ float fog_transmittance(const ShadeP& S, F3 ro, F3 rd, float dist) {
    return expf(-fog_optical_depth(S, ro, rd, dist));
}
__device__ __forceinline__ float fog_invert_segment(const ShadeP& S, F3 ro, F3 rd, float begin, float end, float tau) {
    const float sigma = fog_sigma(S, add(ro, scl(rd, begin)));
    const float k = rd.y * S.fog_falloff;
    if (ro.y + rd.y * ((begin + end) * 0.5f) <= S.fog_height || fabsf(k * (end - begin)) < 0.0001f)
        return fminf(end, begin + tau / fmaxf(sigma, 1e-30f));
    if (sigma > 1e-20f) return clampT(begin - logf(fmaxf(1e-7f, 1.0f - tau * k / sigma)) / k, begin, end);
    const float remaining = fmaxf(0.0f, fog_segment_tau(S, ro, rd, begin, end) - tau);
    const float end_sigma = fog_sigma(S, add(ro, scl(rd, end)));
    return clampT(end + logf(fmaxf(1e-7f, 1.0f + remaining * k / fmaxf(end_sigma, 1e-30f))) / (-k), begin, end);
}
// The first collision along [0, t_end), t_end without one (fog_free_flight).
__device__ __forceinline__ float fog_free_flight(const ShadeP& S, F3 ro, F3 rd, float t_end, uint32_t& rng) {
    if (S.fog_density <= 0.0f) return t_end;
    const float tau = -logf(fmaxf(1.0f - rand01(rng), 1e-7f));
    const float split = fog_split(S, ro, rd, t_end);
    const float first = fog_segment_tau(S, ro, rd, 0.0f, split);
    const float total = first + fog_segment_tau(S, ro, rd, split, t_end);
    if (tau >= total) return t_end;
    if (tau < first) return fog_invert_segment(S, ro, rd, 0.0f, split, tau);
    return fog_invert_segment(S, ro, rd, split, t_end, tau - first);
}
// volume_coordinate (pathtrace.wgsl:43-58): the radical inverse of sample_index + 1 in the
// dimension's prime base, rotated by a per-stream shift.
__device__ __forceinline__ float volume_coordinate(uint32_t sample_index, uint32_t dimension, uint32_t stream) {
    const uint32_t base = dimension == 0u ? 2u : dimension == 1u ? 3u : dimension == 2u ? 5u
                        : dimension == 3u ? 7u : dimension == 4u ? 11u : 13u;
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
    return x - floorf(x);
}
__device__ __forceinline__ float hg_phase(float cos_theta, float g) {  // pathtrace.wgsl:717-720
    const float d = 1.0f + g * g - 2.0f * g * cos_theta;
    return (1.0f - g * g) / (4.0f * PI_F * d * sqrtf(d));
}
__device__ __forceinline__ F3 hg_sample(F3 dir, float g, float r1, float r2) {  // pathtrace.wgsl:722-731
    float cos_t = 1.0f - 2.0f * r1;
    if (fabsf(g) > 1e-3f) {
        const float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * r1);
        cos_t = (1.0f + g * g - sq * sq) / (2.0f * g);
    }
    const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    const float phi = 2.0f * PI_F * r2;
    F3 t, bt;
    tangent_frame(dir, t, bt);
    return frame_mul(t, bt, dir, mk(sin_t * cosf(phi), sin_t * sinf(phi), cos_t));
}
// fog_segment_weight (pathtrace.wgsl:621-633) of an area light
__device__ __forceinline__ float fog_segment_weight(const LightP& l, F3 ro, F3 rd, float end) {
    const F3 center = ld3(l.a);
    const float along = clampT(dot3(sub(center, ro), rd), 0.0f, end);
    const F3 delta = sub(center, add(ro, scl(rd, along)));
    return fmaxf(fmaxf(l.color[0], l.color[1]), l.color[2]) * l.area / (dot3(delta, delta) + l.area);
}
// fog_segment_direct (pathtrace.wgsl:635-715): the emitters' light scattered toward the ray
// along [0, end), each emitter proposing one point chosen by equiangular or distance sampling
// (one-sample MIS of the two), its distance pdf the mean of theirs.  The camera segment
// (primary) takes its selection, emitter point and distance coordinates from the per-pixel
// rotated Halton sequence of the sample index; every random draw happens either way.
template <class Vis>
__device__ __forceinline__ F3 fog_segment_direct(const JobP& P, F3 ro, F3 rd, float end, uint32_t& rng,
                                                 uint32_t point_index, uint32_t point_stream, bool primary,
                                                 Vis&& visible) {
    const ShadeP& S = P.sh;
    if (end <= 0.0f || fmaxf(fmaxf(S.fog_color[0], S.fog_color[1]), S.fog_color[2]) <= 0.0f) return mk(0.0f, 0.0f, 0.0f);
    const float tau_end = fog_optical_depth(S, ro, rd, end);
    if (tau_end <= 0.0f) return mk(0.0f, 0.0f, 0.0f);
    float total = 0.0f;
    LIGHT_LOOP(i)
        if (P.lights[i].kind == SYF_LIGHT_AREA) total += fog_segment_weight(P.lights[i], ro, rd, end);
    F3 result = mk(0.0f, 0.0f, 0.0f);
#pragma unroll 1
    for (uint32_t i = 0; i < P.light_count; ++i) {
        const LightP& l = P.lights[i];
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
        const F3 q = add(add(ld3(l.a), scl(ld3(l.b), u1 - 0.5f)), scl(ld3(l.c), u2 - 0.5f));
        const float tau_begin = 0.0f;  // fog_optical_depth over [0, 0)
        const float tau_finish = fog_optical_depth(S, ro, rd, end);
        const float tau_span = fmaxf(0.0f, tau_finish - tau_begin);
        if (tau_span <= 0.0f) continue;
        float conditional_mass = 1.0f - expf(-tau_span);
        if (tau_span < 0.001f) conditional_mass = tau_span * (1.0f - tau_span * 0.5f + tau_span * tau_span / 6.0f);
        const float interval_mass = expf(-tau_begin) * conditional_mass;
        if (interval_mass < 1e-30f) continue;
        const F3 delta = sub(q, ro);
        const float projection = dot3(delta, rd);
        const float perpendicular = len3(sub(delta, scl(rd, projection)));
        const float theta_a = atan2f(0.0f - projection, fmaxf(perpendicular, 1e-6f));
        const float theta_b = atan2f(end - projection, fmaxf(perpendicular, 1e-6f));
        const float span = theta_b - theta_a;
        const bool regular = perpendicular > 1e-4f && span > 1e-5f;
        float xi = rand01(rng);
        float technique = rand01(rng);
        if (primary) {
            xi = volume_coordinate(point_index, 3u, stream);
            technique = volume_coordinate(point_index, 4u, stream);
        }
        float t = mixf(0.0f, end, xi);
        if (technique < 0.5f) {
            if (regular) t = clampT(projection + perpendicular * tanf(mixf(theta_a, theta_b, xi)), 0.0f, end);
        } else {
            const float tau = tau_begin - logf(fmaxf(1.0f - xi * conditional_mass, 1e-7f));
            const float split = fog_split(S, ro, rd, end);
            const float first = fog_segment_tau(S, ro, rd, 0.0f, split);
            t = tau < first ? fog_invert_segment(S, ro, rd, 0.0f, split, tau)
                            : fog_invert_segment(S, ro, rd, split, end, tau - first);
        }
        const F3 p = add(ro, scl(rd, t));
        const float transmittance = expf(-fog_optical_depth(S, ro, rd, t));
        const float sigma = fog_sigma(S, p);
        float angular_pdf = 1.0f / (end - 0.0f);
        if (regular)
            angular_pdf = perpendicular / (span * (perpendicular * perpendicular + (t - projection) * (t - projection)));
        const float distance_pdf = 0.5f * (angular_pdf + sigma * transmittance / interval_mass);
        const F3 to_light = sub(q, p);
        const float distance2 = dot3(to_light, to_light);
        if (distance2 < 1e-8f) continue;
        const float distance = sqrtf(distance2);
        const F3 wi = mk(to_light.x / distance, to_light.y / distance, to_light.z / distance);
        const float cosine = dot3(ld3(l.n), neg3(wi));
        if (cosine <= 0.0f) continue;
        if (!visible(p, wi, fmaxf(0.0f, distance - fmaxf(1e-5f, distance * 1e-6f)))) continue;
        const float tr = transmittance * expf(-fog_optical_depth(S, p, wi, distance));
        const float area_pdf = distance2 / fmaxf(l.area * cosine, 1e-20f);  // area_direction_pdf
        const float k = sigma * tr * hg_phase(dot3(rd, wi), S.fog_g) * 1.0f * 1.0f /
                        (probability * distance_pdf * area_pdf);
        result = add(result, mk(l.color[0] * k, l.color[1] * k, l.color[2] * k));
    }
    return mk(result.x * S.fog_color[0], result.y * S.fog_color[1], result.z * S.fog_color[2]);
}
// fog_direct_light (pathtrace.wgsl:733-778): the environment and suns at a scattering event,
// weighted by the phase function instead of a BRDF.
template <class Vis>
__device__ __forceinline__ F3 fog_direct_light(const JobP& P, F3 p, F3 rd, float area_total, bool last, uint32_t& rng,
                                               Vis&& visible) {
    const float g = P.sh.fog_g;
    F3 L = mk(0.0f, 0.0f, 0.0f);
#pragma unroll 1
    for (uint32_t i = 0; i < P.light_count; ++i) {
        const LightP& l = P.lights[i];
        if (l.kind == SYF_LIGHT_ENV) {
            const float mx = fmaxf(fmaxf(fmaxf(l.color[0], l.color[1]), l.color[2]),
                                   fmaxf(fmaxf(fmaxf(l.horizon[0], l.horizon[1]), l.horizon[2]),
                                         fmaxf(fmaxf(l.ground[0], l.ground[1]), l.ground[2])));
            if (mx <= 0.0f) continue;
            const float probability = environment_probability(l, area_total);
            if (rand01(rng) >= probability) continue;
            const float z = 1.0f - 2.0f * rand01(rng);
            const float phi = 2.0f * PI_F * rand01(rng);
            const float radius = sqrtf(fmaxf(0.0f, 1.0f - z * z));
            const F3 dir = mk(radius * cosf(phi), z, radius * sinf(phi));
            if (!visible(p, dir, P.t_end)) continue;
            const float tr = fog_transmittance(P.sh, p, dir, P.t_end);
            const float ph = hg_phase(dot3(rd, dir), g);
            const float pl = probability / (4.0f * PI_F);
            const float w = last ? 1.0f : mis_power(pl, ph);
            const F3 e = environment_value(l, dir);
            L = add(L, mk(e.x * ph * tr * w / pl, e.y * ph * tr * w / pl, e.z * ph * tr * w / pl));
        } else if (l.kind == SYF_LIGHT_SUN) {
            const float r1 = rand01(rng);
            const float r2 = rand01(rng);
            const F3 dir = cone_sample(ld3(l.a), l.cos_max, r1, r2);
            if (!visible(p, dir, P.t_end)) continue;
            const float tr = fog_transmittance(P.sh, p, dir, P.t_end);
            const float ph = hg_phase(dot3(rd, dir), g);
            const float pl = sun_pdf(l.cos_max);
            const float w = last ? 1.0f : mis_power(pl, ph);
            L = add(L, mk(l.color[0] * ph * tr * w, l.color[1] * ph * tr * w, l.color[2] * ph * tr * w));
        }
    }
    return L;
}

// Next-event estimation as direct_light (pathtrace.wgsl:471-558): choose one area light by
// weight, then the environment, each sun and the chosen area light each propose one shadow ray.
// The caller supplies the visibility test (a march to the light), so the megakernel and
// the wavefront form share this logic.
// in_fog: the shadow rays' light is attenuated by the fog's transmittance.
template <class Vis>
__device__ __forceinline__ F3 direct_light(const JobP& P, F3 p, const Surf& s, F3 gn, F3 v, float p_spec,
                                           float area_total, bool last, bool in_fog, uint32_t& rng, Vis&& visible) {
    F3 L = mk(0.0f, 0.0f, 0.0f);
    uint32_t selected = NO_LIGHT;
    float selected_weight = 0.0f;
    if (area_total > 0.0f) {
        float threshold = rand01(rng) * area_total;
        float sum = 0.0f;
        LIGHT_LOOP(i) {
            if (P.lights[i].kind != SYF_LIGHT_AREA) continue;
            float w = area_weight(p, s.n, P.lights[i]);
            sum += w;
            if (w > 0.0f) {
                selected = i;
                selected_weight = w;
                if (threshold < sum) break;
            }
        }
    }
    LIGHT_LOOP(i) {
        const LightP& l = P.lights[i];
        if (l.kind == SYF_LIGHT_ENV) {
            float mx = fmaxf(fmaxf(fmaxf(l.color[0], l.color[1]), l.color[2]),
                             fmaxf(fmaxf(fmaxf(l.horizon[0], l.horizon[1]), l.horizon[2]),
                                   fmaxf(fmaxf(l.ground[0], l.ground[1]), l.ground[2])));
            if (mx <= 0.0f) continue;
            float prob = environment_probability(l, area_total);
            if (rand01(rng) >= prob) continue;
            float r1 = rand01(rng);
            float r2 = rand01(rng);
            F3 dir = cosine_sample(s.n, r1, r2);
            float nl = fmaxf(dot3(s.n, dir), 0.0f);
            if (dot3(gn, dir) <= 0.0f || nl <= 0.0f) continue;
            if (!visible(p, dir, P.t_end)) continue;
            const float tr = in_fog ? fog_transmittance(P.sh, p, dir, P.t_end) : 1.0f;
            float pl = prob * nl / PI_F;
            float pb = last ? 0.0f : bsdf_pdf(s, v, dir, p_spec);
            float w = last ? 1.0f : mis_power(pl, pb);
            F3 f = brdf_eval(s, v, dir);
            F3 e = environment_value(l, dir);
            L = add(L, mk(f.x * nl * e.x * tr * w / pl, f.y * nl * e.y * tr * w / pl, f.z * nl * e.z * tr * w / pl));
        } else if (l.kind == SYF_LIGHT_SUN) {
            // colour = irradiance: the disk's radiance is colour * sun_pdf
            const float r1 = rand01(rng);
            const float r2 = rand01(rng);
            const F3 dir = cone_sample(ld3(l.a), l.cos_max, r1, r2);
            const float nl = dot3(s.n, dir);
            if (nl <= 0.0f || dot3(gn, dir) <= 0.0f) continue;
            if (!visible(p, dir, P.t_end)) continue;
            const float tr = in_fog ? fog_transmittance(P.sh, p, dir, P.t_end) : 1.0f;
            const float pl = sun_pdf(l.cos_max);
            const float pb = last ? 0.0f : bsdf_pdf(s, v, dir, p_spec);
            const float w = last ? 1.0f : mis_power(pl, pb);
            const F3 f = brdf_eval(s, v, dir);
            L = add(L, mk(f.x * nl * l.color[0] * w * tr, f.y * nl * l.color[1] * w * tr, f.z * nl * l.color[2] * w * tr));
        } else if (l.kind == SYF_LIGHT_AREA) {
            if (i != selected) continue;
            float u1 = rand01(rng);
            float u2 = rand01(rng);
            F3 q = add(add(ld3(l.a), scl(ld3(l.b), u1 - 0.5f)), scl(ld3(l.c), u2 - 0.5f));
            F3 to_q = sub(q, p);
            float dist2 = dot3(to_q, to_q);
            if (dist2 < 1e-8f) continue;
            float dist = sqrtf(dist2);
            F3 dir = mk(to_q.x / dist, to_q.y / dist, to_q.z / dist);
            float nl = dot3(s.n, dir);
            float cos_l = dot3(ld3(l.n), neg3(dir));
            if (nl <= 0.0f || cos_l <= 0.0f || dot3(gn, dir) <= 0.0f) continue;
            float end = fmaxf(0.0f, dist - fmaxf(1e-5f, dist * 1e-6f));  // shadow_endpoint (lights.wgsl:253-255)
            if (!visible(p, dir, end)) continue;
            const float tr = in_fog ? fog_transmittance(P.sh, p, dir, dist) : 1.0f;
            float pl = (selected_weight / area_total) * (dist2 / fmaxf(l.area * cos_l, 1e-20f));
            float pb = last ? 0.0f : bsdf_pdf(s, v, dir, p_spec);
            float w = last ? 1.0f : mis_power(pl, pb);
            F3 f = brdf_eval(s, v, dir);
            L = add(L, mk(f.x * nl * l.color[0] * tr * w / pl, f.y * nl * l.color[1] * tr * w / pl,
                          f.z * nl * l.color[2] * tr * w / pl));
        }
    }
    return L;
}

// Path state carried between bounces (the wavefront form keeps it in SoA device planes).
struct PathState {
    F3 o, d, throughput, L, prev_n;
    float prev_pdf, prev_total;
    uint32_t rng, bounce;
    bool prev_from_bsdf, alive;
    // fog: the segment starts at a scattering event (its NEE proposals are the phase's), whose
    // emitters the segment's line integral already counted; the camera segment scattered; the
    // light scattered toward the camera segment; the sample index and pixel stream of its
    // Halton coordinates
    bool prev_from_volume, prev_volume_nee, volume_path;
    F3 primary_volume;
    uint32_t sample, point_stream;
    // the ray footprint carried to the hits for finish detail (texture_width, texture_spread,
    // pathtrace.wgsl:811-814): width at the ray origin and its growth per unit length
    float tex_width, tex_spread;
};

// The session planes of one sample (pathtrace.wgsl:1243-1266) of a job with fog: the volume
// radiance (the camera segment's in-scattering, or the whole path when the camera ray
// scattered), its Welford second moments, and the surface radiance r - volume.  vol holds the
// VOLUME, MOMENTS and SURFACE planes, `pixels` entries each.
__device__ __forceinline__ void accumulate_session(float4* __restrict__ vol, size_t pixels, size_t idx, F3 r,
                                                   const PathState& st, bool ok) {
    const F3 volume = st.volume_path ? st.L : st.primary_volume;
    const F3 v = ok ? volume : mk(0.0f, 0.0f, 0.0f);
    const F3 surface = ok ? sub(r, volume) : mk(0.0f, 0.0f, 0.0f);
    const float4 old = vol[idx];
    const float n = fmaxf(old.w, 1.0f);
    const F3 old_mean = mk(old.x / n, old.y / n, old.z / n);
    const F3 delta = sub(v, old_mean);
    const float n1 = old.w + 1.0f;
    const F3 new_mean = add(old_mean, mk(delta.x / n1, delta.y / n1, delta.z / n1));
    float4 m = vol[pixels + idx];
    m.x += fmaxf(0.0f, delta.x * (v.x - new_mean.x));
    m.y += fmaxf(0.0f, delta.y * (v.y - new_mean.y));
    m.z += fmaxf(0.0f, delta.z * (v.z - new_mean.z));
    m.w += 1.0f;
    vol[pixels + idx] = m;
    vol[idx] = make_float4(old.x + v.x, old.y + v.y, old.z + v.z, old.w + 1.0f);
    float4 sf = vol[2 * pixels + idx];
    vol[2 * pixels + idx] = make_float4(sf.x + surface.x, sf.y + surface.y, sf.z + surface.z, sf.w + 1.0f);
}

__device__ __forceinline__ void path_start(const JobP& P, uint32_t px, uint32_t py, uint32_t sample, PathState& st) {
    st.rng = rng_init(px, py, sample, P.seed);
    camera_ray(P.cam, P.width, P.height, px, py, st.rng, st.o, st.d);
    st.throughput = mk(1.0f, 1.0f, 1.0f);
    st.L = mk(0.0f, 0.0f, 0.0f);
    st.prev_n = mk(0.0f, 0.0f, 0.0f);
    st.prev_pdf = 0.0f;
    st.prev_total = 0.0f;
    st.bounce = 0;
    st.prev_from_bsdf = false;
    st.alive = true;
    st.prev_from_volume = false;
    st.prev_volume_nee = false;
    st.volume_path = false;
    st.primary_volume = mk(0.0f, 0.0f, 0.0f);
    st.sample = sample;
    st.point_stream = pcg(pcg(px ^ P.seed) ^ pcg(py ^ 0x27d4eb2du));  // pathtrace.wgsl:822
    st.tex_width = 0.0f;
    st.tex_spread = P.pixel_cone * 0.5f;
}

// One segment's outcome as the march and emitter test leave it.
struct SegHit {
    bool hit, exhausted;
    float t;
    uint32_t steps, light;
    float trap0;
    float t_prev;  // march position before a surface hit: lower end of the FP64 refine bracket
};

// The body of one bounce after its segment was traced (pathtrace.wgsl:1012-1196 subset):
// emission on a miss or an emitter hit, otherwise the surface (fractal or ground), next-event
// estimation and the continuation sample.  `normal` evaluates the field gradient at a surface
// point; `visible` marches a shadow ray; `sample` evaluates the full field sample at a surface
// hit (called only when the shading reads more than the march's trap).  Sets st.alive = false
// when the path ends.  FOG = false: a kernel for jobs without fog or finish detail patterns
// (neither's code is compiled in, so that kernel keeps its registers for the rest); FOG = true
// also carries the ray footprint to the hits for the detail patterns.
template <bool FOG, class Normal, class Vis, class Sample>
__device__ __forceinline__ void path_bounce(const JobP& P, PathState& st, const SegHit& h, float& first_depth,
                                            float4& first_aux, Normal&& normal, Vis&& visible, Sample&& sample) {
    F3 gn = mk(0.0f, 0.0f, 0.0f);
    const bool ground = h.light == GROUND_HIT;
    if (h.hit) {
        if (ground) gn = ground_normal(st.d);
        else if (h.light != NO_LIGHT) gn = light_normal(P, h.light);
        else gn = normal(add(st.o, scl(st.d, h.t)), fmaxf(h.t * P.pixel_cone * 0.5f, 1e-4f) * P.pixel_tolerance);
        if (st.bounce == 0u) {
            first_depth = h.t;
            first_aux = make_float4(gn.x, gn.y, gn.z, float(h.steps));
        }
    }
    const bool in_fog = FOG && P.sh.fog_density > 0.0f;
    if (in_fog) {  // pathtrace.wgsl:980-1010
        const float t_end = h.hit ? h.t : P.t_end;
        const F3 dv = mul3(st.throughput, fog_segment_direct(P, st.o, st.d, t_end, st.rng, st.sample, st.point_stream,
                                                             st.bounce == 0u, visible));
        st.L = add(st.L, dv);
        if (st.bounce == 0u) st.primary_volume = add(st.primary_volume, dv);
        const float t_s = fog_free_flight(P.sh, st.o, st.d, t_end, st.rng);
        if (t_s < t_end) {  // a scattering event before the segment's end
            if (st.bounce == 0u) st.volume_path = true;
            const F3 p_s = add(st.o, scl(st.d, t_s));
            st.tex_width += t_s * st.tex_spread;  // FOG: finish detail is compiled in as well
            st.tex_spread = fmaxf(st.tex_spread, 0.25f);
            if (st.bounce == 0u && !h.hit) first_depth = t_s;
            const float vat = area_total_weight(P, p_s, mk(0.0f, 0.0f, 0.0f));
            const bool last = st.bounce == P.max_bounces;
            const F3 fc = ld3(P.sh.fog_color);
            st.L = add(st.L, mul3(mul3(st.throughput, fc), fog_direct_light(P, p_s, st.d, vat, last, st.rng, visible)));
            st.throughput = mul3(st.throughput, fc);
            if (last) {
                st.alive = false;
                return;
            }
            const float r1 = rand01(st.rng);
            const float r2 = rand01(st.rng);
            const F3 nd = hg_sample(st.d, P.sh.fog_g, r1, r2);
            st.prev_pdf = hg_phase(dot3(st.d, nd), P.sh.fog_g);
            st.prev_from_bsdf = true;
            st.prev_from_volume = true;
            st.prev_volume_nee = true;
            st.prev_n = mk(0.0f, 0.0f, 0.0f);
            st.prev_total = vat;
            st.o = p_s;
            st.d = nd;
            st.bounce++;
            return;
        }
    }
    if (!h.hit) {
        if (!h.exhausted) {
            st.L = add(st.L, mul3(st.throughput, environment_emission(P, st.d, st.prev_pdf, st.prev_from_bsdf, st.prev_n,
                                                                      st.prev_total, st.prev_from_volume)));
            st.L = add(st.L, mul3(st.throughput, sun_radiance(P, st.d, st.prev_pdf, st.prev_from_bsdf)));
        }
        st.alive = false;
        return;
    }
    const F3 p = add(st.o, scl(st.d, h.t));
    if (!ground && h.light != NO_LIGHT) {
        // a fog segment already counted every emitter path by its line integral
        if (!st.prev_volume_nee)
            st.L = add(st.L, mul3(st.throughput, area_emission(P, st.o, st.d, h.t, st.prev_pdf, st.prev_from_bsdf,
                                                               st.prev_n, st.prev_total, st.prev_from_volume)));
        st.alive = false;
        return;
    }
    const F3 n = gn;
    if (st.bounce == 0u && !ground) {
        float t = clampT(float(h.steps) / 96.0f, 0.0f, 1.0f);
        st.throughput = scl(st.throughput, 1.0f - P.cavity_darkness * (t * t * (3.0f - 2.0f * t)));
    }
    const F3 v = neg3(st.d);
    const ShadeSample smp =
        ground ? ground_sample(P) : (P.sh.full_sample ? sample(p) : carried_sample(P, h.trap0));
    F3 emission;
    if (FOG) st.tex_width += h.t * st.tex_spread;
    const Surf s = make_surface<FOG>(P, p, n, smp, fmaxf(st.tex_width, 1e-4f), emission);
    st.L = add(st.L, mul3(st.throughput, emission));
    const float offset = ground ? 1e-3f : fmaxf(h.t * P.pixel_cone * 2.0f, 1e-3f) * P.pixel_tolerance;
    const F3 p_off = add(p, scl(n, offset));
    const float nv = fmaxf(dot3(s.n, v), 1e-4f);
    const float p_spec = spec_probability(s, nv);
    const float area_total = area_total_weight(P, p_off, s.n);
    const bool last = st.bounce == P.max_bounces;
    F3 dl = direct_light(P, p_off, s, n, v, p_spec, area_total, last, in_fog, st.rng, visible);
    st.L = add(st.L, mul3(st.throughput, dl));
    if (last) {
        st.alive = false;
        return;
    }
    float r0 = rand01(st.rng), r1 = rand01(st.rng), r2 = rand01(st.rng);
    F3 l;
    if (r0 < p_spec) l = reflect3(neg3(v), sample_ggx_h(s.n, v, s.alpha, r1, r2));
    else l = cosine_sample(s.n, r1, r2);
    float nl = dot3(s.n, l);
    if (nl <= 0.0f || dot3(n, l) <= 0.0f) {
        st.alive = false;
        return;
    }
    float pdf = bsdf_pdf(s, v, l, p_spec);
    if (pdf <= 1e-6f) {
        st.alive = false;
        return;
    }
    F3 f = brdf_eval(s, v, l);
    st.throughput = mk(st.throughput.x * (f.x * nl / pdf), st.throughput.y * (f.y * nl / pdf),
                       st.throughput.z * (f.z * nl / pdf));
    st.prev_pdf = pdf;
    st.prev_from_bsdf = true;
    st.prev_n = s.n;
    st.prev_total = area_total;
    st.prev_from_volume = false;
    st.prev_volume_nee = false;
    if (st.bounce >= 2u) {
        float q = clampT(fmaxf(st.throughput.x, fmaxf(st.throughput.y, st.throughput.z)), 0.05f, 0.95f);
        if (rand01(st.rng) > q) {
            st.alive = false;
            return;
        }
        st.throughput = mk(st.throughput.x / q, st.throughput.y / q, st.throughput.z / q);
    }
    st.o = p_off;
    st.d = l;
    if (FOG) st.tex_spread = fmaxf(st.tex_spread, s.alpha * 0.25f);
    st.bounce++;
}

// ---------------------------------------------------------------- march in flight
// march() (device.hpp) cut at each field evaluation, for kernels that run many marches per
// lane as state machines (the flattened path tracer, the wavefront intersect stage).

// One march in flight: the state of march() between two evaluations.
struct MarchState {
    F3 ro, rd;
    float t, omega, prev_radius, step_len, cand_t, cand_err, cand_trap0, t_last, t_end;
    uint32_t i;
};

// How a march ended: as Hit<float> of march().
struct MarchEnd {
    bool hit, exhausted;
    float t, t_prev, trap0;
    uint32_t steps;
};

__device__ __forceinline__ void march_reset(MarchState& m, F3 ro, F3 rd, float t_end) {
    m.ro = ro;
    m.rd = rd;
    m.t = 0.0f;
    m.omega = 1.2f;
    m.prev_radius = 0.0f;
    m.step_len = 0.0f;
    m.cand_t = 0.0f;
    m.cand_err = 1e30f;
    m.cand_trap0 = 1.0f;
    m.t_last = 0.0f;
    m.t_end = t_end;
    m.i = 0;
}

// One evaluated march step (the loop body of march()).  Returns true when the march ended.
__device__ __forceinline__ bool march_advance(MarchState& m, const De<float>& s, float safety, float hit_cone,
                                              uint32_t max_steps, MarchEnd& e) {
    const uint32_t steps = m.i + 1u;
    const float radius = fabsf(s.d) * safety;
    const bool sor_fail = m.omega > 1.0f && (radius + m.prev_radius) < m.step_len * 1.0001f;
    if (sor_fail) {
        m.step_len = m.step_len - m.omega * m.step_len;
        m.omega = 1.0f;
    } else {
        m.step_len = radius * m.omega;
    }
    m.prev_radius = radius;
    const float err = radius / fmaxf(m.t, 1e-6f);
    if (!sor_fail && err < m.cand_err) {
        m.cand_t = m.t;
        m.cand_trap0 = s.trap0;
        m.cand_err = err;
    }
    if (!sor_fail && err < hit_cone) {
        e = MarchEnd{true, false, m.t, m.t_last, s.trap0, steps};
        return true;
    }
    if (!sor_fail && m.t + m.step_len >= m.t_end) {
        if (radius >= m.t_end - m.t) {
            e = MarchEnd{false, false, m.t_end, 0.0f, 1.0f, steps};
            return true;
        }
        m.omega = 1.0f;
        m.step_len = radius;
    }
    m.t_last = m.t;
    m.t = m.t + m.step_len;
    ++m.i;
    if (m.i >= max_steps) {
        if (m.cand_err < hit_cone * 4.0f) e = MarchEnd{true, false, m.cand_t, m.cand_t, m.cand_trap0, steps};
        else e = MarchEnd{false, true, m.t_end, 0.0f, 1.0f, steps};
        return true;
    }
    if (m.t >= m.t_end) {
        e = MarchEnd{false, false, m.t_end, 0.0f, 1.0f, steps};
        return true;
    }
    return false;
}

}  // namespace syf
