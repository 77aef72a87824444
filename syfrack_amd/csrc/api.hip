// C ABI of the library: contexts, jobs, lowering of a job description, launches, readback.
//
// Host-side lowering restates Uniforms::from_scene (uniforms.rs:748-853), Camera::basis
// (camera.rs:58-104), GpuLight::from_light (uniforms.rs:640-745) and the validation of
// RenderConfig::validate (renderer.rs:81-138) plus the graph bounds (graph/mod.rs:44-73).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "api_internal.hpp"
#include "kernels.hpp"

namespace syf {

static thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    int code = (e == hipErrorOutOfMemory) ? SYF_E_DEVICE : SYF_E_DEVICE;
    return fail(code, "%s: %s (%s)", what, hipGetErrorName(e), hipGetErrorString(e));
}

void clear_error() { g_error.clear(); }

// ---------------------------------------------------------------- lowering helpers (host f32)

namespace {
struct H3 {
    float x, y, z;
};
inline H3 h3(const float* p) { return {p[0], p[1], p[2]}; }
inline H3 operator+(H3 a, H3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline H3 operator-(H3 a, H3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline H3 operator*(H3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(H3 a, H3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len(H3 a) { return std::sqrt(dot(a, a)); }
inline H3 cross(H3 a, H3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline H3 normalize(H3 a) { return a * (1.0f / len(a)); }
// glam normalize_or(fallback) / try_normalize().unwrap_or(fallback)
inline H3 normalize_or(H3 a, H3 fb) {
    float rcp = 1.0f / len(a);
    return (std::isfinite(rcp) && rcp > 0.0f) ? a * rcp : fb;
}
inline void st3(float* d, H3 v) {
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
}
inline float rad(float deg) { return deg * (3.14159265358979323846f / 180.0f); }
inline bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }
}  // namespace

int lower_fractal(const syf_fractal& f, FractalP& F) {
    if (f.family > SYF_FAMILY_SPHERE) return fail(SYF_E_INVALID, "fractal.family: unknown family %u", f.family);
    if (f.family != SYF_FAMILY_SPHERE) {
        if (f.iterations < 1 || f.iterations > 64)
            return fail(SYF_E_INVALID, "fractal.iterations must be between 1 and 64, got %u", f.iterations);
        if (!(f.bailout >= 2.0f && f.bailout <= 1e6f))
            return fail(SYF_E_INVALID, "fractal.bailout must be between 2 and 1e6, got %g", double(f.bailout));
    }
    if (f.family == SYF_FAMILY_MANDELBULB && !(f.power >= 2.0f && f.power <= 16.0f))
        return fail(SYF_E_INVALID, "fractal.power must be between 2 and 16, got %g", double(f.power));
    if (f.family == SYF_FAMILY_MANDELBOX) {
        if (!(f.scale >= -4.0f && f.scale <= 4.0f))
            return fail(SYF_E_INVALID, "fractal.scale must be between -4 and 4, got %g", double(f.scale));
        if (!(f.min_radius >= 0.05f && f.min_radius <= 1.0f))
            return fail(SYF_E_INVALID, "fractal.min_radius must be between 0.05 and 1, got %g", double(f.min_radius));
        if (!(f.fixed_radius > 0.0f) || !std::isfinite(f.fixed_radius))
            return fail(SYF_E_INVALID, "fractal.fixed_radius must be positive, got %g", double(f.fixed_radius));
        if (!(f.fold_limit > 0.0f) || !std::isfinite(f.fold_limit))
            return fail(SYF_E_INVALID, "fractal.fold_limit must be positive, got %g", double(f.fold_limit));
    }
    if (f.family == SYF_FAMILY_MENGER && !(f.scale >= 2.0f && f.scale <= 4.0f))
        return fail(SYF_E_INVALID, "fractal.scale (Menger) must be between 2 and 4, got %g", double(f.scale));
    if (!(f.safety > 0.0f && f.safety <= 1.0f))
        return fail(SYF_E_INVALID, "fractal.safety must be in (0, 1], got %g", double(f.safety));
    if (!(f.trap_scale > 0.0f) || !std::isfinite(f.trap_scale))
        return fail(SYF_E_INVALID, "fractal.trap_scale must be positive, got %g", double(f.trap_scale));
    if (!finite3(f.offset) || !finite3(f.half) || !finite3(f.plane_normal))
        return fail(SYF_E_INVALID, "fractal: offset/half/plane_normal must be finite");
    std::memset(&F, 0, sizeof F);
    F.family = f.family;
    F.iterations = f.family == SYF_FAMILY_SPHERE ? 0u : f.iterations;
    F.iteration_product = f.family == SYF_FAMILY_SPHERE ? 1u : f.iterations;  // validate.rs:112-137
    F.ipower = (f.power >= 1.0f && f.power <= 16.0f && f.power == std::floor(f.power)) ? uint32_t(f.power) : 0u;
    F.bailout = f.bailout;
    F.power = f.power;
    F.scale = f.scale;
    F.abs_scale = std::fabs(f.scale);
    F.mr2 = f.min_radius * f.min_radius;
    F.fr2 = f.fixed_radius * f.fixed_radius;
    F.fr2_over_mr2 = F.fr2 / F.mr2;
    F.fold_limit = f.fold_limit;
    const float s1 = f.scale - 1.0f;
    for (int i = 0; i < 3; ++i) F.off_s1[i] = f.offset[i] * s1;
    F.menger_shift_threshold = -0.5f * f.offset[2] * s1;
    for (int i = 0; i < 3; ++i) F.half[i] = f.half[i];
    F.radius = f.radius;
    st3(F.plane_n, normalize_or(h3(f.plane_normal), H3{0.0f, 1.0f, 0.0f}));
    F.plane_offset = f.plane_offset;
    F.trap_scale = f.trap_scale;
    F.safety = f.safety;  // step_scale = safety / Lipschitz; the classics have no Twist/Bend
    return SYF_OK;
}

// Colouring, finishes and ground (GpuColoring::from_coloring, uniforms.rs:194-214; GpuRamp,
// :31-55; finishes, :787-817; GpuAtmosphere.ground, :227-230), and whether the shading reads
// more of a surface hit's field sample than the one trap the march returns (ShadeP::full_sample).
int lower_shading(const syf_job_desc& d, bool graph, JobP& P) {
    ShadeP& S = P.sh;
    const syf_coloring& c = d.coloring;
    if (c.slot > SYF_COLOR_ORBIT && c.slot != SYF_COLOR_NONE)
        return fail(SYF_E_INVALID, "coloring.slot must be a trap slot 0..3, SYF_COLOR_ITER, _NORMAL, _ORBIT or _NONE, got %u",
                    c.slot);
    if (c.stop_count > SYF_MAX_STOPS)
        return fail(SYF_E_INVALID, "coloring.stop_count must be at most %d, got %u", SYF_MAX_STOPS, c.stop_count);
    // a ramp without stops is white whatever its mapping (an all-zero coloring is no colouring)
    if (c.stop_count > 0u &&
        (!std::isfinite(c.frequency) || !std::isfinite(c.phase) || !(c.gamma > 0.0f) || !std::isfinite(c.gamma)))
        return fail(SYF_E_INVALID, "coloring: frequency and phase must be finite and gamma positive");
    for (uint32_t i = 0; i < c.stop_count; ++i) {
        if (!finite3(c.stops[i]) || !std::isfinite(c.stops[i][3]))
            return fail(SYF_E_INVALID, "coloring.stops[%u] must be finite", i);
        if (c.positioned && i > 0 && !(c.stops[i][3] > c.stops[i - 1][3]))
            return fail(SYF_E_INVALID, "coloring.stops[%u]: positions must increase strictly", i);
    }
    S.color_slot = c.stop_count > 0u ? c.slot : SYF_COLOR_NONE;
    S.stop_count = c.stop_count;
    S.cyclic = c.cyclic ? 1u : 0u;
    S.smooth = c.smooth ? 1u : 0u;
    S.positioned = c.positioned ? 1u : 0u;
    S.frequency = c.frequency;
    S.phase = c.phase;
    S.gamma = c.gamma;
    const uint32_t n = c.stop_count;
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) S.stops[i][k] = c.stops[i][k];
        // Ramp::position (palette.rs:106-115): given, or evenly spaced
        S.stops[i][3] = c.positioned ? c.stops[i][3]
                                     : float(i) / float(c.cyclic ? n : (n > 1u ? n - 1u : 1u));
    }
    for (int k = 0; k < 4; ++k) S.weights[k] = c.weights[k];
    S.iteration_weight = c.iteration_weight;
    S.normal_weight = c.normal_weight;
    S.logarithmic = c.logarithmic;
    if (d.finish_count > SYF_MAX_FINISHES)
        return fail(SYF_E_INVALID, "finish_count must be at most %d, got %u", SYF_MAX_FINISHES, d.finish_count);
    S.finish_count = d.finish_count;
    for (uint32_t i = 0; i < d.finish_count; ++i) {
        const syf_finish& f = d.finishes[i];
        if (!finite3(f.albedo) || !finite3(f.emission) || !std::isfinite(f.roughness) || !std::isfinite(f.metallic) ||
            !std::isfinite(f.specular))
            return fail(SYF_E_INVALID, "finishes[%u] must be finite", i);
        if (f.oren_nayar > 1u) return fail(SYF_E_INVALID, "finishes[%u].oren_nayar must be 0 or 1", i);
        if (f.pattern > SYF_PATTERN_WET_CONCRETE)
            return fail(SYF_E_INVALID, "finishes[%u].pattern must be 0..%u, got %u", i, SYF_PATTERN_WET_CONCRETE,
                        f.pattern);
        if (f.pattern != SYF_PATTERN_NONE) {
            // FinishDetail bounds (material.rs:223-250)
            if (f.seed > 65535u) return fail(SYF_E_INVALID, "finishes[%u].detail.seed must be 0..65535, got %u", i, f.seed);
            if (!(f.scale >= 0.01f && f.scale <= 200.0f))
                return fail(SYF_E_INVALID, "finishes[%u].detail.scale must be between 0.01 and 200, got %g", i,
                            double(f.scale));
            if (!(f.strength >= 0.0f && f.strength <= 1.0f))
                return fail(SYF_E_INVALID, "finishes[%u].detail.strength must be between 0 and 1, got %g", i,
                            double(f.strength));
            if (!(f.bump >= -1.0f && f.bump <= 1.0f))
                return fail(SYF_E_INVALID, "finishes[%u].detail.bump must be between -1 and 1, got %g", i, double(f.bump));
            if (!(f.roughness_delta >= -1.0f && f.roughness_delta <= 1.0f))
                return fail(SYF_E_INVALID, "finishes[%u].detail.roughness_delta must be between -1 and 1, got %g", i,
                            double(f.roughness_delta));
            for (int k = 0; k < 3; ++k)
                if (!(f.tint[k] >= 0.0f && f.tint[k] <= 1.0f))
                    return fail(SYF_E_INVALID, "finishes[%u].detail.tint must be within [0, 1]", i);
        }
        FinishP& F = S.finishes[i];
        for (int k = 0; k < 3; ++k) {
            F.albedo[k] = f.albedo[k];
            F.emission[k] = f.emission[k];
        }
        F.roughness = f.roughness;
        F.metallic = f.metallic;
        F.specular = f.specular;
        F.sigma_scale = (f.oren_nayar || f.pattern >= SYF_PATTERN_CONCRETE) ? 0.45f : 0.0f;
        F.pattern = f.pattern;
        F.seed = f.pattern ? f.seed : 0u;
        F.scale = f.pattern ? f.scale : 1.0f;
        F.strength = f.pattern ? f.strength : 0.0f;
        F.bump = f.pattern ? f.bump : 0.0f;
        F.roughness_delta = f.pattern ? f.roughness_delta : 0.0f;
        for (int k = 0; k < 3; ++k) F.tint[k] = f.pattern ? f.tint[k] : 0.0f;
        if (f.pattern) S.detail = 1u;
    }
    S.ground_enabled = d.ground.enabled ? 1u : 0u;
    if (d.ground.enabled && !std::isfinite(d.ground.height))
        return fail(SYF_E_INVALID, "ground.height must be finite");
    // the fog's bounds (DENSITY, HEIGHT, FALLOFF, ANISOTROPY: atmosphere.rs:115-118)
    const syf_fog& fog = d.fog;
    if (!(fog.density >= 0.0f && fog.density <= 10.0f))
        return fail(SYF_E_INVALID, "fog.density must be between 0 and 10, got %g", double(fog.density));
    if (fog.density > 0.0f) {
        if (!(fog.height >= -100.0f && fog.height <= 100.0f))
            return fail(SYF_E_INVALID, "fog.height must be between -100 and 100, got %g", double(fog.height));
        if (!(fog.falloff >= 0.0f && fog.falloff <= 10.0f))
            return fail(SYF_E_INVALID, "fog.falloff must be between 0 and 10, got %g", double(fog.falloff));
        if (!(fog.anisotropy >= -0.99f && fog.anisotropy <= 0.99f))
            return fail(SYF_E_INVALID, "fog.anisotropy must be between -0.99 and 0.99, got %g", double(fog.anisotropy));
        if (!finite3(fog.color) || !(fog.color[0] >= 0.0f && fog.color[1] >= 0.0f && fog.color[2] >= 0.0f))
            return fail(SYF_E_INVALID, "fog.color must be finite and non-negative");
        S.fog_density = fog.density;
        S.fog_height = fog.height;
        S.fog_falloff = fog.falloff;
        S.fog_g = fog.anisotropy;
        for (int k = 0; k < 3; ++k) S.fog_color[k] = fog.color[k];
    }
    if (d.ground.enabled && d.ground.finish > SYF_MAX_FINISHES)  // MATERIAL_INDEX bounds
        return fail(SYF_E_INVALID, "ground.finish must be 0..%d, got %u", SYF_MAX_FINISHES, d.ground.finish);
    S.ground_height = d.ground.height;
    S.ground_finish = d.ground.finish;
    // the trap the march's field sample carries: a graph's shading slot, a classic family's slot 0
    // (a classic family's other slots are 1, unwritten, and its material 0)
    const uint32_t carried = graph ? P.f.gshade_slot : 0u;
    const uint32_t cs = S.color_slot;
    const uint32_t ss = d.material.surface_slot;
    const bool color_full = cs == SYF_COLOR_ITER || cs == SYF_COLOR_ORBIT || (graph && cs < 4u && cs != carried);
    const bool surface_full =
        d.material.surface_enabled && (ss == SYF_COLOR_ITER || (graph && ss < 4u && ss != carried));
    const bool material_full = graph && d.finish_count > 0u;
    S.full_sample = (color_full || surface_full || material_full) ? 1u : 0u;
    return SYF_OK;
}

int lower_job(const syf_job_desc& d, JobP& P, const FractalP* graph_fractal) {
    std::memset(&P, 0, sizeof P);
    if (d.fractal.family == SYF_FAMILY_GRAPH && !graph_fractal)
        return fail(SYF_E_INVALID, "fractal.family: a node graph (SYF_FAMILY_GRAPH) is given to syf_job_create_graph, "
                                   "this entry point takes the classic families");
    int rc = graph_fractal ? SYF_OK : lower_fractal(d.fractal, P.f);
    if (graph_fractal) P.f = *graph_fractal;
    if (rc != SYF_OK) return rc;
    if (d.width < 1) return fail(SYF_E_INVALID, "width must be at least 1, got %u", d.width);
    if (d.height < 1) return fail(SYF_E_INVALID, "height must be at least 1, got %u", d.height);
    if (d.width > 32768 || d.height > 32768)
        return fail(SYF_E_INVALID, "width x height (%ux%u) exceeds 32768 on a side", d.width, d.height);
    if (d.max_bounces > 32) return fail(SYF_E_INVALID, "max_bounces must be at most 32, got %u", d.max_bounces);
    if (d.max_steps < 1) return fail(SYF_E_INVALID, "max_steps must be at least 1, got %u", d.max_steps);
    if (!(d.pixel_tolerance >= 0.125f && d.pixel_tolerance <= 1.0f))
        return fail(SYF_E_INVALID, "pixel_tolerance must be between 0.125 and 1.0, got %g", double(d.pixel_tolerance));
    if (!(d.cavity_darkness >= 0.0f && d.cavity_darkness <= 1.0f))
        return fail(SYF_E_INVALID, "cavity_darkness must be between 0 and 1, got %g", double(d.cavity_darkness));
    if (d.march_math != SYF_MATH_FAST && d.march_math != SYF_MATH_EXACT)
        return fail(SYF_E_INVALID, "march_math must be SYF_MATH_FAST (0) or SYF_MATH_EXACT (1), got %u", d.march_math);
    if (d.light_count > SYF_MAX_LIGHTS)
        return fail(SYF_E_INVALID, "light_count must be at most %d, got %u", SYF_MAX_LIGHTS, d.light_count);
    const syf_camera& c = d.camera;
    if (c.mode > SYF_CAMERA_ORBIT) return fail(SYF_E_INVALID, "camera.mode: unknown mode %u", c.mode);
    if (!(c.focal_mm >= 8.0f && c.focal_mm <= 1200.0f))
        return fail(SYF_E_INVALID, "camera.focal_mm must be between 8 and 1200, got %g", double(c.focal_mm));
    if (!(c.f_number >= 0.0f && c.f_number <= 64.0f))
        return fail(SYF_E_INVALID, "camera.f_number must be between 0 and 64, got %g", double(c.f_number));

    // Camera::basis (camera.rs:58-104).
    H3 origin, target;
    if (c.mode == SYF_CAMERA_ORBIT) {
        float az = rad(c.azimuth), el = rad(c.elevation);
        H3 off = H3{std::cos(el) * std::sin(az), std::sin(el), std::cos(el) * std::cos(az)} * c.distance;
        target = h3(c.point);
        origin = target + off;
    } else {
        origin = h3(c.position);
        target = h3(c.target);
    }
    H3 forward = normalize_or(target - origin, H3{0.0f, 0.0f, -1.0f});
    H3 world_up = std::fabs(forward.y) > 0.9999f ? H3{0.0f, 0.0f, 1.0f} : H3{0.0f, 1.0f, 0.0f};
    H3 right0 = normalize(cross(forward, world_up));
    H3 up0 = cross(right0, forward);
    float sr = std::sin(rad(c.roll)), cr = std::cos(rad(c.roll));
    st3(P.cam.origin, origin);
    st3(P.cam.forward, forward);
    st3(P.cam.right, right0 * cr + up0 * sr);
    st3(P.cam.up, up0 * cr - right0 * sr);
    P.cam.tan_half = 18.0f / c.focal_mm;
    P.cam.aspect = float(d.width) / float(d.height);
    P.cam.lens_radius = c.f_number > 0.0f ? (c.focal_mm / 1000.0f) / (2.0f * c.f_number) : 0.0f;
    P.cam.focus = c.focus_distance;

    // Lights (GpuLight::from_light).
    P.light_count = d.light_count;
    for (uint32_t i = 0; i < d.light_count; ++i) {
        const syf_light& l = d.lights[i];
        LightP& L = P.lights[i];
        L.kind = l.kind;
        if (l.kind == SYF_LIGHT_AREA) {
            if (!(l.size[0] > 0.0f && l.size[1] > 0.0f))
                return fail(SYF_E_INVALID, "lights[%u].size must be positive", i);
            H3 n = normalize_or(h3(l.normal), H3{0.0f, 1.0f, 0.0f});
            H3 helper = std::fabs(n.y) > 0.999f ? H3{0.0f, 0.0f, 1.0f} : H3{0.0f, 1.0f, 0.0f};
            H3 u = normalize(cross(helper, n));
            H3 v = cross(n, u);
            st3(L.a, h3(l.position));
            st3(L.b, u * l.size[0]);
            st3(L.c, v * l.size[1]);
            H3 cbc = cross(h3(L.b), h3(L.c));
            float lc = len(cbc), rcp = 1.0f / lc;  // safe_normalize (common.wgsl:242-247)
            st3(L.n, (lc > 0.0f && rcp > 0.0f && rcp <= 3.4028235e38f) ? cbc * rcp : H3{0.0f, 1.0f, 0.0f});
            L.area = lc;
            st3(L.color, h3(l.color) * l.intensity);
        } else if (l.kind == SYF_LIGHT_ENV) {
            st3(L.color, h3(l.zenith) * l.intensity);
            st3(L.horizon, h3(l.horizon) * l.intensity);
            st3(L.ground, h3(l.ground) * l.intensity);
        } else if (l.kind == SYF_LIGHT_SUN) {  // GpuLight::from_light, Light::Sun (uniforms.rs:656-668)
            // ANGULAR_SIZE bounds (scene/light.rs:126)
            if (!(l.size[0] >= 0.01f && l.size[0] <= 90.0f))
                return fail(SYF_E_INVALID, "lights[%u].size[0] (the sun's angular diameter) must be between 0.01 and 90 degrees, got %g",
                            i, double(l.size[0]));
            st3(L.a, normalize_or(h3(l.normal), H3{0.0f, 1.0f, 0.0f}));
            L.cos_max = std::cos(rad(l.size[0] * 0.5f));
            st3(L.color, h3(l.color) * l.intensity);
        } else {
            return fail(SYF_E_INVALID, "lights[%u].kind: unsupported kind %u (sun, area or environment)", i, l.kind);
        }
    }
    const syf_material& m = d.material;
    for (int i = 0; i < 3; ++i) P.mat.albedo[i] = m.albedo[i];
    P.mat.roughness = m.roughness;
    P.mat.metallic = m.metallic;
    P.mat.specular = m.specular;
    P.mat.surface_enabled = m.surface_enabled;
    if (m.surface_enabled && m.surface_slot > SYF_COLOR_NORMAL)
        return fail(SYF_E_INVALID, "material.surface_slot must be a trap slot 0..3, SYF_COLOR_ITER or SYF_COLOR_NORMAL, got %u",
                    m.surface_slot);
    P.mat.surface_slot = m.surface_slot;
    if (graph_fractal) {
        // a graph's field evaluation returns the trap of the shading's slot as its first trap
        P.f.gshade_slot = m.surface_enabled && m.surface_slot < 4u ? m.surface_slot : 0u;
        P.f.trap_scale = P.f.gtrap_scale[P.f.gshade_slot];
    }
    if (int rc = lower_shading(d, graph_fractal != nullptr, P); rc != SYF_OK) return rc;
    P.mat.surface[0] = m.surface_roughness[0];
    P.mat.surface[1] = m.surface_roughness[1];
    P.mat.surface[2] = m.surface_metallic[0];
    P.mat.surface[3] = m.surface_metallic[1];

    P.width = d.width;
    P.height = d.height;
    P.seed = uint32_t(d.seed) ^ uint32_t(d.seed >> 32);  // uniforms.rs:824
    P.max_bounces = d.max_bounces;
    P.max_steps = d.max_steps;
    P.pixel_tolerance = d.pixel_tolerance;
    P.pixel_cone = P.cam.tan_half * 2.0f / float(d.width);
    float p0z = std::fmax(2.0f * d.fractal.bounding_radius_hint + len(origin), 200.0f);  // uniforms.rs:781
    P.t_end = std::fmax(200.0f, p0z);
    P.cavity_darkness = d.cavity_darkness;
    P.march_math = d.march_math;
    return SYF_OK;
}

size_t plane_bytes(uint32_t plane, size_t pixels) {
    switch (plane) {
    case SYF_PLANE_ACCUM: return pixels * 16;
    case SYF_PLANE_DEPTH: return pixels * 4;
    case SYF_PLANE_AUX: return pixels * 16;
    case SYF_PLANE_DEPTH64: return pixels * 8;
    case SYF_PLANE_STEPS: return pixels * 4;
    case SYF_PLANE_VOLUME:
    case SYF_PLANE_MOMENTS:
    case SYF_PLANE_SURFACE: return pixels * 16;
    default: return 0;
    }
}

}  // namespace syf

using namespace syf;

// ================================================================ C ABI

extern "C" {

const char* syf_last_error(void) { return g_error.c_str(); }
int syf_abi_version(void) { return SYF_ABI_VERSION; }

int syf_ctx_create(int device, syf_ctx** out) {
    clear_error();
    if (!out) return fail(SYF_E_INVALID, "out must not be null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(SYF_E_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(SYF_E_NO_DEVICE, "device %d out of range (%d devices)", device, n);
    e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    void* scratch = nullptr;
    if (int rc = noalloc_setup(device, &scratch); rc != SYF_OK) return rc;
    syf_ctx* c = new syf_ctx();
    c->device = device;
    c->scratch = scratch;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->cu_count = prop.multiProcessorCount;
        std::snprintf(c->arch, sizeof c->arch, "%s", prop.gcnArchName);
    }
    *out = c;
    return SYF_OK;
}

int syf_ctx_destroy(syf_ctx* ctx) {
    clear_error();
    delete ctx;
    return SYF_OK;
}

int syf_job_create(syf_ctx* ctx, const syf_job_desc* desc, syf_job** out) {
    clear_error();
    if (!ctx || !desc || !out) return fail(SYF_E_INVALID, "ctx, desc and out must not be null");
    JobP P;
    int rc = lower_job(*desc, P);
    if (rc != SYF_OK) return rc;
    return job_alloc(ctx, *desc, P, out);
}

}  // extern "C"

int syf::job_alloc(syf_ctx* ctx, const syf_job_desc& dref, const JobP& P, syf_job** out) {
    const syf_job_desc* desc = &dref;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    int rc = SYF_OK;
    syf_job* j = new syf_job();
    j->ctx = ctx;
    j->desc = *desc;
    j->P = P;
    j->pixels = size_t(desc->width) * desc->height;
    j->hitq_cap = hitq_capacity(desc->width, desc->height, ctx->cu_count);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
        size_t need = 0;
        for (uint32_t p = 0; p <= SYF_PLANE_STEPS; ++p) need += plane_bytes(p, j->pixels);
        if (need > free_b) {
            delete j;
            return fail(SYF_E_DEVICE, "width x height (%ux%u) needs %zu plane bytes, above the %zu free on the device",
                        desc->width, desc->height, need, free_b);
        }
    }
    if ((e = hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete j;
        return hip_fail(e, "hipStreamCreate");
    }
    for (uint32_t p = 0; p <= SYF_PLANE_STEPS; ++p) {
        if ((e = hipMalloc(&j->plane[p], plane_bytes(p, j->pixels))) != hipSuccess) {
            syf_job_destroy(j);
            return hip_fail(e, "hipMalloc plane");
        }
    }
    // a job with fog accumulates the session planes; without, they follow from ACCUM (syf_readback)
    if (P.sh.fog_density > 0.0f && (e = hipMalloc(&j->vplanes, 3 * j->pixels * 16)) != hipSuccess) {
        syf_job_destroy(j);
        return hip_fail(e, "hipMalloc session planes");
    }
    // the primary trace's counters and the work counters share one allocation (one clear per trace)
    if ((e = hipMalloc(&j->counters, COUNTER_BYTES)) != hipSuccess ||
        (e = hipMalloc(&j->hitq, (HITQ_FIELDS * j->hitq_cap + j->hitq_cap / 64) * sizeof(uint32_t))) != hipSuccess ||
        (e = hipEventCreate(&j->ev0)) != hipSuccess || (e = hipEventCreate(&j->ev1)) != hipSuccess ||
        (e = hipEventCreate(&j->ev2)) != hipSuccess) {
        syf_job_destroy(j);
        return hip_fail(e, "job allocation");
    }
    j->ctr = reinterpret_cast<unsigned long long*>(j->counters + COUNTER_WORDS);
    if ((e = hipMalloc(&j->dP, sizeof(JobP))) != hipSuccess ||
        (e = hipMemcpy(j->dP, &j->P, sizeof(JobP), hipMemcpyHostToDevice)) != hipSuccess) {
        syf_job_destroy(j);
        return hip_fail(e, "job constants");
    }
    rc = syf_job_clear(j);
    if (rc != SYF_OK) {
        syf_job_destroy(j);
        return rc;
    }
    *out = j;
    return SYF_OK;
}

extern "C" {

int syf_job_destroy(syf_job* j) {
    clear_error();
    if (!j) return SYF_OK;
    // teardown: every release is attempted, whatever an earlier one returned
    (void)hipSetDevice(j->ctx->device);
    if (j->stream) (void)hipStreamSynchronize(j->stream);
    for (void*& p : j->plane)
        if (p) (void)hipFree(p);
    if (j->vplanes) (void)hipFree(j->vplanes);
    if (j->hitq) (void)hipFree(j->hitq);
    if (j->counters) (void)hipFree(j->counters);  // holds ctr too
    if (j->ev0) (void)hipEventDestroy(j->ev0);
    if (j->ev1) (void)hipEventDestroy(j->ev1);
    if (j->ev2) (void)hipEventDestroy(j->ev2);
    release_wavefront(j);
    if (j->dP) (void)hipFree(j->dP);
    if (j->gprog) (void)hipFree(j->gprog);
    if (j->stream) (void)hipStreamDestroy(j->stream);
    delete j;
    return SYF_OK;
}

int syf_job_clear(syf_job* j) {
    clear_error();
    if (!j) return fail(SYF_E_INVALID, "job must not be null");
    if (hipError_t de = hipSetDevice(j->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    hipError_t e = hipSuccess;
    for (uint32_t p = 0; p <= SYF_PLANE_STEPS && e == hipSuccess; ++p)
        e = hipMemsetAsync(j->plane[p], 0, plane_bytes(p, j->pixels), j->stream);
    if (e == hipSuccess && j->vplanes) e = hipMemsetAsync(j->vplanes, 0, 3 * j->pixels * 16, j->stream);
    if (e == hipSuccess) e = hipMemsetAsync(j->ctr, 0, sizeof(unsigned long long) * CTR_COUNT, j->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(j->stream);
    return e == hipSuccess ? SYF_OK : hip_fail(e, "syf_job_clear");
}

int syf_job_sync(syf_job* j) {
    clear_error();
    if (!j) return fail(SYF_E_INVALID, "job must not be null");
    hipError_t e = hipStreamSynchronize(j->stream);
    return e == hipSuccess ? SYF_OK : hip_fail(e, "hipStreamSynchronize");
}

int syf_job_stats(syf_job* j, syf_stats* out) {
    clear_error();
    if (!j || !out) return fail(SYF_E_INVALID, "job and out must not be null");
    if (hipError_t de = hipSetDevice(j->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    unsigned long long c[CTR_COUNT];
    uint32_t qerr = 0;
    hipError_t e = hipMemcpyAsync(c, j->ctr, sizeof c, hipMemcpyDeviceToHost, j->stream);
    if (e == hipSuccess && j->traced)
        e = hipMemcpyAsync(&qerr, j->counters + COUNTER_QERR, sizeof qerr, hipMemcpyDeviceToHost, j->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(j->stream);
    if (e != hipSuccess) return hip_fail(e, "syf_job_stats");
    if (qerr)
        return fail(SYF_E_QUEUE, "primary trace hit queue invariant broken: %u blocks above its capacity of %zu entries",
                    qerr, j->hitq_cap);
    float ms = 0.0f, ms_dom = 0.0f;
    if (j->timed && hipEventElapsedTime(&ms, j->ev0, j->ev2) != hipSuccess) ms = 0.0f;
    if (j->timed && hipEventElapsedTime(&ms_dom, j->ev0, j->ev1) != hipSuccess) ms_dom = 0.0f;
    std::memset(out, 0, sizeof *out);
    out->rays = c[CTR_RAYS];
    out->primary_rays = c[CTR_PRIMARY];
    out->hits = c[CTR_HITS];
    out->de_evals = c[CTR_DE_EVALS] + c[CTR_HIT_EVALS];
    out->de_iterations = c[CTR_DE_ITERS] + c[CTR_HIT_ITERS];
    out->hit_de_evals = c[CTR_HIT_EVALS];
    out->hit_de_iterations = c[CTR_HIT_ITERS];
    out->tail_hits = c[CTR_TAIL_HITS];
    out->tail_hit_de_evals = c[CTR_TAIL_EVALS];
    out->tail_hit_de_iterations = c[CTR_TAIL_ITERS];
    out->tail_de64_evals = c[CTR_TAIL_DE64_EVALS];
    out->tail_de64_iterations = c[CTR_TAIL_DE64_ITERS];
    out->de64_evals = c[CTR_DE64_EVALS];
    out->de64_iterations = c[CTR_DE64_ITERS];
    out->steps = c[CTR_STEPS];
    out->kernel_ms = double(ms) + j->extra_ms;
    out->march_ms = double(ms_dom) + j->extra_ms;
    return SYF_OK;
}

int syf_de_probe(syf_job* j, const float* xyz4, size_t n, float* out4, float* traps4) {
    clear_error();
    if (!j || (n && (!xyz4 || !out4 || !traps4))) return fail(SYF_E_INVALID, "job and buffers must not be null");
    if (n == 0) return SYF_OK;
    if (n > (1u << 30)) return fail(SYF_E_INVALID, "n must be below 2^30, got %zu", n);
    if (hipError_t de = hipSetDevice(j->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    float4 *dp = nullptr, *dout = nullptr, *dtr = nullptr;
    hipError_t e = hipMalloc(&dp, n * 16);
    if (e == hipSuccess) e = hipMalloc(&dout, n * 16);
    if (e == hipSuccess) e = hipMalloc(&dtr, n * 16);
    if (e == hipSuccess) e = hipMemcpyAsync(dp, xyz4, n * 16, hipMemcpyHostToDevice, j->stream);
    if (e == hipSuccess) e = launch_de_probe(j->P.f, dp, uint32_t(n), dout, dtr, j->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out4, dout, n * 16, hipMemcpyDeviceToHost, j->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(traps4, dtr, n * 16, hipMemcpyDeviceToHost, j->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(j->stream);
    if (dp) hipFree(dp);
    if (dout) hipFree(dout);
    if (dtr) hipFree(dtr);
    return e == hipSuccess ? SYF_OK : hip_fail(e, "syf_de_probe");
}

int syf_trace_primary(syf_job* j, uint32_t sample) {
    clear_error();
    if (!j) return fail(SYF_E_INVALID, "job must not be null");
    if (hipError_t de = hipSetDevice(j->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    const size_t cap = j->hitq_cap;
    TraceBuffers B{(float*)j->plane[SYF_PLANE_DEPTH],
                   (float4*)j->plane[SYF_PLANE_AUX],
                   (double*)j->plane[SYF_PLANE_DEPTH64],
                   (uint32_t*)j->plane[SYF_PLANE_STEPS],
                   j->counters,
                   HitQueue{j->hitq, j->hitq + cap, j->hitq + 2 * cap, j->hitq + 3 * cap, j->hitq + 4 * cap,
                            j->hitq + 5 * cap, j->counters + COUNTER_HITS, j->counters + COUNTER_CURSOR,
                            j->hitq + HITQ_FIELDS * cap, j->counters + COUNTER_QERR, uint32_t(cap), 0u},
                   j->ctr};
    // ray counters, queue count, cursor, error flag and work counters start at zero
    hipError_t e = hipMemsetAsync(j->counters, 0, COUNTER_BYTES, j->stream);
    // Every queue field must be empty.  The trailing pass of the previous trace emptied the
    // entries it used; after a trace whose trailing pass was not enqueued, or when the done-flag
    // epoch wraps, the whole queue is cleared.  A done flag counts only if it holds this trace's
    // epoch, so the flags are never cleared otherwise.
    if (++j->hitq_epoch == 0u) {
        j->hitq_epoch = 1u;
        j->hitq_empty = false;
    }
    if (!j->hitq_empty) {
        if (e == hipSuccess) e = hipMemsetAsync(j->hitq, 0xff, HITQ_FIELDS * cap * sizeof(uint32_t), j->stream);
        if (e == hipSuccess) e = hipMemsetAsync(j->hitq + HITQ_FIELDS * cap, 0, cap / 64 * sizeof(uint32_t), j->stream);
    }
    B.q.epoch = j->hitq_epoch;
    j->hitq_empty = false;
    j->traced = true;
    if (e == hipSuccess) e = hipEventRecord(j->ev0, j->stream);
    if (e == hipSuccess) e = launch_trace_primary(j->P, sample, B, j->ctx->cu_count, j->stream);
    if (e == hipSuccess) e = hipEventRecord(j->ev1, j->stream);
    if (e == hipSuccess) e = launch_hits_tail(j->P, sample, B, j->ctx->cu_count, j->stream);
    if (e == hipSuccess) j->hitq_empty = true;  // the trailing pass empties the entries the trace used
    if (e == hipSuccess) e = hipEventRecord(j->ev2, j->stream);
    j->timed = true;
    j->extra_ms = 0.0;
    return e == hipSuccess ? SYF_OK : hip_fail(e, "syf_trace_primary");
}

int syf_readback(syf_job* j, uint32_t plane, void* host, size_t bytes) {
    clear_error();
    if (!j || !host) return fail(SYF_E_INVALID, "job and host must not be null");
    if (plane > SYF_PLANE_SURFACE) return fail(SYF_E_INVALID, "plane: unknown plane %u", plane);
    size_t need = plane_bytes(plane, j->pixels);
    if (bytes != need) return fail(SYF_E_INVALID, "bytes must be %zu for plane %u, got %zu", need, plane, bytes);
    if (hipError_t de = hipSetDevice(j->ctx->device); de != hipSuccess) return hip_fail(de, "hipSetDevice");
    const void* src = plane < SYF_PLANE_VOLUME ? j->plane[plane]
                      : j->vplanes ? static_cast<const char*>(j->vplanes) + (plane - SYF_PLANE_VOLUME) * j->pixels * 16
                                   : j->plane[SYF_PLANE_ACCUM];
    hipError_t e = hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, j->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(j->stream);
    if (e != hipSuccess) return hip_fail(e, "syf_readback");
    if (j->vplanes) return SYF_OK;
    // The session planes of pathtrace.wgsl:1243-1266 for a job without a participating
    // medium: every sample's volume estimate is zero, so VOLUME holds (0, 0, 0, n), its Welford
    // moments max(0, delta * (v - mean)) stay (0, 0, 0, n), and SURFACE adds r - 0 = r where
    // ACCUM adds r, in the same order: the same sums bit for bit.
    if (plane == SYF_PLANE_VOLUME || plane == SYF_PLANE_MOMENTS) {
        float* f = static_cast<float*>(host);
        for (size_t i = 0; i < j->pixels; ++i) f[4 * i] = f[4 * i + 1] = f[4 * i + 2] = 0.0f;
    }
    return SYF_OK;
}

}  // extern "C"
