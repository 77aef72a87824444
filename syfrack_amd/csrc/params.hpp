// Per-job constants shared by the host side of the library and its kernels.  A job's
// parameters travel as one kernel argument (scalar constant memory on gfx950), lowered on
// the host from syf_job_desc the way Uniforms::from_scene (uniforms.rs:748-853) lowers a
// scene.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/syfrack_hip.h"

namespace syf {

struct GOp;  // one instruction of a compiled node graph (graph_device.hpp)

// Classic-graph parameters.  Products the reference recomputes per iteration from node
// parameters alone (fr2/mr2, offset*(scale-1), -0.5*offset.z*(scale-1)) are folded once
// on the host with the same f32 operations, so the values are identical.
struct FractalP {
    uint32_t family;
    uint32_t iterations;
    uint32_t iteration_product;
    uint32_t ipower;          // integer power 1..16, or 0 for the trigonometric form
    float bailout;
    float power;
    float scale;
    float abs_scale;
    float mr2;                // min_radius^2
    float fr2;                // fixed_radius^2
    float fr2_over_mr2;
    float fold_limit;
    float off_s1[3];          // MengerFold offset * (scale - 1)
    float menger_shift_threshold;  // -0.5 * offset.z * (scale - 1)
    float half[3];
    float radius;
    float plane_n[3];
    float plane_offset;
    float trap_scale;         // normalisation of the trap the shading reads (a graph: of its shade slot)
    float safety;             // step scale
    // SYF_FAMILY_GRAPH: the compiled program (device memory), the four trap normalisation
    // scales and the trap slot the shading reads
    const GOp* gprog;
    float gtrap_scale[4];
    uint32_t gshade_slot;
    // the run-time compiled kernels of the graph's topology (host memory, never read by a
    // kernel), or null: the graph runs on the interpreter
    const void* gjit;
};

struct CameraP {
    float origin[3];
    float forward[3];
    float right[3];
    float up[3];
    float tan_half;
    float aspect;
    float lens_radius;
    float focus;
};

// Area light as GpuLight (uniforms.rs:676-716): a = centre, b = u * width, c = v * height.
struct LightP {
    uint32_t kind;
    float a[3];
    float b[3];
    float c[3];
    float n[3];               // area_normal = safe_normalize(cross(b, c))
    float color[3];           // colour * intensity (env: zenith)
    float horizon[3];
    float ground[3];
    float area;               // |cross(b, c)|
    float cos_max;            // sun: cosine of the disk's angular radius (GpuLight a.w)
};

// A named finish as make_surface reads it (GpuFinish, uniforms.rs:787-817).
struct FinishP {
    float albedo[3];
    float roughness, metallic, specular;
    float emission[3];
    float sigma_scale;        // Oren-Nayar: sigma = sigma_scale * roughness (0.45, or 0: Lambert)
    uint32_t pattern, seed;   // detail pattern 1..12 (0: none) and its seed (detail_kind)
    float scale, strength, bump, roughness_delta;  // detail_mapping
    float tint[3];            // detail_tint
};

// Colouring ramp, ground plane and finishes (GpuColoring, GpuAtmosphere.ground, finishes).
struct ShadeP {
    uint32_t color_slot;      // 0..3 trap slot, 4 iterations, 5 normal, 6 orbit; above: none
    uint32_t stop_count, cyclic, smooth, positioned;
    float frequency, phase, gamma;
    float stops[SYF_MAX_STOPS][4];
    float weights[4];
    float iteration_weight, normal_weight, logarithmic;
    uint32_t ground_enabled;
    float ground_height;
    uint32_t ground_finish;
    uint32_t finish_count;
    // the shading reads more of a surface hit's field sample than trap slot 0: the other traps,
    // the iteration count or the material (then the sample is evaluated in full at the hit)
    uint32_t full_sample;
    // a finish has a detail pattern (the path carries its ray footprint to the hits)
    uint32_t detail;
    // height fog (GpuAtmosphere.a, b): density 0 = none
    float fog_density, fog_height, fog_falloff, fog_g;
    float fog_color[3];
    FinishP finishes[SYF_MAX_FINISHES];
};

struct MaterialP {
    float albedo[3];
    float roughness;
    float metallic;
    float specular;
    uint32_t surface_enabled;
    uint32_t surface_slot;
    float surface[4];         // roughness t=0, t=1, metallic t=0, t=1
};

struct JobP {
    FractalP f;
    CameraP cam;
    MaterialP mat;
    LightP lights[SYF_MAX_LIGHTS];
    uint32_t light_count;
    uint32_t width;
    uint32_t height;
    uint32_t seed;
    uint32_t max_bounces;
    uint32_t max_steps;
    float pixel_tolerance;
    float pixel_cone;         // tan_half * 2 / width (pathtrace.wgsl:809)
    float t_end;              // ray_extent (common.wgsl:206-208)
    float cavity_darkness;
    uint32_t march_math;      // SYF_MATH_FAST or SYF_MATH_EXACT (the primary trace's FP32 march)
    ShadeP sh;
};

// Kernel specialisation of a family: the Mandelbulb with a non-integer power takes the
// trigonometric form and the power-8 Mandelbulb a compile-time exponent, each instantiated
// separately so the common paths stay lean.
constexpr uint32_t FAM_BULB_TRIG = 4u;
constexpr uint32_t FAM_BULB8 = 5u;
constexpr uint32_t FAM_COUNT = 6u;  // kernel families of the classic graphs: the four, trig and power-8 Mandelbulb
constexpr uint32_t FAM_GRAPH = 6u;  // any node graph, run by the graph interpreter (not batched by the server)
constexpr uint32_t FAM_JIT = 7u;    // a node graph whose topology has run-time compiled kernels
constexpr bool is_bulb(uint32_t fam) {
    return fam == SYF_FAMILY_MANDELBULB || fam == FAM_BULB_TRIG || fam == FAM_BULB8;
}
__host__ __device__ inline uint32_t kernel_family(const FractalP& F) {
    if (F.family == SYF_FAMILY_GRAPH) return F.gjit ? FAM_JIT : FAM_GRAPH;
    if (F.family != SYF_FAMILY_MANDELBULB) return F.family;
    return F.ipower == 0u ? FAM_BULB_TRIG : (F.ipower == 8u ? FAM_BULB8 : SYF_FAMILY_MANDELBULB);
}

enum : uint32_t {
    CTR_DE_EVALS = 0,
    CTR_DE_ITERS = 1,
    CTR_STEPS = 2,
    CTR_DE64_EVALS = 3,
    CTR_DE64_ITERS = 4,
    CTR_HITS = 5,
    CTR_RAYS = 6,
    CTR_PRIMARY = 7,
    CTR_HIT_EVALS = 8,  // FP32 evaluations of the hit kernel (normal taps of the primary trace)
    CTR_HIT_ITERS = 9,
    // of the above: the work of the primary trace's trailing hit pass
    CTR_TAIL_HITS = 10,
    CTR_TAIL_EVALS = 11,
    CTR_TAIL_ITERS = 12,
    CTR_TAIL_DE64_EVALS = 13,
    CTR_TAIL_DE64_ITERS = 14,
    CTR_COUNT = 15,
};

}  // namespace syf
