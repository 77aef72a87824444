/*
 * syfrack_hip.h -- C ABI of the MI355X (gfx950) sphere-traced distance-estimator path.
 *
 * The library owns every device allocation behind opaque handles.  Callers own the host
 * buffers they pass; no pointer is retained after a call returns.  Every entry point returns
 * an int status (SYF_OK or a negative SYF_E_* code) and leaves a human-readable message,
 * naming the offending field where there is one, in a thread-local string read with
 * syf_last_error().
 *
 * Each declaration names the reference interface it replaces (paths relative to the
 * SyFrack source tree).  The Rust host (syfrack-render) binds these through an
 * `extern "C"` block; nothing here uses C++ or torch types.
 */
#ifndef SYFRACK_HIP_H
#define SYFRACK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SYF_ABI_VERSION 6

/* ---------------------------------------------------------------- status codes
 * Replaces anyhow::Result at the seams: config errors (renderer.rs:81-138),
 * device/scope errors (renderer.rs:651-666), TDR timeout (renderer.rs:749-768),
 * wavefront status bits (wavefront.rs:414-417). */
#define SYF_OK 0
#define SYF_E_INVALID (-1)   /* invalid argument or configuration; message names the field */
#define SYF_E_DEVICE (-2)    /* HIP runtime error, device limit or out of device memory */
#define SYF_E_QUEUE (-3)     /* queue invariant broken (capacity overflow, path left alive) */
#define SYF_E_TIMEOUT (-4)   /* watchdog expired or device lost */
#define SYF_E_ADMISSION (-5) /* request refused: its footprint exceeds the HBM budget */
#define SYF_E_NOT_FOUND (-6) /* unknown request id */
#define SYF_E_NO_DEVICE (-7) /* no HIP device with that index */

/* ---------------------------------------------------------------- scene description
 * The flat job description is what Uniforms::from_scene (uniforms.rs:748-853) and
 * ParamLayout::pack (graph_wgsl.rs:72-183) produce for the classic graphs
 * graphs/{mandelbulb,mandelbox,menger,sphere}.ron: one fixed topology per family, its
 * node parameters as plain fields. */
#define SYF_FAMILY_MANDELBULB 0u /* Iterate{Trap, Power, AddOrigin} -> BulbFence */
#define SYF_FAMILY_MANDELBOX 1u  /* Iterate{BoxFold, Trap, SphereFold, Scale, AddOrigin} -> NormFence */
#define SYF_FAMILY_MENGER 2u     /* Iterate{MengerFold, Trap} -> Box */
#define SYF_FAMILY_SPHERE 3u     /* Trap -> Sphere  U  Plane */
#define SYF_FAMILY_GRAPH 4u      /* any node graph, given to syf_job_create_graph */

typedef struct syf_fractal {
    uint32_t family;          /* SYF_FAMILY_* */
    uint32_t iterations;      /* Iterate.iterations (1..64, graph/mod.rs:44-73) */
    float bailout;            /* Iterate.bailout */
    float power;              /* Power.power (Mandelbulb) */
    float scale;              /* Scale.factor (Mandelbox) or MengerFold.scale (Menger) */
    float min_radius;         /* SphereFold.min_radius */
    float fixed_radius;       /* SphereFold.fixed_radius */
    float fold_limit;         /* BoxFold.limit */
    float offset[3];          /* MengerFold.offset */
    float half[3];            /* Box.half (Menger fence) */
    float radius;             /* Sphere.radius (sphere graph) */
    float plane_normal[3];    /* Plane.normal (sphere graph ground) */
    float plane_offset;       /* Plane.offset */
    float trap_scale;         /* Trap.scale of slot 0 */
    float safety;             /* Graph.safety; step scale = safety / Lipschitz (eval.rs:879) */
    float bounding_radius_hint;
} syf_fractal;

#define SYF_CAMERA_FREE 0u
#define SYF_CAMERA_ORBIT 1u
/* scene/camera.rs:5-35; the basis is derived as Camera::basis (camera.rs:58-104). */
typedef struct syf_camera {
    uint32_t mode;            /* SYF_CAMERA_FREE or SYF_CAMERA_ORBIT */
    float position[3];        /* Free */
    float target[3];          /* Free */
    float roll;               /* degrees, both modes */
    float point[3];           /* Orbit */
    float distance;           /* Orbit */
    float azimuth;            /* Orbit, degrees */
    float elevation;          /* Orbit, degrees */
    float focal_mm;           /* 36 mm sensor: tan(half horizontal fov) = 18 / focal_mm */
    float f_number;           /* 0 = pinhole */
    float focus_distance;
} syf_camera;

#define SYF_LIGHT_SUN 0u  /* GpuLight KIND_SUN (uniforms.rs:656-668): a distant disk */
#define SYF_LIGHT_AREA 1u /* GpuLight KIND_AREA (uniforms.rs:676-716) */
#define SYF_LIGHT_ENV 2u  /* GpuLight KIND_ENV (uniforms.rs:732-744) */
#define SYF_MAX_LIGHTS 4
typedef struct syf_light {
    uint32_t kind;            /* SYF_LIGHT_SUN, SYF_LIGHT_AREA or SYF_LIGHT_ENV */
    float position[3];        /* area: centre */
    float normal[3];          /* area: emitting side; sun: direction towards the sun */
    float size[2];            /* area: full width and height; sun: size[0] = angular diameter,
                                 degrees, 0.01..90 */
    float color[3];           /* area, sun: linear colour */
    float intensity;          /* all: multiplies the colour(s); a sun's colour x intensity is
                                 its irradiance */
    float zenith[3];          /* env */
    float horizon[3];         /* env */
    float ground[3];          /* env */
} syf_light;

/* scene/material.rs:138-165 and SurfaceVariation (material.rs:70-91). */
typedef struct syf_material {
    float albedo[3];
    float roughness;
    float metallic;
    float specular;
    uint32_t surface_enabled;      /* SurfaceVariation present */
    uint32_t surface_slot;         /* what drives the variation: trap slot 0..3, SYF_COLOR_ITER or
                                      SYF_COLOR_NORMAL (surface_value, pathtrace.wgsl:265-269) */
    float surface_roughness[2];    /* roughness at t = 0 and t = 1 */
    float surface_metallic[2];     /* metallic at t = 0 and t = 1 */
} syf_material;

/* The shading of a path-traced job beyond the base material (pathtrace.wgsl:271-420):
 * named finishes that MaterialTag nodes (and the ground) select, the colouring ramp, and a
 * flat ground plane.  All zero: no finishes, no colouring, no ground. */
#define SYF_MAX_FINISHES 8   /* GpuFinish slots (uniforms.rs:787) */
#define SYF_MAX_STOPS 16     /* MAX_PALETTE_STOPS (palette.rs:6) */
/* FinishPattern (material.rs:177-191) in the order of GpuFinish.detail_kind (uniforms.rs:797-812). */
#define SYF_PATTERN_NONE 0u
#define SYF_PATTERN_VEINED_STONE 1u
#define SYF_PATTERN_PATINA 2u
#define SYF_PATTERN_BRUSHED_METAL 3u
#define SYF_PATTERN_BASALT 4u
#define SYF_PATTERN_RIPPLES 5u
#define SYF_PATTERN_ROSETTES 6u
#define SYF_PATTERN_COASTAL_ROCK 7u
#define SYF_PATTERN_SAND 8u
#define SYF_PATTERN_DUST 9u
#define SYF_PATTERN_CONCRETE 10u
#define SYF_PATTERN_WEATHERED_CONCRETE 11u
#define SYF_PATTERN_WET_CONCRETE 12u
/* Finish (material.rs:161-174) with its FinishDetail (material.rs:193-221) as GpuFinish
 * (uniforms.rs:787-817).  A detail pattern shapes the finish at the hit (finish_detail.wgsl,
 * make_surface pathtrace.wgsl:306-329): its height blends albedo toward the tint by strength
 * and moves the roughness by roughness_delta, its slope tilts the shading normal by bump, each
 * band faded by the path's ray footprint.  The Concrete, WeatheredConcrete and WetConcrete
 * patterns also give the diffuse lobe Oren-Nayar's sigma = 0.45 x roughness
 * (pathtrace.wgsl:381-384, 407-418); oren_nayar = 1 asks for that lobe without a pattern.
 * The coastal rock's tide line and the sand's wind direction belong to the seascape and the
 * dunes, absent here: the tide is dry and the wind blows along +x. */
typedef struct syf_finish {
    float albedo[3];
    float roughness;
    float metallic;
    float specular;
    float emission[3];        /* linear emitted radiance */
    uint32_t oren_nayar;      /* 1: rough diffuse with sigma = 0.45 x roughness */
    uint32_t pattern;         /* SYF_PATTERN_*; 0: no detail */
    uint32_t seed;            /* 0..65535 */
    float scale;              /* 0.01..200: pattern cells per scene unit */
    float strength;           /* 0..1 */
    float bump;               /* -1..1 */
    float roughness_delta;    /* -1..1 */
    float tint[3];            /* 0..1 */
} syf_finish;
#define SYF_COLOR_ITER 4u     /* COLOR_SLOT_ITER (common.wgsl:13): iterations / iteration product */
#define SYF_COLOR_NORMAL 5u   /* COLOR_SLOT_NORMAL: n.y / 2 + 1/2 */
#define SYF_COLOR_ORBIT 6u    /* COLOR_SLOT_ORBIT: the weighted orbit mix */
#define SYF_COLOR_NONE 0xffffu
/* Coloring (material.rs:12-37, 113-119) with its Ramp (palette.rs:74-115) as GpuColoring
 * (uniforms.rs:194-214) and GpuRamp (uniforms.rs:31-55). */
typedef struct syf_coloring {
    uint32_t slot;            /* trap slot 0..3, SYF_COLOR_ITER, _NORMAL, _ORBIT or _NONE */
    uint32_t stop_count;      /* 0..SYF_MAX_STOPS (0: the ramp is white) */
    uint32_t cyclic;
    uint32_t smooth;          /* smooth-step interpolation between stops */
    uint32_t positioned;      /* the stop positions were given (else evenly spaced) */
    float frequency;
    float phase;
    float gamma;
    float stops[SYF_MAX_STOPS][4]; /* linear rgb and position (Ramp::position) */
    float weights[4];         /* orbit: trap weights */
    float iteration_weight;   /* orbit */
    float normal_weight;      /* orbit */
    float logarithmic;        /* orbit: signed log gain, 0 = linear */
} syf_coloring;
/* The atmosphere's height fog (Atmosphere, atmosphere.rs:7-16; GpuAtmosphere.a/b, uniforms.rs:
 * 218-226): extinction density at and below `height`, decaying above it as exp(-(y - height) x
 * falloff); single-scattering albedo `color`; Henyey-Greenstein anisotropy.  density 0: no fog.
 * A job with fog accumulates the session's VOLUME, MOMENTS and SURFACE planes on the device
 * (pathtrace.wgsl:1243-1266).  The batched server refuses a request with fog; the control
 * machine's preview ignores it. */
typedef struct syf_fog {
    float density;            /* 0..10 */
    float color[3];           /* linear, finite, >= 0 */
    float height;             /* -100..100 */
    float falloff;            /* 0..10 (0: homogeneous) */
    float anisotropy;         /* -0.99..0.99 */
} syf_fog;
/* Atmosphere.ground (GpuAtmosphere.ground, uniforms.rs:227-230): a flat plane (no dunes). */
typedef struct syf_ground {
    uint32_t enabled;
    float height;
    uint32_t finish;          /* its finish: 1..finish_count, 0 = the base material */
} syf_ground;

/* ---------------------------------------------------------------- node graphs
 * A distance field as a node graph (scene/graph/mod.rs:103-257, 427-451): nodes, links between
 * their ports, the Iterate bodies, the graph's safety and bounding-radius hint.  Node kinds are
 * numbered in the order of NodeKind; a node's parameters sit in `v` (floats) and `u`
 * (integers) as listed per kind.  The library validates the graph (a kind this path cannot
 * evaluate, a cycle or a dangling body id is refused with SYF_E_INVALID naming the node),
 * orders and compiles it into a register program for the device, the counterpart of
 * graph_wgsl::compile (graph_wgsl.rs:683-949), and derives its step scale (eval.rs:828-881),
 * trap scales and iteration product. */
#define SYF_NODE_INPUT 0u
#define SYF_NODE_OUTPUT 1u
#define SYF_NODE_TRANSLATE 2u          /* v[0..3] offset */
#define SYF_NODE_ROTATE 3u             /* v[0..3] angles (degrees, XYZ Euler) */
#define SYF_NODE_SCALE 4u              /* v[0] factor */
#define SYF_NODE_SCALE_AXES 5u         /* v[0..3] factors */
#define SYF_NODE_MIRROR 6u             /* v[0..3] normal, v[3] offset */
#define SYF_NODE_BOX_FOLD 7u           /* v[0] limit */
#define SYF_NODE_BOX_FOLD_AXES 8u      /* v[0..3] limit; flags bits 0-2: x, y, z */
#define SYF_NODE_SPHERE_INVERSION 9u   /* v[0] radius, v[1] min_radius */
#define SYF_NODE_SPHERE_FOLD 10u       /* v[0] min_radius, v[1] fixed_radius */
#define SYF_NODE_ABS 11u               /* flags bits 0-2: x, y, z */
#define SYF_NODE_MENGER_FOLD 12u       /* v[0] scale, v[1..4] offset */
#define SYF_NODE_KALEIDO 13u           /* u[0] planes, v[0] phase (degrees) */
#define SYF_NODE_POWER 14u             /* v[0] power */
#define SYF_NODE_ADD_ORIGIN 15u
#define SYF_NODE_ADD_SEED 16u          /* v[0..3] multiplier */
#define SYF_NODE_ADD_CONSTANT 17u      /* v[0..3] multiplier, v[3..6] julia */
#define SYF_NODE_TWIST 18u             /* v[0] rate */
#define SYF_NODE_BEND 19u              /* v[0] rate */
#define SYF_NODE_TILE 20u              /* v[0..3] period */
#define SYF_NODE_REPEAT_FINITE 21u     /* v[0..3] period, u[0..3] count */
#define SYF_NODE_ITERATE 22u           /* u[0] iterations, v[0] bailout, body_first/body_count */
#define SYF_NODE_TRAP 23u              /* u[0] slot, u[1] measure (SYF_MEASURE_*), u[2] mode, v[0] scale */
#define SYF_NODE_SPHERE 24u            /* v[0] radius */
#define SYF_NODE_TORUS 25u             /* v[0] radius, v[1] tube */
#define SYF_NODE_CYLINDER 26u          /* v[0] radius, v[1] half_height */
#define SYF_NODE_BOX 27u               /* v[0..3] half */
#define SYF_NODE_ROUND_BOX 28u         /* v[0..3] half, v[3] radius */
#define SYF_NODE_ROCK 29u              /* v[0..3] half, v[3] erosion, u[0] seed */
#define SYF_NODE_PLANE 30u             /* v[0..3] normal, v[3] offset */
#define SYF_NODE_NORM_FENCE 31u        /* v[0] inset */
#define SYF_NODE_BULB_FENCE 32u
#define SYF_NODE_MATERIAL_TAG 33u      /* u[0] index */
#define SYF_NODE_BOUNDED 34u           /* v[0..3] half, v[3] margin */
#define SYF_NODE_INSTANCES 35u         /* not evaluated on this path: refused */
#define SYF_NODE_UNION 36u
#define SYF_NODE_INTERSECTION 37u
#define SYF_NODE_SUBTRACTION 38u
#define SYF_NODE_SMOOTH_UNION 39u      /* v[0] k */
#define SYF_NODE_SMOOTH_INTERSECTION 40u /* v[0] k */
#define SYF_NODE_KINDS 41u
/* Trap measures and modes (Measure, TrapMode: scene/graph/mod.rs:81-100). */
#define SYF_MEASURE_NORM 0u
#define SYF_MEASURE_AXIS_X 1u
#define SYF_MEASURE_AXIS_Y 2u
#define SYF_MEASURE_AXIS_Z 3u
#define SYF_MEASURE_PLANE_Y 4u
#define SYF_MEASURE_ANGLE_XY 5u
#define SYF_MEASURE_ANGLE_XZ 6u
#define SYF_MEASURE_ANGLE_YZ 7u
#define SYF_MEASURE_LOG_NORM 8u
#define SYF_TRAP_MINIMUM 0u
#define SYF_TRAP_MAXIMUM 1u
#define SYF_TRAP_LAST 2u

typedef struct syf_graph_node {
    uint32_t id;              /* NodeId, unique in the graph */
    uint32_t kind;            /* SYF_NODE_* */
    uint32_t enabled;         /* Node.enabled: 0 = bypassed (scene/graph/mod.rs:432) */
    uint32_t flags;           /* Abs, BoxFoldAxes: bits 0, 1, 2 = x, y, z */
    float v[6];               /* float parameters (per kind, above) */
    uint32_t u[3];            /* integer parameters (per kind, above) */
    uint32_t body_first;      /* Iterate: body ids are bodies[body_first .. body_first + body_count) */
    uint32_t body_count;
} syf_graph_node;

typedef struct syf_graph_link {  /* Link (scene/graph/mod.rs:436-440): output port -> input port */
    uint32_t from_node, from_port, to_node, to_port;
} syf_graph_link;

typedef struct syf_graph {
    const syf_graph_node* nodes;
    uint32_t node_count;
    uint32_t link_count;
    const syf_graph_link* links;
    const uint32_t* bodies;   /* the Iterate bodies, concatenated */
    uint32_t body_len;
    float safety;             /* Graph.safety */
    float bounding_radius_hint;
    uint32_t reserved;
} syf_graph;

/* One render job: the scene plus RenderConfig (renderer.rs:17-43) fields. */
typedef struct syf_job_desc {
    syf_fractal fractal;
    syf_camera camera;
    syf_material material;
    syf_light lights[SYF_MAX_LIGHTS];
    uint32_t light_count;
    uint32_t width;
    uint32_t height;
    uint32_t max_bounces;          /* 0..32 */
    uint32_t max_steps;            /* >= 1 */
    float pixel_tolerance;         /* 0.125..1.0 */
    float cavity_darkness;         /* Atmosphere.cavity_darkness, 0..1 */
    uint32_t march_math;           /* SYF_MATH_FAST (0) or SYF_MATH_EXACT: arithmetic of the primary
                                      trace's FP32 march (no counterpart in the reference) */
    uint64_t seed;                 /* Scene.seed; folded lo ^ hi as uniforms.rs:824 */
    syf_ground ground;             /* Atmosphere.ground */
    syf_coloring coloring;         /* Material.coloring */
    uint32_t finish_count;         /* Material.finishes, 0..SYF_MAX_FINISHES */
    syf_finish finishes[SYF_MAX_FINISHES];
    syf_fog fog;                   /* Atmosphere fog fields */
} syf_job_desc;
/* SYF_MATH_FAST: hardware reciprocal / reciprocal square root / logarithm and FMA contraction in
 * the FP32 march; held to the FP64 oracle's tolerance (>= 99.9 % hit/miss, refined hit distance
 * within 1e-4 relative).  SYF_MATH_EXACT: correctly rounded IEEE operations as written, so the
 * march takes the steps of the CPU restatement of the reference arithmetic. */
#define SYF_MATH_FAST 0u
#define SYF_MATH_EXACT 1u

/* ---------------------------------------------------------------- planes
 * Device planes of a job, read back with syf_readback.  Order follows the reference
 * session planes (session.rs:1217-1243) where they exist. */
#define SYF_PLANE_ACCUM 0u   /* float4 per pixel: sum of radiance, sample count */
#define SYF_PLANE_DEPTH 1u   /* float per pixel: first-hit march distance, -1 on a miss */
#define SYF_PLANE_AUX 2u     /* float4 per pixel: first-hit normal, steps; (0,0,0,-1) on a miss */
#define SYF_PLANE_DEPTH64 3u /* double per pixel: FP64-refined hit distance, -1 on a miss */
#define SYF_PLANE_STEPS 4u   /* uint32 per pixel: march steps of the primary ray, hit or miss */
/* The session's radiance layers (session.rs:1217-1243; accumulated per sample by
 * pathtrace.wgsl:1243-1266).  A job with fog accumulates them on the device: the volume radiance
 * (the camera segment's in-scattering, or the whole path when the camera ray scattered), its
 * Welford second moments, and the surface radiance (the rest).  Without fog the volume estimate
 * of every sample is zero: VOLUME and MOMENTS read (0, 0, 0, samples), SURFACE equals ACCUM, and
 * they are read back only (they follow from ACCUM). */
#define SYF_PLANE_VOLUME 5u  /* float4 per pixel: volume radiance sum, samples */
#define SYF_PLANE_MOMENTS 6u /* float4 per pixel: the volume's Welford second moments, samples */
#define SYF_PLANE_SURFACE 7u /* float4 per pixel: surface radiance sum, samples */

/* Work counters of the last launch on a job, and its device time. */
typedef struct syf_stats {
    uint64_t rays;             /* traced segments: primary + bounce + shadow */
    uint64_t primary_rays;
    uint64_t hits;             /* primary rays that hit the form */
    uint64_t de_evals;         /* FP32 distance evaluations (march steps + normal taps) */
    uint64_t de_iterations;    /* FP32 Iterate-body executions over all evaluations */
    uint64_t de64_evals;       /* FP64 evaluations (final-hit refine) */
    uint64_t de64_iterations;
    uint64_t steps;            /* march steps over all segments */
    double kernel_ms;          /* device time of the last call's launches, HIP events */
    double march_ms;           /* device time of its dominant kernel alone (primary trace: the fused
                                  march + hit-shading kernel, without the trailing hit pass) */
    uint64_t hit_de_evals;     /* of de_evals: the primary trace's normal taps */
    uint64_t hit_de_iterations;
    uint64_t tail_hits;        /* primary trace: hits shaded by the trailing pass, not the fused kernel */
    uint64_t tail_hit_de_evals;     /* of hit_de_evals: those of the trailing pass */
    uint64_t tail_hit_de_iterations;
    uint64_t tail_de64_evals;       /* of de64_evals: those of the trailing pass */
    uint64_t tail_de64_iterations;
} syf_stats;

/* ---------------------------------------------------------------- opaque handles */
typedef struct syf_ctx syf_ctx;
typedef struct syf_job syf_job;
typedef struct syf_planar syf_planar;
typedef struct syf_server syf_server;

/* Thread-local text of the last error on this thread ("" when none). */
const char* syf_last_error(void);
int syf_abi_version(void);

/* Device context: one HIP device and its stream pool.  Replaces GpuContext::new
 * (gpu.rs:15-34) for this path. */
int syf_ctx_create(int device, syf_ctx** out);
int syf_ctx_destroy(syf_ctx* ctx);

/* Render job: validates the description like RenderConfig::validate (renderer.rs:81-138),
 * derives the camera basis (camera.rs:58-104), the step scale (eval.rs:879-881), the ray
 * extent (uniforms.rs:781-782) and allocates its device planes.  Replaces
 * RenderSession::build (session.rs:187-219). */
int syf_job_create(syf_ctx* ctx, const syf_job_desc* desc, syf_job** out);
int syf_job_destroy(syf_job* job);
/* A render job whose distance field is a node graph: desc->fractal.family must be
 * SYF_FAMILY_GRAPH (its other fractal fields are unused; the graph carries safety and bounding
 * hint).  Every job entry point then evaluates the graph (probe, primary trace, megakernel and
 * wavefront path tracing); the batched server takes classic families only.  Replaces the
 * per-topology pipeline of the reference (graph_wgsl::compile, pipeline_cache.rs) for this
 * path.  `mode`: SYF_GRAPH_SPECIALISED runs a graph of a classic family's topology (see
 * syf_graph_info) on that family's kernels, as syf_job_create with the family's parameters, and
 * compiles the kernels of any other topology at run time (hiprtc, a few seconds the first time a
 * topology is seen in the process; graphs of one topology share them whatever their
 * parameters); SYF_GRAPH_INTERPRETED runs the library's graph interpreter (no compilation,
 * slower).  They compute the same values (with SYF_MATH_EXACT; the fast march arithmetic
 * applies to the classic families). */
#define SYF_GRAPH_SPECIALISED 0u
#define SYF_GRAPH_INTERPRETED 1u
int syf_job_create_graph(syf_ctx* ctx, const syf_job_desc* desc, const syf_graph* graph, uint32_t mode, syf_job** out);
/* Derived properties of a graph without a job: the step scale (eval.rs:879-881), the iteration
 * product (validate.rs:112-137), the trap normalisation scales (eval.rs:853-877), and the kernel
 * family a specialised job of it runs on: a classic family (SYF_FAMILY_MANDELBULB ...
 * SYF_FAMILY_SPHERE) for the bundled classic graphs' topologies with those graphs' trap, or
 * SYF_FAMILY_GRAPH.  Validates like syf_job_create_graph.  Any output may be null. */
int syf_graph_info(const syf_graph* graph, float* step_scale, uint32_t* iteration_product, float* trap_scales4,
                   uint32_t* kernel_family);
/* The device source of the graph's specialised topology (what SYF_GRAPH_SPECIALISED compiles),
 * the counterpart of the WGSL text of graph_wgsl::compile (graph_wgsl.rs:188-192, 949): *len is
 * set to the size with its terminating NUL; up to cap bytes are written to buf (may be null). */
int syf_graph_source(const syf_graph* graph, char* buf, size_t cap, size_t* len);
/* Zero the accumulation and first-hit planes. */
int syf_job_clear(syf_job* job);
/* Wait for every launch queued on the job's stream. */
int syf_job_sync(syf_job* job);
/* Counters and device time of the last launch; waits for it first. */
int syf_job_stats(syf_job* job, syf_stats* out);

/* Distance probe: evaluates the field at n points (x, y, z, ignored) and writes
 * (distance, trap0, iterations, material) and the four traps per point, the layout of
 * de_probe.wgsl:13-17.  Replaces parity::probe_scene (parity.rs:219-332) and
 * Renderer::probe_distances (renderer.rs:645-648).  Host buffers; synchronous. */
int syf_de_probe(syf_job* job, const float* xyz4, size_t n, float* out4, float* traps4);

/* Primary-ray sphere trace of sample `sample` over the whole image: FP32 march with the
 * pixel-footprint hit rule (march_segment, pathtrace.wgsl:78-145), tetrahedral normal
 * (common.wgsl:376-387), FP64 refine of each hit.  Fills DEPTH, AUX, DEPTH64 and STEPS.
 * Asynchronous on the job's stream.  This is the first-hit half of Renderer::dispatch
 * (renderer.rs:729-771) with max_bounces = 0. */
int syf_trace_primary(syf_job* job, uint32_t sample);

/* Megakernel path trace of a tile: for each pixel in [x, x+w) x [y, y+h), samples
 * sample_start .. sample_start+sample_count-1 are traced and added to ACCUM in ascending
 * sample order; when sample_start == 0, DEPTH/AUX take sample 0's first hit and DEPTH64 its
 * first surface hit refined in FP64 (as syf_trace_primary; -1 otherwise).  The one-launch form
 * of Renderer::dispatch (renderer.rs:729-771) running pathtrace.wgsl main (:1203-1276).
 * Asynchronous. */
int syf_render_tile(syf_job* job, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                    uint32_t sample_start, uint32_t sample_count);

/* Wavefront path trace of a tile, same result contract as syf_render_tile: per sample,
 * generate -> (march -> shade/extend -> compact) x rounds -> accumulate, with ray state in
 * SoA device queues.  Replaces Wavefront::dispatch (wavefront.rs:310-419).  Synchronous:
 * returns SYF_E_QUEUE when a queue invariant is broken. */
int syf_render_wavefront(syf_job* job, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                         uint32_t sample_start, uint32_t sample_count);

/* Allocation-free primary trace.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  The same trace as syf_trace_primary (the same kernels and planes, bit
 * for bit) of a job described by `desc`, written into caller-owned device planes of
 * width x height entries: depth (float), aux4 (float4, 16-byte aligned), depth64 (double) and
 * steps (uint32).  Enqueued on `hip_stream` (a hipStream_t of ctx's device; NULL = its null
 * stream) and asynchronous; it calls no hipMalloc, hipFree, memset or synchronisation.  Its
 * hit queue and counters live in a fixed device array of the library (31.5 MB per device,
 * held until the process ends), so width x height is limited to SYF_NOALLOC_MAX_PIXELS
 * (SYF_E_INVALID above it, or on a device of more than 256 CUs).  All calls on one device share
 * that array: calls that overlap on one device (two streams, two threads) race. */
#define SYF_NOALLOC_MAX_PIXELS 1048576u
int syf_trace_primary_noalloc(syf_ctx* ctx, const syf_job_desc* desc, uint32_t sample, float* depth, float* aux4,
                              double* depth64, uint32_t* steps, void* hip_stream);

/* Copy a device plane (SYF_PLANE_*) to a host buffer of exactly `bytes` bytes.
 * Replaces the readbacks of RenderSession::readback (session.rs:1217-1243). */
int syf_readback(syf_job* job, uint32_t plane, void* host, size_t bytes);

/* Copy a host buffer of exactly `bytes` bytes into a device plane (ACCUM, DEPTH, AUX, DEPTH64 or
 * STEPS; VOLUME, MOMENTS and SURFACE of a job with fog): the planes of a resumed session, as
 * RenderSession::resume writes its saved buffers back (session.rs:1308-1330).  Synchronous. */
int syf_upload(syf_job* job, uint32_t plane, const void* host, size_t bytes);

/* ---------------------------------------------------------------- checkpoints and export
 * The transport identity of a job: FNV-1a-64 over this library's transport version and the
 * lowered job, the counterpart of the shader fingerprint a session stores in state.ron
 * (shader_source.rs:117-132, session.rs:648-652).  A checkpoint resumes only into a job with the
 * same identity. */
int syf_job_fingerprint(syf_job* job, uint64_t* out);
/* FNV-1a-64 of n bytes continuing from `hash` (start: 0xcbf29ce484222325), the checksum of the
 * export payload (export.rs:1253-1257). */
uint64_t syf_fnv1a64(const void* data, size_t n, uint64_t hash);
/* RenderSession::save (session.rs:1199-1306): writes the planes of the job to `dir` as raw
 * little-endian files (accum.bin, depth.bin, normals.bin, volume.bin, volume-moments.bin,
 * surface.bin, a 48-byte volume-components.bin, and this path's depth64.bin and steps.bin) and
 * the caller's `state_text` as state.ron, into a temporary directory renamed over `dir`; the old
 * checkpoint is moved aside first and removed last, so one complete checkpoint always exists.
 * A directory holding any other file is refused (SYF_E_INVALID naming it). */
int syf_checkpoint_save(syf_job* job, const char* dir, const char* state_text);
/* RenderSession::resume's plane half: reads the planes saved in `dir` (each must have the job's
 * size; depth64.bin and steps.bin when present) into the job.  The caller has rebuilt the job
 * from state.ron and compared its fingerprint. */
int syf_checkpoint_load(syf_job* job, const char* dir);
/* The export payload (export.rs:1226-1262): mean colour (sum / max(n, 1), alpha 1), depth and
 * normals of the job, little-endian, to `path`; returns its byte count and FNV-1a-64 hash. */
int syf_export_write(syf_job* job, const char* path, uint64_t* bytes, uint64_t* hash);
/* export.rs:1263-1324: reads a payload of a width x height image, checking its byte count and
 * hash ("invalid byte count", "checksum mismatch"); any output pointer may be NULL. */
int syf_export_read(const char* path, uint64_t bytes, uint64_t hash, uint32_t width, uint32_t height, float* color4,
                    float* depth, float* normals4);

/* ---------------------------------------------------------------- control machine
 * Interactive geometry without per-edit compilation: ControlSession (control/mod.rs:145-515,
 * shaders/control.wgsl).  A session holds an image of the scene's node graph, evaluated by the
 * library's graph interpreter (never a compiled topology), and advances it in bounded quanta:
 * each pixel is a resumable machine (pixel-centre ray from the camera; the over-relaxed march
 * of march_segment with pixel tolerance 1 and no step limit: a ray continues until it hits or
 * provably leaves the scene; then the four tetrahedral normal taps and a headlight shade,
 * control.wgsl:54-136,267-303).  The image is ACCUM (colour, 1; background (0.025, 0.03, 0.04)
 * until a pixel completes, (0.025, 0.035, 0.05) for a miss), DEPTH (hit distance, -1) and AUX
 * (normal, march steps; (0, 0, 0, -1)).  Scenes without phenomena (the reference's nature
 * field, media and curved light are outside this path). */
typedef struct syf_control syf_control;
#define SYF_CONTROL_QUANTUM 64u /* default field evaluations per wave per dispatch */
/* ControlSession::new (control/mod.rs:172-335): desc gives the camera and the image size
 * (its fractal, lights, material and sampling fields are unused), graph the field; `quantum`
 * bounds one dispatch to that many field evaluations per wave (0 selects SYF_CONTROL_QUANTUM,
 * at most 4096).  The whole requested size is kept (up to 2^26 pixels): the machine's state is
 * 36 bytes per pixel, where the reference shrinks the image to fit a 64 MiB register file. */
int syf_control_create(syf_ctx* ctx, const syf_job_desc* desc, const syf_graph* graph, uint32_t quantum,
                       syf_control** out);
int syf_control_destroy(syf_control* c);
/* ControlSession::set_scene (control/mod.rs:354-401): validates and compiles the new graph
 * (host only), and only then replaces the session's program, parameters and camera and restarts
 * the image; a pending quantum of the old image is discarded.  An invalid graph leaves the
 * session as it was.  Waits for the quantum in flight (a few milliseconds at most). */
int syf_control_set_scene(syf_control* c, const syf_job_desc* desc, const syf_graph* graph);
/* ControlSession::tick (control/mod.rs:404-465): polls the quantum in flight and, once it has
 * completed, submits the next one; never waits.  Returns the session's failure, with the node,
 * after a non-finite distance (or -inf: +inf is empty space) until set_scene replaces the
 * scene. */
int syf_control_tick(syf_control* c);
/* ControlSession::wait_complete (control/mod.rs:502-515): ticks until every pixel is complete
 * (SYF_E_TIMEOUT after timeout_s seconds). */
int syf_control_wait(syf_control* c, double timeout_s);
/* complete / completed_pixels / dispatches / size (control/mod.rs:467-484); any output may be
 * null. */
int syf_control_progress(syf_control* c, uint32_t* completed_pixels, uint32_t* dispatches, uint32_t* complete,
                         uint32_t* width, uint32_t* height);
/* The image planes SYF_PLANE_ACCUM, SYF_PLANE_DEPTH or SYF_PLANE_AUX (accumulation_buffer,
 * control/mod.rs:473); waits for the quantum in flight. */
int syf_control_readback(syf_control* c, uint32_t plane, void* host, size_t bytes);
/* ControlSession::pick (control/mod.rs:488-500): the pixel at (u, v) in [0, 1); *hit = 0 where
 * the depth is not positive, else the depth and AUX. */
int syf_control_pick(syf_control* c, float u, float v, uint32_t* hit, float* depth, float* aux4);
/* Renderer::control_distances (control/mod.rs:530-574): the field of the graph evaluated by the
 * interpreter at n points, in syf_de_probe's layout.  Synchronous. */
int syf_control_probe(syf_ctx* ctx, const syf_graph* graph, const float* xyz4, size_t n, float* out4,
                      float* traps4);

/* ---------------------------------------------------------------- 2D escape time
 * PlanarFractal::sample (planar.rs:77-194) / planar_sample (planar.wgsl:16-91) for the
 * parameter-plane families, plus FP64 and deep-zoom perturbation modes that have no
 * reference counterpart. */
#define SYF_PLANAR_MANDELBROT 2u  /* family numbering of planar.wgsl:29 */
#define SYF_PLANAR_BURNING_SHIP 3u
#define SYF_PLANAR_JULIA 1u
#define SYF_PLANAR_FP32 0u        /* direct FP32 iteration (the reference's precision) */
#define SYF_PLANAR_FP64 1u        /* direct FP64 iteration: counts bit-exact */
#define SYF_PLANAR_PERTURB 2u     /* FP64 reference orbits + FP32 deltas, re-referencing glitches;
                                   * the glitches left after max_references take PERTURB64 */
#define SYF_PLANAR_PERTURB64 3u   /* the view centre's FP64 orbit + FP64 deltas with rebasing
                                   * (Zhuoran 2021): every pixel's c keeps its offset's full
                                   * FP64 precision, whatever the zoom */
typedef struct syf_planar_desc {
    uint32_t family;          /* SYF_PLANAR_* family */
    uint32_t mode;            /* SYF_PLANAR_FP32 / FP64 / PERTURB */
    uint32_t width;
    uint32_t height;
    uint32_t iterations;      /* maximum iterations (budget) */
    uint32_t max_references;  /* PERTURB: FP32-delta reference orbits before the FP64 rebasing
                               * fallback, 1..64 (0 selects 8); ignored by the other modes */
    double center[2];         /* centre of the view in the complex plane */
    double span;              /* width of the view in the complex plane */
    double julia[2];          /* Julia constant */
    float bailout;            /* escape radius */
    float rotation;           /* view rotation, degrees */
} syf_planar_desc;
#define SYF_PLANAR_SMOOTH 0u  /* float per pixel: smooth iteration count */
#define SYF_PLANAR_ITER 1u    /* uint32 per pixel: iteration count */

int syf_planar_create(syf_ctx* ctx, const syf_planar_desc* desc, syf_planar** out);
int syf_planar_destroy(syf_planar* p);
/* Evaluate every pixel (asynchronous; the reference orbit is built on the host first in
 * PERTURB mode). */
int syf_planar_run(syf_planar* p);
int syf_planar_sync(syf_planar* p);
int syf_planar_readback(syf_planar* p, uint32_t plane, void* host, size_t bytes);
/* Iterations evaluated, glitch re-references, and device time of the last run. */
int syf_planar_stats(syf_planar* p, uint64_t* iterations, uint64_t* rebases, double* kernel_ms);
/* Glitch handling of the last PERTURB / PERTURB64 run: reference orbits used by the FP32-delta
 * passes, pixels left to the FP64 rebasing loop after them, and rebasing events of that loop.
 * No reference counterpart (the reference has no deep-zoom mode; planar.rs:77-194 is FP32). */
int syf_planar_glitches(syf_planar* p, uint32_t* references, uint32_t* fallback_pixels, uint64_t* rebasings);

/* ---------------------------------------------------------------- serving
 * A render server on one device: requests of any family, size, spp and bounce count are
 * submitted as they arrive and traced progressively, with their accumulation planes resident
 * in HBM.  It replaces the export scheduling of the reference (TileWorkQueue,
 * export_queue.rs:7-120; two ExportSlots, export.rs:453-464; RenderSession::step,
 * session.rs:694-877) for this path.  In SYF_SERVE_BATCHED mode a scheduler thread admits
 * requests against the HBM budget and runs them together: every launch is one persistent
 * kernel over tiles of all active requests, a request joining at the next launch after it is
 * admitted (continuous batching).  SYF_SERVE_PLAIN runs one launch per request, in arrival
 * order: the one-launch-per-request form.  A request's image does not depend on the mode or
 * on the other requests in flight. */
#define SYF_SERVE_BATCHED 0u
#define SYF_SERVE_PLAIN 1u
typedef struct syf_server_config {
    uint64_t hbm_budget_bytes;   /* admission budget for request planes (0 = 90 % of free HBM) */
    uint32_t mode;               /* SYF_SERVE_BATCHED or SYF_SERVE_PLAIN */
    uint32_t max_active;         /* requests traced concurrently (0 = 4096) */
    uint32_t quantum_samples;    /* samples per pixel per work item of a long request (0 = 1 while
                                    requests arrive, autotuned per family, 8 or 16, once they
                                    have stopped) */
    uint32_t batch_target_us;    /* target device time of one launch while requests arrive
                                    (0 = 30000 us) */
    uint32_t grid_blocks;        /* workgroups of the persistent launch (0 = autotuned per family) */
    uint32_t quiet_target_us;    /* target device time of one launch once no request has arrived
                                    for 200 ms (0 = 32 x batch_target_us; SYF_QUIET_UNBOUNDED =
                                    all remaining work in one launch) */
    uint32_t path_steps;         /* field evaluations per lane per loop trip of the batched kernel:
                                    4, 8 or 16 (0 = autotuned per family) */
    uint32_t tile_width;         /* width of a work item's 64-pixel tile: 8 (8 x 8), 16 (16 x 4) or
                                    32 (32 x 2) (0 = autotuned per family) */
} syf_server_config;
#define SYF_QUIET_UNBOUNDED 0xffffffffu

#define SYF_REQ_QUEUED 0u     /* waiting for admission */
#define SYF_REQ_ACTIVE 1u     /* admitted, being traced */
#define SYF_REQ_DONE 2u       /* all samples accumulated; planes readable */
#define SYF_REQ_CANCELLED 3u
typedef struct syf_request_info {
    uint32_t state;           /* SYF_REQ_* */
    uint32_t samples_done;    /* samples per pixel accumulated so far */
    uint32_t spp;
    uint32_t reserved;
    uint64_t footprint_bytes; /* device bytes the request holds while admitted */
    double submit_ms;         /* host clock, ms since the server was created */
    double start_ms;          /* first launch containing its work */
    double done_ms;           /* completion of the launch holding its last sample */
} syf_request_info;

/* Kernel families of the batched kernel, the index of the per-family tuning arrays below:
 * 0 Mandelbulb (integer power), 1 Mandelbox, 2 Menger, 3 sphere graph, 4 Mandelbulb (non-integer
 * power), 5 Mandelbulb (power 8). */
#define SYF_KERNEL_FAMILIES 6
typedef struct syf_server_counters {
    uint64_t submitted, completed, cancelled, launches;
    uint64_t rays;            /* traced segments (primary + bounce + shadow) */
    uint64_t primary_rays;
    uint64_t hbm_in_use, hbm_peak;
    double device_ms;         /* summed device time of the launches, HIP events */
    double wall_ms;           /* since the server was created */
    uint32_t quantum_samples, batch_target_us, grid_blocks, reserved;
    /* the autotuner's choices (or the configuration's) per kernel family, and its time */
    uint32_t family_grid_blocks[SYF_KERNEL_FAMILIES];
    uint32_t family_path_steps[SYF_KERNEL_FAMILIES];
    double autotune_ms;
    uint32_t family_tile_width[SYF_KERNEL_FAMILIES];
    uint32_t family_quiet_quantum[SYF_KERNEL_FAMILIES];  /* samples per work item in quiet mode */
} syf_server_counters;

int syf_server_create(syf_ctx* ctx, const syf_server_config* cfg, syf_server** out);
/* Stops the scheduler after the launches in flight and frees every request. */
int syf_server_destroy(syf_server* s);
/* Queues a request for `spp` samples per pixel.  SYF_E_ADMISSION when its footprint exceeds
 * the whole budget; SYF_E_INVALID for an invalid description (message names the field). */
int syf_server_submit(syf_server* s, const syf_job_desc* desc, uint32_t spp, uint64_t* id);
int syf_server_poll(syf_server* s, uint64_t id, syf_request_info* out);
/* Waits for a request to finish; SYF_E_TIMEOUT after timeout_ms. */
int syf_server_wait(syf_server* s, uint64_t id, uint32_t timeout_ms);
/* Waits until every submitted request has finished or was cancelled. */
int syf_server_wait_all(syf_server* s, uint32_t timeout_ms);
/* Copies a plane (SYF_PLANE_ACCUM, DEPTH, AUX or DEPTH64) of a finished request to the host. */
int syf_server_result(syf_server* s, uint64_t id, uint32_t plane, void* host, size_t bytes);
/* Frees the device planes of a finished or cancelled request; its id is unknown afterwards. */
int syf_server_release(syf_server* s, uint64_t id);
/* Cancels a queued or active request (its work already launched still completes). */
int syf_server_cancel(syf_server* s, uint64_t id);
int syf_server_stats(syf_server* s, syf_server_counters* out);

#ifdef __cplusplus
}
#endif
#endif /* SYFRACK_HIP_H */
