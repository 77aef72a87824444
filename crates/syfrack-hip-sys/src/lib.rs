//! Bindings of `libsyfrack_hip` (include/syfrack_hip.h): the sphere-traced distance-estimator path of
//! SyFrack on MI355X (gfx950).
//!
//! The raw items mirror the C header one for one: constants, `#[repr(C)]` structures with the header's
//! field order and types, and the `extern "C"` functions.  With the default `scene` feature, the
//! [`scene`] module lowers `syfrack_core` scenes and graphs to those structures and wraps the handles
//! (the seams of `Renderer::dispatch`, `Wavefront::dispatch`, `parity::probe_scene` and
//! `RenderSession::{readback, save, resume}`).
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

pub const SYF_ABI_VERSION: c_int = 6;

pub const SYF_OK: c_int = 0;
pub const SYF_E_INVALID: c_int = -1;
pub const SYF_E_DEVICE: c_int = -2;
pub const SYF_E_QUEUE: c_int = -3;
pub const SYF_E_TIMEOUT: c_int = -4;
pub const SYF_E_ADMISSION: c_int = -5;
pub const SYF_E_NOT_FOUND: c_int = -6;
pub const SYF_E_NO_DEVICE: c_int = -7;

pub const SYF_FAMILY_MANDELBULB: u32 = 0;
pub const SYF_FAMILY_MANDELBOX: u32 = 1;
pub const SYF_FAMILY_MENGER: u32 = 2;
pub const SYF_FAMILY_SPHERE: u32 = 3;
pub const SYF_FAMILY_GRAPH: u32 = 4;

pub const SYF_CAMERA_FREE: u32 = 0;
pub const SYF_CAMERA_ORBIT: u32 = 1;

pub const SYF_LIGHT_SUN: u32 = 0;
pub const SYF_LIGHT_AREA: u32 = 1;
pub const SYF_LIGHT_ENV: u32 = 2;
pub const SYF_MAX_LIGHTS: usize = 4;

pub const SYF_MAX_FINISHES: usize = 8;
/// FinishPattern in the order of GpuFinish.detail_kind (uniforms.rs:797-812); 0: no detail.
pub const SYF_PATTERN_NONE: u32 = 0;
pub const SYF_PATTERN_VEINED_STONE: u32 = 1;
pub const SYF_PATTERN_PATINA: u32 = 2;
pub const SYF_PATTERN_BRUSHED_METAL: u32 = 3;
pub const SYF_PATTERN_BASALT: u32 = 4;
pub const SYF_PATTERN_RIPPLES: u32 = 5;
pub const SYF_PATTERN_ROSETTES: u32 = 6;
pub const SYF_PATTERN_COASTAL_ROCK: u32 = 7;
pub const SYF_PATTERN_SAND: u32 = 8;
pub const SYF_PATTERN_DUST: u32 = 9;
pub const SYF_PATTERN_CONCRETE: u32 = 10;
pub const SYF_PATTERN_WEATHERED_CONCRETE: u32 = 11;
pub const SYF_PATTERN_WET_CONCRETE: u32 = 12;
pub const SYF_MAX_STOPS: usize = 16;
pub const SYF_COLOR_ITER: u32 = 4;
pub const SYF_COLOR_NORMAL: u32 = 5;
pub const SYF_COLOR_ORBIT: u32 = 6;
pub const SYF_COLOR_NONE: u32 = 0xffff;

pub const SYF_NODE_INPUT: u32 = 0;
pub const SYF_NODE_OUTPUT: u32 = 1;
pub const SYF_NODE_TRANSLATE: u32 = 2;
pub const SYF_NODE_ROTATE: u32 = 3;
pub const SYF_NODE_SCALE: u32 = 4;
pub const SYF_NODE_SCALE_AXES: u32 = 5;
pub const SYF_NODE_MIRROR: u32 = 6;
pub const SYF_NODE_BOX_FOLD: u32 = 7;
pub const SYF_NODE_BOX_FOLD_AXES: u32 = 8;
pub const SYF_NODE_SPHERE_INVERSION: u32 = 9;
pub const SYF_NODE_SPHERE_FOLD: u32 = 10;
pub const SYF_NODE_ABS: u32 = 11;
pub const SYF_NODE_MENGER_FOLD: u32 = 12;
pub const SYF_NODE_KALEIDO: u32 = 13;
pub const SYF_NODE_POWER: u32 = 14;
pub const SYF_NODE_ADD_ORIGIN: u32 = 15;
pub const SYF_NODE_ADD_SEED: u32 = 16;
pub const SYF_NODE_ADD_CONSTANT: u32 = 17;
pub const SYF_NODE_TWIST: u32 = 18;
pub const SYF_NODE_BEND: u32 = 19;
pub const SYF_NODE_TILE: u32 = 20;
pub const SYF_NODE_REPEAT_FINITE: u32 = 21;
pub const SYF_NODE_ITERATE: u32 = 22;
pub const SYF_NODE_TRAP: u32 = 23;
pub const SYF_NODE_SPHERE: u32 = 24;
pub const SYF_NODE_TORUS: u32 = 25;
pub const SYF_NODE_CYLINDER: u32 = 26;
pub const SYF_NODE_BOX: u32 = 27;
pub const SYF_NODE_ROUND_BOX: u32 = 28;
pub const SYF_NODE_ROCK: u32 = 29;
pub const SYF_NODE_PLANE: u32 = 30;
pub const SYF_NODE_NORM_FENCE: u32 = 31;
pub const SYF_NODE_BULB_FENCE: u32 = 32;
pub const SYF_NODE_MATERIAL_TAG: u32 = 33;
pub const SYF_NODE_BOUNDED: u32 = 34;
pub const SYF_NODE_INSTANCES: u32 = 35;
pub const SYF_NODE_UNION: u32 = 36;
pub const SYF_NODE_INTERSECTION: u32 = 37;
pub const SYF_NODE_SUBTRACTION: u32 = 38;
pub const SYF_NODE_SMOOTH_UNION: u32 = 39;
pub const SYF_NODE_SMOOTH_INTERSECTION: u32 = 40;
pub const SYF_NODE_KINDS: u32 = 41;

pub const SYF_MEASURE_NORM: u32 = 0;
pub const SYF_MEASURE_AXIS_X: u32 = 1;
pub const SYF_MEASURE_AXIS_Y: u32 = 2;
pub const SYF_MEASURE_AXIS_Z: u32 = 3;
pub const SYF_MEASURE_PLANE_Y: u32 = 4;
pub const SYF_MEASURE_ANGLE_XY: u32 = 5;
pub const SYF_MEASURE_ANGLE_XZ: u32 = 6;
pub const SYF_MEASURE_ANGLE_YZ: u32 = 7;
pub const SYF_MEASURE_LOG_NORM: u32 = 8;
pub const SYF_TRAP_MINIMUM: u32 = 0;
pub const SYF_TRAP_MAXIMUM: u32 = 1;
pub const SYF_TRAP_LAST: u32 = 2;

pub const SYF_GRAPH_SPECIALISED: u32 = 0;
pub const SYF_GRAPH_INTERPRETED: u32 = 1;

pub const SYF_MATH_FAST: u32 = 0;
pub const SYF_MATH_EXACT: u32 = 1;

pub const SYF_PLANE_ACCUM: u32 = 0;
pub const SYF_PLANE_DEPTH: u32 = 1;
pub const SYF_PLANE_AUX: u32 = 2;
pub const SYF_PLANE_DEPTH64: u32 = 3;
pub const SYF_PLANE_STEPS: u32 = 4;
pub const SYF_PLANE_VOLUME: u32 = 5;
pub const SYF_PLANE_MOMENTS: u32 = 6;
pub const SYF_PLANE_SURFACE: u32 = 7;

pub const SYF_NOALLOC_MAX_PIXELS: u32 = 1_048_576;
pub const SYF_CONTROL_QUANTUM: u32 = 64;

pub const SYF_PLANAR_JULIA: u32 = 1;
pub const SYF_PLANAR_MANDELBROT: u32 = 2;
pub const SYF_PLANAR_BURNING_SHIP: u32 = 3;
pub const SYF_PLANAR_FP32: u32 = 0;
pub const SYF_PLANAR_FP64: u32 = 1;
pub const SYF_PLANAR_PERTURB: u32 = 2;
pub const SYF_PLANAR_PERTURB64: u32 = 3;
pub const SYF_PLANAR_SMOOTH: u32 = 0;
pub const SYF_PLANAR_ITER: u32 = 1;

pub const SYF_SERVE_BATCHED: u32 = 0;
pub const SYF_SERVE_PLAIN: u32 = 1;
pub const SYF_QUIET_UNBOUNDED: u32 = 0xffff_ffff;
pub const SYF_REQ_QUEUED: u32 = 0;
pub const SYF_REQ_ACTIVE: u32 = 1;
pub const SYF_REQ_DONE: u32 = 2;
pub const SYF_REQ_CANCELLED: u32 = 3;
pub const SYF_KERNEL_FAMILIES: usize = 6;

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_fractal {
    pub family: u32,
    pub iterations: u32,
    pub bailout: f32,
    pub power: f32,
    pub scale: f32,
    pub min_radius: f32,
    pub fixed_radius: f32,
    pub fold_limit: f32,
    pub offset: [f32; 3],
    pub half: [f32; 3],
    pub radius: f32,
    pub plane_normal: [f32; 3],
    pub plane_offset: f32,
    pub trap_scale: f32,
    pub safety: f32,
    pub bounding_radius_hint: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_camera {
    pub mode: u32,
    pub position: [f32; 3],
    pub target: [f32; 3],
    pub roll: f32,
    pub point: [f32; 3],
    pub distance: f32,
    pub azimuth: f32,
    pub elevation: f32,
    pub focal_mm: f32,
    pub f_number: f32,
    pub focus_distance: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_light {
    pub kind: u32,
    pub position: [f32; 3],
    pub normal: [f32; 3],
    pub size: [f32; 2],
    pub color: [f32; 3],
    pub intensity: f32,
    pub zenith: [f32; 3],
    pub horizon: [f32; 3],
    pub ground: [f32; 3],
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_material {
    pub albedo: [f32; 3],
    pub roughness: f32,
    pub metallic: f32,
    pub specular: f32,
    pub surface_enabled: u32,
    pub surface_slot: u32,
    pub surface_roughness: [f32; 2],
    pub surface_metallic: [f32; 2],
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_finish {
    pub albedo: [f32; 3],
    pub roughness: f32,
    pub metallic: f32,
    pub specular: f32,
    pub emission: [f32; 3],
    pub oren_nayar: u32,
    pub pattern: u32,
    pub seed: u32,
    pub scale: f32,
    pub strength: f32,
    pub bump: f32,
    pub roughness_delta: f32,
    pub tint: [f32; 3],
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_coloring {
    pub slot: u32,
    pub stop_count: u32,
    pub cyclic: u32,
    pub smooth: u32,
    pub positioned: u32,
    pub frequency: f32,
    pub phase: f32,
    pub gamma: f32,
    pub stops: [[f32; 4]; SYF_MAX_STOPS],
    pub weights: [f32; 4],
    pub iteration_weight: f32,
    pub normal_weight: f32,
    pub logarithmic: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_fog {
    pub density: f32,
    pub color: [f32; 3],
    pub height: f32,
    pub falloff: f32,
    pub anisotropy: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_ground {
    pub enabled: u32,
    pub height: f32,
    pub finish: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_graph_node {
    pub id: u32,
    pub kind: u32,
    pub enabled: u32,
    pub flags: u32,
    pub v: [f32; 6],
    pub u: [u32; 3],
    pub body_first: u32,
    pub body_count: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_graph_link {
    pub from_node: u32,
    pub from_port: u32,
    pub to_node: u32,
    pub to_port: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct syf_graph {
    pub nodes: *const syf_graph_node,
    pub node_count: u32,
    pub link_count: u32,
    pub links: *const syf_graph_link,
    pub bodies: *const u32,
    pub body_len: u32,
    pub safety: f32,
    pub bounding_radius_hint: f32,
    pub reserved: u32,
}

impl Default for syf_graph {
    fn default() -> Self {
        Self {
            nodes: std::ptr::null(),
            node_count: 0,
            link_count: 0,
            links: std::ptr::null(),
            bodies: std::ptr::null(),
            body_len: 0,
            safety: 1.0,
            bounding_radius_hint: 2.0,
            reserved: 0,
        }
    }
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_job_desc {
    pub fractal: syf_fractal,
    pub camera: syf_camera,
    pub material: syf_material,
    pub lights: [syf_light; SYF_MAX_LIGHTS],
    pub light_count: u32,
    pub width: u32,
    pub height: u32,
    pub max_bounces: u32,
    pub max_steps: u32,
    pub pixel_tolerance: f32,
    pub cavity_darkness: f32,
    pub march_math: u32,
    pub seed: u64,
    pub ground: syf_ground,
    pub coloring: syf_coloring,
    pub finish_count: u32,
    pub finishes: [syf_finish; SYF_MAX_FINISHES],
    pub fog: syf_fog,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_stats {
    pub rays: u64,
    pub primary_rays: u64,
    pub hits: u64,
    pub de_evals: u64,
    pub de_iterations: u64,
    pub de64_evals: u64,
    pub de64_iterations: u64,
    pub steps: u64,
    pub kernel_ms: f64,
    pub march_ms: f64,
    pub hit_de_evals: u64,
    pub hit_de_iterations: u64,
    pub tail_hits: u64,
    pub tail_hit_de_evals: u64,
    pub tail_hit_de_iterations: u64,
    pub tail_de64_evals: u64,
    pub tail_de64_iterations: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_planar_desc {
    pub family: u32,
    pub mode: u32,
    pub width: u32,
    pub height: u32,
    pub iterations: u32,
    pub max_references: u32,
    pub center: [f64; 2],
    pub span: f64,
    pub julia: [f64; 2],
    pub bailout: f32,
    pub rotation: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_server_config {
    pub hbm_budget_bytes: u64,
    pub mode: u32,
    pub max_active: u32,
    pub quantum_samples: u32,
    pub batch_target_us: u32,
    pub grid_blocks: u32,
    pub quiet_target_us: u32,
    pub path_steps: u32,
    pub tile_width: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_request_info {
    pub state: u32,
    pub samples_done: u32,
    pub spp: u32,
    pub reserved: u32,
    pub footprint_bytes: u64,
    pub submit_ms: f64,
    pub start_ms: f64,
    pub done_ms: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct syf_server_counters {
    pub submitted: u64,
    pub completed: u64,
    pub cancelled: u64,
    pub launches: u64,
    pub rays: u64,
    pub primary_rays: u64,
    pub hbm_in_use: u64,
    pub hbm_peak: u64,
    pub device_ms: f64,
    pub wall_ms: f64,
    pub quantum_samples: u32,
    pub batch_target_us: u32,
    pub grid_blocks: u32,
    pub reserved: u32,
    pub family_grid_blocks: [u32; SYF_KERNEL_FAMILIES],
    pub family_path_steps: [u32; SYF_KERNEL_FAMILIES],
    pub autotune_ms: f64,
    pub family_tile_width: [u32; SYF_KERNEL_FAMILIES],
    pub family_quiet_quantum: [u32; SYF_KERNEL_FAMILIES],
}

#[repr(C)]
pub struct syf_ctx {
    _p: [u8; 0],
}
#[repr(C)]
pub struct syf_job {
    _p: [u8; 0],
}
#[repr(C)]
pub struct syf_planar {
    _p: [u8; 0],
}
#[repr(C)]
pub struct syf_server {
    _p: [u8; 0],
}
#[repr(C)]
pub struct syf_control {
    _p: [u8; 0],
}

extern "C" {
    pub fn syf_last_error() -> *const c_char;
    pub fn syf_abi_version() -> c_int;
    pub fn syf_ctx_create(device: c_int, out: *mut *mut syf_ctx) -> c_int;
    pub fn syf_ctx_destroy(ctx: *mut syf_ctx) -> c_int;
    pub fn syf_job_create(ctx: *mut syf_ctx, desc: *const syf_job_desc, out: *mut *mut syf_job) -> c_int;
    pub fn syf_job_destroy(job: *mut syf_job) -> c_int;
    pub fn syf_job_create_graph(ctx: *mut syf_ctx, desc: *const syf_job_desc, graph: *const syf_graph, mode: u32,
                                out: *mut *mut syf_job) -> c_int;
    pub fn syf_graph_info(graph: *const syf_graph, step_scale: *mut f32, iteration_product: *mut u32,
                          trap_scales4: *mut f32, kernel_family: *mut u32) -> c_int;
    pub fn syf_graph_source(graph: *const syf_graph, buf: *mut c_char, cap: usize, len: *mut usize) -> c_int;
    pub fn syf_job_clear(job: *mut syf_job) -> c_int;
    pub fn syf_job_sync(job: *mut syf_job) -> c_int;
    pub fn syf_job_stats(job: *mut syf_job, out: *mut syf_stats) -> c_int;
    pub fn syf_de_probe(job: *mut syf_job, xyz4: *const f32, n: usize, out4: *mut f32, traps4: *mut f32) -> c_int;
    pub fn syf_trace_primary(job: *mut syf_job, sample: u32) -> c_int;
    pub fn syf_render_tile(job: *mut syf_job, x: u32, y: u32, w: u32, h: u32, sample_start: u32,
                           sample_count: u32) -> c_int;
    pub fn syf_render_wavefront(job: *mut syf_job, x: u32, y: u32, w: u32, h: u32, sample_start: u32,
                                sample_count: u32) -> c_int;
    pub fn syf_trace_primary_noalloc(ctx: *mut syf_ctx, desc: *const syf_job_desc, sample: u32, depth: *mut f32,
                                     aux4: *mut f32, depth64: *mut f64, steps: *mut u32,
                                     hip_stream: *mut c_void) -> c_int;
    pub fn syf_readback(job: *mut syf_job, plane: u32, host: *mut c_void, bytes: usize) -> c_int;
    pub fn syf_upload(job: *mut syf_job, plane: u32, host: *const c_void, bytes: usize) -> c_int;
    pub fn syf_job_fingerprint(job: *mut syf_job, out: *mut u64) -> c_int;
    pub fn syf_fnv1a64(data: *const c_void, n: usize, hash: u64) -> u64;
    pub fn syf_checkpoint_save(job: *mut syf_job, dir: *const c_char, state_text: *const c_char) -> c_int;
    pub fn syf_checkpoint_load(job: *mut syf_job, dir: *const c_char) -> c_int;
    pub fn syf_export_write(job: *mut syf_job, path: *const c_char, bytes: *mut u64, hash: *mut u64) -> c_int;
    pub fn syf_export_read(path: *const c_char, bytes: u64, hash: u64, width: u32, height: u32, color4: *mut f32,
                           depth: *mut f32, normals4: *mut f32) -> c_int;
    pub fn syf_control_create(ctx: *mut syf_ctx, desc: *const syf_job_desc, graph: *const syf_graph, quantum: u32,
                              out: *mut *mut syf_control) -> c_int;
    pub fn syf_control_destroy(c: *mut syf_control) -> c_int;
    pub fn syf_control_set_scene(c: *mut syf_control, desc: *const syf_job_desc, graph: *const syf_graph) -> c_int;
    pub fn syf_control_tick(c: *mut syf_control) -> c_int;
    pub fn syf_control_wait(c: *mut syf_control, timeout_s: f64) -> c_int;
    pub fn syf_control_progress(c: *mut syf_control, completed_pixels: *mut u32, dispatches: *mut u32,
                                complete: *mut u32, width: *mut u32, height: *mut u32) -> c_int;
    pub fn syf_control_readback(c: *mut syf_control, plane: u32, host: *mut c_void, bytes: usize) -> c_int;
    pub fn syf_control_pick(c: *mut syf_control, u: f32, v: f32, hit: *mut u32, depth: *mut f32, aux4: *mut f32)
        -> c_int;
    pub fn syf_control_probe(ctx: *mut syf_ctx, graph: *const syf_graph, xyz4: *const f32, n: usize, out4: *mut f32,
                             traps4: *mut f32) -> c_int;
    pub fn syf_planar_create(ctx: *mut syf_ctx, desc: *const syf_planar_desc, out: *mut *mut syf_planar) -> c_int;
    pub fn syf_planar_destroy(p: *mut syf_planar) -> c_int;
    pub fn syf_planar_run(p: *mut syf_planar) -> c_int;
    pub fn syf_planar_sync(p: *mut syf_planar) -> c_int;
    pub fn syf_planar_readback(p: *mut syf_planar, plane: u32, host: *mut c_void, bytes: usize) -> c_int;
    pub fn syf_planar_stats(p: *mut syf_planar, iterations: *mut u64, rebases: *mut u64, kernel_ms: *mut f64)
        -> c_int;
    pub fn syf_planar_glitches(p: *mut syf_planar, references: *mut u32, fallback_pixels: *mut u32,
        rebasings: *mut u64) -> c_int;
    pub fn syf_server_create(ctx: *mut syf_ctx, cfg: *const syf_server_config, out: *mut *mut syf_server) -> c_int;
    pub fn syf_server_destroy(s: *mut syf_server) -> c_int;
    pub fn syf_server_submit(s: *mut syf_server, desc: *const syf_job_desc, spp: u32, id: *mut u64) -> c_int;
    pub fn syf_server_poll(s: *mut syf_server, id: u64, out: *mut syf_request_info) -> c_int;
    pub fn syf_server_wait(s: *mut syf_server, id: u64, timeout_ms: u32) -> c_int;
    pub fn syf_server_wait_all(s: *mut syf_server, timeout_ms: u32) -> c_int;
    pub fn syf_server_result(s: *mut syf_server, id: u64, plane: u32, host: *mut c_void, bytes: usize) -> c_int;
    pub fn syf_server_release(s: *mut syf_server, id: u64) -> c_int;
    pub fn syf_server_cancel(s: *mut syf_server, id: u64) -> c_int;
    pub fn syf_server_stats(s: *mut syf_server, out: *mut syf_server_counters) -> c_int;
}

#[cfg(feature = "scene")]
pub mod scene;
