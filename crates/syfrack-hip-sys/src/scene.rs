//! Safe layer over the raw bindings: `syfrack_core` scenes and graphs lowered to the C structures, and
//! RAII handles for a device context and a render job.
//!
//! [`job_desc_from_scene`] restates `Uniforms::from_scene` (crates/syfrack-render/src/uniforms.rs:748-853)
//! for this path: camera mode and lens, sun, area and environment lights, the base material with its
//! colouring ramp, surface variation and named finishes, the flat ground, cavity darkness, seed and the
//! render parameters; the library derives the camera basis, light frames, step scale and ray extent
//! itself.  [`GraphBuf::from_graph`] lowers the form graph node for node
//! (crates/syfrack-core/src/scene/graph/mod.rs:103-257) to `syf_graph`; the library validates it with
//! the reference's rules and messages.  What this path does not render is refused with an error naming
//! the scene element: a 2D scene, projector, tube and HDRI lights, the seascape, portal, nebula and
//! dunes, textures, optics and phenomena.  The height
//! fog is lowered.

use std::ffi::{CStr, CString};
use std::os::raw::c_int;

use syfrack_core::{
    Camera, CameraMode, Color, Coloring, Dimension, Finish, FinishPattern, Graph, Light, Measure, NodeKind,
    PaletteInterpolation, Scene, TrapMode,
};

use crate::*;

/// A status of the library with the text of `syf_last_error()`, or a refusal of the lowering.
#[derive(Debug, Clone)]
pub struct Error {
    pub status: c_int,
    pub message: String,
}

impl std::fmt::Display for Error {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        write!(f, "syfrack_hip status {}: {}", self.status, self.message)
    }
}

impl std::error::Error for Error {}

fn refuse(message: impl Into<String>) -> Error {
    Error { status: SYF_E_INVALID, message: message.into() }
}

/// `Ok(())` for `SYF_OK`, else the status with the library's message.
pub fn check(status: c_int) -> Result<(), Error> {
    if status == SYF_OK {
        return Ok(());
    }
    let message = unsafe {
        let p = syf_last_error();
        if p.is_null() { String::new() } else { CStr::from_ptr(p).to_string_lossy().into_owned() }
    };
    Err(Error { status, message })
}

/// The render parameters of a job (the fields of `RenderConfig`, renderer.rs:17-43, that reach the
/// device), filled by the caller from its configuration.
#[derive(Clone, Copy, Debug)]
pub struct RenderParams {
    pub width: u32,
    pub height: u32,
    pub max_bounces: u32,
    pub max_steps: u32,
    pub pixel_tolerance: f32,
    /// `SYF_MATH_FAST` or `SYF_MATH_EXACT` (the primary trace's FP32 march arithmetic).
    pub march_math: u32,
}

fn v3(v: glam::Vec3) -> [f32; 3] {
    [v.x, v.y, v.z]
}

fn camera(c: &Camera) -> syf_camera {
    let mut out = syf_camera {
        focal_mm: c.focal_mm,
        f_number: c.f_number,
        focus_distance: c.focus_distance,
        ..Default::default()
    };
    match &c.mode {
        CameraMode::Free { position, target, roll } => {
            out.mode = SYF_CAMERA_FREE;
            out.position = v3(*position);
            out.target = v3(*target);
            out.roll = *roll;
        }
        CameraMode::Orbit { point, distance, azimuth, elevation, roll } => {
            out.mode = SYF_CAMERA_ORBIT;
            out.point = v3(*point);
            out.distance = *distance;
            out.azimuth = *azimuth;
            out.elevation = *elevation;
            out.roll = *roll;
        }
    }
    out
}

fn light(i: usize, l: &Light) -> Result<syf_light, Error> {
    let rgb = |c: &Color| c.to_linear();
    match l {
        Light::Area { position, normal, size, color, intensity } => Ok(syf_light {
            kind: SYF_LIGHT_AREA,
            position: v3(*position),
            normal: v3(*normal),
            size: *size,
            color: rgb(color),
            intensity: *intensity,
            ..Default::default()
        }),
        Light::Environment { zenith, horizon, ground, intensity, .. } => Ok(syf_light {
            kind: SYF_LIGHT_ENV,
            zenith: rgb(zenith),
            horizon: rgb(horizon),
            ground: rgb(ground),
            intensity: *intensity,
            ..Default::default()
        }),
        // GpuLight::from_light (uniforms.rs:656-668): the library normalises the direction and
        // takes the cosine of the angular radius
        Light::Sun { direction, color, intensity, angular_size } => Ok(syf_light {
            kind: SYF_LIGHT_SUN,
            normal: v3(*direction),
            size: [*angular_size, 0.0],
            color: rgb(color),
            intensity: *intensity,
            ..Default::default()
        }),
        other => {
            let kind = match other {
                Light::Projector { .. } => "Projector",
                Light::Tube { .. } => "Tube",
                Light::Hdri { .. } => "Hdri",
                _ => "this kind of light",
            };
            Err(refuse(format!("lights[{i}]: {kind} is not rendered on this path (sun, area and environment lights are)")))
        }
    }
}

/// Coloring with its Ramp as GpuColoring / GpuRamp (uniforms.rs:31-55, 194-214).
fn coloring(c: &Coloring) -> Result<syf_coloring, Error> {
    let mut out = syf_coloring { slot: SYF_COLOR_NONE, frequency: 1.0, gamma: 1.0, ..Default::default() };
    let (slot, ramp, orbit) = match c {
        Coloring::None => return Ok(out),
        Coloring::Trap { slot, ramp } => (*slot, ramp, None),
        Coloring::Orbit { channels, ramp } => (SYF_COLOR_ORBIT, ramp, Some(channels)),
    };
    if ramp.stops.len() > SYF_MAX_STOPS {
        return Err(refuse(format!("material.coloring: at most {SYF_MAX_STOPS} stops, got {}", ramp.stops.len())));
    }
    if !ramp.positions.is_empty() && ramp.positions.len() != ramp.stops.len() {
        return Err(refuse("material.coloring: one position per stop, or none"));
    }
    out.slot = slot;
    out.stop_count = ramp.stops.len() as u32;
    out.cyclic = ramp.cyclic as u32;
    out.smooth = (ramp.interpolation == PaletteInterpolation::Smooth) as u32;
    out.positioned = (!ramp.positions.is_empty()) as u32;
    out.frequency = ramp.frequency;
    out.phase = ramp.phase;
    out.gamma = ramp.gamma;
    for (i, s) in ramp.stops.iter().enumerate() {
        out.stops[i] = [s[0], s[1], s[2], ramp.position(i)];
    }
    if let Some(o) = orbit {
        out.weights = o.traps;
        out.iteration_weight = o.iterations;
        out.normal_weight = o.normal;
        out.logarithmic = o.logarithmic;
    }
    Ok(out)
}

/// Finish with its FinishDetail as GpuFinish (uniforms.rs:787-817): the pattern numbered as
/// detail_kind, its seed, scale, strength, bump, roughness delta and tint.
fn finish(i: usize, f: &Finish) -> Result<syf_finish, Error> {
    if f.optics.is_some() {
        return Err(refuse(format!("material.finishes[{i}].optics: not rendered on this path")));
    }
    let mut out = syf_finish {
        albedo: f.albedo,
        roughness: f.roughness,
        metallic: f.metallic,
        specular: f.specular,
        emission: f.emission,
        ..Default::default()
    };
    if let Some(d) = &f.detail {
        use FinishPattern::*;
        out.pattern = match d.pattern {
            VeinedStone => 1,
            Patina => 2,
            BrushedMetal => 3,
            Basalt => 4,
            Ripples => 5,
            Rosettes => 6,
            CoastalRock => 7,
            Sand => 8,
            Dust => 9,
            Concrete => 10,
            WeatheredConcrete => 11,
            WetConcrete => 12,
        };
        out.seed = d.seed;
        out.scale = d.scale;
        out.strength = d.strength;
        out.bump = d.bump;
        out.roughness_delta = d.roughness_delta;
        out.tint = d.tint;
    }
    Ok(out)
}

/// The job description of a 3D scene (everything but the form graph, which goes to
/// `syf_job_create_graph` beside it).  Refuses what this path does not render, naming it.
pub fn job_desc_from_scene(scene: &Scene, p: &RenderParams) -> Result<syf_job_desc, Error> {
    if scene.dimension == Dimension::TwoD {
        return Err(refuse("dimension: a 2D scene is rendered with syf_planar_create / syf_planar_run"));
    }
    if scene.phenomena.is_some() {
        return Err(refuse("phenomena: not rendered on this path"));
    }
    let a = &scene.atmosphere;
    if a.seascape.is_some() || a.portal.is_some() || a.nebula.is_some() {
        return Err(refuse("atmosphere: seascape, portal and nebula are not rendered on this path"));
    }
    if a.ground.as_ref().is_some_and(|g| g.dunes.is_some()) {
        return Err(refuse("atmosphere.ground.dunes: not rendered on this path (a flat ground is)"));
    }
    let m = &scene.material;
    if m.texture.is_some() {
        return Err(refuse("material.texture: not rendered on this path"));
    }
    if m.finishes.len() > SYF_MAX_FINISHES {
        return Err(refuse(format!("material.finishes: at most {SYF_MAX_FINISHES}, got {}", m.finishes.len())));
    }
    if m.optics.is_some() {
        return Err(refuse("material.optics: not rendered on this path"));
    }
    if scene.lights.len() > SYF_MAX_LIGHTS {
        return Err(refuse(format!("lights: at most {SYF_MAX_LIGHTS} on this path, got {}", scene.lights.len())));
    }
    let g = scene.fractal.graph();
    let mut d = syf_job_desc::default();
    d.fractal.family = SYF_FAMILY_GRAPH;
    d.fractal.safety = g.safety;
    d.fractal.bounding_radius_hint = g.bounding_radius_hint;
    d.camera = camera(&scene.camera);
    d.material = syf_material {
        albedo: m.albedo,
        roughness: m.roughness,
        metallic: m.metallic,
        specular: m.specular,
        ..Default::default()
    };
    if let Some(s) = &m.surface {
        d.material.surface_enabled = 1;
        d.material.surface_slot = s.slot;
        d.material.surface_roughness = s.roughness;
        d.material.surface_metallic = s.metallic;
    }
    for (i, l) in scene.lights.iter().enumerate() {
        d.lights[i] = light(i, l)?;
    }
    d.light_count = scene.lights.len() as u32;
    d.coloring = coloring(&m.coloring)?;
    for (i, f) in m.finishes.iter().enumerate() {
        d.finishes[i] = finish(i, f)?;
    }
    d.finish_count = m.finishes.len() as u32;
    if let Some(g) = &a.ground {
        d.ground = syf_ground { enabled: 1, height: g.height, finish: g.finish };
    }
    // GpuAtmosphere::from_atmosphere (uniforms.rs:218-226)
    if a.fog_density > 0.0 {
        d.fog = syf_fog {
            density: a.fog_density,
            color: a.fog_color,
            height: a.fog_height,
            falloff: a.fog_falloff,
            anisotropy: a.anisotropy,
        };
    }
    d.width = p.width;
    d.height = p.height;
    d.max_bounces = p.max_bounces;
    d.max_steps = p.max_steps;
    d.pixel_tolerance = p.pixel_tolerance;
    d.cavity_darkness = a.cavity_darkness;
    d.march_math = p.march_math;
    d.seed = scene.seed;
    Ok(d)
}

/// A form graph as `syf_graph` with the arrays it points to.
pub struct GraphBuf {
    nodes: Vec<syf_graph_node>,
    links: Vec<syf_graph_link>,
    bodies: Vec<u32>,
    safety: f32,
    hint: f32,
}

fn measure(m: Measure) -> u32 {
    match m {
        Measure::Norm => SYF_MEASURE_NORM,
        Measure::AxisX => SYF_MEASURE_AXIS_X,
        Measure::AxisY => SYF_MEASURE_AXIS_Y,
        Measure::AxisZ => SYF_MEASURE_AXIS_Z,
        Measure::PlaneY => SYF_MEASURE_PLANE_Y,
        Measure::AngleXY => SYF_MEASURE_ANGLE_XY,
        Measure::AngleXZ => SYF_MEASURE_ANGLE_XZ,
        Measure::AngleYZ => SYF_MEASURE_ANGLE_YZ,
        Measure::LogNorm => SYF_MEASURE_LOG_NORM,
    }
}

fn mode(m: TrapMode) -> u32 {
    match m {
        TrapMode::Minimum => SYF_TRAP_MINIMUM,
        TrapMode::Maximum => SYF_TRAP_MAXIMUM,
        TrapMode::Last => SYF_TRAP_LAST,
    }
}

fn put3(v: &mut [f32; 6], at: usize, x: glam::Vec3) {
    v[at] = x.x;
    v[at + 1] = x.y;
    v[at + 2] = x.z;
}

fn axes(x: bool, y: bool, z: bool) -> u32 {
    (x as u32) | ((y as u32) << 1) | ((z as u32) << 2)
}

impl GraphBuf {
    /// The graph node for node: kinds in `NodeKind` order, parameters where include/syfrack_hip.h lists
    /// them per kind, Iterate bodies as ranges of one body array.
    pub fn from_graph(g: &Graph) -> GraphBuf {
        let mut nodes = Vec::with_capacity(g.nodes.len());
        let mut bodies = Vec::new();
        for n in &g.nodes {
            let mut c = syf_graph_node { id: n.id, enabled: n.enabled as u32, ..Default::default() };
            use NodeKind::*;
            c.kind = match &n.kind {
                Input => SYF_NODE_INPUT,
                Output => SYF_NODE_OUTPUT,
                Translate { offset } => {
                    put3(&mut c.v, 0, *offset);
                    SYF_NODE_TRANSLATE
                }
                Rotate { angles } => {
                    put3(&mut c.v, 0, *angles);
                    SYF_NODE_ROTATE
                }
                Scale { factor } => {
                    c.v[0] = *factor;
                    SYF_NODE_SCALE
                }
                ScaleAxes { factors } => {
                    put3(&mut c.v, 0, *factors);
                    SYF_NODE_SCALE_AXES
                }
                Mirror { normal, offset } => {
                    put3(&mut c.v, 0, *normal);
                    c.v[3] = *offset;
                    SYF_NODE_MIRROR
                }
                BoxFold { limit } => {
                    c.v[0] = *limit;
                    SYF_NODE_BOX_FOLD
                }
                BoxFoldAxes { limit, x, y, z } => {
                    put3(&mut c.v, 0, *limit);
                    c.flags = axes(*x, *y, *z);
                    SYF_NODE_BOX_FOLD_AXES
                }
                SphereInversion { radius, min_radius } => {
                    c.v[0] = *radius;
                    c.v[1] = *min_radius;
                    SYF_NODE_SPHERE_INVERSION
                }
                SphereFold { min_radius, fixed_radius } => {
                    c.v[0] = *min_radius;
                    c.v[1] = *fixed_radius;
                    SYF_NODE_SPHERE_FOLD
                }
                Abs { x, y, z } => {
                    c.flags = axes(*x, *y, *z);
                    SYF_NODE_ABS
                }
                MengerFold { scale, offset } => {
                    c.v[0] = *scale;
                    put3(&mut c.v, 1, *offset);
                    SYF_NODE_MENGER_FOLD
                }
                Kaleido { planes, phase } => {
                    c.u[0] = *planes;
                    c.v[0] = *phase;
                    SYF_NODE_KALEIDO
                }
                Power { power } => {
                    c.v[0] = *power;
                    SYF_NODE_POWER
                }
                AddOrigin => SYF_NODE_ADD_ORIGIN,
                AddSeed { multiplier } => {
                    put3(&mut c.v, 0, *multiplier);
                    SYF_NODE_ADD_SEED
                }
                AddConstant { multiplier, julia } => {
                    put3(&mut c.v, 0, *multiplier);
                    put3(&mut c.v, 3, *julia);
                    SYF_NODE_ADD_CONSTANT
                }
                Twist { rate } => {
                    c.v[0] = *rate;
                    SYF_NODE_TWIST
                }
                Bend { rate } => {
                    c.v[0] = *rate;
                    SYF_NODE_BEND
                }
                Tile { period } => {
                    put3(&mut c.v, 0, *period);
                    SYF_NODE_TILE
                }
                RepeatFinite { period, count } => {
                    put3(&mut c.v, 0, *period);
                    c.u = *count;
                    SYF_NODE_REPEAT_FINITE
                }
                Iterate { iterations, bailout, body } => {
                    c.u[0] = *iterations;
                    c.v[0] = *bailout;
                    c.body_first = bodies.len() as u32;
                    c.body_count = body.len() as u32;
                    bodies.extend_from_slice(body);
                    SYF_NODE_ITERATE
                }
                Trap { slot, measure: m, scale, mode: md } => {
                    c.u = [*slot, measure(*m), mode(*md)];
                    c.v[0] = *scale;
                    SYF_NODE_TRAP
                }
                Sphere { radius } => {
                    c.v[0] = *radius;
                    SYF_NODE_SPHERE
                }
                Torus { radius, tube } => {
                    c.v[0] = *radius;
                    c.v[1] = *tube;
                    SYF_NODE_TORUS
                }
                Cylinder { radius, half_height } => {
                    c.v[0] = *radius;
                    c.v[1] = *half_height;
                    SYF_NODE_CYLINDER
                }
                Box { half } => {
                    put3(&mut c.v, 0, *half);
                    SYF_NODE_BOX
                }
                RoundBox { half, radius } => {
                    put3(&mut c.v, 0, *half);
                    c.v[3] = *radius;
                    SYF_NODE_ROUND_BOX
                }
                Rock { half, erosion, seed } => {
                    put3(&mut c.v, 0, *half);
                    c.v[3] = *erosion;
                    c.u[0] = *seed;
                    SYF_NODE_ROCK
                }
                Plane { normal, offset } => {
                    put3(&mut c.v, 0, *normal);
                    c.v[3] = *offset;
                    SYF_NODE_PLANE
                }
                NormFence { inset } => {
                    c.v[0] = *inset;
                    SYF_NODE_NORM_FENCE
                }
                BulbFence => SYF_NODE_BULB_FENCE,
                MaterialTag { index } => {
                    c.u[0] = *index;
                    SYF_NODE_MATERIAL_TAG
                }
                Bounded { half, margin } => {
                    put3(&mut c.v, 0, *half);
                    c.v[3] = *margin;
                    SYF_NODE_BOUNDED
                }
                // the library refuses an active Instances node by name; a disabled one passes through
                Instances { half, margin, .. } => {
                    put3(&mut c.v, 0, *half);
                    c.v[3] = *margin;
                    SYF_NODE_INSTANCES
                }
                Union => SYF_NODE_UNION,
                Intersection => SYF_NODE_INTERSECTION,
                Subtraction => SYF_NODE_SUBTRACTION,
                SmoothUnion { k } => {
                    c.v[0] = *k;
                    SYF_NODE_SMOOTH_UNION
                }
                SmoothIntersection { k } => {
                    c.v[0] = *k;
                    SYF_NODE_SMOOTH_INTERSECTION
                }
            };
            nodes.push(c);
        }
        let links = g
            .links
            .iter()
            .map(|l| syf_graph_link { from_node: l.from.0, from_port: l.from.1, to_node: l.to.0, to_port: l.to.1 })
            .collect();
        GraphBuf { nodes, links, bodies, safety: g.safety, hint: g.bounding_radius_hint }
    }

    /// The `syf_graph` pointing into this buffer (valid while the buffer lives).
    pub fn raw(&self) -> syf_graph {
        syf_graph {
            nodes: self.nodes.as_ptr(),
            node_count: self.nodes.len() as u32,
            link_count: self.links.len() as u32,
            links: self.links.as_ptr(),
            bodies: self.bodies.as_ptr(),
            body_len: self.bodies.len() as u32,
            safety: self.safety,
            bounding_radius_hint: self.hint,
            reserved: 0,
        }
    }

    /// Validates the graph; returns (step scale, iteration product, trap scales, kernel family).
    pub fn info(&self) -> Result<(f32, u32, [f32; 4], u32), Error> {
        let g = self.raw();
        let (mut step, mut product, mut scales, mut family) = (0.0f32, 0u32, [0.0f32; 4], 0u32);
        check(unsafe { syf_graph_info(&g, &mut step, &mut product, scales.as_mut_ptr(), &mut family) })?;
        Ok((step, product, scales, family))
    }
}

/// One HIP device (`syf_ctx`).
pub struct Context {
    raw: *mut syf_ctx,
}

impl Context {
    pub fn new(device: i32) -> Result<Context, Error> {
        let mut raw = std::ptr::null_mut();
        check(unsafe { syf_ctx_create(device, &mut raw) })?;
        Ok(Context { raw })
    }
    pub fn raw(&self) -> *mut syf_ctx {
        self.raw
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        unsafe {
            syf_ctx_destroy(self.raw);
        }
    }
}

unsafe impl Send for Context {}
unsafe impl Sync for Context {}

/// A render job of one scene (`syf_job`): its device planes and stream.
pub struct Job {
    raw: *mut syf_job,
    pub width: u32,
    pub height: u32,
}

impl Job {
    /// The job of a 3D scene: its graph runs on the kernels of its topology
    /// (`SYF_GRAPH_SPECIALISED`; a classic topology on its family's kernels).
    pub fn from_scene(ctx: &Context, scene: &Scene, p: &RenderParams) -> Result<Job, Error> {
        let desc = job_desc_from_scene(scene, p)?;
        let graph = GraphBuf::from_graph(scene.fractal.graph());
        let g = graph.raw();
        let mut raw = std::ptr::null_mut();
        check(unsafe { syf_job_create_graph(ctx.raw, &desc, &g, SYF_GRAPH_SPECIALISED, &mut raw) })?;
        Ok(Job { raw, width: p.width, height: p.height })
    }
    pub fn raw(&self) -> *mut syf_job {
        self.raw
    }
    /// `Renderer::dispatch` for one tile: samples [sample_start, sample_start + count) accumulated.
    pub fn render_tile(&mut self, x: u32, y: u32, w: u32, h: u32, sample_start: u32, count: u32) -> Result<(), Error> {
        check(unsafe { syf_render_tile(self.raw, x, y, w, h, sample_start, count) })
    }
    pub fn render_wavefront(&mut self, x: u32, y: u32, w: u32, h: u32, sample_start: u32, count: u32)
        -> Result<(), Error> {
        check(unsafe { syf_render_wavefront(self.raw, x, y, w, h, sample_start, count) })
    }
    pub fn trace_primary(&mut self, sample: u32) -> Result<(), Error> {
        check(unsafe { syf_trace_primary(self.raw, sample) })
    }
    pub fn sync(&mut self) -> Result<(), Error> {
        check(unsafe { syf_job_sync(self.raw) })
    }
    /// `RenderSession::readback` of one plane (`SYF_PLANE_*`) into `out` (its exact byte size).
    pub fn readback<T: Copy>(&mut self, plane: u32, out: &mut [T]) -> Result<(), Error> {
        check(unsafe {
            syf_readback(self.raw, plane, out.as_mut_ptr() as *mut std::os::raw::c_void, std::mem::size_of_val(out))
        })
    }
    /// `parity::probe_scene`: the field at points (x, y, z, _); out4 = (distance, trap 0,
    /// iterations, material), traps4 = the four normalised traps.
    pub fn probe(&mut self, xyz4: &[f32], out4: &mut [f32], traps4: &mut [f32]) -> Result<(), Error> {
        let n = xyz4.len() / 4;
        if out4.len() < 4 * n || traps4.len() < 4 * n {
            return Err(refuse("probe: out4 and traps4 need 4 floats per point"));
        }
        check(unsafe { syf_de_probe(self.raw, xyz4.as_ptr(), n, out4.as_mut_ptr(), traps4.as_mut_ptr()) })
    }
    /// `SessionState::shader_fingerprint` of this job's transport.
    pub fn fingerprint(&mut self) -> Result<u64, Error> {
        let mut f = 0u64;
        check(unsafe { syf_job_fingerprint(self.raw, &mut f) })?;
        Ok(f)
    }
    /// `RenderSession::save`: the planes and `state.ron` text, replaced atomically.
    pub fn checkpoint_save(&mut self, dir: &std::path::Path, state_ron: &str) -> Result<(), Error> {
        let d = CString::new(dir.to_string_lossy().as_bytes()).map_err(|_| refuse("checkpoint: NUL in the path"))?;
        let s = CString::new(state_ron).map_err(|_| refuse("checkpoint: NUL in state.ron"))?;
        check(unsafe { syf_checkpoint_save(self.raw, d.as_ptr(), s.as_ptr()) })
    }
    /// `RenderSession::resume`: the planes of a checkpoint of this job's transport.
    pub fn checkpoint_load(&mut self, dir: &std::path::Path) -> Result<(), Error> {
        let d = CString::new(dir.to_string_lossy().as_bytes()).map_err(|_| refuse("checkpoint: NUL in the path"))?;
        check(unsafe { syf_checkpoint_load(self.raw, d.as_ptr()) })
    }
}

impl Drop for Job {
    fn drop(&mut self) {
        unsafe {
            syf_job_destroy(self.raw);
        }
    }
}

unsafe impl Send for Job {}
