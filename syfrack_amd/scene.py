"""Scene -> render-job lowering for the classic fractal graphs.

Mirrors the reference's scene model on this path: the bundled graphs
(graphs/{mandelbulb,mandelbox,menger,sphere}.ron, embedded by
crates/syfrack-core/src/scene/graph/bundled.rs:8-49), their parameter paths
(``fractal.nodes[1].iterations`` ... as Scene::set_param, scene/mod.rs:564), the camera modes
(scene/camera.rs:5-35), sun/area/environment lights (scene/light.rs:66-123), material with
its colouring ramp, surface variation and named finishes (scene/material.rs, palette.rs), the
flat ground plane and height fog (scene/atmosphere.rs:7-38), and RenderConfig (renderer.rs:17-76).  ``Scene.desc()``
produces the flat ``syf_job_desc`` the C ABI takes; the camera basis, step scale and ray
extent are derived inside the library (syf_job_create), as Uniforms::from_scene does.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field
from typing import List, Optional

from . import abi

# Graph bounds (scene/graph/mod.rs:44-73) used to validate set_param.
BOUNDS = {
    "power": (2.0, 16.0),
    "iterations": (1, 64),
    "bailout": (2.0, 1e6),
    "scale": (-4.0, 4.0),
    "min_radius": (0.05, 1.0),
    "menger_scale": (2.0, 4.0),
}


@dataclass
class Fractal:
    """One classic graph with its node parameters (the bundled defaults)."""
    family: int
    iterations: int
    bailout: float
    power: float = 8.0
    scale: float = 2.0
    min_radius: float = 0.5
    fixed_radius: float = 1.0
    fold_limit: float = 1.0
    offset: tuple = (1.0, 1.0, 1.0)
    half: tuple = (1.0, 1.0, 1.0)
    radius: float = 1.0
    plane_normal: tuple = (0.0, 1.0, 0.0)
    plane_offset: float = -1.0
    trap_scale: float = 1.0
    safety: float = 1.0
    bounding_radius_hint: float = 2.5

    @staticmethod
    def bundled(name: str) -> "Fractal":
        """Graph::bundled(name) for the classic graphs."""
        if name == "mandelbulb":  # graphs/mandelbulb.ron
            return Fractal(abi.FAMILY_MANDELBULB, 12, 4.0, power=8.0, trap_scale=4.0, safety=0.7,
                           bounding_radius_hint=2.5)
        if name == "mandelbox":  # graphs/mandelbox.ron
            return Fractal(abi.FAMILY_MANDELBOX, 12, 1000.0, scale=2.0, min_radius=0.5, fixed_radius=1.0,
                           fold_limit=1.0, trap_scale=2.0, safety=0.6, bounding_radius_hint=10.392304)
        if name == "menger":  # graphs/menger.ron
            return Fractal(abi.FAMILY_MENGER, 5, 1000000.0, scale=3.0, offset=(1.0, 1.0, 1.0),
                           half=(1.0, 1.0, 1.0), trap_scale=4.0, safety=0.9, bounding_radius_hint=4.464102)
        if name == "sphere":  # graphs/sphere.ron
            return Fractal(abi.FAMILY_SPHERE, 0, 0.0, radius=1.0, plane_normal=(0.0, 1.0, 0.0), plane_offset=-1.0,
                           trap_scale=1.0, safety=1.0, bounding_radius_hint=100.0)
        raise KeyError(f"unknown bundled graph {name!r}")

    def set_param(self, path: str, value: float) -> None:
        """Scene::set_param for the node parameters of the classic graphs, with the
        graph bounds (scene/graph/mod.rs:44-73); raises ValueError naming the path."""
        key = path.split(".")[-1]
        fam = self.family
        allowed = {
            abi.FAMILY_MANDELBULB: {"iterations", "bailout", "power"},
            abi.FAMILY_MANDELBOX: {"iterations", "bailout", "scale", "min_radius", "fixed_radius", "limit"},
            abi.FAMILY_MENGER: {"iterations", "bailout", "scale"},
            abi.FAMILY_SPHERE: {"radius"},
        }[fam]
        if key not in allowed:
            raise ValueError(f"{path}: no such parameter on this graph")
        bkey = "menger_scale" if (key == "scale" and fam == abi.FAMILY_MENGER) else key
        if bkey in BOUNDS:
            lo, hi = BOUNDS[bkey]
            if not (lo <= value <= hi):
                raise ValueError(f"{path} = {value} is outside [{lo}, {hi}]")
        if key == "iterations":
            self.iterations = int(value)
        elif key == "limit":
            self.fold_limit = float(value)
        else:
            setattr(self, key, float(value))

    def c_struct(self) -> abi.SyfFractal:
        f = abi.SyfFractal()
        f.family = self.family
        f.iterations = self.iterations
        f.bailout = self.bailout
        f.power = self.power
        f.scale = self.scale
        f.min_radius = self.min_radius
        f.fixed_radius = self.fixed_radius
        f.fold_limit = self.fold_limit
        f.offset[:] = list(self.offset)
        f.half[:] = list(self.half)
        f.radius = self.radius
        f.plane_normal[:] = list(self.plane_normal)
        f.plane_offset = self.plane_offset
        f.trap_scale = self.trap_scale
        f.safety = self.safety
        f.bounding_radius_hint = self.bounding_radius_hint
        return f


@dataclass
class Camera:
    """scene/camera.rs Camera: Free{position,target,roll} or Orbit{point,distance,azimuth,
    elevation,roll}; focal_mm on a 36 mm sensor (tan half horizontal fov = 18 / focal_mm)."""
    mode: int = abi.CAMERA_FREE
    position: tuple = (0.0, 0.0, -2.5)
    target: tuple = (0.0, 0.0, 0.0)
    roll: float = 0.0
    point: tuple = (0.0, 0.0, 0.0)
    distance: float = 3.0
    azimuth: float = 0.0
    elevation: float = 0.0
    focal_mm: float = 40.0
    f_number: float = 0.0
    focus_distance: float = 3.0

    @staticmethod
    def free(position, target, focal_mm, roll=0.0) -> "Camera":
        return Camera(abi.CAMERA_FREE, tuple(position), tuple(target), roll, focal_mm=focal_mm)

    @staticmethod
    def orbit(point, distance, azimuth, elevation, focal_mm, roll=0.0) -> "Camera":
        return Camera(abi.CAMERA_ORBIT, point=tuple(point), distance=distance, azimuth=azimuth,
                      elevation=elevation, roll=roll, focal_mm=focal_mm)

    @staticmethod
    def focal_for_vertical_fov(vfov_deg: float, width: int, height: int) -> float:
        """focal_mm giving a vertical field of view on a width x height image."""
        tan_half_h = math.tan(math.radians(vfov_deg) * 0.5) * (width / height)
        return 18.0 / tan_half_h

    def c_struct(self) -> abi.SyfCamera:
        c = abi.SyfCamera()
        c.mode = self.mode
        c.position[:] = list(self.position)
        c.target[:] = list(self.target)
        c.roll = self.roll
        c.point[:] = list(self.point)
        c.distance = self.distance
        c.azimuth = self.azimuth
        c.elevation = self.elevation
        c.focal_mm = self.focal_mm
        c.f_number = self.f_number
        c.focus_distance = self.focus_distance
        return c


@dataclass
class Light:
    kind: int
    position: tuple = (0.0, 0.0, 0.0)
    normal: tuple = (0.0, -1.0, 0.0)
    size: tuple = (1.0, 1.0)
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    zenith: tuple = (1.0, 1.0, 1.0)
    horizon: tuple = (1.0, 1.0, 1.0)
    ground: tuple = (1.0, 1.0, 1.0)

    @staticmethod
    def area(position, normal, size, color=(1.0, 1.0, 1.0), intensity=1.0) -> "Light":
        return Light(abi.LIGHT_AREA, tuple(position), tuple(normal), tuple(size), tuple(color), intensity)

    @staticmethod
    def sun(direction, color=(1.0, 1.0, 1.0), intensity=1.0, angular_size=0.53) -> "Light":
        """Light::Sun: ``direction`` towards the sun, ``intensity`` the irradiance on a
        perpendicular surface, ``angular_size`` the disk's full angular diameter in degrees."""
        return Light(abi.LIGHT_SUN, normal=tuple(direction), size=(angular_size, 0.0), color=tuple(color),
                     intensity=intensity)

    @staticmethod
    def environment(zenith, horizon, ground, intensity=1.0) -> "Light":
        return Light(abi.LIGHT_ENV, intensity=intensity, zenith=tuple(zenith), horizon=tuple(horizon),
                     ground=tuple(ground))

    def c_struct(self) -> abi.SyfLight:
        l = abi.SyfLight()
        l.kind = self.kind
        l.position[:] = list(self.position)
        l.normal[:] = list(self.normal)
        l.size[:] = list(self.size)
        l.color[:] = list(self.color)
        l.intensity = self.intensity
        l.zenith[:] = list(self.zenith)
        l.horizon[:] = list(self.horizon)
        l.ground[:] = list(self.ground)
        return l


@dataclass
class Ramp:
    """Ramp (palette.rs:74-115): linear rgb stops, optional strictly increasing positions."""
    stops: List[tuple] = field(default_factory=lambda: [(0.0, 0.0, 0.0), (1.0, 1.0, 1.0)])
    positions: List[float] = field(default_factory=list)
    cyclic: bool = False
    frequency: float = 1.0
    phase: float = 0.0
    gamma: float = 1.0
    smooth: bool = False  # PaletteInterpolation::Smooth

    def position(self, i: int) -> float:
        if self.positions:
            return float(self.positions[i])
        n = len(self.stops)
        return i / float(n if self.cyclic else max(n - 1, 1))


@dataclass
class Coloring:
    """Coloring (material.rs:12-37): ``slot`` is a trap slot 0..3, COLOR_ITER or COLOR_NORMAL
    for Coloring::Trap; ``orbit`` (trap weights, iterations, normal, logarithmic) makes it
    Coloring::Orbit."""
    ramp: Ramp = field(default_factory=Ramp)
    slot: int = 0
    orbit: Optional[tuple] = None  # ((w0, w1, w2, w3), iterations, normal, logarithmic)

    def c_struct(self) -> abi.SyfColoring:
        c = abi.SyfColoring()
        r = self.ramp
        if len(r.stops) > abi.MAX_STOPS:
            raise ValueError(f"coloring: at most {abi.MAX_STOPS} stops, got {len(r.stops)}")
        if r.positions and len(r.positions) != len(r.stops):
            raise ValueError("coloring: one position per stop, or none")
        c.slot = abi.COLOR_ORBIT if self.orbit is not None else self.slot
        c.stop_count = len(r.stops)
        c.cyclic = int(r.cyclic)
        c.smooth = int(r.smooth)
        c.positioned = int(bool(r.positions))
        c.frequency = r.frequency
        c.phase = r.phase
        c.gamma = r.gamma
        for i, st in enumerate(r.stops):
            c.stops[i][:] = [float(st[0]), float(st[1]), float(st[2]), r.position(i)]
        w, it, nw, lg = self.orbit if self.orbit is not None else ((1.0, 0.0, 0.0, 0.0), 0.0, 0.0, 0.0)
        c.weights[:] = list(w)
        c.iteration_weight = it
        c.normal_weight = nw
        c.logarithmic = lg
        return c


@dataclass
class Finish:
    """Finish (material.rs:161-174) with its FinishDetail (material.rs:193-221): ``pattern`` 1..12
    numbers FinishPattern as abi.PATTERNS (0: no detail); the Concrete patterns (10..12) bring
    Oren-Nayar's rough diffuse lobe, which ``oren_nayar`` also asks for without a pattern."""
    albedo: tuple = (0.8, 0.8, 0.8)
    roughness: float = 0.5
    metallic: float = 0.0
    specular: float = 0.04
    emission: tuple = (0.0, 0.0, 0.0)
    oren_nayar: bool = False
    pattern: int = 0
    seed: int = 7
    scale: float = 2.0
    strength: float = 0.25
    bump: float = 0.004
    roughness_delta: float = 0.15
    tint: tuple = (0.3, 0.27, 0.2)

    def c_struct(self) -> abi.SyfFinish:
        f = abi.SyfFinish()
        f.albedo[:] = list(self.albedo)
        f.roughness = self.roughness
        f.metallic = self.metallic
        f.specular = self.specular
        f.emission[:] = list(self.emission)
        f.oren_nayar = int(self.oren_nayar)
        f.pattern = self.pattern
        if self.pattern:
            f.seed = self.seed
            f.scale = self.scale
            f.strength = self.strength
            f.bump = self.bump
            f.roughness_delta = self.roughness_delta
            f.tint[:] = list(self.tint)
        return f


@dataclass
class Ground:
    """GroundPlane (atmosphere.rs:32-38) without dunes: ``finish`` 0 is the base material,
    1..8 a named finish."""
    height: float = 0.0
    finish: int = 0


@dataclass
class Fog:
    """The atmosphere's height fog (Atmosphere fog fields, atmosphere.rs:7-16): extinction
    ``density`` at and below ``height``, decaying above it by exp(-(y - height) falloff);
    single-scattering albedo ``color``; Henyey-Greenstein ``anisotropy``."""
    density: float = 0.02
    color: tuple = (1.0, 1.0, 1.0)
    height: float = 0.0
    falloff: float = 1.0
    anisotropy: float = 0.0


@dataclass
class Material:
    albedo: tuple = (0.8, 0.8, 0.8)
    roughness: float = 0.4
    metallic: float = 0.0
    specular: float = 0.04
    surface: Optional[tuple] = None  # (slot, (r0, r1), (m0, m1)) : SurfaceVariation
    coloring: Optional[Coloring] = None  # None: Coloring::None
    finishes: List[Finish] = field(default_factory=list)  # MaterialTag index i + 1

    def c_struct(self) -> abi.SyfMaterial:
        m = abi.SyfMaterial()
        m.albedo[:] = list(self.albedo)
        m.roughness = self.roughness
        m.metallic = self.metallic
        m.specular = self.specular
        if self.surface is not None:
            slot, rough, metal = self.surface
            m.surface_enabled = 1
            m.surface_slot = slot
            m.surface_roughness[:] = list(rough)
            m.surface_metallic[:] = list(metal)
        return m


@dataclass
class RenderConfig:
    """renderer.rs:17-76 fields used on this path (defaults of RenderConfig::new)."""
    width: int
    height: int
    samples_per_pixel: int = 1
    max_bounces: int = 4
    max_steps: int = 256
    pixel_tolerance: float = 1.0
    tile_size: int = 256
    # arithmetic of the primary trace's FP32 march: "fast" (hardware reciprocal, reciprocal
    # square root and logarithm, FMA) or "exact" (IEEE operations as written); not a reference field
    march_math: str = "fast"
    # a node graph's kernels: "specialised" (compiled for its topology at run time) or
    # "interpreted" (the library's graph interpreter); not a reference field
    graph_mode: str = "specialised"


@dataclass
class Scene:
    fractal: Fractal  # or a syfrack_amd.graph.Graph
    camera: Camera
    lights: List[Light] = field(default_factory=list)
    material: Material = field(default_factory=Material)
    seed: int = 7
    cavity_darkness: float = 0.0
    ground: Optional[Ground] = None
    fog: Optional[Fog] = None

    @staticmethod
    def from_ron(text: str) -> "Scene":
        """A scene file of the reference (scenes/*.ron) for this path; what it does not render is
        refused with a ValueError naming the scene element (syfrack_amd.scene_file)."""
        from .scene_file import scene_from_ron
        return scene_from_ron(text)

    def with_fractal_param(self, path: str, value: float) -> "Scene":
        s = copy.deepcopy(self)
        s.fractal.set_param(path, value)
        return s

    def desc(self, cfg: RenderConfig) -> abi.SyfJobDesc:
        d = abi.SyfJobDesc()
        if hasattr(self.fractal, "nodes"):  # a node graph: the job is created with syf_job_create_graph
            d.fractal.family = abi.FAMILY_GRAPH
            d.fractal.safety = self.fractal.safety
            d.fractal.bounding_radius_hint = self.fractal.bounding_radius_hint
        else:
            d.fractal = self.fractal.c_struct()
        d.camera = self.camera.c_struct()
        d.material = self.material.c_struct()
        if len(self.lights) > abi.MAX_LIGHTS:
            raise ValueError(f"lights: at most {abi.MAX_LIGHTS} on this path, got {len(self.lights)}")
        for i, l in enumerate(self.lights):
            d.lights[i] = l.c_struct()
        d.light_count = len(self.lights)
        d.width = cfg.width
        d.height = cfg.height
        d.max_bounces = cfg.max_bounces
        d.max_steps = cfg.max_steps
        d.pixel_tolerance = cfg.pixel_tolerance
        d.cavity_darkness = self.cavity_darkness
        if cfg.march_math not in ("fast", "exact"):
            raise ValueError(f"march_math must be 'fast' or 'exact', got {cfg.march_math!r}")
        d.march_math = abi.MATH_FAST if cfg.march_math == "fast" else abi.MATH_EXACT
        d.seed = self.seed & 0xFFFFFFFFFFFFFFFF
        m = self.material
        if m.coloring is not None:
            d.coloring = m.coloring.c_struct()
        else:
            d.coloring.slot = abi.COLOR_NONE
            d.coloring.frequency = 1.0
            d.coloring.gamma = 1.0
        if len(m.finishes) > abi.MAX_FINISHES:
            raise ValueError(f"finishes: at most {abi.MAX_FINISHES}, got {len(m.finishes)}")
        for i, f in enumerate(m.finishes):
            d.finishes[i] = f.c_struct()
        d.finish_count = len(m.finishes)
        if self.ground is not None:
            d.ground.enabled = 1
            d.ground.height = self.ground.height
            d.ground.finish = self.ground.finish
        if self.fog is not None:
            d.fog.density = self.fog.density
            d.fog.color[:] = list(self.fog.color)
            d.fog.height = self.fog.height
            d.fog.falloff = self.fog.falloff
            d.fog.anisotropy = self.fog.anisotropy
        return d


# ---------------------------------------------------------------- named workloads

def mandelbulb_primary(width: int = 320, height: int = 180, seed: int = 7):
    """Mandelbulb primary-ray trace: power 8, 10 iterations, bailout 2.0, camera at
    (0, 0, -2.5) looking at the origin, 60 degree vertical field of view, 1 spp, <= 512
    steps, pixel tolerance 1.0, no bounces."""
    f = Fractal.bundled("mandelbulb")
    f.set_param("fractal.nodes[1].iterations", 10)
    f.set_param("fractal.nodes[1].bailout", 2.0)
    cam = Camera.free((0.0, 0.0, -2.5), (0.0, 0.0, 0.0), Camera.focal_for_vertical_fov(60.0, width, height))
    scene = Scene(f, cam, seed=seed)
    cfg = RenderConfig(width, height, samples_per_pixel=1, max_bounces=0, max_steps=512, pixel_tolerance=1.0)
    return scene, cfg


def mandelbox_pathtrace(width: int = 1920, height: int = 1080, spp: int = 64, seed: int = 11):
    """Mandelbox wavefront path trace: scale 2.0, min radius 0.5, fixed radius 1.0, fold
    limit 1.0, 15 iterations; 4 bounces; Lambert + metal through a trap-driven surface
    variation (metallic 1 -> 0, roughness 0.3); one 6x6 area light at (8, 14, 10) facing the
    origin, on the camera's side of the box; max 512 steps."""
    f = Fractal.bundled("mandelbox")
    f.set_param("fractal.nodes[1].iterations", 15)
    cam = Camera.orbit((0.0, 0.0, 0.0), 11.0, 28.0, 22.0, 35.0)
    light = Light.area((8.0, 14.0, 10.0), (-8.0, -14.0, -10.0), (6.0, 6.0), (1.0, 0.95, 0.9), 40.0)
    mat = Material(albedo=(0.8, 0.8, 0.8), roughness=0.3, metallic=0.0, specular=0.04,
                   surface=(0, (0.3, 0.3), (1.0, 0.0)))
    scene = Scene(f, cam, [light], mat, seed=seed)
    cfg = RenderConfig(width, height, samples_per_pixel=spp, max_bounces=4, max_steps=512, pixel_tolerance=1.0)
    return scene, cfg
