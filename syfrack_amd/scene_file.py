"""Scene files of the reference (``scenes/*.ron``) read into a ``Scene`` for this path.

A scene file is the RON form of ``Scene`` (crates/syfrack-core/src/scene/mod.rs:92-110): seed,
form graph, camera, lights, material, atmosphere and look.  ``scene_from_ron`` reads the fields
this path renders and refuses, with an error naming the scene element, what it does not:

- a 2D scene, a creation graph, phenomena (mod.rs:92-110), a form graph with an Instances node;
- projector, tube and HDRI lights (light.rs:66-123);
- a texture, optics and more than 8 finishes (material.rs:140-255);
- the seascape, portal, nebula and dunes (atmosphere.rs:7-48).

The look (tonemap, grading, grain, vignette) is the display pipeline applied after readback and is
not part of a render job; it is read and ignored.  Light colours are ``Kelvin(t)`` or
``Linear([r, g, b])`` (Color, light.rs:10-26).  A scene without ``fractal`` gets the bundled sphere
(default_form, mod.rs:112-114); a file naming one of the removed hand-written formulas or
colourings is refused with the reference's message (is_pre_plan4, mod.rs:76-89).
"""
from __future__ import annotations

import math
from typing import List

import numpy as np

from .graph import Graph
from .ron import Variant, from_ron
from .scene import Camera, Coloring, Finish, Fog, Ground, Light, Material, Ramp, Scene
from . import abi

LEGACY_MESSAGE = "legacy scene format: use the previous version to convert this scene"
MAX_LIGHTS_REFERENCE = 96  # light.rs:7


class SceneFileError(ValueError):
    """A scene file this path cannot render, or one that does not parse as a scene."""


def srgb_to_linear(v: float) -> float:
    """The sRGB transfer function inverted (IEC 61966-2-1), as srgb_to_linear, light.rs:56-62."""
    return v / 12.92 if v <= 0.04045 else ((v + 0.055) / 1.055) ** 2.4


def kelvin_to_linear(kelvin: float) -> tuple:
    """A black-body colour by Tanner Helland's published fit of the blackbody sRGB colour to
    temperature (the method of kelvin_to_linear, light.rs:28-54): the temperature clamped to
    [1000, 40000] K in hundreds; red 255 up to 6600 K, then a power law in (t - 60); green a
    logarithm up to 6600 K, then a power law; blue 0 up to 1900 K, a logarithm in (t - 10) up to
    6600 K, then 255.  The 8-bit sRGB values are clamped to [0, 1], made linear and scaled so the
    brightest channel is 1.  Computed in float32, as the reference does."""
    f = np.float32
    t = f(min(max(float(kelvin), 1000.0), 40000.0)) / f(100.0)
    if t <= f(66.0):
        r = f(255.0)
        g = f(99.4708025861) * f(math.log(t)) - f(161.1195681661)
    else:
        r = f(329.698727446) * f(float(t - f(60.0)) ** -0.1332047592)
        g = f(288.1221695283) * f(float(t - f(60.0)) ** -0.0755148492)
    if t >= f(66.0):
        b = f(255.0)
    elif t <= f(19.0):
        b = f(0.0)
    else:
        b = f(138.5177312231) * f(math.log(t - f(10.0))) - f(305.0447927307)
    c = [f(srgb_to_linear(float(min(max(x / f(255.0), f(0.0)), f(1.0))))) for x in (r, g, b)]
    m = max(c[0], c[1], c[2], f(1e-6))
    return tuple(float(x / m) for x in c)


def color(v, where: str) -> tuple:
    """Color::to_linear (light.rs:18-26)."""
    if isinstance(v, Variant) and v.variant == "Kelvin" and len(v) == 1:
        return kelvin_to_linear(float(v[0]))
    if isinstance(v, Variant) and v.variant == "Linear" and len(v) == 1 and len(v[0]) == 3:
        return tuple(float(x) for x in v[0])
    raise SceneFileError(f"{where}: a colour is Kelvin(t) or Linear([r, g, b]), got {v!r}")


def _v3(v) -> tuple:
    return tuple(float(x) for x in v)


def _camera(d: dict) -> Camera:
    """Camera with CameraMode Free or Orbit (camera.rs:5-35)."""
    mode = d["mode"]
    kind = mode.get("__struct__") if isinstance(mode, dict) else None
    lens = dict(focal_mm=float(d["focal_mm"]), f_number=float(d["f_number"]),
                focus_distance=float(d["focus_distance"]))
    if kind == "Free":
        return Camera(abi.CAMERA_FREE, position=_v3(mode["position"]), target=_v3(mode["target"]),
                      roll=float(mode.get("roll", 0.0)), **lens)
    if kind == "Orbit":
        return Camera(abi.CAMERA_ORBIT, point=_v3(mode["point"]), distance=float(mode["distance"]),
                      azimuth=float(mode["azimuth"]), elevation=float(mode["elevation"]),
                      roll=float(mode.get("roll", 0.0)), **lens)
    raise SceneFileError(f"camera.mode: Free or Orbit, got {mode!r}")


def _light(i: int, d) -> Light:
    """Light (light.rs:66-123) for the kinds this path renders."""
    kind = d.get("__struct__") if isinstance(d, dict) else None
    where = f"lights[{i}]"
    if kind == "Sun":
        return Light.sun(_v3(d["direction"]), color(d["color"], where + ".color"), float(d["intensity"]),
                         float(d["angular_size"]))
    if kind == "Area":
        return Light.area(_v3(d["position"]), _v3(d["normal"]), tuple(float(x) for x in d["size"]),
                          color(d["color"], where + ".color"), float(d["intensity"]))
    if kind == "Environment":
        # rotation only turns HDR maps; a gradient ignores it (light.rs:105-113)
        return Light.environment(color(d["zenith"], where + ".zenith"), color(d["horizon"], where + ".horizon"),
                                 color(d["ground"], where + ".ground"), float(d["intensity"]))
    if kind in ("Projector", "Tube", "Hdri"):
        raise SceneFileError(f"{where}: {kind} is not rendered on this path (sun, area and environment lights are)")
    raise SceneFileError(f"{where}: unknown light {d!r}")


def _ramp(d: dict) -> Ramp:
    """Ramp (palette.rs:72-100) with its serde defaults."""
    interp = d.get("interpolation", "Linear")
    return Ramp(stops=[_v3(s) for s in d["stops"]], positions=[float(p) for p in d.get("positions", [])],
                cyclic=bool(d.get("cyclic", False)), frequency=float(d.get("frequency", 1.0)),
                phase=float(d.get("phase", 0.0)), gamma=float(d.get("gamma", 1.0)), smooth=interp == "Smooth")


def _coloring(c) -> Coloring:
    """Coloring None, Trap{slot, ramp} or Orbit{channels, ramp} (material.rs:7-37)."""
    if c == "None" or c is None:
        return None
    kind = c.get("__struct__") if isinstance(c, dict) else None
    if kind == "Trap":
        return Coloring(_ramp(c["ramp"]), int(c["slot"]))
    if kind == "Orbit":
        ch = c.get("channels", {})
        traps = tuple(float(x) for x in ch.get("traps", (1.0, 0.0, 0.0, 0.0)))
        return Coloring(_ramp(c["ramp"]), 0, (traps, float(ch.get("iterations", 0.0)), float(ch.get("normal", 0.0)),
                                               float(ch.get("logarithmic", 0.0))))
    raise SceneFileError(f"material.coloring: None, Trap or Orbit, got {c!r}")


def _finish(i: int, d: dict) -> Finish:
    """Finish with its FinishDetail (material.rs:161-221), the pattern numbered as abi.PATTERNS."""
    where = f"material.finishes[{i}]"
    if d.get("optics") is not None:
        raise SceneFileError(f"{where}.optics: not rendered on this path")
    f = Finish(albedo=_v3(d["albedo"]), roughness=float(d["roughness"]), metallic=float(d["metallic"]),
               specular=float(d["specular"]), emission=_v3(d.get("emission", (0.0, 0.0, 0.0))))
    det = d.get("detail")
    if det is not None:
        name = det.get("pattern")
        if name not in abi.PATTERNS:
            raise SceneFileError(f"{where}.detail.pattern: unknown pattern {name!r}")
        f.pattern = abi.PATTERNS.index(name) + 1
        f.seed = int(det["seed"])
        f.scale = float(det["scale"])
        f.strength = float(det["strength"])
        f.bump = float(det["bump"])
        f.roughness_delta = float(det["roughness_delta"])
        f.tint = _v3(det["tint"])
    return f


def _material(d: dict) -> Material:
    if d.get("texture") is not None:
        raise SceneFileError("material.texture: not rendered on this path")
    if d.get("optics") is not None:
        raise SceneFileError("material.optics: not rendered on this path")
    finishes = d.get("finishes", [])
    if len(finishes) > abi.MAX_FINISHES:
        raise SceneFileError(f"material.finishes: at most {abi.MAX_FINISHES}, got {len(finishes)}")
    s = d.get("surface")
    surface = None if s is None else (int(s["slot"]), tuple(float(x) for x in s["roughness"]),
                                      tuple(float(x) for x in s["metallic"]))
    return Material(albedo=_v3(d["albedo"]), roughness=float(d["roughness"]), metallic=float(d["metallic"]),
                    specular=float(d["specular"]), surface=surface, coloring=_coloring(d.get("coloring")),
                    finishes=[_finish(i, f) for i, f in enumerate(finishes)])


def _form(d: dict) -> Graph:
    if d.get("creation") is not None:
        raise SceneFileError("creation: a creation graph is not lowered on this path (a form graph is)")
    fr = d.get("fractal")
    if fr is None:
        return Graph.bundled("sphere")
    if isinstance(fr, Variant) and fr.variant == "Graph" and len(fr) == 1 and isinstance(fr[0], dict):
        g = Graph.from_dict(fr[0])
    elif isinstance(fr, dict) and fr.get("__struct__") == "Graph":
        g = Graph.from_dict(fr)
    else:
        raise SceneFileError(f"fractal: a Graph, got {type(fr).__name__}")
    for n in g.nodes:
        if n.kind == "Instances" and n.enabled:
            raise SceneFileError(f"fractal: node {n.id}: Instances (independent copies) is not evaluated on this path")
    return g


def scene_from_dict(d: dict) -> Scene:
    """The parsed RON of a scene file as a Scene; refuses what this path does not render."""
    if not isinstance(d, dict):
        raise SceneFileError("a scene file holds one Scene struct")
    if d.get("dimension", "ThreeD") != "ThreeD":
        raise SceneFileError("dimension: a 2D scene is rendered with syf_planar_create / syf_planar_run")
    if d.get("phenomena") is not None:
        raise SceneFileError("phenomena: not rendered on this path")
    for key in ("seed", "camera", "lights", "material", "atmosphere"):
        if key not in d:
            raise SceneFileError(f"scene parse error: missing field `{key}`")
    a = d["atmosphere"]
    for key in ("seascape", "portal", "nebula"):
        if a.get(key) is not None:
            raise SceneFileError(f"atmosphere.{key}: not rendered on this path")
    g = a.get("ground")
    if g is not None and g.get("dunes") is not None:
        raise SceneFileError("atmosphere.ground.dunes: not rendered on this path (a flat ground is)")
    lights_d = d["lights"]
    if not lights_d:
        raise SceneFileError("a scene needs at least one light")
    if len(lights_d) > MAX_LIGHTS_REFERENCE:
        raise SceneFileError(f"a scene may have at most {MAX_LIGHTS_REFERENCE} lights, got {len(lights_d)}")
    lights: List[Light] = [_light(i, l) for i, l in enumerate(lights_d)]
    if len(lights) > abi.MAX_LIGHTS:
        raise SceneFileError(f"lights: at most {abi.MAX_LIGHTS} on this path, got {len(lights)}")
    material = _material(d["material"])
    fog = None
    if float(a["fog_density"]) > 0.0:  # GpuAtmosphere::from_atmosphere keeps the fog only when dense
        fog = Fog(density=float(a["fog_density"]), color=_v3(a["fog_color"]), height=float(a["fog_height"]),
                  falloff=float(a["fog_falloff"]), anisotropy=float(a["anisotropy"]))
    ground = None if g is None else Ground(height=float(g["height"]), finish=int(g["finish"]))
    return Scene(_form(d), _camera(d["camera"]), lights, material, seed=int(d["seed"]),
                 cavity_darkness=float(a["cavity_darkness"]), ground=ground, fog=fog)


def _is_legacy(text: str) -> bool:
    """is_pre_plan4 (mod.rs:76-89): the removed formula variants and colourings, whitespace ignored."""
    flat = "".join(text.split())
    return (any(f"fractal:{v}(" in flat for v in ("Sphere", "Mandelbulb", "Mandelbox", "MengerFold"))
            or any(f"coloring:{v}(" in flat for v in ("OrbitTrap", "Normal", "Iteration")))


def scene_from_ron(text: str) -> Scene:
    """A scene file's text as a Scene for this path (Scene::load's parse, mod.rs)."""
    if _is_legacy(text):
        raise SceneFileError(LEGACY_MESSAGE)
    try:
        d = from_ron(text)
    except (ValueError, KeyError, IndexError) as e:
        raise SceneFileError(f"scene parse error: {e}") from e
    return scene_from_dict(d)


def load_scene(path: str) -> Scene:
    with open(path, encoding="utf-8") as fh:
        return scene_from_ron(fh.read())
