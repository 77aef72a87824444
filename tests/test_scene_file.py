"""Scene files of the reference read by syfrack_amd.scene_file: the RON forms they use, the
light colours (Kelvin and linear), the fields lowered and the elements refused by name.

The Kelvin colour follows Tanner Helland's published black-body fit; it is pinned by what the
method fixes (6600 K is white, no blue at or below 1900 K, the clamp to [1000, 40000] K, the
brightest channel scaled to 1, red falling and blue rising with temperature above 6600 K) and by
the fit's widely quoted 8-bit outputs at 1000 K (255, 68, 0), 2700 K (255, 167, 87), 6500 K
(255, 254, 250) and 10000 K (201, 218, 255).  Scaling the brightest channel to 1 is pinned by
nothing: after the clamp to 255, red or blue is always 1.
"""
import glob
import math
import os

import pytest

from syfrack_amd import abi
from syfrack_amd.ron import Variant, from_ron
from syfrack_amd.scene_file import (LEGACY_MESSAGE, SceneFileError, kelvin_to_linear, load_scene, scene_from_ron,
                                    srgb_to_linear)

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = os.path.join(HERE, "golden", "scenes")
REFERENCE_SCENES = "/root/reference/scenes"


def _to_srgb8(v):
    s = v * 12.92 if v <= 0.0031308 else 1.055 * v ** (1 / 2.4) - 0.055
    return s * 255.0


def test_ron_keeps_variant_names_and_reads_short_numbers():
    d = from_ron("(c: Kelvin(4300.0), l: Linear((.2, -.35, 1.)), g: Graph((a: 1)), n: r#None, s: Some(3))")
    assert isinstance(d["c"], Variant) and d["c"].variant == "Kelvin" and d["c"] == (4300.0,)
    assert d["l"].variant == "Linear" and d["l"][0] == (0.2, -0.35, 1.0)
    assert d["g"].variant == "Graph" and d["g"][0] == {"a": 1}
    assert d["n"] is None and d["s"] == 3


def test_srgb_transfer_known_values():
    assert srgb_to_linear(0.0) == 0.0 and srgb_to_linear(1.0) == pytest.approx(1.0)
    assert srgb_to_linear(0.5) == pytest.approx(0.214041, abs=1e-6)
    # the two pieces meet at 0.04045
    assert srgb_to_linear(0.04045) == pytest.approx(((0.04045 + 0.055) / 1.055) ** 2.4, rel=1e-4)


def test_kelvin_colour_pins():
    white = kelvin_to_linear(6600.0)
    assert white == pytest.approx((1.0, 1.0, 1.0))
    for k in (1000.0, 1500.0, 1900.0):
        assert kelvin_to_linear(k)[2] == 0.0
    assert kelvin_to_linear(500.0) == kelvin_to_linear(1000.0)
    assert kelvin_to_linear(90000.0) == kelvin_to_linear(40000.0)
    for k in (1200.0, 2700.0, 4300.0, 6500.0, 9500.0, 20000.0):
        assert max(kelvin_to_linear(k)) == pytest.approx(1.0)
    # the fit's 8-bit outputs
    c = [_to_srgb8(v) for v in kelvin_to_linear(6500.0)]
    assert c == pytest.approx([255.0, 254.0, 250.0], abs=1.0)
    c = [_to_srgb8(v) for v in kelvin_to_linear(1000.0)]
    assert c == pytest.approx([255.0, 68.0, 0.0], abs=1.0)
    c = [_to_srgb8(v) for v in kelvin_to_linear(2700.0)]
    assert c == pytest.approx([255.0, 167.0, 87.0], abs=1.0)
    c = [_to_srgb8(v) for v in kelvin_to_linear(10000.0)]
    assert c == pytest.approx([201.0, 218.0, 255.0], abs=1.0)
    # warmer below 6600 K, bluer above
    ks = [1000.0 + 250.0 * i for i in range(157)]
    ratio = [kelvin_to_linear(k)[2] / kelvin_to_linear(k)[0] for k in ks]
    assert all(b >= a - 1e-7 for a, b in zip(ratio, ratio[1:]))
    hot = [kelvin_to_linear(k) for k in (7000.0, 10000.0, 20000.0, 40000.0)]
    assert all(b[0] < a[0] for a, b in zip(hot, hot[1:]))


def test_bundled_fixtures_lower_every_field():
    s = load_scene(os.path.join(SCENES, "exhibition.ron"))
    assert [l.kind for l in s.lights] == [abi.LIGHT_SUN, abi.LIGHT_AREA, abi.LIGHT_ENV]
    sun, lamp, sky = s.lights
    assert sun.normal == (-0.4, 0.8, 0.3) and sun.intensity == 7.0 and sun.size[0] == 1.5
    assert sun.color == kelvin_to_linear(4300.0) and lamp.color == kelvin_to_linear(9500.0)
    assert lamp.position == (-6.0, 4.0, -8.0) and lamp.size == (4.0, 3.0)
    assert sky.zenith == (0.2, 0.35, 0.7) and sky.ground == (0.15, 0.12, 0.1) and sky.intensity == 0.5
    assert s.fog is not None and (s.fog.density, s.fog.height, s.fog.falloff, s.fog.anisotropy) == (0.025, -2.0, 0.4, 0.55)
    assert s.cavity_darkness == 0.3 and s.seed == 7 and s.ground is None
    assert s.camera.mode == abi.CAMERA_ORBIT and (s.camera.distance, s.camera.azimuth, s.camera.elevation) == (11.0, 28.0, 22.0)
    assert (s.camera.focal_mm, s.camera.f_number, s.camera.focus_distance) == (35.0, 4.0, 4.5)
    c = s.material.coloring
    assert c.slot == 0 and c.orbit is None and len(c.ramp.stops) == 3 and c.ramp.stops[1] == (0.9, 0.85, 0.75)
    assert s.material.albedo == (0.82, 0.78, 0.72) and s.material.roughness == 0.3
    assert len(s.fractal.nodes) == 9 and s.fractal.safety == 0.6
    d = s.desc(__import__("syfrack_amd.scene", fromlist=["RenderConfig"]).RenderConfig(64, 36, 1))
    assert d.light_count == 3 and d.fog.density == pytest.approx(0.025) and d.fractal.family == abi.FAMILY_GRAPH
    for name in ("forme-exposition", "orfevrerie"):
        s = load_scene(os.path.join(SCENES, name + ".ron"))
        assert s.fog is not None and s.lights[0].kind == abi.LIGHT_SUN


BASE = """(
    seed: 3,
    fractal: Graph((nodes: [(id: 0, kind: Input, name: "i", pos: (0.0, 0.0), enabled: true),
                            (id: 1, kind: Sphere(radius: 1.0), name: "s", pos: (0.0, 0.0), enabled: true),
                            (id: 2, kind: Output, name: "o", pos: (0.0, 0.0), enabled: true)],
                    links: [(from: (0, 0), to: (1, 0)), (from: (1, 0), to: (2, 0))],
                    safety: 1.0, bounding_radius_hint: 2.0)),
    camera: (mode: Free(position: (0.0, 0.0, -3.0), target: (0.0, 0.0, 0.0), roll: 0.0), focal_mm: 40.0,
             f_number: 0.0, focus_distance: 3.0),
    lights: [LIGHTS],
    material: (albedo: (0.8, 0.8, 0.8), roughness: 0.4, metallic: 0.0, specular: 0.04, coloring: None MATERIAL),
    atmosphere: (fog_density: 0.0, fog_color: (1.0, 1.0, 1.0), fog_height: 0.0, fog_falloff: 1.0, anisotropy: 0.0,
                 cavity_darkness: 0.0 ATMOSPHERE),
    look: (exposure_ev: 0.0) EXTRA
)"""
SUN = "Sun(direction: (0.0, 1.0, 0.0), color: Linear((1.0, 1.0, 1.0)), intensity: 2.0, angular_size: 0.53)"


def _scene(lights=SUN, material="", atmosphere="", extra=""):
    return (BASE.replace("LIGHTS", lights).replace("MATERIAL", material).replace("ATMOSPHERE", atmosphere)
            .replace("EXTRA", extra))


def test_minimal_scene_and_defaults():
    s = scene_from_ron(_scene())
    assert s.fog is None and s.material.coloring is None and s.camera.mode == abi.CAMERA_FREE
    assert s.lights[0].color == (1.0, 1.0, 1.0)
    text = _scene()
    no_form = text.replace(text[text.index("fractal:"):text.index("camera:")], "")
    assert "fractal" not in no_form and "Sphere" in [n.kind for n in scene_from_ron(no_form).fractal.nodes]
    concrete = (", finishes: [(name: \"c\", albedo: (0.5, 0.5, 0.5), roughness: 0.8, metallic: 0.0, specular: 0.04,"
                " detail: Some((pattern: Concrete, seed: 7, scale: 2.0, strength: 0.0, bump: 0.0,"
                " roughness_delta: 0.0, tint: (0.3, 0.3, 0.3))))]")
    s = scene_from_ron(_scene(material=concrete))
    assert s.material.finishes[0].pattern == 10 and s.material.finishes[0].roughness == 0.8
    orbit = _scene().replace("coloring: None", "coloring: Orbit(channels: (traps: (0.5, 0.5, 0.0, 0.0), iterations: 1.0,"
                             " normal: 0.0, logarithmic: 2.0), ramp: (stops: [(0.0, 0.0, 0.0), (1.0, 1.0, 1.0)],"
                             " interpolation: Smooth, cyclic: true))")
    c = scene_from_ron(orbit).material.coloring
    assert c.orbit == ((0.5, 0.5, 0.0, 0.0), 1.0, 0.0, 2.0) and c.ramp.smooth and c.ramp.cyclic
    veined = (", finishes: [(name: \"v\", albedo: (0.5, 0.5, 0.5), roughness: 0.4, metallic: 0.0, specular: 0.04,"
              " detail: Some((pattern: VeinedStone, seed: 9, scale: 2.5, strength: 0.25, bump: 0.004,"
              " roughness_delta: 0.15, tint: (0.3, 0.27, 0.2))))]")
    f = scene_from_ron(_scene(material=veined)).material.finishes[0]
    assert (f.pattern, f.seed, f.scale, f.strength, f.bump, f.roughness_delta, f.tint) == (
        1, 9, 2.5, 0.25, 0.004, 0.15, (0.3, 0.27, 0.2))
    c = f.c_struct()
    assert c.pattern == 1 and c.seed == 9 and abs(c.bump - 0.004) < 1e-9
    fog = _scene().replace("fog_density: 0.0", "fog_density: 0.1")
    assert scene_from_ron(fog).fog.density == 0.1
    ground = _scene(atmosphere=", ground: Some((height: -1.0, finish: 0))")
    assert scene_from_ron(ground).ground.height == -1.0


@pytest.mark.parametrize("text, names", [
    (_scene(lights="Projector(position: (0.0, 0.0, 0.0), normal: (0.0, -1.0, 0.0), size: (1.0, 1.0), "
                   "color: Kelvin(3000.0), intensity: 1.0, inner_angle: 10.0, outer_angle: 20.0)"), "lights[0]: Projector"),
    (_scene(lights=SUN + ", Tube(start: (0.0, 0.0, 0.0), end: (1.0, 0.0, 0.0), radius: 0.1, color: Kelvin(3000.0),"
                   " intensity: 1.0)"), "lights[1]: Tube"),
    (_scene(lights=SUN + ", Hdri(path: \"sky.hdr\", intensity: 1.0, rotation: 0.0)"), "lights[1]: Hdri"),
    (_scene(lights=""), "at least one light"),
    (_scene(lights=", ".join([SUN] * 5)), "at most 4"),
    (_scene(material=", texture: Some((channel: Albedo))"), "material.texture"),
    (_scene(material=", optics: Some((ior: 1.5))"), "material.optics"),
    (_scene(atmosphere=", seascape: Some((level: 0.0))"), "atmosphere.seascape"),
    (_scene(atmosphere=", nebula: Some((density: 1.0))"), "atmosphere.nebula"),
    (_scene(atmosphere=", portal: Some((energy: 1.0))"), "atmosphere.portal"),
    (_scene(atmosphere=", ground: Some((height: 0.0, finish: 0, dunes: Some((amplitude: 1.0, wavelength: 1.0, "
                       "wind_direction: 0.0, seed: 1))))"), "dunes"),
    (_scene(extra=", phenomena: Some((kind: Aurora))"), "phenomena"),
    (_scene().replace("seed: 3,", "dimension: TwoD, seed: 3,"), "2D scene"),
    (_scene().replace("fractal: Graph(", "creation: Some((nodes: [])), fractal: Graph("), "creation"),
    (_scene().replace("coloring: None", "coloring: OrbitTrap(ramp: 1)"), LEGACY_MESSAGE),
    (_scene().replace("camera:", "kamera:"), "missing field `camera`"),
])
def test_refusals_name_the_element(text, names):
    with pytest.raises(SceneFileError, match=None) as e:
        scene_from_ron(text)
    assert names in str(e.value), str(e.value)


@pytest.mark.skipif(not os.path.isdir(REFERENCE_SCENES), reason="the reference's scene files are not on this machine")
def test_every_reference_scene_loads_or_is_refused_by_name():
    """Every scene file the reference ships parses; each either lowers to a job description or is
    refused with a SceneFileError naming what this path does not render."""
    from syfrack_amd.scene import RenderConfig
    loaded = refused = 0
    for path in sorted(glob.glob(os.path.join(REFERENCE_SCENES, "**", "*.ron"), recursive=True)):
        try:
            s = load_scene(path)
        except SceneFileError as e:
            assert "parse error" not in str(e), (path, str(e))
            refused += 1
            continue
        s.desc(RenderConfig(32, 18, 1))
        loaded += 1
    assert loaded >= 71 and refused > 0
    print(f"{loaded} scene files lowered, {refused} refused by name")


def test_patterned_scene_survives_the_session_state():
    from syfrack_amd.ron import to_ron
    from syfrack_amd.session import scene_from_state
    s = load_scene(os.path.join(SCENES, "3d-01.ron"))
    back = scene_from_state(from_ron(to_ron({"scene": s}))["scene"])
    assert back.material.finishes == s.material.finishes and back.ground == s.ground
    assert [f.pattern for f in back.material.finishes] == [11, 3, 0, 10]
