"""The oracle's sun, flat ground, colouring ramp and named finishes pinned on the CPU.

Each case reduces the estimator to a value fixed without the oracle's own formulas:

  * the ramp (palette_sample, common.wgsl:336-366) against the reference's known answers
    (palette.rs:225-260): the legacy two-stop ramp, the cyclic seam, positioned stops with
    frequency and phase;
  * a Lambert ground under a sun of irradiance E whose disk (angular radius r) is wholly above
    the horizon: L = rho/pi * E * E[cos] with E[cos] = cos(theta) (1 + cos r) / 2 for directions
    uniform over the disk -- pins the cone sampling, the sun's pdf and colour (irradiance) and,
    with one bounce and a wide disk, the MIS between the sun's NEE and BSDF hits of the disk;
  * the ground's finish, its emission (added once, unweighted), and the ground's field sample
    (traps 1, iterations 0, the finish as material: pathtrace.wgsl:86, lights.wgsl:201), which
    makes every colouring slot a closed form: trap and normal slots read 1, the iteration slot
    reads 0, the orbit slot OrbitColoring::value (material.rs:39-53) of those;
  * the Oren-Nayar lobe of a rough finish against the qualitative model in the form of Oren and
    Nayar (1994), A + B max(0, cos(phi_i - phi_r)) sin(alpha) tan(beta), alpha = max(theta_i,
    theta_r), beta = min(theta_i, theta_r), with sigma = 0.45 roughness (pathtrace.wgsl:381-384);
    the code form divides by max(cos) instead of multiplying by tan(min);
  * the ground blocks what lies below it: an emitter under the ground lights nothing above.
Monte Carlo tolerances are several standard errors of the mean, measured from the image.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from syfrack_amd import abi
from syfrack_amd import scene as sc


def _render(scene, w, h, spp, bounces, steps=256):
    cfg = sc.RenderConfig(w, h, spp, bounces, steps)
    acc, dep, aux, _ = oracle.render(scene.desc(cfg), (0, 0, w, h), 0, spp)
    assert np.all(acc[..., 3] == spp)
    return acc[..., :3] / spp, dep, aux, cfg


def _sem(img):
    v = img.mean(axis=-1).ravel()
    return v.std(ddof=1) / math.sqrt(v.size)


def ground_scene(lights, material=None, ground=None, camera=None):
    """A flat ground at y = 0 seen from straight above (camera at (0, 3, 0.5)); the field is a
    0.01 sphere at the origin, out of view, with its own plane out of reach."""
    f = sc.Fractal.bundled("sphere")
    f.radius = 0.01
    f.trap_scale = 0.01
    f.plane_offset = -1000.0
    f.bounding_radius_hint = 1.0
    cam = camera or sc.Camera.free((0.0, 3.0, 0.5), (0.0, 0.0, 0.5), 200.0)
    mat = material or sc.Material(albedo=(0.5,) * 3, roughness=0.5, metallic=0.0, specular=0.0)
    return sc.Scene(f, cam, list(lights), mat, seed=9, ground=ground or sc.Ground(0.0, 0))


def _sun_dir(theta_deg):
    t = math.radians(theta_deg)
    return (math.sin(t), math.cos(t), 0.0)


def _lambert_sun(rho, e, theta_deg, diameter_deg):
    cos_r = math.cos(math.radians(diameter_deg * 0.5))
    return rho / math.pi * e * math.cos(math.radians(theta_deg)) * (1.0 + cos_r) * 0.5


# ---------------------------------------------------------------- the ramp

def _coloring(stops, positions=(), cyclic=False, frequency=1.0, phase=0.0, gamma=1.0, smooth=False, slot=0,
              orbit=None):
    return sc.Coloring(sc.Ramp([tuple(s) for s in stops], list(positions), cyclic, frequency, phase, gamma, smooth),
                       slot, orbit).c_struct()


def test_ramp_known_answers_of_the_reference():
    legacy = _coloring([(0, 0, 0), (1, 1, 1)])
    v = oracle.ramp_sample(legacy, [0.25, -1.0])
    assert np.array_equal(v[0], [0.25] * 3) and np.array_equal(v[1], [0.0] * 3)
    cyc = _coloring([(1, 0, 0), (0, 0, 1)], cyclic=True)
    v = oracle.ramp_sample(cyc, [0.25, -0.25, 0.75, 1.0, 0.0, 1.0 - 1e-6, 1e-6])
    assert np.array_equal(v[0], [0.5, 0.0, 0.5])
    assert np.array_equal(v[1], v[2]) and np.array_equal(v[3], v[4])
    assert abs(v[5][0] - v[6][0]) < 1e-5
    pos = _coloring([(1, 0, 0), (0, 1, 0), (0, 0, 1)], positions=(0.1, 0.2, 0.8), cyclic=True)
    v = oracle.ramp_sample(pos, [0.2, 0.95])
    assert np.array_equal(v[0], [0.0, 1.0, 0.0])
    assert abs(v[1][0] - 0.5) < 1e-6
    pos2 = _coloring([(1, 0, 0), (0, 1, 0), (0, 0, 1)], positions=(0.1, 0.2, 0.8), cyclic=True, frequency=2.0,
                     phase=-0.1)
    assert abs(oracle.ramp_sample(pos2, [0.15])[0][1] - 1.0) < 1e-6


def test_ramp_edges_single_stop_smooth_and_gamma():
    assert np.array_equal(oracle.ramp_sample(_coloring([]), [0.3])[0], [1.0, 1.0, 1.0])  # no stops: white
    assert np.array_equal(oracle.ramp_sample(_coloring([(0.2, 0.4, 0.6)]), [0.9])[0], np.float32([0.2, 0.4, 0.6]))
    # smooth-step interpolation: 3 x^2 - 2 x^3 of the linear fraction
    sm = oracle.ramp_sample(_coloring([(0, 0, 0), (1, 1, 1)], smooth=True), [0.25])[0][0]
    assert abs(sm - (3 * 0.25 ** 2 - 2 * 0.25 ** 3)) < 1e-6
    # gamma maps the coordinate before the lookup
    g = oracle.ramp_sample(_coloring([(0, 0, 0), (1, 1, 1)], gamma=2.0), [0.5])[0][0]
    assert abs(g - 0.25) < 1e-6


# ---------------------------------------------------------------- sun and ground

# the widest disk (90 degrees) gives BSDF hits of the disk a sizeable MIS weight
@pytest.mark.parametrize("bounces,diameter", [(0, 0.53), (0, 30.0), (1, 40.0), (1, 90.0)])
def test_lambert_ground_under_a_sun(bounces, diameter):
    e, theta = 3.0, 40.0
    sun = sc.Light.sun(_sun_dir(theta), (1.0, 1.0, 1.0), e, diameter)
    img, dep, aux, _ = _render(ground_scene([sun]), 8, 8, 512 if bounces else 128, bounces)
    assert np.all(dep > 0) and np.allclose(aux[..., 1], 1.0)  # the ground's normal faces up
    expected = _lambert_sun(0.5, e, theta, diameter)
    got, sem = img.mean(), _sem(img)
    print(f"sun {diameter} deg, {bounces} bounces: {got:.5f} expected {expected:.5f} (sem {sem:.2e})")
    assert abs(got - expected) < 6 * sem + 3e-3 * expected


def test_sun_scales_with_irradiance_and_colour():
    a = _render(ground_scene([sc.Light.sun(_sun_dir(30.0), (1.0, 1.0, 1.0), 1.0, 0.53)]), 4, 4, 16, 0)[0]
    b = _render(ground_scene([sc.Light.sun(_sun_dir(30.0), (0.5, 1.0, 2.0), 4.0, 0.53)]), 4, 4, 16, 0)[0]
    assert np.allclose(b, a * np.float32([2.0, 4.0, 8.0]), rtol=1e-5)


def test_sun_below_the_horizon_lights_nothing():
    sun = sc.Light.sun((0.3, -1.0, 0.0), (1.0, 1.0, 1.0), 5.0, 0.53)
    img = _render(ground_scene([sun]), 4, 4, 16, 1)[0]
    assert np.all(img == 0.0)


def test_ground_finish_and_emission():
    e, theta = 2.0, 35.0
    sun = sc.Light.sun(_sun_dir(theta), (1.0, 1.0, 1.0), e, 0.53)
    glow = (0.2, 0.1, 0.05)
    mat = sc.Material(albedo=(0.9,) * 3, roughness=0.5, specular=0.0,
                      finishes=[sc.Finish((0.3, 0.6, 0.2), 0.5, 0.0, 0.0, glow)])
    img = _render(ground_scene([sun], mat, sc.Ground(0.0, 1)), 4, 4, 64, 0)[0]
    base = np.array([_lambert_sun(r, e, theta, 0.53) for r in (0.3, 0.6, 0.2)])
    got = img.reshape(-1, 3).mean(axis=0)
    assert np.allclose(got, base + np.array(glow), rtol=3e-3), (got, base + np.array(glow))
    # finish 0 is the base material; a finish index beyond the table also falls back to it
    for finish in (0, 2):
        img = _render(ground_scene([sun], mat, sc.Ground(0.0, finish)), 4, 4, 64, 0)[0]
        assert abs(img.mean() - _lambert_sun(0.9, e, theta, 0.53)) < 3e-3 * img.mean()


@pytest.mark.parametrize("slot,expect", [(0, (0, 0, 1)), (3, (0, 0, 1)), (abi.COLOR_NORMAL, (0, 0, 1)),
                                         (abi.COLOR_ITER, (1, 0, 0))])
def test_coloring_slots_on_the_ground(slot, expect):
    sun = sc.Light.sun(_sun_dir(30.0), (1.0, 1.0, 1.0), 2.0, 0.53)
    ramp = sc.Ramp([(1, 0, 0), (0, 1, 0), (0, 0, 1)])
    mat = sc.Material(albedo=(0.5,) * 3, roughness=0.5, specular=0.0, coloring=sc.Coloring(ramp, slot))
    img = _render(ground_scene([sun], mat), 4, 4, 32, 0)[0]
    got = img.reshape(-1, 3).mean(axis=0) / _lambert_sun(0.5, 2.0, 30.0, 0.53)
    assert np.allclose(got, expect, atol=3e-3), got


def _orbit_value(weights, iterations, normal, logarithmic, traps, iters, normal_y):
    """OrbitColoring::value (material.rs:39-53)."""
    v = sum(w * t for w, t in zip(weights, traps)) + iterations * iters + normal * (normal_y * 0.5 + 0.5)
    if logarithmic == 0.0:
        return v
    return math.copysign(1.0, v) * math.log(1.0 + logarithmic * abs(v)) / math.log(1.0 + logarithmic)


@pytest.mark.parametrize("log_gain", [0.0, 3.0])
def test_orbit_coloring_on_the_ground(log_gain):
    sun = sc.Light.sun(_sun_dir(30.0), (1.0, 1.0, 1.0), 2.0, 0.53)
    orbit = ((0.25, 0.125, 0.0, 0.0), 0.3, 0.1, log_gain)
    ramp = sc.Ramp([(1, 0, 0), (0, 1, 0), (0, 0, 1)])
    mat = sc.Material(albedo=(0.5,) * 3, roughness=0.5, specular=0.0, coloring=sc.Coloring(ramp, 0, orbit))
    img = _render(ground_scene([sun], mat), 4, 4, 32, 0)[0]
    value = _orbit_value(*orbit, traps=(1, 1, 1, 1), iters=0.0, normal_y=1.0)
    x = min(max(value, 0.0), 1.0) * 2.0  # three evenly spaced stops
    i = min(int(math.floor(x)), 1)
    f = x - i
    stops = np.array(ramp.stops, np.float64)
    expect = stops[i] * (1 - f) + stops[i + 1] * f
    got = img.reshape(-1, 3).mean(axis=0) / _lambert_sun(0.5, 2.0, 30.0, 0.53)
    assert np.allclose(got, expect, atol=3e-3), (got, expect, value)


def _oren_nayar_paper(rho, sigma, cos_i, cos_r, cos_dphi):
    """Oren-Nayar qualitative model, radiance over irradiance (f times pi / pi)."""
    s2 = sigma * sigma
    a = 1.0 - 0.5 * s2 / (s2 + 0.33)
    b = 0.45 * s2 / (s2 + 0.09)
    ti, tr = np.arccos(cos_i), np.arccos(cos_r)
    alpha, beta = np.maximum(ti, tr), np.minimum(ti, tr)
    return rho / math.pi * (a + b * np.maximum(0.0, cos_dphi) * np.sin(alpha) * np.tan(beta))


@pytest.mark.parametrize("view_deg,dphi_deg", [(30.0, 0.0), (55.0, 0.0), (45.0, 60.0)])
def test_oren_nayar_finish_against_the_qualitative_model(view_deg, dphi_deg):
    e, sun_deg, rho = 2.0, 50.0, 0.6
    sun = sc.Light.sun(_sun_dir(sun_deg), (1.0, 1.0, 1.0), e, 0.01)
    tv, ph = math.radians(view_deg), math.radians(dphi_deg)
    eye = (3.0 * math.sin(tv) * math.cos(ph), 3.0 * math.cos(tv), 0.5 + 3.0 * math.sin(tv) * math.sin(ph))
    cam = sc.Camera.free(eye, (0.0, 0.0, 0.5), 1200.0)
    mat = sc.Material(albedo=(0.8,) * 3, finishes=[sc.Finish((rho,) * 3, 1.0, 0.0, 0.0, (0, 0, 0), True)])
    scene = ground_scene([sun], mat, sc.Ground(0.0, 1), cam)
    w = h = 4
    spp = 16
    img, _, _, cfg = _render(scene, w, h, spp, 0)
    desc = scene.desc(cfg)
    l = np.array(_sun_dir(sun_deg))
    exp = np.zeros(w * h)
    for s in range(spp):
        _, rd = oracle.camera_rays(desc, s)
        v = -rd.astype(np.float64)
        v /= np.linalg.norm(v, axis=1)[:, None]
        cos_dphi = (v[:, 0] * l[0] + v[:, 2] * l[2]) / np.maximum(np.hypot(v[:, 0], v[:, 2]) * math.hypot(l[0], l[2]),
                                                                  1e-12)
        exp += _oren_nayar_paper(rho, 0.45, l[1], v[:, 1], cos_dphi) * e * l[1]
    exp /= spp
    got = img.mean(axis=-1).ravel()
    lam = rho / math.pi * e * l[1]
    print(f"Oren-Nayar view {view_deg}/{dphi_deg}: {got.mean():.5f} expected {exp.mean():.5f} (Lambert {lam:.5f})")
    assert np.allclose(got, exp, rtol=3e-3)


def test_ground_blocks_an_emitter_below_it():
    lamp = sc.Light.area((0.0, -0.5, 0.5), (0.0, 1.0, 0.0), (2.0, 2.0), (1.0, 1.0, 1.0), 20.0)
    img, dep, _, _ = _render(ground_scene([lamp]), 4, 4, 64, 2)
    assert np.all(dep > 0) and np.all(dep < 3.5)  # the ground, not the emitter 0.5 below it
    assert np.all(img == 0.0)
    # without the ground the emitter is seen
    scene = ground_scene([lamp])
    scene.ground = None
    img, _, _, _ = _render(scene, 4, 4, 16, 0)
    assert np.all(img > 1.0)
