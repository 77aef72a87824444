"""The oracle's path-traced radiance pinned against what the mathematics fixes.

CPU only.  Each scene reduces the estimator (next-event estimation, BSDF continuation, MIS,
light selection, Russian roulette) to a value known without the oracle's own formulas:

  * the white furnace (gpu_render.rs:456-475): a grey convex sphere in a uniform white
    environment reflects its albedo, 0.5 +- 0.03, and the background is 1.0 +- 1e-3;
  * the metallic furnace (gpu_render.rs:585-608): a white metal sphere never exceeds 1.02,
    and the furnace rim stays within [0.3, 1.0] (gpu_render.rs:572-583);
  * a Lambert plane under an environment that is Z at the zenith and H at the horizon, with
    the reference's sqrt(elevation) blend: E = 2*pi * int_0^1 (H + (Z-H) sqrt(mu)) mu dmu, so
    L = rho * (H + 0.8 (Z - H)) -- pins the environment value, its cosine-sampled NEE pdf and,
    with one bounce, the MIS between NEE and BSDF continuation;
  * a Lambert plane under a small square emitter: L = rho/pi * E with E from a direct
    quadrature of Le cos_s cos_l / d^2 over the emitter's surface, per camera sample -- pins
    the area pdf (d^2 / (A cos_l)), the emitter's orientation and, with one bounce, the MIS
    weights of emitter hits;
  * two identical emitters render as one of double intensity (gpu_render.rs:397-421): the
    weighted light selection keeps the sum;
  * a rough metal plane in a white environment: the NEE-only estimate (cosine sampling, no
    VNDF) and the MIS estimate with VNDF continuation agree, and neither exceeds 1 -- a wrong
    VNDF frame or pdf biases the second;
  * surface-variation endpoints equal to the uniform material give the same image bit for bit
    (gpu_render.rs:282-306);
  * cavity darkness only darkens, monotonically (gpu_render.rs:840-868);
  * the Mandelbulb's trigonometric power equals the integer spherical product at integer
    powers (eval.rs:462-475 against eval.rs:129-178) and is continuous in the power.
Monte Carlo tolerances are several standard errors of the mean, measured from the image.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from syfrack_amd import scene as sc


def _render(scene, w, h, spp, bounces, steps=256):
    cfg = sc.RenderConfig(w, h, spp, bounces, steps)
    acc, dep, aux, _ = oracle.render(scene.desc(cfg), (0, 0, w, h), 0, spp)
    assert np.all(acc[..., 3] == spp)
    return acc[..., :3] / spp, dep, cfg


def furnace_scene(albedo=0.5, metallic=0.0, specular=0.0, roughness=0.5):
    """gpu_render.rs:62-105: unit sphere (no ground) in a uniform white environment, orbit
    camera at distance 3, 50 mm."""
    f = sc.Fractal.bundled("sphere")
    f.plane_offset = -1000.0  # the ground plane out of reach: the sphere alone
    f.bounding_radius_hint = 1.0
    cam = sc.Camera.orbit((0, 0, 0), 3.0, 0.0, 0.0, 50.0)
    env = sc.Light.environment((1, 1, 1), (1, 1, 1), (1, 1, 1), 1.0)
    mat = sc.Material(albedo=(albedo,) * 3, roughness=roughness, metallic=metallic, specular=specular)
    return sc.Scene(f, cam, [env], mat, seed=3)


def plane_scene(lights, albedo=0.5, metallic=0.0, specular=0.0, roughness=0.5, focal=200.0):
    """A plane y = 0 seen from straight above (camera at (0, 3, 0.5)), a 0.01 sphere at the
    origin (gpu_render.rs:641-693's geometry)."""
    f = sc.Fractal.bundled("sphere")
    f.radius = 0.01
    f.trap_scale = 0.01
    f.plane_offset = 0.0
    cam = sc.Camera.free((0.0, 3.0, 0.5), (0.0, 0.0, 0.5), focal)
    mat = sc.Material(albedo=(albedo,) * 3, roughness=roughness, metallic=metallic, specular=specular)
    return sc.Scene(f, cam, list(lights), mat, seed=5)


def _sem(img):
    """Standard error of the image mean from the per-pixel spread."""
    v = img.mean(axis=-1).ravel()
    return v.std(ddof=1) / math.sqrt(v.size)


def _rim(depth):
    """Deepest pixel whose 3x3 neighbourhood is all sphere (gpu_render.rs:439-455)."""
    h, w = depth.shape
    best, best_d = (w // 2, h // 2), -1.0
    for y in range(1, h - 1):
        for x in range(1, w - 1):
            if np.all(depth[y - 1:y + 2, x - 1:x + 2] >= 0) and depth[y, x] > best_d:
                best_d, best = depth[y, x], (x, y)
    return best


def test_white_furnace():
    img, dep, _ = _render(furnace_scene(), 48, 48, 64, 4)
    centre = img[22:26, 22:26].mean()
    corner = img[0:3, 0:3].mean()
    print(f"furnace centre {centre:.4f} corner {corner:.6f}")
    assert abs(centre - 0.5) < 0.03
    assert abs(corner - 1.0) < 1e-3
    assert abs(dep[24, 24] - 2.0) < 0.05


def test_furnace_rim_and_metallic_furnace_do_not_gain_energy():
    img, dep, _ = _render(furnace_scene(), 48, 48, 256, 4)
    rx, ry = _rim(dep)
    rim = img[ry - 1:ry + 2, rx - 1:rx + 2].mean()
    assert 0.3 <= rim <= 1.0, (rim, rx, ry)
    img, dep, _ = _render(furnace_scene(albedo=1.0, metallic=1.0, specular=0.04), 48, 48, 256, 4)
    rx, ry = _rim(dep)
    centre = img[22:26, 22:26].mean()
    rim = img[ry - 1:ry + 2, rx - 1:rx + 2].mean()
    print(f"metallic furnace centre {centre:.4f} rim {rim:.4f}")
    assert centre <= 1.02 and rim <= 1.02
    assert centre > 0.8  # a white metal in a white furnace keeps most of the energy


@pytest.mark.parametrize("bounces", [0, 1])
def test_lambert_plane_under_gradient_environment(bounces):
    zen, hor = 1.0, 0.25
    env = sc.Light.environment((zen,) * 3, (hor,) * 3, (0.0,) * 3, 1.0)
    img, dep, _ = _render(plane_scene([env]), 16, 16, 256, bounces)
    assert np.all(dep >= 0)
    expected = 0.5 * (hor + 0.8 * (zen - hor))
    got, sem = img.mean(), _sem(img)
    print(f"gradient env: {got:.5f} expected {expected:.5f} (sem {sem:.2e})")
    assert abs(got - expected) < max(6 * sem, 2e-3 * expected) + 2e-3 * expected


def _emitter_irradiance(light, points, n=32):
    """E at plane points (normal +y) from a one-sided parallelogram emitter, by midpoint
    quadrature over its surface; the emitter frame is the reference's GpuLight::from_light
    (uniforms.rs:676-716)."""
    nrm = np.asarray(light.normal, np.float64)
    nrm = nrm / np.linalg.norm(nrm)
    helper = np.array([0, 0, 1.0]) if abs(nrm[1]) > 0.999 else np.array([0, 1.0, 0])
    u = np.cross(helper, nrm)
    u /= np.linalg.norm(u)
    v = np.cross(nrm, u)
    b = u * light.size[0]
    c = v * light.size[1]
    area = np.linalg.norm(np.cross(b, c))
    s = (np.arange(n) + 0.5) / n - 0.5
    q = (np.asarray(light.position)[None, None, :] + s[:, None, None] * b[None, None, :]
         + s[None, :, None] * c[None, None, :]).reshape(-1, 3)
    le = light.intensity * np.asarray(light.color)[0]
    out = np.zeros(len(points))
    for k in range(0, len(points), 512):
        p = points[k:k + 512]
        d = q[None, :, :] - p[:, None, :]
        d2 = np.sum(d * d, axis=-1)
        dist = np.sqrt(d2)
        cos_s = np.maximum(d[..., 1] / dist, 0.0)
        cos_l = np.maximum(-np.sum(d * nrm[None, None, :], axis=-1) / dist, 0.0)
        out[k:k + 512] = le * np.sum(cos_s * cos_l / d2, axis=-1) * (area / len(q))
    return out


@pytest.mark.parametrize("bounces", [0, 1])
def test_lambert_plane_under_area_light_is_rho_e_over_pi(bounces):
    # facing straight down from off to the side: cos_l ~ 0.8 and cos_s ~ 0.8 at the view
    lamp = sc.Light.area((1.5, 2.0, 0.5), (0.0, -1.0, 0.0), (0.5, 0.5), (1.0, 1.0, 1.0), 6.0)
    scene = plane_scene([lamp])
    w = h = 8
    spp = 512 if bounces == 0 else 2048
    img, dep, cfg = _render(scene, w, h, spp, bounces)
    assert np.all(dep >= 0)
    # expected value of each pixel: the mean over its camera samples of rho/pi * E at the
    # sample's shading point, the plane point lifted along the normal by the pixel footprint
    # max(2 t pixel_cone, 1e-3) * pixel_tolerance (pathtrace.wgsl:1085,1109)
    desc = scene.desc(cfg)
    cone = (18.0 / scene.camera.focal_mm) * 2.0 / w
    exp = np.zeros(w * h)
    ns = 512  # the same camera samples as the first 512 of the render
    for s in range(ns):
        ro, rd = oracle.camera_rays(desc, s)
        t = -ro[:, 1].astype(np.float64) / rd[:, 1]
        pts = ro.astype(np.float64) + rd.astype(np.float64) * t[:, None]
        pts[:, 1] += np.maximum(t * cone * 2.0, 1e-3)
        exp += 0.5 / math.pi * _emitter_irradiance(lamp, pts)
    exp /= ns
    res = img.mean(axis=-1).ravel() - exp
    sem = res.std(ddof=1) / math.sqrt(res.size)
    print(f"area light, {bounces} bounces: {img.mean():.5f} expected {exp.mean():.5f} "
          f"(residual {res.mean():.2e}, sem {sem:.2e})")
    assert abs(res.mean()) < 6 * sem + 2e-3 * exp.mean()
    if bounces == 0:  # NEE alone on a small emitter: nearly noiseless per pixel
        rel = np.abs(img.mean(axis=-1).ravel() / exp - 1.0)
        print(f"per-pixel max relative deviation {rel.max():.2e}")
        assert rel.max() < 3e-2


@pytest.mark.parametrize("bounces", [0, 2])
def test_identical_emitters_sum_like_one_of_double_intensity(bounces):
    lamp = sc.Light.area((0.0, 3.0, 3.0), (0.0, -1.0, -1.0), (3.0, 3.0), (1.0, 1.0, 1.0), 4.0)
    scene = furnace_scene()
    scene.lights = [lamp, lamp]
    pair, _, _ = _render(scene, 24, 24, 32, bounces)
    scene.lights = [sc.Light.area((0.0, 3.0, 3.0), (0.0, -1.0, -1.0), (3.0, 3.0), (1.0, 1.0, 1.0), 8.0)]
    one, _, _ = _render(scene, 24, 24, 32, bounces)
    if bounces == 0:  # gpu_render.rs:397-421: the same random stream, the same values
        assert np.abs(pair - one).max() < 1e-5
    else:  # emitter hits carry different MIS weights: equal in expectation
        assert abs(pair.mean() - one.mean()) < 6 * math.hypot(_sem(pair), _sem(one)) + 1e-3


@pytest.mark.parametrize("view", ["normal", "oblique"])
def test_rough_metal_nee_and_vndf_estimates_agree(view):
    env = sc.Light.environment((1, 1, 1), (1, 1, 1), (1, 1, 1), 1.0)
    scene = plane_scene([env], albedo=1.0, metallic=1.0, specular=0.04, roughness=0.6)
    if view == "oblique":  # 69 degrees off the normal, where Smith's G1(v) is well below 1
        scene.camera = sc.Camera.free((0.0, 1.0, 4.6), (0.0, 0.0, 2.0), 200.0)
    nee, _, _ = _render(scene, 8, 8, 4096, 0)
    mis, _, _ = _render(scene, 8, 8, 4096, 1)
    a, b = nee.mean(), mis.mean()
    tol = 6 * math.hypot(_sem(nee), _sem(mis)) + 2e-3
    print(f"metal plane: NEE {a:.5f} MIS+VNDF {b:.5f} (tol {tol:.2e})")
    assert a <= 1.0 + tol and b <= 1.0 + tol
    assert abs(a - b) < tol


def test_surface_variation_endpoints_match_uniform_materials():
    scene = furnace_scene(specular=0.04)
    for slot, rough, metal in [(0, 0.25, 0.0), (3, 0.35, 1.0)]:
        scene.material.roughness = rough
        scene.material.metallic = metal
        scene.material.surface = None
        uni, _, _ = _render(scene, 24, 24, 8, 4)
        scene.material.surface = (slot, (rough, rough), (metal, metal))
        mapped, _, _ = _render(scene, 24, 24, 8, 4)
        assert np.array_equal(uni, mapped)
        scene.material.surface = (slot, (rough, rough), (0.1, 0.9))
        varied, _, _ = _render(scene, 24, 24, 8, 4)
        assert np.all(np.isfinite(varied))
        if slot == 0:  # slot 0 carries the sphere's trap; the others read 1 (unwritten)
            assert not np.array_equal(mapped, varied)


def test_cavity_darkness_only_darkens_monotonically():
    scene, cfg = sc.mandelbox_pathtrace(32, 32, 4)
    out = []
    for cav in (0.0, 0.4, 0.8):
        scene.cavity_darkness = cav
        img, _, _ = _render(scene, 32, 32, 4, 2, 256)
        out.append(img[..., 1])
    a, mid, b = out
    assert b.sum() < mid.sum() < a.sum()
    assert np.all(b <= a + 1e-4)


def _pow_points():
    rng = np.random.default_rng(12341)
    pts = rng.uniform(-1.5, 1.5, (2000, 3)).astype(np.float32)
    return np.concatenate([pts, np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1], [1e-6, 0, 1]], np.float32)])


@pytest.mark.parametrize("n", list(range(2, 13)))
def test_trig_power_equals_spherical_product_at_integer_powers(n):
    p = _pow_points()
    r = np.linalg.norm(p.astype(np.float64), axis=1)
    spher, dsph = oracle.spherical_power(p, n)
    for prec, tol in ((64, 2e-6), (32, 2e-5)):
        trig, dtrig = oracle.trig_power(p, float(n), prec)
        scale = np.maximum(r ** n, 1e-20)
        err = np.linalg.norm(trig - spher, axis=1) / scale
        assert err.max() < tol, (prec, err.max())
        assert np.max(np.abs(dtrig - dsph) / np.maximum(n * r ** (n - 1), 1e-20)) < tol
        # |z^n| = r^n whatever the angles
        assert np.allclose(np.linalg.norm(trig, axis=1), r ** n, rtol=tol, atol=1e-20)


def test_trig_power_is_continuous_in_the_power():
    p = _pow_points()
    a, da = oracle.trig_power(p, 8.0, 64)
    b, db = oracle.trig_power(p, 8.0 + 1e-4, 64)
    r = np.linalg.norm(p.astype(np.float64), axis=1)
    # d/dn of r^n e^{i n angle} is bounded by r^n (|ln r| + pi): a step of 1e-4 moves it by ~1e-4
    bound = 1e-4 * (np.abs(np.log(np.maximum(r, 1e-12))) + 2 * math.pi + 1) * np.maximum(r ** 8, 1e-20)
    assert np.all(np.linalg.norm(a - b, axis=1) <= bound * 1.01 + 1e-12)
