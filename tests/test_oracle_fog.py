"""The oracle's height fog pinned on the CPU.

Each case reduces the fog estimator to a value fixed without the oracle's own formulas:

  * the Halton coordinates: the radical inverse of sample + 1 in the dimension's prime base
    (2, 3, 5, 7, 11, 13), rotated by pcg(stream ^ dimension * 0x9e3779b9) / 2^32 (24 bits), mod 1;
  * optical depth against a fine quadrature of sigma(y) = density exp(min(0, -(y - height)
    falloff)) along rays that stay below, stay above, and cross the fog height either way;
  * free flight: the collision distances' distribution function is 1 - exp(-tau(t)), with the
    atom exp(-tau(t_end)) at t_end (Kolmogorov-Smirnov distance);
  * Henyey-Greenstein: the phase integrates to 1 over the sphere, its samples have mean cosine g
    and follow the phase (histogram);
  * an absorbing fog (albedo 0) shows the sky attenuated by exp(-tau) along the camera ray, for a
    homogeneous and for a height fog (tau by quadrature);
  * the volume white furnace: a non-absorbing fog (albedo 1) in a uniform sky has radiance equal
    to the sky's everywhere, whatever the density and anisotropy (L = T + (1 - T) for a
    homogeneous medium) -- pins the free flight, the phase sampling, the sky's next-event
    estimation at scattering events and its MIS against phase-sampled escapes;
  * single scattering of a small emitter's light, against a quadrature of
    int sigma e^(-sigma t) int phase Le cos_l / d^2 e^(-sigma d) dA dt along the camera ray --
    pins the equiangular and distance sampling of fog_segment_direct and their combination;
  * the session planes: VOLUME + SURFACE = ACCUM per pixel, MOMENTS >= 0, and without fog VOLUME
    stays zero;
  * an emitter's light is counted once after a scattering event: with one bounce and no sky, the
    radiance is a S1 + a^2 S2 in the fog albedo a for the same random draws.
Monte Carlo tolerances are several standard errors of the mean, measured from the image.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from syfrack_amd import scene as sc


def _render(scene, w, h, spp, bounces, session=False):
    cfg = sc.RenderConfig(w, h, spp, bounces, 256)
    out = oracle.render(scene.desc(cfg), (0, 0, w, h), 0, spp, session=session)
    assert np.all(out[0][..., 3] == spp)
    return (out[0][..., :3] / spp,) + tuple(out[1:]) + (cfg,)


def _sem(img):
    v = img.mean(axis=-1).ravel()
    return v.std(ddof=1) / math.sqrt(v.size)


def open_sky(lights, fog, camera=None):
    """Nothing in view: a 0.01 sphere at the origin behind the camera, its plane out of reach."""
    f = sc.Fractal.bundled("sphere")
    f.radius = 0.01
    f.trap_scale = 0.01
    f.plane_offset = -1000.0
    f.bounding_radius_hint = 1.0
    cam = camera or sc.Camera.free((0.0, 0.0, 5.0), (0.0, 0.0, 10.0), 50.0)
    mat = sc.Material(albedo=(0.5,) * 3, roughness=0.5, specular=0.0)
    return sc.Scene(f, cam, list(lights), mat, seed=17, fog=fog)


def _desc(fog):
    return open_sky([], fog).desc(sc.RenderConfig(4, 4, 1, 0, 64))


# ---------------------------------------------------------------- the pieces

def _radical_inverse(i, base):
    v, inv = 0.0, 1.0 / base
    while i:
        v += (i % base) * inv
        i //= base
        inv /= base
    return v


@pytest.mark.parametrize("dimension,base", [(0, 2), (1, 3), (2, 5), (3, 7), (4, 11), (5, 13)])
def test_volume_coordinate_is_a_rotated_radical_inverse(dimension, base):
    lib = oracle.load()
    for stream in (0, 1, 0x27d4eb2d, 0xdeadbeef):
        shift = (lib.orc_pcg((stream ^ (dimension * 0x9e3779b9)) & 0xFFFFFFFF) >> 8) / 16777216.0
        for i in (0, 1, 2, 3, 7, 100, 4097):
            expect = (_radical_inverse(i + 1, base) + shift) % 1.0
            got = oracle.volume_coordinate(i, dimension, stream)
            assert abs(got - expect) < 2e-6 or abs(abs(got - expect) - 1.0) < 2e-6, (i, stream, got, expect)


def _tau_quadrature(fog, ro, rd, dist, n=200001):
    t = np.linspace(0.0, dist, n)
    y = ro[1] + rd[1] * t
    sigma = fog.density * np.exp(np.minimum(0.0, -(y - fog.height) * fog.falloff))
    return np.trapezoid(sigma, t)


def test_optical_depth_against_quadrature():
    fog = sc.Fog(0.3, (1, 1, 1), 0.5, 0.8, 0.0)
    d = _desc(fog)
    rng = np.random.default_rng(4)
    rays = [((0, 0, 0), (1, 0, 0), 7.0), ((0, 0, 0), (0, 1, 0), 9.0), ((0, 3, 0), (0, -1, 0), 9.0),
            ((0, -2, 0), (0.6, 0.8, 0), 12.0), ((0, 4, 0), (0.6, -0.8, 0), 12.0), ((0, 2, 0), (1, 0, 0), 5.0),
            ((0, 0.5, 0), (0, 1, 0), 3.0)]
    for _ in range(40):
        ro = rng.uniform(-3, 3, 3)
        rd = rng.normal(size=3)
        rays.append((ro, rd / np.linalg.norm(rd), rng.uniform(0.1, 20.0)))
    ro = np.array([r[0] for r in rays], np.float32)
    rd = np.array([r[1] for r in rays], np.float32)
    dist = np.array([r[2] for r in rays], np.float32)
    got = oracle.fog_optical_depth(d, ro, rd, dist)
    for k, (o, v, t) in enumerate(rays):
        expect = _tau_quadrature(fog, np.float64(ro[k]), np.float64(rd[k]), float(dist[k]))
        assert abs(got[k] - expect) <= 2e-4 * expect + 1e-6, (k, got[k], expect)
    # homogeneous: density x distance
    d0 = _desc(sc.Fog(0.3, (1, 1, 1), 0.5, 0.0, 0.0))
    assert np.allclose(oracle.fog_optical_depth(d0, ro, rd, dist), 0.3 * dist, rtol=1e-6)


@pytest.mark.parametrize("falloff,ro,rd", [(0.0, (0, 0, 0), (1, 0, 0)), (0.7, (0, -1, 0), (0.6, 0.8, 0)),
                                           (0.7, (0, 3, 0), (0.6, -0.8, 0))])
def test_free_flight_follows_the_optical_depth(falloff, ro, rd):
    fog = sc.Fog(0.15, (1, 1, 1), 0.0, falloff, 0.0)
    d = _desc(fog)
    t_end = 12.0
    n = 200000
    t = np.sort(oracle.fog_free_flight(d, ro, rd, t_end, n).astype(np.float64))
    assert t.min() >= 0.0 and t.max() <= t_end
    grid = np.linspace(0.0, t_end, 121)[1:-1]
    tau = oracle.fog_optical_depth(d, np.tile(ro, (len(grid), 1)), np.tile(rd, (len(grid), 1)), grid).astype(np.float64)
    cdf = 1.0 - np.exp(-tau)
    emp = np.searchsorted(t, grid, side="right") / n
    ks = np.abs(emp - cdf).max()
    tau_end = oracle.fog_optical_depth(d, [ro], [rd], [t_end])[0]
    atom = np.mean(t == np.float32(t_end))
    print(f"free flight KS {ks:.4f}, escape {atom:.4f} vs {math.exp(-tau_end):.4f}")
    assert ks < 5.0 / math.sqrt(n)
    assert abs(atom - math.exp(-tau_end)) < 5.0 * math.sqrt(atom * (1 - atom) / n) + 1e-4


@pytest.mark.parametrize("g", [0.0, 0.3, 0.75, -0.5])
def test_henyey_greenstein_phase_and_samples(g):
    mu = np.linspace(-1.0, 1.0, 20001)
    p = np.array([oracle.hg_phase(float(m), g) for m in mu[::10]])
    assert abs(2 * math.pi * np.trapezoid(p, mu[::10]) - 1.0) < 2e-3
    n = 100000
    rng = np.random.default_rng(7)
    r = rng.random((n, 2)).astype(np.float32)
    d = np.array([0.3, -0.5, 0.81], np.float32)
    d /= np.linalg.norm(d)
    s = oracle.hg_sample(g, d, r).astype(np.float64)
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-5)
    cos = s @ d.astype(np.float64)
    assert abs(cos.mean() - g) < 5 * cos.std() / math.sqrt(n) + 1e-3
    hist, edges = np.histogram(cos, bins=20, range=(-1, 1))
    expect = np.array([2 * math.pi * np.trapezoid([oracle.hg_phase(float(x), g) for x in np.linspace(a, b, 41)],
                                                  np.linspace(a, b, 41)) for a, b in zip(edges[:-1], edges[1:])]) * n
    assert np.all(np.abs(hist - expect) < 6 * np.sqrt(expect) + 5), (hist, expect.round())


# ---------------------------------------------------------------- radiance

def test_absorbing_fog_attenuates_the_sky():
    sky = sc.Light.environment((0.7,) * 3, (0.7,) * 3, (0.7,) * 3, 1.0)
    for fog in (sc.Fog(0.006, (0, 0, 0), 0.0, 0.0, 0.0), sc.Fog(0.05, (0, 0, 0), -1.0, 0.3, 0.0)):
        scene = open_sky([sky], fog, sc.Camera.free((0.0, 0.0, 5.0), (0.0, 2.0, 10.0), 200.0))
        img, _, _, _, cfg = _render(scene, 4, 4, 512, 2)
        desc = scene.desc(cfg)
        ro, rd = oracle.camera_rays(desc, 0)
        t_end = 200.0  # the ray extent of this scene
        tau = np.array([_tau_quadrature(fog, np.float64(o), np.float64(v), t_end) for o, v in zip(ro, rd)])
        expect = 0.7 * np.exp(-tau).mean()
        got, sem = img.mean(), _sem(img)
        print(f"absorbing fog {fog.density}/{fog.falloff}: {got:.5f} expected {expect:.5f} (sem {sem:.2e})")
        assert abs(got - expect) < 6 * sem + 2e-3 * expect


@pytest.mark.parametrize("density,g", [(0.004, 0.0), (0.004, 0.6), (0.01, -0.3)])
def test_volume_white_furnace(density, g):
    sky = sc.Light.environment((1, 1, 1), (1, 1, 1), (1, 1, 1), 1.0)
    scene = open_sky([sky], sc.Fog(density, (1, 1, 1), 0.0, 0.0, g))
    img, _, _, _, _ = _render(scene, 8, 8, 64, 32)
    got, sem = img.mean(), _sem(img)
    print(f"volume furnace {density} g {g}: {got:.5f} (sem {sem:.2e})")
    assert abs(got - 1.0) < 6 * sem + 2e-3


def _single_scatter(fog, lamp, ro, rd, t_end, nt=1500, na=24):
    """int_0^t_end sigma e^(-sigma t) int_A phase Le cos_l / d^2 e^(-sigma d) dA dt, homogeneous."""
    nrm = np.asarray(lamp.normal, np.float64)
    nrm /= np.linalg.norm(nrm)
    helper = np.array([0, 0, 1.0]) if abs(nrm[1]) > 0.999 else np.array([0, 1.0, 0])
    u = np.cross(helper, nrm)
    u /= np.linalg.norm(u)
    v = np.cross(nrm, u)
    b, c = u * lamp.size[0], v * lamp.size[1]
    area = np.linalg.norm(np.cross(b, c))
    s = (np.arange(na) + 0.5) / na - 0.5
    q = (np.asarray(lamp.position)[None, None] + s[:, None, None] * b + s[None, :, None] * c).reshape(-1, 3)
    t = (np.arange(nt) + 0.5) / nt * t_end
    p = ro[None] + rd[None] * t[:, None]
    dvec = q[None] - p[:, None]
    d2 = np.sum(dvec * dvec, axis=-1)
    dist = np.sqrt(d2)
    wi = dvec / dist[..., None]
    cos_l = np.maximum(-np.sum(wi * nrm, axis=-1), 0.0)
    g = fog.anisotropy
    mu = np.sum(wi * rd[None, None], axis=-1)
    den = 1 + g * g - 2 * g * mu
    phase = (1 - g * g) / (4 * math.pi * den * np.sqrt(den))
    sig = fog.density
    inner = np.sum(phase * cos_l / d2 * np.exp(-sig * dist), axis=1) * (area / len(q))
    le = lamp.intensity * lamp.color[0]
    return le * np.sum(sig * np.exp(-sig * t) * inner) * (t_end / nt)


@pytest.mark.parametrize("g", [0.0, 0.7])
def test_single_scattering_of_an_emitter(g):
    fog = sc.Fog(0.05, (1, 1, 1), 0.0, 0.0, g)
    lamp = sc.Light.area((1.5, 0.0, 9.0), (-1.0, 0.0, 0.0), (0.4, 0.4), (1, 1, 1), 40.0)
    scene = open_sky([lamp], fog, sc.Camera.free((0.0, 0.0, 5.0), (0.0, 0.0, 10.0), 400.0))
    w = h = 4
    spp = 256
    img, _, _, _, cfg = _render(scene, w, h, spp, 0)
    desc = scene.desc(cfg)
    exp = 0.0
    for smp in range(0, spp, 32):
        ro, rd = oracle.camera_rays(desc, smp)
        exp += np.mean([_single_scatter(fog, lamp, np.float64(o), np.float64(v), 200.0) for o, v in zip(ro, rd)])
    exp /= len(range(0, spp, 32))
    got, sem = img.mean(), _sem(img)
    print(f"single scatter g {g}: {got:.5f} expected {exp:.5f} (sem {sem:.2e})")
    assert abs(got - exp) < 6 * sem + 1e-2 * exp


def test_session_planes_partition_the_radiance():
    sky = sc.Light.environment((0.6, 0.7, 0.9), (0.5, 0.5, 0.5), (0.1, 0.1, 0.1), 1.0)
    lamp = sc.Light.area((1.5, 0.0, 9.0), (-1.0, 0.0, 0.0), (0.4, 0.4), (1, 1, 1), 40.0)
    f = sc.Fractal.bundled("mandelbulb")
    f.set_param("fractal.nodes[1].iterations", 6)
    cam = sc.Camera.orbit((0, 0, 0), 3.5, 20.0, 15.0, 35.0)
    scene = sc.Scene(f, cam, [sky, lamp], sc.Material(), seed=3, fog=sc.Fog(0.08, (0.9, 0.9, 0.9), 0.0, 0.5, 0.4))
    acc, _, _, _, planes, _ = _render(scene, 12, 8, 8, 3, session=True)
    vol, mom, surf = planes
    assert np.all(vol[..., 3] == 8) and np.all(mom[..., 3] == 8) and np.all(surf[..., 3] == 8)
    assert np.allclose((vol[..., :3] + surf[..., :3]) / 8, acc, rtol=1e-4, atol=1e-5)
    assert np.all(mom[..., :3] >= 0)
    assert vol[..., :3].sum() > 0
    # without fog the in-scattering is zero: the volume plane stays zero
    scene.fog = None
    acc0, _, _, _, planes0, _ = _render(scene, 12, 8, 4, 3, session=True)
    assert np.all(planes0[0][..., :3] == 0) and np.allclose(planes0[2][..., :3] / 4, acc0, rtol=1e-6)


def test_emitter_light_is_counted_once_after_a_scattering_event():
    """With one bounce and no sky, a path's radiance is a S1 + a^2 S2 in the fog albedo a
    (single scattering of the lamp on the camera segment, then on the segment after one phase
    sample), with the same random draws whatever a.  A phase-sampled segment that hits the lamp
    adds nothing: the line integral of its segment already counted the lamp.  Counting it again
    would add a term linear in a."""
    lamp = sc.Light.area((1.5, 0.0, 8.0), (-1.0, 0.0, 0.0), (4.0, 4.0), (1, 1, 1), 5.0)
    cam = sc.Camera.free((0.0, 0.0, 5.0), (0.0, 0.0, 10.0), 200.0)
    second = {}
    for a in (1.0, 0.5, 0.25):
        fog = sc.Fog(0.3, (a, a, a), 0.0, 0.0, 0.3)
        one = _render(open_sky([lamp], fog, cam), 4, 4, 16, 1)[0]
        zero = _render(open_sky([lamp], fog, cam), 4, 4, 16, 0)[0]
        second[a] = (one - zero) / (a * a)
    assert second[1.0].mean() > 1e-3  # the second scattering contributes
    for a in (0.5, 0.25):
        assert np.allclose(second[a], second[1.0], rtol=1e-3, atol=1e-6), a
