"""GPU parity of the distance probe and the primary-ray trace against the oracle.

Runs on an MI355X through the C ABI (libsyfrack_hip.so).  Tolerances:
  * DE probe vs the reference's 64-point fixtures: distance 1e-5 relative (absolute below
    1), traps 1e-5, iterations exact (graph_parity.rs:34-65);
  * DE probe vs the FP32 oracle on 4096 LCG points (gpu_de.rs:40-91 corpus, seed 11):
    iterations exact, distance within 1e-5 relative;
  * primary trace vs the FP64 oracle: >= 99.9 % hit/miss agreement and refined hit
    distance within 1e-4 relative (per config, the observed error is printed).
"""
import numpy as np
import pytest

from helpers import lcg_points, load_de_fixture

pytestmark = pytest.mark.gpu

NAMES = ["sphere", "mandelbulb", "mandelbox", "menger"]


@pytest.fixture(scope="module")
def renderer():
    from syfrack_amd.render import Renderer
    return Renderer(0)


@pytest.mark.parametrize("name", NAMES)
def test_probe_matches_reference_fixtures(renderer, name):
    from syfrack_amd.scene import Camera, Scene, Fractal
    doc, p, d, traps, iters = load_de_fixture(name)
    s = renderer.probe_distances(Scene(Fractal.bundled(name), Camera()), p)
    err = np.abs(s.distance - d) / np.maximum(np.abs(d), 1.0)
    print(f"{name}: worst relative distance error {err.max():.3e}")
    assert err.max() <= 1e-5
    assert np.abs(s.traps - traps).max() <= 1e-5
    assert np.array_equal(s.iterations, iters)


@pytest.mark.parametrize("name,extent", [("mandelbulb", 1.6), ("mandelbox", 8.0), ("menger", 1.6), ("sphere", 3.0)])
def test_probe_matches_oracle_on_corpus(renderer, name, extent):
    from oracle import oracle
    from syfrack_amd.scene import Camera, Scene, Fractal
    f = Fractal.bundled(name)
    p = lcg_points(4096, extent, 11)
    s = renderer.probe_distances(Scene(f, Camera()), p)
    dist, traps, iters, _ = oracle.de_probe(f.c_struct(), p, 32)
    assert np.array_equal(s.iterations, iters)
    err = np.abs(s.distance - dist) / np.maximum(np.abs(dist), 1.0)
    assert err.max() <= 1e-5, err.max()
    assert np.abs(s.traps[:, 0] - traps[:, 0]).max() <= 1e-5


@pytest.mark.parametrize("power", [2, 3, 6, 8, 16, 6.25])
def test_probe_mandelbulb_powers(renderer, power):
    from oracle import oracle
    from syfrack_amd.scene import Camera, Scene, Fractal
    f = Fractal.bundled("mandelbulb")
    f.power = float(power)
    p = lcg_points(2048, 1.6, 11)
    p = np.concatenate([p, np.array([[0, 0, 1], [0, 0, -1], [1e-6, 0, 1], [0, 0, 0.5]], np.float32)])
    s = renderer.probe_distances(Scene(f, Camera()), p)
    dist, traps, iters, _ = oracle.de_probe(f.c_struct(), p, 32)
    if float(power).is_integer():
        # exact spherical product: IEEE arithmetic on both sides
        assert np.array_equal(s.iterations, iters)
        err = np.abs(s.distance - dist) / np.maximum(np.abs(dist), 1.0)
        assert err.max() <= 1e-5, err.max()
    else:
        # trigonometric form: device libm vs glibc differ by ulps that the orbit amplifies;
        # the reference's own GPU/CPU gate (gpu_de.rs:40-91): points with |d| < 0.02 skipped,
        # |delta d| <= 1e-3 + 2 % |d|, iterations equal, at most 0.5 % of points off
        keep = np.abs(dist) >= 0.02
        bad = (np.abs(s.distance - dist) > 1e-3 + 0.02 * np.abs(dist)) | (s.iterations != iters)
        assert np.mean(bad[keep]) <= 0.005, np.mean(bad[keep])


def _trace_vs_oracle(renderer, scene, cfg, rect_rows=None):
    from oracle import oracle
    out = renderer.trace_primary(scene, cfg)
    desc = scene.desc(cfg)
    if rect_rows is None:
        ref64 = oracle.trace_primary(desc, 0, 64)
        ref32 = oracle.trace_primary(desc, 0, 32)
        sl = slice(None)
    else:
        y0, h = rect_rows
        ref64 = oracle.trace_primary(desc, 0, 64, rect=(0, y0, cfg.width, h))
        ref32 = oracle.trace_primary(desc, 0, 32, rect=(0, y0, cfg.width, h))
        sl = slice(y0, y0 + h)
    return out, ref32, ref64, sl


def _angle(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    c = np.sum(a * b, axis=-1) / (np.linalg.norm(a, axis=-1) * np.linalg.norm(b, axis=-1))
    return np.arccos(np.clip(c, -1.0, 1.0))


def _check_normals(name, out, ref32, ref64, sl, exact=True):
    """AUX plane (normal, steps) against the oracle: where the device march took the FP32
    twin's steps the normals carry the same step count and agree to 1e-4 rad on >= 99.9 % of the
    pixels (at most 1e-3: a one-ulp libm difference in a tap's field value, divided by a tap
    spacing of a few 1e-4, moves a normal by up to a few 1e-4 rad); against the FP64 twin the
    angle is <= 1e-3 rad on >= 99.9 % of the pixels both hit."""
    g = out.normals[sl]
    hit_g = out.depth[sl] >= 0
    both32 = hit_g & (ref32["depth"] >= 0) & (out.steps[sl] == ref32["steps"])
    a32 = _angle(g[both32][:, :3], ref32["aux"][both32][:, :3])
    assert np.array_equal(g[both32][:, 3], ref32["aux"][both32][:, 3])
    both64 = hit_g & (ref64["depth"] >= 0)
    a64 = _angle(g[both64][:, :3], ref64["aux"][both64][:, :3])
    print(f"{name} normals: vs FP32 twin max {a32.max():.2e} rad over {both32.sum()}; vs FP64 twin "
          f"p99 {np.percentile(a64, 99):.2e} max {a64.max():.2e} rad over {both64.sum()}")
    if exact:  # the exact march takes the FP32 twin's steps; the fast one only nearly
        assert both32.sum() > 0.95 * both64.sum()
        assert np.mean(a32 <= 1e-4) >= 0.999 and a32.max() <= 1e-3
    assert np.mean(a64 <= 1e-3) >= 0.999


@pytest.mark.parametrize("math", ["fast", "exact"])
def test_mandelbulb_primary_320x180_matches_fp64_oracle(renderer, math):
    from syfrack_amd import scene as sc
    scene, cfg = sc.mandelbulb_primary(320, 180)
    cfg.march_math = math
    out, ref32, ref64, sl = _trace_vs_oracle(renderer, scene, cfg)
    hit_g, hit_o = out.depth >= 0, ref64["depth"] >= 0
    agree = np.mean(hit_g == hit_o)
    both = hit_g & hit_o
    rel = np.abs(out.depth64[both] - ref64["depth64"][both]) / ref64["depth64"][both]
    relm = np.abs(out.depth[both] - ref64["depth"][both]) / ref64["depth"][both]
    print(f"320x180: hit/miss agreement {agree:.5f}; refined distance rel err max {rel.max():.3e} "
          f"p99 {np.percentile(rel, 99):.3e}; march distance rel err max {relm.max():.3e}")
    assert agree >= 0.999
    assert np.mean(rel <= 1e-4) >= 0.999
    same = np.mean(out.steps == ref32["steps"])
    print(f"{math} march: {same:.4f} of the pixels take the FP32 twin's step count")
    if math == "exact":  # the exact march follows the FP32 restatement step for step
        assert same >= 0.99
    hits = out.depth >= 0
    n = out.normals[hits][:, :3]
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-4)
    assert np.all(out.normals[~hits] == np.array([0, 0, 0, -1], np.float32))
    assert out.stats["primary_rays"] == 0 or out.stats["primary_rays"] == 320 * 180
    _check_normals(f"mandelbulb 320x180 {math}", out, ref32, ref64, sl, exact=(math == "exact"))


@pytest.mark.parametrize("math", ["fast", "exact"])
@pytest.mark.parametrize("name", ["mandelbox", "menger"])
def test_other_families_primary_match_fp64_oracle(renderer, name, math):
    from syfrack_amd import scene as sc
    f = sc.Fractal.bundled(name)
    if name == "mandelbox":
        cam = sc.Camera.orbit((0, 0, 0), 11.0, 28.0, 22.0, 35.0)
    else:
        cam = sc.Camera.orbit((0, 0, 0), 3.6, 40.0, 30.0, 40.0)
    scene = sc.Scene(f, cam, seed=5)
    cfg = sc.RenderConfig(160, 90, 1, 0, 512, march_math=math)
    out, ref32, ref64, sl = _trace_vs_oracle(renderer, scene, cfg)
    hit_g, hit_o = out.depth >= 0, ref64["depth"] >= 0
    agree = np.mean(hit_g == hit_o)
    both = hit_g & hit_o
    rel = np.abs(out.depth64[both] - ref64["depth64"][both]) / ref64["depth64"][both]
    print(f"{name}: hits {both.sum()}, agreement {agree:.5f}, refined rel err max {rel.max():.3e}")
    assert both.sum() > 100
    assert agree >= 0.999
    assert np.mean(rel <= 1e-4) >= 0.999
    _check_normals(f"{name} {math}", out, ref32, ref64, sl, exact=(math == "exact"))


def test_mandelbulb_1080p_properties_and_centre_rows(renderer):
    """Full 1920x1080: every ray accounted for, unit normals, steps >= 1, misses flagged,
    and the centre rows against the FP64 oracle."""
    from syfrack_amd import scene as sc
    scene, cfg = sc.mandelbulb_primary(1920, 1080)
    out, ref32, ref64, sl = _trace_vs_oracle(renderer, scene, cfg, rect_rows=(536, 8))
    assert np.all(out.steps >= 1) and np.all(out.steps <= 512)
    hits = out.depth >= 0
    assert 0.2 < hits.mean() < 0.6
    assert np.allclose(np.linalg.norm(out.normals[hits][:, :3], axis=1), 1.0, atol=1e-4)
    assert np.all(out.depth64[~hits] == -1.0)
    g = out.depth[sl] >= 0
    o = ref64["depth"] >= 0
    assert np.mean(g == o) >= 0.999
    both = g & o
    rel = np.abs(out.depth64[sl][both] - ref64["depth64"][both]) / ref64["depth64"][both]
    assert np.mean(rel <= 1e-4) >= 0.999
    _check_normals("mandelbulb 1080p centre rows (fast)", out, ref32, ref64, sl, exact=False)
    st = out.stats
    assert st["hits"] == int(hits.sum())
    assert st["steps"] == int(out.steps.astype(np.int64).sum())
    assert st["de_evals"] == st["steps"] + 4 * st["hits"]
    # every hit was shaded once, by the fused kernel or the trailing pass: its normal carries
    # its step count and its FP64 depth is set
    assert 0 <= st["tail_hits"] <= st["hits"]
    assert st["tail_hit_de_evals"] == 4 * st["tail_hits"]
    assert np.all(out.normals[hits][:, 3] == out.steps[hits].astype(np.float32))
    assert np.all(out.depth64[hits] > 0.0)
    print(f"1080p: {st['hits']} hits, {st['tail_hits']} shaded by the trailing pass")


def test_primary_trace_is_deterministic(renderer):
    """Which wave shades which hit varies from run to run; the planes do not."""
    from syfrack_amd import scene as sc
    scene, cfg = sc.mandelbulb_primary(1920, 1080)
    a = renderer.trace_primary(scene, cfg)
    b = renderer.trace_primary(scene, cfg)
    for x, y in ((a.depth, b.depth), (a.depth64, b.depth64), (a.normals, b.normals), (a.steps, b.steps)):
        assert np.array_equal(x, y)


def test_repeated_traces_on_one_job_reuse_the_hit_queue(renderer):
    """A job's hit queue is cleared once; each trace's trailing pass empties the entries it used
    and done flags count only for their own trace.  Traces of several samples on one job, then
    the first sample again, must give the planes of a fresh job, with no queue error."""
    from syfrack_amd import abi, scene as sc
    scene, cfg = sc.mandelbulb_primary(640, 360)
    job = renderer.job(scene, cfg)
    try:
        planes = (abi.PLANE_DEPTH, abi.PLANE_DEPTH64, abi.PLANE_AUX, abi.PLANE_STEPS)
        runs = []
        for sample in (0, 1, 2, 0):
            job.trace_primary(sample)
            job.sync()
            st = job.stats()
            assert st["hits"] > 0
            runs.append([job.readback(p).copy() for p in planes])
        fresh = renderer.trace_primary(scene, cfg, sample=1)
    finally:
        job.close()
    for x, y in zip(runs[0], runs[3]):
        assert np.array_equal(x, y)
    assert not np.array_equal(runs[0][0], runs[1][0])  # another sample jitters the rays
    for x, y in zip(runs[1], (fresh.depth, fresh.depth64, fresh.normals, fresh.steps)):
        assert np.array_equal(x.reshape(y.shape), y)


def test_sphere_scene_centre_depth_and_tolerance(renderer):
    """gpu_render.rs:479-490 and :113-145 on the device: a unit sphere seen from distance 3
    has centre depth 2 +- 0.05, and a finer pixel tolerance lowers the hit-radius error."""
    from syfrack_amd import scene as sc
    f = sc.Fractal.bundled("sphere")
    cam = sc.Camera.free((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 50.0)
    scene = sc.Scene(f, cam, seed=3)
    errs = []
    for tol in (1.0, 0.25):
        cfg = sc.RenderConfig(64, 64, 1, 0, 512, pixel_tolerance=tol)
        out = renderer.trace_primary(scene, cfg)
        assert abs(out.depth[32, 32] - 2.0) < 0.05
        t = out.depth[20:44, 20:44].astype(np.float64)
        nz = out.normals[20:44, 20:44, 2].astype(np.float64)
        assert np.all(t > 0) and np.all(nz > 0)
        radius = 3.0 * nz - np.sqrt(np.maximum(t * t - 9.0 * (1.0 - nz * nz), 0.0))
        errs.append(np.mean(np.abs(radius - 1.0)))
        # the FP64 refine lands on the hit rule's crossing |p| - 1 = cone * t; head-on that
        # is t = 2 / (1 + cone)
        cone = float(np.float32(np.float32(np.float32(18.0) / np.float32(50.0)) * np.float32(2.0)
                                / np.float32(64)) * np.float32(tol))
        assert abs(out.depth64[32, 32] - 2.0 / (1.0 + cone)) < 2e-4
    assert errs[0] > 1e-5 and errs[1] < errs[0] * 0.7


def test_invalid_config_names_the_field(renderer):
    from syfrack_amd import abi, scene as sc
    scene, cfg = sc.mandelbulb_primary(32, 32)
    cfg.pixel_tolerance = 2.0
    with pytest.raises(abi.SyfError) as e:
        renderer.trace_primary(scene, cfg)
    assert e.value.code == abi.SYF_E_INVALID and "pixel_tolerance" in e.value.message
