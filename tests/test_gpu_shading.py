"""GPU parity of the shading beyond the base material: suns, the flat ground, the colouring ramp
and named finishes (emission, Oren-Nayar), against the oracle.

Tolerances are those of test_gpu_pathtrace.py:
  * megakernel and wavefront forms against the oracle's FP32 radiance of the same random stream:
    per-pixel radiance within 1e-3 relative on >= 98 % of the pixels, the image mean within
    1e-3, depth within 1e-5 relative where both hit;
  * the wavefront form against the megakernel: the same sums to 1e-6 relative;
  * the flattened kernel of the server against the megakernel: bit for bit.
The scenes cover the full-sample path (colouring by iterations or the orbit mix, a graph's
MaterialTag finishes: the field is evaluated again at the hit) and the carried path (normal
colouring, trap-driven surface variation: the march's own trap).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    from syfrack_amd.render import Renderer
    return Renderer(0)


def _finishes():
    from syfrack_amd import scene as sc
    return [sc.Finish((0.55, 0.52, 0.48), 0.8, 0.0, 0.04, (0.0, 0.0, 0.0), True),   # concrete: Oren-Nayar
            sc.Finish((0.9, 0.7, 0.3), 0.25, 1.0, 0.04),                          # gold
            sc.Finish((0.1, 0.1, 0.1), 0.6, 0.0, 0.04, (0.6, 0.3, 0.1))]          # glowing


def shaded_scenes():
    from syfrack_amd import abi
    from syfrack_amd import scene as sc
    sky = sc.Light.environment((0.35, 0.4, 0.5), (0.25, 0.25, 0.28), (0.05, 0.05, 0.05), 1.0)
    sun = sc.Light.sun((0.4, 0.8, 0.3), (1.0, 0.95, 0.85), 3.0, 0.53)
    big_sun = sc.Light.sun((-0.5, 0.6, 0.4), (1.0, 0.9, 0.8), 2.0, 12.0)
    lamp = sc.Light.area((3.0, 5.0, -4.0), (-3.0, -5.0, 4.0), (2.0, 2.0), (1.0, 0.95, 0.9), 12.0)
    cyc = sc.Ramp([(0.9, 0.2, 0.1), (0.95, 0.8, 0.2), (0.2, 0.7, 0.3), (0.1, 0.3, 0.9), (0.6, 0.2, 0.8)],
                  [], True, 3.0, 0.1, 1.2, True)
    pos = sc.Ramp([(0.1, 0.1, 0.4), (0.8, 0.4, 0.1), (1.0, 0.95, 0.8)], [0.0, 0.35, 1.0])
    out = {}
    f = sc.Fractal.bundled("mandelbulb")
    f.set_param("fractal.nodes[1].iterations", 10)
    mat = sc.Material(albedo=(0.9, 0.9, 0.9), roughness=0.35, specular=0.04,
                      coloring=sc.Coloring(cyc, 0, ((1.0, 0.0, 0.0, 0.0), 0.5, 0.2, 2.0)), finishes=_finishes())
    out["bulb_orbit_sun_ground"] = sc.Scene(f, sc.Camera.orbit((0, 0, 0), 3.4, 30.0, 20.0, 35.0), [sky, sun], mat,
                                            seed=41, ground=sc.Ground(-1.2, 1))
    f = sc.Fractal.bundled("mandelbox")
    mat = sc.Material(albedo=(0.85, 0.85, 0.85), roughness=0.4, specular=0.04,
                      coloring=sc.Coloring(pos, abi.COLOR_ITER), finishes=_finishes())
    out["box_iter_sun_lamp_glow"] = sc.Scene(f, sc.Camera.orbit((0, 0, 0), 11.0, 120.0, 25.0, 35.0),
                                             [sun, lamp], mat, seed=42, ground=sc.Ground(-4.5, 3))
    f = sc.Fractal.bundled("menger")
    f.set_param("fractal.nodes[1].iterations", 5)
    mat = sc.Material(albedo=(0.8, 0.8, 0.8), roughness=0.5, specular=0.04, surface=(0, (0.2, 0.7), (0.8, 0.0)),
                      coloring=sc.Coloring(pos, abi.COLOR_NORMAL))
    out["menger_normal_bigsun"] = sc.Scene(f, sc.Camera.orbit((0, 0, 0), 4.2, 40.0, 30.0, 35.0), [sky, big_sun], mat,
                                           seed=43, ground=sc.Ground(-1.5, 0))
    return out


def tagged_graph():
    """A trap-coloured sphere tagged finish 1 (concrete) beside a box at x = 2.2 (Translate moves
    the point by its offset) tagged finish 2 (gold)."""
    from syfrack_amd.graph import Graph, Node
    nodes = [Node(0, "Input"), Node(1, "Trap", {"slot": 0, "measure": "Norm", "scale": 2.0}),
             Node(2, "Sphere", {"radius": 1.0}), Node(3, "MaterialTag", {"index": 1}),
             Node(4, "Translate", {"offset": (-2.2, 0.0, 0.0)}), Node(5, "Box", {"half": (0.7, 0.7, 0.7)}),
             Node(6, "MaterialTag", {"index": 2}), Node(7, "Union"), Node(8, "Output")]
    links = [((0, 0), (1, 0)), ((1, 0), (2, 0)), ((2, 0), (3, 0)), ((0, 0), (4, 0)), ((4, 0), (5, 0)),
             ((5, 0), (6, 0)), ((3, 0), (7, 0)), ((6, 0), (7, 1)), ((7, 0), (8, 0))]
    return Graph(nodes, links, 1.0, 3.5)


def _oracle(scene, cfg, graph=None):
    from oracle import oracle
    if graph is not None:
        oracle.graph_use(graph)
    return oracle.render(scene.desc(cfg), (0, 0, cfg.width, cfg.height), 0, cfg.samples_per_pixel)


@pytest.mark.parametrize("name", ["bulb_orbit_sun_ground", "box_iter_sun_lamp_glow", "menger_normal_bigsun"])
@pytest.mark.parametrize("form", ["megakernel", "wavefront"])
def test_shading_matches_oracle(renderer, name, form):
    from syfrack_amd import scene as sc
    from test_gpu_pathtrace import _against_oracle
    scene = shaded_scenes()[name]
    cfg = sc.RenderConfig(96, 54, 8, 4, 256)
    cfg.tile_size = 32
    out = renderer.render(scene, cfg, wavefront=(form == "wavefront"))
    acc, dep, aux, _ = _oracle(scene, cfg)
    assert acc[..., :3].mean() > 1e-3
    _against_oracle(f"{name} {form}", out, acc, dep, aux)


@pytest.mark.parametrize("name", ["bulb_orbit_sun_ground", "menger_normal_bigsun"])
def test_shaded_wavefront_equals_megakernel(renderer, name):
    from syfrack_amd import scene as sc
    scene = shaded_scenes()[name]
    cfg = sc.RenderConfig(80, 48, 4, 4, 256)
    mk = renderer.render(scene, cfg)
    wf = renderer.render(scene, cfg, wavefront=True)
    assert np.allclose(wf.accum, mk.accum, rtol=1e-6, atol=1e-7)
    assert np.array_equal(wf.depth, mk.depth)


@pytest.mark.parametrize("name", ["bulb_orbit_sun_ground", "box_iter_sun_lamp_glow"])
def test_shaded_scene_served_equals_megakernel(renderer, name):
    from syfrack_amd import abi, serve
    from syfrack_amd import scene as sc
    from syfrack_amd.render import Job
    scene = shaded_scenes()[name]
    cfg = sc.RenderConfig(96, 64, 4, 4, 256)
    ctx = renderer.ctx
    j = Job(ctx, scene.desc(cfg))
    try:
        j.render_tile(0, 0, cfg.width, cfg.height, 0, cfg.samples_per_pixel)
        j.sync()
        ref = [j.readback(p) for p in (abi.PLANE_ACCUM, abi.PLANE_DEPTH, abi.PLANE_AUX, abi.PLANE_DEPTH64)]
    finally:
        j.close()
    with serve.Server(ctx, "batched", quantum_samples=1, batch_target_us=1, grid_blocks=256) as s:
        rid = s.submit(scene, cfg)
        s.wait(rid, 120)
        got = [s.plane(rid, p) for p in (abi.PLANE_ACCUM, abi.PLANE_DEPTH, abi.PLANE_AUX, abi.PLANE_DEPTH64)]
        s.release(rid)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", ["specialised", "interpreted"])
@pytest.mark.parametrize("form", ["megakernel", "wavefront"])
def test_graph_finishes_match_oracle(renderer, mode, form):
    from syfrack_amd import scene as sc
    from test_gpu_pathtrace import _against_oracle
    g = tagged_graph()
    sky = sc.Light.environment((0.35, 0.4, 0.5), (0.25, 0.25, 0.28), (0.05, 0.05, 0.05), 1.0)
    sun = sc.Light.sun((0.4, 0.8, 0.3), (1.0, 0.95, 0.85), 3.0, 2.0)
    ramp = sc.Ramp([(0.9, 0.2, 0.1), (0.2, 0.7, 0.3), (0.1, 0.3, 0.9)])
    mat = sc.Material(albedo=(0.8, 0.8, 0.8), roughness=0.4, specular=0.04, coloring=sc.Coloring(ramp, 0),
                      finishes=_finishes())
    scene = sc.Scene(g, sc.Camera.orbit((1.1, 0, 0), 6.0, 20.0, 25.0, 35.0), [sky, sun], mat, seed=44,
                     ground=sc.Ground(-1.0, 3))
    cfg = sc.RenderConfig(96, 54, 8, 4, 256, graph_mode=mode)
    cfg.tile_size = 32
    out = renderer.render(scene, cfg, wavefront=(form == "wavefront"))
    acc, dep, aux, _ = _oracle(scene, cfg, g)
    _against_oracle(f"tagged graph {mode} {form}", out, acc, dep, aux)


def test_shading_validation(renderer):
    from syfrack_amd import abi
    from syfrack_amd import scene as sc
    from syfrack_amd.render import Job
    scene = shaded_scenes()["bulb_orbit_sun_ground"]
    cfg = sc.RenderConfig(16, 16, 1, 1, 64)

    def refused(mutate):
        d = scene.desc(cfg)
        mutate(d)
        with pytest.raises(abi.SyfError) as e:
            Job(renderer.ctx, d).close()
        assert e.value.code == abi.SYF_E_INVALID
        return str(e.value)

    def sun_size(d, v):
        d.lights[1].size[0] = v
    assert "angular diameter" in refused(lambda d: sun_size(d, 0.0))
    refused(lambda d: sun_size(d, 91.0))
    refused(lambda d: setattr(d.coloring, "slot", 7))
    refused(lambda d: setattr(d.coloring, "stop_count", abi.MAX_STOPS + 1))
    refused(lambda d: setattr(d.coloring, "gamma", 0.0))
    refused(lambda d: setattr(d, "finish_count", abi.MAX_FINISHES + 1))
    refused(lambda d: setattr(d.ground, "finish", abi.MAX_FINISHES + 1))
    refused(lambda d: setattr(d.ground, "height", float("nan")))

    def unordered(d):
        d.coloring.positioned = 1
        d.coloring.stops[1][3] = d.coloring.stops[0][3]
    assert "increase" in refused(unordered)

    def bad_finish(d):
        d.finishes[0].oren_nayar = 2
    refused(bad_finish)
    # an all-zero colouring is no colouring
    d = scene.desc(cfg)
    d.coloring = abi.SyfColoring()
    Job(renderer.ctx, d).close()


# ---------------------------------------------------------------- height fog

def fog_scenes():
    """The reference's exhibition-like atmosphere: a height fog under sun, sky and an emitter,
    over a Mandelbulb and over a Menger sponge on the ground."""
    from syfrack_amd import scene as sc
    sky = sc.Light.environment((0.35, 0.4, 0.5), (0.25, 0.25, 0.28), (0.05, 0.05, 0.05), 1.0)
    sun = sc.Light.sun((0.4, 0.8, 0.3), (1.0, 0.95, 0.85), 3.0, 0.53)
    lamp = sc.Light.area((2.0, 1.5, -2.0), (-2.0, -1.5, 2.0), (1.0, 1.0), (1.0, 0.9, 0.8), 15.0)
    out = {}
    f = sc.Fractal.bundled("mandelbulb")
    f.set_param("fractal.nodes[1].iterations", 8)
    out["bulb_fog_sun_lamp"] = sc.Scene(f, sc.Camera.orbit((0, 0, 0), 3.4, 30.0, 20.0, 35.0), [sky, sun, lamp],
                                        sc.Material(albedo=(0.8, 0.75, 0.7), roughness=0.4), seed=51,
                                        fog=sc.Fog(0.12, (0.85, 0.8, 0.78), -0.5, 0.4, 0.55))
    f = sc.Fractal.bundled("menger")
    f.set_param("fractal.nodes[1].iterations", 4)
    out["menger_fog_ground"] = sc.Scene(f, sc.Camera.orbit((0, 0, 0), 4.5, 40.0, 20.0, 35.0), [sky, sun, lamp],
                                        sc.Material(albedo=(0.8, 0.8, 0.8), roughness=0.5), seed=52,
                                        ground=sc.Ground(-1.2, 0), fog=sc.Fog(0.2, (0.82, 0.76, 0.68), -1.0, 0.35, 0.75))
    return out


@pytest.mark.parametrize("name", ["bulb_fog_sun_lamp", "menger_fog_ground"])
@pytest.mark.parametrize("form", ["megakernel", "wavefront"])
def test_fog_matches_oracle(renderer, name, form):
    """Radiance and the session planes (VOLUME, MOMENTS, SURFACE) of a fog scene against the
    oracle's, with the family tolerance (the fog's equiangular sampling calls atan2 and tan, whose
    device results differ from glibc's by an ulp now and then)."""
    from syfrack_amd import abi
    from syfrack_amd import scene as sc
    from oracle import oracle
    from test_gpu_pathtrace import _against_oracle
    scene = fog_scenes()[name]
    cfg = sc.RenderConfig(96, 54, 8, 4, 256)
    cfg.tile_size = 32
    job = renderer.job(scene, cfg)
    try:
        for s0 in range(0, cfg.samples_per_pixel, 4):
            for y in range(0, cfg.height, 32):
                for x in range(0, cfg.width, 32):
                    w, h = min(32, cfg.width - x), min(32, cfg.height - y)
                    if form == "wavefront":
                        job.render_wavefront(x, y, w, h, s0, 4)
                    else:
                        job.render_tile(x, y, w, h, s0, 4)
        job.sync()
        from syfrack_amd.render import RenderOutput
        out = RenderOutput(job.readback(abi.PLANE_ACCUM), job.readback(abi.PLANE_DEPTH), job.readback(abi.PLANE_AUX),
                           depth64=job.readback(abi.PLANE_DEPTH64))
        planes = [job.readback(p) for p in (abi.PLANE_VOLUME, abi.PLANE_MOMENTS, abi.PLANE_SURFACE)]
    finally:
        job.close()
    acc, dep, aux, _, sess = oracle.render(scene.desc(cfg), (0, 0, cfg.width, cfg.height), 0, cfg.samples_per_pixel,
                                           session=True)
    assert acc[..., :3].mean() > 1e-3
    _against_oracle(f"{name} {form}", out, acc, dep, aux)
    vol, mom, surf = planes
    assert np.array_equal(vol[..., 3], sess[0][..., 3]) and np.array_equal(surf[..., 3], sess[2][..., 3])
    assert np.allclose(vol[..., :3] + surf[..., :3], out.accum[..., :3], rtol=1e-4, atol=1e-5)
    for k, (g, o) in enumerate(((vol, sess[0]), (surf, sess[2]))):
        close = np.abs(g[..., :3] - o[..., :3]) <= 1e-3 * np.abs(o[..., :3]) + 1e-4
        frac = np.mean(np.all(close, axis=-1))
        print(f"{name} {form} session plane {k}: {frac:.4f} within 1e-3")
        assert frac >= 0.98
    assert np.all(mom[..., :3] >= 0)


def test_fog_wavefront_equals_megakernel(renderer):
    from syfrack_amd import abi
    from syfrack_amd import scene as sc
    scene = fog_scenes()["bulb_fog_sun_lamp"]
    cfg = sc.RenderConfig(64, 40, 4, 4, 256)
    mk = renderer.render(scene, cfg)
    wf = renderer.render(scene, cfg, wavefront=True)
    assert np.allclose(wf.accum, mk.accum, rtol=1e-6, atol=1e-7)
    assert np.array_equal(wf.depth, mk.depth)


def test_fog_refused_by_the_server_and_resumed_by_checkpoints(renderer, tmp_path):
    from syfrack_amd import abi, serve
    from syfrack_amd import scene as sc
    from syfrack_amd.render import Job
    scene = fog_scenes()["bulb_fog_sun_lamp"]
    cfg = sc.RenderConfig(32, 24, 2, 2, 128)
    with serve.Server(renderer.ctx, "batched", path_steps=8, grid_blocks=512, tile_width=8, quantum_samples=1) as s:
        with pytest.raises(abi.SyfError) as e:
            s.submit(scene, cfg)
        assert e.value.code == abi.SYF_E_INVALID and "fog" in str(e.value)
    # a checkpoint of a fog job resumes its session planes
    a = renderer.job(scene, cfg)
    try:
        a.render_tile(0, 0, cfg.width, cfg.height, 0, 2)
        a.sync()
        a.checkpoint_save(str(tmp_path / "ck"), "(samples_done: 2)")
        ref = [a.readback(p) for p in (abi.PLANE_ACCUM, abi.PLANE_VOLUME, abi.PLANE_MOMENTS, abi.PLANE_SURFACE)]
    finally:
        a.close()
    b = renderer.job(scene, cfg)
    try:
        b.checkpoint_load(str(tmp_path / "ck"))
        got = [b.readback(p) for p in (abi.PLANE_ACCUM, abi.PLANE_VOLUME, abi.PLANE_MOMENTS, abi.PLANE_SURFACE)]
    finally:
        b.close()
    for x, y in zip(ref, got):
        assert np.array_equal(x, y)
