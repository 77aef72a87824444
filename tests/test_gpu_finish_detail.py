"""GPU parity of the finish detail patterns (finish_detail.wgsl through make_surface) against the
oracle: every pattern on the tagged sphere, box and ground, with bump, roughness delta and tint,
and two of the reference's scene files that use them (3d-01: brushed metal and the concrete
finishes on a ground; 3d-volumes-01: brushed metal), in the megakernel and wavefront forms.  Tolerances are those of test_gpu_pathtrace.py.
"""
import os

import pytest

pytestmark = pytest.mark.gpu

SCENES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")
_ORACLE = {}


@pytest.fixture(scope="module")
def renderer():
    from syfrack_amd.render import Renderer
    return Renderer(0)


def _pattern_scene(patterns):
    from syfrack_amd import scene as sc
    from test_gpu_shading import tagged_graph
    tints = [(0.6, 0.3, 0.2), (0.2, 0.5, 0.3), (0.25, 0.3, 0.6)]
    finishes = [sc.Finish((0.7, 0.68, 0.62), 0.45, 0.3 if i == 1 else 0.0, 0.04, pattern=k, seed=11 + i,
                          scale=(3.0, 1.5, 0.8)[i], strength=0.6, bump=(0.05, -0.03, 0.08)[i], roughness_delta=0.2,
                          tint=tints[i]) for i, k in enumerate(patterns)]
    sky = sc.Light.environment((0.35, 0.4, 0.5), (0.25, 0.25, 0.28), (0.05, 0.05, 0.05), 1.0)
    sun = sc.Light.sun((0.4, 0.8, 0.3), (1.0, 0.95, 0.85), 3.0, 2.0)
    mat = sc.Material(albedo=(0.8, 0.8, 0.8), roughness=0.4, specular=0.04, finishes=finishes)
    return sc.Scene(tagged_graph(), sc.Camera.orbit((1.1, 0, 0), 6.0, 20.0, 25.0, 35.0), [sky, sun], mat, seed=61,
                    ground=sc.Ground(-1.0, 3))


def _check(renderer, scene, cfg, form, key, label):
    from oracle import oracle
    from test_gpu_pathtrace import _against_oracle
    print(f"{label}, {form}, {cfg.width}x{cfg.height}, {cfg.samples_per_pixel} spp", flush=True)
    out = renderer.render(scene, cfg, wavefront=(form == "wavefront"))
    if key not in _ORACLE:
        oracle.graph_use(scene.fractal)
        _ORACLE[key] = oracle.render(scene.desc(cfg), (0, 0, cfg.width, cfg.height), 0, cfg.samples_per_pixel)
    acc, dep, aux, _ = _ORACLE[key]
    assert acc[..., :3].mean() > 1e-2
    _against_oracle(f"{label} {form}", out, acc, dep, aux)


@pytest.mark.parametrize("patterns", [(1, 2, 3), (4, 5, 6), (7, 8, 9), (10, 11, 12)])
@pytest.mark.parametrize("form", ["megakernel", "wavefront"])
def test_patterns_match_oracle(renderer, patterns, form):
    from syfrack_amd import scene as sc
    cfg = sc.RenderConfig(96, 54, 8, 4, 256)
    cfg.tile_size = 32
    _check(renderer, _pattern_scene(patterns), cfg, form, ("patterns", patterns), f"patterns {patterns}")


@pytest.mark.parametrize("name, size, spp", [("3d-01", (96, 54), 8), ("3d-volumes-01", (96, 54), 8)])
@pytest.mark.parametrize("form", ["megakernel", "wavefront"])
def test_reference_pattern_scenes_match_oracle(renderer, name, size, spp, form):
    from syfrack_amd import scene as sc
    from syfrack_amd.scene_file import load_scene
    scene = load_scene(os.path.join(SCENES, name + ".ron"))
    cfg = sc.RenderConfig(size[0], size[1], spp, 4, 256)
    cfg.tile_size = 32
    _check(renderer, scene, cfg, form, ("scene", name), f"scene file {name}")


def test_detail_validation_and_server_refusal(renderer):
    from syfrack_amd import abi, serve
    from syfrack_amd import scene as sc
    scene = _pattern_scene((1, 2, 3))
    cfg = sc.RenderConfig(32, 18, 1, 2, 128)
    for field, value, text in [("pattern", 13, "pattern must be 0..12"), ("scale", 0.0, "detail.scale"),
                               ("strength", 1.5, "detail.strength"), ("bump", -2.0, "detail.bump"),
                               ("roughness_delta", 3.0, "detail.roughness_delta"), ("tint", (0.5, 2.0, 0.5), "tint")]:
        bad = _pattern_scene((1, 2, 3))
        setattr(bad.material.finishes[1], field, value)
        with pytest.raises(RuntimeError, match=text):
            renderer.job(bad, cfg).close()
    classic = sc.Scene(sc.Fractal.bundled("mandelbulb"), scene.camera, scene.lights, scene.material, seed=5,
                       ground=sc.Ground(-1.2, 1))
    with serve.Server(renderer.ctx, "batched", path_steps=8, grid_blocks=512, tile_width=8, quantum_samples=1) as s:
        with pytest.raises(abi.SyfError) as e:
            s.submit(classic, cfg)
        assert e.value.code == abi.SYF_E_INVALID and "finish detail" in str(e.value)
