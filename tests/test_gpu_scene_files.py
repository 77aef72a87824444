"""GPU parity on the reference's own scene files: exhibition, forme-exposition and orfevrerie
(tests/golden/scenes, copied from the reference's scenes/), read by syfrack_amd.scene_file and
rendered in the megakernel and wavefront forms against the oracle.

Each combines a Kelvin-coloured sun, area emitters, a sky gradient and the height fog over a form
graph with a trap colouring ramp, a thin lens (exhibition) and cavity darkness.  Tolerances are
those of test_gpu_pathtrace.py: per-pixel radiance within 1e-3 relative on >= 98 % of the pixels,
the image mean within 1e-3, depth within 1e-5 relative where both hit.
"""
import os

import pytest

pytestmark = pytest.mark.gpu

SCENES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


@pytest.fixture(scope="module")
def renderer():
    from syfrack_amd.render import Renderer
    return Renderer(0)


@pytest.mark.parametrize("name", ["exhibition", "forme-exposition", "orfevrerie"])
@pytest.mark.parametrize("form", ["megakernel", "wavefront"])
def test_reference_scene_matches_oracle(renderer, name, form):
    from syfrack_amd import scene as sc
    from syfrack_amd.scene_file import load_scene
    from oracle import oracle
    from test_gpu_pathtrace import _against_oracle
    scene = load_scene(os.path.join(SCENES, name + ".ron"))
    cfg = sc.RenderConfig(96, 54, 8, 4, 256)
    cfg.tile_size = 32
    print(f"scene file {name}, {form}, 96x54, 8 spp", flush=True)
    out = renderer.render(scene, cfg, wavefront=(form == "wavefront"))
    oracle.graph_use(scene.fractal)
    acc, dep, aux, _ = oracle.render(scene.desc(cfg), (0, 0, cfg.width, cfg.height), 0, cfg.samples_per_pixel)
    assert acc[..., :3].mean() > 1e-2
    _against_oracle(f"{name} {form}", out, acc, dep, aux)
