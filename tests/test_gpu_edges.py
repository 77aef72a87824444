"""Edge cases of the HIP path: ragged and tiny images, rays that run out of steps, empty
inputs, invalid sizes, and served requests whose tiles overhang the image.

Tolerances:
  * primary trace on ragged sizes against the FP64 oracle: hit/miss agrees on all but
    n // 1000 + 1 pixels, the refined distance within 1e-4 relative where both hit; the planes
    of every ray are written (steps >= 1, misses flagged, FP64 depth -1 on misses);
  * rays cut at max_steps (march_segment's best-candidate rule, pathtrace.wgsl:137-143) against
    the FP32 oracle, which runs the same FP32 march: hit/miss and step counts agree on >= 99.9 %
    of the pixels, best-candidate distances within 1e-6 relative (the device's and the host's
    logf may differ by an ulp); a best-candidate hit has no refine bracket, so its FP64 depth is
    its march t exactly;
  * FP64 escape counts on a ragged planar image: bit-exact against the oracle's FP64 loop;
  * a request served in batched launches equals the plain megakernel's planes bit for bit on
    an image whose 8x8 tiles overhang both edges.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    from syfrack_amd.render import Renderer
    return Renderer(0)


@pytest.mark.parametrize("width,height", [(1, 1), (7, 3), (33, 17), (131, 67)])
def test_primary_ragged_sizes_match_fp64_oracle(renderer, width, height):
    from oracle import oracle
    from syfrack_amd import scene as sc
    scene, cfg = sc.mandelbulb_primary(width, height)
    print(f"primary trace {width}x{height}", flush=True)
    out = renderer.trace_primary(scene, cfg)
    ref = oracle.trace_primary(scene.desc(cfg), 0, 64)
    n = width * height
    hit_g = out.depth >= 0
    hit_o = ref["depth"] >= 0
    assert int(np.sum(hit_g != hit_o)) <= n // 1000 + 1
    both = hit_g & hit_o
    if both.any():
        rel = np.abs(out.depth64[both] - ref["depth64"][both]) / ref["depth64"][both]
        assert np.all(rel <= 1e-4)
    assert np.all(out.steps >= 1) and np.all(out.steps <= cfg.max_steps)
    assert np.all(out.normals[~hit_g] == np.array([0, 0, 0, -1], np.float32))
    assert np.all(out.depth64[~hit_g] == -1.0)
    assert np.all(out.depth64[hit_g] > 0.0)
    st = out.stats
    assert st["primary_rays"] == n and st["hits"] == int(hit_g.sum())
    assert st["de_evals"] == st["steps"] + 4 * st["hits"]
    assert st["tail_hits"] <= st["hits"]


def test_rays_cut_at_max_steps_take_the_best_candidate(renderer):
    from oracle import oracle
    from syfrack_amd import scene as sc
    scene, cfg = sc.mandelbulb_primary(160, 90)
    cfg.max_steps = 6
    out = renderer.trace_primary(scene, cfg)
    ref = oracle.trace_primary(scene.desc(cfg), 0, 32)
    hit_g = out.depth >= 0
    hit_o = ref["depth"] >= 0
    assert np.mean(hit_g == hit_o) >= 0.999
    assert np.mean(out.steps == ref["steps"]) >= 0.999
    cut = out.steps == cfg.max_steps
    # the oracle's best-candidate hits: cut at max_steps, and its FP64 depth is its march t (a
    # hit by the ordinary rule on the last step has a bracket and is refined)
    cand = cut & hit_g & hit_o & (out.steps == ref["steps"]) & (ref["depth64"] == ref["depth"])
    print(f"max_steps 6: {int(cut.sum())} rays cut, {int(cand.sum())} best-candidate hits", flush=True)
    assert cut.mean() > 0.2 and cand.sum() > 0
    # a best-candidate hit keeps its march t (no bracket to refine)
    assert np.all(out.depth64[cand] == out.depth[cand].astype(np.float64))
    assert np.allclose(out.depth[cand], ref["depth"][cand], rtol=1e-6, atol=0.0)


def test_probe_of_no_points_and_zero_sized_jobs(renderer):
    from syfrack_amd import abi, scene as sc
    scene, cfg = sc.mandelbulb_primary(8, 8)
    s = renderer.probe_distances(scene, np.zeros((0, 3), np.float32))
    assert s.distance.shape == (0,)
    for w, h, field in ((0, 8, "width"), (8, 0, "height")):
        bad = sc.RenderConfig(w, h, 1, 0, 512)
        with pytest.raises(abi.SyfError) as e:
            renderer.trace_primary(scene, bad)
        assert e.value.code == abi.SYF_E_INVALID and field in e.value.message


def test_planar_fp64_ragged_image_is_bit_exact():
    from oracle import oracle
    from syfrack_amd import abi
    from syfrack_amd.planar import Planar, PlanarView
    from syfrack_amd.render import Context
    v = PlanarView(family=abi.PLANAR_MANDELBROT, mode=abi.PLANAR_FP64, width=37, height=23, iterations=2000,
                   center=(-0.75, 0.1), span=2.5, bailout=16.0)
    p = Planar(Context(0), v)
    try:
        p.run()
        p.sync()
        it, _ = p.readback()
    finally:
        p.close()
    ref_it, _ = oracle.escape_f64(v.c_struct())
    assert it.shape == (23, 37)
    assert np.array_equal(it, ref_it)


def test_served_request_with_overhanging_tiles_equals_megakernel(renderer):
    from syfrack_amd import abi, serve
    from syfrack_amd.scene import RenderConfig
    q = serve.uneven_requests(1, seed=17)[0]
    q.cfg = RenderConfig(37, 21, samples_per_pixel=5, max_bounces=2, max_steps=256)
    planes = {}
    for mode, kw in (("plain", {}), ("batched", {"quantum_samples": 1, "batch_target_us": 1, "grid_blocks": 64})):
        with serve.Server(renderer.ctx, mode, **kw) as s:
            rid = s.submit(q.scene, q.cfg)
            s.wait(rid, 120)
            out = s.result(rid)
            planes[mode] = (out.accum, out.depth, out.normals, s.plane(rid, abi.PLANE_DEPTH64))
            s.release(rid)
    assert np.all(planes["plain"][0][..., 3] == 5)
    for a, b in zip(planes["plain"], planes["batched"]):
        assert np.array_equal(a, b)
