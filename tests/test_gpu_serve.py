"""The render server on the device: continuous batching must not change any request's image.

Tolerances:
  * a request's accumulation, depth, normal and FP64-depth planes are bit-identical whether it
    runs alone (plain mode: one launch per request), in a mix of uneven requests (batched mode,
    several launches each holding work of several requests), or through the job API
    (syf_render_tile in 64-pixel tiles, 8 samples per dispatch): every form runs the same
    per-pixel arithmetic and adds the samples in ascending order;
  * the served FP64-refined first-hit distance agrees with the FP64 oracle (its FP64 march and
    refine) within 1e-4 relative on >= 99.9 % of the pixels both hit, and hit/miss agrees on
    >= 99.9 % of the pixels;
  * admission: a request above the budget is refused with SYF_E_ADMISSION; requests that
    together exceed it wait until earlier ones are released, and the peak stays within it.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from syfrack_amd.render import Context
    return Context(0)


def _planes(out):
    return out.accum, out.depth, out.normals


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_result_same_alone_and_in_the_mix(ctx):
    from syfrack_amd import abi, serve
    from syfrack_amd.render import Job
    from syfrack_amd.scene import RenderConfig
    reqs = serve.uneven_requests(10, seed=9, min_side=48, max_side=200)
    # one request too large for a launch: it runs one sample per launch, 8 launches, beside
    # the others (which each fit in a launch and go in one piece)
    reqs[4].cfg = RenderConfig(512, 512, samples_per_pixel=8, max_bounces=1, max_steps=256)
    print("requests:", [(q.cfg.width, q.spp, q.cfg.max_bounces, q.scene.fractal.family) for q in reqs], flush=True)
    alone = []
    with serve.Server(ctx, "plain") as s:
        for q in reqs:
            rid = s.submit(q.scene, q.cfg)
            s.wait(rid, 120)
            alone.append((_planes(s.result(rid)), s.plane(rid, abi.PLANE_DEPTH64)))
            s.release(rid)
        assert s.stats()["launches"] == len(reqs)
    # batched: one sample per item and a tiny launch target, so the large request is split
    # over many launches that also hold other requests' work; the second half arrives later
    with serve.Server(ctx, "batched", quantum_samples=1, batch_target_us=1, grid_blocks=256) as s:
        ids = [s.submit(q.scene, q.cfg) for q in reqs[:5]]
        time.sleep(0.02)
        ids += [s.submit(q.scene, q.cfg) for q in reqs[5:]]
        s.wait_all(300)
        st = s.stats()
        print("batched stats:", st, flush=True)
        assert st["completed"] == len(reqs) and st["launches"] >= 8
        infos = [s.poll(r) for r in ids]
        assert all(i.state == "done" and i.samples_done == q.spp for i, q in zip(infos, reqs))
        overlaps = sum(1 for i in range(len(infos)) for j in range(i + 1, len(infos))
                       if infos[i].start_ms < infos[j].done_ms and infos[j].start_ms < infos[i].done_ms)
        assert overlaps > 0
        for k, (q, rid) in enumerate(zip(reqs, ids)):
            got = _planes(s.result(rid))
            assert _same(got, alone[k][0]), k
            assert np.array_equal(s.plane(rid, abi.PLANE_DEPTH64), alone[k][1]), k
            s.release(rid)
        assert s.stats()["hbm_in_use"] == 0
    # the job API, tile by tile
    for k in (0, 4, 7):
        q = reqs[k]
        j = Job(ctx, q.scene.desc(q.cfg))
        try:
            for s0 in range(0, q.spp, 8):
                n = min(8, q.spp - s0)
                for y in range(0, q.cfg.height, 64):
                    for x in range(0, q.cfg.width, 64):
                        j.render_tile(x, y, min(64, q.cfg.width - x), min(64, q.cfg.height - y), s0, n)
            j.sync()
            got = (j.readback(abi.PLANE_ACCUM), j.readback(abi.PLANE_DEPTH), j.readback(abi.PLANE_AUX))
            assert _same(got, alone[k][0]), k
            assert np.array_equal(j.readback(abi.PLANE_DEPTH64), alone[k][1]), k
        finally:
            j.close()


def test_served_depth_matches_fp64_oracle(ctx):
    from oracle import oracle
    from syfrack_amd import abi, serve
    reqs = serve.uneven_requests(3, seed=21, min_side=96, max_side=97)
    with serve.Server(ctx, "batched") as s:
        ids = [s.submit(q.scene, q.cfg) for q in reqs]
        s.wait_all(120)
        for q, rid in zip(reqs, ids):
            d64 = s.plane(rid, abi.PLANE_DEPTH64)
            ref = oracle.trace_primary(q.scene.desc(q.cfg), 0, 64)
            hit_g, hit_o = d64 >= 0, ref["depth64"] >= 0
            both = hit_g & hit_o
            rel = np.abs(d64[both] - ref["depth64"][both]) / ref["depth64"][both]
            print(f"family {q.scene.fractal.family}: hits {both.sum()}, agreement {np.mean(hit_g == hit_o):.5f}, "
                  f"rel max {rel.max() if both.any() else 0:.3e}", flush=True)
            assert np.mean(hit_g == hit_o) >= 0.999
            if both.any():
                assert np.mean(rel <= 1e-4) >= 0.999


def test_quiet_mode_launches_equal_plain(ctx):
    """Once submissions stop (200 ms), the batched server plans everything left in one launch
    with whole-request items; every image still equals its plain render bit for bit."""
    from syfrack_amd import abi, serve
    reqs = serve.uneven_requests(8, seed=13, min_side=64, max_side=256)
    alone = []
    with serve.Server(ctx, "plain") as s:
        for q in reqs:
            rid = s.submit(q.scene, q.cfg)
            s.wait(rid, 120)
            alone.append((_planes(s.result(rid)), s.plane(rid, abi.PLANE_DEPTH64)))
            s.release(rid)
    for quiet in (abi.QUIET_UNBOUNDED, 1):  # everything at once; the smallest quiet launches
        with serve.Server(ctx, "batched", quiet_target_us=quiet) as s:
            ids = [s.submit(q.scene, q.cfg) for q in reqs]
            s.wait_all(300)
            st = s.stats()
            print(f"quiet_target_us {quiet}: {st['launches']} launches", flush=True)
            for k, rid in enumerate(ids):
                assert _same(_planes(s.result(rid)), alone[k][0]), (quiet, k)
                assert np.array_equal(s.plane(rid, abi.PLANE_DEPTH64), alone[k][1]), (quiet, k)
                s.release(rid)


def test_admission_against_the_budget(ctx):
    from syfrack_amd import abi, serve
    from syfrack_amd.scene import RenderConfig
    q = serve.uneven_requests(1, seed=4)[0]
    cfg = RenderConfig(128, 128, samples_per_pixel=2, max_bounces=1, max_steps=128)
    fp = serve.footprint(cfg)
    with serve.Server(ctx, "batched", hbm_budget_bytes=int(fp * 1.5)) as s:
        with pytest.raises(abi.SyfError) as e:
            s.submit(q.scene, RenderConfig(256, 256, 1, 1, 128))
        assert e.value.code == abi.SYF_E_ADMISSION
        a = s.submit(q.scene, cfg)
        b = s.submit(q.scene, cfg)
        s.wait(a, 60)
        time.sleep(0.05)
        assert s.poll(b).state == "queued"  # a still holds its planes
        with pytest.raises(abi.SyfError) as e:
            s.wait(b, 0.05)
        assert e.value.code == abi.SYF_E_TIMEOUT
        s.release(a)
        s.wait(b, 60)
        st = s.stats()
        assert st["hbm_peak"] <= int(fp * 1.5) and st["completed"] == 2
        s.release(b)


def test_cancel_release_and_unknown_ids(ctx):
    from syfrack_amd import abi, serve
    from syfrack_amd.scene import RenderConfig
    q = serve.uneven_requests(1, seed=6)[0]
    with serve.Server(ctx, "batched", quantum_samples=1, batch_target_us=500) as s:
        big = s.submit(q.scene, RenderConfig(512, 512, 64, 2, 256))
        small = s.submit(q.scene, RenderConfig(64, 64, 1, 1, 128))
        s.cancel(big)
        assert s.poll(big).state == "cancelled"
        s.wait(small, 60)
        s.wait_all(60)
        with pytest.raises(abi.SyfError) as e:
            s.plane(big, abi.PLANE_ACCUM)
        assert e.value.code == abi.SYF_E_INVALID
        s.release(big)
        s.release(small)
        for fn in (lambda: s.poll(big), lambda: s.release(small), lambda: s.cancel(12345)):
            with pytest.raises(abi.SyfError) as e:
                fn()
            assert e.value.code == abi.SYF_E_NOT_FOUND
        assert s.stats()["hbm_in_use"] == 0


def test_area_light_scene_served_equals_megakernel(ctx):
    """The Mandelbox path-trace scene (an area light: light sampling, shadow rays an emitter
    decides, emitter hits) served by the flattened kernel, split over several launches, gives
    the megakernel's planes bit for bit."""
    from syfrack_amd import abi, serve
    from syfrack_amd import scene as sc
    from syfrack_amd.render import Job
    scene, cfg = sc.mandelbox_pathtrace(160, 96, 6)
    j = Job(ctx, scene.desc(cfg))
    try:
        j.render_tile(0, 0, cfg.width, cfg.height, 0, cfg.samples_per_pixel)
        j.sync()
        ref = [j.readback(p) for p in (abi.PLANE_ACCUM, abi.PLANE_DEPTH, abi.PLANE_AUX, abi.PLANE_DEPTH64)]
    finally:
        j.close()
    assert ref[0][..., :3].mean() > 1e-3  # the light reaches the visible surface
    with serve.Server(ctx, "batched", quantum_samples=1, batch_target_us=1, grid_blocks=256) as s:
        rid = s.submit(scene, cfg)
        s.wait(rid, 120)
        got = [s.plane(rid, p) for p in (abi.PLANE_ACCUM, abi.PLANE_DEPTH, abi.PLANE_AUX, abi.PLANE_DEPTH64)]
        assert s.stats()["launches"] >= cfg.samples_per_pixel
        s.release(rid)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_poisson_serving_mix_scaled(ctx):
    """The serving loop (Poisson arrivals, 70 % previews of 256^2 at 1 spp, 30 % progressive
    renders), scaled to renders of 448^2-512^2 at 32-64 spp: every served result equals the same
    request rendered alone bit for bit; previews' refined depth agrees with the FP64 oracle; and
    a preview finishes before every progressive render that arrived earlier and still had work
    for more than 300 ms when the preview arrived (a render in its last launches may end first:
    up to three launches of about 30 ms are in flight)."""
    from oracle import oracle
    from syfrack_amd import abi, serve
    reqs = serve.serving_mix(30, rate_per_s=100.0, seed=17, big_sides=(448, 512), big_spp=(32, 64))
    assert 5 <= sum(not q.preview for q in reqs) <= 25
    print("stream:", [("P" if q.preview else f"R{q.cfg.width}x{q.spp}", round(q.arrival_s, 3)) for q in reqs],
          flush=True)
    planes = (abi.PLANE_ACCUM, abi.PLANE_DEPTH, abi.PLANE_AUX, abi.PLANE_DEPTH64)
    with serve.Server(ctx, "batched") as s:
        t0 = time.perf_counter()
        ids = []
        for q in reqs:
            dt = q.arrival_s - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(s.submit(q.scene, q.cfg))
        s.wait_all(300)
        infos = [s.poll(i) for i in ids]
        served = [[s.plane(i, p) for p in planes] for i in ids]
        for i in ids:
            s.release(i)
    assert all(i.state == "done" for i in infos)
    with serve.Server(ctx, "plain") as s:
        for k, q in enumerate(reqs):
            rid = s.submit(q.scene, q.cfg)
            s.wait(rid, 300)
            alone = [s.plane(rid, p) for p in planes]
            s.release(rid)
            assert all(np.array_equal(a, b) for a, b in zip(served[k], alone)), k
    # previews against the FP64 oracle
    for k in [k for k, q in enumerate(reqs) if q.preview][:3]:
        q = reqs[k]
        twin = oracle.trace_primary(q.scene.desc(q.cfg), 0, 64)["depth64"]
        d64 = served[k][3]
        hg, ho = d64 >= 0, twin >= 0
        assert np.mean(hg == ho) >= 0.999
        both = hg & ho
        if both.any():
            rel = np.abs(d64[both] - twin[both]) / twin[both]
            assert np.mean(rel <= 1e-4) >= 0.999
    # ordering
    pairs = ok = 0
    for i, q in zip(infos, reqs):
        if not q.preview:
            continue
        for j, r in zip(infos, reqs):
            if r.preview or r.arrival_s >= q.arrival_s or j.done_ms <= i.submit_ms + 300.0:
                continue
            pairs += 1
            ok += i.done_ms < j.done_ms
    print(f"previews ahead of earlier renders with more than 300 ms left: {ok}/{pairs}")
    assert pairs > 0 and ok == pairs


def test_autotuner_choices_leave_images_unchanged(ctx):
    """The autotuner picks, per kernel family, the persistent grid (2 or 3 workgroups per CU),
    the field evaluations per loop trip (4, 8, 16), the tile shape (8 x 8, 16 x 4, 32 x 2) and
    the quiet-mode samples per work item (8, 16) of the batched kernel; every choice only
    schedules the same per-pixel arithmetic, so forced choices give the same planes bit for bit."""
    from syfrack_amd import abi, serve
    reqs = [q for q in serve.uneven_requests(40, seed=13, min_side=64, max_side=160)]
    seen, pick = set(), []
    for q in reqs:
        fam = (q.scene.fractal.family, q.scene.fractal.power == 8.0)
        if fam not in seen:
            seen.add(fam)
            pick.append(q)
    planes = (abi.PLANE_ACCUM, abi.PLANE_DEPTH, abi.PLANE_AUX, abi.PLANE_DEPTH64)

    def serve_all(**kw):
        with serve.Server(ctx, "batched", **kw) as s:
            ids = [s.submit(q.scene, q.cfg) for q in pick]
            s.wait_all(300)
            out = [[s.plane(i, p) for p in planes] for i in ids]
            st = s.stats()
        return out, st

    tuned, st = serve_all()
    print("autotuner:", st["family_grid_blocks"], st["family_path_steps"], st["family_tile_width"],
          st["family_quiet_quantum"], f"{st['autotune_ms']:.0f} ms")
    assert st["autotune_ms"] > 0
    assert all(k in (4, 8, 16) for k in st["family_path_steps"])
    assert all(g > 0 and g % 256 == 0 and g // 256 in (2, 3) for g in st["family_grid_blocks"])
    assert all(w in (8, 16, 32) for w in st["family_tile_width"])
    assert all(q in (8, 16) for q in st["family_quiet_quantum"])
    for kw in (dict(path_steps=4, grid_blocks=256, tile_width=32, quantum_samples=3),
               dict(path_steps=16, grid_blocks=768, tile_width=16, quantum_samples=2),
               dict(path_steps=8, grid_blocks=512, tile_width=8, quantum_samples=1)):
        forced, st2 = serve_all(**kw)
        assert st2["family_path_steps"] == [kw["path_steps"]] * 6 and st2["autotune_ms"] == 0
        assert st2["family_tile_width"] == [kw["tile_width"]] * 6
        assert st2["family_quiet_quantum"] == [kw["quantum_samples"]] * 6
        for a, b in zip(tuned, forced):
            assert all(np.array_equal(x, y) for x, y in zip(a, b))
