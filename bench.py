"""Benchmark of the sphere-traced distance-estimator path on one MI355X.

Workload: Mandelbulb primary-ray sphere trace at 1920x1080 (power 8, 10 iterations,
bailout 2.0, camera (0, 0, -2.5) -> origin, 60 degree vertical fov, 1 spp, <= 512 steps,
pixel tolerance 1.0) -- the FP32 march with the pixel-footprint hit rule, tetrahedral
normals, and the FP64 refine of every hit.  One step = one full frame (2,073,600 primary
rays) traced on the device; the job's planes stay in HBM, nothing is copied per step.

Prints one JSON line: Mrays/s (primary rays per second over the timed steps), DE evals/s,
the roofline of the dominant kernel (FP32 VALU, from device work counters and HIP-event
timing on the kernel's own stream), the CPU baseline (the oracle's restatement of the
same trace on a bounded sample of rows, on the host cores), and the serving legs: the
render server in batched mode against the plain one-launch-per-request form on an uneven
batch of 256 requests and on a Poisson serving loop, with rays/s (all traced segments,
primary + bounce + shadow) and p50/p99 submit-to-completion latency.  The same line records,
outside `value`, the other workloads of this path: the Mandelbox 1080p path trace in its
three kernel forms (`pathtrace`) and the 4096^2 perturbation deep zoom (`deep_zoom`).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--width 1920 --height 1080]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak (packed); non-packed is half
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def cpu_baseline(width: int, height: int, budget_s: float = 10.0) -> dict:
    """The oracle (CPU restatement of the same trace: FP32 march + normals + FP64 refine), in
    its baseline build (g++ -O3 -march=native -ffp-contract=off, OpenMP over the host's cores,
    built on this host), on whole rows from the image centre, the sample growing to whole
    frames (repeated) until it takes about budget_s."""
    from oracle import oracle  # the checker, used here only as the CPU baseline
    from syfrack_amd import scene as sc

    try:
        lib, build = oracle.load_native(), "g++ -O3 -march=native -ffp-contract=off"
    except Exception as e:  # no compiler on this host: the test build, and say so
        progress(f"CPU baseline: native build failed ({e}); timing the -O2 test build")
        lib, build = oracle.load(), "g++ -O2 -ffp-contract=off"
    scene, cfg = sc.mandelbulb_primary(width, height)
    desc = scene.desc(cfg)
    rows, frames = 4, 1
    while True:
        y0 = max(0, height // 2 - rows // 2)
        t0 = time.perf_counter()
        for f in range(frames):
            oracle.trace_primary(desc, f, 32, rect=(0, y0, width, rows), lib=lib)
        elapsed = time.perf_counter() - t0
        if elapsed > budget_s / 4 or frames >= 64:
            break
        if rows < height:
            rows = min(height, rows * 4)
        else:
            frames = min(64, max(frames * 2, int(frames * budget_s / max(elapsed, 1e-3))))
    rays = width * rows * frames
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    what = f"{frames} whole frames" if rows == height else f"{rows} centre rows x {width} px"
    return {"value": rays / elapsed / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"{what} of the same trace ({rays} primary rays, {elapsed:.2f} s); oracle C++ restatement, "
                      f"OpenMP, {build}"}


def _rel_stats(a, b, mask):
    import numpy as np
    a = np.asarray(a, np.float64)[mask]
    b = np.asarray(b, np.float64)[mask]
    if a.size == 0:
        return {"n": 0}
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
    return {"n": int(a.size), "max": float(rel.max()), "p99": float(np.percentile(rel, 99)),
            "frac_within_1e-4": float(np.mean(rel <= 1e-4))}


def error_report(r, march_math: str) -> dict:
    """The error of each configuration against the oracle (computed with the oracle, so here in
    the CPU leg, outside the timed region): hit/miss agreement and the relative error of the
    FP64-refined hit distance against the FP64 twin; path-traced radiance against the oracle's
    FP32 radiance; deep-zoom counts against binary128 where FP32 deltas resolve the point;
    served FP64 depth against the FP64 twin."""
    import numpy as np
    from oracle import oracle
    from syfrack_amd import abi, scene as sc, serve
    from syfrack_amd.planar import Planar, PlanarView, boundary_centre
    out = {}

    def primary(scene, cfg, rect=None):
        g = r.trace_primary(scene, cfg)
        ref = oracle.trace_primary(scene.desc(cfg), 0, 64, rect=rect)
        sl = (slice(None), slice(None)) if rect is None else (slice(rect[1], rect[1] + rect[3]),
                                                            slice(rect[0], rect[0] + rect[2]))
        gd, gd64 = g.depth[sl], g.depth64[sl]
        hg, ho = gd >= 0, ref["depth"] >= 0
        both = hg & ho
        return {"pixels": int(hg.size), "hit_miss_agreement": float(np.mean(hg == ho)),
                "refined_distance_rel_err": _rel_stats(gd64, ref["depth64"], both),
                "march_distance_rel_err": _rel_stats(gd, ref["depth"], both)}

    progress("errors: primary trace")
    s1, c1 = sc.mandelbulb_primary(320, 180)
    c1.march_math = march_math
    out["mandelbulb_primary_320x180"] = primary(s1, c1)
    s1, c1 = sc.mandelbulb_primary(1920, 1080)
    c1.march_math = march_math
    out["mandelbulb_primary_1080p_centre_32_rows"] = primary(s1, c1, rect=(0, 524, 1920, 32))

    progress("errors: path trace")
    s2, c2 = sc.mandelbox_pathtrace(1920, 1080, 64)
    rect = (896, 520, 128, 16)
    job = r.job(s2, c2)
    job.render_tile(*rect, 0, 64)
    job.sync()
    acc = job.readback(abi.PLANE_ACCUM)[rect[1]:rect[1] + rect[3], rect[0]:rect[0] + rect[2]]
    d64 = job.readback(abi.PLANE_DEPTH64)[rect[1]:rect[1] + rect[3], rect[0]:rect[0] + rect[2]]
    job.close()
    oacc, _, _, _ = oracle.render(s2.desc(c2), rect, 0, 64)
    twin = oracle.trace_primary(s2.desc(c2), 0, 64, rect=rect)["depth64"]
    g, o = acc[..., :3], oacc[..., :3]
    out["mandelbox_pathtrace_1080p_64spp_128x16_crop"] = {
        "pixels": int(g.shape[0] * g.shape[1]),
        "radiance_frac_within_1e-3": float(np.mean(np.all(np.abs(g - o) <= 1e-3 * np.abs(o) + 1e-4, axis=-1))),
        "image_mean_rel_err": float(abs(g.mean() - o.mean()) / abs(o.mean())),
        "first_hit_refined_distance_rel_err": _rel_stats(d64, twin, (d64 >= 0) & (twin >= 0))}

    progress("errors: deep zoom")
    views = []
    rng = np.random.default_rng(0)
    for fam, name in ((abi.PLANAR_MANDELBROT, "mandelbrot"), (abi.PLANAR_BURNING_SHIP, "burning_ship")):
        c = boundary_centre(r.ctx, fam, seed=0, iterations=2000)
        for span, mode in ((1e-6, abi.PLANAR_PERTURB), (1e-12, abi.PLANAR_PERTURB), (1e-12, abi.PLANAR_PERTURB64)):
            v = PlanarView(family=fam, mode=mode, width=4096, height=4096, iterations=10000,
                           center=c, span=span, bailout=16.0)
            p = Planar(r.ctx, v)
            p.run()
            it, _ = p.readback()
            st = p.stats()
            p.close()
            xs = rng.integers(0, 4096, 192)
            ys = rng.integers(0, 4096, 192)
            # PERTURB follows c = reference + f32 offset: binary128 there, resolvable under 16 f32
            # ulps of the offset; PERTURB64 follows the exact FP64 offset: binary128 there,
            # resolvable under 64 FP64 ulps
            truth, nudge = ((oracle.escape_f128, 16.0) if mode == abi.PLANAR_PERTURB
                            else (oracle.escape_f128_exact, 64.0))
            q = truth(v.c_struct(), xs, ys)[0].astype(np.int64)
            stable = np.ones(len(xs), bool)
            for n in (nudge, -nudge):
                stable &= np.abs(truth(v.c_struct(), xs, ys, n)[0].astype(np.int64) - q) <= 1
            near = np.abs(it[ys, xs].astype(np.int64) - q) <= 1
            views.append({"family": name, "span": span,
                          "mode": "perturb" if mode == abi.PLANAR_PERTURB else "perturb64",
                          "sampled": int(len(xs)), "resolvable": int(stable.sum()),
                          "count_within_1_of_binary128_on_resolvable": float(near[stable].mean())
                          if stable.any() else None,
                          "fallback_pixels": st["fallback_pixels"], "references": st["references"]})
    out["deep_zoom_4096_10000_iterations"] = views

    progress("errors: served requests")
    reqs = [q for q in serve.uneven_requests(64, seed=3) if q.cfg.width <= 384][:4]
    served = []
    with serve.Server(r.ctx, "batched") as srv:
        ids = [srv.submit(q.scene, q.cfg) for q in reqs]
        for i in ids:
            srv.wait(i, 300)
        for q, i in zip(reqs, ids):
            d64 = srv.plane(i, abi.PLANE_DEPTH64)
            srv.release(i)
            twin = oracle.trace_primary(q.scene.desc(q.cfg), 0, 64)["depth64"]
            hg, ho = d64 >= 0, twin >= 0
            served.append({"family": int(q.scene.fractal.family), "side": q.cfg.width,
                           "hit_miss_agreement": float(np.mean(hg == ho)),
                           "refined_distance_rel_err": _rel_stats(d64, twin, hg & ho)})
    out["served_requests_first_hit"] = served
    return out


def read_traffic(kernel: str, width: int, height: int):
    """HBM bytes per launch of a kernel at this frame size, from the committed rocprofv3 counter
    summary (``traffic_bytes: <kernel> <width>x<height> <bytes>`` lines), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path, encoding="utf-8"):
        f = line.split()
        if len(f) == 4 and f[0] == "traffic_bytes:" and f[1] == kernel and f[2] == f"{width}x{height}":
            return float(f[3])
    return None


def read_counters(kernel: str, width: int, height: int) -> dict:
    """VALU counters of a kernel at this frame size, from the committed rocprofv3 summary
    (``<name>: <kernel> <width>x<height> <value>`` lines); empty when absent."""
    path = os.path.join(ROOT, "profiles", "counters.txt")
    out = {}
    if os.path.exists(path):
        for line in open(path, encoding="utf-8"):
            f = line.split()
            if len(f) == 4 and f[0].endswith(":") and f[1] == kernel and f[2] == f"{width}x{height}":
                out[f[0][:-1]] = float(f[3])
    return out


def calibrate_copy() -> int:
    """A plain copy of 256 MiB (torch's device copy kernel), three times: the reference
    reading for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters into bytes."""
    import torch
    n = 256 * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 3}), flush=True)
    return 0


def pathtrace_leg(spp: int) -> int:
    """The path-trace leg alone: prints its JSON line."""
    print(json.dumps(pathtrace_results(spp)), flush=True)
    return 0


def pathtrace_results(spp: int) -> dict:
    """Mandelbox path trace at 1920x1080 (15 iterations, 4 bounces, one area light, <= 512
    steps) in its three forms: the megakernel (syf_render_tile over 256-pixel tiles, 8 samples
    per dispatch), the wavefront stages (syf_render_wavefront) and the flattened lane-persistent
    kernel (one request through the batched server).  Mrays/s (all traced segments: primary,
    bounce and shadow) and device time per form."""
    from syfrack_amd import scene as sc, serve
    from syfrack_amd.render import Renderer
    scene, cfg = sc.mandelbox_pathtrace(1920, 1080, spp)
    r = Renderer(0)
    out = {"workload": f"Mandelbox path trace 1920x1080, {spp} spp, 4 bounces, one area light, <=512 steps"}
    for name, wf in (("megakernel", False), ("wavefront", True)):
        progress(f"path trace: {name}")
        job = r.job(scene, cfg)
        rays, dev_ms = 0, 0.0
        t0 = time.perf_counter()
        for s0 in range(0, spp, 8):  # the whole frame per dispatch, 8 samples at a time
            n = min(8, spp - s0)
            if wf:
                job.render_wavefront(0, 0, cfg.width, cfg.height, s0, n)
            else:
                job.render_tile(0, 0, cfg.width, cfg.height, s0, n)
            st = job.stats()  # waits; counters and device time of this dispatch
            rays += st["rays"]
            dev_ms += st["kernel_ms"]
        el = time.perf_counter() - t0
        job.close()
        out[name] = {"s": el, "device_ms": dev_ms, "rays": rays, "mrays_s": rays / el / 1e6}
    progress("path trace: flattened (server)")
    with serve.Server(r.ctx, "batched") as srv:
        t0 = time.perf_counter()
        rid = srv.submit(scene, cfg)
        srv.wait(rid, 900)
        el = time.perf_counter() - t0
        st = srv.stats()
        srv.release(rid)
    out["flattened"] = {"s": el, "rays": st["rays"], "mrays_s": st["rays"] / el / 1e6, "launches": st["launches"]}
    return out


def control_results() -> dict:
    """The control machine (interactive preview of a node graph, no compilation): for the bundled
    Mandelbulb, Mandelbox and Menger graphs and the "cathedrale" graph at 1280x720 under an orbit
    camera, the wall time from set_scene to a complete image as an interactive client sees it
    (tick polled every 0.1 ms, never waiting) and the quanta it took (each bounded to 64 field
    evaluations per wave).  Scenes are set on a warm session."""
    from syfrack_amd.graph import Graph
    from syfrack_amd.render import Renderer
    from syfrack_amd.scene import Camera, Scene
    r = Renderer(0)
    size = (1280, 720)
    radii = {"mandelbulb": 1.25, "mandelbox": 4.0, "menger": 1.6, "cathedrale": 1.6}
    rows = []
    s = r.control_session(Scene(Graph.new_minimal(), Camera.orbit((0, 0, 0), 2.5, 35.0, 25.0, 35.0)), size)
    s.wait_complete()
    for name, radius in radii.items():
        progress(f"control: {name}")
        scene = Scene(Graph.bundled(name), Camera.orbit((0, 0, 0), 2.5 * radius, 35.0, 25.0, 35.0))
        t0 = time.perf_counter()
        s.set_scene(scene)
        while True:
            s.tick()
            if s.complete():
                break
            time.sleep(1e-4)
        el = time.perf_counter() - t0
        rows.append({"graph": name, "width": size[0], "height": size[1], "s": el, "quanta": s.dispatches(),
                     "mpix_s": size[0] * size[1] / el / 1e6})
    s.close()
    return {"workload": "control machine preview, 1280x720, bundled graphs, quantum 64 evaluations per wave",
            "views": rows}


def planar_leg(size: int, seeds: int) -> int:
    """The deep-zoom leg alone: prints its JSON line."""
    print(json.dumps(planar_results(size, seeds)), flush=True)
    return 0


def planar_results(size: int, seeds: int) -> dict:
    """2D deep zoom: size x size Mandelbrot and Burning Ship images at centres drawn in the
    boundary band (seeded), spans 1e-6 and 1e-12, 10000 iterations: PERTURB (double-double
    reference orbits + FP32 deltas with glitch re-reference, the glitches left after 8
    references taking FP64 deltas with rebasing) at both spans, and PERTURB64 (FP64 deltas
    with rebasing on every pixel) at 1e-12.  Mpixel/s and iterations/s per view, with the
    references used and the pixels left to the rebasing fallback."""
    from syfrack_amd import abi
    from syfrack_amd.planar import Planar, PlanarView, boundary_centre
    from syfrack_amd.render import Context
    ctx = Context(0)
    rows = []
    for fam, name in ((abi.PLANAR_MANDELBROT, "mandelbrot"), (abi.PLANAR_BURNING_SHIP, "burning_ship")):
        for seed in range(seeds):
            c = boundary_centre(ctx, fam, seed=seed, iterations=2000)
            for span, mode, mname in ((1e-6, abi.PLANAR_PERTURB, "perturb"), (1e-12, abi.PLANAR_PERTURB, "perturb"),
                                      (1e-12, abi.PLANAR_PERTURB64, "perturb64")):
                progress(f"planar: {name} seed {seed} span {span:g} {mname}")
                v = PlanarView(family=fam, mode=mode, width=size, height=size, iterations=10000,
                               center=c, span=span, bailout=16.0)
                p = Planar(ctx, v)
                t0 = time.perf_counter()
                p.run()
                p.sync()
                el = time.perf_counter() - t0
                st = p.stats()
                p.close()
                rows.append({"family": name, "seed": seed, "mode": mname, "centre": list(c), "span": span, "s": el,
                             "mpix_s": size * size / el / 1e6, "giter_s": st["iterations"] / el / 1e9,
                             "iterations": st["iterations"], "glitch_restarts": st["rebases"],
                             "references": st["references"], "fallback_pixels": st["fallback_pixels"],
                             "rebasings": st["rebasings"], "kernel_ms": st["kernel_ms"]})
    return {"workload": f"{size}x{size} deep zoom, 10000 iterations, perturbation", "views": rows}


def progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if xs else None


def serve_stream(ctx, reqs, mode: str, realtime: bool, budget: int = 200 * 10**9, **server_kw) -> dict:
    """Runs a request stream through a render server; wall time from the first submission to
    the last completion.  With ``realtime`` each request is submitted at its arrival time."""
    from syfrack_amd import serve
    descs = [(q.scene, q.cfg) for q in reqs]
    with serve.Server(ctx, mode, hbm_budget_bytes=budget, **server_kw) as s:
        t0 = time.perf_counter()
        ids = []
        for q, (sc_, cfg) in zip(reqs, descs):
            if realtime:
                dt = q.arrival_s - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
            ids.append(s.submit(sc_, cfg))
        for k, i in enumerate(ids):
            s.wait(i, 900)
            if (k + 1) % max(1, len(ids) // 4) == 0:
                progress(f"{mode}: {k + 1}/{len(ids)} requests done at {time.perf_counter() - t0:.1f} s")
        wall = time.perf_counter() - t0
        st = s.stats()
        infos = [s.poll(i) for i in ids]
        for i in ids:
            s.release(i)
    prev = [i.latency_ms for i, q in zip(infos, reqs) if q.preview]
    rest = [i.latency_ms for i, q in zip(infos, reqs) if not q.preview]
    return {"mode": mode, "wall_s": wall, "rays": st["rays"], "primary_rays": st["primary_rays"],
            "mrays_s": st["rays"] / wall / 1e6, "device_ms": st["device_ms"], "launches": st["launches"],
            "quantum_samples": st["quantum_samples"], "grid_blocks": st["grid_blocks"],
            "batch_target_us": st["batch_target_us"], "hbm_peak_gb": st["hbm_peak"] / 1e9,
            "preview_p50_ms": _pct(prev, 50), "preview_p99_ms": _pct(prev, 99),
            "render_p50_ms": _pct(rest, 50), "render_p99_ms": _pct(rest, 99),
            "all_p50_ms": _pct([i.latency_ms for i in infos], 50), "all_p99_ms": _pct([i.latency_ms for i in infos], 99)}


def serve_window(ctx, reqs, mode: str, window_s: float, budget: int = 200 * 10**9) -> dict:
    """A serving loop over a fixed arrival window: requests are submitted at their arrival times
    until ``window_s``; then every progressive render not finished is cancelled (its launched
    work still completes) and the queued previews are left to finish.  Aggregate Mrays/s = rays
    traced / wall time from the first submission to the drain; preview latency = submit to
    completion."""
    from syfrack_amd import serve
    reqs = [q for q in reqs if q.arrival_s <= window_s]
    with serve.Server(ctx, mode, hbm_budget_bytes=budget) as s:
        t0 = time.perf_counter()
        ids = []
        next_note = 10.0
        for q in reqs:
            dt = q.arrival_s - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(s.submit(q.scene, q.cfg))
            if time.perf_counter() - t0 > next_note:
                progress(f"{mode} loop: {len(ids)}/{len(reqs)} submitted at {time.perf_counter() - t0:.0f} s")
                next_note += 10.0
        dt = window_s - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        t_end = time.perf_counter() - t0
        renders_done_in_window = 0
        cancelled = 0
        for i, q in zip(ids, reqs):
            if q.preview:
                continue
            if s.poll(i).state == "done":
                renders_done_in_window += 1
            else:
                s.cancel(i)
                cancelled += 1
        s.wait_all(900)
        wall = time.perf_counter() - t0
        st = s.stats()
        infos = [s.poll(i) for i in ids]
        for i in ids:
            s.release(i)
    prev = [i.latency_ms for i, q in zip(infos, reqs) if q.preview and i.state == "done"]
    prev_in_window = [i.latency_ms for i, q in zip(infos, reqs)
                      if q.preview and i.state == "done" and i.done_ms <= t_end * 1e3 + infos[0].submit_ms]
    # previews that finished before the progressive renders that arrived earlier and still had
    # more than 300 ms of work left when the preview arrived (unfinished renders count as later)
    rend = [(q.arrival_s, i) for i, q in zip(infos, reqs) if not q.preview]
    over, pairs = 0, 0
    for i, q in zip(infos, reqs):
        if not q.preview or i.state != "done":
            continue
        for a, ri in rend:
            if a < q.arrival_s and (ri.state != "done" or ri.done_ms > i.submit_ms + 300.0):
                pairs += 1
                over += (ri.state != "done") or (i.done_ms < ri.done_ms)
    return {"mode": mode, "window_s": window_s, "requests": len(reqs), "previews": sum(q.preview for q in reqs),
            "renders": sum(not q.preview for q in reqs), "renders_done_in_window": renders_done_in_window,
            "renders_cancelled": cancelled, "drain_s": wall - t_end, "wall_s": wall, "rays": st["rays"],
            "mrays_s": st["rays"] / wall / 1e6, "device_ms": st["device_ms"], "launches": st["launches"],
            "hbm_peak_gb": st["hbm_peak"] / 1e9, "previews_done": len(prev),
            "preview_p50_ms": _pct(prev, 50), "preview_p99_ms": _pct(prev, 99),
            "previews_done_in_window": len(prev_in_window),
            "preview_overtakes_earlier_renders": over / pairs if pairs else None}


def serving_loop_bench(ctx, rates, plain_window_s: float) -> dict:
    """The Poisson serving loop at each rate: batched over the window that brings 1,000 previews,
    the plain form over the same window capped at plain_window_s."""
    from syfrack_amd import serve
    out = {"workload": "Poisson arrivals, 70% preview tiles 256^2 1 spp / 30% progressive renders 1024^2-2048^2 "
                       "16-64 spp over the fractal mix (bounces 1-4, <=256 steps); each rate's window brings 1,000 "
                       "previews; at the window's end unfinished renders are cancelled and queued previews "
                       "finish; 200 GB admission budget", "rates": []}
    for rate in rates:
        reqs = serve.serving_window(rate, 1000, seed=5)
        window = reqs[-1].arrival_s + 1e-3
        progress(f"serving loop at {rate:g}/s: {len(reqs)} requests over {window:.1f} s")
        b = serve_window(ctx, reqs, "batched", window)
        pw = min(window, plain_window_s)
        p = serve_window(ctx, reqs, "plain", pw)
        out["rates"].append({"rate_per_s": rate, "batched": b, "plain": p, "speedup": b["mrays_s"] / p["mrays_s"]})
    return out


def serving_bench(ctx, n_uneven: int) -> dict:
    """Batched serving against the plain one-launch-per-request form on the uneven batch."""
    from syfrack_amd import abi, serve
    out = {}
    uneven = serve.uneven_requests(n_uneven, seed=3)
    progress(f"uneven batch of {n_uneven} requests")
    # a batch submitted at once: once the submissions stop, everything left goes in one
    # persistent launch (quiet_target_us unbounded)
    b = serve_stream(ctx, uneven, "batched", realtime=False, quiet_target_us=abi.QUIET_UNBOUNDED)
    p = serve_stream(ctx, uneven, "plain", realtime=False)
    out["uneven_batch"] = {
        "workload": f"{n_uneven} requests submitted at once: square resolutions log-uniform 128-1024, "
                    "Mandelbulb (power 2-12) / Mandelbox (scale -2.5..3) / Menger (4-7 levels), spp 1-16, "
                    "bounces 1-4, <=256 steps, sky light; batched: 30 ms launches while requests arrive, then "
                    "all remaining work in one persistent launch",
        "batched": b, "plain": p, "speedup": b["mrays_s"] / p["mrays_s"], "same_rays": b["rays"] == p["rays"]}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving (batched vs plain), path-trace and deep-zoom legs")
    ap.add_argument("--serve-requests", type=int, default=256, help="requests of the uneven batch")
    ap.add_argument("--loop-rates", default="20,200",
                    help="arrival rates (per second) of the Poisson serving loop, comma-separated")
    ap.add_argument("--loop-plain-window", type=float, default=20.0,
                    help="seconds of each rate's stream the plain form serves (the batched form serves the "
                         "whole window that brings 1,000 previews)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 256 MiB device copy (reference for the byte counters)")
    ap.add_argument("--pathtrace", action="store_true",
                    help="run only the Mandelbox 1080p path-trace leg (three kernel forms)")
    ap.add_argument("--spp", type=int, default=64, help="samples per pixel of the path-trace leg")
    ap.add_argument("--planar", action="store_true", help="run only the 2D deep-zoom leg")
    ap.add_argument("--control", action="store_true", help="run only the control-machine leg")
    ap.add_argument("--planar-size", type=int, default=4096)
    ap.add_argument("--planar-seeds", type=int, default=2)
    ap.add_argument("--march-math", choices=["fast", "exact"], default="fast",
                    help="arithmetic of the primary trace's FP32 march (fast: hardware rcp/rsq/log + FMA; "
                         "exact: IEEE operations as written, step for step with the oracle's FP32 twin)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.control:
        print(json.dumps(control_results()), flush=True)
        return 0
    if args.pathtrace:
        return pathtrace_leg(args.spp)
    if args.planar:
        return planar_leg(args.planar_size, args.planar_seeds)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this path shards by whole "
              "requests or tiles with no exchange, see the design notes", file=sys.stderr)
        return 2

    from syfrack_amd import abi, flops
    from syfrack_amd import scene as sc
    from syfrack_amd.render import Renderer

    progress("primary trace")
    scene, cfg = sc.mandelbulb_primary(args.width, args.height)
    cfg.march_math = args.march_math
    r = Renderer(0)
    job = r.job(scene, cfg)
    fam = scene.fractal.family
    for i in range(args.warmup):
        job.trace_primary(i)
        job.sync()
    # Timed: K frames (samples W .. W+K-1) queued back to back on the job's stream, as a
    # renderer producing a frame sequence does; synchronised before and after.
    job.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        job.trace_primary(args.warmup + k)
    job.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    # The same K frames again, one at a time, for each frame's HIP-event kernel times and device
    # work counters (read after every frame; not part of the timed region).
    per = []
    for k in range(args.steps):
        job.trace_primary(args.warmup + k)
        per.append(job.stats())
    rays = args.width * args.height * args.steps
    de_evals = sum(p["de_evals"] + p["de64_evals"] for p in per)
    fused_ms = sum(p["march_ms"] for p in per) / len(per)   # k_primary alone (HIP events on its stream)
    kernel_ms = sum(p["kernel_ms"] for p in per) / len(per)  # k_primary + k_hits_tail
    fl32, fl64 = (sum(x) / len(per) for x in zip(*(flops.fused_flops(fam, p) for p in per)))
    fl_march = sum(flops.march_flops(fam, p) for p in per) / len(per)
    achieved = (fl32 + fl64) / (fused_ms * 1e-3) / 1e12
    hits = per[-1]["hits"]
    tail_hits = sum(p["tail_hits"] for p in per) / len(per)
    tail_ms = kernel_ms - fused_ms
    traffic = read_traffic("k_primary", args.width, args.height)
    out = {
        "metric": "Mrays/s + DE evals/s (Mandelbulb 1080p); p50/p99 tile latency vs plain",
        "value": rays / elapsed / 1e6,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (march) + f64 (final-hit refine)",
        "data": "synthetic: procedural scene, rays from the PCG stream",
        "config": {"workload": f"Mandelbulb primary-ray sphere trace {args.width}x{args.height}, power 8, "
                               "10 iterations, bailout 2, 1 spp, <=512 steps, pixel tolerance 1, FP64 hit refine",
                   "width": args.width, "height": args.height, "spp": 1, "max_steps": 512,
                   "march_math": args.march_math,
                   "hit_rule": "pixel cone |d| safety / t < pixel_cone * tolerance (the reference's rule); "
                               "FP64 refine to 1e-9 t of its crossing",
                   "parallelism": "1 GPU"},
        "de_evals_per_s": de_evals / elapsed,
        "de_evals_fp32_per_step": per[-1]["de_evals"],
        "de_evals_fp64_per_step": per[-1]["de64_evals"],
        "hits_per_step": hits,
        "device_ms_per_step": kernel_ms,
        "fused_kernel_ms": fused_ms,
        "tail_kernel_ms": tail_ms,
        "tail_hits_per_step": tail_hits,
        "fused_kernel_fp32_tflops": fl32 / (fused_ms * 1e-3) / 1e12,
        "fused_kernel_fp64_tflops": fl64 / (fused_ms * 1e-3) / 1e12,
        "march_fp32_flops_per_step": fl_march,
        "roofline": {"bound": "valu", "kernel": "k_primary<power-8 Mandelbulb> (march + hit shading)",
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                     "flops_per_launch": fl32 + fl64, "fp32_flops_per_launch": fl32, "fp64_flops_per_launch": fl64,
                     "note": "FP32 VALU bound (no MFMA; HBM traffic a few tens of bytes per ray); algorithmic "
                             "FLOPs of the fused kernel alone from device counters (march: Iterate bodies, "
                             "evaluations, steps; its hits: normal taps in FP32, refine evaluations in FP64, which "
                             "issue at the FP32 non-packed rate on MI355X), special ops (div/sqrt/log) not counted "
                             "in achieved but given apart (special_ops_per_launch, special_ops_per_s); "
                             "packed FP32 is disabled in the build, so the reachable ceiling is 78.6 TFLOP/s; "
                             "traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a "
                             "plain copy kernel; valu_counters = VALU busy, lane utilisation and resident waves per "
                             "CU of the same kernel from a committed rocprofv3 --pmc run"},
    }
    sp = [flops.special_ops(fam, p) for p in per]
    if all(x is not None for x in sp):
        # the correctly rounded division, square root and logarithm the FLOP count leaves out;
        # each issues as a sequence of about ten or more VALU instructions
        avg = {prec: {op: sum(x[prec][op] for x in sp) / len(sp) for op in sp[0][prec]} for prec in sp[0]}
        out["roofline"]["special_ops_per_launch"] = avg
        out["roofline"]["special_ops_per_s"] = sum(v for d in avg.values() for v in d.values()) / (fused_ms * 1e-3)
    counters = read_counters("k_primary", args.width, args.height)
    if counters:
        out["roofline"]["valu_counters"] = counters
    # the frame job's stream holds one of the process's hardware queues: release it before the
    # servers create their launch streams, so that each of those gets a queue of its own
    job.close()
    if not args.no_serve:
        sv = serving_bench(r.ctx, args.serve_requests)
        sv["serving_loop"] = serving_loop_bench(r.ctx, [float(x) for x in args.loop_rates.split(",") if x.strip()],
                                                args.loop_plain_window)
        out["serving"] = sv
        rates = sv["serving_loop"]["rates"]
        out["latency"] = {
            "per_rate": [{"rate_per_s": x["rate_per_s"], "previews": x["batched"]["previews_done"],
                          "preview_p50_ms": x["batched"]["preview_p50_ms"],
                          "preview_p99_ms": x["batched"]["preview_p99_ms"],
                          "plain_previews": x["plain"]["previews_done"],
                          "plain_preview_p50_ms": x["plain"]["preview_p50_ms"],
                          "plain_preview_p99_ms": x["plain"]["preview_p99_ms"]} for x in rates],
            "note": "submit-to-completion latency of the 256^2 preview requests in the Poisson serving loop, "
                    "batched server vs the plain one-launch-per-request form"}
        out["batched_vs_plain_rays"] = {"uneven_batch": sv["uneven_batch"]["speedup"],
                                        "serving_loop": {str(x["rate_per_s"]): x["speedup"] for x in rates}}
        # the other configurations' workloads, for the record (not part of `value`)
        out["pathtrace"] = pathtrace_results(args.spp)
        out["deep_zoom"] = planar_results(args.planar_size, args.planar_seeds)
        out["control"] = control_results()
    if not args.no_cpu:
        progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(args.width, args.height)
        out["errors"] = error_report(r, args.march_math)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
