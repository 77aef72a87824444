"""Render server over the C ABI: continuous batching of uneven requests on one device.

For this path it takes the place of the reference's export scheduling — the tile queue
(TileWorkQueue, export_queue.rs:7-120), the two in-flight export slots (export.rs:453-464) and
the progressive tile loop of RenderSession::step (session.rs:694-877).  Requests (scene,
resolution, spp) are submitted as they arrive; the library admits them against an HBM budget,
keeps their accumulation planes resident and traces them in shared persistent launches
(``mode="batched"``) or one launch per request in arrival order (``mode="plain"``).  A
request's image does not depend on the mode or on the other requests in flight.

The two request streams the throughput and latency figures are quoted on are generated here
with a seed: ``uneven_requests`` (resolutions log-uniform 128^2..1024^2 over the fractal mix,
spp 1-16, bounces 1-4) and ``serving_mix`` (Poisson arrivals; 70 % preview tiles of 256^2 at
1 spp, 30 % progressive renders of 1024^2..2048^2 at 16-64 spp).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import abi
from .render import Context, RenderOutput, _p
from .scene import Camera, Fractal, Light, Material, RenderConfig, Scene

BYTES_PER_PIXEL = 16 + 4 + 16 + 8  # accumulation, depth, aux and FP64 depth planes

STATE_NAMES = {abi.REQ_QUEUED: "queued", abi.REQ_ACTIVE: "active", abi.REQ_DONE: "done",
               abi.REQ_CANCELLED: "cancelled"}


def footprint(cfg: RenderConfig) -> int:
    """Device bytes a request holds while admitted."""
    return cfg.width * cfg.height * BYTES_PER_PIXEL


@dataclass
class RequestInfo:
    state: str
    samples_done: int
    spp: int
    footprint_bytes: int
    submit_ms: float
    start_ms: float
    done_ms: float

    @property
    def latency_ms(self) -> float:
        return self.done_ms - self.submit_ms


class Server:
    """syf_server: requests in, resident accumulation planes out."""

    def __init__(self, ctx: Context, mode: str = "batched", hbm_budget_bytes: int = 0, max_active: int = 0,
                 quantum_samples: int = 0, batch_target_us: int = 0, grid_blocks: int = 0,
                 quiet_target_us: int = 0, path_steps: int = 0, tile_width: int = 0):
        if mode not in ("batched", "plain"):
            raise ValueError(f"mode must be 'batched' or 'plain', got {mode!r}")
        self.ctx = ctx
        self.lib = ctx.lib
        cfg = abi.SyfServerConfig()
        cfg.hbm_budget_bytes = hbm_budget_bytes
        cfg.mode = abi.SERVE_BATCHED if mode == "batched" else abi.SERVE_PLAIN
        cfg.max_active = max_active
        cfg.quantum_samples = quantum_samples
        cfg.batch_target_us = batch_target_us
        cfg.grid_blocks = grid_blocks
        cfg.quiet_target_us = quiet_target_us
        cfg.path_steps = path_steps
        cfg.tile_width = tile_width
        h = C.c_void_p()
        abi.check(self.lib.syf_server_create(ctx.handle, C.byref(cfg), C.byref(h)))
        self.handle = h
        self.mode = mode
        self._shape = {}

    def close(self):
        if self.handle:
            self.lib.syf_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, scene: Scene, cfg: RenderConfig, spp: Optional[int] = None) -> int:
        """Queues a render of ``spp`` samples per pixel (default cfg.samples_per_pixel)."""
        desc = scene.desc(cfg)
        rid = C.c_uint64()
        n = cfg.samples_per_pixel if spp is None else spp
        abi.check(self.lib.syf_server_submit(self.handle, C.byref(desc), n, C.byref(rid)))
        self._shape[rid.value] = (cfg.height, cfg.width)
        return rid.value

    def poll(self, rid: int) -> RequestInfo:
        info = abi.SyfRequestInfo()
        abi.check(self.lib.syf_server_poll(self.handle, rid, C.byref(info)))
        return RequestInfo(STATE_NAMES[info.state], info.samples_done, info.spp, info.footprint_bytes,
                           info.submit_ms, info.start_ms, info.done_ms)

    def wait(self, rid: int, timeout_s: float = 60.0) -> None:
        abi.check(self.lib.syf_server_wait(self.handle, rid, int(timeout_s * 1000)))

    def wait_all(self, timeout_s: float = 600.0) -> None:
        abi.check(self.lib.syf_server_wait_all(self.handle, int(timeout_s * 1000)))

    def plane(self, rid: int, plane: int) -> np.ndarray:
        h, w = self._shape[rid]
        if plane == abi.PLANE_DEPTH64:
            out = np.empty((h, w), np.float64)
        else:
            out = np.empty((h, w, 4) if plane in (abi.PLANE_ACCUM, abi.PLANE_AUX) else (h, w), np.float32)
        abi.check(self.lib.syf_server_result(self.handle, rid, plane, _p(out), out.nbytes))
        return out

    def result(self, rid: int) -> RenderOutput:
        """The accumulation, depth, normals and FP64 depth planes of a finished request (as
        Renderer.render)."""
        return RenderOutput(self.plane(rid, abi.PLANE_ACCUM), self.plane(rid, abi.PLANE_DEPTH),
                            self.plane(rid, abi.PLANE_AUX), self.plane(rid, abi.PLANE_DEPTH64))

    def release(self, rid: int) -> None:
        abi.check(self.lib.syf_server_release(self.handle, rid))
        self._shape.pop(rid, None)

    def cancel(self, rid: int) -> None:
        abi.check(self.lib.syf_server_cancel(self.handle, rid))

    def stats(self) -> dict:
        c = abi.SyfServerCounters()
        abi.check(self.lib.syf_server_stats(self.handle, C.byref(c)))
        out = {name: getattr(c, name) for name, _ in c._fields_ if name != "reserved"}
        out["family_grid_blocks"] = list(c.family_grid_blocks)
        out["family_path_steps"] = list(c.family_path_steps)
        out["family_tile_width"] = list(c.family_tile_width)
        out["family_quiet_quantum"] = list(c.family_quiet_quantum)
        return out


# ---------------------------------------------------------------- request streams

def _environment() -> Light:
    return Light.environment((0.85, 0.9, 1.0), (0.55, 0.55, 0.6), (0.15, 0.13, 0.12), 1.0)


def random_scene(rng: np.random.Generator, seed: int) -> Scene:
    """One scene of the fractal mix: Mandelbulb (integer power 2-12), Mandelbox (scale -2.5
    to 3) or Menger (4-7 levels), chosen uniformly, seen from a random orbit under a sky."""
    kind = int(rng.integers(0, 3))
    az, el = float(rng.uniform(0.0, 360.0)), float(rng.uniform(5.0, 40.0))
    if kind == 0:
        f = Fractal.bundled("mandelbulb")
        f.set_param("fractal.nodes[1].power", float(rng.integers(2, 13)))
        f.set_param("fractal.nodes[1].iterations", 10)
        cam = Camera.orbit((0.0, 0.0, 0.0), 3.2, az, el, 35.0)
    elif kind == 1:
        f = Fractal.bundled("mandelbox")
        f.set_param("fractal.nodes[1].scale", round(float(rng.uniform(-2.5, 3.0)), 3))
        f.set_param("fractal.nodes[1].iterations", 12)
        cam = Camera.orbit((0.0, 0.0, 0.0), 11.0, az, el, 35.0)
    else:
        f = Fractal.bundled("menger")
        f.set_param("fractal.nodes[1].iterations", int(rng.integers(4, 8)))
        cam = Camera.orbit((0.0, 0.0, 0.0), 4.2, az, el, 35.0)
    mat = Material(albedo=(0.8, 0.75, 0.7), roughness=0.35, metallic=0.0, specular=0.04,
                   surface=(0, (0.35, 0.35), (1.0, 0.0)))
    return Scene(f, cam, [_environment()], mat, seed=seed)


@dataclass
class ServeRequest:
    scene: Scene
    cfg: RenderConfig
    arrival_s: float = 0.0
    preview: bool = False

    @property
    def spp(self) -> int:
        return self.cfg.samples_per_pixel


def uneven_requests(n: int = 256, seed: int = 3, min_side: int = 128, max_side: int = 1024) -> List[ServeRequest]:
    """Uneven render jobs: square resolutions log-uniform in [min_side, max_side] (multiples of
    8), the fractal mix above, spp uniform 1-16, bounces uniform 1-4, max 256 steps."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(round(math.exp(rng.uniform(math.log(min_side), math.log(max_side))) / 8.0)) * 8
        spp = int(rng.integers(1, 17))
        bounces = int(rng.integers(1, 5))
        cfg = RenderConfig(side, side, samples_per_pixel=spp, max_bounces=bounces, max_steps=256)
        out.append(ServeRequest(random_scene(rng, seed * 100003 + i), cfg))
    return out


def serving_window(rate_per_s: float, min_previews: int = 1000, seed: int = 5, preview_fraction: float = 0.7,
                   big_sides=(1024, 2048), big_spp=(16, 64)) -> List[ServeRequest]:
    """The serving mix of ``serving_mix`` over a fixed arrival window: Poisson arrivals at
    ``rate_per_s`` until ``min_previews`` preview tiles have arrived (the window is the last
    arrival time)."""
    out: List[ServeRequest] = []
    n = 0
    while sum(q.preview for q in out) < min_previews:
        n = max(2 * n, int(min_previews / preview_fraction * 1.05) + 8)
        out = serving_mix(n, rate_per_s, seed, preview_fraction, big_sides, big_spp)
        if sum(q.preview for q in out) >= min_previews:
            cut = [i for i, q in enumerate(out) if q.preview][min_previews - 1]
            out = out[:cut + 1]
    return out


def serving_mix(n: int = 64, rate_per_s: float = 100.0, seed: int = 5, preview_fraction: float = 0.7,
                big_sides=(1024, 2048), big_spp=(16, 64)) -> List[ServeRequest]:
    """Poisson arrivals at ``rate_per_s``: a ``preview_fraction`` of preview tiles (256^2, 1 spp)
    and progressive renders (square side uniform in ``big_sides`` rounded to 64, spp uniform in
    ``big_spp``), over the fractal mix above with bounces 1-4."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate_per_s))
        bounces = int(rng.integers(1, 5))
        if rng.uniform() < preview_fraction:
            cfg = RenderConfig(256, 256, samples_per_pixel=1, max_bounces=bounces, max_steps=256)
            preview = True
        else:
            side = int(round(rng.uniform(*big_sides) / 64.0)) * 64
            cfg = RenderConfig(side, side, samples_per_pixel=int(rng.integers(big_spp[0], big_spp[1] + 1)),
                               max_bounces=bounces, max_steps=256)
            preview = False
        out.append(ServeRequest(random_scene(rng, seed * 100003 + i), cfg, t, preview))
    return out
