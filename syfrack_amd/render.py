"""Host-side mirror of the reference's Renderer API for this path, over the C ABI.

``Renderer.probe_distances`` mirrors Renderer::probe_distances (renderer.rs:645-648),
``Renderer.render`` mirrors Renderer::render (renderer.rs:668-693) returning a
``RenderOutput`` with accumulation, depth and normals planes (renderer.rs:151-165), and
``Renderer.trace_primary`` is its first-hit half (max_bounces = 0) with the FP64-refined hit
distance.  A scene whose fractal is a node graph (syfrack_amd.graph.Graph) becomes a graph job
(syf_job_create_graph).  Every call runs the HIP kernels of libsyfrack_hip.so; there is no
other path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import abi
from .graph import Graph
from .scene import RenderConfig, Scene


@dataclass
class DeSample:
    distance: np.ndarray    # (n,) f32
    traps: np.ndarray       # (n, 4) f32, normalised to [0, 1]
    iterations: np.ndarray  # (n,) u32
    material: np.ndarray    # (n,) u32


@dataclass
class RenderOutput:
    accum: Optional[np.ndarray]   # (h, w, 4): radiance sums and sample counts
    depth: np.ndarray             # (h, w): first-hit distance, -1 on a miss
    normals: np.ndarray           # (h, w, 4): normal xyz and steps; (0, 0, 0, -1) on a miss
    depth64: Optional[np.ndarray] = None  # (h, w): first surface hit refined in FP64, -1 otherwise
    steps: Optional[np.ndarray] = None    # (h, w): primary march steps
    stats: Optional[dict] = None


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One HIP device (syf_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = abi.load()
        h = C.c_void_p()
        abi.check(self.lib.syf_ctx_create(device, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self.lib.syf_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Job:
    """A render job (syf_job): the lowered scene plus its device planes."""

    def __init__(self, ctx: Context, desc: abi.SyfJobDesc, graph: Optional[Graph] = None,
                 graph_mode: str = "specialised"):
        self.ctx = ctx
        self.lib = ctx.lib
        self.desc = desc
        h = C.c_void_p()
        if graph is not None:
            if graph_mode not in ("specialised", "interpreted"):
                raise ValueError(f"graph_mode must be 'specialised' or 'interpreted', got {graph_mode!r}")
            mode = abi.GRAPH_SPECIALISED if graph_mode == "specialised" else abi.GRAPH_INTERPRETED
            cg = graph.c_struct()
            abi.check(self.lib.syf_job_create_graph(ctx.handle, C.byref(desc), cg.ref(), mode, C.byref(h)))
        else:
            abi.check(self.lib.syf_job_create(ctx.handle, C.byref(desc), C.byref(h)))
        self.handle = h
        self.width = desc.width
        self.height = desc.height

    def close(self):
        if self.handle:
            self.lib.syf_job_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        abi.check(self.lib.syf_job_clear(self.handle))

    def sync(self):
        abi.check(self.lib.syf_job_sync(self.handle))

    def stats(self) -> dict:
        s = abi.SyfStats()
        abi.check(self.lib.syf_job_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in abi.SyfStats._fields_}

    def trace_primary(self, sample: int = 0):
        abi.check(self.lib.syf_trace_primary(self.handle, sample))

    def render_tile(self, x, y, w, h, sample_start, sample_count):
        abi.check(self.lib.syf_render_tile(self.handle, x, y, w, h, sample_start, sample_count))

    def render_wavefront(self, x, y, w, h, sample_start, sample_count):
        abi.check(self.lib.syf_render_wavefront(self.handle, x, y, w, h, sample_start, sample_count))

    def readback(self, plane: int) -> np.ndarray:
        h, w = self.height, self.width
        shapes = {
            abi.PLANE_ACCUM: ((h, w, 4), np.float32),
            abi.PLANE_DEPTH: ((h, w), np.float32),
            abi.PLANE_AUX: ((h, w, 4), np.float32),
            abi.PLANE_DEPTH64: ((h, w), np.float64),
            abi.PLANE_STEPS: ((h, w), np.uint32),
            abi.PLANE_VOLUME: ((h, w, 4), np.float32),
            abi.PLANE_MOMENTS: ((h, w, 4), np.float32),
            abi.PLANE_SURFACE: ((h, w, 4), np.float32),
        }
        shape, dt = shapes[plane]
        out = np.empty(shape, dt)
        abi.check(self.lib.syf_readback(self.handle, plane, _p(out), out.nbytes))
        return out

    def upload(self, plane: int, data: np.ndarray) -> None:
        data = np.ascontiguousarray(data)
        abi.check(self.lib.syf_upload(self.handle, plane, _p(data), data.nbytes))

    def fingerprint(self) -> int:
        out = C.c_uint64()
        abi.check(self.lib.syf_job_fingerprint(self.handle, C.byref(out)))
        return out.value

    def checkpoint_save(self, directory: str, state_text: str) -> None:
        abi.check(self.lib.syf_checkpoint_save(self.handle, directory.encode(), state_text.encode()))

    def checkpoint_load(self, directory: str) -> None:
        abi.check(self.lib.syf_checkpoint_load(self.handle, directory.encode()))

    def export_write(self, path: str):
        """Writes the export payload; returns (bytes, FNV-1a-64 hash)."""
        n, h = C.c_uint64(), C.c_uint64()
        abi.check(self.lib.syf_export_write(self.handle, path.encode(), C.byref(n), C.byref(h)))
        return n.value, h.value

    def probe(self, points: np.ndarray) -> DeSample:
        pts = np.zeros((len(points), 4), np.float32)
        pts[:, :3] = np.asarray(points, np.float32)[:, :3]
        out = np.zeros((len(pts), 4), np.float32)
        traps = np.zeros((len(pts), 4), np.float32)
        abi.check(self.lib.syf_de_probe(self.handle, _p(pts), len(pts), _p(out), _p(traps)))
        return DeSample(out[:, 0].copy(), traps, out[:, 2].astype(np.uint32), out[:, 3].astype(np.uint32))


class Renderer:
    """Renderer (renderer.rs:254-266) for the distance-estimator path on one HIP device."""

    def __init__(self, device: int = 0):
        self.ctx = Context(device)

    def job(self, scene: Scene, cfg: RenderConfig) -> Job:
        graph = scene.fractal if isinstance(scene.fractal, Graph) else None
        return Job(self.ctx, scene.desc(cfg), graph, cfg.graph_mode)

    def probe_distances(self, scene: Scene, points: np.ndarray) -> DeSample:
        cfg = RenderConfig(1, 1, 1, 0, 1)
        j = self.job(scene, cfg)
        try:
            return j.probe(points)
        finally:
            j.close()

    def control_distances(self, scene: Scene, points: np.ndarray) -> DeSample:
        """Renderer::control_distances (control/mod.rs:530-574): the scene's graph evaluated by
        the control machine's interpreter, without compiling its topology."""
        from .control import control_distances
        return control_distances(self.ctx, scene, points)

    def control_session(self, scene: Scene, size, quantum: int = 0):
        """ControlSession::new(renderer, scene, size) (control/mod.rs:172)."""
        from .control import ControlSession
        return ControlSession(self.ctx, scene, size, quantum)

    def estimate_bounding_radius(self, graph: Graph) -> float:
        """Graph::estimate_bounding_radius (eval.rs:883-922) with the field evaluated on the
        device: from 100 units out along the 26 directions of the unit cube's neighbours, walk
        inward while the field stays positive and above 5 % of the radius; directions whose far
        point already lies within that margin (a ground plane) are skipped.  The walks of all
        directions advance together, one probe per step."""
        from .scene import Camera
        dirs = []
        for x in (-1, 0, 1):
            for y in (-1, 0, 1):
                for z in (-1, 0, 1):
                    if (x, y, z) != (0, 0, 0):
                        v = np.array([x, y, z], np.float32)
                        dirs.append(v * (np.float32(1.0) / np.float32(np.sqrt(np.float32(np.dot(v, v))))))
        dirs = np.array(dirs, np.float32)
        j = self.job(Scene(graph, Camera()), RenderConfig(1, 1, 1, 0, 1))
        try:
            f32 = np.float32
            r = np.full(len(dirs), f32(100.0), np.float32)
            far = j.probe(dirs * r[:, None]).distance
            walking = np.isfinite(far) & (far > f32(0.05) * r)
            if not walking.any():
                return 100.0
            active = walking.copy()
            while active.any():
                idx = np.nonzero(active)[0]
                idx = idx[r[idx] > f32(0.5)]
                active[:] = False
                if len(idx) == 0:
                    break
                dist = j.probe(dirs[idx] * r[idx, None]).distance
                go = np.isfinite(dist) & (dist > f32(0.05) * r[idx])
                step = np.maximum(np.minimum(dist[go], r[idx][go] * f32(0.5)), f32(0.01)).astype(np.float32)
                r[idx[go]] = (r[idx[go]] - step).astype(np.float32)
                active[idx[go]] = True
            radius = max(f32(0.5), r[walking].max())
            return float(min(max(radius, f32(0.5)), f32(100.0)))
        finally:
            j.close()

    def trace_primary(self, scene: Scene, cfg: RenderConfig, sample: int = 0) -> RenderOutput:
        j = self.job(scene, cfg)
        try:
            j.trace_primary(sample)
            j.sync()
            return RenderOutput(None, j.readback(abi.PLANE_DEPTH), j.readback(abi.PLANE_AUX),
                                j.readback(abi.PLANE_DEPTH64), j.readback(abi.PLANE_STEPS), j.stats())
        finally:
            j.close()

    def render(self, scene: Scene, cfg: RenderConfig, wavefront: bool = False) -> RenderOutput:
        """Progressive tile loop like RenderSession::step (session.rs:694-877): every tile of
        ``cfg.tile_size`` gets all samples, dispatched in chunks of 8 samples."""
        j = self.job(scene, cfg)
        try:
            ts = max(1, cfg.tile_size)
            spd = 8
            for s0 in range(0, cfg.samples_per_pixel, spd):
                n = min(spd, cfg.samples_per_pixel - s0)
                for y in range(0, cfg.height, ts):
                    for x in range(0, cfg.width, ts):
                        w, h = min(ts, cfg.width - x), min(ts, cfg.height - y)
                        if wavefront:
                            j.render_wavefront(x, y, w, h, s0, n)
                        else:
                            j.render_tile(x, y, w, h, s0, n)
            j.sync()
            return RenderOutput(j.readback(abi.PLANE_ACCUM), j.readback(abi.PLANE_DEPTH), j.readback(abi.PLANE_AUX),
                                depth64=j.readback(abi.PLANE_DEPTH64), stats=j.stats())
        finally:
            j.close()
