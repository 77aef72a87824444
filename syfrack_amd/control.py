"""The control machine over the C ABI: ControlSession (crates/syfrack-render/src/control/mod.rs:
145-515) and Renderer::control_distances (:530-574).

A session keeps an image of a node-graph scene that the library advances in bounded quanta,
evaluating the graph with its interpreter: editing the scene (``set_scene``) uploads data and
never compiles a topology.  ``tick`` polls and submits at most one quantum and never waits;
``wait_complete`` is the blocking diagnostic.  The image is the machine's preview: pixel-centre
rays, the march without a step limit, tetrahedral normals and a headlight shade.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import Optional, Tuple

import numpy as np

from . import abi
from .graph import Graph
from .scene import RenderConfig, Scene


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _graph_of(scene: Scene) -> Graph:
    if not isinstance(scene.fractal, Graph):
        raise TypeError("the control machine evaluates node graphs: give the scene a syfrack_amd.graph.Graph")
    return scene.fractal


class ControlSession:
    """ControlSession::new(renderer, scene, size) and its methods."""

    def __init__(self, ctx, scene: Scene, size: Tuple[int, int], quantum: int = 0):
        self.ctx = ctx  # the context must outlive the session
        self.lib = ctx.lib
        self._requested = (int(size[0]), int(size[1]))
        self._started = time.perf_counter()
        desc = scene.desc(RenderConfig(self._requested[0], self._requested[1], 1, 0, 1))
        cg = _graph_of(scene).c_struct()
        h = C.c_void_p()
        abi.check(self.lib.syf_control_create(ctx.handle, C.byref(desc), C.byref(cg.s), quantum, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self.lib.syf_control_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scene(self, scene: Scene) -> None:
        """Upload a changed scene without compiling anything; the image restarts."""
        desc = scene.desc(RenderConfig(self._requested[0], self._requested[1], 1, 0, 1))
        cg = _graph_of(scene).c_struct()
        abi.check(self.lib.syf_control_set_scene(self.handle, C.byref(desc), C.byref(cg.s)))
        self._started = time.perf_counter()

    def tick(self) -> None:
        abi.check(self.lib.syf_control_tick(self.handle))

    def wait_complete(self, timeout_s: float = 120.0) -> None:
        abi.check(self.lib.syf_control_wait(self.handle, C.c_double(timeout_s)))

    def _progress(self):
        v = [C.c_uint32() for _ in range(5)]
        abi.check(self.lib.syf_control_progress(self.handle, *[C.byref(x) for x in v]))
        return [x.value for x in v]

    def complete(self) -> bool:
        return bool(self._progress()[2])

    def completed_pixels(self) -> int:
        return self._progress()[0]

    def dispatches(self) -> int:
        return self._progress()[1]

    def size(self) -> Tuple[int, int]:
        p = self._progress()
        return p[3], p[4]

    def requested_size(self) -> Tuple[int, int]:
        return self._requested

    def elapsed(self) -> float:
        return time.perf_counter() - self._started

    def readback(self):
        """(accum (h, w, 4), depth (h, w), aux (h, w, 4)) of the current image."""
        w, h = self.size()
        acc = np.empty((h, w, 4), np.float32)
        dep = np.empty((h, w), np.float32)
        aux = np.empty((h, w, 4), np.float32)
        abi.check(self.lib.syf_control_readback(self.handle, abi.PLANE_ACCUM, _p(acc), acc.nbytes))
        abi.check(self.lib.syf_control_readback(self.handle, abi.PLANE_DEPTH, _p(dep), dep.nbytes))
        abi.check(self.lib.syf_control_readback(self.handle, abi.PLANE_AUX, _p(aux), aux.nbytes))
        return acc, dep, aux

    def pick(self, u: float, v: float) -> Optional[Tuple[float, np.ndarray]]:
        hit = C.c_uint32()
        d = C.c_float()
        aux = np.zeros(4, np.float32)
        abi.check(self.lib.syf_control_pick(self.handle, C.c_float(u), C.c_float(v), C.byref(hit), C.byref(d), _p(aux)))
        return (d.value, aux) if hit.value else None


def control_distances(ctx, scene: Scene, points: np.ndarray):
    """Renderer::control_distances: the scene's graph evaluated by the interpreter at points
    (n, 3); returns a render.DeSample."""
    from .render import DeSample
    pts = np.zeros((len(points), 4), np.float32)
    pts[:, :3] = np.asarray(points, np.float32)[:, :3]
    out = np.zeros((len(points), 4), np.float32)
    traps = np.zeros((len(points), 4), np.float32)
    cg = _graph_of(scene).c_struct()
    abi.check(ctx.lib.syf_control_probe(ctx.handle, C.byref(cg.s), _p(pts), len(points), _p(out), _p(traps)))
    return DeSample(out[:, 0].copy(), traps, out[:, 2].astype(np.uint32), out[:, 3].astype(np.uint32))
