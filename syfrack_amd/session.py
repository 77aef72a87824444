"""RenderSession over the C ABI: progressive rendering, readback, checkpoint and resume.

Mirrors crates/syfrack-render/src/session.rs for this path: ``step`` adds the next samples per
pixel to every tile in ascending order (RenderSession::step, session.rs:694-877, with the tile
grid of tiles.rs); ``readback`` gives the mean colour, depth and normals (decode_readback,
session.rs:917-1000; this path has no participating medium, so no radiance layers); ``save``
and ``resume`` follow session.rs:1199-1330: the library writes the device planes and replaces
the checkpoint directory atomically (syf_checkpoint_save), this module writes ``state.ron``
with the scene, the render configuration, the progress and the transport fingerprint, and a
resume rebuilds the job from it, refuses a checkpoint of another transport and reloads the
planes; ``export`` writes the export payload with its FNV-1a-64 checksum (export.rs:1226-1324).
"""
from __future__ import annotations

import dataclasses
import os
import time
from typing import Any

import numpy as np

from . import abi
from .ron import from_ron, to_ron  # noqa: F401  (state.ron)
from .graph import Graph
from .render import Renderer, RenderOutput
from .scene import Camera, Coloring, Finish, Fog, Fractal, Ground, Light, Material, Ramp, RenderConfig, Scene

STATE_FILE = "state.ron"


def _build(cls, d: dict):
    kw = {}
    for f in dataclasses.fields(cls):
        if f.name not in d:
            continue
        v = d[f.name]
        if isinstance(v, list):
            v = tuple(v)
        kw[f.name] = v
    return cls(**kw)


def scene_from_state(d: dict) -> Scene:
    s = d
    lights = [_build(Light, l) for l in s.get("lights", [])]
    md = s["material"]
    surface = md.get("surface")
    m = _build(Material, {k: v for k, v in md.items() if k not in ("surface", "coloring", "finishes")})
    m.surface = None if surface is None else (int(surface[0]), tuple(surface[1]), tuple(surface[2]))
    c = md.get("coloring")
    if c is not None:
        r = c["ramp"]
        ramp = _build(Ramp, {k: v for k, v in r.items() if k not in ("stops", "positions")})
        ramp.stops = [tuple(x) for x in r["stops"]]
        ramp.positions = [float(x) for x in r.get("positions", [])]
        orbit = c.get("orbit")
        m.coloring = Coloring(ramp, int(c["slot"]),
                              None if orbit is None else (tuple(orbit[0]), orbit[1], orbit[2], orbit[3]))
    m.finishes = [_build(Finish, f) for f in md.get("finishes", [])]
    g = s.get("ground")
    fog = s.get("fog")
    fr = s["fractal"]
    fractal = Graph.from_dict(fr) if fr.get("__struct__") == "Graph" else _build(Fractal, fr)
    return Scene(fractal, _build(Camera, s["camera"]), lights, m, seed=int(s["seed"]),
                 cavity_darkness=float(s["cavity_darkness"]), ground=None if g is None else _build(Ground, g),
                 fog=None if fog is None else _build(Fog, fog))


# ---------------------------------------------------------------- the session

class RenderSession:
    """A progressive render of one scene on the device (session.rs:164-1330)."""

    def __init__(self, renderer: Renderer, scene: Scene, cfg: RenderConfig, samples_per_step: int = 8):
        self.renderer = renderer
        self.scene = scene
        self.cfg = cfg
        self.spd = max(1, samples_per_step)
        self.job = renderer.job(scene, cfg)
        self.samples_done = 0
        self.elapsed = 0.0

    def close(self):
        self.job.close()

    def done(self) -> bool:
        return self.samples_done >= self.cfg.samples_per_pixel

    def step(self) -> int:
        """Adds the next samples (at most samples_per_step) to every tile; returns the count."""
        n = min(self.spd, self.cfg.samples_per_pixel - self.samples_done)
        if n <= 0:
            return 0
        t0 = time.perf_counter()
        ts = max(1, self.cfg.tile_size)
        for y in range(0, self.cfg.height, ts):
            for x in range(0, self.cfg.width, ts):
                self.job.render_tile(x, y, min(ts, self.cfg.width - x), min(ts, self.cfg.height - y),
                                     self.samples_done, n)
        self.job.sync()
        self.elapsed += time.perf_counter() - t0
        self.samples_done += n
        return n

    def readback(self) -> RenderOutput:
        return RenderOutput(self.job.readback(abi.PLANE_ACCUM), self.job.readback(abi.PLANE_DEPTH),
                            self.job.readback(abi.PLANE_AUX), self.job.readback(abi.PLANE_DEPTH64))

    def state(self) -> dict:
        return {"shader_fingerprint": self.job.fingerprint(), "scene": self.scene, "cfg": self.cfg,
                "samples_done": self.samples_done, "spd": self.spd, "elapsed_secs": self.elapsed}

    def save(self, directory: str) -> None:
        """A checkpoint in `directory` (replaced atomically)."""
        self.job.checkpoint_save(directory, to_ron(self.state()) + "\n")

    @staticmethod
    def resume(renderer: Renderer, directory: str) -> "RenderSession":
        with open(os.path.join(directory, STATE_FILE), encoding="utf-8") as fh:
            st = from_ron(fh.read())
        scene = scene_from_state(st["scene"])
        cfg = _build(RenderConfig, st["cfg"])
        s = RenderSession(renderer, scene, cfg, int(st.get("spd", 8)))
        done = int(st["samples_done"])
        if done > 0 and st.get("shader_fingerprint") != s.job.fingerprint():
            s.close()
            raise ValueError("this checkpoint was rendered with a different transport or scene; start a new render "
                             "(the saved files have not been changed)")
        s.job.checkpoint_load(directory)
        s.samples_done = done
        s.elapsed = float(st.get("elapsed_secs", 0.0))
        return s

    def export(self, path: str):
        """The export payload (mean colour, depth, normals) and its (bytes, FNV-1a-64 hash)."""
        return self.job.export_write(path)


def export_read(path: str, nbytes: int, digest: int, width: int, height: int):
    """Reads an export payload back, checking its byte count and checksum."""
    lib = abi.load()
    color = np.empty((height, width, 4), np.float32)
    depth = np.empty((height, width), np.float32)
    normals = np.empty((height, width, 4), np.float32)
    abi.check(lib.syf_export_read(path.encode(), nbytes, digest, width, height, color.ctypes.data,
                                  depth.ctypes.data, normals.ctypes.data))
    return color, depth, normals


def fnv1a64(data: bytes, h: int = 0xcbf29ce484222325) -> int:
    """The library's FNV-1a-64 (syf_fnv1a64) of a byte string."""
    buf = np.frombuffer(data, np.uint8) if data else np.zeros(1, np.uint8)
    return int(abi.load().syf_fnv1a64(buf.ctypes.data, len(data), h))
