"""2D escape time over the C ABI: PlanarFractal::sample (planar.rs:77-194) for the
parameter-plane families in FP32 (reference precision), FP64 (bit-exact counts) and deep-zoom
perturbation (FP64 reference orbits + FP32 deltas, the glitches left after the reference budget
taking FP64 deltas with rebasing; or FP64 rebasing for every pixel)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi


@dataclass
class PlanarView:
    family: int = abi.PLANAR_MANDELBROT
    mode: int = abi.PLANAR_FP64
    width: int = 512
    height: int = 512
    iterations: int = 1000
    center: tuple = (-0.75, 0.0)
    span: float = 3.0
    julia: tuple = (-0.74543, 0.11301)
    bailout: float = 16.0
    rotation: float = 0.0
    max_references: int = 0  # PERTURB: FP32-delta references before the FP64 rebasing fallback (0: 8)

    def c_struct(self) -> abi.SyfPlanarDesc:
        d = abi.SyfPlanarDesc()
        d.family = self.family
        d.mode = self.mode
        d.width = self.width
        d.height = self.height
        d.iterations = self.iterations
        d.center[:] = list(self.center)
        d.span = self.span
        d.julia[:] = list(self.julia)
        d.bailout = self.bailout
        d.rotation = self.rotation
        d.max_references = self.max_references
        return d


class Planar:
    """A device-resident 2D escape-time image (syf_planar)."""

    def __init__(self, ctx, view: PlanarView):
        self.ctx = ctx  # the context must outlive the image
        self.lib = ctx.lib
        self.view = view
        self.desc = view.c_struct()
        h = C.c_void_p()
        abi.check(self.lib.syf_planar_create(ctx.handle, C.byref(self.desc), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self.lib.syf_planar_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self):
        abi.check(self.lib.syf_planar_run(self.handle))

    def sync(self):
        abi.check(self.lib.syf_planar_sync(self.handle))

    def readback(self):
        n = self.view.width * self.view.height
        sm = np.empty(n, np.float32)
        it = np.empty(n, np.uint32)
        abi.check(self.lib.syf_planar_readback(self.handle, abi.PLANAR_SMOOTH, sm.ctypes.data_as(C.c_void_p), sm.nbytes))
        abi.check(self.lib.syf_planar_readback(self.handle, abi.PLANAR_ITER, it.ctypes.data_as(C.c_void_p), it.nbytes))
        return it.reshape(self.view.height, self.view.width), sm.reshape(self.view.height, self.view.width)

    def stats(self):
        a, b, ms = C.c_uint64(), C.c_uint64(), C.c_double()
        abi.check(self.lib.syf_planar_stats(self.handle, C.byref(a), C.byref(b), C.byref(ms)))
        refs, fallback, rebasings = C.c_uint32(), C.c_uint32(), C.c_uint64()
        abi.check(self.lib.syf_planar_glitches(self.handle, C.byref(refs), C.byref(fallback), C.byref(rebasings)))
        return {"iterations": a.value, "rebases": b.value, "kernel_ms": ms.value, "references": refs.value,
                "fallback_pixels": fallback.value, "rebasings": rebasings.value}


def boundary_centre(ctx, family: int, seed: int, iterations: int, grid: int = 256):
    """A view centre in the boundary band: on a coarse FP64 grid over the family's main view,
    the points whose escape count lies in [iterations/2, iterations), drawn with numpy's
    default_rng(seed)."""
    base = {abi.PLANAR_MANDELBROT: ((-0.75, 0.0), 3.0), abi.PLANAR_BURNING_SHIP: ((-0.5, -0.5), 3.5)}[family]
    v = PlanarView(family, abi.PLANAR_FP64, grid, grid, iterations, base[0], base[1], bailout=16.0)
    p = Planar(ctx, v)
    p.run()
    it, _ = p.readback()
    p.close()
    ys, xs = np.nonzero((it >= iterations // 2) & (it < iterations))
    if len(xs) == 0:
        ys, xs = np.nonzero(it == np.max(it[it < iterations]))
    rng = np.random.default_rng(seed)
    k = int(rng.integers(len(xs)))
    fx = ((xs[k] + 0.5) / grid - 0.5) * base[1]
    fy = (0.5 - (ys[k] + 0.5) / grid) * base[1]
    return (base[0][0] + fx, base[0][1] + fy)
