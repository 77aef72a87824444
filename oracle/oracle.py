"""ctypes binding of the oracle library (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py, always as the checker.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
NATIVE_PATH = os.path.join(HERE, "_build", "liboracle_native.so")
_LIB = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load_native() -> C.CDLL:
    """The CPU-baseline build (-O3 -march=native) of the same source, built on the calling host
    (make native) when absent."""
    if not os.path.exists(NATIVE_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "native"], check=True)
    return _bind(C.CDLL(NATIVE_PATH))


def load() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        build()
    _LIB = _bind(C.CDLL(LIB_PATH))
    return _LIB


def _bind(lib: C.CDLL) -> C.CDLL:
    P = C.c_void_p
    lib.orc_pcg.restype = C.c_uint32
    lib.orc_pcg.argtypes = [C.c_uint32]
    lib.orc_rng_init.restype = C.c_uint32
    lib.orc_rng_init.argtypes = [C.c_uint32] * 4
    lib.orc_fold_seed.restype = C.c_uint32
    lib.orc_fold_seed.argtypes = [C.c_uint64]
    lib.orc_camera_basis.restype = None
    lib.orc_camera_basis.argtypes = [P, P]
    lib.orc_camera_rays.restype = None
    lib.orc_camera_rays.argtypes = [P, C.c_uint32, P, P]
    lib.orc_de_probe.restype = C.c_int
    lib.orc_de_probe.argtypes = [P, C.c_int, P, C.c_size_t, P, P, P, P]
    lib.orc_spherical_power.restype = None
    lib.orc_spherical_power.argtypes = [P, C.c_size_t, C.c_uint32, P]
    lib.orc_trace_primary.restype = C.c_int
    lib.orc_trace_primary.argtypes = [P, C.c_uint32, C.c_int] + [C.c_uint32] * 4 + [P] * 5
    lib.orc_first_root_bruteforce.restype = C.c_double
    lib.orc_first_root_bruteforce.argtypes = [P, P, P, C.c_double, C.c_double]
    lib.orc_planar_sample_f32.restype = C.c_uint32
    lib.orc_planar_sample_f32.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, P, C.c_float, C.c_float, C.c_float,
                                          C.c_float, C.c_float, P]
    lib.orc_planar_offsets.restype = None
    lib.orc_planar_offsets.argtypes = [P, P, P]
    lib.orc_escape_f64.restype = None
    lib.orc_escape_f64.argtypes = [P] + [C.c_uint32] * 4 + [P, P]
    lib.orc_escape_f128.restype = None
    lib.orc_escape_f128.argtypes = [P, P, P, C.c_size_t, C.c_double, P, P]
    lib.orc_ramp_sample.restype = None
    lib.orc_ramp_sample.argtypes = [P, P, C.c_size_t, P]
    lib.orc_fog_optical_depth.restype = None
    lib.orc_fog_optical_depth.argtypes = [P, P, P, P, C.c_size_t, P]
    lib.orc_fog_free_flight.restype = None
    lib.orc_fog_free_flight.argtypes = [P, P, P, C.c_float, C.c_size_t, P]
    lib.orc_volume_coordinate.restype = C.c_float
    lib.orc_volume_coordinate.argtypes = [C.c_uint32] * 3
    lib.orc_finish_relief.restype = None
    lib.orc_finish_relief.argtypes = [C.c_uint32, C.c_uint32, C.c_float, P, C.c_int, C.c_float, P]
    lib.orc_hg_phase.restype = C.c_float
    lib.orc_hg_phase.argtypes = [C.c_float, C.c_float]
    lib.orc_hg_sample.restype = None
    lib.orc_hg_sample.argtypes = [C.c_float, P, P, C.c_size_t, P]
    if hasattr(lib, "orc_render"):
        lib.orc_render.restype = C.c_int
        lib.orc_render.argtypes = [P] + [C.c_uint32] * 6 + [P, P, P, P, P]
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def de_probe(fractal, points: np.ndarray, precision: int = 32):
    """Distance, traps (n, 4), iterations and body executions at points (n, 3 or 4)."""
    lib = load()
    pts = np.zeros((len(points), 4), np.float32)
    pts[:, :3] = np.asarray(points, np.float32)[:, :3]
    n = len(pts)
    dist = np.zeros(n, np.float64)
    traps = np.zeros((n, 4), np.float64)
    iters = np.zeros(n, np.uint32)
    runs = np.zeros(n, np.uint32)
    rc = lib.orc_de_probe(C.byref(fractal), precision, _ptr(pts), n, _ptr(dist), _ptr(traps), _ptr(iters), _ptr(runs))
    assert rc == 0
    return dist, traps, iters, runs


def spherical_power(points: np.ndarray, power: int):
    lib = load()
    pts = np.zeros((len(points), 4), np.float32)
    pts[:, :3] = np.asarray(points, np.float32)[:, :3]
    out = np.zeros((len(pts), 4), np.float32)
    lib.orc_spherical_power(_ptr(pts), len(pts), power, _ptr(out))
    return out[:, :3], out[:, 3]


def trig_power(points: np.ndarray, power: float, precision: int = 32):
    """The Power node's trigonometric branch forced at every point: (z^n (n, 3), dr (n,))."""
    lib = load()
    pts = np.zeros((len(points), 4), np.float32)
    pts[:, :3] = np.asarray(points, np.float32)[:, :3]
    out = np.zeros((len(pts), 4), np.float64)
    lib.orc_trig_power.restype = None
    lib.orc_trig_power.argtypes = [C.c_void_p, C.c_size_t, C.c_float, C.c_int, C.c_void_p]
    lib.orc_trig_power(_ptr(pts), len(pts), power, precision, _ptr(out))
    return out[:, :3], out[:, 3]


def camera_basis(desc) -> np.ndarray:
    out = np.zeros(16, np.float32)
    load().orc_camera_basis(C.byref(desc), _ptr(out))
    return out


def camera_rays(desc, sample: int):
    n = desc.width * desc.height
    ro = np.zeros((n, 3), np.float32)
    rd = np.zeros((n, 3), np.float32)
    load().orc_camera_rays(C.byref(desc), sample, _ptr(ro), _ptr(rd))
    return ro, rd


def trace_primary(desc, sample: int = 0, precision: int = 32, rect=None, lib=None):
    """Primary trace of a pixel rectangle (x0, y0, w, h); default the whole image.  ``lib``: another
    build of the oracle (load_native()), else the test build."""
    lib = lib or load()
    x0, y0, w, h = rect if rect is not None else (0, 0, desc.width, desc.height)
    n = w * h
    depth = np.zeros(n, np.float64)
    depth64 = np.zeros(n, np.float64)
    aux = np.zeros((n, 4), np.float32)
    steps = np.zeros(n, np.uint32)
    counts = np.zeros(5, np.uint64)
    rc = lib.orc_trace_primary(C.byref(desc), sample, precision, x0, y0, w, h, _ptr(depth), _ptr(depth64), _ptr(aux),
                               _ptr(steps), _ptr(counts))
    assert rc == 0
    return {"depth": depth.reshape(h, w), "depth64": depth64.reshape(h, w), "aux": aux.reshape(h, w, 4),
            "steps": steps.reshape(h, w), "counts": dict(zip(
                ["de_evals", "de_iterations", "steps", "de64_evals", "de64_iterations"], counts.tolist()))}


def first_root(fractal, ro, rd, t_max: float, dt: float) -> float:
    a = np.asarray(ro, np.float64)
    b = np.asarray(rd, np.float64)
    return load().orc_first_root_bruteforce(C.byref(fractal), _ptr(a), _ptr(b), t_max, dt)


def planar_sample_f32(family, power, iterations, julia, bailout, trap_radius, stripes, x, y):
    out = np.zeros(5, np.float32)
    j = np.asarray(julia, np.float32)
    it = load().orc_planar_sample_f32(family, power, iterations, _ptr(j), bailout, trap_radius, stripes, x, y,
                                      _ptr(out))
    return int(it), out


def planar_offsets(desc):
    fx = np.zeros(desc.width, np.float64)
    fy = np.zeros(desc.height, np.float64)
    load().orc_planar_offsets(C.byref(desc), _ptr(fx), _ptr(fy))
    return fx, fy


def escape_f64(desc, rect=None):
    x0, y0, w, h = rect if rect is not None else (0, 0, desc.width, desc.height)
    iters = np.zeros(w * h, np.uint32)
    smooth = np.zeros(w * h, np.float32)
    load().orc_escape_f64(C.byref(desc), x0, y0, w, h, _ptr(iters), _ptr(smooth))
    return iters.reshape(h, w), smooth.reshape(h, w)


def escape_f128(desc, px: np.ndarray, py: np.ndarray, nudge: float = 0.0):
    """binary128 counts at c = centre + f32 offset (+ nudge f32 ulps of the offset)."""
    px = np.ascontiguousarray(px, np.uint32)
    py = np.ascontiguousarray(py, np.uint32)
    iters = np.zeros(len(px), np.uint32)
    smooth = np.zeros(len(px), np.float32)
    load().orc_escape_f128(C.byref(desc), _ptr(px), _ptr(py), C.c_size_t(len(px)), C.c_double(nudge), _ptr(iters),
                           _ptr(smooth))
    return iters, smooth


def escape_f128_exact(desc, px: np.ndarray, py: np.ndarray, nudge: float = 0.0):
    """binary128 counts at c = centre + the pixel's FP64 offset (+ nudge FP64 ulps of it)."""
    px = np.ascontiguousarray(px, np.uint32)
    py = np.ascontiguousarray(py, np.uint32)
    iters = np.zeros(len(px), np.uint32)
    smooth = np.zeros(len(px), np.float32)
    lib = load()
    lib.orc_escape_f128_exact.restype = None
    lib.orc_escape_f128_exact(C.byref(desc), _ptr(px), _ptr(py), C.c_size_t(len(px)), C.c_double(nudge),
                              _ptr(iters), _ptr(smooth))
    return iters, smooth


def reference_orbit(desc, ox: float = 0.0, oy: float = 0.0, cap: int = 100000, quad: bool = False):
    """The deep-zoom reference orbit of C = centre + (ox, oy): double-double iteration rounded
    to FP64 (quad=False, the product's orbit), or binary128 iteration rounded to FP64."""
    zx = np.zeros(cap, np.float64)
    zy = np.zeros(cap, np.float64)
    lib = load()
    fn = lib.orc_reference_orbit_f128 if quad else lib.orc_reference_orbit
    fn.restype = C.c_uint32
    last = fn(C.byref(desc), C.c_double(ox), C.c_double(oy), _ptr(zx), _ptr(zy), C.c_uint32(cap))
    return zx[:last + 1], zy[:last + 1]


def perturb64(desc, px: np.ndarray, py: np.ndarray):
    """The FP64 rebasing perturbation loop (PERTURB64) at listed pixels: counts, smooth."""
    px = np.ascontiguousarray(px, np.uint32)
    py = np.ascontiguousarray(py, np.uint32)
    iters = np.zeros(len(px), np.uint32)
    smooth = np.zeros(len(px), np.float32)
    lib = load()
    lib.orc_perturb64.restype = None
    lib.orc_perturb64(C.byref(desc), _ptr(px), _ptr(py), C.c_size_t(len(px)), _ptr(iters), _ptr(smooth))
    return iters, smooth


def control_trace(desc, rect=None, max_evals: int = 1 << 20):
    """The control machine's image of the job's field (graph loaded with graph_use): accum
    (h, w, 4), depth (h, w), aux (h, w, 4), status (h, w; 0 done, 1 undefined distance,
    2 over max_evals evaluations)."""
    x0, y0, w, h = rect if rect is not None else (0, 0, desc.width, desc.height)
    acc = np.zeros((h, w, 4), np.float32)
    dep = np.zeros((h, w), np.float32)
    aux = np.zeros((h, w, 4), np.float32)
    st = np.zeros((h, w), np.uint32)
    lib = load()
    lib.orc_control_trace.restype = C.c_int
    rc = lib.orc_control_trace(C.byref(desc), C.c_uint32(x0), C.c_uint32(y0), C.c_uint32(w), C.c_uint32(h),
                               C.c_uint32(max_evals), _ptr(acc), _ptr(dep), _ptr(aux), _ptr(st))
    if rc != 0:
        raise RuntimeError("orc_control_trace failed")
    return acc, dep, aux, st


def escape_f64_points(desc, px: np.ndarray, py: np.ndarray):
    """FP64 iteration counts at the points escape_f128 evaluates (centre + f32 offset)."""
    px = np.ascontiguousarray(px, np.uint32)
    py = np.ascontiguousarray(py, np.uint32)
    iters = np.zeros(len(px), np.uint32)
    lib = load()
    lib.orc_escape_f64_points.restype = None
    lib.orc_escape_f64_points(C.byref(desc), _ptr(px), _ptr(py), C.c_size_t(len(px)), _ptr(iters))
    return iters


def ramp_sample(coloring, values) -> np.ndarray:
    """palette_sample of a syf_coloring's ramp at each value: (n, 3) linear rgb."""
    v = np.ascontiguousarray(np.asarray(values, np.float32).ravel())
    out = np.zeros((len(v), 3), np.float32)
    load().orc_ramp_sample(C.byref(coloring), _ptr(v), len(v), _ptr(out))
    return out


def render(desc, rect, sample_start: int, sample_count: int, session: bool = False):
    """Path-traced accumulation (float4 per pixel), first-hit depth, aux and FP64-refined
    first-hit depth of a rectangle; with session, also the VOLUME, MOMENTS and SURFACE planes
    (3, h, w, 4) accumulated from zero."""
    lib = load()
    x0, y0, w, h = rect
    accum = np.zeros((h, w, 4), np.float32)
    depth = np.zeros((h, w), np.float32)
    aux = np.zeros((h, w, 4), np.float32)
    depth64 = np.zeros((h, w), np.float64)
    planes = np.zeros((3, h, w, 4), np.float32) if session else None
    rc = lib.orc_render(C.byref(desc), x0, y0, w, h, sample_start, sample_count, _ptr(accum), _ptr(depth), _ptr(aux),
                        _ptr(depth64), _ptr(planes) if session else None)
    assert rc == 0
    if session:
        return accum, depth, aux, depth64, planes
    return accum, depth, aux, depth64


def fog_optical_depth(desc, ro, rd, dist) -> np.ndarray:
    ro = np.ascontiguousarray(np.asarray(ro, np.float32).reshape(-1, 3))
    rd = np.ascontiguousarray(np.asarray(rd, np.float32).reshape(-1, 3))
    dist = np.ascontiguousarray(np.broadcast_to(np.asarray(dist, np.float32), (len(ro),)))
    out = np.zeros(len(ro), np.float32)
    load().orc_fog_optical_depth(C.byref(desc), _ptr(ro), _ptr(rd), _ptr(dist), len(ro), _ptr(out))
    return out


def fog_free_flight(desc, ro, rd, t_end: float, n: int) -> np.ndarray:
    ro = np.ascontiguousarray(np.asarray(ro, np.float32))
    rd = np.ascontiguousarray(np.asarray(rd, np.float32))
    out = np.zeros(n, np.float32)
    load().orc_fog_free_flight(C.byref(desc), _ptr(ro), _ptr(rd), C.c_float(t_end), n, _ptr(out))
    return out


def volume_coordinate(sample_index: int, dimension: int, stream: int) -> float:
    return load().orc_volume_coordinate(sample_index, dimension, stream & 0xFFFFFFFF)


def finish_relief(pattern: int, seed: int, scale: float, points, footprint: float) -> np.ndarray:
    """finish_eval (finish_detail.wgsl:38-151) at each point: rows (slope x, y, z, height)."""
    p = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
    out = np.zeros((len(p), 4), np.float32)
    load().orc_finish_relief(pattern, seed, C.c_float(scale), p.ctypes.data, len(p), C.c_float(footprint),
                             out.ctypes.data)
    return out


def hg_phase(cos_theta: float, g: float) -> float:
    return load().orc_hg_phase(C.c_float(cos_theta), C.c_float(g))


def hg_sample(g: float, direction, r12) -> np.ndarray:
    d = np.ascontiguousarray(np.asarray(direction, np.float32))
    r = np.ascontiguousarray(np.asarray(r12, np.float32).reshape(-1, 2))
    out = np.zeros((len(r), 3), np.float32)
    load().orc_hg_sample(C.c_float(g), _ptr(d), _ptr(r), len(r), _ptr(out))
    return out


def planar_f32(desc, rect=None):
    x0, y0, w, h = rect if rect is not None else (0, 0, desc.width, desc.height)
    iters = np.zeros(w * h, np.uint32)
    smooth = np.zeros(w * h, np.float32)
    lib = load()
    lib.orc_planar_f32.restype = None
    lib.orc_planar_f32.argtypes = [C.c_void_p] + [C.c_uint32] * 4 + [C.c_void_p, C.c_void_p]
    lib.orc_planar_f32(C.byref(desc), x0, y0, w, h, _ptr(iters), _ptr(smooth))
    return iters.reshape(h, w), smooth.reshape(h, w)


# ---------------------------------------------------------------- node graphs

def graph_use(graph) -> None:
    """Loads a graph (syfrack_amd.graph.Graph) for graph_probe and for jobs of the graph family."""
    lib = load()
    lib.orc_graph_use.restype = C.c_int
    lib.orc_graph_use.argtypes = [C.c_void_p]
    cg = graph.c_struct()
    if lib.orc_graph_use(C.addressof(cg.s)) != 0:
        raise ValueError("oracle: the graph has an unknown node id or a cycle")


def graph_props(radius: float = None):
    """(step scale, iteration product, trap scales, Lipschitz factor at radius, plan) of the loaded
    graph; radius defaults to nothing meaningful (pass the bounding hint for the step scale's own)."""
    lib = load()
    lib.orc_graph_props.restype = C.c_int
    lib.orc_graph_props.argtypes = [C.c_float] + [C.c_void_p] * 5 + [C.c_uint32, C.c_void_p]
    step, lip = C.c_float(), C.c_float()
    prod, plen = C.c_uint32(), C.c_uint32()
    scales = (C.c_float * 4)()
    plan = (C.c_uint32 * 4096)()
    if lib.orc_graph_props(0.0 if radius is None else radius, C.byref(step), C.byref(prod), scales, C.byref(lip),
                           plan, 4096, C.byref(plen)) != 0:
        raise ValueError("oracle: no graph loaded")
    return step.value, prod.value, tuple(scales), lip.value, list(plan[:plen.value])


def graph_probe(points: np.ndarray, precision: int = 32):
    """The loaded graph at points (n, 3): distance, traps (n, 4), iterations, materials, body runs."""
    lib = load()
    lib.orc_graph_probe.restype = C.c_int
    lib.orc_graph_probe.argtypes = [C.c_int, C.c_void_p, C.c_size_t] + [C.c_void_p] * 5
    p = np.ascontiguousarray(points, np.float32)
    n = p.shape[0]
    xyz4 = np.zeros((n, 4), np.float32)
    xyz4[:, :3] = p
    dist = np.empty(n, np.float64)
    traps = np.empty((n, 4), np.float64)
    iters = np.empty(n, np.uint32)
    mats = np.empty(n, np.uint32)
    runs = np.empty(n, np.uint32)
    if lib.orc_graph_probe(precision, _ptr(xyz4), n, _ptr(dist), _ptr(traps), _ptr(iters), _ptr(mats), _ptr(runs)) != 0:
        raise ValueError("oracle: no graph loaded or bad precision")
    return dist, traps, iters, mats, runs


def graph_point(node_id: int, points: np.ndarray, precision: int = 32):
    """One point node of the loaded graph applied to (p, dr 1) at points (n, 3): (p' (n, 3), dr (n,))."""
    lib = load()
    lib.orc_graph_point.restype = C.c_int
    lib.orc_graph_point.argtypes = [C.c_uint32, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p]
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    xyz4 = np.zeros((len(p), 4), np.float32)
    xyz4[:, :3] = p
    out = np.empty((len(p), 4), np.float64)
    if lib.orc_graph_point(node_id, precision, _ptr(xyz4), len(p), _ptr(out)) != 0:
        raise ValueError(f"oracle: no node {node_id} in the loaded graph")
    return out[:, :3], out[:, 3]
