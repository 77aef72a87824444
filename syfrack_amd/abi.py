"""ctypes mirror of include/syfrack_hip.h and the loader of the HIP library.

The library is built in-tree (``syfrack_amd/libsyfrack_hip.so``).  There is no fallback:
when the library is missing, ``load()`` raises, so nothing on the product path can run on
anything but the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsyfrack_hip.so")

SYF_OK = 0
SYF_E_INVALID = -1
SYF_E_DEVICE = -2
SYF_E_QUEUE = -3
SYF_E_TIMEOUT = -4
SYF_E_ADMISSION = -5
SYF_E_NOT_FOUND = -6
SYF_E_NO_DEVICE = -7

FAMILY_MANDELBULB = 0
FAMILY_MANDELBOX = 1
FAMILY_MENGER = 2
FAMILY_SPHERE = 3
FAMILY_GRAPH = 4

CAMERA_FREE = 0
CAMERA_ORBIT = 1

LIGHT_SUN = 0
LIGHT_AREA = 1
LIGHT_ENV = 2
MAX_LIGHTS = 4

MAX_FINISHES = 8
# FinishPattern (material.rs:177-191) numbered as GpuFinish.detail_kind (uniforms.rs:797-812)
PATTERNS = ("VeinedStone", "Patina", "BrushedMetal", "Basalt", "Ripples", "Rosettes", "CoastalRock", "Sand", "Dust",
            "Concrete", "WeatheredConcrete", "WetConcrete")
MAX_STOPS = 16
COLOR_ITER = 4
COLOR_NORMAL = 5
COLOR_ORBIT = 6
COLOR_NONE = 0xFFFF

PLANE_ACCUM = 0
PLANE_DEPTH = 1
PLANE_AUX = 2
PLANE_DEPTH64 = 3
PLANE_STEPS = 4
PLANE_VOLUME = 5
PLANE_MOMENTS = 6
PLANE_SURFACE = 7

PLANAR_JULIA = 1
PLANAR_MANDELBROT = 2
PLANAR_BURNING_SHIP = 3
PLANAR_FP32 = 0
PLANAR_FP64 = 1
PLANAR_PERTURB = 2
PLANAR_PERTURB64 = 3
PLANAR_SMOOTH = 0
PLANAR_ITER = 1

SERVE_BATCHED = 0
SERVE_PLAIN = 1
REQ_QUEUED = 0
REQ_ACTIVE = 1
REQ_DONE = 2
REQ_CANCELLED = 3

F3 = C.c_float * 3
F2 = C.c_float * 2


class SyfFractal(C.Structure):
    _fields_ = [
        ("family", C.c_uint32),
        ("iterations", C.c_uint32),
        ("bailout", C.c_float),
        ("power", C.c_float),
        ("scale", C.c_float),
        ("min_radius", C.c_float),
        ("fixed_radius", C.c_float),
        ("fold_limit", C.c_float),
        ("offset", F3),
        ("half", F3),
        ("radius", C.c_float),
        ("plane_normal", F3),
        ("plane_offset", C.c_float),
        ("trap_scale", C.c_float),
        ("safety", C.c_float),
        ("bounding_radius_hint", C.c_float),
    ]


class SyfCamera(C.Structure):
    _fields_ = [
        ("mode", C.c_uint32),
        ("position", F3),
        ("target", F3),
        ("roll", C.c_float),
        ("point", F3),
        ("distance", C.c_float),
        ("azimuth", C.c_float),
        ("elevation", C.c_float),
        ("focal_mm", C.c_float),
        ("f_number", C.c_float),
        ("focus_distance", C.c_float),
    ]


class SyfLight(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("position", F3),
        ("normal", F3),
        ("size", F2),
        ("color", F3),
        ("intensity", C.c_float),
        ("zenith", F3),
        ("horizon", F3),
        ("ground", F3),
    ]


class SyfMaterial(C.Structure):
    _fields_ = [
        ("albedo", F3),
        ("roughness", C.c_float),
        ("metallic", C.c_float),
        ("specular", C.c_float),
        ("surface_enabled", C.c_uint32),
        ("surface_slot", C.c_uint32),
        ("surface_roughness", F2),
        ("surface_metallic", F2),
    ]


class SyfFinish(C.Structure):
    _fields_ = [
        ("albedo", F3),
        ("roughness", C.c_float),
        ("metallic", C.c_float),
        ("specular", C.c_float),
        ("emission", F3),
        ("oren_nayar", C.c_uint32),
        ("pattern", C.c_uint32),
        ("seed", C.c_uint32),
        ("scale", C.c_float),
        ("strength", C.c_float),
        ("bump", C.c_float),
        ("roughness_delta", C.c_float),
        ("tint", F3),
    ]


class SyfColoring(C.Structure):
    _fields_ = [
        ("slot", C.c_uint32),
        ("stop_count", C.c_uint32),
        ("cyclic", C.c_uint32),
        ("smooth", C.c_uint32),
        ("positioned", C.c_uint32),
        ("frequency", C.c_float),
        ("phase", C.c_float),
        ("gamma", C.c_float),
        ("stops", (C.c_float * 4) * MAX_STOPS),
        ("weights", C.c_float * 4),
        ("iteration_weight", C.c_float),
        ("normal_weight", C.c_float),
        ("logarithmic", C.c_float),
    ]


class SyfGround(C.Structure):
    _fields_ = [
        ("enabled", C.c_uint32),
        ("height", C.c_float),
        ("finish", C.c_uint32),
    ]


class SyfFog(C.Structure):
    _fields_ = [
        ("density", C.c_float),
        ("color", F3),
        ("height", C.c_float),
        ("falloff", C.c_float),
        ("anisotropy", C.c_float),
    ]


class SyfJobDesc(C.Structure):
    _fields_ = [
        ("fractal", SyfFractal),
        ("camera", SyfCamera),
        ("material", SyfMaterial),
        ("lights", SyfLight * MAX_LIGHTS),
        ("light_count", C.c_uint32),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("max_bounces", C.c_uint32),
        ("max_steps", C.c_uint32),
        ("pixel_tolerance", C.c_float),
        ("cavity_darkness", C.c_float),
        ("march_math", C.c_uint32),
        ("seed", C.c_uint64),
        ("ground", SyfGround),
        ("coloring", SyfColoring),
        ("finish_count", C.c_uint32),
        ("finishes", SyfFinish * MAX_FINISHES),
        ("fog", SyfFog),
    ]


class SyfStats(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64),
        ("primary_rays", C.c_uint64),
        ("hits", C.c_uint64),
        ("de_evals", C.c_uint64),
        ("de_iterations", C.c_uint64),
        ("de64_evals", C.c_uint64),
        ("de64_iterations", C.c_uint64),
        ("steps", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("march_ms", C.c_double),
        ("hit_de_evals", C.c_uint64),
        ("hit_de_iterations", C.c_uint64),
        ("tail_hits", C.c_uint64),
        ("tail_hit_de_evals", C.c_uint64),
        ("tail_hit_de_iterations", C.c_uint64),
        ("tail_de64_evals", C.c_uint64),
        ("tail_de64_iterations", C.c_uint64),
    ]


class SyfPlanarDesc(C.Structure):
    _fields_ = [
        ("family", C.c_uint32),
        ("mode", C.c_uint32),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("iterations", C.c_uint32),
        ("max_references", C.c_uint32),
        ("center", C.c_double * 2),
        ("span", C.c_double),
        ("julia", C.c_double * 2),
        ("bailout", C.c_float),
        ("rotation", C.c_float),
    ]


class SyfServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_uint64),
        ("mode", C.c_uint32),
        ("max_active", C.c_uint32),
        ("quantum_samples", C.c_uint32),
        ("batch_target_us", C.c_uint32),
        ("grid_blocks", C.c_uint32),
        ("quiet_target_us", C.c_uint32),
        ("path_steps", C.c_uint32),
        ("tile_width", C.c_uint32),
    ]


QUIET_UNBOUNDED = 0xFFFFFFFF


class SyfRequestInfo(C.Structure):
    _fields_ = [
        ("state", C.c_uint32),
        ("samples_done", C.c_uint32),
        ("spp", C.c_uint32),
        ("reserved", C.c_uint32),
        ("footprint_bytes", C.c_uint64),
        ("submit_ms", C.c_double),
        ("start_ms", C.c_double),
        ("done_ms", C.c_double),
    ]


class SyfServerCounters(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("cancelled", C.c_uint64),
        ("launches", C.c_uint64),
        ("rays", C.c_uint64),
        ("primary_rays", C.c_uint64),
        ("hbm_in_use", C.c_uint64),
        ("hbm_peak", C.c_uint64),
        ("device_ms", C.c_double),
        ("wall_ms", C.c_double),
        ("quantum_samples", C.c_uint32),
        ("batch_target_us", C.c_uint32),
        ("grid_blocks", C.c_uint32),
        ("reserved", C.c_uint32),
        ("family_grid_blocks", C.c_uint32 * 6),
        ("family_path_steps", C.c_uint32 * 6),
        ("autotune_ms", C.c_double),
        ("family_tile_width", C.c_uint32 * 6),
        ("family_quiet_quantum", C.c_uint32 * 6),
    ]


class SyfGraphNode(C.Structure):
    _fields_ = [
        ("id", C.c_uint32),
        ("kind", C.c_uint32),
        ("enabled", C.c_uint32),
        ("flags", C.c_uint32),
        ("v", C.c_float * 6),
        ("u", C.c_uint32 * 3),
        ("body_first", C.c_uint32),
        ("body_count", C.c_uint32),
    ]


class SyfGraphLink(C.Structure):
    _fields_ = [("from_node", C.c_uint32), ("from_port", C.c_uint32), ("to_node", C.c_uint32), ("to_port", C.c_uint32)]


class SyfGraph(C.Structure):
    _fields_ = [
        ("nodes", C.c_void_p),
        ("node_count", C.c_uint32),
        ("link_count", C.c_uint32),
        ("links", C.c_void_p),
        ("bodies", C.c_void_p),
        ("body_len", C.c_uint32),
        ("safety", C.c_float),
        ("bounding_radius_hint", C.c_float),
        ("reserved", C.c_uint32),
    ]


NOALLOC_MAX_PIXELS = 1048576  # SYF_NOALLOC_MAX_PIXELS
CONTROL_QUANTUM = 64  # SYF_CONTROL_QUANTUM
MATH_FAST = 0  # SYF_MATH_FAST
MATH_EXACT = 1  # SYF_MATH_EXACT
GRAPH_SPECIALISED = 0  # SYF_GRAPH_SPECIALISED
GRAPH_INTERPRETED = 1  # SYF_GRAPH_INTERPRETED

# Every symbol the header declares, with its ctypes signature.
_P = C.c_void_p
SIGNATURES = {
    "syf_last_error": (C.c_char_p, []),
    "syf_abi_version": (C.c_int, []),
    "syf_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "syf_ctx_destroy": (C.c_int, [_P]),
    "syf_job_create": (C.c_int, [_P, C.POINTER(SyfJobDesc), C.POINTER(_P)]),
    "syf_job_destroy": (C.c_int, [_P]),
    "syf_job_create_graph": (C.c_int, [_P, C.POINTER(SyfJobDesc), C.POINTER(SyfGraph), C.c_uint32, C.POINTER(_P)]),
    "syf_graph_info": (C.c_int, [C.POINTER(SyfGraph), C.POINTER(C.c_float), C.POINTER(C.c_uint32), _P,
                                 C.POINTER(C.c_uint32)]),
    "syf_graph_source": (C.c_int, [C.POINTER(SyfGraph), C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "syf_job_clear": (C.c_int, [_P]),
    "syf_job_sync": (C.c_int, [_P]),
    "syf_job_stats": (C.c_int, [_P, C.POINTER(SyfStats)]),
    "syf_de_probe": (C.c_int, [_P, _P, C.c_size_t, _P, _P]),
    "syf_trace_primary": (C.c_int, [_P, C.c_uint32]),
    "syf_trace_primary_noalloc": (C.c_int, [_P, C.POINTER(SyfJobDesc), C.c_uint32, _P, _P, _P, _P, _P]),
    "syf_render_tile": (C.c_int, [_P] + [C.c_uint32] * 6),
    "syf_render_wavefront": (C.c_int, [_P] + [C.c_uint32] * 6),
    "syf_readback": (C.c_int, [_P, C.c_uint32, _P, C.c_size_t]),
    "syf_upload": (C.c_int, [_P, C.c_uint32, _P, C.c_size_t]),
    "syf_job_fingerprint": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "syf_fnv1a64": (C.c_uint64, [_P, C.c_size_t, C.c_uint64]),
    "syf_checkpoint_save": (C.c_int, [_P, C.c_char_p, C.c_char_p]),
    "syf_checkpoint_load": (C.c_int, [_P, C.c_char_p]),
    "syf_export_write": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "syf_export_read": (C.c_int, [C.c_char_p, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, _P, _P, _P]),
    "syf_control_create": (C.c_int, [_P, C.POINTER(SyfJobDesc), _P, C.c_uint32, C.POINTER(_P)]),
    "syf_control_destroy": (C.c_int, [_P]),
    "syf_control_set_scene": (C.c_int, [_P, C.POINTER(SyfJobDesc), _P]),
    "syf_control_tick": (C.c_int, [_P]),
    "syf_control_wait": (C.c_int, [_P, C.c_double]),
    "syf_control_progress": (C.c_int, [_P] + [C.POINTER(C.c_uint32)] * 5),
    "syf_control_readback": (C.c_int, [_P, C.c_uint32, _P, C.c_size_t]),
    "syf_control_pick": (C.c_int, [_P, C.c_float, C.c_float, C.POINTER(C.c_uint32), C.POINTER(C.c_float), _P]),
    "syf_control_probe": (C.c_int, [_P, _P, _P, C.c_size_t, _P, _P]),
    "syf_planar_create": (C.c_int, [_P, C.POINTER(SyfPlanarDesc), C.POINTER(_P)]),
    "syf_planar_destroy": (C.c_int, [_P]),
    "syf_planar_run": (C.c_int, [_P]),
    "syf_planar_sync": (C.c_int, [_P]),
    "syf_planar_readback": (C.c_int, [_P, C.c_uint32, _P, C.c_size_t]),
    "syf_planar_stats": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "syf_planar_glitches": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    "syf_server_create": (C.c_int, [_P, C.POINTER(SyfServerConfig), C.POINTER(_P)]),
    "syf_server_destroy": (C.c_int, [_P]),
    "syf_server_submit": (C.c_int, [_P, C.POINTER(SyfJobDesc), C.c_uint32, C.POINTER(C.c_uint64)]),
    "syf_server_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(SyfRequestInfo)]),
    "syf_server_wait": (C.c_int, [_P, C.c_uint64, C.c_uint32]),
    "syf_server_wait_all": (C.c_int, [_P, C.c_uint32]),
    "syf_server_result": (C.c_int, [_P, C.c_uint64, C.c_uint32, _P, C.c_size_t]),
    "syf_server_release": (C.c_int, [_P, C.c_uint64]),
    "syf_server_cancel": (C.c_int, [_P, C.c_uint64]),
    "syf_server_stats": (C.c_int, [_P, C.POINTER(SyfServerCounters)]),
}


class SyfError(RuntimeError):
    """A negative status from the library, with its syf_last_error() text."""

    def __init__(self, code: int, message: str):
        super().__init__(f"syfrack_hip error {code}: {message}")
        self.code = code
        self.message = message


_LIB = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library (once).  Raises when it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the HIP path has no fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(code: int) -> None:
    if code != SYF_OK:
        msg = load().syf_last_error().decode("utf-8", "replace")
        raise SyfError(code, msg)


def header_symbols(header_path: str) -> list:
    """Names of the functions declared in a C header (``int syf_x(`` style)."""
    import re
    text = open(header_path, encoding="utf-8").read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(syf_\w+)\s*\(", text, re.M)))
