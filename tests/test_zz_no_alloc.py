"""The allocation-free primary trace (syf_trace_primary_noalloc) against syf_trace_primary.

Through the C ABI on an MI355X: caller-owned device planes and a caller's stream, made with the
HIP runtime the library links; the planes must equal the job path's bit for bit, on two calls in a row (the
second smaller, of another family and sample, so the shared scratch's hit queue is reused at a
different size), and a size above SYF_NOALLOC_MAX_PIXELS is refused with a message naming the
maximum before anything is enqueued.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    from syfrack_amd.render import Renderer
    return Renderer(0)


def _hip():
    """The HIP runtime libsyfrack_hip.so itself links (found in this process's mappings after the
    library is loaded); torch's wheel carries a runtime of its own, which must not be mixed in."""
    from syfrack_amd import abi
    abi.load()
    with open("/proc/self/maps") as fh:
        paths = {ln.split()[-1] for ln in fh if "libamdhip64" in ln and "/torch/" not in ln}
    assert paths, "libamdhip64 not mapped"
    hip = C.CDLL(sorted(paths)[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


def _noalloc(renderer, scene, cfg, sample):
    """One allocation-free trace into planes this test allocates (filled with 0xff bytes first:
    every pixel must be written), on a stream of its own."""
    from syfrack_amd import abi
    hip = _hip()
    n = cfg.width * cfg.height
    sizes = [4 * n, 16 * n, 8 * n, 4 * n]  # depth, aux4, depth64, steps
    ptrs = [C.c_void_p() for _ in sizes]
    stream = C.c_void_p()
    try:
        for p, b in zip(ptrs, sizes):
            assert hip.hipMalloc(C.byref(p), b) == 0
            assert hip.hipMemset(p, 0xff, b) == 0
        assert hip.hipStreamCreate(C.byref(stream)) == 0
        desc = scene.desc(cfg)
        lib = renderer.ctx.lib
        abi.check(lib.syf_trace_primary_noalloc(renderer.ctx.handle, C.byref(desc), sample, *ptrs, stream))
        assert hip.hipStreamSynchronize(stream) == 0
        h, w = cfg.height, cfg.width
        out = [np.empty((h, w), np.float32), np.empty((h, w, 4), np.float32), np.empty((h, w), np.float64),
               np.empty((h, w), np.uint32)]
        for a, p, b in zip(out, ptrs, sizes):
            assert hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, b, 2) == 0  # hipMemcpyDeviceToHost
        return out
    finally:
        if stream.value:
            hip.hipStreamDestroy(stream)
        for p in ptrs:
            if p.value:
                hip.hipFree(p)


def test_noalloc_trace_equals_job_trace_on_two_calls(renderer):
    from syfrack_amd import scene as sc
    big, big_cfg = sc.mandelbulb_primary(640, 360)
    box = sc.Scene(sc.Fractal.bundled("mandelbox"), sc.Camera.orbit((0, 0, 0), 11.0, 28.0, 22.0, 35.0), seed=5)
    box_cfg = sc.RenderConfig(320, 180, 1, 0, 512)
    for scene, cfg, sample in ((big, big_cfg, 0), (box, box_cfg, 1)):
        print(f"noalloc call: {cfg.width}x{cfg.height} family {scene.fractal.family} sample {sample}", flush=True)
        d, aux, d64, st = _noalloc(renderer, scene, cfg, sample)
        ref = renderer.trace_primary(scene, cfg, sample=sample)
        hits = ref.depth >= 0
        assert hits.sum() > 1000
        assert np.array_equal(d, ref.depth)
        assert np.array_equal(d64, ref.depth64)
        assert np.array_equal(aux, ref.normals)
        assert np.array_equal(st, ref.steps)


def test_noalloc_trace_refuses_sizes_above_its_maximum(renderer):
    from syfrack_amd import abi, scene as sc
    scene, cfg = sc.mandelbulb_primary(2048, 1024)
    desc = scene.desc(cfg)
    lib = renderer.ctx.lib
    one = C.c_void_p(256)  # never dereferenced: the size check comes first
    rc = lib.syf_trace_primary_noalloc(renderer.ctx.handle, C.byref(desc), 0, one, one, one, one, None)
    assert rc == abi.SYF_E_INVALID
    msg = lib.syf_last_error().decode()
    assert str(abi.NOALLOC_MAX_PIXELS) in msg and "width x height" in msg
