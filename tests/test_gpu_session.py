"""Sessions on the device: radiance layers, checkpoint and resume, export payload.

  * the session planes of pathtrace.wgsl:1243-1266 on this path (no participating medium):
    SURFACE equals ACCUM bit for bit, VOLUME and MOMENTS are (0, 0, 0, samples);
  * a session saved after some samples, resumed into a fresh job and finished, equals one
    rendered without interruption bit for bit (every pixel adds its samples in ascending order);
  * a checkpoint of another scene is refused by its transport fingerprint, and a directory with
    a foreign file is never replaced (session.rs:48-65);
  * the export payload holds the readback's mean colour, depth and normals, and its checksum is
    the FNV-1a-64 of those bytes.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    from syfrack_amd.render import Renderer
    return Renderer(0)


def _scene():
    from syfrack_amd import scene as sc
    scene, cfg = sc.mandelbox_pathtrace(72, 40, 6)
    cfg.tile_size = 32
    return scene, cfg


def test_session_planes_follow_accum(renderer):
    from syfrack_amd import abi
    from syfrack_amd.session import RenderSession
    scene, cfg = _scene()
    s = RenderSession(renderer, scene, cfg, samples_per_step=3)
    try:
        while not s.done():
            s.step()
        acc = s.job.readback(abi.PLANE_ACCUM)
        assert np.all(acc[..., 3] == 6)
        assert np.array_equal(s.job.readback(abi.PLANE_SURFACE), acc)
        for p in (abi.PLANE_VOLUME, abi.PLANE_MOMENTS):
            v = s.job.readback(p)
            assert np.all(v[..., :3] == 0.0) and np.array_equal(v[..., 3], acc[..., 3])
    finally:
        s.close()


def test_checkpoint_resume_equals_uninterrupted_render(renderer, tmp_path):
    from syfrack_amd import abi
    from syfrack_amd.session import RenderSession
    scene, cfg = _scene()
    whole = RenderSession(renderer, scene, cfg, samples_per_step=2)
    while not whole.done():
        whole.step()
    ref = [whole.job.readback(p) for p in (abi.PLANE_ACCUM, abi.PLANE_DEPTH, abi.PLANE_AUX, abi.PLANE_DEPTH64)]
    whole.close()
    part = RenderSession(renderer, scene, cfg, samples_per_step=2)
    part.step()
    part.step()
    d = str(tmp_path / "ckpt")
    part.save(d)
    part.save(d)  # replacing a checkpoint keeps one complete copy
    assert sorted(os.listdir(d)) == sorted(["state.ron", "accum.bin", "depth.bin", "normals.bin", "volume.bin",
                                            "volume-moments.bin", "surface.bin", "volume-components.bin",
                                            "depth64.bin", "steps.bin"])
    assert os.path.getsize(os.path.join(d, "accum.bin")) == 72 * 40 * 16
    assert os.path.getsize(os.path.join(d, "volume-components.bin")) == 48
    part.close()
    res = RenderSession.resume(renderer, d)
    try:
        assert res.samples_done == 4
        while not res.done():
            res.step()
        got = [res.job.readback(p) for p in (abi.PLANE_ACCUM, abi.PLANE_DEPTH, abi.PLANE_AUX, abi.PLANE_DEPTH64)]
    finally:
        res.close()
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_resume_refuses_another_transport_and_foreign_directories(renderer, tmp_path):
    from syfrack_amd import abi
    from syfrack_amd.session import RenderSession, STATE_FILE
    scene, cfg = _scene()
    s = RenderSession(renderer, scene, cfg, samples_per_step=2)
    s.step()
    d = str(tmp_path / "ckpt")
    s.save(d)
    # a fingerprint that does not match the rebuilt job: refused, files untouched
    p = os.path.join(d, STATE_FILE)
    text = open(p, encoding="utf-8").read()
    fp = s.job.fingerprint()
    open(p, "w", encoding="utf-8").write(text.replace(f"shader_fingerprint: {fp}", f"shader_fingerprint: {fp ^ 1}"))
    with pytest.raises(ValueError, match="different transport"):
        RenderSession.resume(renderer, d)
    # a directory holding anything else is never replaced
    open(os.path.join(d, "notes.txt"), "w").write("mine")
    with pytest.raises(abi.SyfError, match="not a SyFrack state directory"):
        s.save(d)
    assert os.path.exists(os.path.join(d, "notes.txt"))
    s.close()


def test_export_payload_and_checksum(renderer, tmp_path):
    from syfrack_amd import abi
    from syfrack_amd.session import RenderSession, export_read, fnv1a64
    scene, cfg = _scene()
    s = RenderSession(renderer, scene, cfg, samples_per_step=6)
    s.step()
    path = str(tmp_path / "export.bin")
    nbytes, digest = s.export(path)
    acc = s.job.readback(abi.PLANE_ACCUM)
    mean = np.concatenate([acc[..., :3] / np.maximum(acc[..., 3:4], 1.0), np.ones_like(acc[..., :1])], axis=-1)
    depth, aux = s.job.readback(abi.PLANE_DEPTH), s.job.readback(abi.PLANE_AUX)
    s.close()
    payload = open(path, "rb").read()
    assert nbytes == len(payload) == 72 * 40 * 36
    assert digest == fnv1a64(payload)
    c, d, n = export_read(path, nbytes, digest, 72, 40)
    assert np.array_equal(c, mean.astype(np.float32)) and np.array_equal(d, depth) and np.array_equal(n, aux)
