"""The session's host side on the CPU: state.ron round trips, the FNV-1a-64 of the export
checksum against its published test vectors, and the export payload's byte-count and checksum
checks (export.rs:1263-1324), through the C ABI's host-only entry points."""
import os

import numpy as np
import pytest

from syfrack_amd import abi, scene as sc, session


def test_state_ron_round_trip_rebuilds_the_scene():
    scene, cfg = sc.mandelbox_pathtrace(64, 36, 4)
    scene.lights.append(sc.Light.environment((0.85, 0.9, 1.0), (0.55, 0.55, 0.6), (0.15, 0.13, 0.12), 1.0))
    scene.lights.append(sc.Light.sun((0.3, 0.8, 0.5), (1.0, 0.95, 0.9), 3.0, 0.53))
    scene.material.coloring = sc.Coloring(sc.Ramp([(1, 0, 0), (0, 1, 0), (0, 0, 1)], [0.1, 0.2, 0.8], True, 2.0, -0.1,
                                                  1.5, True), 0, ((0.5, 0.25, 0.0, 0.0), 0.1, 0.2, 3.0))
    scene.material.finishes = [sc.Finish((0.2, 0.3, 0.4), 0.6, 0.1, 0.05, (1.0, 0.5, 0.0), True), sc.Finish()]
    scene.ground = sc.Ground(-1.5, 2)
    scene.fog = sc.Fog(0.025, (0.85, 0.8, 0.78), -2.0, 0.4, 0.55)
    cfg.march_math = "exact"
    state = {"shader_fingerprint": 0xfedcba9876543210, "scene": scene, "cfg": cfg, "samples_done": 3, "spd": 2,
             "elapsed_secs": 0.1 + 0.2, "note": None}
    text = session.to_ron(state)
    back = session.from_ron(text)
    assert back["shader_fingerprint"] == 0xfedcba9876543210 and back["samples_done"] == 3
    assert back["elapsed_secs"] == 0.1 + 0.2 and back["note"] is None
    s2 = session.scene_from_state(back["scene"])
    assert s2 == scene
    assert session._build(sc.RenderConfig, back["cfg"]) == cfg
    assert back["scene"]["__struct__"] == "Scene"


def test_ron_parser_reads_comments_options_and_nested_groups():
    v = session.from_ron('(a: Some(1.5), // comment\n b: [1, 2, (3, "x\\"y")], c: Name(d: true), e: -inf)')
    assert v["a"] == 1.5 and v["b"] == [1, 2, (3, 'x"y')] and v["c"] == {"d": True, "__struct__": "Name"}
    assert v["e"] == float("-inf")
    with pytest.raises(ValueError):
        session.from_ron("(a: 1,")


def test_fnv1a64_known_answers():
    # the published FNV-1a 64-bit test vectors (offset basis 0xcbf29ce484222325, prime 0x100000001b3)
    assert session.fnv1a64(b"") == 0xcbf29ce484222325
    assert session.fnv1a64(b"a") == 0xaf63dc4c8601ec8c
    assert session.fnv1a64(b"foobar") == 0x85944171f73967e8
    # continuing a hash over a second buffer equals hashing the concatenation
    assert session.fnv1a64(b"bar", session.fnv1a64(b"foo")) == session.fnv1a64(b"foobar")


def test_export_payload_checks_byte_count_and_checksum(tmp_path):
    w, h = 5, 3
    rng = np.random.default_rng(1)
    color = rng.random((h, w, 4), dtype=np.float32)
    depth = rng.random((h, w), dtype=np.float32)
    normals = rng.random((h, w, 4), dtype=np.float32)
    payload = color.tobytes() + depth.tobytes() + normals.tobytes()
    path = str(tmp_path / "export.bin")
    with open(path, "wb") as fh:
        fh.write(payload)
    digest = session.fnv1a64(payload)
    c, d, n = session.export_read(path, len(payload), digest, w, h)
    assert np.array_equal(c, color) and np.array_equal(d, depth) and np.array_equal(n, normals)
    with pytest.raises(abi.SyfError, match="checksum mismatch"):
        session.export_read(path, len(payload), digest ^ 1, w, h)
    with pytest.raises(abi.SyfError, match="invalid byte count"):
        session.export_read(path, len(payload) - 4, digest, w, h)
    with open(path, "ab") as fh:
        fh.write(b"\0")
    with pytest.raises(abi.SyfError):
        session.export_read(path, len(payload), digest, w, h)
