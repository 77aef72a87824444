"""The control machine's image restated on the CPU (oracle.control_trace), pinned by what the
mathematics and the reference's own control tests fix (crates/syfrack-render/tests/
gpu_control.rs), and the host side of the session (validation before any device work).

  * a unit sphere seen from distance 3 through an odd image: the centre pixel's ray is the view
    axis, its depth is 2 less the pixel-cone tolerance, its normal points back at the camera
    and its colour is the headlight shade of that normal;
  * every hit's normal is the sphere's radial direction at the hit point;
  * the grazing plane of gpu_control.rs:219-246 (safety 0.1, camera 0.1 above the plane looking
    along it, 1200 mm lens): the ray is not truncated at 512 steps and hits between 40 and 100;
  * the undefined field of gpu_control.rs:249-280 (Power 8 then NormFence, probed at the
    origin): status 1, where the sphere of the same camera is complete;
  * a disabled fence is empty space (+inf), a miss, not an error.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from syfrack_amd import abi
from syfrack_amd.graph import Graph, Node
from syfrack_amd.scene import Camera, RenderConfig, Scene


def _trace(graph, cam, w, h, max_evals=1 << 20):
    oracle.graph_use(graph)
    desc = Scene(graph, cam).desc(RenderConfig(w, h, 1, 0, 1))
    return oracle.control_trace(desc, max_evals=max_evals)


def test_unit_sphere_centre_depth_normal_and_shade():
    cam = Camera.free((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), 35.0)
    acc, dep, aux, st = _trace(Graph.new_minimal(), cam, 33, 19)
    assert np.all(st == 0)
    c = dep[9, 16]
    # hit when |d| / t < cone: d = 3 - 1 - t along the axis, so t = 2 / (1 + cone) at the latest
    cone = (18.0 / 35.0) * 2.0 / 33.0
    assert 2.0 / (1.0 + cone) - 1e-3 <= c <= 2.0
    n = aux[9, 16, :3]
    assert np.allclose(n, (0.0, 0.0, -1.0), atol=1e-3)
    assert aux[9, 16, 3] >= 1
    light = np.array([0.4, 0.7, 0.6]) / math.sqrt(0.4 ** 2 + 0.7 ** 2 + 0.6 ** 2)
    lit = 0.2 + 0.8 * max(float(np.dot(n, light)), 0.0)
    assert np.allclose(acc[9, 16, :3], np.array([0.68, 0.70, 0.73]) * lit, atol=1e-5)
    # misses: the corner rays pass the sphere
    assert dep[0, 0] == -1.0 and np.allclose(acc[0, 0], (0.025, 0.035, 0.05, 1.0))
    assert np.array_equal(aux[0, 0], np.float32([0, 0, 0, -1]))


def test_sphere_normals_are_radial():
    cam = Camera.orbit((0.0, 0.0, 0.0), 3.2, 30.0, 20.0, 35.0)
    acc, dep, aux, st = _trace(Graph.new_minimal(), cam, 48, 32)
    desc = Scene(Graph.new_minimal(), cam).desc(RenderConfig(48, 32, 1, 0, 1))
    basis = oracle.camera_basis(desc)
    origin, fwd, right, up, tan_half, aspect = basis[0:3], basis[3:6], basis[6:9], basis[9:12], basis[12], basis[13]
    hits = np.argwhere(dep > 0)
    assert len(hits) > 200
    for y, x in hits:
        ndc = ((np.array([x, y]) + 0.5) / np.array([48, 32])) * 2 - 1
        rd = fwd + right * ndc[0] * tan_half + up * (-ndc[1] * tan_half / aspect)
        rd = rd / np.linalg.norm(rd)
        p = origin + rd * dep[y, x]
        assert np.dot(aux[y, x, :3], p / np.linalg.norm(p)) > 0.999


def test_grazing_ray_continues_past_512_steps():
    g = Graph.new_minimal()
    g.nodes[1] = Node(1, "Plane", {"normal": (0.0, 1.0, 0.0), "offset": 0.0})
    g.safety = 0.1
    cam = Camera.free((0.0, 0.1, 0.0), (100.0, 0.0, 0.0), 1200.0)
    acc, dep, aux, st = _trace(g, cam, 257, 1)
    assert st[0, 128] == 0
    assert 40.0 < dep[0, 128] < 100.0
    assert aux[0, 128, 1] > 0.999
    assert aux[0, 128, 3] > 512


def test_undefined_field_stops_the_pixel():
    g = Graph([Node(0, "Input"), Node(1, "Power", {"power": 8.0}), Node(2, "NormFence", {"inset": 0.5}),
               Node(3, "Output")], [((0, 0), (1, 0)), ((1, 0), (2, 0)), ((2, 0), (3, 0))])
    cam = Camera.free((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), 35.0)
    _, _, _, st = _trace(g, cam, 1, 1)
    assert st[0, 0] == 1
    _, dep, _, st = _trace(Graph.new_minimal(), cam, 1, 1)
    assert st[0, 0] == 0 and dep[0, 0] > 0


def test_disabled_fence_is_empty_space():
    g = Graph.new_minimal()
    g.nodes[1].enabled = False
    cam = Camera.free((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), 35.0)
    acc, dep, aux, st = _trace(g, cam, 5, 5)
    assert np.all(st == 0) and np.all(dep == -1.0)
    assert np.allclose(acc[..., :3], (0.025, 0.035, 0.05))


def test_session_refuses_invalid_arguments_before_device_work():
    lib = abi.load()
    assert lib.syf_control_create(None, None, None, 0, None) == abi.SYF_E_INVALID
    assert b"null" in lib.syf_last_error()
    assert lib.syf_control_tick(None) == abi.SYF_E_INVALID
    assert lib.syf_control_probe(None, None, None, 0, None, None) == abi.SYF_E_INVALID
