"""Host side of the render server: the request streams are seeded and follow their stated
distributions, and footprints are what admission charges.  No device is used."""
import math

import numpy as np

from syfrack_amd import abi, serve


def test_uneven_requests_distribution_and_determinism():
    a = serve.uneven_requests(256, seed=3)
    b = serve.uneven_requests(256, seed=3)
    assert [(q.cfg.width, q.spp, q.cfg.max_bounces) for q in a] == [(q.cfg.width, q.spp, q.cfg.max_bounces) for q in b]
    sides = np.array([q.cfg.width for q in a])
    assert sides.min() >= 128 and sides.max() <= 1024 and np.all(sides % 8 == 0)
    assert all(q.cfg.width == q.cfg.height for q in a)
    # log-uniform: the median side sits near the geometric mean of the range
    assert abs(math.log(np.median(sides)) - 0.5 * (math.log(128) + math.log(1024))) < 0.25
    spp = np.array([q.spp for q in a])
    assert spp.min() == 1 and spp.max() == 16
    bounces = {q.cfg.max_bounces for q in a}
    assert bounces == {1, 2, 3, 4}
    fams = [q.scene.fractal.family for q in a]
    for f in (abi.FAMILY_MANDELBULB, abi.FAMILY_MANDELBOX, abi.FAMILY_MENGER):
        assert 60 <= fams.count(f) <= 110
    for q in a:
        f = q.scene.fractal
        if f.family == abi.FAMILY_MANDELBULB:
            assert 2 <= f.power <= 12 and f.power == int(f.power)
        elif f.family == abi.FAMILY_MANDELBOX:
            assert -2.5 <= f.scale <= 3.0
        else:
            assert 4 <= f.iterations <= 7
        q.scene.desc(q.cfg)  # lowers without error


def test_serving_mix_arrivals_and_shares():
    m = serve.serving_mix(400, rate_per_s=100.0, seed=5)
    t = np.array([q.arrival_s for q in m])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 0.01) < 0.002  # Poisson at 100/s: mean gap 10 ms
    prev = [q for q in m if q.preview]
    assert 0.62 < len(prev) / len(m) < 0.78
    assert all(q.cfg.width == 256 and q.spp == 1 for q in prev)
    big = [q for q in m if not q.preview]
    assert all(1024 <= q.cfg.width <= 2048 and 16 <= q.spp <= 64 for q in big)


def test_footprint_is_four_planes():
    from syfrack_amd.scene import RenderConfig
    assert serve.footprint(RenderConfig(256, 128)) == 256 * 128 * (16 + 4 + 16 + 8)
