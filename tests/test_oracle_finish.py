"""The oracle's finish detail patterns (finish_detail.wgsl) pinned by what the mathematics fixes.

- Patterns 1-7, 9 and 10 return the analytic gradient of their height: it must match a central
  finite difference of the height at >= 95 % of random points (the rest sit on the kinks of abs,
  sign, fract and clamp).
- Every band fades to its mean where the footprint passes its wavelength: at a huge footprint the
  lattice noise is gone and each pattern's height is its closed-form constant.
- Sand (8), WeatheredConcrete (11) and WetConcrete (12) return shading slopes, not gradients; each
  is pinned in a footprint range where its parts separate: sand's ripple slope is the gradient of
  its height / 0.70 once the grain has faded; weathered concrete's height gradient is
  0.72 (g11 - g10) / 0.12 + 0.28 g10 / 0.18 once aggregate, pores and boards have faded (g10 is
  plain concrete's gradient at the same point, its stains alone); wet concrete's slope is
  wet x (2.8, 0, 4.1) cos(2.8 qx + 4.1 qz) 0.018 visible x scale once its grit has faded.
  Still unpinned: the grain weights of sand (2.5) and wet concrete (16) at small footprints.
"""
import numpy as np
import pytest

from oracle import oracle

SCALE = 1.7
SEED = 7


def _points(n=400, seed=3, extent=3.0):
    return np.random.default_rng(seed).uniform(-extent, extent, (n, 3)).astype(np.float32)


def _fd(kind, p, footprint, scale=SCALE, eps=5e-4):
    h = lambda q: oracle.finish_relief(kind, SEED, scale, q, footprint)[:, 3].astype(np.float64)
    g = np.zeros((len(p), 3))
    for a in range(3):
        d = np.zeros(3, np.float32)
        d[a] = eps / scale
        g[:, a] = (h(p + d) - h(p - d)) / (2 * eps / scale)
    return g


def _agree(a, b, rel=0.03):
    err = np.linalg.norm(a - b, axis=1)
    tol = rel * np.maximum(np.linalg.norm(b, axis=1), np.median(np.linalg.norm(b, axis=1)))
    return np.mean(err <= tol + 1e-3)


@pytest.mark.parametrize("kind", [1, 2, 3, 4, 5, 6, 7, 9, 10])
def test_slope_is_the_gradient_of_the_height(kind):
    p = _points()
    out = oracle.finish_relief(kind, SEED, SCALE, p, 1e-4)
    assert np.all(out[:, 3] >= -1e-6) and np.all(out[:, 3] <= 1.0 + 1e-6)
    frac = _agree(out[:, :3].astype(np.float64), _fd(kind, p, 1e-4))
    print(f"pattern {kind}: {frac:.3f} of the points agree")
    assert frac >= 0.95


def test_bands_fade_to_their_means_at_a_huge_footprint():
    p = _points(50)
    s = lambda a, b, x: (lambda t: t * t * (3 - 2 * t))(min(max((x - a) / (b - a), 0.0), 1.0))
    expect = {1: 0.5, 2: s(0.35, 0.67, 0.5), 3: 0.5, 4: 0.5, 5: 0.5, 6: 0.2, 7: 0.5, 8: 0.5,
              9: s(0.36, 0.64, 0.5) * (0.85 + 0.15 * 0.5), 10: 0.18 * 0.5 + 0.65 * 0.5 - 0.14 * 0.2 + 0.12,
              11: 0.72 * s(0.33, 0.68, 0.5) + 0.28 * 0.5, 12: s(0.47, 0.59, 0.5)}
    for kind, h in expect.items():
        out = oracle.finish_relief(kind, SEED, SCALE, p, 1e4)
        assert np.allclose(out[:, 3], h, atol=1e-6), kind
        if kind != 12:
            assert np.allclose(out[:, :3], 0.0, atol=1e-6), kind


def test_sand_ripple_slope_once_the_grain_has_faded():
    p = _points()
    fp, scale = 0.012, 0.5  # grain (x90) faded, ripples visible
    out = oracle.finish_relief(8, SEED, scale, p, fp)
    assert _agree(0.70 * out[:, :3].astype(np.float64), _fd(8, p, fp, scale)) >= 0.95


def test_weathered_concrete_slope_against_its_height():
    p = _points()
    fp = 0.2 / SCALE  # width 0.2: aggregate, pores and board joints faded; stains and rain visible
    g10 = oracle.finish_relief(10, SEED, SCALE, p, fp)[:, :3].astype(np.float64)
    g11 = oracle.finish_relief(11, SEED, SCALE, p, fp)[:, :3].astype(np.float64)
    assert _agree(g10, _fd(10, p, fp)) >= 0.95
    predicted = 0.72 * (g11 - g10) / 0.12 + 0.28 * g10 / 0.18
    assert _agree(predicted, _fd(11, p, fp)) >= 0.95


def test_wet_concrete_pools_carry_the_ripple_slope():
    p = _points()
    fp = 0.1 / SCALE  # width 0.1: grit (x16) faded, basins visible
    out = oracle.finish_relief(12, SEED, SCALE, p, fp).astype(np.float64)
    q = p.astype(np.float64) * SCALE
    width = 0.1
    t = min(max((width * 3 - 0.15) / 0.45, 0.0), 1.0)
    vis = 1 - t * t * (3 - 2 * t)
    c = np.cos(q[:, 0] * 2.8 + q[:, 2] * 4.1) * 0.018 * vis * SCALE
    expect = out[:, 3:4] * np.stack([2.8 * c, 0 * c, 4.1 * c], axis=1)
    assert np.allclose(out[:, :3], expect, rtol=1e-4, atol=1e-6)
    assert 0.05 < np.mean(out[:, 3] > 0.5) < 0.95  # both dry and wet points were tested


def test_lattice_values_are_uniform_and_seeded():
    p = _points(2000, seed=5, extent=40.0)
    h = oracle.finish_relief(4, SEED, 1.0, p, 1e4)  # constant: the bands are gone
    a = oracle.finish_relief(3, 11, 1.0, p, 1e-4)[:, 3]
    b = oracle.finish_relief(3, 12, 1.0, p, 1e-4)[:, 3]
    assert np.all(h[:, 3] == h[0, 3]) and not np.allclose(a, b)
    assert 0.4 < a.mean() < 0.6 and a.min() >= 0.0 and a.max() <= 1.0
