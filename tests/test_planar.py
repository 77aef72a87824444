"""2D escape time: the oracle pinned on the CPU, then the HIP kernels against it.

CPU (oracle) pins:
  * planar.rs:275-345 known answers: complex powers, c = 0 never escapes, (2, 1) escapes
    within 5 iterations, stripe/smooth continuity across an escape boundary, finite traps;
  * conjugate symmetry of the Mandelbrot set (iter(c) = iter(conj c)) and of the Burning
    Ship's mirror (iter(x, y) = iter(x, y) only) checked against the FP64 loop;
  * FP64 and quad-precision loops agree where FP64 resolves the pixel grid;
  * the deep-zoom reference orbit (double-double) equals the binary128 orbit rounded to FP64
    where a plain FP64 orbit already drifts, and the PERTURB64 loop (FP64 deltas with rebasing
    against it) agrees with binary128 at the exact pixel points of 1e-12 views, Burning Ship's
    chaotic bands included, where the FP64 direct loop does not.
GPU (marked gpu):
  * FP64 mode: iteration counts bit-exact against the oracle's FP64 loop;
  * FP32 mode: counts equal for >= 99.9 % of pixels against the f32 restatement;
  * perturbation: against binary128 evaluations of the same points at span 1e-6 (calm and
    chaotic views, on the points whose count survives a 16-ulp move of the f32 offset) and
    1e-12 (a boundary-band centre), and against FP64 direct on a whole-set view.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from syfrack_amd import abi
from syfrack_amd.planar import PlanarView

MB, BS, JU = abi.PLANAR_MANDELBROT, abi.PLANAR_BURNING_SHIP, abi.PLANAR_JULIA


def test_reference_known_answers():
    # c = 0 never escapes; (2, 1) escapes within 5 (planar.rs:318-327), default bailout 16
    it, _ = oracle.planar_sample_f32(MB, 2, 128, (0, 0), 16.0, 0.5, 12.0, 0.0, 0.0)
    assert it == 128
    it, _ = oracle.planar_sample_f32(MB, 2, 128, (0, 0), 16.0, 0.5, 12.0, 2.0, 1.0)
    assert it < 5
    # Julia c = 0 from |z| = 4 -+ 1e-5: the escape count changes, stripe and smooth do not jump
    a = math.pi / 32
    d = (math.cos(a), math.sin(a))
    i1, o1 = oracle.planar_sample_f32(JU, 2, 128, (0, 0), 16.0, 0.5, 12.0, d[0] * (4 - 1e-5), d[1] * (4 - 1e-5))
    i2, o2 = oracle.planar_sample_f32(JU, 2, 128, (0, 0), 16.0, 0.5, 12.0, d[0] * (4 + 1e-5), d[1] * (4 + 1e-5))
    assert i1 != i2
    assert abs(o1[3] - o2[3]) < 1e-4
    assert abs(o1[1] - o2[1]) < 1e-5
    for fam in (JU, MB, BS):
        for i in range(100):
            it, o = oracle.planar_sample_f32(fam, 2, 128, (-0.74543, 0.11301), 16.0, 0.5, 12.0,
                                             (i % 10) * 0.37 - 1.6, (i // 10) * 0.31 - 1.5)
            assert np.all(np.isfinite(o)) and np.all((o[1:] >= 0) & (o[1:] <= 1))


def _view(**kw):
    v = PlanarView(**kw)
    return v.c_struct()


def test_fp64_loop_mandelbrot_symmetry_and_trivial_points():
    d = _view(family=MB, width=64, height=64, iterations=500, center=(-0.5, 0.0), span=3.0, bailout=2.0)
    it, sm = oracle.escape_f64(d)
    # the view is symmetric about the real axis: row j and row 63-j see conjugate points
    assert np.array_equal(it, it[::-1, :])
    # far points escape at once (|c| > 2 => one iteration), the main cardioid never escapes
    d2 = _view(family=MB, width=8, height=8, iterations=500, center=(-0.2, 0.0), span=0.1, bailout=2.0)
    it2, _ = oracle.escape_f64(d2)
    assert np.all(it2 == 500)
    d3 = _view(family=MB, width=8, height=8, iterations=500, center=(3.0, 3.0), span=0.1, bailout=2.0)
    it3, _ = oracle.escape_f64(d3)
    assert np.all(it3 == 1)


def test_fp64_and_quad_loops_agree_at_moderate_zoom():
    d = _view(family=MB, width=64, height=64, iterations=2000, center=(-0.743643887, 0.131825904), span=1e-4,
              bailout=16.0)
    it64, sm64 = oracle.escape_f64(d)
    ys, xs = np.mgrid[0:64:4, 0:64:4]
    # quad evaluates centre + (float) offset; at span 1e-4 that rounding is below 1e-11
    itq, smq = oracle.escape_f128(d, xs.ravel(), ys.ravel())
    a = it64[ys.ravel(), xs.ravel()]
    assert np.mean(np.abs(a.astype(int) - itq.astype(int)) <= 1) >= 0.95


# Deep-zoom views at span 1e-12: the Mandelbrot seahorse valley, Burning Ship's antenna band and
# the Burning Ship boundary-band centre of seed 0 (boundary_centre(family 3, seed 0, 2000)).
DEEP_VIEWS = [(MB, (-0.743643887, 0.131825904)), (BS, (-1.62, -0.01)), (BS, (0.6826171875, -1.1083984375))]


def _exact_resolvable(desc, xs, ys, nudge=64.0):
    """binary128 counts and smooth at the exact pixel points, and the points whose count moves
    by at most 1 (and, separately, whose smooth moves by at most 1e-3) when the FP64 offset is
    moved by `nudge` of its ulps either way: the points an FP64-delta method can resolve."""
    q, qs = oracle.escape_f128_exact(desc, xs, ys)
    stable = np.ones(len(q), bool)
    smooth_stable = np.ones(len(q), bool)
    for n in (nudge, -nudge):
        qn, qsn = oracle.escape_f128_exact(desc, xs, ys, n)
        stable &= np.abs(qn.astype(int) - q.astype(int)) <= 1
        smooth_stable &= (qn == q) & (np.abs(qsn - qs) <= 1e-3)
    return q, qs, stable, smooth_stable


def test_reference_orbit_double_double_tracks_binary128():
    """The deep-zoom reference orbit is iterated in double-double: on the seahorse centre every
    entry equals the binary128 orbit rounded to FP64, while a plain FP64 iteration of the same
    recurrence leaves it within two steps (so an orbit that drops the low words fails here)."""
    c = (-0.743643887, 0.131825904)
    v = _view(family=MB, mode=abi.PLANAR_PERTURB64, width=4096, height=4096, iterations=10000, center=c,
              span=1e-12, bailout=16.0)
    dx, dy = oracle.reference_orbit(v)
    qx, qy = oracle.reference_orbit(v, quad=True)
    assert len(dx) == len(qx) > 1000
    assert np.array_equal(dx, qx) and np.array_equal(dy, qy)
    x = y = 0.0
    drift = None
    for n in range(1, len(qx)):
        x, y = x * x - y * y + c[0], 2.0 * x * y + c[1]
        if x != qx[n] or y != qy[n]:
            drift = n
            break
    assert drift is not None and drift < 10
    # Burning Ship: the |.| fold in double-double, against binary128 over the calm start
    vb = _view(family=BS, mode=abi.PLANAR_PERTURB64, width=4096, height=4096, iterations=10000, center=(-1.62, -0.01),
               span=1e-12, bailout=16.0)
    bx, by = oracle.reference_orbit(vb, 2e-13, -1e-13)
    cx, cy = oracle.reference_orbit(vb, 2e-13, -1e-13, quad=True)
    assert np.array_equal(bx[:60], cx[:60]) and np.array_equal(by[:60], cy[:60])


@pytest.mark.parametrize("family,centre", DEEP_VIEWS)
def test_perturb64_matches_binary128_at_deep_zoom(family, centre):
    """PERTURB64 restated (FP64 deltas, rebasing, double-double orbit) against binary128 at the
    exact pixel points of a 4096^2 view at span 1e-12 with 10,000 iterations: within 1 on every
    resolvable point, smooth within 2e-3 where it is resolvable.  The FP64 direct loop, whose
    c = fl(centre + offset) is off by up to half a pixel here, fails the same comparison on the
    Burning Ship views."""
    v = _view(family=family, mode=abi.PLANAR_PERTURB64, width=4096, height=4096, iterations=10000, center=centre,
              span=1e-12, bailout=16.0)
    rng = np.random.default_rng(2)
    xs = rng.integers(0, 4096, 200)
    ys = rng.integers(0, 4096, 200)
    q, qs, stable, sstable = _exact_resolvable(v, xs, ys)
    it, sm = oracle.perturb64(v, xs, ys)
    near = np.abs(it.astype(int) - q.astype(int)) <= 1
    assert stable.sum() >= 40
    assert near[stable].mean() >= 0.99
    esc = sstable & (q < 10000) & (it == q)
    assert esc.sum() >= 20 and np.all(np.abs(sm[esc] - qs[esc]) <= 2e-3)
    if family == BS:
        direct = np.array([oracle.escape_f64(v, rect=(int(x), int(y), 1, 1))[0][0, 0] for x, y in zip(xs, ys)])
        assert (np.abs(direct.astype(int) - q.astype(int)) <= 1)[stable].mean() < 0.6


def test_perturb64_agrees_with_fp64_where_fp64_resolves():
    """A whole-set view (rebasing fires constantly there): counts within 1 of FP64 direct."""
    v = _view(family=MB, mode=abi.PLANAR_PERTURB64, width=128, height=96, iterations=1000, center=(-0.5, 0.0),
              span=3.0, bailout=16.0)
    ys, xs = np.mgrid[0:96, 0:128]
    it, _ = oracle.perturb64(v, xs.ravel(), ys.ravel())
    ref, _ = oracle.escape_f64(v)
    assert np.mean(np.abs(it.astype(int) - ref.ravel().astype(int)) <= 1) >= 0.99


@pytest.mark.gpu
class TestPlanarGPU:
    @pytest.fixture(scope="class")
    def ctx(self):
        from syfrack_amd.render import Context
        return Context(0)

    def _run(self, ctx, view):
        from syfrack_amd.planar import Planar
        p = Planar(ctx, view)
        p.run()
        p.sync()
        it, sm = p.readback()
        st = p.stats()
        p.close()
        return it, sm, st

    @pytest.mark.parametrize("family", [MB, BS, JU])
    def test_fp64_counts_bit_exact(self, ctx, family):
        v = PlanarView(family=family, mode=abi.PLANAR_FP64, width=256, height=192, iterations=3000,
                       center=(-0.75, 0.1) if family != BS else (-1.75, -0.03), span=2.5 if family != BS else 0.2,
                       bailout=16.0, rotation=10.0)
        it, sm, st = self._run(ctx, v)
        ref_it, ref_sm = oracle.escape_f64(v.c_struct())
        assert np.array_equal(it, ref_it)
        assert st["iterations"] == int(it.astype(np.int64).sum())
        esc = ref_it < 3000
        assert np.allclose(sm[esc], ref_sm[esc], rtol=1e-5)

    @pytest.mark.parametrize("family", [MB, BS])
    def test_fp32_matches_reference_precision_loop(self, ctx, family):
        v = PlanarView(family=family, mode=abi.PLANAR_FP32, width=256, height=256, iterations=512,
                       center=(-0.75, 0.0) if family == MB else (-0.5, -0.5), span=3.0, bailout=16.0)
        it, sm, _ = self._run(ctx, v)
        ref_it, ref_sm = oracle.planar_f32(v.c_struct())
        assert np.mean(it == ref_it) >= 0.999
        esc = (ref_it < 512) & (it == ref_it)
        assert np.allclose(sm[esc], ref_sm[esc], rtol=1e-4, atol=1e-4)

    @pytest.mark.parametrize("family,centre,span", [
        (MB, (-0.743643887, 0.131825904), 1e-6),   # Seahorse valley
        (BS, (-1.7548, -0.0284), 1e-6),            # the ship in the antenna
        (BS, (-1.62, -0.01), 1e-6),                # chaotic antenna band
    ])
    def test_perturbation_matches_quad_on_the_same_points(self, ctx, family, centre, span):
        """FP32 deltas against a binary128 evaluation of c = centre + (f32) offset, the point
        the first reference's deltas represent, on the points an FP32-delta method can resolve.
        An FP32 delta carries a rounding error of about one ulp of itself per iteration; where
        the dynamics stretch errors (|delta_k| growing as the amplification), each iteration's
        rounding reaches the escape as about one ulp of the initial offset, so after n
        iterations the orbit is known to about sqrt(n) ulps of the offset: 16 for the few
        hundred iterations these points run.  The resolvable points are those whose binary128
        count moves by at most 1 when the f32 offset is moved by 16 of its ulps either way
        (near the set boundary, counts are chaotic below that: nearly half of the chaotic band
        view).  On them: within 1 of binary128 for >= 99 %."""
        v = PlanarView(family=family, mode=abi.PLANAR_PERTURB, width=128, height=128, iterations=2000,
                       center=centre, span=span, bailout=16.0)
        it, sm, st = self._run(ctx, v)
        ys, xs = np.mgrid[0:128:4, 0:128:4]
        xs, ys = xs.ravel(), ys.ravel()
        q = oracle.escape_f128(v.c_struct(), xs, ys)[0].astype(int)
        stable = np.ones(len(q), bool)
        for nudge in (16.0, -16.0):
            stable &= np.abs(oracle.escape_f128(v.c_struct(), xs, ys, nudge)[0].astype(int) - q) <= 1
        near = np.abs(it[ys, xs].astype(int) - q) <= 1
        print(f"family {family} centre {centre} span {span}: resolvable {stable.mean():.4f}, within 1 of quad there "
              f"{near[stable].mean():.4f} (all points {near.mean():.4f}), glitch restarts {st['rebases']}")
        assert stable.mean() >= 0.5
        assert near[stable].mean() >= 0.99

    def test_perturbation_deep_zoom_matches_quad(self, ctx):
        """Mandelbrot at 1e-12 around a boundary-band centre (seed 5, the deep-zoom case)."""
        from syfrack_amd.planar import boundary_centre
        c = boundary_centre(ctx, MB, seed=5, iterations=2000)
        v = PlanarView(family=MB, mode=abi.PLANAR_PERTURB, width=512, height=512, iterations=10000,
                       center=c, span=1e-12, bailout=16.0)
        it, sm, st = self._run(ctx, v)
        rng = np.random.default_rng(0)
        xs = rng.integers(0, 512, 300)
        ys = rng.integers(0, 512, 300)
        q_it, q_sm = oracle.escape_f128(v.c_struct(), xs, ys)
        g = it[ys, xs]
        near = np.mean(np.abs(g.astype(int) - q_it.astype(int)) <= 1)
        print(f"deep zoom 1e-12 centre {c}: within 1 of quad: {near:.4f}; glitch restarts {st['rebases']}")
        assert near >= 0.95

    def test_perturbation_deep_zoom_burning_ship_matches_quad(self, ctx):
        """Burning Ship at 1e-12 around its boundary-band centre (seed 0), 10,000 iterations:
        counts within 1 of binary128 (at c = centre + f32 offset) on >= 99 % of the points whose
        count survives a 16-ulp move of the f32 offset, smooth within 2e-3 where the smooth value
        survives it too; pixels left to the FP64 rebasing fallback are reported and bounded."""
        from syfrack_amd.planar import boundary_centre
        c = boundary_centre(ctx, BS, seed=0, iterations=2000)
        v = PlanarView(family=BS, mode=abi.PLANAR_PERTURB, width=512, height=512, iterations=10000,
                       center=c, span=1e-12, bailout=16.0)
        it, sm, st = self._run(ctx, v)
        rng = np.random.default_rng(4)
        xs = rng.integers(0, 512, 1500)
        ys = rng.integers(0, 512, 1500)
        q, qs = oracle.escape_f128(v.c_struct(), xs, ys)
        stable = np.ones(len(q), bool)
        sstable = np.ones(len(q), bool)
        for nudge in (16.0, -16.0):
            qn, qsn = oracle.escape_f128(v.c_struct(), xs, ys, nudge)
            stable &= np.abs(qn.astype(int) - q.astype(int)) <= 1
            sstable &= (qn == q) & (np.abs(qsn - qs) <= 1e-3)
        g, gs = it[ys, xs], sm[ys, xs]
        near = np.abs(g.astype(int) - q.astype(int)) <= 1
        esc = sstable & (q < 10000) & (g == q)
        print(f"Burning Ship 1e-12 centre {c}: resolvable {int(stable.sum())}, within 1 there {near[stable].mean():.4f}; "
              f"references {st['references']}, fallback pixels {st['fallback_pixels']}, smooth points {int(esc.sum())}")
        assert stable.sum() >= 20
        assert near[stable].mean() >= 0.99
        assert np.all(np.abs(gs[esc] - qs[esc]) <= 2e-3)
        assert st["fallback_pixels"] <= 0.01 * 512 * 512

    def test_deep_zoom_smooth_plane_matches_quad(self, ctx):
        """The SMOOTH plane of the Mandelbrot 1e-12 view against binary128's smooth count."""
        from syfrack_amd.planar import boundary_centre
        c = boundary_centre(ctx, MB, seed=5, iterations=2000)
        v = PlanarView(family=MB, mode=abi.PLANAR_PERTURB, width=512, height=512, iterations=10000,
                       center=c, span=1e-12, bailout=16.0)
        it, sm, st = self._run(ctx, v)
        rng = np.random.default_rng(6)
        xs = rng.integers(0, 512, 300)
        ys = rng.integers(0, 512, 300)
        q, qs = oracle.escape_f128(v.c_struct(), xs, ys)
        esc = (q < 10000) & (it[ys, xs] == q)
        assert esc.mean() >= 0.9
        assert np.percentile(np.abs(sm[ys, xs][esc] - qs[esc]), 99) <= 1e-3

    @pytest.mark.parametrize("family,centre", DEEP_VIEWS)
    def test_perturb64_bit_exact_against_oracle(self, ctx, family, centre):
        """PERTURB64 (the view centre's double-double orbit, FP64 deltas with rebasing) against
        the oracle's restatement of the same loop on every pixel of a 96x96 view at 1e-12 with
        10,000 iterations: counts bit-exact (IEEE adds and multiplies rounded as written on both
        sides), smooth counts within 1e-6 relative (their logarithms are the two sides' own)."""
        v = PlanarView(family=family, mode=abi.PLANAR_PERTURB64, width=96, height=96, iterations=10000,
                       center=centre, span=1e-12, bailout=16.0)
        it, sm, st = self._run(ctx, v)
        ys, xs = np.mgrid[0:96, 0:96]
        ref_it, ref_sm = oracle.perturb64(v.c_struct(), xs.ravel(), ys.ravel())
        assert np.array_equal(it.ravel(), ref_it)
        assert np.allclose(sm.ravel(), ref_sm, rtol=1e-6)
        assert st["iterations"] == int(it.astype(np.int64).sum())
        assert st["references"] == 1 and st["fallback_pixels"] == 0

    def test_rebasing_fallback_matches_quad(self, ctx):
        """PERTURB with a budget of one reference: every glitch of the first pass goes to the
        FP64 rebasing loop over the glitch list.  Those pixels equal the oracle's PERTURB64 on
        the same pixels exactly; the image is within 1 of binary128 on the resolvable points."""
        centre = (-1.62, -0.01)
        v = PlanarView(family=BS, mode=abi.PLANAR_PERTURB, width=256, height=256, iterations=10000,
                       center=centre, span=1e-9, bailout=16.0, max_references=1)
        it, sm, st = self._run(ctx, v)
        assert st["references"] == 1 and st["fallback_pixels"] > 0
        v64 = PlanarView(**{**v.__dict__, "mode": abi.PLANAR_PERTURB64, "max_references": 0})
        it64, _, _ = self._run(ctx, v64)
        rng = np.random.default_rng(8)
        xs = rng.integers(0, 256, 400)
        ys = rng.integers(0, 256, 400)
        q, _ = oracle.escape_f128(v.c_struct(), xs, ys)
        stable = np.ones(len(q), bool)
        for nudge in (16.0, -16.0):
            stable &= np.abs(oracle.escape_f128(v.c_struct(), xs, ys, nudge)[0].astype(int) - q.astype(int)) <= 1
        near = np.abs(it[ys, xs].astype(int) - q.astype(int)) <= 1
        print(f"fallback pixels {st['fallback_pixels']}, rebasings {st['rebasings']}, resolvable {int(stable.sum())}, "
              f"within 1 there {near[stable].mean():.4f}, equal to PERTURB64 {np.mean(it == it64):.4f}")
        assert near[stable].mean() >= 0.99
        # pixels of the fallback took the PERTURB64 loop itself
        assert np.mean(it == it64) >= st["fallback_pixels"] / (256 * 256)

    def test_planar_refuses_bad_reference_budget(self, ctx):
        from syfrack_amd.planar import Planar
        v = PlanarView(family=MB, mode=abi.PLANAR_PERTURB, width=64, height=64, iterations=100, max_references=65)
        with pytest.raises(abi.SyfError, match="max_references"):
            Planar(ctx, v)
        with pytest.raises(abi.SyfError, match="parameter-plane"):
            Planar(ctx, PlanarView(family=JU, mode=abi.PLANAR_PERTURB64, width=64, height=64, iterations=100))

    @pytest.mark.parametrize("iterations", [7, 37, 1000])
    def test_perturbation_agrees_with_fp64_where_fp64_resolves(self, ctx, iterations):
        """A view of the whole set: FP64 direct resolves every pixel, perturbation must match.
        Iteration counts below and between the kernel's 16-entry orbit chunks included."""
        v = PlanarView(family=MB, mode=abi.PLANAR_PERTURB, width=256, height=192, iterations=iterations,
                       center=(-0.5, 0.0), span=3.0, bailout=16.0)
        it, sm, st = self._run(ctx, v)
        v64 = PlanarView(**{**v.__dict__, "mode": abi.PLANAR_FP64})
        ref_it, _ = oracle.escape_f64(v64.c_struct())
        assert np.mean(np.abs(it.astype(int) - ref_it.astype(int)) <= 1) >= 0.99
