"""Pins the CPU oracle (oracle/) against what fixes each of its functions
independently of its own code: the reference's defect functions through the
golden vectors in tests/golden/, known answers derived by hand from the
reference's config values, closed forms, invariants and brute force.

No GPU needed.  Each test names the oracle function and the reference
file:line it restates.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import ROOT, rel_rmse

GOLDEN = ROOT / "tests" / "golden" / "defects.npz"


def rofex_geom(risa, **kw):
    """The reference's example geometry (RISA/example/config.cfg:1-107)."""
    g = risa.geom_from_config(risa.config_path("c0"))
    # second plane as in config.cfg:36-42 (sourceDiameter, deltaX, deltaZ, sourceAngle)
    g = g.replace(sourceDiameter=(365.0, 370.0), deltaX=(815.0, 815.0), deltaZ=(1417.0, 1430.0),
                  sourceAngle=(240.0, 240.0))
    return g.replace(**kw) if kw else g


# --------------------------------------------------------------- a3 defects --
def golden_cases():
    z = np.load(GOLDEN, allow_pickle=False)
    return [(str(n), z[f"{n}__data"], z[f"{n}__thresh"], z[f"{n}__flags"], z[f"{n}__repaired"])
            for n in z["names"]]


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_defects_oracle_matches_reference_vectors(orc, case):
    """orc_find_defects / orc_interp_defects vs the reference's own
    interpolationFunctions.h:28-97 (vectors from oracle/_ref): bit-exact."""
    name, data, thr, flags, repaired = case
    got = orc.find_defects(data, thr[0], thr[1])
    np.testing.assert_array_equal(got, flags)
    np.testing.assert_array_equal(orc.interp_defects(data, flags), repaired)


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_defects_product_host_matches_reference_vectors(risa, case):
    """The product's host defect repair (risa_find_defect_detectors /
    risa_interpolate_defect_detectors) against the same vectors."""
    from risa_amd import _lib

    name, data, thr, flags, repaired = case
    P, D = data.shape
    lib = _lib.load()
    got = np.zeros(D, np.int32)
    d = np.ascontiguousarray(data)
    _lib.check(lib.risa_find_defect_detectors(_lib.ptr(d), D, P, thr[0], thr[1], _lib.ptr(got)))
    np.testing.assert_array_equal(got, flags)
    rep = d.copy()
    f = np.ascontiguousarray(flags, np.int32)
    _lib.check(lib.risa_interpolate_defect_detectors(_lib.ptr(rep), _lib.ptr(f), D, P))
    np.testing.assert_array_equal(rep, repaired)


def test_defect_vectors_exercise_the_wraps():
    """The fixture covers the unsigned-wrap neighbours (Q17): a flag at 0 and
    at D-1, and detectors changed by repair at the seam."""
    cases = {c[0]: c for c in golden_cases()}
    _, data, _, flags, rep = cases["flicker_edges_D72"]
    assert flags[0] == 1 and flags[71] == 1
    _, data, _, flags, rep = cases["dead_seam_D72"]
    assert flags[0] == 1 and flags[70] == 1 and flags[71] == 1
    # run 70..71 continues into 0 across the seam: all three repaired
    assert (rep[:, [0, 70, 71]] != data[:, [0, 70, 71]]).all()


# -------------------------------------------------------------- a4 area mask --
def test_area_mask_known_answers_rofex(orc, risa):
    """relevantAreaMask (Attenuation.cu:327-381) at the reference's ROFEX values
    (config.cfg:63-70: offsets 47/313, sourceOffset 23.2, xa..xf = 43, 285,
    175, 362, 0, 431; 432 x 500).  Hand-derived:
      lowerLimit*500 = 97.5 -> ya = yb = 98; upperLimit*500 = 466.94 -> yc = yd = 467
      m = (98 - 467) / (43 - 362) = 1.156740; ye = round(467 - 175 m) = 265
      whole-buffer fills: [0, trunc(0.195*432*500) = 42120) and [trunc(0.93389*500*432) = 201720, end)."""
    g = rofex_geom(risa, xa=43, xb=285, xc=175, xd=362, xe=0, xf=431,
                   lowerLimOffset=47.0, upperLimOffset=313.0, sourceOffset=23.2)
    m = orc.area_mask(g)
    assert m.shape == (500, 432)
    assert set(np.unique(m)) <= {0.0, 1.0}
    flat = m.ravel()
    lo, hi = 42120, 201720
    assert (flat[:lo] == 0).all() and (flat[hi:] == 0).all()
    # row 97 is cut in the middle by the whole-buffer fill: 42120 - 97*432 = 216
    assert (m[97, :216] == 0).all() and (m[97, 216:] == 1).all()
    # row 466 (last kept row by the fill) is cut at 201720 - 466*432 = 408
    assert m[466, 407] in (0.0, 1.0) and (m[466, 408:] == 0).all()
    slope = (98.0 - 467.0) / (43.0 - 362.0)
    # first sweep, x in [0, xa]: rows [98, round(265 + m x)) zeroed
    for x in (0, 20, 43):
        ymax = int(math.floor(265 + slope * x + 0.5))
        assert (m[98:ymax, x] == 0).all() and m[ymax, x] == 1.0, x
    assert (m[98:265, 0] == 0).all() and m[265, 0] == 1.0
    # third sweep, x in [xc, xd]: rows [round(98 + m (x - 43)), 467) zeroed, rows above kept
    for x in (175, 250, 362):
        ymin = int(math.floor(98 + slope * (x - 43) + 0.5))
        assert (m[ymin:467, x] == 0).all() and m[ymin - 1, x] == 1.0, x
    # fourth sweep, x in [xb, xf]: rows [98, round(98 + m (x - 285))) zeroed
    x = 431
    ymax = int(math.floor(98 + slope * (431 - 285) + 0.5))
    assert ymax == 267
    assert (m[98:267, x] == 0).all()
    # a column between the sweeps, e.g. x = 100 (second sweep): zero band is
    # [round(98 + m*57), round(265 + m*100)) = [164, 381)
    assert (m[164:381, 100] == 0).all() and m[163, 100] == 1.0 and m[381, 100] == 1.0


def test_area_mask_product_host_bit_exact(orc, risa, geom_c0, geom_c1):
    from risa_amd import _lib

    lib = _lib.load()
    for g in (geom_c0, geom_c1):
        out = np.empty(g.fan_shape, np.float32)
        _lib.check(lib.risa_attenuation_area_mask(g, _lib.ptr(out)))
        np.testing.assert_array_equal(out, orc.area_mask(g))


# ------------------------------------------------------ a2 reference average --
def test_reference_average_constant_blocks_and_block_index(orc, risa):
    """computeAverage (Attenuation.cu:236-251) reads block (i + plane), not
    (i*planes + plane) (Q2); the dark average is identically 0 (Q1,
    :222-223).  Frames filled with distinct constants (no column is flagged
    as defective: every column has zero variation and an equal neighbourhood)
    make the expected average a known answer."""
    g = rofex_geom(risa, numberOfFanDetectors=64, numberOfFanProjections=20, numberOfPlanes=2,
                   xa=0, xb=0, xc=0, xd=1, xe=0, xf=0)
    R = 3
    vals = [1000, 2000, 4000, 8000, 16000, 32000]      # block b holds vals[b]
    frames = np.stack([np.full((20, 64), v, np.uint16) for v in vals])
    avg_ref, avg_dark, repaired = orc.reference_average(g, frames, R)
    np.testing.assert_array_equal(repaired, frames)
    assert (avg_dark == 0).all()
    f = np.float32(1.0 / np.float32(R))
    for k in range(2):
        exp = np.float32(0)
        for i in range(R):
            exp = np.float32(exp + np.float32(vals[i + k]) * f)
        assert (avg_ref[k] == exp).all(), k
    # plane 0 averages blocks 0,1,2 = 7000/3; plane 1 blocks 1,2,3 = 14000/3
    assert abs(float(avg_ref[0, 0, 0]) - 7000 / 3) < 1e-3
    assert abs(float(avg_ref[1, 0, 0]) - 14000 / 3) < 1e-3


def test_reference_average_product_host_bit_exact(orc, risa, geom_c0):
    from risa_amd import _lib, synth

    g = geom_c0
    ref = synth.reference_frames(g, 3, seed=11)
    a_ref, a_dark, _ = orc.reference_average(g, ref.copy(), 3)
    p_ref, p_dark = synth.prepare_reference(g, ref.copy())
    np.testing.assert_array_equal(p_ref, a_ref)
    np.testing.assert_array_equal(p_dark, a_dark)


# ------------------------------------------------------------ a1 attenuation --
def test_attenuation_closed_form_and_clamps(orc, risa):
    """computeAttenuation (Attenuation.cu:416-444): -log(max(raw-dark, 1e-5) /
    max(ref-dark, 1e-5)) * mask, per element, plane offsets k*D and k*D*P."""
    g = rofex_geom(risa, numberOfFanDetectors=16, numberOfFanProjections=6, numberOfPlanes=2)
    rng = np.random.default_rng(3)
    P, D = 6, 16
    raw = rng.integers(0, 60000, (P, D)).astype(np.uint16)
    ref = rng.uniform(100, 60000, (2, P, D)).astype(np.float32)
    dark = rng.uniform(0, 300, (2, D)).astype(np.float32)
    mask = (rng.random((P, D)) > 0.2).astype(np.float32)
    raw[0, 0] = 0                        # raw - dark < eps
    ref[1, 1, 1] = dark[1, 1] - 5.0      # ref - dark < eps
    raw[1, 1] = 5000
    for k in range(2):
        num = np.maximum(raw.astype(np.float64) - dark[k][None, :], 1e-5)
        den = np.maximum(ref[k].astype(np.float64) - dark[k][None, :], 1e-5)
        exp = -np.log(num / den) * mask
        got64 = orc.attenuation_f64(g, raw, k, ref, dark, mask)
        got32 = orc.attenuation_f32(g, raw, k, ref, dark, mask)
        np.testing.assert_allclose(got64, exp, rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose(got32, exp, rtol=2e-6, atol=2e-6)
        assert (got32[mask == 0] == 0).all()
    # the clamped denominator: -log((5000 - dark) / 1e-5) at plane 1
    assert abs(orc.attenuation_f64(g, raw, 1, ref, dark, mask)[1, 1]
               - (-math.log((5000 - float(dark[1, 1])) / 1e-5)) * mask[1, 1]) < 1e-6


# ------------------------------------------------------- f3 reordering --
@pytest.mark.parametrize("P,D", [(7, 432), (5, 768), (1, 432)])
def test_reorder_is_the_module_transpose(orc, P, D):
    """orc_reorder_u16 (Reordering.cu:180-204).  With 27 modules of 16
    detectors the table moves [module][projection][16] to
    [projection][module*16 + detector] (a numpy transpose, independent of the
    table loops); wider frames keep that order over the first P*432 samples
    and read sample 0 past it (the zero-initialised table)."""
    rng = np.random.default_rng(P * 1000 + D)
    raw = rng.integers(1, 65535, size=(2, P * D), dtype=np.uint16)
    got = orc.reorder_u16(raw, P, D)
    for f in range(2):
        head = raw[f, :P * 432].reshape(27, P, 16).transpose(1, 0, 2).reshape(-1)
        flat = got[f].reshape(-1)
        np.testing.assert_array_equal(flat[:P * 432], head)
        assert (flat[P * 432:] == raw[f, 0]).all()
    if D == 432:
        # a permutation: every sample appears once
        ids = np.arange(P * D, dtype=np.uint16)
        np.testing.assert_array_equal(np.sort(orc.reorder_u16(ids, P, D).reshape(-1)), ids)
    with pytest.raises(ValueError):
        orc.reorder_u16(np.zeros(P * 400, np.uint16), P, 400)


# ---------------------------------------------------- a5 rebinning tables --
def test_ellipse_map_fixed_points_and_identity(orc):
    """ellipse_kreis_uwe (cuda_kernels_fan2para.h:38-62): for any tilt, 0 and
    pi are fixed points and f(2pi - a) = 2pi - f(a); with deltaX = 0 the map is
    the identity (eps = acos p1 = acos p2, gamma = 0, ae = cos(alpha))."""
    for dx, dz, diam in ((815.0, 1417.0, 365.0), (815.0, 1430.0, 370.0), (300.0, 900.0, 500.0)):
        assert abs(orc.ellipse_f64(1e-12, dx, dz, diam)) < 1e-5
        assert abs(orc.ellipse_f64(math.pi, dx, dz, diam) - math.pi) < 1e-6
        for a in (0.3, 1.0, 2.5):
            fa = orc.ellipse_f64(a, dx, dz, diam)
            fb = orc.ellipse_f64(2 * math.pi - a, dx, dz, diam)
            assert abs(fa + fb - 2 * math.pi) < 1e-9
            assert abs(orc.ellipse_f32(a, dx, dz, diam, 0) - fa) < 1e-4
            assert abs(orc.ellipse_f32(a, dx, dz, diam, 1) - fa) < 1e-4
        # monotone increasing on (0, pi)
        xs = np.linspace(0.05, math.pi - 0.05, 40)
        ys = [orc.ellipse_f64(x, dx, dz, diam) for x in xs]
        assert all(b > a for a, b in zip(ys, ys[1:]))
    for a in (0.1, 1.2, 3.0, 4.5, 6.0):
        assert abs(orc.ellipse_f64(a, 0.0, 1417.0, 365.0) - a) < 1e-7


def ellipse_from_geometry(alpha, dx, dz, diam):
    """The tilted-ring map built from its geometry, not from the reference's
    text: the source ring point S(a) = (R cos a, R sin a, 0), the tilt centre
    Q = (-dX, 0, -dZ); eps(a) is the 3-D angle at Q between Q->O and Q->S(a)
    (atan2 of |cross| and dot), gamma = (eps(pi) - eps(0)) / 2, and the map is
    the polar angle of (eps cos a + gamma, eps sin a) in [0, 2 pi)."""
    a = np.asarray(alpha, np.float64)
    R = diam / 2.0
    u = np.array([dx, 0.0, dz])                                        # Q -> O

    def eps(al):
        v = np.stack([R * np.cos(al) + dx, R * np.sin(al), np.full_like(al, dz)], -1)   # Q -> S
        cr = np.cross(np.broadcast_to(u, v.shape), v)
        return np.arctan2(np.linalg.norm(cr, axis=-1), v @ u)

    e = eps(a)
    gam = 0.5 * (eps(np.array([math.pi]))[0] - eps(np.array([0.0]))[0])
    out = np.arctan2(e * np.sin(a), e * np.cos(a) + gam)
    return np.where(out < 0, out + 2 * math.pi, out)


@pytest.mark.parametrize("dx,dz,diam", [(815.0, 1417.0, 365.0), (815.0, 1430.0, 370.0), (300.0, 900.0, 500.0),
                                        (-400.0, 1200.0, 300.0)])
def test_ellipse_map_matches_its_geometry(orc, dx, dz, diam):
    """orc_ellipse_kreis_f64 / _f32 (cuda_kernels_fan2para.h:38-62) against the
    map constructed from the tilted-ring geometry.  A sign slip in eps's
    numerator (L^2 + R dX cos a), in gamma's order, or in the (eps cos a +
    gamma, eps sin a) vector moves the result by far more than the bounds.
    FP64: 1e-12 rad where |sin a| > 0.01; acos near +-1 (outputs near 0 and
    pi) loses up to sqrt(2 eps) ~ 2e-8 there.  FP32 (both libm bindings):
    2e-5 rad where |sin a| > 0.05, 1e-3 everywhere (float acos near +-1)."""
    a = np.concatenate([np.linspace(1e-4, 2 * math.pi - 1e-4, 2501), [math.pi / 2, math.pi, 1.5 * math.pi]])
    want = ellipse_from_geometry(a, dx, dz, diam)
    got = np.array([orc.ellipse_f64(x, dx, dz, diam) for x in a])
    away = np.abs(np.sin(a)) > 0.01
    assert np.abs(got - want)[away].max() < 1e-12
    assert np.abs(got - want).max() < 1e-7
    far = np.abs(np.sin(a)) > 0.05
    for variant in (0, 1):
        g32 = np.array([orc.ellipse_f32(np.float32(x), dx, dz, diam, variant) for x in a], np.float64)
        w32 = ellipse_from_geometry(a.astype(np.float32).astype(np.float64), dx, dz, diam)
        err = np.abs(g32 - w32)
        err = np.minimum(err, 2 * math.pi - err)
        assert err[far].max() < 2e-5, variant
        assert err.max() < 1e-3, variant
    # the map is not the identity under tilt (the invariants alone would pass that)
    if dx != 0:
        assert np.abs(want - np.where(a < 2 * math.pi, a, a - 2 * math.pi)).max() > 1e-3


def small_geom(risa, **kw):
    base = dict(numberOfFanDetectors=96, numberOfFanProjections=120, numberOfParallelProjections=60,
                numberOfParallelDetectors=64, numberOfPlanes=2)
    base.update(kw)
    return rofex_geom(risa, **base)


def brute_force_tables(tab, k):
    """Literal brute force of computeAngles' searches (Fan2Para.cu:401-462),
    vectorised: float differences, strict improvement from dif_best = fl32(pi),
    best_x carried from ray 1 into ray 2 (Q5)."""
    theta = tab.theta.astype(np.float32)
    gamma = tab.gamma.astype(np.float32)
    Pf, Df = len(theta), len(gamma)
    pi32 = np.float32(math.pi)
    out = {}
    carried = np.zeros(tab.t.P2 * tab.t.D, np.int64)
    for r in (1, 2):
        goal = tab.plane(f"thetaGoal{r}", k).astype(np.float32)
        flag = tab.plane(f"ray{r}", k).astype(bool)
        dif = (theta[None, :] - goal[:, None]).astype(np.float32)
        ok = (goal[:, None] <= theta[None, :]) & (dif < pi32)
        d = np.where(ok, dif, np.float32(np.inf))
        best = np.argmin(d, axis=1)                      # first occurrence on ties
        found = np.isfinite(d[np.arange(len(goal)), best])
        best = np.where(found, best, carried)
        carried = np.where(flag, best, carried)
        tb = np.where(best == 0, Pf - 1, best - 1)
        gg = tab.plane(f"gammaGoal{r}", k).astype(np.float32)
        ge = gg[:, None] <= gamma[None, :]
        x = np.argmax(ge, axis=1)
        gb = np.where(x == 0, Df - 1, x - 1)
        ga = x
        over = gg > gamma[-1]
        gb = np.where(over, Df - 1, gb)
        ga = np.where(over, 0, ga)
        z = np.zeros_like(best)
        out[f"thetaBefore{r}"] = np.where(flag, tb, z)
        out[f"thetaAfter{r}"] = np.where(flag, best, z)
        out[f"gammaBefore{r}"] = np.where(flag, gb, z)
        out[f"gammaAfter{r}"] = np.where(flag, ga, z)
    return out


@pytest.mark.parametrize("kw", [{}, {"sourceOffset": 0.0}, {"sourceOffset": 71.3, "imageWidth": 150.0}],
                         ids=["rofex", "offset0", "offset71"])
def test_rebin_tables_brute_force_and_invariants(orc, risa, kw):
    """computeFan2ParaTransp / computeAngles (Fan2Para.cu:251-480): indices
    equal a brute-force scan; flags only where |s/r_D| <= 1; alpha_0 = 0 (Q4);
    theta rotated by sourceOffset; theta/gamma 'after' = 'before' + 1 mod n."""
    g = small_geom(risa, **kw)
    tab = orc.Tables(g, 0)
    P2, D, Pf, Df = tab.t.P2, tab.t.D, tab.t.Pf, tab.t.Df
    assert P2 == 2 * g.numberOfParallelProjections
    assert tab.alpha[0] == 0.0
    j = np.arange(1, P2)
    # alpha is stored as float before the > 2 pi test: fl32(2 pi) > 2 pi, so
    # the entry at exactly 2 pi wraps to ~1.7e-7 (Fan2Para.cu:303-309)
    a32 = ((2 * math.pi / 360.0) * (j * (360.0 / P2)).astype(np.float32) + math.pi / 2).astype(np.float32)
    exp_alpha = np.where(a32.astype(np.float64) > 2 * math.pi, a32 - 2 * math.pi, a32)
    np.testing.assert_allclose(tab.alpha[1:], exp_alpha, atol=2e-6)
    if P2 % 4 == 0:
        assert tab.alpha[3 * P2 // 4] < 1e-6
    th = np.deg2rad(np.arange(Pf) * 360.0 / Pf - g.sourceOffset)
    th = np.where(th < 0, th + 2 * math.pi, th)
    np.testing.assert_allclose(tab.theta, th, atol=2e-6)
    s = -g.imageWidth / 2 + (0.5 + np.arange(D)) * g.imageWidth / D
    np.testing.assert_allclose(tab.s, s, atol=1e-4)
    in_range = (np.abs(s) / (g.detectorDiameter / 2) <= 1)[None, :].repeat(P2, 0).ravel()
    for k in range(g.numberOfPlanes):
        bf = brute_force_tables(tab, k)
        for name, arr in bf.items():
            np.testing.assert_array_equal(tab.plane(name, k), arr, err_msg=f"{name} plane {k}")
        for r in (1, 2):
            flag = tab.plane(f"ray{r}", k).astype(bool)
            assert not (flag & ~in_range).any()
            tb, ta = tab.plane(f"thetaBefore{r}", k), tab.plane(f"thetaAfter{r}", k)
            gb, ga = tab.plane(f"gammaBefore{r}", k), tab.plane(f"gammaAfter{r}", k)
            assert ((ta[flag] - (tb[flag] + 1) % Pf) == 0).all()
            assert ((ga[flag] - (gb[flag] + 1) % Df) == 0).all()
        assert tab.plane("ray1", k).sum() > 0 and tab.plane("ray2", k).sum() > 0
    tab.close()


def test_rebin_goal_angles_identity_tilt(orc, risa):
    """With deltaX = 0 the tilt map is the identity, so the goal angles have the
    closed form thetaGoal1 = wrap(alpha - asin(s/r_T)), thetaGoal2 =
    wrap(alpha + asin(s/r_T) - pi) (reflected at 2 pi), gammaGoal1 = wrap(asin(s/r_D) + alpha -
    1.5 pi), gammaGoal2 = wrap(-asin(s/r_D) + alpha - pi/2)
    (Fan2Para.cu:366-399, 441-443)."""
    g = small_geom(risa, sourceAngle=(360.0,) * 8)
    g.deltaX[0] = 0.0
    tab = orc.Tables(g, 0)
    P2, D = tab.t.P2, tab.t.D
    a = tab.alpha.astype(np.float64)[:, None]
    s = tab.s.astype(np.float64)[None, :]
    rT, rD = g.sourceDiameter[0] / 2, g.detectorDiameter / 2
    wrap = lambda x: np.where(x < 0, x + 2 * math.pi, x)   # noqa: E731
    # only negative goals are wrapped; one above 2 pi meets the map's
    # alpha > pi branch, 2 pi - acos(cos alpha) = 4 pi - alpha
    refl = lambda x: np.where(x > 2 * math.pi, 4 * math.pi - x, x)   # noqa: E731
    t1 = refl(wrap(a - np.arcsin(s / rT))).ravel()
    t2 = refl(wrap(a + np.arcsin(s / rT) - math.pi)).ravel()
    g1 = wrap(np.arcsin(s / rD) + a - 1.5 * math.pi)
    g1 = np.where(g1 > 2 * math.pi, g1 - 2 * math.pi, g1).ravel()
    g2 = wrap(-np.arcsin(s / rD) + a - math.pi / 2).ravel()
    r1 = tab.plane("ray1", 0).astype(bool)
    r2 = tab.plane("ray2", 0).astype(bool)

    def close(x, y):
        d = np.abs(x - y)
        return np.minimum(d, 2 * math.pi - d)            # angles equal modulo the wrap point

    # float acos near +-1 (goals near 0 and pi) loses up to sqrt(2 ulp) ~ 5e-4
    for got, exp in ((tab.plane("thetaGoal1", 0), t1), (tab.plane("thetaGoal2", 0), t2)):
        err = close(got, exp)
        assert err.max() < 5e-4 and np.quantile(err, 0.99) < 1e-5
    assert close(tab.plane("gammaGoal1", 0)[r1], g1[r1]).max() < 2e-5
    assert close(tab.plane("gammaGoal2", 0)[r2], g2[r2]).max() < 2e-5
    tab.close()


def test_rebin_tables_product_host_bit_exact(orc, risa, geom_c0):
    """The product's table builder (binary search) against the oracle's
    literal scan: every integer and float entry equal."""
    from risa_amd import _lib

    lib = _lib.load()
    for g in (geom_c0, small_geom(risa, sourceOffset=71.3)):
        tab = orc.Tables(g, 0)
        n = tab.t.P2 * tab.t.D
        for k in range(g.numberOfPlanes):
            ints = np.empty((10, n), np.int32)
            flts = np.empty((4, n), np.float32)
            _lib.check(lib.risa_fan2para_tables(g, k, _lib.ptr(ints), _lib.ptr(flts)))
            for i, name in enumerate(orc.Tables.INT_NAMES):
                np.testing.assert_array_equal(ints[i], tab.plane(name, k), err_msg=name)
            for i, name in enumerate(orc.Tables.FLOAT_NAMES):
                np.testing.assert_array_equal(flts[i], tab.plane(name, k), err_msg=name)
        tab.close()


def test_rebin_tables_product_host_bit_exact_c1_both_planes(orc, risa, geom_c1):
    """The production geometry (768 -> 512 detectors, 2000 internal angles,
    two tilted planes): the product's binary-search tables equal the
    oracle's literal scan entry for entry on both planes.  Also counts the
    integer entries that change when the reference's unqualified libm calls
    bind to the double overloads (the one choice only nvcc could settle)."""
    from risa_amd import _lib

    lib = _lib.load()
    g = geom_c1
    tab = orc.Tables(g, 0)
    tab64 = orc.Tables(g, 1)
    n = tab.t.P2 * tab.t.D
    differing = []
    for k in range(g.numberOfPlanes):
        ints = np.empty((10, n), np.int32)
        flts = np.empty((4, n), np.float32)
        _lib.check(lib.risa_fan2para_tables(g, k, _lib.ptr(ints), _lib.ptr(flts)))
        for i, name in enumerate(orc.Tables.INT_NAMES):
            np.testing.assert_array_equal(ints[i], tab.plane(name, k), err_msg=f"{name} plane {k}")
        for i, name in enumerate(orc.Tables.FLOAT_NAMES):
            np.testing.assert_array_equal(flts[i], tab.plane(name, k), err_msg=f"{name} plane {k}")
        assert tab.plane("ray1", k).sum() > 0 and tab.plane("ray2", k).sum() > 0
        differing.append(int(sum((tab.plane(nm, k) != tab64.plane(nm, k)).sum() for nm in orc.Tables.INT_NAMES)))
    print(f"C1 integer entries changed by the double libm binding: plane 0 {differing[0]}, plane 1 {differing[1]} "
          f"of {10 * n} each")
    assert max(differing) < 1e-4 * 10 * n
    tab.close()
    tab64.close()


# ------------------------------------------------------------------ a6 gather --
def test_gather_reproduces_linear_functions(orc, risa):
    """interpolation (cuda_kernels_fan2para.h:64-177): bilinear interpolation
    in (theta, gamma) reproduces a function linear in the angle values exactly
    wherever neither bracket wraps; rays are flag-weighted and the fold adds
    0.5*Z(j, i) + 0.5*Z(j+P, D-1-i)."""
    g = small_geom(risa)
    tab = orc.Tables(g, 0)
    P2, D = tab.t.P2, tab.t.D
    P = P2 // 2
    A, B, Cc = 0.7, 1.3, -2.1
    fan = (A + B * tab.theta.astype(np.float64)[:, None] + Cc * tab.gamma.astype(np.float64)[None, :])
    th, ga = tab.theta.astype(np.float64), tab.gamma.astype(np.float64)
    for k in range(2):
        Z = np.zeros(P2 * D)
        good = np.ones(P2 * D, bool)
        w = np.zeros(P2 * D)
        for r in (1, 2):
            f = tab.plane(f"ray{r}", k).astype(bool)
            tb, ta = tab.plane(f"thetaBefore{r}", k), tab.plane(f"thetaAfter{r}", k)
            gb, gaa = tab.plane(f"gammaBefore{r}", k), tab.plane(f"gammaAfter{r}", k)
            good &= ~f | ((th[ta] > th[tb]) & (ga[gaa] > ga[gb]))
            zr = A + B * tab.plane(f"thetaGoal{r}", k).astype(np.float64) \
                + Cc * tab.plane(f"gammaGoal{r}", k).astype(np.float64)
            Z += np.where(f, zr, 0.0)
            w += f
        Z = np.where(w > 0, Z / np.maximum(w, 1), 0.0).reshape(P2, D)
        good = good.reshape(P2, D)
        exp = 0.5 * Z[:P] + 0.5 * Z[P:, ::-1]
        ok = good[:P] & good[P:, ::-1]
        assert ok.mean() > 0.9
        got32 = orc.gather_f32(g, tab, k, fan.astype(np.float32))
        got64 = orc.gather_f64(g, tab, k, fan)
        assert np.abs(got32 - exp)[ok].max() < 5e-5
        # the FP64 gather recomputes the goal angles in double; they differ from
        # the float table's by the float acos error near 0 and pi (< 5e-4 rad)
        e64 = np.abs(got64 - exp)[ok]
        assert e64.max() < 1e-3 and np.quantile(e64, 0.99) < 1e-5
    tab.close()


# -------------------------------------------------------------------- a7 taps --
@pytest.mark.parametrize("D", [256, 384, 432, 512, 768])
def test_taps_known_answers(orc, risa, D):
    """designFilter (Filter.cu:169-191): h[0] = 0; the Nyquist tap is 0 at
    cutoff 1.0 because fl32(pi) > pi (Q8); ramp h[i] = fl32(fl32(2i/D)/D);
    windows multiply the ramp by sinc, cos, Hamming, Hann of w/2 resp. w."""
    g = rofex_geom(risa, numberOfParallelDetectors=D, cutoffFraction=1.0)
    i = np.arange(1, D // 2 + 1)
    w = (2 * math.pi * i / D).astype(np.float32).astype(np.float64)     # w is a float (Filter.cu:175)
    w64 = 2 * math.pi * i / D
    ramp = (2.0 * i / D).astype(np.float32)
    h = orc.taps_f32(g.replace(filterType=0))
    assert h[0] == 0.0 and h[-1] == 0.0
    np.testing.assert_array_equal(h[1:-1], (ramp / np.float32(D)).astype(np.float32)[:-1])
    def windows(w):
        return {1: np.sinc(w / (2 * math.pi)), 2: np.cos(w / 2), 3: 0.54 + 0.46 * np.cos(w),
                4: 0.5 * (1 + np.cos(w))}

    for ft in (1, 2, 3, 4):
        h = orc.taps_f32(g.replace(filterType=ft))
        h64 = orc.taps_f64(g.replace(filterType=ft))
        exp = 2.0 * i / D * windows(w)[ft] / D
        exp64 = 2.0 * i / D * windows(w64)[ft] / D
        assert h[0] == 0.0 and h[-1] == 0.0
        # float storage and the float cos of Hamming/Hann: 1e-6 of the largest tap
        np.testing.assert_allclose(h[1:-1], exp[:-1], rtol=2e-6, atol=1e-6 * exp.max())
        np.testing.assert_allclose(h64[1:-1], exp64[:-1], rtol=1e-12, atol=1e-15)
    # cutoff 0.5 zeroes every tap with w > pi/2
    h = orc.taps_f32(g.replace(filterType=0, cutoffFraction=0.5))
    assert (h[D // 4 + 1:] == 0).all() and (h[1:D // 4] > 0).all()


def test_taps_product_host_bit_exact(orc, risa):
    from risa_amd import _lib

    lib = _lib.load()
    for D in (256, 512, 768):
        for ft in range(5):
            for cut in (1.0, 0.7):
                g = rofex_geom(risa, numberOfParallelDetectors=D, filterType=ft, cutoffFraction=cut)
                out = np.empty(D // 2 + 1, np.float32)
                _lib.check(lib.risa_filter_taps(g, _lib.ptr(out)))
                np.testing.assert_array_equal(out, orc.taps_f32(g))


# ------------------------------------------------------------------ a8 filter --
@pytest.mark.parametrize("D", [256, 384, 500, 512, 768])
def test_filter_rows_match_dft_definition(orc, D):
    """Filter::processor (Filter.cu:102-143): y = C2R_D(h . R2C_D(x)),
    unnormalised, length D (circular).  numpy's rfft/irfft define the DFT
    independently of the oracle's direct circulant and Stockham FFT."""
    rng = np.random.default_rng(D)
    x = rng.standard_normal((7, D))
    h = rng.uniform(0, 1, D // 2 + 1)
    exp = np.fft.irfft(np.fft.rfft(x, axis=1) * h[None, :], n=D, axis=1) * D
    got64 = orc.filter_rows_f64(h, x)
    np.testing.assert_allclose(got64, exp, rtol=1e-9, atol=1e-9)
    got32 = orc.filter_rows_f32(h.astype(np.float32), x.astype(np.float32))
    assert rel_rmse(got32, exp) < 1e-6
    # the circulant kernel c[d] is the inverse DFT of the symmetric response
    c = orc.filter_kernel_f64(h, D)
    full = np.concatenate([h, h[1:D - D // 2][::-1]]) if D % 2 == 0 else np.concatenate([h, h[1:][::-1]])
    np.testing.assert_allclose(c, np.fft.ifft(full[:D]).real * D, atol=1e-9)


def test_filter_grad_taps_finite_differences(orc):
    """d<filter(h, x), g>/dh: the loss is linear in h, so central differences
    at any step are exact up to rounding."""
    rng = np.random.default_rng(5)
    D = 64
    x = rng.standard_normal((5, D))
    gr = rng.standard_normal((5, D))
    h = rng.uniform(0, 1, D // 2 + 1)
    dh = orc.grad_taps_f64(x, gr)
    for k in (0, 1, 7, D // 2 - 1, D // 2):
        e = np.zeros_like(h)
        e[k] = 1e-3
        fd = (np.sum(orc.filter_rows_f64(h + e, x) * gr) - np.sum(orc.filter_rows_f64(h - e, x) * gr)) / 2e-3
        assert abs(fd - dh[k]) < 1e-8 * max(1.0, abs(fd)), k


# ----------------------------------------------------------- a9 / a10 backprojection --
def test_bp_lut_known_answers(orc, risa):
    """Backprojection.cu:115-148: theta_i = fl32(i*pi/P + rot*pi/180), +2pi while
    negative; sin/cos in float; scale = D/N, norm = pi/P/scale, centre (N-1)/2."""
    for rot, D, N, P in ((0.0, 256, 256, 500), (-30.0, 512, 384, 360), (12.5, 300, 512, 720)):
        g = rofex_geom(risa, numberOfParallelDetectors=D, numberOfPixels=N, numberOfParallelProjections=P,
                       rotationOffset=rot)
        s, c, k = orc.bp_lut_f32(g)
        th = (np.arange(P) * math.pi / P + rot * math.pi / 180).astype(np.float32).astype(np.float64)
        th = np.where(th < 0, th + 2 * math.pi, th)
        ulp = np.spacing(np.float32(1.0))
        assert np.abs(s - np.sin(th)).max() <= 2 * ulp
        assert np.abs(c - np.cos(th)).max() <= 2 * ulp
        assert k[0] == np.float32(D / N)
        assert abs(k[1] - math.pi / P / (D / N)) < 1e-7
        assert k[2] == np.float32((N - 1) / 2)
        from risa_amd import _lib

        ps, pc, pk = (np.empty(P, np.float32), np.empty(P, np.float32), np.empty(3, np.float32))
        _lib.check(_lib.load().risa_bp_lut(g, _lib.ptr(ps), _lib.ptr(pc), _lib.ptr(pk)))
        np.testing.assert_array_equal(ps, s)
        np.testing.assert_array_equal(pc, c)
        np.testing.assert_array_equal(pk, k)


def test_bp_constant_sinogram_known_answer(orc, risa):
    """backProjectLinear (Backprojection.cu:238-271): a sinogram constant at c
    gives c*pi*N/D wherever both taps stay inside the detector for every angle
    (the two weights sum to 1 and norm = pi/P * N/D)."""
    for D, N, P in ((64, 64, 90), (96, 64, 50), (64, 80, 33)):
        g = rofex_geom(risa, numberOfParallelDetectors=D, numberOfPixels=N, numberOfParallelProjections=P)
        c = 0.75
        img = orc.bp_linear_f64(g, np.full((P, D), c))
        img32 = orc.bp_linear_f32(g, np.full((P, D), c, np.float32))
        scale = D / N
        y, x = np.mgrid[0:N, 0:N]
        r = np.hypot((x - (N - 1) / 2) * scale, (y - (N - 1) / 2) * scale)
        inner = r < D // 2 - 2
        np.testing.assert_allclose(img[inner], c * math.pi * N / D, rtol=1e-12)
        np.testing.assert_allclose(img32[inner], c * math.pi * N / D, rtol=2e-6)


def test_bp_brute_force_tiny(orc, risa):
    """Per-pixel brute force of backProjectLinear in pure Python on a tiny grid,
    with the reference's index convention k = floor(t) + D//2 (Q10)."""
    D, N, P = 11, 7, 5
    g = rofex_geom(risa, numberOfParallelDetectors=D, numberOfPixels=N, numberOfParallelProjections=P,
                   rotationOffset=7.0)
    rng = np.random.default_rng(9)
    sino = rng.standard_normal((P, D))
    th = np.arange(P) * math.pi / P + 7.0 * math.pi / 180      # FP64 angles, as the FP64 form uses
    sT, cT = np.sin(th), np.cos(th)
    scale, cen = D / N, (N - 1) / 2
    norm = math.pi / P / scale
    exp = np.zeros((N, N))
    for yy in range(N):
        for xx in range(N):
            xp, yp = (xx - cen) * scale, (yy - cen) * scale
            acc = 0.0
            for p in range(P):
                t = xp * float(cT[p]) + yp * float(sT[p])
                a = math.floor(t)
                kk = a + D // 2
                if 0 <= kk < D:
                    acc += (a + 1 - t) * sino[p, kk]
                if 0 <= kk + 1 < D:
                    acc += (t - a) * sino[p, kk + 1]
            exp[yy, xx] = acc * norm
    np.testing.assert_allclose(orc.bp_linear_f64(g, sino), exp, rtol=1e-12, atol=1e-12)
    assert rel_rmse(orc.bp_linear_f32(g, sino.astype(np.float32)), exp) < 1e-6


def _f32_round(fr):
    """Round an exact rational to the nearest float32 (ties to even), the way
    an IEEE single-precision operation rounds its exact result."""
    from fractions import Fraction

    if fr == 0:
        return 0.0
    sign = -1 if fr < 0 else 1
    m = abs(Fraction(fr))
    e = m.numerator.bit_length() - m.denominator.bit_length()
    if Fraction(2) ** e > m:
        e -= 1
    q = m / Fraction(2) ** (e - 23)                  # in [2^23, 2^24)
    n, r = divmod(q.numerator, q.denominator)
    if 2 * r > q.denominator or (2 * r == q.denominator and n % 2 == 1):
        n += 1
    return sign * float(Fraction(n) * Fraction(2) ** (e - 23))


def _f32_ops():
    from fractions import Fraction

    F = Fraction
    mul = lambda a, b: _f32_round(F(a) * F(b))               # noqa: E731
    add = lambda a, b: _f32_round(F(a) + F(b))               # noqa: E731
    fma = lambda a, b, c: _f32_round(F(a) * F(b) + F(c))     # noqa: E731
    return mul, add, fma


def _round_half_away(v):
    return math.floor(v + 0.5) if v >= 0 else math.ceil(v - 0.5)


@pytest.mark.parametrize("N,P,D,rot", [(7, 11, 5, 0.0), (8, 6, 9, 17.0), (6, 13, 4, -40.0), (9, 7, 8, 90.0),
                                       (8, 4, 8, 0.0)])
def test_bp_nearest_and_tex_brute_force(orc, risa, N, P, D, rot):
    """backProjectNearest and backProjectTex (Backprojection.cu:273-336) by
    per-pixel brute force in pure Python with every float32 operation rounded
    exactly (rational arithmetic): nearest t = round-half-away-from-zero of
    xp cos + yp sin (the multiply-add contracted as fma(yp, sin, xp cos), the
    oracle's stated reading of nvcc's default contraction) plus floor(D/2),
    xp = (x - (N-1)/2) / (N/D) in double, result sum * pi / P * (N/D) in
    double; tex t = int() truncation toward zero of xp cos + fma(yp, sin, D/2),
    xp = (x - c) * (D/N) in float, result sum * norm.  Bit for bit; truncation
    vs floor, half-to-even rounding, a swapped scale or a missing bounds test
    each change pixels here (t reaches -D/2-1 .. D/2+1)."""
    g = rofex_geom(risa, numberOfParallelDetectors=D, numberOfPixels=N, numberOfParallelProjections=P,
                   rotationOffset=rot)
    s, c, k = orc.bp_lut_f32(g)
    rng = np.random.default_rng(N * 100 + P * 10 + D)
    sino = rng.standard_normal((P, D)).astype(np.float32)
    mul, add, fma = _f32_ops()
    f = lambda v: float(np.float32(v))                      # noqa: E731
    near = np.zeros((N, N), np.float32)
    tex = np.zeros((N, N), np.float32)
    from fractions import Fraction

    scale_n = _f32_round(Fraction(N) / Fraction(D))         # (float)N / (float)D
    cI_int = int(D * 0.5)
    cI_f = f(D * 0.5)
    for y in range(N):
        for x in range(N):
            # nearest
            xp = f((x - (N - 1.0) * 0.5) / scale_n)
            yp = f((y - (N - 1.0) * 0.5) / scale_n)
            acc = 0.0
            for p in range(P):
                tt = fma(yp, float(s[p]), mul(xp, float(c[p])))
                t = int(add(float(_round_half_away(tt)), float(cI_int)))
                if 0 <= t < D:
                    acc = add(acc, float(sino[p, t]))
            near[y, x] = f(acc * math.pi / P * scale_n)
            # tex
            xq = mul(add(float(x), -float(k[2])), float(k[0]))
            yq = mul(add(float(y), -float(k[2])), float(k[0]))
            acc = 0.0
            for p in range(P):
                tt = add(mul(xq, float(c[p])), fma(yq, float(s[p]), cI_f))
                t = int(tt)
                if 0 <= t < D:
                    acc = add(acc, float(sino[p, t]))
            tex[y, x] = mul(acc, float(k[1]))
    np.testing.assert_array_equal(orc.bp_nearest_f32(g, sino), near)
    np.testing.assert_array_equal(orc.bp_tex_f32(g, sino), tex)


def test_bp_nearest_and_tex_constant_sinogram(orc, risa):
    """A sinogram constant at c: every in-range sample adds c, so wherever the
    picked sample stays inside the detector for all angles both nearest forms
    give P c * pi / P * N/D = c pi N/D (to float rounding of the sum)."""
    for D, N, P in ((64, 64, 90), (96, 64, 50), (64, 80, 33)):
        g = rofex_geom(risa, numberOfParallelDetectors=D, numberOfPixels=N, numberOfParallelProjections=P)
        cval = 0.75
        sino = np.full((P, D), cval, np.float32)
        scale = D / N
        y, x = np.mgrid[0:N, 0:N]
        r = np.hypot((x - (N - 1) / 2) * scale, (y - (N - 1) / 2) * scale)
        inner = r < D // 2 - 2
        for img in (orc.bp_nearest_f32(g, sino), orc.bp_tex_f32(g, sino)):
            np.testing.assert_allclose(img[inner], cval * math.pi * N / D, rtol=2e-6)
            # outside the detector's reach no sample is picked at all for some angle
            assert (img[r > D / 2 + 2] < cval * math.pi * N / D * (1 - 0.5 / P)).all()


def test_bp_adjoint_dot_product(orc, risa):
    """<BP s, x> = <s, BP^T x> for the ray-driven adjoint, in FP64."""
    D, N, P = 64, 48, 40
    g = rofex_geom(risa, numberOfParallelDetectors=D, numberOfPixels=N, numberOfParallelProjections=P,
                   rotationOffset=3.0)
    rng = np.random.default_rng(2)
    s = rng.standard_normal((P, D))
    x = rng.standard_normal((N, N))
    lhs = float(np.sum(orc.bp_linear_f64(g, s) * x))
    rhs = float(np.sum(s * orc.bp_adjoint_f64(g, x)))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


# -------------------------------------------------------------- a12 masking --
def test_masking_closed_form(orc, risa):
    """Masking (Masking.cu:113-169): pixels with (x-c)^2 + (y-c)^2 > N^2/4
    become maskingValue; normalisation maps [min, max] to [0, 1] first."""
    N = 33
    g = rofex_geom(risa, numberOfPixels=N, maskingValue=-7.0, normalization=0)
    rng = np.random.default_rng(4)
    img = rng.standard_normal((N, N)).astype(np.float32)
    c = (N - 1) / 2
    y, x = np.mgrid[0:N, 0:N]
    outside = (x - c) ** 2 + (y - c) ** 2 > N * N / 4
    got = orc.masking_f32(g, img)
    assert (got[outside] == -7.0).all()
    np.testing.assert_array_equal(got[~outside], img[~outside])
    gn = g.replace(normalization=1)
    got = orc.masking_f64(gn, img.astype(np.float64))
    exp = (img.astype(np.float64) - img.min()) / (img.max() - img.min())
    np.testing.assert_allclose(got[~outside], exp[~outside], atol=1e-12)
    assert (got[outside] == -7.0).all()


# ----------------------------------------------------------------- chain --
def test_chain_f32_vs_f64_small(orc, risa, geom_c0, frames_c0):
    """The faithful FP32 chain and the FP64 chain agree on a C0 frame (the
    same formulas, the same integer tables).  The gap, ~5e-5 relative RMSE, is
    the reference arithmetic's own: float interpolation weights from float goal
    angles, which lose up to ~5e-4 rad through float acos near 0 and pi."""
    fs = frames_c0
    ch = orc.Chain(geom_c0, fs.avg_ref, fs.avg_dark)
    img32 = ch.f32(fs.raw[:1], fs.planes[:1])[0]
    img64 = ch.f64(fs.raw[0], int(fs.planes[0]))
    from conftest import disk_mask

    m = disk_mask(geom_c0.numberOfPixels)
    assert rel_rmse(img32, img64, m) < 1e-4
    ch.close()
