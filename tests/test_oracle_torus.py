"""Oracle torus against the reference's known answers (TestTorus2.cpp) and
what the geometry fixes: the implicit surface (rho - R)^2 + z^2 = r^2, the
first crossing of it along a ray found by bisection, and its gradient."""
import math

import numpy as np
import pytest

import oracle
from tests.golden import kats

R_TOR, R_MAX, R_MIN = 100.0, 90.0, 10.0


def _pv(geo, params):
    return geo.place(geo.logical_volume(geo.torus(*params)))


def _q(geo, pv, method, p, d=None):
    pts = [np.atleast_1d(np.asarray(c, dtype=np.float64)) for c in p]
    dirs = None if d is None else [np.atleast_1d(np.asarray(c, dtype=np.float64)) for c in d]
    return geo.query(pv, method, pts, dirs)


def _radsq(p, rtor):
    rho = np.sqrt(p[0] ** 2 + p[1] ** 2)
    return (rho - rtor) ** 2 + p[2] ** 2


@pytest.mark.parametrize("case", kats.TORUS_NORMAL_KATS)
def test_torus_normal_kat(case):
    params, p, want = case
    geo = oracle.Geometry()
    nx, ny, nz, valid = geo.normal(_pv(geo, params), [np.array([float(c)]) for c in p])
    assert valid[0]
    assert all(kats.approx_equal(a, b) for a, b in zip((nx[0], ny[0], nz[0]), want)), (nx, ny, nz, want)


def _random_rays(rng, n, box):
    p = rng.uniform(-box, box, size=(3, n))
    d = rng.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    return p, d


def _first_crossing(p, d, rtor, rad, inside_start, tmax=600.0, nstep=60000):
    """First t at which the ray crosses (rho - rtor)^2 + z^2 = rad^2: a fine
    march, then bisection on the bracketing step."""
    t = np.linspace(0.0, tmax, nstep)
    pts = p[:, None] + d[:, None] * t[None, :]
    f = _radsq(pts, rtor) - rad * rad
    sign = f < 0 if inside_start else f > 0
    k = np.argmin(sign)
    if sign[k]:
        return math.inf
    lo, hi = t[max(k - 1, 0)], t[k]
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        fm = _radsq(p + d * mid, rtor) - rad * rad
        if (fm < 0) == inside_start:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def test_full_torus_inside_is_the_definition():
    geo = oracle.Geometry()
    rng = np.random.default_rng(11)
    p = rng.uniform(-200, 200, size=(3, 20000))
    p[2] *= 0.5
    got = _q(geo, _pv(geo, (R_MIN, R_MAX, R_TOR, 0.0, 2 * np.pi)), "Inside", p)
    r2 = _radsq(p, R_TOR)
    tol = 100e-9
    want = np.full(p.shape[1], 2)
    want[(r2 > R_MAX * R_MAX + tol * R_MAX) | (r2 < R_MIN * R_MIN - tol * R_MIN)] = 3
    want[(r2 < R_MAX * R_MAX - tol * R_MAX) & (r2 > R_MIN * R_MIN + tol * R_MIN)] = 1
    np.testing.assert_array_equal(got, want)


def test_full_torus_distances_match_bisection():
    """The quartic + Newton distances against the first crossing of the
    implicit surface found by marching and bisection."""
    geo = oracle.Geometry()
    rtor, rmax = 100.0, 40.0
    pv = _pv(geo, (0.0, rmax, rtor, 0.0, 2 * np.pi))
    rng = np.random.default_rng(12)
    p, d = _random_rays(rng, 600, 180.0)
    # half of the rays aimed at a point of the solid
    phi = rng.uniform(0, 2 * np.pi, 300)
    tgt = np.stack([rtor * np.cos(phi), rtor * np.sin(phi), rng.uniform(-20, 20, 300)])
    d[:, :300] = tgt - p[:, :300]
    d[:, :300] /= np.linalg.norm(d[:, :300], axis=0)
    inside = _q(geo, pv, "Inside", p)
    din = _q(geo, pv, "DistanceToIn", p, d)
    dout = _q(geo, pv, "DistanceToOut", p, d)
    n_in = n_out = 0
    for i in range(p.shape[1]):
        if inside[i] == 3:
            want = _first_crossing(p[:, i], d[:, i], rtor, rmax, False)
            if math.isinf(want):
                assert din[i] >= kats.KINF
            else:
                assert abs(din[i] - want) < 1e-7 * max(1.0, want), (i, din[i], want)
                n_in += 1
        elif inside[i] == 1:
            want = _first_crossing(p[:, i], d[:, i], rtor, rmax, True)
            assert abs(dout[i] - want) < 1e-7 * max(1.0, want), (i, dout[i], want)
            n_out += 1
    assert n_in > 150 and n_out > 20


def test_hollow_torus_exit_through_the_inner_surface():
    geo = oracle.Geometry()
    pv = _pv(geo, (R_MIN, R_MAX, R_TOR, 0.0, 2 * np.pi))
    # halfway between the inner (10) and outer (90) tube radii: 40 either way
    assert _q(geo, pv, "DistanceToOut", (R_TOR + 50, 0.0, 0.0), (1.0, 0.0, 0.0))[0] == pytest.approx(40.0, abs=1e-9)
    assert _q(geo, pv, "DistanceToOut", (R_TOR + 50, 0.0, 0.0), (-1.0, 0.0, 0.0))[0] == pytest.approx(40.0, abs=1e-9)
    # from the tube's centre circle, in the inner cavity, up to the inner surface
    assert _q(geo, pv, "Inside", (R_TOR, 0.0, 0.0))[0] == 3
    assert _q(geo, pv, "DistanceToIn", (R_TOR, 0.0, 0.0), (0.0, 0.0, 1.0))[0] == pytest.approx(10.0, abs=1e-9)


def test_clad_core_chain():
    """TestTorus2.cpp:223-245: along +y through a clad (r 10) around a core
    (r 5), both half tori of radius 100."""
    geo = oracle.Geometry()
    clad = _pv(geo, (0.0, 10.0, 100.0, 0.0, np.pi))
    core = _pv(geo, (0.0, 5.0, 100.0, 0.0, np.pi))
    vy = (0.0, 1.0, 0.0)
    y = 0.0
    steps = [(clad, "DistanceToIn", 90.0, (3, 2)), (core, "DistanceToIn", 5.0, (2, 1)),
             (core, "DistanceToOut", 10.0, (2, 1)), (clad, "DistanceToOut", 5.0, (3, 2))]
    for pv, m, want, (s_core, s_clad) in steps:
        dist = _q(geo, pv, m, (0.0, y, 0.0), vy)[0]
        assert dist == pytest.approx(want, abs=1e-9)
        y += dist
        assert _q(geo, core, "Inside", (0.0, y, 0.0))[0] == s_core
        assert _q(geo, clad, "Inside", (0.0, y, 0.0))[0] == s_clad


@pytest.mark.parametrize("z", [i / 10 for i in range(12)])
def test_ring_from_far_outside(z):
    """TestTorus2.cpp:247-253: a thin ring (r 100, R 1000) entered along +x."""
    geo = oracle.Geometry()
    pv = _pv(geo, (0.0, 100.0, 1000.0, 0.0, 2 * np.pi))
    p, d = (-1200.0, 0.0, z), (1.0, 0.0, 0.0)
    assert _q(geo, pv, "DistanceToIn", p, d)[0] == pytest.approx(200.0 - math.sqrt(1e4 - z * z), abs=1e-9)
    assert _q(geo, pv, "DistanceToOut", p, d)[0] == -1.0


def test_normal_is_the_gradient_on_the_tube():
    """Away from the phi planes the outer-surface normal points from the
    tube's centre circle to the point."""
    geo = oracle.Geometry()
    rng = np.random.default_rng(13)
    phi = rng.uniform(0.2, 1.3, 300)
    alpha = rng.uniform(0, 2 * np.pi, 300)
    rho = R_TOR + R_MAX * np.cos(alpha)
    p = np.stack([rho * np.cos(phi), rho * np.sin(phi), R_MAX * np.sin(alpha)])
    nx, ny, nz, valid = geo.normal(_pv(geo, (R_MIN, R_MAX, R_TOR, 0.0, np.pi / 2)), p)
    assert valid.all()
    want = np.stack([np.cos(alpha) * np.cos(phi), np.cos(alpha) * np.sin(phi), np.sin(alpha)])
    np.testing.assert_allclose(np.stack([nx, ny, nz]), want, atol=1e-12)


def test_contains_ignores_the_phi_cut():
    """The reference's Contains skips the wedge (GenericKernelForContainsAndInside
    <true, false>, TorusImplementation2.h:520-529) while Inside applies it."""
    geo = oracle.Geometry()
    pv = _pv(geo, (R_MIN, R_MAX, R_TOR, 0.0, np.pi / 2))
    p = (-R_TOR - 50, 0.0, 0.0)  # in the tube, at phi = pi
    assert _q(geo, pv, "Inside", p)[0] == 3
    assert _q(geo, pv, "Contains", p)[0] == 1


def test_safety_bounds_distance():
    geo = oracle.Geometry()
    rng = np.random.default_rng(14)
    for params in [(R_MIN, R_MAX, R_TOR, 0.0, 2 * np.pi), (R_MIN, R_MAX, R_TOR, 0.3, 1.7),
                   (0.0, 30.0, 100.0, -0.5, 4.0)]:
        pv = _pv(geo, params)
        p, d = _random_rays(rng, 3000, 220.0)
        inside = _q(geo, pv, "Inside", p)
        sin_, sout = _q(geo, pv, "SafetyToIn", p), _q(geo, pv, "SafetyToOut", p)
        din, dout = _q(geo, pv, "DistanceToIn", p, d), _q(geo, pv, "DistanceToOut", p, d)
        o, i = inside == 3, inside == 1
        assert np.all(sin_[o] <= din[o] + 1e-9)
        assert np.all(sout[i] <= dout[i] + 1e-9)
        assert np.all(dout[i] >= 0)
