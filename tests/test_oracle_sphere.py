"""Oracle sphere pinned beyond the reference's known answers: the full
sphere against the orb (the same solid by another algorithm), boundary hits
on the surface, and safeties below the distances (ShapeTester conventions)."""
import math

import numpy as np
import pytest

import oracle
from vecgeom_amd import sampling

PI = math.pi
CUT_SPHERES = [
    (0.0, 50.0, 0.3, 1.2 * PI, 0.0, PI),           # phi sector > pi... (1.2 pi)
    (10.0, 50.0, -0.5, 0.8 * PI, 0.0, PI),         # hollow, phi < pi
    (0.0, 40.0, 0.0, 2 * PI, 0.3, 0.9),            # theta band above the equator
    (5.0, 40.0, 0.0, 2 * PI, 1.2, 1.5),            # theta band across the equator
    (20.0, 40.0, 0.4, 1.1, 1.9, 0.8),              # patch below the equator
    (0.0, 30.0, 0.0, 2 * PI, 0.5 * PI, 0.5 * PI),  # lower hemisphere
]


def _pv(geo, kind, *args):
    return geo.place(geo.logical_volume(getattr(geo, kind)(*args)))


def _inputs(n, half, seed):
    pts = sampling.points_in_box(seed, 1, n, (half, half, half), 1.3)
    dirs = sampling.isotropic_directions(seed, 2, n)
    return pts, dirs


def test_full_sphere_is_the_orb():
    geo = oracle.Geometry()
    sph, orb = _pv(geo, "sphere", 0.0, 25.0), _pv(geo, "orb", 25.0)
    pts, dirs = _inputs(20000, 25.0, 3)
    ins_s, ins_o = geo.query(sph, "Inside", pts), geo.query(orb, "Inside", pts)
    assert (ins_s == ins_o).all()
    for m in ("DistanceToIn", "DistanceToOut"):
        a, b = geo.query(sph, m, pts, dirs), geo.query(orb, m, pts, dirs)
        fin = (a < 1e30) & (b < 1e30)
        assert ((a >= 1e30) == (b >= 1e30)).all()
        np.testing.assert_allclose(a[fin], b[fin], rtol=0, atol=1e-9)
    for m in ("SafetyToIn", "SafetyToOut"):
        np.testing.assert_allclose(geo.query(sph, m, pts), geo.query(orb, m, pts), rtol=0, atol=1e-9)


@pytest.mark.parametrize("params", CUT_SPHERES)
def test_hits_on_surface_and_safety_below_distance(params):
    geo = oracle.Geometry()
    pv = _pv(geo, "sphere", *params)
    pts, dirs = _inputs(20000, params[1], 5)
    inside = geo.query(pv, "Inside", pts)
    din = geo.query(pv, "DistanceToIn", pts, dirs)
    dout = geo.query(pv, "DistanceToOut", pts, dirs)
    out, ins = inside == 3, inside == 1
    hit = out & (din < 1e30)
    assert hit.sum() > 100
    hp = [pts[k][hit] + din[hit] * dirs[k][hit] for k in range(3)]
    # the hit point is on the surface: both safeties there vanish to rounding
    s_in, s_out = geo.query(pv, "SafetyToIn", hp), geo.query(pv, "SafetyToOut", hp)
    assert (np.minimum(np.abs(s_in), np.abs(s_out)) < 1e-7).mean() > 0.999
    # The reference's theta cone does not reject a negative second root when
    # the end angle lies below the equator (ThetaCone_Evolution.h, the
    # fETheta > pi/2 branch of DistanceToOut), so a sphere whose theta range
    # lies below the equator can return a negative DistanceToOut; the oracle
    # keeps that, and those exits are left out here
    if params[4] <= 0.5 * PI:
        assert (dout[ins] >= 0).all()
    ins &= dout >= 0
    ep = [pts[k][ins] + dout[ins] * dirs[k][ins] for k in range(3)]
    s_in, s_out = geo.query(pv, "SafetyToIn", ep), geo.query(pv, "SafetyToOut", ep)
    assert (np.minimum(np.abs(s_in), np.abs(s_out)) < 1e-7).mean() > 0.999
    # safeties are lower bounds of the distances
    assert (geo.query(pv, "SafetyToIn", pts)[out] <= din[out] + 1e-9).all()
    assert (geo.query(pv, "SafetyToOut", pts)[ins] <= dout[ins] + 1e-9).all()



@pytest.mark.parametrize("params", CUT_SPHERES)
def test_inside_matches_the_definition(params):
    """Away from the surface, Inside is the set definition in spherical coordinates."""
    rmin, rmax, sphi, dphi, sth, dth = params
    geo = oracle.Geometry()
    pv = _pv(geo, "sphere", *params)
    pts, _ = _inputs(20000, rmax, 7)
    x, y, z = pts
    r = np.sqrt(x * x + y * y + z * z)
    th = np.arccos(np.clip(z / np.maximum(r, 1e-300), -1, 1))
    ph = np.mod(np.arctan2(y, x) - sphi, 2 * PI)
    eth = min(sth + dth, PI)
    want = (r > rmin) & (r < rmax) & (th > sth) & (th < eth) & ((ph < dphi) if dphi < 2 * PI else True)
    # margin from every boundary
    far = ((np.abs(r - rmin) > 1e-6) & (np.abs(r - rmax) > 1e-6) & (np.abs(th - sth) > 1e-6) &
           (np.abs(th - eth) > 1e-6) & (r > 1e-3))
    if dphi < 2 * PI:
        far &= (np.abs(ph) > 1e-6) & (np.abs(ph - dphi) > 1e-6) & (np.abs(ph - 2 * PI) > 1e-6)
    got = geo.query(pv, "Inside", pts)
    assert (got[far] == np.where(want[far], 1, 3)).all()
