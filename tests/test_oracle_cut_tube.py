"""The oracle's cut tube (oracle/cuttube.cpp) pinned on the CPU.

The reference has no cut tube unit test with known answers, so the oracle is
pinned by what the geometry fixes:
- flat cuts reduce it to the tube of the same half-length: the safeties agree
  bit for bit and the classifications and distances agree away from the
  surface;
- a ray along z from inside leaves through the top cut plane at the closed
  form distance (n . ((0,0,dz) - p)) / n_z, and a ray entering from below
  through the bottom plane likewise;
- hit points lie on the surface, safeties never exceed distances, and every
  contained point lies inside the extent, whose faces the contained points
  reach.
"""
import math

import numpy as np
import pytest

import oracle
from vecgeom_amd import sampling
from tests.shapes_spec import SHAPES, cut_normals

CUTS = sorted(k for k in SHAPES if k.startswith("cut_tube"))


def _placed(geo, kind, args):
    s = getattr(geo, kind)(*args)
    return s, geo.place(geo.logical_volume(s))


def _inputs(geo, s, n, seed, scale=1.3):
    mn, mx = geo.extent(s)
    pts = sampling.points_in_box(seed, 1, n, 0.5 * (mx - mn), scale, 0.5 * (mx + mn))
    dirs = sampling.isotropic_directions(seed, 2, n)
    return pts, dirs


@pytest.mark.parametrize("rmin, sphi, dphi", [(0.0, 0.0, 2 * math.pi), (3.0, 0.4, 2.0), (2.0, -1.0, 4.5)])
def test_flat_cuts_are_the_tube(rmin, sphi, dphi):
    geo = oracle.Geometry()
    _, ct = _placed(geo, "cut_tube", (rmin, 5.0, 10.0, sphi, dphi, (0.0, 0.0, -1.0), (0.0, 0.0, 1.0)))
    tb_s, tb = _placed(geo, "tube", (rmin, 5.0, 10.0, sphi, dphi))
    pts, dirs = _inputs(geo, tb_s, 20000, 3)
    for q in ("SafetyToIn", "SafetyToOut"):
        assert np.array_equal(geo.query(ct, q, pts, dirs), geo.query(tb, q, pts, dirs)), q
    # away from the surface band the classification and the distances agree
    saf = np.abs(geo.query(tb, "SafetyToIn", pts, dirs))
    saf = np.minimum(saf, np.abs(geo.query(tb, "SafetyToOut", pts, dirs)))
    far = saf > 1e-6
    for q in ("Contains", "Inside"):
        a, b = geo.query(ct, q, pts, dirs), geo.query(tb, q, pts, dirs)
        assert np.array_equal(a[far], b[far]), q
    for q in ("DistanceToIn", "DistanceToOut"):
        a, b = geo.query(ct, q, pts, dirs), geo.query(tb, q, pts, dirs)
        ok = far & (b > 0) & (b < 1e30)
        assert ok.sum() > 1000
        np.testing.assert_allclose(a[ok], b[ok], rtol=1e-9, atol=1e-9, err_msg=q)
        assert np.array_equal(a[far & (b >= 1e30)] >= 1e30, np.ones(int((far & (b >= 1e30)).sum()), bool)), q


@pytest.mark.parametrize("tilt", [0.3, 0.7])
def test_axial_ray_meets_the_cut_plane(tilt):
    bot, top = cut_normals(math.pi - tilt, 1.1, tilt, 2.5)
    dz = 10.0
    geo = oracle.Geometry()
    _, ct = _placed(geo, "cut_tube", (0.0, 5.0, dz, 0.0, 2 * math.pi, bot, top))
    rng = np.random.default_rng(7)
    r = 3.0 * np.sqrt(rng.random(500))
    ph = rng.random(500) * 2 * math.pi
    x, y, z = r * np.cos(ph), r * np.sin(ph), rng.uniform(-2.0, 2.0, 500)
    one, zero = np.ones(500), np.zeros(500)
    nt, nb = np.array(top), np.array(bot)
    # out through the top plane along +z
    got = geo.query(ct, "DistanceToOut", (x, y, z), (zero, zero, one))
    want = (nt[0] * (0 - x) + nt[1] * (0 - y) + nt[2] * (dz - z)) / nt[2]
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    # in through the bottom plane along +z from below
    zb = -dz - 40.0
    got = geo.query(ct, "DistanceToIn", (x, y, zb + zero), (zero, zero, one))
    want = (nb[0] * (0 - x) + nb[1] * (0 - y) + nb[2] * (-dz - zb)) / nb[2]
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", CUTS)
def test_hits_on_surface_and_safety_below_distance(name):
    kind, args = SHAPES[name]
    geo = oracle.Geometry()
    s, pv = _placed(geo, kind, args)
    pts, dirs = _inputs(geo, s, 20000, 11)
    inside = geo.query(pv, "Inside", pts, dirs)
    din = geo.query(pv, "DistanceToIn", pts, dirs)
    dout = geo.query(pv, "DistanceToOut", pts, dirs)
    sin_ = geo.query(pv, "SafetyToIn", pts, dirs)
    sout = geo.query(pv, "SafetyToOut", pts, dirs)
    P, D = np.array(pts), np.array(dirs)
    outs, ins = inside == 3, inside == 1
    assert outs.sum() > 1000 and ins.sum() > 1000
    assert np.all(sin_[outs] <= din[outs] + 1e-9)
    assert np.all(sout[ins] <= dout[ins] + 1e-9)
    for sel, dist in ((outs & (din < 1e30), din), (ins, dout)):
        q = P[:, sel] + dist[sel] * D[:, sel]
        on = geo.query(pv, "Inside", tuple(q), tuple(D[:, sel]))
        assert np.mean(on == 2) > 0.999, name
    # misses: the ray never meets the solid (sampled along it)
    miss = outs & (din >= 1e30)
    for t in np.linspace(0.0, 200.0, 81):
        q = P[:, miss] + t * D[:, miss]
        assert not np.any(geo.query(pv, "Inside", tuple(q), tuple(D[:, miss])) == 1)


@pytest.mark.parametrize("name", CUTS)
def test_extent_bounds_the_solid(name):
    kind, args = SHAPES[name]
    geo = oracle.Geometry()
    s, pv = _placed(geo, kind, args)
    mn, mx = geo.extent(s)
    rmax, dz = args[1], args[2]
    big = max(rmax, dz) * 4
    pts = sampling.points_in_box(5, 1, 400000, np.array([big, big, big]), 1.0, np.zeros(3))
    c = geo.query(pv, "Contains", pts, sampling.isotropic_directions(5, 2, 400000)).astype(bool)
    P = np.array(pts)[:, c]
    assert np.all(P.min(axis=1) >= mn - 1e-12) and np.all(P.max(axis=1) <= mx + 1e-12)
    # tight: the contained points come within a few percent of each face
    span = mx - mn
    assert np.all(P.min(axis=1) - mn < 0.08 * span) and np.all(mx - P.max(axis=1) < 0.08 * span)


def test_illegal_normals_are_rejected():
    geo = oracle.Geometry()
    with pytest.raises(Exception):
        geo.cut_tube(0.0, 5.0, 10.0, 0.0, 2 * math.pi, (0.0, 0.0, 1.0), (0.0, 0.0, 1.0))
