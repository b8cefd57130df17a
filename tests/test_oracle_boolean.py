"""The oracle's boolean solids (oracle/boolean.cpp) pinned on the CPU.

The reference's TestBooleans.cpp checks only that its boolean solids build
and place; so the oracle is pinned by what the operations fix:
- a subtraction of a disjoint solid classifies like the left solid, and its
  DistanceToIn is the left solid's plus kPushTolerance (the push the
  subtraction adds to every entry);
- an intersection with a box that contains the solid gives, from points
  inside that box, the solid's own DistanceToIn (at least kTolerance);
- closed forms: a ray through the bore of a box minus a tube misses, a ray
  onto the box face enters at the face plus the push, and a ray along the
  axis of two overlapping boxes enters their intersection where the second
  box begins.
"""
import math

import numpy as np

import oracle
from vecgeom_amd import sampling
from tests.shapes_spec import transform12

PUSH, TOL = 1e-6, 1e-9


def _pv(geo, s):
    return geo.place(geo.logical_volume(s))


def _inputs(mn, mx, n, seed):
    pts = sampling.points_in_box(seed, 1, n, 0.5 * (mx - mn), 1.4, 0.5 * (mx + mn))
    return pts, sampling.isotropic_directions(seed, 2, n)


def test_subtracting_a_disjoint_solid():
    geo = oracle.Geometry()
    left = geo.tube(2.0, 10.0, 15.0, 0.3, 4.0)
    far = transform12((200.0, 0.0, 0.0))
    b = _pv(geo, geo.boolean("subtraction", left, ("box", (5.0, 5.0, 5.0)), None, far))
    a = _pv(geo, left)
    pts, dirs = _inputs(*geo.extent(left), 20000, 3)
    for q in ("Contains", "Inside", "SafetyToIn"):
        assert np.array_equal(geo.query(b, q, pts, dirs), geo.query(a, q, pts, dirs)), q
    da, db = geo.query(a, "DistanceToIn", pts, dirs), geo.query(b, "DistanceToIn", pts, dirs)
    hit = (da >= 0) & (da < 1e30)
    assert hit.sum() > 1000 and np.array_equal(db[hit], da[hit] + PUSH)
    assert np.all(db[da >= 1e30] >= 1e30)


def test_intersecting_with_a_containing_box():
    geo = oracle.Geometry()
    inner = geo.cone(0.0, 10.0, 5.0, 25.0, 40.0, 0.0, 2 * math.pi)
    b = _pv(geo, geo.boolean("intersection", inner, ("box", (60.0, 60.0, 60.0))))
    a = _pv(geo, inner)
    pts, dirs = _inputs(*geo.extent(inner), 20000, 5)
    assert np.array_equal(geo.query(b, "Contains", pts, dirs), geo.query(a, "Contains", pts, dirs))
    da, db = geo.query(a, "DistanceToIn", pts, dirs), geo.query(b, "DistanceToIn", pts, dirs)
    hit = (da >= TOL) & (da < 1e30)
    assert hit.sum() > 1000 and np.array_equal(db[hit], da[hit])


def test_closed_forms():
    geo = oracle.Geometry()
    bore = _pv(geo, geo.boolean("subtraction", ("box", (10.0, 10.0, 10.0)), ("tube", (0.0, 5.0, 20.0, 0.0, 2 * math.pi))))
    one, zero = np.ones(1), np.zeros(1)
    # down the bore: no entry
    d = geo.query(bore, "DistanceToIn", (zero, zero, 30 + zero), (zero, zero, -one))
    assert d[0] >= 1e30
    # onto the -x face at x = -10
    d = geo.query(bore, "DistanceToIn", (-20 + zero, 7 + zero, zero), (one, zero, zero))
    assert d[0] == 10.0 + PUSH
    # two boxes overlapping in x in [-5, 10]
    inter = _pv(geo, geo.boolean("intersection", ("box", (10.0, 10.0, 10.0)), ("box", (10.0, 10.0, 10.0)), None,
                                 transform12((5.0, 0.0, 0.0), angle=0.0)))
    d = geo.query(inter, "DistanceToIn", (-30 + zero, zero, zero), (one, zero, zero))
    assert abs(d[0] - 25.0) < 1e-12
    d = geo.query(inter, "DistanceToOut", (zero, zero, zero), (one, zero, zero))
    assert abs(d[0] - 10.0) < 1e-12
    d = geo.query(inter, "DistanceToOut", (zero, zero, zero), (-one, zero, zero))
    assert abs(d[0] - 5.0) < 1e-12
