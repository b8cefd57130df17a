"""The oracle's boolean solids (oracle/boolean.cpp) pinned on the CPU.

By the reference's TestBooleans.cpp known answers (a union of two tube
holes, a box minus that union placed off-centre, the holes joined by a
turned slab; nested booleans, which the oracle builds and the library does
not), and by what the operations fix:
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


def _approx(a, b, tol=1e-6):
    """ApproxEqual<double> (test/unit_tests/ApproxEqual.h)."""
    if a == b:
        return True
    if a * b == 0.0:
        return abs(a - b) < tol
    return abs(a - b) / (abs(a) + abs(b)) < tol


def _shift(t):
    """A translation with the identity rotation (12 doubles)."""
    return np.array([t[0], t[1], t[2], 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0])


def test_reference_boolean_known_answers():
    """TestBooleans.cpp:14-190: L = 10; a tube of radius 0.9 L / 4 and
    half-length L placed at (-L/4, -L/4, 0) and (L/4, L/4, 0) as a union
    ("holes"); a box of half-size L minus the holes placed at (-L/2, -L/2, 0);
    the holes joined by a slab (L/4 sqrt2, L/10, L/10) turned by 45 degrees
    about z (its long axis along (1, 1, 0))."""
    L = 10.0
    geo = oracle.Geometry()
    tube = geo.tube(0.0, 0.9 * L / 4.0, L, 0.0, 2 * np.pi)
    rmax = 0.9 * L / 4.0
    holes = geo.boolean("union", tube, tube, _shift((-L / 4, -L / 4, 0.0)), _shift((L / 4, L / 4, 0.0)))
    box = geo.box(L, L, L)
    combined = geo.boolean("subtraction", box, holes, None, _shift((-L / 2, -L / 2, 0.0)))
    c, s = math.cos(math.pi / 4), math.sin(math.pi / 4)
    slab_tr = np.array([0.0, 0.0, 0.0, c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0])
    slabunion = geo.boolean("union", holes, geo.box(L / 4 * math.sqrt(2.0), L / 10, L / 10), None, slab_tr)
    ph, pc, ps = _pv(geo, holes), _pv(geo, combined), _pv(geo, slabunion)
    lower = geo.place(geo.logical_volume(tube), _shift((L / 4, L / 4, 0.0)))

    def q(pv, m, p, d=None):
        dd = None if d is None else [np.array([v]) for v in np.asarray(d, float) / np.linalg.norm(d)]
        return geo.query(pv, m, [np.array([v]) for v in p], dd)[0]

    quarter = math.sqrt(2.0) * L / 2.0
    assert q(ph, "Contains", (0, 0, 0)) == 0
    assert q(ph, "Contains", (-L / 4, -L / 4, 0)) == 1 and q(ph, "Contains", (L / 4, L / 4, 0)) == 1
    for d in ((0, 0, 1), (1, 0, 0), (0, 1, 0)):
        assert q(ph, "DistanceToIn", (0, 0, 0), d) >= 1e300
    d1, d2 = q(ph, "DistanceToIn", (0, 0, 0), (1, 1, 0)), q(ph, "DistanceToIn", (0, 0, 0), (-1, -1, 0))
    assert _approx(q(lower, "DistanceToIn", (0, 0, 0), (1, 1, 0)), d1) and _approx(d1, d2)
    assert _approx(d1, quarter / 2.0 - rmax)
    assert _approx(q(ph, "PlacedDistanceToOut", (L / 4, L / 4, 0), (1, 1, 0)), rmax)
    assert _approx(q(ph, "PlacedDistanceToOut", (-L / 4, -L / 4, 0), (1, 1, 0)), rmax)
    for p, want in (((0, 0, 0), 1), ((-L / 2, -L / 2, 0), 1), ((-3 * L / 4, -3 * L / 4, 0), 0),
                    ((-L / 4, -L / 4, 0), 0), ((L / 4, L / 4, 0), 1), ((3 * L / 4, 3 * L / 4, 0), 1)):
        assert q(pc, "Contains", p) == want, p
    for d in ((1, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1)):
        assert _approx(q(pc, "DistanceToOut", (0, 0, 0), d), 10.0)
    assert _approx(q(pc, "DistanceToOut", (-L / 4, 0, 0), (0, -1, 0)), (L / 2 - 2 * rmax) / 2)
    assert _approx(q(pc, "DistanceToOut", (-L / 4, 0, 0), (0, 1, 0)), 10.0)
    assert _approx(q(pc, "DistanceToOut", (-3 * L / 4, 0, 0), (0, -1, 0)), L / 2 + (L / 2 - 2 * rmax) / 2)
    assert _approx(q(pc, "DistanceToOut", (-L / 2, -L / 2, 0), (1, 1, 0)), quarter / 2.0 - rmax)
    d = q(pc, "DistanceToOut", (L, L, 0), (-1, -1, 0))
    assert math.sqrt(2) * L < d < math.sqrt(2) * L * 1.5
    assert _approx(q(pc, "DistanceToIn", (-20, 0, 0), (1, 0, 0)), 10.0)
    assert _approx(q(pc, "DistanceToIn", (0, -20, 0), (0, 1, 0)), 10.0)
    assert _approx(q(pc, "DistanceToIn", (20, 20, 20), (-1, -1, -1)), 10.0 * math.sqrt(3.0))
    assert _approx(q(pc, "DistanceToIn", (-L / 4, -L / 4, 0), (-1, 0, 0)), rmax)
    assert _approx(q(pc, "DistanceToIn", (-L / 4 - rmax, -L / 4, 0), (1, 0, 0)), 2 * rmax)
    assert _approx(q(pc, "DistanceToIn", (-L / 4, -L / 4, 0), (0, -1, 0)), rmax)
    assert q(pc, "DistanceToIn", (-L / 4, -L / 4, 0), (0, 0, -1)) >= 1e300
    r2 = math.sqrt(2.0)
    assert _approx(q(ps, "PlacedDistanceToOut", (0, 0, 0), (1, 1, 0)), L / 4 * r2 + rmax)
    assert _approx(q(ps, "PlacedDistanceToOut", (-L / 4, -L / 4, 0), (1, 1, 0)), L / 2 * r2 + rmax)
    assert _approx(q(ps, "PlacedDistanceToOut", (L / 4, L / 4, 0), (1, 1, 0)), rmax)
    assert _approx(q(ps, "PlacedDistanceToOut", (-L / 4, -L / 4, 0), (-1, -1, 0)), rmax)
    assert _approx(q(ps, "PlacedDistanceToOut", (L / 4, L / 4, 0), (-1, -1, 0)), L / 2 * r2 + rmax)
