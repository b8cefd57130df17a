"""Oracle vs the reference's own known-answer tests (CPU)."""
import numpy as np
import pytest

import oracle
from tests.golden import kats

CASES = ([("box", c) for c in kats.BOX_KATS] + [("tube", c) for c in kats.TUBE_KATS]
         + [("cone", c) for c in kats.CONE_KATS] + [("cone", c) for c in kats.CONE_T1_KATS]
         + [("trd", c) for c in kats.TRD_KATS]
         + [("trapezoid", c) for c in kats.TRAP_KATS] + [("polycone", c) for c in kats.POLYCONE_KATS]
         + [("polyhedron", c) for c in kats.POLYHEDRON_KATS]
         + [("tessellated", c) for c in kats.TESS_KATS] + [("orb", c) for c in kats.ORB_KATS]
         + [("sphere", c) for c in kats.SPHERE_KATS] + [("torus", c) for c in kats.TORUS_KATS]
         + [("sphere", c) for c in kats.WEDGE_THETA_KATS])


def _run(geo, kind, params, method, p, d):
    shape = getattr(geo, kind)(*params)
    pv = geo.place(geo.logical_volume(shape))
    pts = [np.array([p[k]], dtype=np.float64) for k in range(3)]
    dirs = None if d is None else [np.array([d[k]], dtype=np.float64) for k in range(3)]
    return geo.query(pv, method, pts, dirs)[0]


@pytest.mark.parametrize("kind,case", CASES, ids=[f"{k}-{i}-{c[1]}" for i, (k, c) in enumerate(CASES)])
def test_kat(kind, case):
    params, method, p, d, expect = case
    geo = oracle.Geometry()
    r = _run(geo, kind, params, method, p, d)
    assert kats.check(r, expect), (params, method, p, d, r, expect)


def test_tube_bug810_chain():
    """TestTube.cpp:432-472: walk along +y through a clad/core pair of half tubes."""
    geo = oracle.Geometry()
    clad = geo.place(geo.logical_volume(geo.tube(90.0, 110.0, 105.0, 0.0, np.pi)))
    core = geo.place(geo.logical_volume(geo.tube(95.0, 105.0, 100.0, 0.0, np.pi)))
    vy = [np.array([0.0]), np.array([1.0]), np.array([0.0])]

    def q(pv, m, p, d=None):
        return geo.query(pv, m, [np.array([c]) for c in p], d)[0]

    p = np.zeros(3)
    p = p + q(clad, "DistanceToIn", p, vy) * np.array([0, 1.0, 0])
    assert q(core, "Inside", p) == 3 and q(clad, "Inside", p) == 2
    p = p + q(core, "DistanceToIn", p, vy) * np.array([0, 1.0, 0])
    assert q(core, "Inside", p) == 2 and q(clad, "Inside", p) == 1
    p = p + q(core, "DistanceToOut", p, vy) * np.array([0, 1.0, 0])
    assert q(core, "Inside", p) == 2 and q(clad, "Inside", p) == 1
    p = p + q(clad, "DistanceToOut", p, vy) * np.array([0, 1.0, 0])
    assert q(core, "Inside", p) == 3 and q(clad, "Inside", p) == 2


@pytest.mark.parametrize("case", kats.CONE_SOFT_KATS, ids=[f"cone-soft-{i}" for i in range(len(kats.CONE_SOFT_KATS))])
def test_cone_soft_kat(case):
    params, method, p, d, wanted = case
    geo = oracle.Geometry()
    r = float(_run(geo, "cone", params, method, p, d))
    if wanted >= kats.KINF:
        assert r >= kats.KINF, r
    else:
        assert abs(r - wanted) <= 0.01, (r, wanted)


def test_polyhedron_either_kat():
    """TestPolyhedra.cpp:187: DistanceToIn(p2, +y) is ~0 or beyond 1e10."""
    geo = oracle.Geometry()
    for params, method, p, d in kats.POLYHEDRON_EITHER:
        r = float(_run(geo, "polyhedron", params, method, p, d))
        assert abs(r) < 1e-9 or abs(r) > 1e10, r


def test_polyhedron_extent_kat():
    """TestPolyhedra.cpp:170-173: Extent equals the expected corner box."""
    geo = oracle.Geometry()
    for params, lo, hi in kats.POLYHEDRON_EXTENT:
        mn, mx = geo.extent(geo.polyhedron(*params))
        for a, b in zip(list(mn) + list(mx), list(lo) + list(hi)):
            assert kats.approx_equal(float(a), float(b)), (mn, mx, lo, hi)


@pytest.mark.parametrize("brute", [False, True], ids=["tree", "brute"])
def test_tessellated_extent_kat(brute):
    """TestTessellated.cpp:232-244."""
    geo = oracle.Geometry()
    for params, lo, hi in kats.TESS_EXTENT:
        mn, mx = geo.extent(geo.tessellated(*params, brute_force=brute))
        assert np.allclose(mn, lo) and np.allclose(mx, hi)


def test_trapezoid_extent_kat():
    """TestTrap.cpp:640-655: trap1's extent is the box (20, 30, 40); trap2's
    is (30, 40, 40), its wider +z face."""
    geo = oracle.Geometry()
    for params, lo, hi in ((kats.TRAP1, (-20, -30, -40), (20, 30, 40)), (kats.TRAP2, (-30, -40, -40), (30, 40, 40))):
        mn, mx = geo.extent(geo.trapezoid(*params))
        assert np.allclose(mn, lo, rtol=0, atol=1e-6) and np.allclose(mx, hi, rtol=0, atol=1e-6), (params, mn, mx)
