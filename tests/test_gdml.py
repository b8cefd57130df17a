"""GDML reader (vecgeom_amd/gdml.py) on the reference's GDML test geometries,
on the CPU: the constructor arguments it derives (units, halves, z planes,
facets, placements), and the oracle built from the files agreeing with the
oracle built by hand.  The GPU side is tests/test_gpu_gdml.py."""
import math
import os

import numpy as np
import pytest

import oracle
from vecgeom_amd import gdml, sampling

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gdml")


def path(name):
    return os.path.join(HERE, name + ".gdml")


class Recorder:
    """Geometry stand-in that records the builder calls."""

    def __init__(self):
        self.calls = []

    def _rec(self, *call):
        self.calls.append(call)
        return len(self.calls) - 1

    def __getattr__(self, name):
        if name in ("box", "orb", "sphere", "torus", "tube", "cut_tube", "boolean", "cone", "trd", "trapezoid", "polycone",
                    "polyhedron", "tessellated",
                    "logical_volume", "place", "add_daughter"):
            return lambda *a: self._rec(name, *a)
        raise AttributeError(name)

    def of(self, kind):
        return [c for c in self.calls if c[0] == kind]


def read(name, mm_unit=0.1):
    r = Recorder()
    world, solids, volumes = gdml.load(path(name), r, mm_unit=mm_unit)
    return r, world, solids, volumes


def test_units_and_halves():
    # lunit="cm" with the reference's millimetre of 0.1: 1 cm = 1.0
    r, *_ = read("oneBox")
    assert r.of("box") == [("box", 50.0, 50.0, 50.0)]
    r, *_ = read("oneBox", mm_unit=1.0)
    assert r.of("box") == [("box", 500.0, 500.0, 500.0)]
    r, *_ = read("oneTube")  # deltaphi 360 deg
    (_, rmin, rmax, dz, sphi, dphi), = r.of("tube")
    assert (rmin, rmax, dz, sphi) == (0.0, 100.0, 50.0, 0.0) and dphi == pytest.approx(2 * math.pi, rel=1e-15)
    r, *_ = read("oneConeSegment")
    assert r.of("cone") == [("cone", 15.0, 30.0, 70.0, 100.0, 25.0, 1.0, 4.0)]
    r, *_ = read("oneTrapezoid")
    assert r.of("trd") == [("trd", 45.0, 40.0, 30.0, 25.0, 50.0)]
    r, *_ = read("oneGeneralTrapezoid")
    assert r.of("trapezoid") == [("trapezoid", 50.0, 1.0, 2.0, 75.0, 50.0, 5.0, 1.0, 75.0, 5.0, 5.0, 1.0)]


def test_z_planes():
    r, *_ = read("onePolycone")  # lunit mm: 0.1 per mm
    (_, sphi, dphi, z, rmin, rmax), = r.of("polycone")
    assert (sphi, dphi) == (1.0, 4.0)
    assert np.allclose(z, [5.0, 8.0], rtol=0, atol=1e-15) and np.allclose(rmin, [1.5, 7.0], rtol=0, atol=1e-15)
    assert np.allclose(rmax, [3.0, 10.0], rtol=0, atol=1e-15)
    r, *_ = read("onePolyhedron")
    (_, sphi, dphi, nsides, z, rmin, rmax), = r.of("polyhedron")
    assert (sphi, dphi, nsides) == (1.0, 4.0, 11) and len(z) == 2


def test_tessellated_facets():
    r, *_ = read("oneTesselatedSolid")
    (_, tris), = r.of("tessellated")
    # six triangles and one quadrangle split into two
    assert len(tris) == 8
    v1, v2, v3, v4 = (10.0, 10.0, 0.0), (-10.0, 10.0, 0.0), (-10.0, -10.0, 0.0), (10.0, -10.0, 0.0)
    assert [tuple(t) for t in tris[6:]] == [(v4, v3, v2), (v4, v2, v1)]


def test_relative_vertices_accumulate():
    text = open(path("oneTesselatedSolid")).read().replace('type="ABSOLUTE"', 'type="RELATIVE"')
    r = Recorder()
    gdml.load(text, r)
    (_, tris), = r.of("tessellated")
    v1, v2, v6 = np.array([10.0, 10.0, 0.0]), np.array([-10.0, 10.0, 0.0]), np.array([-3.0, 7.0, 5.0])
    # vt0, vt0 + vt1, vt0 + vt1 + vt2
    assert np.array_equal(np.array(tris[0]), np.array([v1, v1 + v2, v1 + v2 + v6]))


def test_placements():
    r, world, _, volumes = read("aPosition")
    (_, lv, tr), = [c for c in r.of("place") if len(c) > 2]
    assert np.array_equal(tr, [7.0, 17.0, 37.0, 1, 0, 0, 0, 1, 0, 0, 0, 1])
    assert r.calls[world] == ("place", volumes["World"])
    # a rotation of the frame about z by 37 deg alone: the transposed matrix
    text = open(path("aRotation")).read().replace('x="7" y="17" z="37"', 'x="0" y="0" z="37"')
    r = Recorder()
    gdml.load(text, r)
    (_, lv, tr), = [c for c in r.of("place") if len(c) > 2]
    c, s = math.cos(math.radians(37.0)), math.sin(math.radians(37.0))
    assert np.allclose(tr, [0, 0, 0, c, s, 0, -s, c, 0, 0, 0, 1], rtol=0, atol=1e-15)
    # three angles: a proper rotation
    r, *_ = read("aRotation")
    (_, lv, tr), = [c for c in r.of("place") if len(c) > 2]
    R = np.array(tr[3:]).reshape(3, 3)
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-15) and np.linalg.det(R) == pytest.approx(1.0, abs=1e-14)


def test_numbers_read_like_stod():
    text = open(path("oneBox")).read()
    r = Recorder()
    gdml.load(text.replace('x="100"', 'x="2*pi"'), r)  # the numeric prefix: 2
    assert r.of("box")[0][1] == 1.0
    r = Recorder()
    gdml.load(text.replace('x="100"', 'x="half"').replace("<define>", '<define><constant name="half" value="30"/>'), r)
    assert r.of("box")[0][1] == 15.0
    r = Recorder()
    gdml.load(text.replace('x="100"', 'x="unknown"'), r)  # an unknown name reads 0
    assert r.of("box")[0][1] == 0.0


@pytest.mark.parametrize("name, what", [("scaledPhysvol", "scale")])
def test_unsupported_is_an_error(name, what):
    with pytest.raises(gdml.GdmlError, match=what):
        read(name)
    text = open(path("oneOrb")).read().replace("<orb ", "<ellipsoid ")
    with pytest.raises(gdml.GdmlError, match="ellipsoid"):
        gdml.load(text, Recorder())


def test_cut_tube_reads_normals_as_lengths():
    r, *_ = read("oneCutTube")  # lunit cm: 1 cm = 1.0
    (_, rmin, rmax, dz, sphi, dphi, low, high), = r.of("cut_tube")
    assert (rmin, rmax, dz, sphi, dphi) == (1.0, 5.0, 10.0, 1.0, 4.0)
    assert low == (-15.0, -15.0, -15.0) and high == (10.0, 10.0, 10.0)


def test_boolean_operands_and_placement():
    r, world, solids, _ = read("solidSubtraction")
    (_, op, first, second, ltr, rtr), = r.of("boolean")
    assert (op, first, second, ltr) == ("subtraction", solids["minuend"], solids["subtrahend"], None)
    assert rtr[:3] == [40.0, 35.0, 30.0]  # positionref in cm
    R = np.array(rtr[3:]).reshape(3, 3)
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-15) and not np.allclose(R, np.eye(3))
    assert r.of("box")[0] == ("box", 50.0, 50.0, 50.0)


def test_union_operands_and_placement():
    r, world, solids, _ = read("solidUnion")
    (_, op, first, second, ltr, rtr), = r.of("boolean")
    assert (op, first, second, ltr) == ("union", solids["firstHalf"], solids["secondHalf"], None)
    assert [b[1:] for b in r.of("box")] == [(50.0, 50.0, 50.0), (35.0, 35.0, 35.0)]


def test_orb_radius_and_constant():
    # the radius is not halved; aConstant names it through a <constant>
    for name in ("oneOrb", "aConstant"):
        r, *_ = read(name)
        assert r.of("orb") == [("orb", 100.0)]


@pytest.mark.parametrize("name", ["compositeMaterial", "mixtureElement", "mixtureMaterial", "simpleMaterial"])
def test_materials_are_read_and_ignored(name):
    """Elements, isotopes, composites and mixtures carry no geometry: each
    file is one orb of radius 100 cm, as oneOrb.gdml."""
    r, world, solids, volumes = read(name)
    assert r.of("orb") == read("oneOrb")[0].of("orb") == [("orb", 100.0)]
    assert list(volumes) == ["World"] and not r.of("add_daughter")


def test_union_world_with_a_rotated_detector():
    """unionWIthDetector.gdml: a union of two boxes as the world's solid,
    holding the second box placed with the union's own position and
    rotation."""
    r, world, solids, volumes = read("unionWIthDetector")
    assert set(solids) == {"firstHalf", "secondHalf", "world"}
    assert [c[1:] for c in r.of("box")] == [(50.0, 50.0, 50.0), (35.0, 35.0, 35.0)]
    (_, mother, pv), = r.of("add_daughter")
    assert mother == volumes["World"]
    (_, op, left, right, ltr, rtr), = r.of("boolean")
    assert (op, left, right, ltr) == ("union", solids["firstHalf"], solids["secondHalf"], None)
    # the detector sits exactly where the union puts its second box
    placed = [c for c in r.of("place") if c[1] == volumes["Detector"]]
    assert len(placed) == 1 and placed[0][2] == rtr
    assert not np.allclose(np.asarray(rtr)[3:12], np.eye(3).ravel())  # a real rotation


def test_two_boxes_structure():
    r, world, solids, volumes = read("twoBoxes")
    assert set(solids) == {"detector", "world"} and set(volumes) == {"Detector", "World"}
    (_, mother, pv), = r.of("add_daughter")
    assert mother == volumes["World"]


@pytest.mark.parametrize("name, build", [
    ("oneBox", lambda g: g.box(50.0, 50.0, 50.0)),
    ("oneOrb", lambda g: g.orb(100.0)),
    ("oneConeSegment", lambda g: g.cone(15.0, 30.0, 70.0, 100.0, 25.0, 1.0, 4.0)),
    ("oneTrapezoid", lambda g: g.trd(45.0, 40.0, 30.0, 25.0, 50.0)),
    ("onePolyhedron", None),
])
def test_oracle_from_gdml_equals_oracle_by_hand(name, build):
    """The oracle built from the file answers like the oracle built from the
    same constructor arguments."""
    geo = oracle.Geometry()
    world, solids, _ = gdml.load(path(name), geo)
    shape = solids["world"]
    geo2 = oracle.Geometry()
    if build is None:
        r, *_ = read(name)
        (_, *args), = r.of("polyhedron")
        s2 = geo2.polyhedron(*args)
    else:
        s2 = build(geo2)
    pv2 = geo2.place(geo2.logical_volume(s2))
    mn, mx = geo.extent(shape)
    pts = sampling.points_in_box(3, 1, 2000, 0.5 * (mx - mn), 1.3, 0.5 * (mx + mn))
    dirs = sampling.isotropic_directions(3, 2, 2000)
    for q in ("Inside", "DistanceToIn", "DistanceToOut", "SafetyToIn"):
        assert np.array_equal(geo.query(world, q, pts, dirs), geo2.query(pv2, q, pts, dirs)), q


def test_torus_lengths_and_angles():
    r, *_ = read("oneTorus")  # lunit cm, aunit rad
    (_, rmin, rmax, rtor, sphi, dphi), = r.of("torus")
    assert (rmin, rmax, rtor) == (10.0, 100.0, 20.0)
    assert (sphi, dphi) == (1.0, 3.0)


def test_sphere_lengths_and_angles():
    r, *_ = read("oneSphere")  # lunit cm, aunit deg
    (_, rmin, rmax, sphi, dphi, sth, dth), = r.of("sphere")
    assert (rmin, rmax) == (0.0, 100.0)
    assert (sphi, sth) == (0.0, 0.0)
    assert abs(dphi - 2 * np.pi) < 1e-15 and abs(dth - np.pi) < 1e-15
