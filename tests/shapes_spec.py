"""Shape specifications shared by the oracle and GPU parity tests.

Each spec is (builder name, constructor arguments); both vecgeom_amd.Context and
oracle.Geometry expose builders with the same names and argument order.
"""
import math

import numpy as np

PI = math.pi


def rotation_matrix(axis, angle):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    x, y, z = a
    c, s = math.cos(angle), math.sin(angle)
    C = 1 - c
    return np.array([
        [c + x * x * C, x * y * C - z * s, x * z * C + y * s],
        [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
        [z * x * C - y * s, z * y * C + x * s, c + z * z * C],
    ])


def transform12(translation, axis=(1.0, 2.0, 3.0), angle=0.7):
    """Translation + rotation about an oblique axis, in the reference's 12-double order."""
    return np.concatenate([np.asarray(translation, dtype=np.float64), rotation_matrix(axis, angle).reshape(-1)])


SHAPES = {
    # config 1
    "box_c1": ("box", (10.0, 20.0, 30.0)),
    "box_thin": ("box", (0.15, 24.707, 22.7)),
    # config 2 tube and cone
    "tube_c2": ("tube", (5.0, 20.0, 30.0, 0.0, 1.5 * PI)),
    "cone_c2": ("cone", (0.0, 10.0, 5.0, 25.0, 40.0, 0.0, 2 * PI)),
    # other tube and cone branches (phi <= pi, solid, thin)
    "tube_solid": ("tube", (0.0, 50.0, 50.0, 0.0, 2 * PI)),
    "tube_sector": ("tube", (0.0, 50.0, 50.0, 0.5 * PI, 0.5 * PI)),
    "tube_hollow_sector": ("tube", (45.0, 50.0, 50.0, PI / 4.0, PI / 8.0)),
    "tube_cross0": ("tube", (2200.0, 3200.0, 2500.0, -0.68977164349384879, 3.831364227270472)),
    "cone_cut": ("cone", (50.0, 100.0, 50.0, 200.0, 50.0, -PI / 6, PI / 3)),
    "cone_ctest10": ("cone", (20.0, 60.0, 80.0, 140.0, 100.0, 10 * PI / 180.0, 300 * PI / 180.0)),
    "cone_thin": ("cone", (95.0, 100.0, 95.0, 100.0, 50.0, 0.0, 2 * PI)),
    "cone_inv": ("cone", (100.0, 150.0, 50.0, 100.0, 50.0, 0.0, 2 * PI)),
    "cone_exotic": ("cone", (50.0, 60.0, 0.0, 10.0, 50.0, 0.0, 2 * PI)),
    "cone_c5": ("cone", (25.0, 50.0, 75.0, 150.0, 50.0, 0.0, 1.5 * PI)),
}

SHAPES.update({
    # spheres: TestSphere.cpp's b4, b8, b9, s3, s41, sn13 and cut spheres on both
    # sides of the equator (the theta cone's branches)
    "sph_b4": ("sphere", (6.0, 8.0, 0.0, 2 * PI, 0.0, PI)),
    "sph_b8": ("sphere", (6.0, 8.0, PI / 6, PI / 6, 0.0, PI)),
    "sph_b9": ("sphere", (6.0, 8.0, PI / 6, PI / 6, PI / 6, PI / 6)),
    "sph_s3": ("sphere", (45.0, 50.0, 0.0, 2 * PI, PI / 4, 0.5 * PI)),
    "sph_s41": ("sphere", (5.0, 50.0, -PI, 1.5 * PI, 0.0, 2 * PI)),
    "sph_sn13": ("sphere", (0.0, 50.0, 0.0, 2 * PI, 0.75 * PI, 0.25 * PI)),
    "sph_solid": ("sphere", (0.0, 25.0, 0.0, 2 * PI, 0.0, PI)),
    "sph_band": ("sphere", (5.0, 40.0, 0.0, 2 * PI, 1.2, 1.5)),
    "sph_patch_low": ("sphere", (20.0, 40.0, 0.4, 1.1, 1.9, 0.8)),
    "sph_wide_phi": ("sphere", (0.0, 50.0, 0.3, 1.2 * PI, 0.3, 0.9)),
})

SHAPES.update({
    # tori: TestTorus2.cpp's t1, t2, tn3, t3, t6, the ring and cmsEECool3
    "torus_t1": ("torus", (0.0, 90.0, 100.0, 0.0, 2 * PI)),
    "torus_t2": ("torus", (10.0, 90.0, 100.0, 0.0, PI / 2)),
    "torus_tn3": ("torus", (10.0, 90.0, 100.0, PI / 2, 3 * PI / 2)),
    "torus_t3": ("torus", (40.0, 90.0, 100.0, PI / 2 - PI / 24, PI / 12)),
    "torus_t6": ("torus", (100.0, 150.0, 200.0, 0.0, PI / 3)),
    "torus_ring": ("torus", (0.0, 100.0, 1000.0, 0.0, 2 * PI)),
    "torus_cms": ("torus", (0.0, 6.0, 1640.0, 0.09599, 1.4312)),
})

SHAPES.update({
    # TestOrb.cpp radii
    "orb_9": ("orb", (9.0,)),
    "orb_6": ("orb", (6.0,)),
    "orb_big": ("orb", (3000.0,)),
})

def cut_normals(thb, phib, tht, phit):
    """Bottom and top cut normals from polar and azimuthal angles, as
    test/shape_tester/shape_testCutTube.cpp:50-51 builds them."""
    b = (math.sin(thb) * math.cos(phib), math.sin(thb) * math.sin(phib), math.cos(thb))
    t = (math.sin(tht) * math.cos(phit), math.sin(tht) * math.sin(phit), math.cos(tht))
    return b, t


_CT = cut_normals(3 * PI / 4, PI / 3, PI / 4, 2 * PI / 3)  # shape_testCutTube.cpp:23-31
SHAPES.update({
    # the four shape_testCutTube.cpp cases (:33-45)
    "cut_tube_t0": ("cut_tube", (0.0, 5.0, 10.0, 0.0, 2 * PI) + _CT),
    "cut_tube_t1": ("cut_tube", (3.0, 5.0, 10.0, 0.0, 2 * PI) + _CT),
    "cut_tube_t2": ("cut_tube", (0.0, 5.0, 10.0, 0.0, 2 * PI / 3) + _CT),
    "cut_tube_t3": ("cut_tube", (3.0, 5.0, 10.0, 0.0, 2 * PI / 3) + _CT),
    # shallow cuts on a wide sector with a bore
    "cut_tube_wide": ("cut_tube", (20.0, 60.0, 40.0, -0.5, 4.0) + cut_normals(PI - 0.2, 0.3, 0.15, 2.0)),
})

SHAPES.update({
    # booleans: (operation, left (kind, args), right (kind, args), left placement, right placement)
    "bool_box_minus_tube": ("boolean", ("subtraction", ("box", (20.0, 20.0, 20.0)),
                                        ("tube", (0.0, 8.0, 30.0, 0.0, 2 * PI)), None, None)),
    "bool_box_minus_rot_box": ("boolean", ("subtraction", ("box", (20.0, 15.0, 10.0)), ("box", (8.0, 8.0, 8.0)),
                                           None, transform12((10.0, 5.0, 0.0), angle=0.4))),
    "bool_tube_and_box": ("boolean", ("intersection", ("tube", (5.0, 25.0, 30.0, 0.0, 2 * PI)),
                                      ("box", (20.0, 20.0, 40.0)), None, transform12((3.0, 0.0, 0.0)))),
    "bool_cone_and_orb": ("boolean", ("intersection", ("cone", (0.0, 10.0, 5.0, 25.0, 40.0, 0.0, 2 * PI)),
                                      ("orb", (30.0,)), None, transform12((0.0, 0.0, 15.0), angle=0.0))),
    # unions: a rod in a sleeve (they touch along r = 10: the Normal branch of
    # Inside), a box with a rotated box through one face, a sector and a
    # polycone side by side
    "bool_rod_or_sleeve": ("boolean", ("union", ("tube", (0.0, 10.0, 20.0, 0.0, 2 * PI)),
                                       ("tube", (10.0, 20.0, 15.0, 0.0, 2 * PI)), None, None)),
    "bool_box_or_rot_box": ("boolean", ("union", ("box", (20.0, 15.0, 10.0)), ("box", (8.0, 8.0, 8.0)),
                                        None, transform12((20.0, 5.0, 0.0), angle=0.4))),
    "bool_sector_or_pcon": ("boolean", ("union", ("tube", (5.0, 25.0, 30.0, 0.3, 1.2 * PI)),
                                        ("polycone", (0.0, 2 * PI, (-10.0, 0.0, 15.0), (0.0, 2.0, 0.0),
                                                      (12.0, 18.0, 9.0))),
                                        None, transform12((-10.0, 4.0, 0.0), angle=0.0))),
    # TestBooleans.cpp's two holes (tubes of radius 2.25, 5 mm apart along x
    # and y), shifted so that the left one sits at the origin: the union's
    # SafetyToOut evaluates the left operand at the untransformed point
    # (BooleanUnionImplementation.h:214), which holds for an unplaced left one
    "bool_two_holes": ("boolean", ("union", ("tube", (0.0, 2.25, 10.0, 0.0, 2 * PI)),
                                   ("tube", (0.0, 2.25, 10.0, 0.0, 2 * PI)), None,
                                   transform12((5.0, 5.0, 0.0), angle=0.0))),
})

SHAPES.update({
    "trd_box": ("trd", (20.0, 20.0, 30.0, 30.0, 40.0)),
    "trd_2": ("trd", (10.0, 30.0, 20.0, 40.0, 40.0)),
    "trd_1": ("trd", (10.0, 30.0, 20.0, 20.0, 40.0)),
    "trap_box": ("trapezoid", (40.0, 0.0, 0.0, 30.0, 20.0, 20.0, 0.0, 30.0, 20.0, 20.0, 0.0)),
    "trap_skew": ("trapezoid", (50.0, 0.0, 0.0, 50.0, 50.0, 50.0, PI / 4, 50.0, 50.0, 50.0, PI / 4)),
    "trap_general": ("trapezoid", (60.0, 0.3, 0.4, 20.0, 15.0, 25.0, 0.2, 30.0, 22.0, 33.0, 0.2)),
})

SHAPES.update({
    "pcon_poly1": ("polycone", (0.0, 2 * PI, (-1.0, -0.5, 0.5, 2.0), (0.1, 0.0, 0.0, 0.4), (1.0, 2.0, 2.0, 1.5))),
    "pcon_simple": ("polycone", (0.0, 2 * PI, (-10.0, 0.0, 10.0), (0.0, 0.0, 0.0), (70.0, 70.0, 80.0))),
    "pcon_mypcone": ("polycone", (-10.0 * PI / 180.0, 10.0 * PI / 180.0,
                                  (-20.0, -10.0, -10.0, 0.0, 10.0, 20.0, 30.0, 40.0),
                                  (30.0, 30.0, 0.0, 0.0, 0.0, 0.0, 40.0, 40.0),
                                  (70.0, 70.0, 70.0, 40.0, 40.0, 80.0, 80.0, 60.0))),
    "pcon_trak": ("polycone", (0.0, 2 * PI, (-2935., -1899., -1899., -1899., 1899., 1899., 1899., 2935.),
                               (74., 34., 31., 31., 31., 31., 34., 74.), (1233.,) * 8)),
    "pcon_10planes": ("polycone", (0.3, 4.5, tuple(float(z) for z in range(-50, 50, 10)),
                                   (5.0, 6.0, 4.0, 8.0, 8.0, 3.0, 2.0, 5.0, 6.0, 4.0),
                                   (20.0, 25.0, 22.0, 30.0, 28.0, 26.0, 35.0, 30.0, 24.0, 20.0))),
})

# Polyhedra: (sphi, dphi, nsides, z, rmin, rmax); radii are to the sides.
SHAPES.update({
    "pgon_mypgon": ("polyhedron", (0.0, PI / 4.0, 5, (-30.0, -20.0, -10.0, 0.0, 10.0, 20.0, 30.0, 40.0),
                                   (30.0, 30.0, 0.0, 0.0, 0.0, 0.0, 40.0, 40.0),
                                   (70.0, 70.0, 70.0, 40.0, 40.0, 80.0, 80.0, 60.0))),
    "pgon_hex": ("polyhedron", (0.0, 2 * PI, 6, (-10.0, 10.0), (0.0, 0.0), (12.99, 12.99))),
    "pgon_8x6": ("polyhedron", (0.1, 2 * PI, 8, (-40.0, -25.0, -5.0, 5.0, 25.0, 40.0),
                                (5.0, 8.0, 8.0, 12.0, 6.0, 6.0), (20.0, 30.0, 25.0, 35.0, 30.0, 18.0))),
    "pgon_bigcut": ("polyhedron", (1.0, 4.0, 7, (-20.0, 0.0, 30.0), (0.0, 5.0, 10.0), (25.0, 30.0, 20.0))),
    # a zero-height middle segment (two planes at z = 0): the border case
    "pgon_smallcut": ("polyhedron", (-0.4, 1.2, 3, (-15.0, 0.0, 0.0, 20.0), (0.0, 2.0, 6.0, 0.0),
                                     (20.0, 25.0, 30.0, 15.0))),
})


def _tess(kind, *args):
    from tests.golden import kats
    from vecgeom_amd import workloads
    if kind == "trd":
        return (kats.trd_like_triangles(*args),)
    return (workloads.ellipsoid_hull(*args),)


# Tessellated solids: the reference's trd-like KAT constructions and convex
# hulls of points on an ellipsoid.
SHAPES.update({
    "tess_box": ("tessellated", _tess("trd", 20.0, 20.0, 30.0, 30.0, 40.0)),
    "tess_trd": ("tessellated", _tess("trd", 10.0, 30.0, 20.0, 40.0, 40.0)),
    "tess_hull_300": ("tessellated", _tess("hull", 300, (30.0, 20.0, 45.0), 3)),
    "tess_hull_2k": ("tessellated", _tess("hull", 2000, (12.0, 50.0, 30.0), 4)),
})
