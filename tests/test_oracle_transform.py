"""Placement convention of the oracle (CPU), pinned by answers derived by hand.

The reference stores a rotation row-major (Transformation3D.h:54-63) and maps
a master point to the local frame by local_i = sum_j (m_j - t_j) r[3j + i]
(the generic branch of Transformation3D::Transform, Transformation3D.h:660-673):
local = R^T (m - t), i.e. master = R local + t.  Its Euler constructor
SetRotation(phi, theta, psi) in degrees (source/Transformation3D.cpp:129-148)
fills r[0..8] as written below in `set_rotation`.

Every case places an asymmetric solid and asks Inside / Contains at points
whose answer follows from the geometry alone: where the rotated solid lies in
the master frame.  A transposed rotation, a translation with the wrong sign
or mixed-up Euler angles puts the solid elsewhere and fails.
"""
import math

import numpy as np
import pytest

import oracle

kInside, kSurface, kOutside = 1, 2, 3


def set_rotation(phi, theta, psi):
    """Transformation3D::SetRotation(phi, theta, psi), angles in degrees
    (source/Transformation3D.cpp:129-148), as the 9 stored entries."""
    d = math.pi / 180.0
    sf, cf = math.sin(d * phi), math.cos(d * phi)
    st, ct = math.sin(d * theta), math.cos(d * theta)
    sp, cp = math.sin(d * psi), math.cos(d * psi)
    return [cp * cf - ct * sf * sp, -sp * cf - ct * sf * cp, st * sf,
            cp * sf + ct * cf * sp, -sp * sf + ct * cf * cp, -st * cf,
            sp * st, cp * st, ct]


def tr12(t, rot9):
    return np.array(list(t) + list(rot9), dtype=np.float64)


def inside(geo, pv, pts):
    p = np.asarray(pts, dtype=np.float64)
    return list(geo.query(pv, "Inside", (p[:, 0], p[:, 1], p[:, 2])))


def test_translation_kat():
    """Transformation3DTest.cpp:14-20: t1(-2,-2,-2) maps (-1,1,2) to (1,3,4):
    a unit box placed there contains master (-2.5,-1.5,-2) (local
    (-0.5,0.5,0)) and not (-1,1,2) (local (1,3,4))."""
    g = oracle.Geometry()
    pv = g.place(g.logical_volume(g.box(1.0, 1.0, 1.0)), tr12((-2.0, -2.0, -2.0), set_rotation(0, 0, 0)))
    assert inside(g, pv, [(-2.5, -1.5, -2.0), (-1.0, 1.0, 2.0), (-2.0, -2.0, -3.0)]) == [kInside, kOutside,
                                                                                         kSurface]


def test_rotation_about_z_moves_a_quadrant_sector():
    """A first-quadrant tube sector (phi 0..90 deg) rotated by SetRotation(90,0,0)
    and moved to (100,0,0) occupies the second quadrant around (100,0,0)."""
    g = oracle.Geometry()
    s = g.tube(0.0, 10.0, 5.0, 0.0, 0.5 * math.pi)
    pv = g.place(g.logical_volume(s), tr12((100.0, 0.0, 0.0), set_rotation(90, 0, 0)))
    got = inside(g, pv, [(95.0, 5.0, 0.0), (105.0, 5.0, 0.0), (95.0, -5.0, 0.0), (105.0, -5.0, 0.0),
                         (100.0, 7.0, 4.0), (93.0, 0.0, 0.0)])
    # second quadrant inside; first, third and fourth outside; on the local
    # x axis edge (master +y axis) and the local y axis edge (master -x axis)
    assert got == [kInside, kOutside, kOutside, kOutside, kSurface, kSurface]


def test_euler_composition_about_z():
    """With theta = 0 the Euler rotations about z add up: (30,0,60) is the
    same placement as (90,0,0) and (0,0,90)."""
    g = oracle.Geometry()
    s = g.tube(0.0, 10.0, 5.0, 0.0, 0.5 * math.pi)
    lv = g.logical_volume(s)
    pts = [(-5.0, 5.0, 0.0), (5.0, 5.0, 0.0), (-5.0, -5.0, 0.0), (3.0, -6.0, 1.0)]
    res = [inside(g, g.place(lv, tr12((0, 0, 0), set_rotation(*a))), pts) for a in ((90, 0, 0), (30, 0, 60),
                                                                                  (0, 0, 90))]
    assert res[0] == [kInside, kOutside, kOutside, kOutside]
    assert res[0] == res[1] == res[2]


def test_rotation_with_theta_permutes_axes():
    """SetRotation(90, 90, 0) is the cyclic permutation master = (l_z, l_x,
    l_y): a box of half-lengths (1, 2, 3) then extends 3 along master x, 1
    along y and 2 along z."""
    g = oracle.Geometry()
    rot = set_rotation(90, 90, 0)
    assert np.allclose(np.array(rot).reshape(3, 3), [[0, 0, 1], [1, 0, 0], [0, 1, 0]], atol=1e-15)
    pv = g.place(g.logical_volume(g.box(1.0, 2.0, 3.0)), tr12((0, 0, 0), rot))
    got = inside(g, pv, [(2.5, 0.0, 0.0), (0.0, 0.5, 1.5), (0.0, 1.5, 0.0), (0.0, 0.0, 2.5), (0.0, 0.0, 2.0)])
    assert got == [kInside, kInside, kOutside, kOutside, kSurface]


def test_rotated_distance_and_safety():
    """Distances and safeties go through the same placement: the permuted box
    seen from (10, 0, 0) along -x is 7 away (its face at master x = 3), and
    its SafetyToIn there is 7; from (0, 5, 0) along -y it is 4 (face y = 1)."""
    g = oracle.Geometry()
    pv = g.place(g.logical_volume(g.box(1.0, 2.0, 3.0)), tr12((0, 0, 0), set_rotation(90, 90, 0)))
    p = (np.array([10.0, 0.0]), np.array([0.0, 5.0]), np.array([0.0, 0.0]))
    d = (np.array([-1.0, 0.0]), np.array([0.0, -1.0]), np.array([0.0, 0.0]))
    din = g.query(pv, "DistanceToIn", p, d)
    sin_ = g.query(pv, "SafetyToIn", p)
    assert din[0] == pytest.approx(7.0, abs=1e-12) and din[1] == pytest.approx(4.0, abs=1e-12)
    assert sin_[0] == pytest.approx(7.0, abs=1e-12) and sin_[1] == pytest.approx(4.0, abs=1e-12)
