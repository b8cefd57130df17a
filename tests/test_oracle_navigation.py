"""Navigation oracle (CPU): linear daughter scans equal the BVH walks on the
barrel geometry, and closed-form steps and safeties along the x axis."""
import numpy as np
import pytest

import oracle
from vecgeom_amd import workloads


@pytest.fixture(scope="module")
def barrels():
    out = []
    for bvh in (False, True):
        g = oracle.Geometry()
        w = workloads.build_barrel(g)
        g.nav_close(w, use_bvh=bvh)
        out.append(g)
    return out


def test_node_numbering(barrels):
    g = barrels[1]
    assert g.nav_node(1) == (g.nav_node(1)[0], 0, 0)  # world: level 0, no mother
    pv, lvl, mo = g.nav_node(2)
    assert lvl == 1 and mo == 1  # first layer
    pv, lvl, mo = g.nav_node(3)
    assert lvl == 2 and mo == 2  # its first module


def test_linear_equals_bvh(barrels):
    lin, tree = barrels
    pts, dirs = workloads.barrel_tracks(20000, seed=3)
    s_lin, s_tree = lin.nav_locate(pts), tree.nav_locate(pts)
    assert np.array_equal(s_lin, s_tree)
    res = []
    for g, s in ((lin, s_lin), (tree, s_tree)):
        x, y, z = (a.copy() for a in pts)
        st = s.copy()
        last = np.zeros_like(st)
        bnd = np.zeros(len(st), np.uint8)
        steps = np.zeros(len(st), np.int32)
        g.nav_transport(x, y, z, *dirs, st, last, bnd, 200, steps)
        res.append((x, y, z, st, last, bnd, steps))
    for a, b in zip(*res):
        assert np.array_equal(a, b)


def test_steps_along_x_axis(barrels):
    """From the origin along +x: world -> layer i (15 mm gaps, 100 mm first),
    module k=0 of each layer spans [rmin+10, rmin+70], layer [rmin, rmin+80]."""
    g = barrels[1]
    p = [np.zeros(1), np.zeros(1), np.zeros(1)]
    d = [np.ones(1), np.zeros(1), np.zeros(1)]
    st = g.nav_locate(p)
    last = np.zeros(1, np.uint32)
    bnd = np.zeros(1, np.uint8)
    x = 0.0
    expected = [100.0]
    for i in range(20):
        expected += [10.0, 60.0, 10.0, 15.0]
    levels = []
    for k, e in enumerate(expected):
        s = g.nav_step([np.array([x]), p[1], p[2]], d, st, last, bnd)[0]
        assert abs(s - e) < 1e-9, (k, s, e)
        x += s
        levels.append(g.nav_node(int(st[0]))[1] if st[0] else -1)
    assert st[0] == 0  # left the world at x = 2000
    assert abs(x - 2000.0) < 1e-6
    assert levels[:5] == [1, 2, 1, 0, 1]


def test_safety_at_origin(barrels):
    g = barrels[1]
    p = [np.zeros(1), np.zeros(1), np.zeros(1)]
    st = g.nav_locate(p)
    assert g.nav_node(int(st[0]))[1] == 0
    assert abs(g.nav_safety(p, st)[0] - 100.0) < 1e-9


@pytest.mark.parametrize("k", [3, 13, 50])
def test_steps_along_a_rotated_module_ray(barrels, k):
    """The ray from the origin at phi_k = 2 pi k / 64 runs through the middle
    of module k of layer 0 and of the module at the same angle in every other
    layer (their counts are multiples of 64), each placed by a rotation of
    phi about z: radially they span [rmin + 10, rmin + 70] exactly as the
    unrotated modules on +x do.  A transposed rotation would tilt every
    module by 2 phi against the ray and change these crossings."""
    g = barrels[1]
    phi = 2.0 * np.pi * k / 64
    c, s = np.cos(phi), np.sin(phi)
    d = [np.array([c]), np.array([s]), np.zeros(1)]
    p = [np.zeros(1), np.zeros(1), np.zeros(1)]
    st = g.nav_locate(p)
    last = np.zeros(1, np.uint32)
    bnd = np.zeros(1, np.uint8)
    r = 0.0
    expected = [100.0]
    for i in range(20):
        expected += [10.0, 60.0, 10.0, 15.0]
    levels = []
    for j, e in enumerate(expected[:-1]):
        s_ = g.nav_step([np.array([r * c]), np.array([r * s]), np.zeros(1)], d, st, last, bnd)[0]
        assert abs(s_ - e) < 1e-9, (j, s_, e)
        r += s_
        levels.append(g.nav_node(int(st[0]))[1] if st[0] else -1)
    assert levels[:5] == [1, 2, 1, 0, 1]
    assert abs(r - 1985.0) < 1e-6


def test_compute_step_is_the_step_without_relocation(barrels):
    """ComputeStep (VNavigator.h:494-540) gives the step lengths of
    ComputeStepAndPropagatedState; its state is PrepareOutState's: the same
    state off the boundary, the current state when the mother is left, and a
    daughter of it when one is entered, with the current state as last."""
    g = barrels[1]
    pts, dirs = workloads.barrel_tracks(20000, seed=4)
    st = g.nav_locate(pts)
    s1, l1, b1 = st.copy(), np.zeros_like(st), np.zeros(len(st), np.uint8)
    s2, l2, b2 = st.copy(), np.zeros_like(st), np.zeros(len(st), np.uint8)
    lim = np.where(np.arange(len(st)) % 3 == 0, 20.0, 1e300)
    step_r = g.nav_step(pts, dirs, s1, l1, b1, step_limit=lim)
    step_c = g.nav_compute_step(pts, dirs, s2, l2, b2, step_limit=lim)
    assert np.array_equal(step_r, step_c)
    assert np.array_equal(b1, b2)
    off = (b2 == 0) & (st != 0)
    assert off.sum() > 1000 and np.array_equal(s2[off], st[off]) and np.array_equal(s1[off], st[off])
    on = (b2 == 1) & (st != 0)
    assert np.array_equal(l2[on], st[on])
    stay = on & (s2 == st)
    enter = on & (s2 != st)
    assert stay.sum() > 100 and enter.sum() > 100
    assert all(g.nav_node(int(a))[2] == int(b) for a, b in zip(s2[enter], st[enter]))
    # the relocated state of an exit is never the volume that was left
    assert np.all(s1[stay] != st[stay])
