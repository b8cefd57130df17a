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
