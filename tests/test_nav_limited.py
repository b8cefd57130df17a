"""Transport under a physics step with the safety first, and
BVH::ApproachNextDaughter, on the CPU oracle (oracle/navigation.cpp).

Pinned by what the navigation fixes:
- with an unlimited physics step the limited transport is the plain
  straight-line transport, bit for bit, every step computing a distance;
- carrying the safety left after a safety-only move makes the same decisions
  as recomputing it (the carried value never exceeds the fresh one and a
  fresh one is taken whenever the carried one does not cover the step), so
  both runs agree bit for bit while computing fewer safeties;
- the approach to an axis-aligned daughter box is the slab distance to its
  near face, and the BVH traversal returns what a scan of all daughters
  returns.
The GPU side is in tests/test_gpu_navigation.py.
"""
import numpy as np
import pytest

import oracle
from vecgeom_amd import sampling, workloads

N = 20000


@pytest.fixture(scope="module")
def barrel():
    geo = oracle.Geometry()
    w = workloads.build_barrel(geo)
    geo.nav_close(w, use_bvh=True)
    pts, dirs = workloads.barrel_tracks(N, seed=4)
    return geo, pts, dirs, geo.nav_locate(pts)


def _run_limited(geo, pts, dirs, st, L, nsteps, reuse):
    x, y, z = (a.copy() for a in pts)
    s, last, b = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8)
    steps, geo_steps = np.zeros(N, np.int32), np.zeros(N, np.int32)
    geo.nav_transport_limited(x, y, z, *dirs, np.full(N, L), s, last, b, nsteps, steps, geo_steps, reuse_safety=reuse)
    return x, y, z, s, last, b, steps, geo_steps


def test_unlimited_step_is_plain_transport(barrel):
    geo, pts, dirs, st = barrel
    got = _run_limited(geo, pts, dirs, st, np.finfo(np.float64).max, 40, False)
    x, y, z = (a.copy() for a in pts)
    s, last, b, steps = st.copy(), np.zeros(N, np.uint32), np.zeros(N, np.uint8), np.zeros(N, np.int32)
    geo.nav_transport(x, y, z, *dirs, s, last, b, 40, steps)
    for a, r in zip(got[:7], (x, y, z, s, last, b, steps)):
        assert np.array_equal(a, r)
    assert np.array_equal(got[7], got[6])


@pytest.mark.parametrize("L", [2.0, 25.0])
def test_carried_safety_makes_the_same_decisions(barrel, L):
    geo, pts, dirs, st = barrel
    fresh = _run_limited(geo, pts, dirs, st, L, 60, False)
    carried = _run_limited(geo, pts, dirs, st, L, 60, True)
    for a, r in zip(fresh, carried):
        assert np.array_equal(a, r)
    steps, geo_steps = fresh[6], fresh[7]
    assert np.all(geo_steps <= steps) and steps.sum() > 40 * N
    # in a barrel of centimetre-scale modules most short steps need no distance
    assert geo_steps.sum() < (0.5 if L == 2.0 else 0.9) * steps.sum()
    # a track that moved only by safety never crossed: its state is unchanged
    still = (geo_steps == 0) & (steps == 60)
    assert np.array_equal(fresh[3][still], st[still])
    for k in range(3):
        np.testing.assert_allclose(fresh[k][still], pts[k][still] + 60 * L * dirs[k][still], rtol=0, atol=1e-9)


def _box_world(use_bvh):
    geo = oracle.Geometry()
    world = geo.logical_volume(geo.box(100.0, 100.0, 100.0))
    boxes = [((5.0, 6.0, 7.0), (20.0, -3.0, 4.0)), ((3.0, 3.0, 3.0), (-30.0, 10.0, -20.0)),
             ((8.0, 2.0, 4.0), (0.0, 40.0, 30.0))]
    for half, t in boxes:
        lv = geo.logical_volume(geo.box(*half))
        geo.add_daughter(world, geo.place(lv, np.array(list(t) + [1, 0, 0, 0, 1, 0, 0, 0, 1], float)))
    w = geo.place(world)
    geo.nav_close(w, use_bvh=use_bvh)
    return geo, boxes


def test_approach_is_the_slab_distance():
    geo, boxes = _box_world(True)
    n = 20000
    pts = sampling.points_in_box(8, 1, n, np.array([95.0] * 3), 1.0, np.zeros(3))
    # rays aimed near the box centres, so that most of them hit one
    rng = np.random.default_rng(8)
    tgt = np.array([t for _, t in boxes])[rng.integers(0, len(boxes), n)] + rng.uniform(-6, 6, (n, 3))
    D = tgt - np.array(pts).T
    D /= np.linalg.norm(D, axis=1)[:, None]
    dirs = tuple(np.ascontiguousarray(D[:, k]) for k in range(3))
    st = geo.nav_locate(pts)
    world_only = st == 1
    step = np.full(n, np.finfo(np.float64).max)
    geo.nav_approach(pts, dirs, st, np.zeros(n, np.uint32), step)
    P, D = np.array(pts), np.array(dirs)
    want = np.full(n, np.inf)
    for half, t in boxes:
        lo, hi = np.array(t)[:, None] - np.array(half)[:, None], np.array(t)[:, None] + np.array(half)[:, None]
        with np.errstate(divide="ignore", invalid="ignore"):
            t1, t2 = (lo - P) / D, (hi - P) / D
        tmin, tmax = np.minimum(t1, t2).max(axis=0), np.maximum(t1, t2).min(axis=0)
        hit = (tmin <= tmax) & (tmax > 0)
        want = np.where(hit, np.minimum(want, tmin), want)
    sel = world_only & np.isfinite(want)
    assert sel.sum() > 1000
    np.testing.assert_allclose(step[sel], want[sel], rtol=1e-12, atol=1e-12)
    miss = world_only & ~np.isfinite(want)
    assert np.all(step[miss] == np.finfo(np.float64).max)


def test_approach_bvh_equals_scan(barrel):
    geo, pts, dirs, st = barrel
    scan = oracle.Geometry()
    scan.nav_close(workloads.build_barrel(scan), use_bvh=False)
    a = np.full(N, np.finfo(np.float64).max)
    b = a.copy()
    geo.nav_approach(pts, dirs, st, np.zeros(N, np.uint32), a)
    scan.nav_approach(pts, dirs, st, np.zeros(N, np.uint32), b)
    assert np.mean(a == b) > 0.9999 and np.mean(a < np.finfo(np.float64).max) > 0.2
