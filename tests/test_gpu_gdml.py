"""Geometries read from the reference's GDML test files, on the GPU against
the CPU oracle built from the same files, bit-exact: the six SOA3D queries
of every solid on uniform and surface-band points, then locate, one step
with safeties and a 50-step transport through the placement tree.  A torus's
distances, and the steps and positions that follow from them, are held to the
torus's root-refinement tolerance instead (as in the shape parity test)."""
import os

import numpy as np
import pytest

import oracle
import vecgeom_amd as vg
from vecgeom_amd import gdml, sampling
from tests.test_gpu_shapes import TORUS_REL_TOL, compare, compare_refined, gpu_query, make_inputs

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gdml")
FILES = ["aConstant", "aPosition", "aRotation", "auxiliary", "oneBox", "oneConeSegment", "oneCutTube", "oneOrb", "oneGeneralTrapezoid", "onePolycone",
         "onePolyhedron", "oneTesselatedSolid", "oneTrapezoid", "oneTube", "solidIntersection", "solidSubtraction", "solidUnion", "twoBoxes",
         "oneSphere", "oneTorus", "compositeMaterial", "mixtureElement", "mixtureMaterial", "simpleMaterial",
         "unionWIthDetector"]
QUERIES = ["Contains", "Inside", "DistanceToIn", "DistanceToOut", "SafetyToIn", "SafetyToOut"]


@pytest.mark.parametrize("name", FILES)
def test_gdml_geometry(ctx, name):
    print(f"[gdml] {name}", flush=True)
    path = os.path.join(HERE, name + ".gdml")
    refined = name == "oneTorus"
    geo = oracle.Geometry()
    w_o, solids_o, _ = gdml.load(path, geo)
    w_g, solids_g, _ = gdml.load(path, ctx)
    assert sorted(solids_o) == sorted(solids_g)
    for sname in sorted(solids_o):
        pv_o = geo.place(geo.logical_volume(solids_o[sname]))
        pv_g = ctx.place(ctx.logical_volume(solids_g[sname]))
        mn, mx = ctx.extent(solids_g[sname])
        pts, dirs = make_inputs(geo, pv_o, mn, mx, 4000, seed=5)
        for q in QUERIES:
            ref = geo.query(pv_o, q, pts, dirs)
            got = gpu_query(ctx, pv_g, q, pts, dirs)
            bad = compare_refined(ref, got) if refined and "Distance" in q else compare(ref, got)
            assert bad.size == 0, (name, sname, q, bad[:5])
    # navigation through the file's placement tree
    nav = vg.Navigation(ctx, w_g)
    assert geo.nav_close(w_o, use_bvh=True) == nav.nnodes
    mn, mx = ctx.extent(solids_g[[k for k in solids_g if k.lower().startswith("world")][0]])
    n = 5000
    pts = sampling.points_in_box(7, 1, n, 0.5 * (mx - mn), 0.9, 0.5 * (mx + mn))
    dirs = sampling.isotropic_directions(7, 2, n)
    st = np.zeros(n, np.uint32)
    nav.locate(*pts, st)
    st_o = geo.nav_locate(pts)
    assert np.array_equal(st, st_o)
    s_g, l_g, b_g = st.copy(), np.zeros(n, np.uint32), np.zeros(n, np.uint8)
    step, saf = np.zeros(n), np.zeros(n)
    nav.step_safety(*pts, *dirs, s_g, l_g, b_g, step, saf)
    s_o, l_o, b_o = st_o.copy(), np.zeros(n, np.uint32), np.zeros(n, np.uint8)
    assert np.array_equal(saf, geo.nav_safety(pts, st_o))
    step_o = geo.nav_step(pts, dirs, s_o, l_o, b_o)
    assert (compare_refined(step_o, step) if refined else compare(step_o, step)).size == 0
    assert np.array_equal(s_g, s_o) and np.array_equal(l_g, l_o) and np.array_equal(b_g, b_o)
    x, y, z = (np.array(a, dtype=np.float64) for a in pts)
    s_g, l_g, b_g = st.copy(), np.zeros(n, np.uint32), np.zeros(n, np.uint8)
    steps = np.zeros(n, np.int32)
    nav.transport(x, y, z, *dirs, s_g, l_g, b_g, 50, steps)
    xo, yo, zo = (np.array(a, dtype=np.float64) for a in pts)
    s_o, l_o, b_o = st_o.copy(), np.zeros(n, np.uint32), np.zeros(n, np.uint8)
    steps_o = np.zeros(n, np.int32)
    geo.nav_transport(xo, yo, zo, *dirs, s_o, l_o, b_o, 50, steps_o)
    if refined:
        for a, b in ((x, xo), (y, yo), (z, zo)):
            # one refined step length of error per step, at most the world's diagonal each
            assert np.all(np.abs(a - b) <= 50 * TORUS_REL_TOL * float(np.linalg.norm(mx - mn)))
    else:
        assert np.array_equal(x, xo) and np.array_equal(y, yo) and np.array_equal(z, zo)
    assert np.array_equal(s_g, s_o) and np.array_equal(steps, steps_o)
