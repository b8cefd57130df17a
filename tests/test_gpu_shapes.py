"""GPU parity: every SOA3D query of every shape vs the FP64 oracle, bit-exact.
The torus's distances are the one exception: they solve a quartic through
cbrt, pow, cos and atan2 (SolveCubic, TorusImplementation2.h:24-64), whose
device and host versions may round differently, and then refine the root by
Newton steps until one is below kTolerance; they are held to
|gpu - oracle| <= 1e-9 * max(1, |oracle|) (the refinement's tolerance),
with the same hit/miss and sign everywhere, and the share that is bit-exact
is printed.

Inputs: uniform points in 1.5x the bounding box plus a surface band (entry and
exit points of random rays, pushed by U(-1e-6, 1e-6) mm along the ray), with
isotropic directions; local frame and placed with a rotated transformation.
"""
import numpy as np
import pytest

import oracle
import vecgeom_amd as vg
from vecgeom_amd import sampling
from tests.shapes_spec import SHAPES, transform12

pytestmark = pytest.mark.gpu

METHODS = ["Contains", "Inside", "DistanceToIn", "DistanceToOut", "SafetyToIn", "SafetyToOut",
           "PlacedDistanceToOut"]
SHIFT = np.array([3.0, -2.0, 1.5])


def gpu_query(ctx, pv, method, pts, dirs):
    P = vg.SOA3D(*pts)
    D = vg.SOA3D(*dirs)
    fn = {
        "Contains": lambda: ctx.contains(pv, P),
        "Inside": lambda: ctx.inside(pv, P),
        "SafetyToIn": lambda: ctx.safety_to_in(pv, P),
        "SafetyToOut": lambda: ctx.safety_to_out(pv, P),
        "DistanceToIn": lambda: ctx.distance_to_in(pv, P, D),
        "DistanceToOut": lambda: ctx.distance_to_out(pv, P, D),
        "PlacedDistanceToOut": lambda: ctx.placed_distance_to_out(pv, P, D),
    }[method]
    return fn()


def surface_band(geo, pv_local, pts, dirs, seed):
    """Entry/exit points of the rays (oracle distances), jittered by +-1e-6 mm along the ray."""
    din = geo.query(pv_local, "DistanceToIn", pts, dirs)
    dout = geo.query(pv_local, "DistanceToOut", pts, dirs)
    t = np.where((din > 0) & (din < 1e30), din, np.where((dout > 0) & (dout < 1e30), dout, 0.0))
    jit = (2.0 * sampling.uniform(seed, 9, len(t)) - 1.0) * 1e-6
    return tuple(pts[k] + (t + jit) * dirs[k] for k in range(3))


def make_inputs(geo, pv_local, mn, mx, n, seed=42):
    half, cen = 0.5 * (mx - mn), 0.5 * (mx + mn)
    pts = sampling.points_in_box(seed, 1, n, half, 1.5, cen)
    dirs = sampling.isotropic_directions(seed, 2, n)
    band = surface_band(geo, pv_local, pts, dirs, seed)
    dirs2 = sampling.isotropic_directions(seed, 3, n)
    P = tuple(np.concatenate([pts[k], band[k]]) for k in range(3))
    D = tuple(np.concatenate([dirs[k], dirs2[k]]) for k in range(3))
    return P, D


TORUS_REL_TOL = 1e-9


def compare_refined(ref, got, rel=TORUS_REL_TOL):
    """Indices where two distance arrays disagree beyond the root
    refinement's tolerance, in hit/miss or in sign."""
    ref = np.asarray(ref, dtype=np.float64)
    got = np.asarray(got, dtype=np.float64)
    big_r, big_g = ref >= 1e300, got >= 1e300
    close = np.abs(ref - got) <= rel * np.maximum(1.0, np.abs(ref))
    ok = (ref == got) | ((big_r == big_g) & ~big_r & close & (np.sign(ref) == np.sign(got)))
    return np.flatnonzero(~ok)


def compare(ref, got):
    ref = np.asarray(ref)
    got = np.asarray(got)
    same = (ref == got)
    if ref.dtype.kind == "f":
        same |= np.isnan(ref) & np.isnan(got)
    return np.flatnonzero(~same)


@pytest.mark.parametrize("placed", [False, True], ids=["local", "placed"])
@pytest.mark.parametrize("name", sorted(SHAPES))
def test_shape_parity(ctx, name, placed):
    print(f"[parity] shape={name} placed={placed}", flush=True)
    kind, args = SHAPES[name]
    geo = oracle.Geometry()
    s_o = getattr(geo, kind)(*args)
    s_g = getattr(ctx, kind)(*args)
    lv_o, lv_g = geo.logical_volume(s_o), ctx.logical_volume(s_g)
    pv_local = geo.place(lv_o)
    tr = transform12(SHIFT) if placed else None
    pv_o, pv_g = geo.place(lv_o, tr), ctx.place(lv_g, tr)
    mn, mx = geo.extent(s_o)
    pts, dirs = make_inputs(geo, pv_local, mn, mx, 100_000)
    if placed:
        # the same local points expressed in the master frame
        T = transform12(SHIFT)
        R = T[3:].reshape(3, 3)
        loc = np.stack(pts)
        # local = R^T-convention of Transformation3D: local_i = sum_j (m_j - t_j) r[3j+i] -> m = t + R l
        pts = tuple(np.ascontiguousarray(c) for c in (SHIFT[:, None] + R @ loc))
        dirs = tuple(np.ascontiguousarray(c) for c in (R @ np.stack(dirs)))
    for m in METHODS:
        ref = geo.query(pv_o, m, pts, dirs if "Distance" in m else None)
        got = gpu_query(ctx, pv_g, m, pts, dirs)
        if kind == "torus" and "Distance" in m:
            inexact = np.flatnonzero(np.asarray(ref) != np.asarray(got))
            err = np.abs(ref[inexact] - got[inexact]) if inexact.size else np.zeros(1)
            print(f"[parity] torus {m}: {inexact.size} of {len(ref)} not bit-exact, max |diff| {err.max():.3g}",
                  flush=True)
            bad = compare_refined(ref, got)
        else:
            bad = compare(ref, got)
        assert bad.size == 0, (m, bad.size, bad[:5], ref[bad[:5]], got[bad[:5]])
    if True:  # VPlacedVolume::Normal, every family
        # the inputs above plus the rays' surface points without the jitter
        # (on the surface to rounding, where Normal is valid)
        half, cen = 0.5 * (mx - mn), 0.5 * (mx + mn)
        p0 = sampling.points_in_box(7, 1, 50_000, half, 1.5, cen)
        d0 = sampling.isotropic_directions(7, 2, 50_000)
        din = geo.query(pv_local, "DistanceToIn", p0, d0)
        dout = geo.query(pv_local, "DistanceToOut", p0, d0)
        t = np.where((din > 0) & (din < 1e30), din, np.where((dout > 0) & (dout < 1e30), dout, 0.0))
        on = np.stack([p0[k] + t * d0[k] for k in range(3)])
        if placed:
            on = SHIFT[:, None] + transform12(SHIFT)[3:].reshape(3, 3) @ on
        npts = tuple(np.ascontiguousarray(np.concatenate([pts[k], on[k]])) for k in range(3))
        ref = geo.normal(pv_o, npts)
        got = ctx.normal(pv_g, vg.SOA3D(*npts))
        for k, c in enumerate("xyzv"):
            bad = compare(ref[k], got[k])
            assert bad.size == 0, ("Normal", c, bad.size, bad[:5], ref[k][bad[:5]], got[k][bad[:5]])


def test_boolean_operand_families(ctx):
    """Operands outside the families the boolean kernels dispatch to (booleans,
    tessellated solids, tori) are refused when the solid is built, with the
    message saying so; a primitive pair builds."""
    box = ctx.box(10.0, 10.0, 10.0)
    torus = ctx.torus(0.0, 5.0, 20.0)
    tess = ctx.tessellated(*SHAPES["tess_box"][1])
    ok = ctx.boolean("union", box, ctx.sphere(0.0, 12.0))
    for bad in (torus, tess, ok):
        for left, right in ((box, bad), (bad, box)):
            with pytest.raises(vg.VgmError, match="primitive solids"):
                ctx.boolean("subtraction", left, right)
