"""GPU parity: every SOA3D query of every shape vs the FP64 oracle, bit-exact."""
import numpy as np
import pytest

import oracle
import vecgeom_amd as vg
from vecgeom_amd import sampling
from tests.shapes_spec import SHAPES, transform12

pytestmark = pytest.mark.gpu

METHODS = ["Contains", "Inside", "DistanceToIn", "DistanceToOut", "SafetyToIn", "SafetyToOut",
           "PlacedDistanceToOut"]


def _gpu_query(ctx, pv, method, pts, dirs):
    P = vg.SOA3D(*pts)
    D = vg.SOA3D(*dirs)
    if method == "Contains":
        return ctx.contains(pv, P)
    if method == "Inside":
        return ctx.inside(pv, P)
    if method == "SafetyToIn":
        return ctx.safety_to_in(pv, P)
    if method == "SafetyToOut":
        return ctx.safety_to_out(pv, P)
    if method == "DistanceToIn":
        return ctx.distance_to_in(pv, P, D)
    if method == "DistanceToOut":
        return ctx.distance_to_out(pv, P, D)
    if method == "PlacedDistanceToOut":
        return ctx.placed_distance_to_out(pv, P, D)
    raise ValueError(method)


def _setup(ctx, name, placed):
    kind, args = SHAPES[name]
    geo = oracle.Geometry()
    s_o = getattr(geo, kind)(*args)
    s_g = getattr(ctx, kind)(*args)
    tr = transform12((3.0, -2.0, 1.5)) if placed else None
    pv_o = geo.place(geo.logical_volume(s_o), tr)
    pv_g = ctx.place(ctx.logical_volume(s_g), tr)
    mn, mx = geo.extent(s_o)
    return geo, pv_o, pv_g, mn, mx


@pytest.mark.parametrize("placed", [False, True], ids=["local", "placed"])
@pytest.mark.parametrize("name", sorted(SHAPES))
def test_shape_parity(ctx, name, placed):
    print(f"[parity] shape={name} placed={placed}", flush=True)
    geo, pv_o, pv_g, mn, mx = _setup(ctx, name, placed)
    n = 200_000
    half = 0.5 * (mx - mn)
    center = 0.5 * (mx + mn)
    pts = sampling.points_in_box(42, 1, n, half, 1.5, center)
    dirs = sampling.isotropic_directions(42, 2, n)
    if placed:
        # put the sample in the master frame around the placed solid
        pts = tuple(p + t for p, t in zip(pts, (3.0, -2.0, 1.5)))
    for m in METHODS:
        ref = geo.query(pv_o, m, pts, dirs if "Distance" in m else None)
        got = _gpu_query(ctx, pv_g, m, pts, dirs)
        bad = np.flatnonzero(~((ref == got) | (np.isnan(ref.astype(np.float64)) & np.isnan(got.astype(np.float64)))))
        assert bad.size == 0, (m, bad.size, bad[:5], ref[bad[:5]], got[bad[:5]])
