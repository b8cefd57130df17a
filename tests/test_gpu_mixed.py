"""Mixed-precision SOA3D queries (VGM_MIXED) against the FP64 path on the GPU,
for every test shape, on uniform points and on a +-1e-6 mm surface band.

Contains and Inside keep FP64 decisions near the surface, so they must equal
the FP64 classification.  Distances and safeties are float results with FP64
only near zero; their error against FP64 is reported per query and bounded
relative to the size of the solid."""
import numpy as np
import pytest

import oracle
from tests.shapes_spec import SHAPES
from tests.test_gpu_shapes import gpu_query, make_inputs

pytestmark = pytest.mark.gpu

QUERIES = ["Contains", "Inside", "DistanceToIn", "DistanceToOut", "SafetyToIn", "SafetyToOut"]


def _mixed_query(ctx, pv, method, pts, dirs):
    import vecgeom_amd as vg

    P, D = vg.SOA3D(*pts), vg.SOA3D(*dirs)
    return {
        "Contains": lambda: ctx.contains(pv, P, mixed=True),
        "Inside": lambda: ctx.inside(pv, P, mixed=True),
        "SafetyToIn": lambda: ctx.safety_to_in(pv, P, mixed=True),
        "SafetyToOut": lambda: ctx.safety_to_out(pv, P, mixed=True),
        "DistanceToIn": lambda: ctx.distance_to_in(pv, P, D, mixed=True),
        "DistanceToOut": lambda: ctx.distance_to_out(pv, P, D, mixed=True),
    }[method]()


def error_report(ref, got, scale):
    """Counts and error of float results against FP64 (finite on both sides)."""
    ref, got = np.asarray(ref, dtype=np.float64), np.asarray(got, dtype=np.float64)
    big_r, big_g = ref >= 1e300, got >= 1e300
    fin = ~big_r & ~big_g
    err = np.abs(ref[fin] - got[fin])
    return {"n": len(ref), "inf_disagree": int(np.sum(big_r != big_g)),
            "max_abs": float(err.max(initial=0.0)), "mean_abs": float(err.mean()) if err.size else 0.0,
            "p999_abs": float(np.quantile(err, 0.999)) if err.size else 0.0,
            "exact": float(np.mean(ref[fin] == got[fin])) if err.size else 1.0,
            "max_rel_scale": float(err.max(initial=0.0) / scale)}


@pytest.mark.parametrize("name", sorted(SHAPES))
def test_mixed_queries(ctx, name):
    print(f"[mixed] shape={name}", flush=True)
    kind, args = SHAPES[name]
    geo = oracle.Geometry()
    s_o, s_g = getattr(geo, kind)(*args), getattr(ctx, kind)(*args)
    pv_o = geo.place(geo.logical_volume(s_o))
    pv_g = ctx.place(ctx.logical_volume(s_g))
    mn, mx = ctx.extent(s_g)
    scale = float(np.linalg.norm(mx - mn))
    pts, dirs = make_inputs(geo, pv_o, mn, mx, 20000, seed=7)
    for q in QUERIES:
        ref = gpu_query(ctx, pv_g, q, pts, dirs)
        got = _mixed_query(ctx, pv_g, q, pts, dirs)
        if q in ("Contains", "Inside"):
            bad = np.flatnonzero(np.asarray(ref) != np.asarray(got))
            print(f"[mixed] {name} {q}: {bad.size} classifications differ", flush=True)
            assert bad.size == 0, (q, bad[:5])
            continue
        rep = error_report(ref, got, scale)
        print(f"[mixed] {name} {q}: {rep}", flush=True)
        # float rounding: errors of a few float ulps of the solid's size; hit
        # and miss can flip for grazing rays, and a grazing float ray can reach
        # a farther surface: rare, counted
        assert rep["inf_disagree"] <= max(4, 2e-3 * rep["n"]), rep
        assert rep["p999_abs"] < 1e-4 * scale, rep
