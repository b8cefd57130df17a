"""Mixed-precision SOA3D queries (VGM_MIXED) against the FP64 oracle on the GPU,
for every test shape, on uniform points and on a +-1e-6 mm surface band.

The mixed path lets float select and FP64 decide: classifications are float
away from the surface and FP64 near it; tube and cone distances are selected
in float among their surfaces and the selected surface is evaluated with the
FP64 kernel's own expressions (uncertain selections take the FP64 kernel);
safeties and the other families' distances are FP64.  The bar is the north
star's: every finite distance and safety within 1e-9 mm of the FP64 oracle,
no hit/miss disagreement, the same classification everywhere (surface band
included).  The fraction of results that are bit-identical is reported."""
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


TOL_MM = 1e-9


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
        ref = geo.query(pv_o, q, pts, dirs if "Distance" in q else None)
        got = _mixed_query(ctx, pv_g, q, pts, dirs)
        if q in ("Contains", "Inside"):
            bad = np.flatnonzero(np.asarray(ref) != np.asarray(got))
            print(f"[mixed] {name} {q}: {bad.size} classifications differ", flush=True)
            assert bad.size == 0, (q, bad[:5])
            continue
        rep = error_report(ref, got, scale)
        print(f"[mixed] {name} {q}: {rep}", flush=True)
        assert rep["inf_disagree"] == 0, rep
        assert rep["max_abs"] <= TOL_MM, rep
        # the sign conventions (-1 on the wrong side, 0 on the surface) too
        assert np.array_equal(np.sign(np.asarray(ref)), np.sign(np.asarray(got))), q


@pytest.mark.parametrize("name", ["tube_c2", "cone_c2"])
def test_mixed_distances_bit_exact_config2_shapes(ctx, name):
    """The config-2 tube and cone: the selected surface is evaluated by the
    FP64 kernel's expressions, so the mixed distances are the FP64 ones bit
    for bit (200k uniform + 200k surface-band points, placed and local)."""
    print(f"[mixed] bit-exact {name}", flush=True)
    import vecgeom_amd as vg

    kind, args = SHAPES[name]
    geo = oracle.Geometry()
    s_o, s_g = getattr(geo, kind)(*args), getattr(ctx, kind)(*args)
    pv_o, pv_g = geo.place(geo.logical_volume(s_o)), ctx.place(ctx.logical_volume(s_g))
    mn, mx = ctx.extent(s_g)
    pts, dirs = make_inputs(geo, pv_o, mn, mx, 200_000, seed=21)
    for q in ("DistanceToIn", "DistanceToOut"):
        ref = geo.query(pv_o, q, pts, dirs)
        got = _mixed_query(ctx, pv_g, q, pts, dirs)
        bad = np.flatnonzero(ref != got)
        assert bad.size == 0, (q, bad.size, bad[:5], ref[bad[:5]], got[bad[:5]])
