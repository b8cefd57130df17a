"""The allocation-free query (vgm_query_noalloc) against the allocating entry
points, through the C ABI on cuda:0.

Every method of three shapes (a box, a hollow tube sector, a tessellated
hull), placed with a rotation, is called twice in a row through
vgm_query_noalloc, the second time with a smaller, different input, and
compared bit for bit with vgm_contains / vgm_inside / vgm_distance_to_in /
vgm_distance_to_out / vgm_placed_distance_to_out / vgm_safety_to_in /
vgm_safety_to_out on the same points; the FP64 and the mixed-precision
paths both.  Also: the size limit and the stale-geometry check fail cleanly
before any work.
"""
import numpy as np
import pytest

import vecgeom_amd as vg
from vecgeom_amd import sampling, workloads
from tests.shapes_spec import transform12

pytestmark = pytest.mark.gpu

SHIFT = np.array([3.0, -2.0, 1.5])


def allocating(ctx, method, pv, P, D, step, mixed):
    return {
        "Contains": lambda: ctx.contains(pv, P, mixed=mixed),
        "Inside": lambda: ctx.inside(pv, P, mixed=mixed),
        "DistanceToIn": lambda: ctx.distance_to_in(pv, P, D, step, mixed=mixed),
        "DistanceToOut": lambda: ctx.distance_to_out(pv, P, D, step, mixed=mixed),
        "PlacedDistanceToOut": lambda: ctx.placed_distance_to_out(pv, P, D, step, mixed=mixed),
        "SafetyToIn": lambda: ctx.safety_to_in(pv, P, mixed=mixed),
        "SafetyToOut": lambda: ctx.safety_to_out(pv, P, mixed=mixed),
    }[method]()


def inputs(seed, n, half):
    pts = sampling.points_in_box(seed, 1, n, half, 1.5)
    dirs = sampling.isotropic_directions(seed, 2, n)
    step = np.full(n, 1e30) if seed % 2 else None
    return vg.SOA3D(*pts), vg.SOA3D(*dirs), step


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
@pytest.mark.parametrize("kind", ["box", "tube", "tessellated"])
def test_noalloc_matches_allocating_entry_points(ctx, kind, mixed):
    print(f"[noalloc] shape={kind} mixed={mixed}", flush=True)
    if kind == "box":
        shape, half = ctx.box(10.0, 20.0, 30.0), (10.0, 20.0, 30.0)
    elif kind == "tube":
        shape, half = ctx.tube(5.0, 20.0, 30.0, 0.0, 1.5 * np.pi), (20.0, 20.0, 30.0)
    else:
        shape, half = ctx.tessellated(workloads.ellipsoid_hull(400, (30.0, 20.0, 45.0), 3)), (30.0, 20.0, 45.0)
    half = tuple(h + 4.0 for h in half)  # the placement's shift stays inside the sampled box
    pv = ctx.place(ctx.logical_volume(shape), transform12(SHIFT))
    for call, n in enumerate((200_000, 70_001)):  # the second call is smaller and different
        P, D, step = inputs(11 + call, n, half)
        for method in vg.NOALLOC_METHODS:
            ref = allocating(ctx, method, pv, P, D, step, mixed)
            dirs = D if "Distance" in method else None
            got = ctx.query_noalloc(method, pv, P, dirs, step if dirs is not None else None, mixed=mixed)
            assert got.dtype == ref.dtype and got.shape == ref.shape
            bad = np.flatnonzero(got != ref)
            assert bad.size == 0, (kind, method, call, bad.size, bad[:4], ref[bad[:4]], got[bad[:4]])


def test_noalloc_rejects_oversize_and_stale_geometry(ctx):
    print("[noalloc] limits", flush=True)
    L = vg.lib()
    shape = ctx.box(1.0, 2.0, 3.0)
    pv = ctx.place(ctx.logical_volume(shape))
    P, D, _ = inputs(3, 1000, (1.0, 2.0, 3.0))
    ctx.contains(pv, P)  # uploads the geometry
    # 7 arrays of n doubles + n doubles out: above the limit
    n_big = vg.NOALLOC_SCRATCH_BYTES // 64 + 1024
    x = np.zeros(n_big)
    out = np.empty(n_big)
    rc = L.vgm_query_noalloc(ctx.handle, pv, 2, x.ctypes.data, x.ctypes.data, x.ctypes.data, x.ctypes.data,
                             x.ctypes.data, x.ctypes.data, x.ctypes.data, n_big, out.ctypes.data, 0)
    assert rc == -1 and b"VGM_NOALLOC_SCRATCH_BYTES" in L.vgm_last_error()
    # a new volume after the last upload: the geometry on the device is stale
    pv2 = ctx.place(ctx.logical_volume(ctx.box(2.0, 2.0, 2.0)))
    o8 = np.empty(1000, np.uint8)
    rc = L.vgm_query_noalloc(ctx.handle, pv2, 0, P.x.ctypes.data, P.y.ctypes.data, P.z.ctypes.data, None, None,
                             None, None, 1000, o8.ctypes.data, 0)
    assert rc == -3, L.vgm_last_error()
    ref = ctx.contains(pv2, P)  # an allocating call uploads it again
    got = ctx.query_noalloc("Contains", pv2, P)
    assert np.array_equal(got, ref)
    # device pointers are refused
    rc = L.vgm_query_noalloc(ctx.handle, pv, 0, P.x.ctypes.data, P.y.ctypes.data, P.z.ctypes.data, None, None,
                             None, None, 1000, o8.ctypes.data, vg.DEVICE_PTRS)
    assert rc == -1
