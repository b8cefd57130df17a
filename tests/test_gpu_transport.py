"""GPU batched transport vs the CPU oracle.

FP64 mode must reproduce the oracle's trajectories bit for bit (positions,
directions, sides, crossing counts), and so must the mixed, distance-first,
culled and tuned modes, which take the same decisions.
"""
import numpy as np
import pytest

import oracle
import vecgeom_amd as vg
from vecgeom_amd import workloads

pytestmark = pytest.mark.gpu

FAMILIES = ("box", "tube", "cone", "trd", "trapezoid", "polycone", "polyhedron", "tessellated")


def _setup(ctx, total, nsteps):
    import torch

    specs, counts = workloads.batched_solids(total, families=FAMILIES, per_family=4, seed=11)
    geo = oracle.Geometry()
    shapes_o, offs, steps, arr = workloads.build(geo, specs, counts, seed=11)
    shapes_g, offs_g, steps_g, _ = workloads.build(ctx, specs, counts, seed=11)
    assert np.array_equal(offs, offs_g) and np.array_equal(steps, steps_g)
    n = int(offs[-1])
    ref = [a.copy() for a in arr]
    side_o = np.zeros(n, np.int8)
    cr_o = np.zeros(n, np.int32)
    geo.batch_transport(shapes_o, offs, steps, *ref, side_o, cr_o, nsteps, step0=0, seed=5, classify=True)
    dev = [torch.from_numpy(a.copy()).cuda() for a in arr]
    side = torch.zeros(n, dtype=torch.int8, device="cuda")
    cr = torch.zeros(n, dtype=torch.int32, device="cuda")
    b = vg.Batch(ctx, shapes_g, offs_g, steps_g)
    torch.cuda.synchronize()  # torch's fills run on its stream, the library on its own
    b.classify(dev[0], dev[1], dev[2], side)
    return b, dev, side, cr, ref, side_o, cr_o, arr


@pytest.mark.parametrize("launch", [None, (128, 64, 512), (256, 1024, 8192), (64, 0, 2048), (256, 0, 512, 1)],
                         ids=["default", "small", "large", "per-tile", "one-stream"])
def test_transport_fp64_bit_exact(ctx, launch):
    print(f"[transport] fp64 launch={launch}", flush=True)
    b, dev, side, cr, ref, side_o, cr_o, _ = _setup(ctx, 120_000, 30)
    if launch:
        b.set_launch(*launch[:3])
        if len(launch) > 3:
            b.set_streams(launch[3])
            assert b.streams() == launch[3]
    b.transport(*dev, side, cr, 30, step0=0, seed=5)
    for k in range(6):
        got = dev[k].cpu().numpy()
        bad = np.flatnonzero(got != ref[k])
        assert bad.size == 0, (k, bad.size, bad[:5])
    assert np.array_equal(side.cpu().numpy(), side_o)
    assert np.array_equal(cr.cpu().numpy(), cr_o)
    b.close()


def test_transport_split_runs_equal_single_run(ctx):
    """Two launches of 15 steps (step0 = 0, 15) equal one launch of 30."""
    print("[transport] split", flush=True)
    b, dev, side, cr, ref, side_o, cr_o, _ = _setup(ctx, 60_000, 30)
    b.transport(*dev, side, cr, 15, step0=0, seed=5)
    b.transport(*dev, side, cr, 15, step0=15, seed=5)
    for k in range(6):
        assert np.array_equal(dev[k].cpu().numpy(), ref[k]), k
    assert np.array_equal(cr.cpu().numpy(), cr_o)
    b.close()


def test_transport_mixed_follows_fp64(ctx):
    """Mixed mode takes the FP64 step's decisions (a guarded float safety
    settles the steps it covers; boundary distances come from the mixed
    kernels, which return the FP64 value), so every trajectory is the
    oracle's bit for bit: 0 divergent tracks."""
    print("[transport] mixed", flush=True)
    b, dev, side, cr, ref, side_o, cr_o, _ = _setup(ctx, 120_000, 30)
    nfp64 = b.transport(*dev, side, cr, 30, step0=0, seed=5, mixed=True, count_fp64=True)
    pos = np.stack([dev[k].cpu().numpy() for k in range(3)])
    rpos = np.stack(ref[:3])
    diverged = np.any(pos != rpos, axis=0) | (side.cpu().numpy() != side_o)
    n = pos.shape[1]
    print(f"[transport] mixed: diverged {diverged.sum()} / {n}, steps not settled in float {nfp64} / {30 * n}",
          flush=True)
    assert diverged.sum() == 0
    for k in range(3, 6):
        assert np.array_equal(dev[k].cpu().numpy(), ref[k]), k
    assert np.array_equal(cr.cpu().numpy(), cr_o)
    assert 0 < nfp64 < 30 * n
    b.close()


def test_transport_mixed_equals_fp64_long_runs(ctx):
    """The benchmark's solids at 1M tracks, four launches of 50 track-steps:
    the mixed run (float safeties carried across the steps of a launch) ends
    on the FP64 run's state bit for bit.  Divergences here are rare events
    (a float safety near a facet plane), so the run is long."""
    import torch

    print("[transport] mixed long runs", flush=True)
    specs, counts = workloads.batched_solids(1_000_000, families=FAMILIES, per_family=8, seed=42)
    shapes, offs, steps, arr = workloads.build(ctx, specs, counts, seed=42)
    n = int(offs[-1])
    b = vg.Batch(ctx, shapes, offs, steps)
    init = [torch.from_numpy(a).cuda() for a in arr]
    side0 = torch.zeros(n, dtype=torch.int8, device="cuda")
    torch.cuda.synchronize()
    b.classify(init[0], init[1], init[2], side0)
    finals = []
    for mixed in (False, True):
        st = [t.clone() for t in init]
        side, cr = side0.clone(), torch.zeros(n, dtype=torch.int32, device="cuda")
        for k in range(4):
            b.transport(*st, side, cr, 50, step0=50 * k, seed=7, mixed=mixed)
        torch.cuda.synchronize()
        finals.append(st + [side, cr])
    diff = torch.zeros(n, dtype=torch.bool, device="cuda")
    for x, y in zip(*finals):
        diff |= (x != y)
    print(f"[transport] mixed long runs: {int(diff.sum())} of {n} tracks differ", flush=True)
    assert int(diff.sum()) == 0
    b.close()


@pytest.mark.parametrize("mode,launch", [("distance_first", None), ("cull", None),
                                         ("distance_first", (64, 0, 2048)), ("cull", (128, 64, 512)),
                                         ("cull", (256, 1024, 8192))],
                         ids=["distance-first", "cull", "distance-first-per-tile", "cull-small", "cull-large"])
def test_transport_distance_first_bit_exact(ctx, mode, launch):
    """Boundary distance before the safety: the same decisions as the FP64
    step (SafetyToIn/Out never exceed DistanceToIn/Out beyond the window), so
    trajectories match the oracle bit for bit for every family.  The culled
    mode also moves outside tracks whose step stays clear of the bounding box
    without any query: still bit for bit.  The launch shapes give a thread
    one track of a tile or many (the tessellated kernel's lanes run on from
    one track to their next without waiting for the wave)."""
    print(f"[transport] {mode} launch={launch}", flush=True)
    b, dev, side, cr, ref, side_o, cr_o, _ = _setup(ctx, 120_000, 30)
    if launch:
        b.set_launch(*launch)
    b.transport(*dev, side, cr, 30, step0=0, seed=5, **{mode: True})
    for k in range(6):
        got = dev[k].cpu().numpy()
        bad = np.flatnonzero(got != ref[k])
        assert bad.size == 0, (k, bad.size, bad[:5])
    assert np.array_equal(side.cpu().numpy(), side_o)
    assert np.array_equal(cr.cpu().numpy(), cr_o)
    b.close()


def test_transport_tuned_modes_exact(ctx):
    """Autotuned per-family modes without mixed precision (FP64,
    distance-first, culled) reproduce the oracle bit for bit."""
    print("[transport] tuned", flush=True)
    b, dev, side, cr, ref, side_o, cr_o, _ = _setup(ctx, 120_000, 30)
    b.autotune(*dev, side, nsteps=2, distance_first=True, cull=True)
    modes = b.modes()
    assert set(modes.values()) <= {"fp64", "distance-first", "culled"}, modes
    b.transport(*dev, side, cr, 30, step0=0, seed=5, tuned=True)
    for k in range(6):
        assert np.array_equal(dev[k].cpu().numpy(), ref[k]), k
    assert np.array_equal(cr.cpu().numpy(), cr_o)
    b.close()
