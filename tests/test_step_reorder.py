"""Exact reorderings of the track-step, on the CPU.

The shape kernels are host+device templates; the FLOP counter
(vecgeom_amd/data/count_flops.cpp) runs them on the host in the three step
orders of the transport kernels -- the FP64 step, distance-first, and
culled (distance-first plus the bounding-box cull of outside moves) -- and
reports when a reordered run leaves the plain FP64 trajectory of any track
(position bits or inside flag after every step sequence).  Two solids of
every family of the batched workload, 300 tracks each, 50 track-steps.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "vecgeom_amd", "data"))

from vecgeom_amd import workloads  # noqa: E402


@pytest.fixture(scope="module")
def counter():
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "vecgeom_amd", "data")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    import make_flops

    return make_flops.FlopGeometry()


def _solids():
    specs, counts = workloads.batched_solids(40_000_000, per_family=8, seed=42)
    offsets = np.concatenate([[0], np.cumsum(counts)])
    out = []
    for fam in workloads.FAMILIES:
        idx = [i for i, (k, _) in enumerate(specs) if k == fam][:2]
        out += [(specs[i], int(counts[i]), int(offsets[i]), i) for i in idx]
    return out


@pytest.mark.parametrize("mode", [1, 3], ids=["distance-first", "culled"])
def test_reordered_steps_follow_fp64(counter, mode):
    g = counter
    for (kind, args), count, first, j in _solids():
        s = workloads.make_shape(g, kind, args)
        ext = g.extent(s)
        pts, dirs = workloads.tracks_for(ext, count, 42, j)
        k = min(count, 300)
        soa = [pts[0][:k], pts[1][:k], pts[2][:k], dirs[0][:k], dirs[1][:k], dirs[2][:k]]
        # raises when the reordered run leaves the FP64 trajectory of a track
        flops = g.count(s, workloads.mean_step(ext), 7, first, soa, 50, mode=mode)
        assert flops > 0, kind
