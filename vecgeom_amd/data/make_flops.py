"""Writes track_step_flops.json: algorithmic FP64 FLOPs per track-step of
every solid of the batched heterogeneous benchmark workload (64 solids, 8 per
family, seed 42, 40M tracks, 50 track-steps per launch, transport seed 7).

The counts come from count_flops.cpp (built by the Makefile in this folder):
the FP64 track-step run on a scalar type that counts +, -, *, /, sqrt, over
the first --sample tracks of each solid, checked to follow the plain-double
trajectory bit for bit.

    make -C vecgeom_amd/data && python vecgeom_amd/data/make_flops.py
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from vecgeom_amd import workloads  # noqa: E402

LIB = os.path.join(ROOT, "_build", "flops", "libflops.so")
_P = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_P)


class FlopGeometry:
    """The shape constructors workloads.make_shape calls, over the counter."""

    def __init__(self):
        L = ctypes.CDLL(LIB)
        L.cf_create.restype = ctypes.c_void_p
        L.cf_error.restype = ctypes.c_char_p
        L.cf_error.argtypes = [ctypes.c_void_p]
        for name in ("cf_box", "cf_tube", "cf_cone", "cf_trd"):
            getattr(L, name).argtypes = [ctypes.c_void_p] + [ctypes.c_double] * {"cf_box": 3, "cf_tube": 5,
                                                                                "cf_cone": 7, "cf_trd": 5}[name]
        L.cf_trapezoid.argtypes = [ctypes.c_void_p, _P]
        L.cf_polycone.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, _P, _P, _P]
        L.cf_polyhedron.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    _P, _P, _P]
        L.cf_tessellated.argtypes = [ctypes.c_void_p, ctypes.c_int, _P]
        L.cf_extent.argtypes = [ctypes.c_void_p, ctypes.c_int, _P, _P]
        L.cf_lv.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.cf_place.argtypes = [ctypes.c_void_p, ctypes.c_int, _P]
        L.cf_add_daughter.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.cf_nav_close.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.cf_nav_count.argtypes = [ctypes.c_void_p, ctypes.c_size_t] + [_P] * 6 + [
            ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
        L.cf_count.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64,
                               ctypes.c_size_t] + [_P] * 6 + [ctypes.c_int, ctypes.c_int,
                                                              ctypes.POINTER(ctypes.c_uint64)]
        self.L = L
        self.h = ctypes.c_void_p(L.cf_create())

    def _chk(self, v):
        if v < 0:
            raise ValueError(self.L.cf_error(self.h).decode())
        return v

    def box(self, *a):
        return self._chk(self.L.cf_box(self.h, *a))

    def tube(self, *a):
        return self._chk(self.L.cf_tube(self.h, *a))

    def cone(self, *a):
        return self._chk(self.L.cf_cone(self.h, *a))

    def trd(self, *a):
        return self._chk(self.L.cf_trd(self.h, *a))

    def trapezoid(self, *a):
        return self._chk(self.L.cf_trapezoid(self.h, _p(np.asarray(a, dtype=np.float64))))

    def polycone(self, sphi, dphi, z, rmin, rmax):
        return self._chk(self.L.cf_polycone(self.h, sphi, dphi, len(z), _p(z), _p(rmin), _p(rmax)))

    def polyhedron(self, sphi, dphi, nsides, z, rmin, rmax):
        return self._chk(self.L.cf_polyhedron(self.h, sphi, dphi, nsides, len(z), _p(z), _p(rmin), _p(rmax)))

    def tessellated(self, tris):
        t = np.ascontiguousarray(tris, dtype=np.float64)
        return self._chk(self.L.cf_tessellated(self.h, len(t), _p(t)))

    def extent(self, s):
        mn, mx = np.zeros(3), np.zeros(3)
        self._chk(self.L.cf_extent(self.h, s, _p(mn), _p(mx)))
        return mn, mx

    # the placement tree (workloads.build_barrel's calls)
    def logical_volume(self, shape):
        return self._chk(self.L.cf_lv(self.h, shape))

    def place(self, lv, transform=None):
        t = None if transform is None else _p(np.asarray(transform, dtype=np.float64))
        return self._chk(self.L.cf_place(self.h, lv, t))

    def add_daughter(self, mother_lv, pv):
        return self._chk(self.L.cf_add_daughter(self.h, mother_lv, pv))

    def nav_count(self, world_pv, pts, dirs, nsteps):
        """(FP64 flops, steps) of the straight-line navigation transport of the
        given tracks, checked to follow the plain-double run bit for bit."""
        self._chk(self.L.cf_nav_close(self.h, world_pv))
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in list(pts) + list(dirs)]
        f, st = ctypes.c_uint64(0), ctypes.c_uint64(0)
        rc = self.L.cf_nav_count(self.h, len(arrs[0]), *[a.ctypes.data_as(_P) for a in arrs], nsteps,
                                 ctypes.byref(f), ctypes.byref(st))
        if rc == -2:
            raise RuntimeError("counted navigation left the FP64 path")
        self._chk(rc)
        return int(f.value), int(st.value)

    def count(self, shape, lam, seed, first, soa, nsteps, mode=0):
        """mode: 0 FP64 step, 1 distance-first, 3 culled (distance-first and bounding-box cull)."""
        out = ctypes.c_uint64(0)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in soa]
        rc = self.L.cf_count(self.h, shape, lam, seed, first, len(arrs[0]), *[a.ctypes.data_as(_P) for a in arrs],
                             nsteps, mode, ctypes.byref(out))
        if rc == -2:
            raise RuntimeError("counted run left the FP64 trajectory")
        self._chk(rc)
        return int(out.value)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--tracks", type=int, default=40_000_000)
    ap.add_argument("--track-steps", type=int, default=50)
    ap.add_argument("--sample", type=int, default=2000, help="tracks counted per solid")
    ap.add_argument("--out", default=os.path.join(HERE, "track_step_flops.json"))
    ap.add_argument("--nav", action="store_true",
                    help="count the barrel navigation instead (writes nav_step_flops.json)")
    ap.add_argument("--nav-sample", type=int, default=20000)
    ap.add_argument("--nav-steps", type=int, default=200)
    args = ap.parse_args()
    if args.nav:
        return nav_main(args)
    g = FlopGeometry()
    specs, counts = workloads.batched_solids(args.tracks, per_family=8, seed=42)
    offsets = np.concatenate([[0], np.cumsum(counts)])
    per_solid, per_solid_df, per_solid_cu = [], [], []
    t0 = time.time()
    for j, ((kind, a), c) in enumerate(zip(specs, counts)):
        s = workloads.make_shape(g, kind, a)
        ext = g.extent(s)
        pts, dirs = workloads.tracks_for(ext, int(c), 42, j)
        k = min(int(c), args.sample)
        soa = [pts[0][:k], pts[1][:k], pts[2][:k], dirs[0][:k], dirs[1][:k], dirs[2][:k]]
        lam = workloads.mean_step(ext)
        f = g.count(s, lam, 7, int(offsets[j]), soa, args.track_steps)
        fd = g.count(s, lam, 7, int(offsets[j]), soa, args.track_steps, mode=1)
        fc = g.count(s, lam, 7, int(offsets[j]), soa, args.track_steps, mode=3)
        per_solid.append(f / (k * args.track_steps))
        per_solid_df.append(fd / (k * args.track_steps))
        per_solid_cu.append(fc / (k * args.track_steps))
        print(f"{j:2d} {kind:12s} {per_solid[-1]:10.1f} flops/track-step, distance-first {per_solid_df[-1]:10.1f} "
              f"culled {per_solid_cu[-1]:10.1f} "
              f"({time.time() - t0:.0f}s)", flush=True)
    def by_family(vals):
        acc = {}
        for (kind, _), c, f in zip(specs, counts, vals):
            d = acc.setdefault(kind, [0.0, 0])
            d[0] += f * int(c)
            d[1] += int(c)
        return {k: v[0] / v[1] for k, v in acc.items()}

    total = sum(f * int(c) for f, c in zip(per_solid, counts)) / float(sum(counts))
    out = {
        "description": "algorithmic FP64 flops (+ - * / sqrt) per track-step of the batched transport, "
                       "steps 0..track_steps-1 of the first `sample` tracks of each solid; *_distance_first: the "
                       "distance-first step (same decisions; families without kSafetyBelowDistance count "
                       "the FP64 step); *_culled: the distance-first step, and outside moves clear of the "
                       "solid's bounding box skip the queries (the bounding-box test is counted)",
        "workload": {"tracks": int(sum(counts)), "solids": len(specs), "seed": 42, "transport_seed": 7,
                     "track_steps": args.track_steps, "sample": args.sample},
        "families": [kind for kind, _ in specs],
        "counts": [int(c) for c in counts],
        "per_solid": per_solid,
        "per_family": by_family(per_solid),
        "per_solid_distance_first": per_solid_df,
        "per_family_distance_first": by_family(per_solid_df),
        "per_solid_culled": per_solid_cu,
        "per_family_culled": by_family(per_solid_cu),
        "fp64_flops_per_track_step": total,
    }
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"mean {total:.1f} flops/track-step -> {args.out}")


def nav_main(args):
    """FP64 flops per step of the barrel navigation benchmark: the first
    --nav-sample of its tracks (seed 42), --nav-steps straight-line steps."""
    g = FlopGeometry()
    world = workloads.build_barrel(g)
    pts, dirs = workloads.barrel_tracks(args.nav_sample, seed=42)
    t0 = time.time()
    f, st = g.nav_count(world, pts, dirs, args.nav_steps)
    out = {"description": "algorithmic FP64 flops (+ - * / sqrt) per step of the straight-line navigation "
                          "transport on the barrel (ComputeStepAndPropagatedState: mother DistanceToOut, BVH "
                          "daughter walk with its box tests, candidate DistanceToIn, relocation, and the move of "
                          "the global point), the first `sample` tracks of the benchmark's track set",
           "workload": {"geometry": "barrel (20 tube layers, box modules)", "sample": args.nav_sample,
                        "max_steps": args.nav_steps, "seed": 42},
           "steps": st, "flops": f, "flops_per_step": f / st}
    path = os.path.join(HERE, "nav_step_flops.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"{f / st:.1f} flops per navigation step over {st} steps ({time.time() - t0:.0f}s) -> {path}")


if __name__ == "__main__":
    main()
