"""Synthetic workloads: geometry and track populations for the benchmarks.

Batched heterogeneous solids: 8 solids of each shape family with parameters
drawn log-uniform in [1, 100] mm (subject to validity constraints), polycones
with 10 z planes, polyhedra with 8 sides and 6 z planes, tessellated convex
hulls; per-solid track counts log-uniform in [1k, 4M], rescaled to the
requested total; points uniform in 1.3x the solid's bounding box, isotropic
directions; mean physics step 0.05 x bounding-box diagonal.

Every solid is described as (builder name, constructor arguments) so the same
list can be instantiated in any object exposing builders of those names.
"""
from __future__ import annotations

import math

import numpy as np

from . import sampling

FAMILIES = ("box", "tube", "cone", "trd", "trapezoid", "polycone", "polyhedron", "tessellated")


def _lu(u, lo=1.0, hi=100.0):
    """log-uniform in [lo, hi] from u in [0, 1)."""
    return float(lo * math.exp(u * math.log(hi / lo)))


def solid_spec(family: str, seed: int, index: int):
    """Constructor arguments of the index-th solid of a family (deterministic in seed)."""
    u = sampling.uniform(seed, 1000 + 16 * FAMILIES.index(family) + index, 64)
    it = iter(u)
    nx = lambda: float(next(it))
    two_pi = 2.0 * math.pi
    if family == "box":
        return ("box", (_lu(nx()), _lu(nx()), _lu(nx())))
    if family == "tube":
        rmax = _lu(nx())
        rmin = rmax * 0.8 * nx() if nx() < 0.6 else 0.0
        dz = _lu(nx())
        dphi = two_pi if nx() < 0.5 else 0.5 + (two_pi - 0.6) * nx()
        return ("tube", (rmin, rmax, dz, two_pi * nx(), dphi))
    if family == "cone":
        rmax1, rmax2 = _lu(nx()), _lu(nx())
        hollow = nx() < 0.5
        rmin1 = rmax1 * 0.7 * nx() if hollow else 0.0
        rmin2 = rmax2 * 0.7 * nx() if hollow else 0.0
        dz = _lu(nx())
        dphi = two_pi if nx() < 0.5 else 0.5 + (two_pi - 0.6) * nx()
        return ("cone", (rmin1, rmax1, rmin2, rmax2, dz, two_pi * nx(), dphi))
    if family == "trd":
        return ("trd", (_lu(nx()), _lu(nx()), _lu(nx()), _lu(nx()), _lu(nx())))
    if family == "trapezoid":
        dz = _lu(nx())
        theta, phi = 0.4 * nx(), two_pi * nx()
        dy1, dy2 = _lu(nx()), _lu(nx())
        dx1, dx2 = _lu(nx()), _lu(nx())
        # planar side faces need dx3/dx4 = dx1/dx2 scaled like dy2/dy1
        k = dy2 / dy1
        alpha = 0.5 * (nx() - 0.5)
        return ("trapezoid", (dz, theta, phi, dy1, dx1, dx2, alpha, dy2, dx1 * k, dx2 * k, alpha))
    if family == "polycone":
        nz = 10
        dz = _lu(nx(), 1.0, 20.0)
        z = [-0.5 * nz * dz + i * dz for i in range(nz)]
        rmax = [_lu(nx(), 5.0, 100.0) for _ in range(nz)]
        rmin = [r * 0.5 * nx() if i % 3 else 0.0 for i, r in enumerate(rmax)]
        dphi = two_pi if nx() < 0.5 else 1.0 + (two_pi - 1.1) * nx()
        return ("polycone", (two_pi * nx() if dphi < two_pi else 0.0, dphi, tuple(z), tuple(rmin), tuple(rmax)))
    if family == "polyhedron":
        nz = 6
        dz = _lu(nx(), 2.0, 30.0)
        z = [-0.5 * nz * dz + i * dz for i in range(nz)]
        rmax = [_lu(nx(), 5.0, 100.0) for _ in range(nz)]
        rmin = [r * 0.4 * nx() if i % 2 else 0.0 for i, r in enumerate(rmax)]
        # full phi, or a cut below 180 degrees: above it the reference's Inside
        # treats the two phi planes as an intersection (non-convex wedge)
        dphi = two_pi if nx() < 0.5 else 0.8 + 2.3 * nx()
        return ("polyhedron", (0.0, dphi, 8, tuple(z), tuple(rmin), tuple(rmax)))
    if family == "tessellated":
        # convex hull of points on an ellipsoid; triangle count log-uniform in [1k, 20k]
        ntri = int(_lu(nx(), 1000.0, 20000.0))
        axes = (_lu(nx(), 5.0, 100.0), _lu(nx(), 5.0, 100.0), _lu(nx(), 5.0, 100.0))
        return ("tessellated", ("ellipsoid_hull", ntri, axes, int(seed * 100 + index)))
    raise ValueError(family)


def track_counts(nsolids: int, total: int, seed: int = 42, lo=1_000, hi=4_000_000):
    u = sampling.uniform(seed, 77, nsolids)
    c = np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))
    c = np.maximum(1, np.round(c * total / c.sum())).astype(np.int64)
    return c


def batched_solids(total_tracks: int, families=FAMILIES, per_family: int = 8, seed: int = 42):
    """[(kind, args)], track counts per solid."""
    specs = [solid_spec(f, seed, i) for f in families for i in range(per_family)]
    counts = track_counts(len(specs), total_tracks, seed)
    return specs, counts


def tracks_for(extent, n, seed, stream, scale=1.3):
    mn, mx = extent
    mn, mx = np.asarray(mn), np.asarray(mx)
    half, cen = 0.5 * (mx - mn), 0.5 * (mx + mn)
    pts = sampling.points_in_box(seed, 2 * stream, n, half, scale, cen)
    dirs = sampling.isotropic_directions(seed, 2 * stream + 1, n)
    return pts, dirs


def mean_step(extent) -> float:
    mn, mx = extent
    return 0.05 * float(np.linalg.norm(np.asarray(mx) - np.asarray(mn)))


def build(geom, specs, counts, seed=42):
    """Instantiate the solids in `geom` (library context or oracle) and draw the tracks.

    Returns shape ids, offsets (nsolids+1), mean steps and six float64 arrays.
    """
    shapes, steps = [], []
    for kind, args in specs:
        s = make_shape(geom, kind, args)
        shapes.append(s)
        steps.append(mean_step(geom.extent(s)))
    offsets = np.zeros(len(specs) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(counts)
    n = int(offsets[-1])
    arrays = [np.empty(n) for _ in range(6)]
    for i, s in enumerate(shapes):
        a, b = offsets[i], offsets[i + 1]
        pts, dirs = tracks_for(geom.extent(s), int(b - a), seed, i)
        for k in range(3):
            arrays[k][a:b] = pts[k]
            arrays[3 + k][a:b] = dirs[k]
    return np.asarray(shapes, dtype=np.int32), offsets, np.asarray(steps), arrays


def make_shape(geom, kind, args):
    if kind == "polyhedron":
        sphi, dphi, nsides, z, rmin, rmax = args
        return geom.polyhedron(sphi, dphi, nsides, z, rmin, rmax)
    if kind == "polycone":
        sphi, dphi, z, rmin, rmax = args
        return geom.polycone(sphi, dphi, z, rmin, rmax)
    if kind == "tessellated":
        _, ntri, axes, s = args
        return geom.tessellated(ellipsoid_hull(ntri, axes, s))
    return getattr(geom, kind)(*args)


def ellipsoid_hull(ntri: int, axes, seed: int):
    """Triangles (n, 3, 3) of the convex hull of points on an ellipsoid.

    A hull of m points in convex position has 2m - 4 triangles, so m = ntri/2 + 2.
    """
    from scipy.spatial import ConvexHull

    m = ntri // 2 + 2
    u = sampling.uniform(seed, 5, m)
    v = sampling.uniform(seed, 6, m)
    z = 1.0 - 2.0 * u
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    ph = 2.0 * np.pi * v
    pts = np.stack([axes[0] * r * np.cos(ph), axes[1] * r * np.sin(ph), axes[2] * z], axis=1)
    hull = ConvexHull(pts)
    tris = pts[hull.simplices]
    # orient outward: normal . (centroid) > 0 for a hull around the origin
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    flip = np.einsum("ij,ij->i", n, tris.mean(axis=1)) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return np.ascontiguousarray(tris)


# ---- navigation workload: a barrel of tube layers holding box modules --------
def barrel_layers(nlayers: int = 20):
    """Layer table of the navigation benchmark geometry.

    World box half-lengths (2000, 2000, 3000) mm.  Layer i is a full tube with
    rmin = 100 + 95 i, thickness 80 and half-length 2800 mm, placed at the
    origin.  It holds n_i = min(64 * 2**floor(3 i / 20), 512) box modules on a
    ring of radius rmin + 40, module k rotated about z by 2 pi k / n_i.  The
    modules are 60 mm thick radially, 2700 mm long, and 0.8 of the ring pitch
    wide at their inner face, so neighbours never touch and all corners stay
    inside the layer.
    """
    layers = []
    for i in range(nlayers):
        rmin = 100.0 + 95.0 * i
        n = min(64 * 2 ** ((3 * i) // 20), 512)
        rc, hr = rmin + 40.0, 30.0
        dphi = 2.0 * np.pi / n
        hw = 0.8 * (rc - hr) * np.sin(0.5 * dphi)
        # corner inside rmax = rmin + 80
        hw = min(hw, 0.9 * np.sqrt((rmin + 80.0) ** 2 - (rc + hr) ** 2))
        layers.append({"rmin": rmin, "rmax": rmin + 80.0, "dz": 2800.0, "n": n, "rc": rc,
                       "half": (hr, hw, 2700.0)})
    return layers


def rotz_transform(phi: float, t=(0.0, 0.0, 0.0)):
    """Translation + rotation about z in the 12-double placement layout
    (local = R^T (master - t), rotation stored row-major)."""
    c, s = np.cos(phi), np.sin(phi)
    return np.array([t[0], t[1], t[2], c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0], dtype=np.float64)


def build_barrel(geom, nlayers: int = 20):
    """Build the barrel on `geom` (library context or oracle); returns the world placed volume."""
    world_lv = geom.logical_volume(geom.box(2000.0, 2000.0, 3000.0))
    for L in barrel_layers(nlayers):
        tube_lv = geom.logical_volume(geom.tube(L["rmin"], L["rmax"], L["dz"], 0.0, 2.0 * np.pi))
        mod_lv = geom.logical_volume(geom.box(*L["half"]))
        for k in range(L["n"]):
            phi = 2.0 * np.pi * k / L["n"]
            tr = rotz_transform(phi, (L["rc"] * np.cos(phi), L["rc"] * np.sin(phi), 0.0))
            geom.add_daughter(tube_lv, geom.place(mod_lv, tr))
        geom.add_daughter(world_lv, geom.place(tube_lv))
    return geom.place(world_lv)


def barrel_tracks(n: int, seed: int = 42):
    """Points uniform in the world box, isotropic directions."""
    pts = sampling.points_in_box(seed, 11, n, np.array([2000.0, 2000.0, 3000.0]), 1.0, np.zeros(3))
    dirs = sampling.isotropic_directions(seed, 12, n)
    return pts, dirs
