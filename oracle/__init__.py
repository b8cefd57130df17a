"""FP64 CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/liborc.so, a scalar C++ restatement of the reference
algorithms (see the file headers under oracle/ for the reference file:line each
part follows).  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg use it, always as the checker and never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

# Direction of the Inside/Contains test ray of tessellated solids, the same
# default as the library's (sin 1.1 cos 0.7, sin 1.1 sin 0.7, cos 1.1).
TESS_TEST_DIR = (0.681632986593423, 0.5741315443479861, 0.4535961214255773)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborc.so")

CONTAINS, INSIDE, DISTANCE_TO_IN, DISTANCE_TO_OUT, SAFETY_TO_IN, SAFETY_TO_OUT, PLACED_DISTANCE_TO_OUT = range(7)
METHODS = {
    "Contains": CONTAINS,
    "Inside": INSIDE,
    "DistanceToIn": DISTANCE_TO_IN,
    "DistanceToOut": DISTANCE_TO_OUT,
    "SafetyToIn": SAFETY_TO_IN,
    "SafetyToOut": SAFETY_TO_OUT,
    "PlacedDistanceToOut": PLACED_DISTANCE_TO_OUT,
}

_lib = None
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_SZ = ctypes.c_size_t


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_geom_create.restype = _P
        L.orc_geom_create.argtypes = []
        L.orc_geom_destroy.argtypes = [_P]
        L.orc_geom_destroy.restype = None
        sig = {
            "orc_box_create": [_P, _D, _D, _D],
            "orc_orb_create": [_P, _D],
            "orc_sphere_create": [_P, _D, _D, _D, _D, _D, _D],
            "orc_torus_create": [_P, _D, _D, _D, _D, _D],
            "orc_cut_tube_create": [_P] + [_D] * 11,
            "orc_boolean_create": [_P, _I, _I, _P, _I, _P],
            "orc_tube_create": [_P, _D, _D, _D, _D, _D],
            "orc_cone_create": [_P, _D, _D, _D, _D, _D, _D, _D],
            "orc_trd_create": [_P, _D, _D, _D, _D, _D],
            "orc_polycone_create": [_P, _D, _D, _I, _P, _P, _P],
            "orc_polyhedron_create": [_P, _D, _D, _I, _I, _P, _P, _P],
            "orc_tessellated_create": [_P, _I, _P, _P, _I],
            "orc_nav_close": [_P, _I, _I],
            "orc_nav_node": [_P, ctypes.c_uint, _P, _P, _P],
            "orc_nav_locate": [_P, _P, _P, _P, ctypes.c_size_t, _P, _I],
            "orc_nav_step": [_P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_size_t, _P, _P, _P, _P, _I],
            "orc_nav_step2": [_P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_size_t, _P, _P, _P, _P, _I, _I],
            "orc_nav_transport": [_P, _P, _P, _P, _P, _P, _P, ctypes.c_size_t, _P, _P, _P, _I, _P, _I],
            "orc_nav_transport_limited": [_P] * 8 + [ctypes.c_size_t, _P, _P, _P, _I, _I, _P, _P, _I],
            "orc_nav_approach": [_P] * 7 + [ctypes.c_size_t, _P, _P, _P, _I],
            "orc_nav_safety": [_P, _P, _P, _P, ctypes.c_size_t, _P, _P, _I],
            "orc_nav_level_locate": [_P] * 7 + [ctypes.c_size_t] + [_P] * 6 + [_I],
            "orc_trapezoid_create": [_P] + [_D] * 11,
            "orc_shape_extent": [_P, _I, _P, _P],
            "orc_lv_create": [_P, _I],
            "orc_pv_create": [_P, _I, _P],
            "orc_lv_add_daughter": [_P, _I, _I],
            "orc_query": [_P, _I, _I, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _I],
            "orc_normal": [_P, _I, _P, _P, _P, _SZ, _P, _P, _P, _P, _I],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _I
        L.orc_batch_transport.argtypes = [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I,
                                          ctypes.c_uint64, _I, _I]
        L.orc_batch_transport.restype = _I
        L.orc_physics_step.argtypes = [_D, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]
        L.orc_physics_step.restype = _D
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


class Geometry:
    """Oracle geometry: shapes, logical volumes, placed volumes."""

    def __init__(self):
        self._L = lib()
        self.h = self._L.orc_geom_create()

    def __del__(self):
        try:
            if self.h:
                self._L.orc_geom_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _chk(self, v):
        if v < 0:
            raise RuntimeError("oracle rejected the shape parameters")
        return v

    def box(self, dx, dy, dz):
        return self._chk(self._L.orc_box_create(self.h, dx, dy, dz))

    def orb(self, r):
        return self._chk(self._L.orc_orb_create(self.h, r))

    def sphere(self, rmin, rmax, sphi=0.0, dphi=2 * np.pi, stheta=0.0, dtheta=np.pi):
        return self._chk(self._L.orc_sphere_create(self.h, rmin, rmax, sphi, dphi, stheta, dtheta))

    def torus(self, rmin, rmax, rtor, sphi=0.0, dphi=2 * np.pi):
        return self._chk(self._L.orc_torus_create(self.h, rmin, rmax, rtor, sphi, dphi))

    def boolean(self, op, left, right, left_tr=None, right_tr=None):
        ops = {"intersection": 1, "subtraction": 2, "union": 3}
        ids = [getattr(self, s[0])(*s[1]) if isinstance(s, tuple) else s for s in (left, right)]
        trs = [None if t is None else np.ascontiguousarray(t, dtype=np.float64) for t in (left_tr, right_tr)]
        return self._chk(self._L.orc_boolean_create(self.h, ops[op], ids[0], None if trs[0] is None else _p(trs[0]),
                                                    ids[1], None if trs[1] is None else _p(trs[1])))

    def cut_tube(self, rmin, rmax, z, sphi, dphi, bottom, top):
        return self._chk(self._L.orc_cut_tube_create(self.h, rmin, rmax, z, sphi, dphi, *bottom, *top))

    def tube(self, rmin, rmax, dz, sphi, dphi):
        return self._chk(self._L.orc_tube_create(self.h, rmin, rmax, dz, sphi, dphi))

    def cone(self, rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi):
        return self._chk(self._L.orc_cone_create(self.h, rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi))

    def trd(self, x1, x2, y1, y2, z):
        return self._chk(self._L.orc_trd_create(self.h, x1, x2, y1, y2, z))

    def trapezoid(self, dz, theta, phi, dy1, dx1, dx2, alpha1, dy2, dx3, dx4, alpha2):
        return self._chk(self._L.orc_trapezoid_create(self.h, dz, theta, phi, dy1, dx1, dx2, alpha1, dy2, dx3,
                                                      dx4, alpha2))

    def polycone(self, sphi, dphi, z, rmin, rmax):
        z, rmin, rmax = (np.ascontiguousarray(a, dtype=np.float64) for a in (z, rmin, rmax))
        return self._chk(self._L.orc_polycone_create(self.h, sphi, dphi, len(z), _p(z), _p(rmin), _p(rmax)))

    def polyhedron(self, sphi, dphi, nsides, z, rmin, rmax):
        z, rmin, rmax = (np.ascontiguousarray(a, dtype=np.float64) for a in (z, rmin, rmax))
        return self._chk(self._L.orc_polyhedron_create(self.h, sphi, dphi, nsides, len(z), _p(z), _p(rmin),
                                                       _p(rmax)))

    def tessellated(self, triangles, test_dir=None, brute_force=False):
        """Triangles (n, 3, 3): vertices anti-clockwise seen from outside."""
        tri = np.ascontiguousarray(triangles, dtype=np.float64).reshape(-1, 9)
        td = np.ascontiguousarray(TESS_TEST_DIR if test_dir is None else test_dir, dtype=np.float64)
        return self._chk(self._L.orc_tessellated_create(self.h, len(tri), _p(tri), _p(td), int(brute_force)))

    # ---- navigation -----------------------------------------------------------
    def nav_close(self, world_pv, use_bvh=False):
        """Build navigation tables; returns the number of state nodes."""
        return self._chk(self._L.orc_nav_close(self.h, world_pv, int(use_bvh)))

    def nav_node(self, node):
        pv, lvl, mo = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        if self._L.orc_nav_node(self.h, node, ctypes.byref(pv), ctypes.byref(lvl), ctypes.byref(mo)) != 0:
            raise ValueError("invalid node")
        return pv.value, lvl.value, mo.value

    def nav_locate(self, pts, nthreads=0):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in pts)
        st = np.zeros(len(x), np.uint32)
        self._L.orc_nav_locate(self.h, _p(x), _p(y), _p(z), len(x), _p(st), nthreads)
        return st

    def nav_step(self, pts, dirs, state, last, boundary, step_limit=None, nthreads=0):
        """One ComputeStepAndPropagatedState per track; updates the state arrays, returns steps."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in pts)
        dx, dy, dz = (np.ascontiguousarray(a, dtype=np.float64) for a in dirs)
        sl = None if step_limit is None else np.ascontiguousarray(step_limit, dtype=np.float64)
        out = np.zeros(len(x))
        self._L.orc_nav_step(self.h, _p(x), _p(y), _p(z), _p(dx), _p(dy), _p(dz), None if sl is None else _p(sl),
                             len(x), _p(state), _p(last), _p(boundary), _p(out), nthreads)
        return out

    def nav_compute_step(self, pts, dirs, state, last, boundary, step_limit=None, nthreads=0):
        """One ComputeStep per track (no relocation); updates the state arrays, returns steps."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in pts)
        dx, dy, dz = (np.ascontiguousarray(a, dtype=np.float64) for a in dirs)
        sl = None if step_limit is None else np.ascontiguousarray(step_limit, dtype=np.float64)
        out = np.zeros(len(x))
        self._L.orc_nav_step2(self.h, _p(x), _p(y), _p(z), _p(dx), _p(dy), _p(dz), None if sl is None else _p(sl),
                              len(x), _p(state), _p(last), _p(boundary), _p(out), 0, nthreads)
        return out

    def nav_transport(self, x, y, z, dx, dy, dz, state, last, boundary, nsteps, steps, nthreads=0):
        self._L.orc_nav_transport(self.h, _p(x), _p(y), _p(z), _p(dx), _p(dy), _p(dz), len(x), _p(state), _p(last),
                                  _p(boundary), nsteps, _p(steps), nthreads)

    def nav_transport_limited(self, x, y, z, dx, dy, dz, step_limit, state, last, boundary, nsteps, steps, geo_steps,
                              reuse_safety=False, nthreads=0):
        sl = np.ascontiguousarray(step_limit, dtype=np.float64)
        self._L.orc_nav_transport_limited(self.h, _p(x), _p(y), _p(z), _p(dx), _p(dy), _p(dz), _p(sl), len(x),
                                          _p(state), _p(last), _p(boundary), nsteps, 1 if reuse_safety else 0,
                                          _p(steps), _p(geo_steps), nthreads)

    def nav_approach(self, pts, dirs, state, last, step, nthreads=0):
        """BVH::ApproachNextDaughter per track; step is updated in place."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in pts)
        dx, dy, dz = (np.ascontiguousarray(a, dtype=np.float64) for a in dirs)
        self._L.orc_nav_approach(self.h, _p(x), _p(y), _p(z), _p(dx), _p(dy), _p(dz), len(x), _p(state), _p(last),
                                 _p(step), nthreads)

    def nav_level_locate(self, pts, state, dirs=None, excl=None, nthreads=0):
        """LevelLocate in the volume of each state (directional with dirs): (child state, daughter-frame point)."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in pts)
        st = np.ascontiguousarray(state, dtype=np.uint32)
        d = [None] * 3 if dirs is None else [np.ascontiguousarray(a, dtype=np.float64) for a in dirs]
        ex = None if excl is None else np.ascontiguousarray(excl, dtype=np.uint32)
        child = np.zeros(len(x), np.uint32)
        q = [np.zeros(len(x)) for _ in range(3)]
        self._L.orc_nav_level_locate(self.h, _p(x), _p(y), _p(z), _p(d[0]), _p(d[1]), _p(d[2]), len(x), _p(st),
                                     _p(ex), _p(child), _p(q[0]), _p(q[1]), _p(q[2]), nthreads)
        return child, q

    def nav_safety(self, pts, state, nthreads=0):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in pts)
        out = np.zeros(len(x))
        self._L.orc_nav_safety(self.h, _p(x), _p(y), _p(z), len(x), _p(state), _p(out), nthreads)
        return out

    def extent(self, shape):
        mn, mx = np.zeros(3), np.zeros(3)
        self._L.orc_shape_extent(self.h, shape, mn.ctypes.data, mx.ctypes.data)
        return mn, mx

    def logical_volume(self, shape):
        return self._L.orc_lv_create(self.h, shape)

    def place(self, lv, transform=None):
        tr = None if transform is None else np.ascontiguousarray(transform, dtype=np.float64)
        return self._L.orc_pv_create(self.h, lv, _p(tr))

    def add_daughter(self, mother_lv, pv):
        self._L.orc_lv_add_daughter(self.h, mother_lv, pv)

    def batch_transport(self, shapes, offsets, mean_step, x, y, z, dx, dy, dz, side, crossings, nsteps, step0=0,
                        seed=1, classify=False, nthreads=0):
        """Track-step loop over solid i's tracks [offsets[i], offsets[i+1]); arrays updated in place."""
        shapes = np.ascontiguousarray(shapes, dtype=np.int32)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        mean_step = np.ascontiguousarray(mean_step, dtype=np.float64)
        self._L.orc_batch_transport(self.h, len(shapes), _p(shapes), _p(offsets), _p(mean_step), _p(x), _p(y), _p(z),
                                    _p(dx), _p(dy), _p(dz), _p(side), _p(crossings), nsteps, step0, seed,
                                    1 if classify else 0, nthreads)

    def query(self, pv, method, pts, dirs=None, step_max=None, nthreads=0):
        m = METHODS[method] if isinstance(method, str) else method
        n = len(pts[0])
        dt = np.uint8 if m == CONTAINS else (np.int32 if m == INSIDE else np.float64)
        out = np.empty(n, dtype=dt)
        px, py, pz = (np.ascontiguousarray(a, dtype=np.float64) for a in pts)
        if dirs is not None:
            dx, dy, dz = (np.ascontiguousarray(a, dtype=np.float64) for a in dirs)
        else:
            dx = dy = dz = None
        st = None if step_max is None else np.ascontiguousarray(step_max, dtype=np.float64)
        self._L.orc_query(self.h, pv, m, _p(px), _p(py), _p(pz), _p(dx), _p(dy), _p(dz), _p(st), n,
                          out.ctypes.data, nthreads)
        return out

    def normal(self, pv, pts, nthreads=0):
        """VPlacedVolume::Normal of master-frame points: (nx, ny, nz, valid)."""
        px, py, pz = (np.ascontiguousarray(a, dtype=np.float64) for a in pts)
        n = len(px)
        nx, ny, nz = (np.empty(n, np.float64) for _ in range(3))
        valid = np.empty(n, np.uint8)
        r = self._L.orc_normal(self.h, pv, _p(px), _p(py), _p(pz), n, _p(nx), _p(ny), _p(nz), _p(valid), nthreads)
        if r != 0:
            raise RuntimeError("no Normal for this shape family")
        return nx, ny, nz, valid
