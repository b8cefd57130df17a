"""vecgeom_amd — Python host binding of the gfx950 geometry library (libvgm.so).

The library itself is C++/HIP behind the C ABI declared in include/vgm/vgm.h;
this module is a thin ctypes layer over it for tests and benchmarks.  There is
no CPU fallback: if libvgm.so has not been built, importing `lib()` raises.

Inputs are SOA3D batches (three coordinate arrays).  Host numpy arrays use the
synchronous host-pointer entry points; torch tensors on the GPU are passed as
device pointers.
"""
from __future__ import annotations

import ctypes
import weakref
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvgm.so")

HOST_PTRS = 0
DEVICE_PTRS = 1
ASYNC = 2
MIXED = 4
TIMING = 8
LEARNED = 16
AUDIT = 32
DISTANCE_FIRST = 64
TUNED = 128
CULL = 256
GRAPH = 512
STEP_KERNELS = 1024
FAMILY_MODES = ("fp64", "mixed", "distance-first", "learned", "culled")
MLP_PARAMS = 56321
SHAPE_TYPES = ("box", "tube", "cone", "trd", "trapezoid", "polycone", "polyhedron", "tessellated", "orb", "cut_tube", "boolean",
               "sphere", "torus")

# vgm_query_noalloc method codes (VGM_Q_*) and output types
NOALLOC_METHODS = {
    "Contains": (0, np.uint8), "Inside": (1, np.int32), "DistanceToIn": (2, np.float64),
    "DistanceToOut": (3, np.float64), "PlacedDistanceToOut": (4, np.float64), "SafetyToIn": (5, np.float64),
    "SafetyToOut": (6, np.float64),
}
NOALLOC_SCRATCH_BYTES = 32 << 20

# Inside_t (VecGeom/base/Global.h:71-88)
kInside, kSurface, kOutside = 1, 2, 3
kInfLength = np.finfo(np.float64).max

_lib = None


class VgmError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """Load libvgm.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VgmError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        # torch ships its own libamdhip64 with the same soname as the system
        # one.  Loading torch first makes libvgm.so bind to that single HIP
        # runtime, so device pointers and streams are shared with torch;
        # loading libvgm.so first would put two runtimes in the process and
        # torch would then find no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_SZ = ctypes.c_size_t
_U = ctypes.c_uint


def _declare(L):
    L.vgm_last_error.restype = ctypes.c_char_p
    L.vgm_last_error.argtypes = []
    sig = {
        "vgm_ctx_create": [_I, _P, ctypes.POINTER(_P)],
        "vgm_ctx_destroy": [_P],
        "vgm_ctx_synchronize": [_P],
        "vgm_box_create": [_P, _D, _D, _D, ctypes.POINTER(_I)],
        "vgm_orb_create": [_P, _D, ctypes.POINTER(_I)],
        "vgm_sphere_create": [_P, _D, _D, _D, _D, _D, _D, ctypes.POINTER(_I)],
        "vgm_torus_create": [_P, _D, _D, _D, _D, _D, ctypes.POINTER(_I)],
        "vgm_cut_tube_create": [_P] + [_D] * 11 + [ctypes.POINTER(_I)],
        "vgm_boolean_create": [_P, _I, _I, _P, _I, _P, ctypes.POINTER(_I)],
        "vgm_tube_create": [_P, _D, _D, _D, _D, _D, ctypes.POINTER(_I)],
        "vgm_cone_create": [_P, _D, _D, _D, _D, _D, _D, _D, ctypes.POINTER(_I)],
        "vgm_trd_create": [_P, _D, _D, _D, _D, _D, ctypes.POINTER(_I)],
        "vgm_polycone_create": [_P, _D, _D, _I, _P, _P, _P, ctypes.POINTER(_I)],
        "vgm_polyhedron_create": [_P, _D, _D, _I, _I, _P, _P, _P, ctypes.POINTER(_I)],
        "vgm_tessellated_create": [_P, _I, _P, _P, ctypes.POINTER(_I)],
        "vgm_trapezoid_create": [_P] + [_D] * 11 + [ctypes.POINTER(_I)],
        "vgm_shape_extent": [_P, _I, _P, _P],
        "vgm_lv_create": [_P, _I, ctypes.POINTER(_I)],
        "vgm_pv_create": [_P, _I, _P, ctypes.POINTER(_I)],
        "vgm_lv_add_daughter": [_P, _I, _I],
        "vgm_contains": [_P, _I, _P, _P, _P, _SZ, _P, _U],
        "vgm_inside": [_P, _I, _P, _P, _P, _SZ, _P, _U],
        "vgm_safety_to_in": [_P, _I, _P, _P, _P, _SZ, _P, _U],
        "vgm_safety_to_out": [_P, _I, _P, _P, _P, _SZ, _P, _U],
        "vgm_normal": [_P, _I, _P, _P, _P, _SZ, _P, _P, _P, _P, _U],
        "vgm_distance_to_in": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _U],
        "vgm_distance_to_out": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _U],
        "vgm_placed_distance_to_out": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _U],
        "vgm_query_noalloc": [_P, _I, _I, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _U],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _I
    L.vgm_ctx_stream.argtypes = [_P]
    L.vgm_ctx_stream.restype = _P
    nsig = {
        "vgm_nav_close": [_P, _I, ctypes.POINTER(_I)],
        "vgm_nav_node": [_P, ctypes.c_uint32, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)],
        "vgm_nav_locate": [_P, _P, _P, _P, _SZ, _P, _U],
        "vgm_nav_step": [_P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _P, _P, _P, _U],
        "vgm_nav_compute_step": [_P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _P, _P, _P, _U],
        "vgm_nav_compute_step_and_next_volume": [_P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _P, _P, _P, _U],
        "vgm_nav_step_safety": [_P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _P, _P, _P, _P, _P, _U],
        "vgm_nav_path_to_state": [_P, _P, _I, ctypes.POINTER(ctypes.c_uint32)],
        "vgm_nav_state_to_path": [_P, ctypes.c_uint32, _P, _I, ctypes.POINTER(_I)],
        "vgm_nav_transport": [_P, _P, _P, _P, _P, _P, _P, _SZ, _P, _P, _P, _I, _P, _U],
        "vgm_nav_transport_limited": [_P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _P, _P, _I, _I, _P, _P, _U],
        "vgm_nav_approach": [_P, _P, _P, _P, _P, _P, _P, _SZ, _P, _P, _P, _U],
        "vgm_nav_safety": [_P, _P, _P, _P, _SZ, _P, _P, _U],
        "vgm_nav_level_locate": [_P] * 7 + [_SZ] + [_P] * 6 + [_U],
    }
    for name, args in nsig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _I
    bsig = {
        "vgm_batch_create": [_P, _I, _P, _P, _P, ctypes.POINTER(_P)],
        "vgm_batch_destroy": [_P],
        "vgm_batch_classify": [_P, _P, _P, _P, _P],
        "vgm_batch_transport": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, ctypes.c_uint64, _U,
                                ctypes.POINTER(ctypes.c_ulonglong)],
        "vgm_batch_set_launch": [_P, _I, _I, _I],
        "vgm_batch_get_launch": [_P, _P],
        "vgm_batch_set_streams": [_P, _I],
        "vgm_batch_get_streams": [_P, _P],
        "vgm_batch_family_ms": [_P, _P],
        "vgm_batch_learned_stats": [_P, _P],
        "vgm_batch_set_fallback": [_P, _I],
        "vgm_batch_get_modes": [_P, _P],
        "vgm_batch_set_modes": [_P, _P],
        "vgm_batch_get_weights": [_P, _P],
        "vgm_mlp_train": [_P, _I, _P, _P],
        "vgm_mlp_predict": [_P, _I, _P, _P, _P, _SZ, _P, _U],
        "vgm_mlp_bound": [_P, _I, _P, _P, _P, _SZ, _P, _U],
        "vgm_mlp_get": [_P, _I, _P, _P],
        "vgm_mlp_set": [_P, _I, _P, _P],
        "vgm_mlp_gradient": [_P, _P, _P, _P, _SZ, ctypes.c_float, _P, _P],
        "vgm_batch_autotune": [_P, _P, _P, _P, _P, _P, _P, _P, _I, _U, _P],
    }
    for name, args in bsig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _I
    L.vgm_mlp_config_default.argtypes = [_P]
    L.vgm_mlp_config_default.restype = None


class MlpConfig(ctypes.Structure):
    """vgm_mlp_config (include/vgm/vgm.h)."""
    _fields_ = [("samples", ctypes.c_int64), ("validation", ctypes.c_int64), ("batch", ctypes.c_int),
                ("epochs", ctypes.c_int), ("lr", ctypes.c_float), ("over_weight", ctypes.c_float),
                ("margin_factor", ctypes.c_double), ("margin_abs", ctypes.c_double), ("domain", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


class MlpReport(ctypes.Structure):
    """vgm_mlp_report (include/vgm/vgm.h)."""
    _fields_ = [("margin", ctypes.c_double), ("max_over", ctypes.c_double), ("mean_abs_err", ctypes.c_double),
                ("rms_err", ctypes.c_double), ("final_loss", ctypes.c_double), ("train_ms", ctypes.c_double),
                ("gflop", ctypes.c_double), ("scale", ctypes.c_double), ("steps", ctypes.c_int64),
                ("samples", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def check(rc: int):
    if rc != 0:
        raise VgmError(f"vgm error {rc}: {lib().vgm_last_error().decode()}")


def _ptr(a):
    """Raw address of a numpy array or a torch tensor (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def _is_device(a) -> bool:
    return not isinstance(a, np.ndarray) and getattr(a, "is_cuda", False)


class SOA3D:
    """Three coordinate arrays of equal length (SOA3D<double>)."""

    def __init__(self, x, y, z):
        self.x, self.y, self.z = x, y, z
        self.n = len(x)
        if not (len(y) == self.n and len(z) == self.n):
            raise ValueError("SOA3D components differ in length")

    @classmethod
    def from_array(cls, a):
        """From an (n,3) or (3,n) float64 numpy array."""
        a = np.asarray(a, dtype=np.float64)
        if a.ndim != 2:
            raise ValueError("expected a 2-D array")
        if a.shape[0] == 3 and a.shape[1] != 3:
            return cls(*(np.ascontiguousarray(a[k]) for k in range(3)))
        return cls(*(np.ascontiguousarray(a[:, k]) for k in range(3)))


class Context:
    """One device, one stream, one geometry (vgm_ctx)."""

    def __init__(self, device: int = 0, stream: Optional[int] = None):
        self._L = lib()
        h = _P()
        check(self._L.vgm_ctx_create(device, stream, ctypes.byref(h)))
        self.handle = h
        self._batches = weakref.WeakSet()  # batches live on the context: closed before it

    def close(self):
        if self.handle:
            for b in list(self._batches):
                b.close()
            self._L.vgm_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(self._L.vgm_ctx_synchronize(self.handle))

    @property
    def stream(self) -> int:
        return self._L.vgm_ctx_stream(self.handle)

    # ---- shapes -------------------------------------------------------------
    def _new(self, fn, *args) -> int:
        out = _I()
        check(fn(self.handle, *args, ctypes.byref(out)))
        return out.value

    def box(self, dx, dy, dz) -> int:
        return self._new(self._L.vgm_box_create, dx, dy, dz)

    def orb(self, r) -> int:
        return self._new(self._L.vgm_orb_create, r)

    def sphere(self, rmin, rmax, sphi=0.0, dphi=2 * np.pi, stheta=0.0, dtheta=np.pi) -> int:
        """UnplacedSphere(rmin, rmax, sphi, dphi, stheta, dtheta)."""
        return self._new(self._L.vgm_sphere_create, rmin, rmax, sphi, dphi, stheta, dtheta)

    def torus(self, rmin, rmax, rtor, sphi=0.0, dphi=2 * np.pi) -> int:
        """UnplacedTorus2(rmin, rmax, rtor, sphi, dphi)."""
        return self._new(self._L.vgm_torus_create, rmin, rmax, rtor, sphi, dphi)

    def boolean(self, op, left, right, left_tr=None, right_tr=None) -> int:
        """UnplacedBooleanVolume: op "union", "intersection" or "subtraction";
        left and right are shape ids or (kind, args) pairs built here first."""
        ops = {"union": 0, "intersection": 1, "subtraction": 2}
        ids = [getattr(self, s[0])(*s[1]) if isinstance(s, tuple) else s for s in (left, right)]
        trs = [None if t is None else np.ascontiguousarray(t, dtype=np.float64) for t in (left_tr, right_tr)]
        return self._new(self._L.vgm_boolean_create, ops[op], ids[0], _ptr(trs[0]), ids[1], _ptr(trs[1]))

    def cut_tube(self, rmin, rmax, z, sphi, dphi, bottom, top) -> int:
        """UnplacedCutTube(rmin, rmax, z, sphi, dphi, bottomNormal, topNormal)."""
        return self._new(self._L.vgm_cut_tube_create, rmin, rmax, z, sphi, dphi, *bottom, *top)

    def tube(self, rmin, rmax, dz, sphi, dphi) -> int:
        return self._new(self._L.vgm_tube_create, rmin, rmax, dz, sphi, dphi)

    def cone(self, rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi) -> int:
        return self._new(self._L.vgm_cone_create, rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi)

    def trd(self, x1, x2, y1, y2, z) -> int:
        return self._new(self._L.vgm_trd_create, x1, x2, y1, y2, z)

    def trapezoid(self, dz, theta, phi, dy1, dx1, dx2, alpha1, dy2, dx3, dx4, alpha2) -> int:
        return self._new(self._L.vgm_trapezoid_create, dz, theta, phi, dy1, dx1, dx2, alpha1, dy2, dx3, dx4, alpha2)

    def polycone(self, sphi, dphi, z, rmin, rmax) -> int:
        z, rmin, rmax = (np.ascontiguousarray(a, dtype=np.float64) for a in (z, rmin, rmax))
        if not (len(z) == len(rmin) == len(rmax)):
            raise ValueError("z, rmin and rmax differ in length")
        return self._new(self._L.vgm_polycone_create, sphi, dphi, len(z), _ptr(z), _ptr(rmin), _ptr(rmax))

    def polyhedron(self, sphi, dphi, nsides, z, rmin, rmax) -> int:
        z, rmin, rmax = (np.ascontiguousarray(a, dtype=np.float64) for a in (z, rmin, rmax))
        return self._new(self._L.vgm_polyhedron_create, sphi, dphi, nsides, len(z), _ptr(z), _ptr(rmin),
                         _ptr(rmax))

    def tessellated(self, triangles, test_dir=None) -> int:
        """Triangles (n, 3, 3), anti-clockwise seen from outside."""
        tri = np.ascontiguousarray(triangles, dtype=np.float64).reshape(-1, 9)
        td = None if test_dir is None else np.ascontiguousarray(test_dir, dtype=np.float64)
        return self._new(self._L.vgm_tessellated_create, len(tri), _ptr(tri), _ptr(td))

    def extent(self, shape: int):
        mn = np.zeros(3)
        mx = np.zeros(3)
        check(self._L.vgm_shape_extent(self.handle, shape, mn.ctypes.data, mx.ctypes.data))
        return mn, mx

    # ---- learned safety -----------------------------------------------------
    def mlp_train(self, shape: int, **cfg) -> dict:
        """Train the shape's safety network on the GPU (vgm_mlp_train); keyword
        arguments override vgm_mlp_config fields.  Returns the report."""
        c = MlpConfig()
        self._L.vgm_mlp_config_default(ctypes.byref(c))
        for k, v in cfg.items():
            if not hasattr(c, k):
                raise ValueError(f"unknown mlp config field {k}")
            setattr(c, k, v)
        r = MlpReport()
        check(self._L.vgm_mlp_train(self.handle, shape, ctypes.byref(c), ctypes.byref(r)))
        return r.as_dict()

    def _mlp_query(self, fn, shape, x, y, z, out):
        dev = _is_device(x)
        if out is None:
            out = self._alloc(x, len(x), np.float64)
        check(fn(self.handle, shape, _ptr(x), _ptr(y), _ptr(z), len(x), _ptr(out), DEVICE_PTRS if dev else HOST_PTRS))
        return out

    def mlp_predict(self, shape: int, x, y, z, out=None):
        """Network prediction (mm) at points in the shape's frame."""
        return self._mlp_query(self._L.vgm_mlp_predict, shape, x, y, z, out)

    def mlp_bound(self, shape: int, x, y, z, out=None):
        """Served bound max(0, prediction - margin)."""
        return self._mlp_query(self._L.vgm_mlp_bound, shape, x, y, z, out)

    def mlp_get(self, shape: int):
        """(params[56321] float32, geom[8]: centre, half-widths, scale, margin)."""
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        g = np.zeros(8)
        check(self._L.vgm_mlp_get(self.handle, shape, p.ctypes.data, g.ctypes.data))
        return p, g

    def mlp_set(self, shape: int, params, geom):
        p = np.ascontiguousarray(params, dtype=np.float32)
        g = np.ascontiguousarray(geom, dtype=np.float64)
        if p.size != MLP_PARAMS or g.size != 8:
            raise ValueError("expected 56321 parameters and 8 geometry values")
        check(self._L.vgm_mlp_set(self.handle, shape, p.ctypes.data, g.ctypes.data))

    def mlp_gradient(self, params, u, target, over_weight=4.0):
        """Loss gradient of one batch from the training kernels (no update)."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        uu = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        check(self._L.vgm_mlp_gradient(self.handle, p.ctypes.data, uu.ctypes.data, t.ctypes.data, len(t),
                                       float(over_weight), g.ctypes.data, loss.ctypes.data))
        return g, float(loss[0])

    # ---- volumes ------------------------------------------------------------
    def logical_volume(self, shape: int) -> int:
        return self._new(self._L.vgm_lv_create, shape)

    def place(self, lv: int, transform: Optional[Sequence[float]] = None) -> int:
        tr = None if transform is None else np.ascontiguousarray(transform, dtype=np.float64)
        if tr is not None and tr.size != 12:
            raise ValueError("transform is translation[3] + rotation[9]")
        out = _I()
        check(self._L.vgm_pv_create(self.handle, lv, _ptr(tr), ctypes.byref(out)))
        return out.value

    def add_daughter(self, mother_lv: int, pv: int):
        check(self._L.vgm_lv_add_daughter(self.handle, mother_lv, pv))

    # ---- SOA3D queries ------------------------------------------------------
    def _flags(self, arr, mixed: bool, asynchronous: bool = False) -> int:
        f = DEVICE_PTRS if _is_device(arr) else HOST_PTRS
        if f == DEVICE_PTRS and asynchronous:
            f |= ASYNC
        if mixed:
            f |= MIXED
        return f

    def _alloc(self, like, n, dtype):
        if _is_device(like):
            import torch
            tdt = {np.uint8: torch.uint8, np.int32: torch.int32, np.float64: torch.float64}[dtype]
            return torch.empty(n, dtype=tdt, device=like.device)
        return np.empty(n, dtype=dtype)

    def _q1(self, fn, pv, pts: SOA3D, dtype, out, mixed, asynchronous=False):
        if out is None:
            out = self._alloc(pts.x, pts.n, dtype)
        check(fn(self.handle, pv, _ptr(pts.x), _ptr(pts.y), _ptr(pts.z), pts.n, _ptr(out),
                 self._flags(pts.x, mixed, asynchronous)))
        return out

    def _q2(self, fn, pv, pts: SOA3D, dirs: SOA3D, step_max, out, mixed, asynchronous=False):
        if dirs.n != pts.n:
            raise ValueError("points and directions differ in length")
        if out is None:
            out = self._alloc(pts.x, pts.n, np.float64)
        check(fn(self.handle, pv, _ptr(pts.x), _ptr(pts.y), _ptr(pts.z), _ptr(dirs.x), _ptr(dirs.y),
                 _ptr(dirs.z), _ptr(step_max), pts.n, _ptr(out), self._flags(pts.x, mixed, asynchronous)))
        return out

    def contains(self, pv, pts, out=None, mixed=False, asynchronous=False):
        return self._q1(self._L.vgm_contains, pv, pts, np.uint8, out, mixed, asynchronous)

    def inside(self, pv, pts, out=None, mixed=False, asynchronous=False):
        return self._q1(self._L.vgm_inside, pv, pts, np.int32, out, mixed, asynchronous)

    def safety_to_in(self, pv, pts, out=None, mixed=False, asynchronous=False):
        return self._q1(self._L.vgm_safety_to_in, pv, pts, np.float64, out, mixed, asynchronous)

    def safety_to_out(self, pv, pts, out=None, mixed=False, asynchronous=False):
        return self._q1(self._L.vgm_safety_to_out, pv, pts, np.float64, out, mixed, asynchronous)

    def normal(self, pv, pts: SOA3D, asynchronous=False):
        """VPlacedVolume::Normal of master-frame points: (nx, ny, nz, valid)."""
        nx, ny, nz = (self._alloc(pts.x, pts.n, np.float64) for _ in range(3))
        valid = self._alloc(pts.x, pts.n, np.uint8)
        check(self._L.vgm_normal(self.handle, pv, _ptr(pts.x), _ptr(pts.y), _ptr(pts.z), pts.n, _ptr(nx), _ptr(ny),
                                 _ptr(nz), _ptr(valid), self._flags(pts.x, False, asynchronous)))
        return nx, ny, nz, valid

    def distance_to_in(self, pv, pts, dirs, step_max=None, out=None, mixed=False, asynchronous=False):
        return self._q2(self._L.vgm_distance_to_in, pv, pts, dirs, step_max, out, mixed, asynchronous)

    def distance_to_out(self, pv, pts, dirs, step_max=None, out=None, mixed=False, asynchronous=False):
        return self._q2(self._L.vgm_distance_to_out, pv, pts, dirs, step_max, out, mixed, asynchronous)

    def placed_distance_to_out(self, pv, pts, dirs, step_max=None, out=None, mixed=False, asynchronous=False):
        return self._q2(self._L.vgm_placed_distance_to_out, pv, pts, dirs, step_max, out, mixed, asynchronous)

    def query_noalloc(self, method: str, pv, pts: SOA3D, dirs: Optional[SOA3D] = None, step_max=None, out=None,
                      mixed=False):
        """vgm_query_noalloc on host arrays: the query named by `method` (a
        key of NOALLOC_METHODS) staged through the library's fixed device
        array, with no device allocation."""
        code, dtype = NOALLOC_METHODS[method]
        if out is None:
            out = np.empty(pts.n, dtype=dtype)
        dx, dy, dz = (None, None, None) if dirs is None else (dirs.x, dirs.y, dirs.z)
        check(self._L.vgm_query_noalloc(
            self.handle, pv, code, _ptr(pts.x), _ptr(pts.y), _ptr(pts.z), _ptr(dx), _ptr(dy), _ptr(dz),
            _ptr(step_max), pts.n, _ptr(out), MIXED if mixed else HOST_PTRS))
        return out


class Navigation:
    """Navigation on a closed geometry (vgm_nav_*): states are uint32 node
    indices of the placement tree (1 = world, 0 = outside)."""

    def __init__(self, ctx: "Context", world_pv: int):
        self.ctx = ctx
        self._L = ctx._L
        n = _I()
        check(self._L.vgm_nav_close(ctx.handle, world_pv, ctypes.byref(n)))
        self.nnodes = n.value

    def node(self, state: int):
        pv, lvl, mo = _I(), _I(), _I()
        check(self._L.vgm_nav_node(self.ctx.handle, state, ctypes.byref(pv), ctypes.byref(lvl), ctypes.byref(mo)))
        return pv.value, lvl.value, mo.value

    def _fl(self, a, asynchronous=False):
        return (DEVICE_PTRS if _is_device(a) else HOST_PTRS) | (ASYNC if asynchronous else 0)

    def locate(self, x, y, z, state):
        check(self._L.vgm_nav_locate(self.ctx.handle, _ptr(x), _ptr(y), _ptr(z), len(x), _ptr(state), self._fl(x)))

    def step(self, x, y, z, dx, dy, dz, state, last, boundary, step, step_limit=None):
        check(self._L.vgm_nav_step(self.ctx.handle, _ptr(x), _ptr(y), _ptr(z), _ptr(dx), _ptr(dy), _ptr(dz),
                                   _ptr(step_limit), len(x), _ptr(state), _ptr(last), _ptr(boundary), _ptr(step),
                                   self._fl(x)))

    def compute_step(self, x, y, z, dx, dy, dz, state, last, boundary, step, step_limit=None):
        """VNavigator::ComputeStep: the step and PrepareOutState's state, no relocation."""
        check(self._L.vgm_nav_compute_step(self.ctx.handle, _ptr(x), _ptr(y), _ptr(z), _ptr(dx), _ptr(dy),
                                           _ptr(dz), _ptr(step_limit), len(x), _ptr(state), _ptr(last),
                                           _ptr(boundary), _ptr(step), self._fl(x)))

    def compute_step_and_next_volume(self, x, y, z, dx, dy, dz, state, last, boundary, step, step_limit=None):
        """vgm_nav_compute_step_and_next_volume (the same call as step())."""
        check(self._L.vgm_nav_compute_step_and_next_volume(
            self.ctx.handle, _ptr(x), _ptr(y), _ptr(z), _ptr(dx), _ptr(dy), _ptr(dz), _ptr(step_limit), len(x),
            _ptr(state), _ptr(last), _ptr(boundary), _ptr(step), self._fl(x)))

    def step_safety(self, x, y, z, dx, dy, dz, state, last, boundary, step, safety, calcsafety=None, step_limit=None):
        """ComputeStepsAndSafetiesAndPropagatedStates: safety[i] before the step where calcsafety[i] (default all)."""
        check(self._L.vgm_nav_step_safety(self.ctx.handle, _ptr(x), _ptr(y), _ptr(z), _ptr(dx), _ptr(dy), _ptr(dz),
                                          _ptr(step_limit), len(x), _ptr(state), _ptr(last), _ptr(boundary),
                                          _ptr(step), _ptr(calcsafety), _ptr(safety), self._fl(x)))

    def path_to_state(self, pvs) -> int:
        """NavStatePath (placed volumes from the world down) -> state index."""
        a = np.ascontiguousarray(pvs, dtype=np.int32)
        out = ctypes.c_uint32(0)
        check(self._L.vgm_nav_path_to_state(self.ctx.handle, _ptr(a), len(a), ctypes.byref(out)))
        return int(out.value)

    def state_to_path(self, state: int):
        """State index -> placed volumes from the world down to the top volume."""
        buf = np.zeros(64, dtype=np.int32)
        d = _I()
        check(self._L.vgm_nav_state_to_path(self.ctx.handle, int(state), _ptr(buf), len(buf), ctypes.byref(d)))
        return [int(v) for v in buf[:d.value]]

    def transport(self, x, y, z, dx, dy, dz, state, last, boundary, nsteps, steps, asynchronous=False,
                  graph=False, step_kernels=False):
        """nsteps straight-line steps per track: one persistent kernel (default),
        a captured HIP graph of one-step kernels (graph=True) or those kernels
        launched one by one (step_kernels=True); device arrays for the last two."""
        fl = self._fl(x, asynchronous) | (GRAPH if graph else 0) | (STEP_KERNELS if step_kernels else 0)
        check(self._L.vgm_nav_transport(self.ctx.handle, _ptr(x), _ptr(y), _ptr(z), _ptr(dx), _ptr(dy), _ptr(dz),
                                        len(x), _ptr(state), _ptr(last), _ptr(boundary), nsteps, _ptr(steps), fl))

    def transport_limited(self, x, y, z, dx, dy, dz, step_limit, state, last, boundary, nsteps, steps, geo_steps,
                          reuse_safety=False):
        """nsteps of ComputeStepAndSafetyAndPropagatedState under the physics step step_limit[i]."""
        check(self._L.vgm_nav_transport_limited(self.ctx.handle, _ptr(x), _ptr(y), _ptr(z), _ptr(dx), _ptr(dy),
                                                _ptr(dz), _ptr(step_limit), len(x), _ptr(state), _ptr(last),
                                                _ptr(boundary), nsteps, 1 if reuse_safety else 0, _ptr(steps),
                                                _ptr(geo_steps), self._fl(x)))

    def approach(self, x, y, z, dx, dy, dz, state, last, step):
        """BVH::ApproachNextDaughter: step shortened to the nearest daughter bounding box."""
        check(self._L.vgm_nav_approach(self.ctx.handle, _ptr(x), _ptr(y), _ptr(z), _ptr(dx), _ptr(dy), _ptr(dz),
                                       len(x), _ptr(state), _ptr(last), _ptr(step), self._fl(x)))

    def level_locate(self, x, y, z, state, child, qx, qy, qz, dirs=None, excl=None):
        """LevelLocate in the volume of each state, points in that volume's frame
        (VLevelLocator::LevelLocateExclVol; directional with dirs = (dx, dy, dz)):
        child = the daughter's state or 0, (qx, qy, qz) = the point in its frame."""
        d = (None, None, None) if dirs is None else dirs
        check(self._L.vgm_nav_level_locate(self.ctx.handle, _ptr(x), _ptr(y), _ptr(z), _ptr(d[0]), _ptr(d[1]),
                                           _ptr(d[2]), len(x), _ptr(state), _ptr(excl), _ptr(child), _ptr(qx),
                                           _ptr(qy), _ptr(qz), self._fl(x)))

    def safety(self, x, y, z, state, out):
        check(self._L.vgm_nav_safety(self.ctx.handle, _ptr(x), _ptr(y), _ptr(z), len(x), _ptr(state), _ptr(out),
                                     self._fl(x)))


class Batch:
    """Batched heterogeneous track-step transport over many solids (vgm_batch).

    Track state lives in device tensors (float64 x, y, z, dx, dy, dz; int8 side;
    int32 crossings), solid i owning tracks [offsets[i], offsets[i+1]).
    """

    def __init__(self, ctx: Context, shapes, offsets, mean_step):
        self.ctx = ctx
        self._L = ctx._L
        self._shapes = np.ascontiguousarray(shapes, dtype=np.int32)
        self._offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self._steps = np.ascontiguousarray(mean_step, dtype=np.float64)
        h = _P()
        check(self._L.vgm_batch_create(ctx.handle, len(self._shapes), _ptr(self._shapes), _ptr(self._offsets),
                                       _ptr(self._steps), ctypes.byref(h)))
        self.handle = h
        ctx._batches.add(self)

    def close(self):
        if self.handle:
            self._L.vgm_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def classify(self, x, y, z, side):
        check(self._L.vgm_batch_classify(self.handle, _ptr(x), _ptr(y), _ptr(z), _ptr(side)))

    def transport(self, x, y, z, dx, dy, dz, side, crossings, nsteps, step0=0, seed=1, mixed=False,
                  asynchronous=False, count_fp64=False, timing=False, learned=False, audit=False,
                  distance_first=False, tuned=False, cull=False, graph=False):
        flags = (MIXED if mixed else 0) | (ASYNC if asynchronous else 0) | (TIMING if timing else 0)
        flags |= GRAPH if graph else 0
        flags |= (LEARNED if learned else 0) | (AUDIT if audit else 0) | (DISTANCE_FIRST if distance_first else 0)
        flags |= (TUNED if tuned else 0) | (CULL if cull else 0)
        cnt = ctypes.c_ulonglong(0)
        check(self._L.vgm_batch_transport(self.handle, _ptr(x), _ptr(y), _ptr(z), _ptr(dx), _ptr(dy), _ptr(dz),
                                          _ptr(side), _ptr(crossings), nsteps, step0, seed, flags,
                                          ctypes.byref(cnt) if count_fp64 else None))
        return cnt.value

    def set_launch(self, block, grid, tile_tracks):
        check(self._L.vgm_batch_set_launch(self.handle, block, grid, tile_tracks))

    def set_streams(self, streams: int):
        """1: families back to back on the context stream; 4: concurrently on the family streams."""
        check(self._L.vgm_batch_set_streams(self.handle, int(streams)))

    def streams(self) -> int:
        out = ctypes.c_int(0)
        check(self._L.vgm_batch_get_streams(self.handle, ctypes.byref(out)))
        return int(out.value)

    def learned_stats(self):
        """Counters of the last learned transport."""
        out = np.zeros(4, dtype=np.uint64)
        check(self._L.vgm_batch_learned_stats(self.handle, out.ctypes.data))
        return {"queries": int(out[0]), "served": int(out[1]), "fallbacks": int(out[2]), "violations": int(out[3])}

    def set_fallback(self, pct: int):
        check(self._L.vgm_batch_set_fallback(self.handle, int(pct)))

    def family_ms(self):
        """{family: kernel ms} of the last transport run with timing=True."""
        out = np.zeros(len(SHAPE_TYPES), dtype=np.float32)
        check(self._L.vgm_batch_family_ms(self.handle, out.ctypes.data))
        return {SHAPE_TYPES[t]: float(v) for t, v in enumerate(out) if v >= 0}

    def launch(self):
        out = np.zeros(3, dtype=np.int32)
        check(self._L.vgm_batch_get_launch(self.handle, out.ctypes.data))
        return tuple(int(v) for v in out)

    def modes(self):
        """{family: mode} chosen by the last autotune."""
        out = np.zeros(len(SHAPE_TYPES), dtype=np.int32)
        check(self._L.vgm_batch_get_modes(self.handle, out.ctypes.data))
        return {SHAPE_TYPES[t]: FAMILY_MODES[m] for t, m in enumerate(out) if m >= 0}

    def weights(self):
        """{family: weight}: standalone kernel ms after an autotune."""
        out = np.zeros(len(SHAPE_TYPES))
        check(self._L.vgm_batch_get_weights(self.handle, out.ctypes.data))
        return {SHAPE_TYPES[t]: float(v) for t, v in enumerate(out) if v > 0}

    def set_modes(self, modes):
        """modes: {family: mode name}; families not given are left untuned."""
        out = np.full(len(SHAPE_TYPES), -1, dtype=np.int32)
        for f, m in modes.items():
            out[SHAPE_TYPES.index(f)] = FAMILY_MODES.index(m)
        check(self._L.vgm_batch_set_modes(self.handle, out.ctypes.data))

    def autotune(self, x, y, z, dx, dy, dz, side, nsteps, mixed=False, learned=False, distance_first=False,
                 cull=False):
        out = np.zeros(3, dtype=np.int32)
        flags = (MIXED if mixed else 0) | (LEARNED if learned else 0) | (DISTANCE_FIRST if distance_first else 0)
        flags |= CULL if cull else 0
        check(self._L.vgm_batch_autotune(self.handle, _ptr(x), _ptr(y), _ptr(z), _ptr(dx), _ptr(dy), _ptr(dz),
                                         _ptr(side), nsteps, flags, out.ctypes.data))
        return tuple(int(v) for v in out)
