"""Known-answer vectors transcribed (as data) from the reference's unit tests.

Each entry: (method, point, direction-or-None, expectation) where expectation is
  ("eq", value)        ApproxEqual(result, value)  (test/unit_tests/ApproxEqual.h)
  ("inside", code)     Inside() == code (1 inside, 2 surface, 3 outside)
  ("le", value)        result <= value
  ("ge", value)        result >= value
  ("inf",)             result >= kInfLength (reference: `Dist >= kInfLength`)
Shapes are given by their constructor arguments.  Directions are normalised
with Vector3D::Unit's rule (divide by sqrt(mag2 + DBL_MIN)) where the test does
so; axis and diagonal vectors are used as written.
"""
import math

import numpy as np

S2 = 1 / math.sqrt(2.0)
DBL_MIN = np.finfo(np.float64).tiny
KINF = np.finfo(np.float64).max
kHalfTolerance = 0.5e-9


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return tuple(float(c) for c in v / np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + DBL_MIN))


def approx_equal(x, y, tol=1e-6):
    """ApproxEqual<double> (test/unit_tests/ApproxEqual.h)."""
    if x == y:
        return True
    if x * y == 0.0:
        return abs(x - y) < tol
    if abs(x) > 1e100 or abs(y) > 1e100:
        return (x * y > 0) and abs(x) > 1e100 and abs(y) > 1e100
    return abs(x - y) / (abs(x) + abs(y)) < tol


def check(result, expect):
    kind = expect[0]
    if kind == "eq":
        return approx_equal(float(result), float(expect[1]))
    if kind == "inside":
        return int(result) == expect[1]
    if kind == "le":
        return float(result) <= expect[1]
    if kind == "ge":
        return float(result) >= expect[1]
    if kind == "inf":
        return float(result) >= KINF
    raise ValueError(kind)


O = (0.0, 0.0, 0.0)
VX, VY, VZ = (1.0, 0, 0), (0, 1.0, 0), (0, 0, 1.0)
VMX, VMY, VMZ = (-1.0, 0, 0), (0, -1.0, 0), (0, 0, -1.0)
VXY, VMXY = (S2, S2, 0.0), (-S2, S2, 0.0)
VXMZ = (S2, 0.0, -S2)

# ---- Box: test/unit_tests/TestBox.cpp ---------------------------------------
# b1 = (20,30,40), b2 = (10,10,10), box3 = BABAR box, b4 = (5,5,5), bx = (1,200,200)
BOX_B1 = (20.0, 30.0, 40.0)
BOX_B2 = (10.0, 10.0, 10.0)
BOX_BABAR = (0.14999999999999999, 24.707000000000001, 22.699999999999999)

_vecgeom431_in = unit(np.array([0.0, 1.0, 1.0]) - 4.0e-06 * np.array([1.0, 0, 0]))
_vecgeom431_in_m = unit(np.array([0.0, 1.0, 1.0]) + 4.0e-06 * np.array([1.0, 0, 0]))

BOX_KATS = [
    # TestBox.cpp:196-226 — DistanceToOut from the origin
    (BOX_B1, "DistanceToOut", O, VX, ("eq", 20)),
    (BOX_B1, "DistanceToOut", O, VMX, ("eq", 20)),
    (BOX_B1, "DistanceToOut", O, VY, ("eq", 30)),
    (BOX_B1, "DistanceToOut", O, VMY, ("eq", 30)),
    (BOX_B1, "DistanceToOut", O, VZ, ("eq", 40)),
    (BOX_B1, "DistanceToOut", O, VMZ, ("eq", 40)),
    (BOX_B1, "DistanceToOut", O, VXY, ("eq", math.sqrt(800.0))),
    (BOX_B1, "DistanceToOut", O, (0.0, 0.0, -1.0), ("eq", 40)),
    (BOX_B1, "DistanceToOut", O, (0.0, 0.0, 1.0), ("eq", 40)),
    (BOX_B1, "DistanceToOut", O, (0.0, 1.0, 0.0), ("eq", 30)),
    (BOX_B1, "DistanceToOut", O, (1.0, 0.0, 0.0), ("eq", 20)),
    (BOX_B1, "DistanceToOut", O, (-1.0, 0.0, 0.0), ("eq", 20)),
    # TestBox.cpp:228-248 — from the faces
    (BOX_B1, "DistanceToOut", (20.0, 0, 0), VX, ("eq", 0)),
    (BOX_B1, "DistanceToOut", (20.0, 0, 0), VMX, ("eq", 40)),
    (BOX_B1, "DistanceToOut", (-20.0, 0, 0), VMX, ("eq", 0)),
    (BOX_B1, "DistanceToOut", (0, 30.0, 0), VY, ("eq", 0)),
    (BOX_B1, "DistanceToOut", (0, -30.0, 0), VMY, ("eq", 0)),
    (BOX_B1, "DistanceToOut", (0, 0, 40.0), VZ, ("eq", 0)),
    (BOX_B1, "DistanceToOut", (0, 0, -40.0), VMZ, ("eq", 0)),
    # TestBox.cpp:250-266 — Inside
    (BOX_B1, "Inside", O, None, ("inside", 1)),
    (BOX_B1, "Inside", (0, 0, 100.0), None, ("inside", 3)),
    (BOX_B1, "Inside", (20.0, 0, 0), None, ("inside", 2)),
    (BOX_B1, "Inside", (0, 30.0, 0), None, ("inside", 2)),
    (BOX_B1, "Inside", (0, 0, 40.0), None, ("inside", 2)),
    (BOX_B2, "Inside", O, None, ("inside", 1)),
    (BOX_B2, "Inside", (0, 0, 100.0), None, ("inside", 3)),
    (BOX_B2, "Inside", (20.0, 0, 0), None, ("inside", 3)),
    (BOX_B2, "Inside", (0, 30.0, 0), None, ("inside", 3)),
    (BOX_B2, "Inside", (0, 0, 40.0), None, ("inside", 3)),
    (BOX_B2, "Inside", (10.0, 0, 0), None, ("inside", 2)),
    (BOX_B2, "Inside", (0, 10.0, 0), None, ("inside", 2)),
    (BOX_B2, "Inside", (0, 0, 10.0), None, ("inside", 2)),
    (BOX_B2, "Inside", (20.0, 20.0, 10.0), None, ("inside", 3)),
    (BOX_B2, "Inside", (100.0, 10.0, 30.0), None, ("inside", 3)),
    (BOX_B2, "Inside", (10.0, 20.0, 20.0), None, ("inside", 3)),
    # TestBox.cpp:268-275 — SafetyToOut
    (BOX_B1, "SafetyToOut", O, None, ("eq", 20)),
    (BOX_B1, "SafetyToOut", VX, None, ("eq", 19)),
    (BOX_B1, "SafetyToOut", VY, None, ("eq", 20)),
    (BOX_B1, "SafetyToOut", VZ, None, ("eq", 20)),
    # TestBox.cpp:307-318 — SafetyToIn
    (BOX_B1, "SafetyToIn", (100.0, 0, 0), None, ("eq", 80)),
    (BOX_B1, "SafetyToIn", (-100.0, 0, 0), None, ("eq", 80)),
    (BOX_B1, "SafetyToIn", (0, 100.0, 0), None, ("eq", 70)),
    (BOX_B1, "SafetyToIn", (0, -100.0, 0), None, ("eq", 70)),
    (BOX_B1, "SafetyToIn", (0, 0, 100.0), None, ("eq", 60)),
    (BOX_B1, "SafetyToIn", (0, 0, -100.0), None, ("eq", 60)),
    # TestBox.cpp:320-331 — DistanceToIn along axes
    (BOX_B1, "DistanceToIn", (100.0, 0, 0), VMX, ("eq", 80)),
    (BOX_B1, "DistanceToIn", (-100.0, 0, 0), VX, ("eq", 80)),
    (BOX_B1, "DistanceToIn", (0, 100.0, 0), VMY, ("eq", 70)),
    (BOX_B1, "DistanceToIn", (0, -100.0, 0), VY, ("eq", 70)),
    (BOX_B1, "DistanceToIn", (0, 0, 100.0), VMZ, ("eq", 60)),
    (BOX_B1, "DistanceToIn", (0, 0, -100.0), VZ, ("eq", 60)),
    # TestBox.cpp:333-350
    (BOX_B1, "DistanceToIn", (0, 0, -50.0), (0.0, 0.0, 1.0), ("eq", 10)),
    (BOX_B1, "DistanceToIn", (0, 0, 50.0), (0.0, 0.0, -1.0), ("eq", 10)),
    (BOX_B1, "DistanceToIn", (0, -40.0, 0), (0.0, 1.0, 0.0), ("eq", 10)),
    (BOX_B1, "DistanceToIn", (-30.0, 0, 0), (1.0, 0.0, 0.0), ("eq", 10)),
    (BOX_B1, "DistanceToIn", (30.0, 0, 0), (-1.0, 0.0, 0.0), ("eq", 10)),
    # TestBox.cpp:352-385 — misses and grazing
    (BOX_B1, "DistanceToIn", (100.0, 0, 0), VXY, ("inf",)),
    (BOX_B1, "DistanceToIn", (-100.0, 0, 0), VXY, ("inf",)),
    (BOX_B2, "DistanceToIn", (9.0, 0, 12.0), VXMZ, ("inf",)),
    (BOX_B2, "DistanceToIn", (12.0, 9.0, 0), VMXY, ("inf",)),
    (BOX_B2, "DistanceToIn", (12.0, 9.0, 0), VMX, ("eq", 2)),
    (BOX_B2, "DistanceToIn", (32.0, -11.0, 0), VMXY, ("inf",)),
    (BOX_B1, "DistanceToIn", (-25.0, -35.0, 0), VX, ("inf",)),
    (BOX_B1, "DistanceToIn", (-25.0, -35.0, 0), VY, ("inf",)),
    # TestBox.cpp:389-392 — BABAR box
    (BOX_BABAR, "DistanceToIn", (0.15000000000000185, -22.048743592955137, 2.4268539333219472),
     unit((-0.76165597579890043, 0.64364445891356026, -0.074515708658524193)), ("eq", 0.0)),
    # TestBox.cpp:395-398 — tolerance of DistanceToIn
    ((5.0, 5.0, 5.0), "DistanceToIn", (-3.0087437277453119577, -4.9999999999999928946, 4.8935648380409944025),
     unit((0.76315134679548990437, 0.53698876104646497964, -0.35950395323836459305)), ("le", 0.0)),
    # TestBox.cpp:400-419 — leaving surface point
    (BOX_B2, "DistanceToIn", (10.0, 0, 0), (1.0, 0, 0), ("inf",)),
    (BOX_B2, "DistanceToIn", (10 - 0.9 * kHalfTolerance, 0, 0), (1.0, 0, 0), ("inf",)),
    (BOX_B2, "Inside", (10 - 0.9 * kHalfTolerance, 0, 0), None, ("inside", 2)),
    # TestBox.cpp:437-466 — VECGEOM-431
    ((1.0, 200.0, 200.0), "DistanceToIn", (1.0 + 6.0e-15, 0, 0), _vecgeom431_in, ("eq", 0.0)),
    ((1.0, 200.0, 200.0), "DistanceToOut", (1.0 + 6.0e-15, 0, 0), _vecgeom431_in, ("eq", 200.0 * math.sqrt(2.0))),
    ((1.0, 200.0, 200.0), "DistanceToOut", (1.0 + 6.0e-15, 0, 0), VX, ("eq", 0.0)),
    ((1.0, 200.0, 200.0), "DistanceToIn", (-1.0 - 6.0e-15, 0, 0), _vecgeom431_in_m, ("eq", 0.0)),
]
