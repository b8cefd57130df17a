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
    if kind == "exact":
        return float(result) == float(expect[1])
    if kind == "near":
        return abs(float(result) - expect[1]) < expect[2]
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


def mag_unit(v):
    """v /= v.Mag() (Vector3D operator/=)."""
    v = np.asarray(v, dtype=np.float64)
    m = np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return tuple(float(c) for c in v / m)


# ---- Tube: test/unit_tests/TestTube.cpp ---------------------------------------
PI = math.pi
T1 = (0.0, 50.0, 50.0, 0.0, 2 * PI)
T1A = (0.0, 50.0, 50.0, 0.0, 0.5 * PI)
T2 = (45.0, 50.0, 50.0, 0.0, 2 * PI)
T2A = (5.0, 50.0, 50.0, 0.0, 2 * PI)
T3 = (0.0, 50.0, 50.0, 0.5 * PI, 0.5 * PI)
T5 = (50.0, 100.0, 50.0, 0.0, 1.5 * PI)
TUBE6 = (750.0, 760.0, 350.0, 0.31415926535897931, 5.6548667764616276)
TUBE7 = (2200.0, 3200.0, 2500.0, -0.68977164349384879, 3.831364227270472)
TUBE8 = (2550.0, 2580.0, 2000.0, 0.0, 2 * PI)
TUBE9 = (1150.0, 1180.0, 2000.0, 0.0, 2 * PI)
DEG = PI / 180.0
JIRA174 = (57.555599999999998, 205.55599999999998, 348.0, -5.5747247 * DEG, 5.5747247 * DEG)
_d174 = mag_unit((0.303163, -0.492481, 0.815815))
VMXMY = (-S2, -S2, 0.0)
VXMY = (S2, -S2, 0.0)
P0 = (0.0, 0.0, 0.0)
PBX, PBMX, PBY, PBMY, PBZ, PBMZ = (100.0, 0, 0), (-100.0, 0, 0), (0, 100.0, 0), (0, -100.0, 0), (0, 0, 100.0), (0, 0, -100.0)
PONX, PONY, PONZ = (50.0, 0, 0), (0, 50.0, 0), (0, 0, 50.0)

TUBE_KATS = [
    # TestTube.cpp:85-105 (jira 174)
    (JIRA174, "DistanceToIn", (92.4733, 5.32907e-15, 19.943), _d174, ("eq", 0.0)),
    (JIRA174, "DistanceToOut", (92.4733, 5.32907e-15, 19.943), _d174, ("eq", 19.499)),
    (JIRA174, "DistanceToIn", (92.4733, -5.32907e-15, 19.943), tuple(-c for c in _d174), ("inf",)),
    (JIRA174, "DistanceToOut", (92.4733, -5.32907e-15, 19.943), tuple(-c for c in _d174), ("eq", 0.0)),
    # :116-121 (VECGEOM-206)
    (T1, "DistanceToOut", (44.991, 21.816, 4.677), unit((0.16636, 0.64765, -0.74356)), ("le", -1e-300)),
    # :124-130 (VECGEOM-222)
    (T5, "DistanceToIn", (60.0, -60.0, 10.0), unit((-1.0, 1.0, 0.001)), ("eq", 134.8528474556)),
    # :134-153 Inside
    (T1, "Inside", P0, None, ("inside", 1)),
    (T1, "Inside", PBZ, None, ("inside", 3)),
    (T1, "Inside", PONX, None, ("inside", 2)),
    (T1, "Inside", PONY, None, ("inside", 2)),
    (T1, "Inside", PONZ, None, ("inside", 2)),
    (T1A, "Inside", P0, None, ("inside", 2)),
    (T1A, "Inside", PBZ, None, ("inside", 3)),
    (T1A, "Inside", PONX, None, ("inside", 2)),
    (T1A, "Inside", PONY, None, ("inside", 2)),
    (T1A, "Inside", PONZ, None, ("inside", 2)),
    (T2, "Inside", P0, None, ("inside", 3)),
    (T2, "Inside", PBZ, None, ("inside", 3)),
    (T2, "Inside", PONX, None, ("inside", 2)),
    (T2, "Inside", PONY, None, ("inside", 2)),
    (T2, "Inside", PONZ, None, ("inside", 3)),
    (T2A, "Inside", P0, None, ("inside", 3)),
    (T2A, "Inside", PONX, None, ("inside", 2)),
    (T2A, "Inside", PONZ, None, ("inside", 3)),
    # :179-196 SafetyToOut / DistanceToOut from the centre
    (T1, "SafetyToOut", P0, None, ("eq", 50)),
    (T1, "DistanceToOut", P0, VX, ("eq", 50)),
    (T1, "DistanceToOut", P0, VMX, ("eq", 50)),
    (T1, "DistanceToOut", P0, VY, ("eq", 50)),
    (T1, "DistanceToOut", P0, VMY, ("eq", 50)),
    (T1, "DistanceToOut", P0, VZ, ("eq", 50)),
    (T1, "DistanceToOut", P0, VMZ, ("eq", 50)),
    (T1, "DistanceToOut", P0, VXY, ("eq", 50)),
    # :239-262 sector
    (T3, "DistanceToOut", (0.0, 10.0, 0.0), VX, ("eq", 0)),
    (T3, "Inside", (0.5, 10.0, 0.0), None, ("inside", 3)),
    (T3, "DistanceToOut", (0.5, 10.0, 0.0), VX, ("le", -1e-300)),
    (T3, "DistanceToOut", (-0.5, 9.0, 0.0), VX, ("eq", 0.5)),
    (T3, "DistanceToOut", (-5.0, 9.5, 0.0), VX, ("eq", 5)),
    (T3, "DistanceToOut", (-5.0, 9.5, 0.0), VMY, ("eq", 9.5)),
    (T3, "DistanceToOut", (-5.0, 9.0, 0.0), VXMY, ("eq", 7.0710678)),
    # :265-297 SafetyToIn / DistanceToIn
    (T1, "SafetyToIn", PBX, None, ("eq", 50)),
    (T1, "SafetyToIn", PBMX, None, ("eq", 50)),
    (T1, "SafetyToIn", PBY, None, ("eq", 50)),
    (T1, "SafetyToIn", PBMY, None, ("eq", 50)),
    (T1, "SafetyToIn", PBZ, None, ("eq", 50)),
    (T1, "SafetyToIn", PBMZ, None, ("eq", 50)),
    (T1, "DistanceToIn", PBX, VMX, ("eq", 50)),
    (T1, "DistanceToIn", PBMX, VX, ("eq", 50)),
    (T1, "DistanceToIn", PBY, VMY, ("eq", 50)),
    (T1, "DistanceToIn", PBMY, VY, ("eq", 50)),
    (T1, "DistanceToIn", PBZ, VMZ, ("eq", 50)),
    (T1, "DistanceToIn", PBMZ, VZ, ("eq", 50)),
    (T1, "DistanceToIn", PBX, VXY, ("inf",)),
    (T1A, "DistanceToIn", PBZ, VMZ, ("eq", 50)),
    # :315-356
    (T3, "DistanceToIn", (49.5, 5.0, 0.0), VMX, ("eq", 49.5)),
    (T3, "DistanceToIn", (49.5, -0.5, 0.0), VMX, ("inf",)),
    (T5, "DistanceToIn", (30.0, -20.0, 0.0), VXY, ("eq", 28.284271)),
    (T5, "DistanceToIn", (30.0, -20.0, 0.0), VMXMY, ("eq", 42.426407)),
    (T5, "DistanceToIn", (50.0, -20.0, 0.0), VY, ("eq", 20)),
    (T5, "DistanceToIn", (30.0, -50.0, 0.0), VMX, ("eq", 30)),
    # :404-428 (bugs 76, 91 and others)
    (TUBE6, "DistanceToOut", (-388.20504321896431, -641.71398957741451, 332.85995254027955),
     (-0.47312863350457468, -0.782046391443315, 0.40565100491504164), ("eq", 10.940583)),
    (TUBE7, "DistanceToOut", (-2460.0, 1030.0, -2500.0),
     (-0.086580540180167642, 0.070084247882560638, 0.9937766390194761), ("eq", 4950.348576972614)),
    (TUBE8, "DistanceToOut", (6.71645645882942, 2579.415860329989, -1.519530725281157),
     (-0.6305220496340839, -0.07780451841562354, 0.7722618738739774), ("eq", 1022.64931421)),
    (TUBE9, "DistanceToOut", (2.267347771505638, 1170.164934028592, 4.820317321984064),
     (-0.1443054266272111, -0.01508874701037938, 0.9894181489944458), ("eq", 2016.51817758)),
    (T1A, "DistanceToOut", (0.0, 0.0, 50.0), VX, ("eq", 50)),
    (T1A, "DistanceToOut", (0.0, 5.0, 50.0), VMY, ("eq", 5)),
    # :475-479
    ((45.0, 50.0, 50.0, PI / 4.0, PI / 8.0), "Inside", (33.315052227388207, 37.284142675357259, 33.366096020078537),
     None, ("inside", 2)),
]


# ---- Cone: test/unit_tests/TestCons.cpp -----------------------------------------
# Inside asserts (:183-221) are hard KATs; the OutRange(dist, wanted) checks
# (DELTA = 0.01, :28-35) only print in the reference and are kept as soft KATs.
C1 = (50.0, 100.0, 50.0, 100.0, 50.0, 0.0, 2 * PI)
C2 = (50.0, 100.0, 50.0, 200.0, 50.0, -1.0, 2 * PI)
C3 = (50.0, 100.0, 50.0, 100.0, 50.0, -PI / 6, PI / 3)
C4 = (50.0, 100.0, 50.0, 200.0, 50.0, -PI / 6, PI / 3)
C5 = (25.0, 50.0, 75.0, 150.0, 50.0, 0.0, 1.5 * PI)
C6 = (0.0, 150.0, 0.0, 150.0, 50.0, 0.0, 2 * PI)
C7 = (95.0, 100.0, 95.0, 100.0, 50.0, 0.0, 2 * PI)
C8A = (0.0, 100.0, 0.0, 150.0, 50.0, 0.0, 2 * PI)
C8B = (50.0, 100.0, 100.0, 150.0, 50.0, 0.0, 2 * PI)
C8C = (100.0, 150.0, 50.0, 100.0, 50.0, 0.0, 2 * PI)
C9 = (50.0, 60.0, 0.0, 10.0, 50.0, 0.0, 2 * PI)
CTEST10 = (20.0, 60.0, 80.0, 140.0, 100.0, 10 * PI / 180.0, 300 * PI / 180.0)
PPLX, PPLY, PPLZ = (120.0, 0, 0), (0, 120.0, 0), (0, 0, 120.0)
PMIY, PMIZ = (0, -120.0, 0), (0, 0, -120.0)
PONPHI1 = (60 * math.cos(PI / 6), -60 * math.sin(PI / 6), 0.0)
PNEARPLZ, PNEARMIZ = (45.0, 45.0, 45.0), (45.0, 45.0, -45.0)
PONR1 = (math.sqrt(50 * 50 / 2.0), math.sqrt(50 * 50 / 2.0), 0.0)
PONR2 = (math.sqrt(100 * 100 / 2.0), math.sqrt(100 * 100 / 2.0), 0.0)

CONE_KATS = [
    (CTEST10, "Inside", (-691.0 - 500, 174.0, 404.0), None, ("inside", 3)),
    (CTEST10, "Inside", (400.0 - 500, 20.9, 5.89), None, ("inside", 1)),
    (CTEST10, "Inside", (456.0 - 500, 13.0, -14.7), None, ("inside", 1)),
    (CTEST10, "Inside", (537.0 - 500, 1.67, -44.1), None, ("inside", 3)),
    (CTEST10, "Inside", (537.0, 1.67, -44.1), None, ("inside", 3)),
    (CTEST10, "Inside", (1e3, -63.5, -213.0), None, ("inside", 3)),
    (CTEST10, "Inside", (-50.0, 0.0, 0.0), None, ("inside", 2)),
    (CTEST10, "Inside", (6.454731216775542, -90.42080754048007, 100.0), None, ("inside", 2)),
    (CTEST10, "Inside", (22.65282328600368, -69.34877585931267, 76.51600623610082), None, ("inside", 2)),
    (CTEST10, "Inside", (51.28206938732319, -32.10510677306267, 35.00932544708616), None, ("inside", 3)),
]

CONE_SOFT_KATS = [  # (shape, method, p, d, wanted) with |result - wanted| <= 0.01
    (C4, "SafetyToOut", PONPHI1, None, 0.0),
    (C1, "SafetyToOut", PONPHI1, None, 10.0),
    (C1, "SafetyToOut", PNEARPLZ, None, 5.0),
    (C1, "SafetyToOut", PNEARMIZ, None, 5.0),
    (C1, "SafetyToOut", PONR1, None, 0.0),
    (C1, "SafetyToOut", PONR2, None, 0.0),
    (C6, "SafetyToOut", P0, None, 50.0),
    (C5, "SafetyToOut", (0.0, -70.0, 0.0), None, 0.0),
    (C4, "DistanceToOut", PPLX, VMX, 70.0),
    (C2, "DistanceToOut", PPLX, VMX, 70.0),
    (C1, "SafetyToIn", P0, None, 50.0),
    (C1, "SafetyToIn", PPLX, None, 20.0),
    (C1, "SafetyToIn", PPLY, None, 20.0),
    (C1, "SafetyToIn", PPLZ, None, 70.0),
    (C6, "DistanceToIn", PPLZ, VMZ, 70.0),
    (C8A, "DistanceToIn", PPLZ, VMZ, 70.0),
    (C9, "DistanceToIn", PPLZ, VMZ, 70.0),
    (C9, "DistanceToIn", (0.0, 0.0, 50.0), VMZ, 0.0),
    (C6, "DistanceToIn", PMIZ, VZ, 70.0),
    (C8A, "DistanceToIn", PMIZ, VZ, 70.0),
    (C1, "DistanceToIn", PPLX, VMX, 20.0),
] + [(c, "DistanceToIn", PPLZ, VMZ, KINF) for c in (C1, C2, C3, C4, C5, C7, C8B, C8C)] \
  + [(c, "DistanceToIn", PMIZ, VZ, KINF) for c in (C1, C2, C3, C4, C5, C7, C8B, C8C, C9)]


# ---- Trd: test/unit_tests/TestTrd.cpp -------------------------------------------
TRD1 = (20.0, 20.0, 30.0, 30.0, 40.0)
TRD2 = (10.0, 30.0, 20.0, 40.0, 40.0)
TRD3 = (0.14999999999999999, 0.14999999999999999, 24.707000000000001, 24.707000000000001, 22.699999999999999)
COSA, SINA = 4 / math.sqrt(17.0), 1 / math.sqrt(17.0)
PONX1, PONY1, PONZ1 = (20.0, 0, 0), (0, 30.0, 0), (0, 0, 40.0)
PONMX1, PONMY1, PONMZ1 = (-20.0, 0, 0), (0, -30.0, 0), (0, 0, -40.0)

TRD_KATS = [
    (TRD1, "Inside", P0, None, ("inside", 1)),
    (TRD1, "Inside", PBZ, None, ("inside", 3)),
    (TRD1, "Inside", PONX1, None, ("inside", 2)),
    (TRD1, "Inside", PONY1, None, ("inside", 2)),
    (TRD1, "Inside", PONZ1, None, ("inside", 2)),
    (TRD2, "Inside", P0, None, ("inside", 1)),
    (TRD2, "Inside", PBZ, None, ("inside", 3)),
    (TRD2, "Inside", PONX1, None, ("inside", 2)),
    (TRD2, "Inside", PONY1, None, ("inside", 2)),
    (TRD2, "Inside", PONZ1, None, ("inside", 2)),
    (TRD1, "SafetyToOut", P0, None, ("eq", 20)),
    (TRD1, "SafetyToOut", VX, None, ("eq", 19)),
    (TRD1, "SafetyToOut", VY, None, ("eq", 20)),
    (TRD1, "SafetyToOut", VZ, None, ("eq", 20)),
    (TRD2, "SafetyToOut", P0, None, ("eq", 20 * COSA)),
    (TRD2, "SafetyToOut", VX, None, ("eq", 19 * COSA)),
    (TRD2, "SafetyToOut", VY, None, ("eq", 20 * COSA)),
    (TRD2, "SafetyToOut", VZ, None, ("eq", 20 * COSA + SINA)),
    (TRD1, "DistanceToOut", P0, VX, ("eq", 20)),
    (TRD1, "DistanceToOut", P0, VMX, ("eq", 20)),
    (TRD1, "DistanceToOut", P0, VY, ("eq", 30)),
    (TRD1, "DistanceToOut", P0, VMY, ("eq", 30)),
    (TRD1, "DistanceToOut", P0, VZ, ("eq", 40)),
    (TRD1, "DistanceToOut", P0, VMZ, ("eq", 40)),
    (TRD1, "DistanceToOut", P0, VXY, ("eq", math.sqrt(800.0))),
    (TRD1, "DistanceToOut", PONX1, VX, ("eq", 0)),
    (TRD1, "DistanceToOut", PONMX1, VMX, ("eq", 0)),
    (TRD1, "DistanceToOut", PONY1, VY, ("eq", 0)),
    (TRD1, "DistanceToOut", PONMY1, VMY, ("eq", 0)),
    (TRD1, "DistanceToOut", PONZ1, VZ, ("eq", 0)),
    (TRD1, "DistanceToOut", PONMZ1, VMZ, ("eq", 0)),
    (TRD1, "SafetyToIn", PBX, None, ("eq", 80)),
    (TRD1, "SafetyToIn", PBMX, None, ("eq", 80)),
    (TRD1, "SafetyToIn", PBY, None, ("eq", 70)),
    (TRD1, "SafetyToIn", PBMY, None, ("eq", 70)),
    (TRD1, "SafetyToIn", PBZ, None, ("eq", 60)),
    (TRD1, "SafetyToIn", PBMZ, None, ("eq", 60)),
    (TRD2, "SafetyToIn", PBX, None, ("eq", 80 * COSA)),
    (TRD2, "SafetyToIn", PBMX, None, ("eq", 80 * COSA)),
    (TRD2, "SafetyToIn", PBY, None, ("eq", 70 * COSA)),
    (TRD2, "SafetyToIn", PBMY, None, ("eq", 70 * COSA)),
    (TRD2, "SafetyToIn", PBZ, None, ("eq", 60)),
    (TRD2, "SafetyToIn", PBMZ, None, ("eq", 60)),
] + [(t, "DistanceToIn", p, d, ("eq", e)) for t in (TRD1, TRD2) for p, d, e in
     ((PBX, VMX, 80), (PBMX, VX, 80), (PBY, VMY, 70), (PBMY, VY, 70), (PBZ, VMZ, 60), (PBMZ, VZ, 60))] \
  + [(t, "DistanceToIn", p, VXY, ("inf",)) for t in (TRD1, TRD2) for p in (PBX, PBMX)] + [
    (TRD3, "DistanceToIn", (0.15000000000000185, -22.048743592955137, 2.4268539333219472),
     unit((-0.76165597579890043, 0.64364445891356026, -0.074515708658524193)), ("eq", 0.0)),
]

# ---- Cone: the second test of test/unit_tests/TestCons.cpp (:996-1087), hard asserts
# t1 = tube-like cone (0, 50, 0, 50, 50, 0, 2 pi); test10 (:1009)
CONE_T1 = (0.0, 50.0, 0.0, 50.0, 50.0, 0.0, 2.0 * PI)
CONE_TEST10 = (20.0, 80.0, 60.0, 140.0, 100.0, 0.17453292519943, 5.235987755983)
CONE_T1_KATS = [
    (CONE_T1, "Inside", (0.0, 0.0, 0.0), None, ("inside", 1)),
    (CONE_T1, "Inside", (100.0, 0, 0), None, ("inside", 3)),
    (CONE_T1, "SafetyToOut", (0.0, 0.0, 0.0), None, ("eq", 50)),
] + [(CONE_T1, "DistanceToOut", (0.0, 0.0, 0.0), d, ("eq", 50)) for d in (VX, VMX, VY, VMY, VZ, VMZ, VXY)] + [
    (CONE_T1, "SafetyToIn", p, None, ("eq", 50)) for p in ((100.0, 0, 0), (-100.0, 0, 0), (0, 100.0, 0),
                                                            (0, -100.0, 0), (0, 0, 100.0), (0, 0, -100.0))] + [
    (CONE_T1, "DistanceToIn", p, d, ("eq", 50)) for p, d in (((100.0, 0, 0), VMX), ((-100.0, 0, 0), VX),
                                                             ((0, 100.0, 0), VMY), ((0, -100.0, 0), VY),
                                                             ((0, 0, 100.0), VMZ), ((0, 0, -100.0), VZ))] + [
    (CONE_T1, "DistanceToIn", (100.0, 0, 0), VXY, ("inf",)),
    # a point on (rmin, z = -100) moving out of test10: DistanceToOut is 0 (:1080-1084)
    (CONE_TEST10, "DistanceToOut", (19.218716967888, 5.5354239324172, -100.0),
     (-0.25644483536346, -0.073799216676426, 0.96373737191901), ("eq", 0)),
]


# ---- Trapezoid: test/unit_tests/TestTrap.cpp (box-like trap1) ---------------------
TRAP1 = (40.0, 0.0, 0.0, 30.0, 20.0, 20.0, 0.0, 30.0, 20.0, 20.0, 0.0)
TRAP_KATS = [
    (TRAP1, "Inside", P0, None, ("inside", 1)),
    (TRAP1, "Inside", PBZ, None, ("inside", 3)),
    (TRAP1, "Inside", PONX1, None, ("inside", 2)),
    (TRAP1, "Inside", PONY1, None, ("inside", 2)),
    (TRAP1, "Inside", PONZ1, None, ("inside", 2)),
    (TRAP1, "SafetyToOut", P0, None, ("eq", 20)),
    (TRAP1, "SafetyToOut", VX, None, ("eq", 19)),
    (TRAP1, "SafetyToOut", VY, None, ("eq", 20)),
    (TRAP1, "SafetyToOut", VZ, None, ("eq", 20)),
    (TRAP1, "SafetyToIn", (100.0, 100.0, 100.0), None, ("eq", 80)),
    (TRAP1, "SafetyToIn", PBX, None, ("eq", 80)),
    (TRAP1, "SafetyToIn", PBMX, None, ("eq", 80)),
    (TRAP1, "SafetyToIn", PBY, None, ("eq", 70)),
    (TRAP1, "SafetyToIn", PBMY, None, ("eq", 70)),
    (TRAP1, "SafetyToIn", PBZ, None, ("eq", 60)),
    (TRAP1, "SafetyToIn", PBMZ, None, ("eq", 60)),
    (TRAP1, "SafetyToIn", (20.11, 0, 0), None, ("eq", 0.11)),
    (TRAP1, "SafetyToOut", (20.11, 0, 0), None, ("le", 0.0)),
    (TRAP1, "SafetyToIn", (19.89, 0, 0), None, ("le", 0.0)),
    (TRAP1, "SafetyToOut", (19.89, 0, 0), None, ("eq", 0.11)),
    (TRAP1, "SafetyToOut", (-19.89, 0, 0), None, ("ge", 0.0)),
    (TRAP1, "SafetyToIn", (-20.11, 0, 0), None, ("eq", 0.11)),
    (TRAP1, "SafetyToOut", (-20.11, 0, 0), None, ("le", 0.0)),
    (TRAP1, "SafetyToIn", (0, 30.11, 0), None, ("eq", 0.11)),
    (TRAP1, "SafetyToOut", (0, 30.11, 0), None, ("le", 0.0)),
    (TRAP1, "SafetyToOut", (0, 29.89, 0), None, ("eq", 0.11)),
]

# TestTrap.cpp, the rest of trap1's asserts and trap2 (the trd-like trap
# built from 8 vertices, :63-70; its parameter form is the commented line
# :68), trap3 and trap4 (parallelepipeds sheared by alpha = +-pi/4, :72-73).
TRAP2 = (40.0, 0.0, 0.0, 20.0, 10.0, 10.0, 0.0, 40.0, 30.0, 30.0, 0.0)
TRAP3 = (50.0, 0.0, 0.0, 50.0, 50.0, 50.0, PI / 4, 50.0, 50.0, 50.0, PI / 4)
TRAP4 = (50.0, 0.0, 0.0, 50.0, 50.0, 50.0, -PI / 4, 50.0, 50.0, 50.0, -PI / 4)
_TV = 0.11  # testValue (:447)
VYMZ = (0.0, S2, -S2)
_T_OUT = ((P0, VX, 20), (P0, VMX, 20), (P0, VY, 30), (P0, VMY, 30), (P0, VZ, 40), (P0, VMZ, 40),
          (P0, VXY, math.sqrt(800.0)), (PONX1, VX, 0), (PONMX1, VMX, 0), (PONY1, VY, 0), (PONMY1, VMY, 0),
          (PONZ1, VZ, 0), (PONMZ1, VMZ, 0))
_T_IN = ((PBX, VMX, 80), (PBMX, VX, 80), (PBY, VMY, 70), (PBMY, VY, 70), (PBZ, VMZ, 60), (PBMZ, VZ, 60))
TRAP_KATS += [(TRAP1, "DistanceToOut", p, d, ("eq", e)) for p, d, e in _T_OUT] + [
    (TRAP1, "DistanceToOut", p, d, ("eq", e)) for p, d, e in ((PONX1, VMX, 40), (PONMX1, VX, 40),
                                                             (PONY1, VMY, 60), (PONMY1, VY, 60),
                                                             (PONZ1, VMZ, 80), (PONMZ1, VZ, 80))] + [
    (TRAP1, "SafetyToOut", (0, -30.0 + _TV, 0), None, ("eq", _TV)),
    (TRAP1, "SafetyToIn", (0, -30.0 - _TV, 0), None, ("eq", _TV)),
    (TRAP1, "SafetyToOut", (0, -30.0 - _TV, 0), None, ("le", 0.0)),
    (TRAP1, "SafetyToIn", (0, 0, 40.0 + _TV), None, ("eq", _TV)),
    (TRAP1, "SafetyToOut", (0, 0, 40.0 + _TV), None, ("le", 0.0)),
    (TRAP1, "SafetyToIn", (0, 0, 40.0 - _TV), None, ("le", 0.0)),
    (TRAP1, "SafetyToOut", (0, 0, 40.0 - _TV), None, ("eq", _TV)),
    (TRAP1, "SafetyToIn", (0, 0, -40.0 + _TV), None, ("le", 0.0)),
    (TRAP1, "SafetyToOut", (0, 0, -40.0 + _TV), None, ("eq", _TV)),
    (TRAP1, "SafetyToIn", (0, 0, -40.0 - _TV), None, ("eq", _TV)),
    (TRAP1, "SafetyToOut", (0, 0, -40.0 - _TV), None, ("le", 0.0)),
] + [(TRAP1, "DistanceToIn", p, d, ("eq", e)) for p, d, e in _T_IN] + [
    (TRAP1, "DistanceToIn", PBX, VXY, ("inf",)),
    (TRAP1, "DistanceToIn", PBMX, VXY, ("inf",)),
    (TRAP1, "DistanceToIn", (0, 60.0, 0), VXMY, ("inf",)),
    (TRAP1, "DistanceToIn", (0, 50.0, 0), VXMY, ("eq", math.sqrt(800.0))),
    (TRAP1, "DistanceToIn", (0, 40.0, 0), VXMY, ("eq", 10.0 * math.sqrt(2.0))),
    (TRAP1, "DistanceToIn", (0, 40.0, 50.0), VXMY, ("inf",)),
    (TRAP1, "DistanceToIn", (40.0, 60.0, 0), VMX, ("inf",)),
    (TRAP1, "DistanceToIn", (40.0, 60.0, 0), VMY, ("inf",)),
    (TRAP1, "DistanceToIn", (40.0, 60.0, 50.0), VMZ, ("inf",)),
    (TRAP1, "DistanceToIn", (0, 0, 50.0), VYMZ, ("eq", 10.0 * math.sqrt(2.0))),
    (TRAP1, "DistanceToIn", (0, 0, 80.0), VYMZ, ("inf",)),
    (TRAP1, "DistanceToIn", (0, 0, 70.0), VYMZ, ("eq", 30.0 * math.sqrt(2.0))),
    (TRAP2, "Inside", P0, None, ("inside", 1)),
    (TRAP2, "Inside", PBZ, None, ("inside", 3)),
    (TRAP2, "Inside", PONX1, None, ("inside", 2)),
    (TRAP2, "Inside", PONY1, None, ("inside", 2)),
    (TRAP2, "Inside", PONZ1, None, ("inside", 2)),
    (TRAP2, "SafetyToOut", P0, None, ("eq", 20 * COSA)),
    (TRAP2, "SafetyToOut", VX, None, ("eq", 19 * COSA)),
    (TRAP2, "SafetyToOut", VY, None, ("eq", 20 * COSA)),
    (TRAP2, "SafetyToOut", VZ, None, ("eq", 20 * COSA + SINA)),
] + [(TRAP2, "DistanceToOut", p, d, ("eq", e)) for p, d, e in _T_OUT] + [
    (TRAP2, "SafetyToIn", p, None, ("eq", e)) for p, e in ((PBX, 80 * COSA), (PBMX, 80 * COSA), (PBY, 70 * COSA),
                                                           (PBMY, 70 * COSA), (PBZ, 60), (PBMZ, 60))] + [
    (TRAP2, "DistanceToIn", p, d, ("eq", e)) for p, d, e in _T_IN] + [
    (TRAP2, "DistanceToIn", PBX, VXY, ("inf",)),
    (TRAP2, "DistanceToIn", PBMX, VXY, ("inf",)),
    (TRAP3, "DistanceToIn", (50.0, -50.0, 0), VY, ("eq", 50)),
    (TRAP3, "DistanceToIn", (50.0, -50.0, 0), VMY, ("inf",)),
    (TRAP4, "DistanceToIn", (50.0, 50.0, 0), VY, ("inf",)),
    (TRAP4, "DistanceToIn", (50.0, 50.0, 0), VMY, ("eq", 50)),
]


# ---- Polycone: test/unit_tests/TestPolycone.cpp -----------------------------------
# shapes are (sphi, dphi, z, rmin, rmax); poly1 asserts at :57-88 are exact (==).
POLY1 = (0.0, 360.0 * PI / 180.0, (-1.0, -0.5, 0.5, 2.0), (0.1, 0.0, 0.0, 0.4), (1.0, 2.0, 2.0, 1.5))
PSIMPLE = (0.0, 360.0 * PI / 180.0, (-10.0, 0.0, 10.0), (0.0, 0.0, 0.0), (70.0, 70.0, 80.0))
MYPCONE = (-10.0 * PI / 180.0, 10.0 * PI / 180.0, (-20.0, -10.0, -10.0, 0.0, 10.0, 20.0, 30.0, 40.0),
           (30.0, 30.0, 0.0, 0.0, 0.0, 0.0, 40.0, 40.0), (70.0, 70.0, 70.0, 40.0, 40.0, 80.0, 80.0, 60.0))


def _exact(v):
    return ("exact", v)


POLYCONE_KATS = [
    (POLY1, "Contains", (0.0, 0.0, 0.0), None, _exact(1)),
    (POLY1, "Contains", (0.0, 0.0, -2.0), None, _exact(0)),
    (POLY1, "Contains", (0.0, 0.0, -0.8), None, _exact(0)),
    (POLY1, "Contains", (0.0, 0.0, -1.8), None, _exact(0)),
    (POLY1, "Contains", (0.0, 0.0, 10.0), None, _exact(0)),
    (POLY1, "Contains", (0.0, 0.0, 1.8), None, _exact(0)),
    (POLY1, "DistanceToIn", (0.0, 0.0, -3.0), (0.0, 0.0, 1.0), _exact(2.5)),
    (POLY1, "DistanceToIn", (0.0, 0.0, -2.0), (0.0, 0.0, -1.0), _exact(KINF)),
    (POLY1, "DistanceToIn", (0.0, 0.0, 3.0), (0.0, 0.0, -1.0), _exact(2.5)),
    (POLY1, "DistanceToIn", (0.0, 0.0, 3.0), (0.0, 0.0, 1.0), _exact(KINF)),
    (POLY1, "DistanceToIn", (3.0, 0.0, 0.0), (-1.0, 0.0, 0.0), _exact(1.0)),
    (POLY1, "DistanceToIn", (0.0, 0.0, 1.9999999), (1.0, 0.0, 0.0), ("near", 0.4, 1e-6)),
    (POLY1, "SafetyToIn", (0.0, 0.0, -3.0), None, _exact(2.0)),
    (POLY1, "SafetyToIn", (0.5, 0.0, -1.0), None, _exact(0.0)),
    (POLY1, "SafetyToIn", (0.0, 0.0, 3.0), None, _exact(1.0)),
    (POLY1, "SafetyToIn", (2.0, 0.0, 0.1), None, _exact(0.0)),
    (POLY1, "SafetyToOut", (0.0, 0.0, 0.0), None, _exact(0.5)),
    (POLY1, "SafetyToOut", (0.0, 0.0, 0.5), None, _exact(0.0)),
    (POLY1, "SafetyToOut", (1.9, 0.0, 0.0), None, ("near", 0.1, 1e-6)),
    (POLY1, "SafetyToOut", (0.2, 0.0, -1.0), None, _exact(0.0)),
    (POLY1, "SafetyToOut", (1.4, 0.0, 2.0), None, _exact(0.0)),
    (POLY1, "DistanceToOut", (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), _exact(0.5)),
    (POLY1, "DistanceToOut", (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), _exact(0.5)),
    (POLY1, "DistanceToOut", (2.0, 0.0, 0.0), (1.0, 0.0, 0.0), _exact(0.0)),
    (POLY1, "DistanceToOut", (2.0, 0.0, 0.0), (-1.0, 0.0, 0.0), _exact(4.0)),
    (POLY1, "DistanceToOut", (1.0, 0.0, 2.0), (0.0, 0.0, 1.0), _exact(0.0)),
    (POLY1, "DistanceToOut", (0.5, 0.0, -1.0), (0.0, 0.0, -1.0), _exact(0.0)),
    (POLY1, "DistanceToOut", (0.5, 0.0, -1.0), (0.0, 0.0, 1.0), _exact(3.0)),
    # :158-171 Inside of the simple tube+cone
    (PSIMPLE, "Inside", (0.0, 0.0, 0.0), None, ("inside", 1)),
    (PSIMPLE, "Inside", (0.0, 0.0, 100.0), None, ("inside", 3)),
    (PSIMPLE, "Inside", (100.0, 0.0, 0.0), None, ("inside", 3)),
    (PSIMPLE, "Inside", (0.0, 100.0, 0.0), None, ("inside", 3)),
    (PSIMPLE, "Inside", (70.0, 0.0, -5.0), None, ("inside", 2)),
    (PSIMPLE, "Inside", (0.0, 70.0, -5.0), None, ("inside", 2)),
    (PSIMPLE, "Inside", (70.0, 0.0, 10.0), None, ("inside", 2)),
    (PSIMPLE, "Inside", (-70.0, 0.0, -5.0), None, ("inside", 2)),
    (PSIMPLE, "Inside", (0.0, -70.0, -5.0), None, ("inside", 2)),
    (PSIMPLE, "Inside", (0.0, 0.0, -10.0), None, ("inside", 2)),
    (PSIMPLE, "Inside", (0.0, 25.0, 0.0), None, ("inside", 1)),
    (PSIMPLE, "Inside", (4.0, 25.0, 0.0), None, ("inside", 1)),
    # :209-216 SafetyToOut, :236-251 DistanceToIn
    (PSIMPLE, "SafetyToOut", (5.0, 5.0, -5.0), None, ("eq", 5)),
    (PSIMPLE, "SafetyToOut", (5.0, 5.0, 7.0), None, ("eq", 3)),
    (PSIMPLE, "SafetyToOut", (69.0, 0.0, -5.0), None, ("eq", 1)),
    (PSIMPLE, "SafetyToOut", (-3.0, -3.0, 8.0), None, ("eq", 2)),
    (PSIMPLE, "DistanceToIn", (100.0, 0.0, -1.0), VMX, ("eq", 30)),
    (PSIMPLE, "DistanceToIn", (-100.0, 0.0, -1.0), VX, ("eq", 30)),
    (PSIMPLE, "DistanceToIn", (0.0, 100.0, -5.0), VMY, ("eq", 30)),
    (PSIMPLE, "DistanceToIn", (0.0, -100.0, -5.0), VY, ("eq", 30)),
    (PSIMPLE, "DistanceToIn", (0.0, 0.0, 100.0), VMZ, ("eq", 90)),
    (PSIMPLE, "DistanceToIn", (0.0, 0.0, -100.0), VZ, ("eq", 90)),
    (PSIMPLE, "DistanceToIn", (100.0, 0.0, 0.0), VXY, ("inf",)),
    (PSIMPLE, "DistanceToIn", (-100.0, 0.0, 0.0), VMXY, ("inf",)),
]

# ---- Polyhedron: test/unit_tests/TestPolyhedra.cpp ---------------------------
# shapes are (sphi, dphi, nsides, z, rmin, rmax).  The asserts at :183-224 are
# `fabs(result - value) < kTolerance` ("near" with 1e-9) or exact Inside codes.
MYPGON = (0.0, PI / 4.0, 5, (-30.0, -20.0, -10.0, 0.0, 10.0, 20.0, 30.0, 40.0),
          (30.0, 30.0, 0.0, 0.0, 0.0, 0.0, 40.0, 40.0), (70.0, 70.0, 70.0, 40.0, 40.0, 80.0, 80.0, 60.0))
# :97-101 SimplePolyhedron "Hexagonal prism" from the (r, z) contour
# {(15,10), (15,-10), (0,-10), (0,10)}.  The contour constructor takes corner
# radii and converts them to side radii r * cos(sidePhi / 2)
# (VecGeom/volumes/PolyhedronStruct.h:204-222): z planes (-10, 10), rmin 0,
# rmax 15 cos(pi/6), six sides, full phi.
_HEX_R = 15.0 * math.cos(0.5 * (2.0 * PI / 6))
HEXPRISM = (0.0, 2.0 * PI, 6, (-10.0, 10.0), (0.0, 0.0), (_HEX_R, _HEX_R))
_PG_TOL = 1e-9
_PG_P1, _PG_P2, _PG_P3 = (0.0, 0.0, -5.0), (50.0, 0.0, 40.0), (5.0, 1.0, 20.0)
_PG_P4, _PG_P5, _PG_P6 = (45.0, 5.0, 30.0), (0.0, 0.0, 30.0), (41.0, 0.0, 10.0)


def _near(v):
    return ("near", v, _PG_TOL)


POLYHEDRON_KATS = [
    (MYPGON, "Inside", _PG_P1, None, ("inside", 2)),
    (MYPGON, "Inside", _PG_P2, None, ("inside", 2)),
    (MYPGON, "Inside", _PG_P3, None, ("inside", 1)),
    (MYPGON, "Inside", _PG_P4, None, ("inside", 1)),
    (MYPGON, "Inside", _PG_P5, None, ("inside", 3)),
    (MYPGON, "Inside", _PG_P6, None, ("inside", 3)),
    (HEXPRISM, "Inside", O, None, ("inside", 1)),
    (HEXPRISM, "Inside", (15.0, 0.0, 0.0), None, ("inside", 2)),
    (MYPGON, "Contains", (20.0 + 1e-9, 20.0, 5.0), None, _exact(1)),
    (MYPGON, "Inside", (20.0 + 1e-9, 20.0, 5.0), None, ("inside", 1)),
    (MYPGON, "Contains", (20.0, 20.0, 5.0), None, _exact(0)),
    (MYPGON, "Inside", (20.0, 20.0, 5.0), None, ("inside", 2)),
    (MYPGON, "Contains", (20.0 - 1e-9, 20.0, 5.0), None, _exact(0)),
    (MYPGON, "Inside", (20.0 - 1e-9, 20.0, 5.0), None, ("inside", 3)),
    (MYPGON, "DistanceToIn", _PG_P1, VX, _near(0.0)),
    (MYPGON, "DistanceToIn", _PG_P1, VMY, _near(0.0)),
    (MYPGON, "DistanceToIn", _PG_P5, VX, _near(40.12368793931)),
    (MYPGON, "DistanceToIn", (41.0, 0.0, 10.0), VMX, _near(0.87631206069)),
    (MYPGON, "DistanceToIn", (41.0, 0.0, 10.0), VZ, _near(0.218402670765)),
    (MYPGON, "DistanceToOut", _PG_P1, VMX, _near(0.0)),
    (MYPGON, "DistanceToOut", _PG_P3, VMY, _near(1.0)),
    (MYPGON, "DistanceToOut", _PG_P3, VZ, _near(1.27382374146)),
    (MYPGON, "DistanceToOut", _PG_P4, VZ, _near(10.0)),
    (MYPGON, "DistanceToOut", _PG_P4, VX, _near(34.8538673445)),
    (MYPGON, "DistanceToOut", _PG_P4, VY, _near(40.0)),
    (HEXPRISM, "DistanceToOut", O, VX, ("ge", 1e-300)),
    (HEXPRISM, "DistanceToOut", O, VY, ("ge", 1e-300)),
    (HEXPRISM, "DistanceToOut", O, VZ, ("ge", 1e-300)),
]
# :187 DistanceToIn(p2, +y) is either ~0 or beyond 1e10.
POLYHEDRON_EITHER = [(MYPGON, "DistanceToIn", _PG_P2, VY)]
# :170-173 Extent == bounding box == ((0,0,-30), (80/cos(h), 40 sqrt(2)/cos(h), 40)), h = dphi/10.
_H = 0.5 * (PI / 4.0) / 5.0
POLYHEDRON_EXTENT = [(MYPGON, (0.0, 0.0, -30.0), (80.0 / math.cos(_H), 40.0 * math.sqrt(2.0) / math.cos(_H), 40.0))]

# ---- Tessellated: test/unit_tests/TestTessellated.cpp ------------------------
def trd_like_triangles(x1, x2, y1, y2, z):
    """CreateTrdLikeTessellated (TestTessellated.cpp:15-37): six quadrilateral
    facets, each added as the triangles (v0, v1, v2) and (v0, v2, v3)
    (TessellatedStruct::AddQuadrilateralFacet, VecGeom/volumes/TessellatedStruct.h:626-650)."""
    quads = [
        ((-x2, y2, z), (-x2, -y2, z), (x2, -y2, z), (x2, y2, z)),
        ((-x1, y1, -z), (x1, y1, -z), (x1, -y1, -z), (-x1, -y1, -z)),
        ((-x2, -y2, z), (-x1, -y1, -z), (x1, -y1, -z), (x2, -y2, z)),
        ((x2, -y2, z), (x1, -y1, -z), (x1, y1, -z), (x2, y2, z)),
        ((x2, y2, z), (x1, y1, -z), (-x1, y1, -z), (-x2, y2, z)),
        ((-x2, y2, z), (-x1, y1, -z), (-x1, -y1, -z), (-x2, -y2, z)),
    ]
    tris = []
    for a, b, c, d in quads:
        tris.append((a, b, c))
        tris.append((a, c, d))
    return np.array(tris, dtype=np.float64)


TSL1 = (trd_like_triangles(20.0, 20.0, 30.0, 30.0, 40.0),)
TSL2 = (trd_like_triangles(10.0, 30.0, 20.0, 40.0, 40.0),)
TSL3 = (trd_like_triangles(0.14999999999999999, 0.14999999999999999, 24.707000000000001, 24.707000000000001,
                           22.699999999999999),)
_COSA, _SINA = 4 / math.sqrt(17.0), 1 / math.sqrt(17.0)
_PBIG = {"x": (100.0, 0.0, 0.0), "y": (0.0, 100.0, 0.0), "z": (0.0, 0.0, 100.0),
         "mx": (-100.0, 0.0, 0.0), "my": (0.0, -100.0, 0.0), "mz": (0.0, 0.0, -100.0)}

TESS_KATS = [
    # :55-77 Inside
    (TSL1, "Inside", O, None, ("inside", 1)),
    (TSL1, "Inside", _PBIG["z"], None, ("inside", 3)),
    (TSL1, "Inside", (20.0, 0, 0), None, ("inside", 2)),
    (TSL1, "Inside", (0, 30.0, 0), None, ("inside", 2)),
    (TSL1, "Inside", (0, 0, 40.0), None, ("inside", 2)),
    (TSL1, "Inside", (20.0, 30.0, 40.0), None, ("inside", 2)),
    (TSL1, "Inside", (-20.0, 30.0, 40.0), None, ("inside", 2)),
    (TSL1, "Inside", (20.0, -30.0, 40.0), None, ("inside", 2)),
    (TSL1, "Inside", (20.0, 30.0, -40.0), None, ("inside", 2)),
    (TSL1, "Inside", (20.0, 30.0, 0.0), None, ("inside", 2)),
    (TSL1, "Inside", (0.0, 30.0, 40.0), None, ("inside", 2)),
    (TSL1, "Inside", (20.0, 0.0, 40.0), None, ("inside", 2)),
    (TSL1, "Inside", (-20.0, -30.0, -40.0), None, ("inside", 2)),
    (TSL2, "Inside", O, None, ("inside", 1)),
    (TSL2, "Inside", _PBIG["z"], None, ("inside", 3)),
    (TSL2, "Inside", (20.0, 0, 0), None, ("inside", 2)),
    (TSL2, "Inside", (0, 30.0, 0), None, ("inside", 2)),
    (TSL2, "Inside", (0, 0, 40.0), None, ("inside", 2)),
    # :97-113 SafetyToOut
    (TSL1, "SafetyToOut", O, None, ("eq", 20.0)),
    (TSL1, "SafetyToOut", VX, None, ("eq", 19.0)),
    (TSL1, "SafetyToOut", VY, None, ("eq", 20.0)),
    (TSL1, "SafetyToOut", VZ, None, ("eq", 20.0)),
    (TSL2, "SafetyToOut", O, None, ("eq", 20 * _COSA)),
    (TSL2, "SafetyToOut", VX, None, ("eq", 19 * _COSA)),
    (TSL2, "SafetyToOut", VY, None, ("eq", 20 * _COSA)),
    (TSL2, "SafetyToOut", VZ, None, ("eq", 20 * _COSA + _SINA)),
    # :115-165 DistanceToOut
    (TSL1, "DistanceToOut", O, VX, ("eq", 20.0)),
    (TSL1, "DistanceToOut", O, VMX, ("eq", 20.0)),
    (TSL1, "DistanceToOut", O, VY, ("eq", 30.0)),
    (TSL1, "DistanceToOut", O, VMY, ("eq", 30.0)),
    (TSL1, "DistanceToOut", O, VZ, ("eq", 40.0)),
    (TSL1, "DistanceToOut", O, VMZ, ("eq", 40.0)),
    (TSL1, "DistanceToOut", O, VXY, ("eq", math.sqrt(800.0))),
    (TSL1, "DistanceToOut", (20.0, 0, 0), VX, ("eq", 0.0)),
    (TSL1, "DistanceToOut", (-20.0, 0, 0), VMX, ("eq", 0.0)),
    (TSL1, "DistanceToOut", (0, 30.0, 0), VY, ("eq", 0.0)),
    (TSL1, "DistanceToOut", (0, -30.0, 0), VMY, ("eq", 0.0)),
    (TSL1, "DistanceToOut", (0, 0, 40.0), VZ, ("eq", 0.0)),
    (TSL1, "DistanceToOut", (0, 0, -40.0), VMZ, ("eq", 0.0)),
    (TSL2, "DistanceToOut", O, VX, ("eq", 20.0)),
    (TSL2, "DistanceToOut", O, VMX, ("eq", 20.0)),
    (TSL2, "DistanceToOut", O, VY, ("eq", 30.0)),
    (TSL2, "DistanceToOut", O, VMY, ("eq", 30.0)),
    (TSL2, "DistanceToOut", O, VZ, ("eq", 40.0)),
    (TSL2, "DistanceToOut", O, VMZ, ("eq", 40.0)),
    (TSL2, "DistanceToOut", O, VXY, ("eq", math.sqrt(800.0))),
    (TSL2, "DistanceToOut", (20.0, 0, 0), VX, ("eq", 0.0)),
    (TSL2, "DistanceToOut", (-20.0, 0, 0), VMX, ("eq", 0.0)),
    (TSL2, "DistanceToOut", (0, 30.0, 0), VY, ("eq", 0.0)),
    (TSL2, "DistanceToOut", (0, -30.0, 0), VMY, ("eq", 0.0)),
    (TSL2, "DistanceToOut", (0, 0, 40.0), VZ, ("eq", 0.0)),
    (TSL2, "DistanceToOut", (0, 0, -40.0), VMZ, ("eq", 0.0)),
    # :166-190 SafetyToIn
    (TSL1, "SafetyToIn", _PBIG["x"], None, ("eq", 80.0)),
    (TSL1, "SafetyToIn", _PBIG["mx"], None, ("eq", 80.0)),
    (TSL1, "SafetyToIn", _PBIG["y"], None, ("eq", 70.0)),
    (TSL1, "SafetyToIn", _PBIG["my"], None, ("eq", 70.0)),
    (TSL1, "SafetyToIn", _PBIG["z"], None, ("eq", 60.0)),
    (TSL1, "SafetyToIn", _PBIG["mz"], None, ("eq", 60.0)),
    (TSL2, "SafetyToIn", _PBIG["x"], None, ("eq", 80 * _COSA)),
    (TSL2, "SafetyToIn", _PBIG["mx"], None, ("eq", 80 * _COSA)),
    (TSL2, "SafetyToIn", _PBIG["y"], None, ("eq", 70 * _COSA)),
    (TSL2, "SafetyToIn", _PBIG["my"], None, ("eq", 70 * _COSA)),
    (TSL2, "SafetyToIn", _PBIG["z"], None, ("eq", 60.0)),
    (TSL2, "SafetyToIn", _PBIG["mz"], None, ("eq", 60.0)),
    # :191-230 DistanceToIn
    (TSL1, "DistanceToIn", _PBIG["x"], VMX, ("eq", 80.0)),
    (TSL1, "DistanceToIn", _PBIG["mx"], VX, ("eq", 80.0)),
    (TSL1, "DistanceToIn", _PBIG["y"], VMY, ("eq", 70.0)),
    (TSL1, "DistanceToIn", _PBIG["my"], VY, ("eq", 70.0)),
    (TSL1, "DistanceToIn", _PBIG["z"], VMZ, ("eq", 60.0)),
    (TSL1, "DistanceToIn", _PBIG["mz"], VZ, ("eq", 60.0)),
    (TSL1, "DistanceToIn", _PBIG["x"], VXY, ("eq", KINF)),
    (TSL1, "DistanceToIn", _PBIG["mx"], VXY, ("eq", KINF)),
    (TSL2, "DistanceToIn", _PBIG["x"], VMX, ("eq", 80.0)),
    (TSL2, "DistanceToIn", _PBIG["mx"], VX, ("eq", 80.0)),
    (TSL2, "DistanceToIn", _PBIG["y"], VMY, ("eq", 70.0)),
    (TSL2, "DistanceToIn", _PBIG["my"], VY, ("eq", 70.0)),
    (TSL2, "DistanceToIn", _PBIG["z"], VMZ, ("eq", 60.0)),
    (TSL2, "DistanceToIn", _PBIG["mz"], VZ, ("eq", 60.0)),
    (TSL2, "DistanceToIn", _PBIG["x"], VXY, ("eq", KINF)),
    (TSL2, "DistanceToIn", _PBIG["mx"], VXY, ("eq", KINF)),
    (TSL3, "DistanceToIn", (0.15000000000000185, -22.048743592955137, 2.4268539333219472),
     unit((-0.76165597579890043, 0.64364445891356026, -0.074515708658524193)), ("eq", 0.0)),
]
# :232-244 Extent
TESS_EXTENT = [(TSL1, (-20.0, -30.0, -40.0), (20.0, 30.0, 40.0)), (TSL2, (-30.0, -40.0, -40.0), (30.0, 40.0, 40.0))]


# ---- orb: test/unit_tests/TestOrb.cpp -------------------------------------------
ORB_B1, ORB_B2, ORB_B3 = (9.0,), (6.0,), (3.0,)
_ORB_PO = (-8.363470934547895, 2.754420966126675, -2.665617952433236)
_ORB_DO = tuple(c / 9.2 for c in _ORB_PO)
_ORB_DI = tuple(-c / 9.2 for c in _ORB_PO)
_ORB_POT = (8.884242447222299, 0.134875592787852, -1.432495973274375)
_ORB_PI = (-3.618498437781364, 2.401810108299175, -6.718465394675017)

ORB_KATS = (
    # :110-133 DistanceToOut from the origin along the axes
    [(ORB_B1, "DistanceToOut", O, v, ("eq", 9.0)) for v in (VX, VMX, VY, VMY, VZ, VMZ)]
    # :134-181 from the surface points, outwards and across
    + [(ORB_B1, "DistanceToOut", p, v, ("eq", w)) for p, v, w in (
        ((9.0, 0, 0), VX, 0.0), ((9.0, 0, 0), VMX, 18.0), ((-9.0, 0, 0), VX, 18.0), ((-9.0, 0, 0), VMX, 0.0),
        ((0, 9.0, 0), VY, 0.0), ((0, 9.0, 0), VMY, 18.0), ((0, -9.0, 0), VY, 18.0), ((0, -9.0, 0), VMY, 0.0),
        ((0, 0, 9.0), VZ, 0.0), ((0, 0, 9.0), VMZ, 18.0), ((0, 0, -9.0), VZ, 18.0), ((0, 0, -9.0), VMZ, 0.0))]
    # :184-199 Inside
    + [(ORB_B1, "Inside", p, None, ("inside", c)) for p, c in (
        (O, 1), ((100.0, 0, 0), 3), ((0, 100.0, 0), 3), ((0, 0, 100.0), 3),
        ((9.0, 0, 0), 2), ((0, 9.0, 0), 2), ((0, 0, 9.0), 2))]
    + [(ORB_B2, "Inside", p, None, ("inside", c)) for p, c in (
        (O, 1), ((9.0, 0, 0), 3), ((0, 9.0, 0), 3), ((0, 0, 9.0), 3),
        ((6.0, 0, 0), 2), ((0, 6.0, 0), 2), ((0, 0, 6.0), 2))]
    # :201-223 safeties
    + [(ORB_B1, "SafetyToOut", p, None, ("eq", w)) for p, w in ((O, 9.0), (VX, 8.0), (VY, 8.0), (VZ, 8.0))]
    + [(ORB_B1, "SafetyToIn", p, None, ("eq", 91.0)) for p in (
        (100.0, 0, 0), (-100.0, 0, 0), (0, 100.0, 0), (0, -100.0, 0), (0, 0, 100.0), (0, 0, -100.0))]
    # :225-276 DistanceToIn
    + [(ORB_B1, "DistanceToIn", p, v, ("eq", 91.0)) for p, v in (
        ((100.0, 0, 0), VMX), ((-100.0, 0, 0), VX), ((0, 100.0, 0), VMY), ((0, -100.0, 0), VY),
        ((0, 0, 100.0), VMZ), ((0, 0, -100.0), VZ))]
    + [(ORB_B1, "DistanceToIn", (100.0, 0, 0), VXY, ("inf",)), (ORB_B1, "DistanceToIn", (-100.0, 0, 0), VXY, ("inf",)),
       (ORB_B2, "DistanceToIn", (9.0, 0, 12.0), VXMZ, ("inf",)), (ORB_B2, "DistanceToIn", (12.0, 9.0, 0), VMXY, ("inf",)),
       (ORB_B2, "DistanceToIn", (12.0, 9.0, 0), VMX, ("inf",)), (ORB_B2, "DistanceToIn", (8.0, 0, 0), VMX, ("eq", 2.0)),
       (ORB_B2, "DistanceToIn", (7.0, 5.0, 0), VMX, ("eq", 3.6833752)),
       (ORB_B1, "DistanceToIn", (-25.0, -35.0, 0), VX, ("inf",)),
       (ORB_B1, "DistanceToIn", (-25.0, -35.0, 0), VY, ("inf",)),
       # :278-330 points near and inside the surface
       (ORB_B1, "DistanceToIn", _ORB_PO, _ORB_DO, ("inf",)),
       (ORB_B1, "DistanceToOut", _ORB_PO, _ORB_DO, ("le", 0.0)),
       (ORB_B1, "DistanceToIn", _ORB_PO, _ORB_DI, ("eq", 0.2)),
       (ORB_B1, "DistanceToIn", _ORB_POT, tuple(c / 9 for c in _ORB_POT), ("inf",)),
       (ORB_B1, "DistanceToIn", _ORB_POT, tuple(-c / 9 for c in _ORB_POT), ("eq", 0.0)),
       (ORB_B1, "DistanceToIn", _ORB_PI, tuple(c / 8 for c in _ORB_PI), ("le", -1e-300)),
       (ORB_B1, "DistanceToOut", _ORB_PI, tuple(c / 8 for c in _ORB_PI), ("eq", 1.0)),
       (ORB_B1, "DistanceToIn", _ORB_PI, tuple(-c / 8 for c in _ORB_PI), ("le", -1e-300)),
       (ORB_B1, "DistanceToOut", _ORB_PI, tuple(-c / 8 for c in _ORB_PI), ("eq", 17.0)),
       # :337-341
       (ORB_B3, "DistanceToIn", (-1.704652541027918744, 0.92532982039202238411, -2.2886512267834184797),
        (0.55992127966252225324, -0.73207465430510332283, 0.38801399601708530529), ("eq", 0.0))]
)


# ---- Normal (VPlacedVolume::Normal): (family, shape, point, expected normal, expected valid or None) ----
_P2, _P3 = 1.0 / math.sqrt(2.0), 1.0 / math.sqrt(3.0)
_BOX1 = (20.0, 30.0, 40.0)
_FACES = [((20, 0, 0), (1, 0, 0)), ((-20, 0, 0), (-1, 0, 0)), ((0, 30, 0), (0, 1, 0)), ((0, -30, 0), (0, -1, 0)),
          ((0, 0, 40), (0, 0, 1)), ((0, 0, -40), (0, 0, -1)), ((0, 25, 40), (0, 0, 1)), ((0, 25, -40), (0, 0, -1))]
_EDGES = [((20, 30, 0), (_P2, _P2, 0)), ((-20, -30, 0), (-_P2, -_P2, 0)), ((20, -30, 0), (_P2, -_P2, 0)),
          ((-20, 30, 0), (-_P2, _P2, 0)), ((20, 0, 40), (_P2, 0, _P2)), ((-20, 0, -40), (-_P2, 0, -_P2)),
          ((20, 0, -40), (_P2, 0, -_P2)), ((-20, 0, 40), (-_P2, 0, _P2)), ((0, 30, 40), (0, _P2, _P2)),
          ((0, -30, -40), (0, -_P2, -_P2)), ((0, 30, -40), (0, _P2, -_P2)), ((0, -30, 40), (0, -_P2, _P2))]
_CORNERS = [((sx * 20, sy * 30, sz * 40), (sx * _P3, sy * _P3, sz * _P3))
            for sx in (1, -1) for sy in (1, -1) for sz in (1, -1)]
_COSA, _SINA = 4 / math.sqrt(17.0), 1 / math.sqrt(17.0)
_T4 = (45.0, 50.0, 50.0, 0.5 * PI, 0.5 * PI)
_T4_CASES = [((0., 50., 0.), (_P2, _P2, 0.)), ((0., 45., 0.), (_P2, -_P2, 0.)), ((0., 45., 50.), (_P3, -_P3, _P3)),
             ((0., 45., -50.), (_P3, -_P3, -_P3)), ((-50., 0., -50.), (-_P3, -_P3, -_P3)),
             ((-50., 0., 0.), (-_P2, -_P2, 0.))]
NORMAL_KATS = (
    # TestBox.cpp:81-167
    [("box", _BOX1, p, n, True) for p, n in _FACES + _EDGES + _CORNERS]
    # TestTrd.cpp:115-226: trd1 is the box, trd2 has inclined x and y faces
    + [("trd", (20.0, 20.0, 30.0, 30.0, 40.0), p, n, True) for p, n in _FACES + _EDGES + _CORNERS]
    + [("trd", (10.0, 30.0, 20.0, 40.0, 40.0), p, n, None) for p, n in [
        ((20, 0, 0), (_COSA, 0, -_SINA)), ((-20, 0, 0), (-_COSA, 0, -_SINA)), ((0, 30, 0), (0, _COSA, -_SINA)),
        ((0, -30, 0), (0, -_COSA, -_SINA)), ((0, 0, 40), (0, 0, 1)), ((0, 0, -40), (0, 0, -1)),
        ((0, 25, 40), (0, 0, 1)), ((0, 25, -40), (0, 0, -1))]]
    # TestTube.cpp:160-175
    + [("tube", (0.0, 50.0, 50.0, 0.0, 2 * PI), (50, 0, 0), (1, 0, 0), None)]
    + [("tube", _T4, p, n, None) for p, n in _T4_CASES]
    + [("tube", (0.0, 50.0, 50.0, 0.5 * PI, 1.5 * PI), (0., 0., 0.), (_P2, _P2, 0.), None)]
    # TestCons.cpp:277-316 (printed checks): cn1 is the tube t4 as a cone
    + [("cone", (45.0, 50.0, 45.0, 50.0, 50.0, 0.5 * PI, 0.5 * PI), p, n, None) for p, n in _T4_CASES]
    + [("cone", (45.0, 50.0, 45.0, 50.0, 50.0, 0.5 * PI, 1.5 * PI), (50., 0., 0.), (_P2, _P2, 0.), None),
       ("cone", (0.0, 150.0, 0.0, 150.0, 50.0, 0.0, 2 * PI), (0., 0., 50.), (0., 0., 1.), None)]
    # TestOrb.cpp:87-108
    + [("orb", (9.0,), p, n, True) for p, n in [((9, 0, 0), (1, 0, 0)), ((0, 9, 0), (0, 1, 0)), ((0, 0, 9), (0, 0, 1)),
                                               ((-9, 0, 0), (-1, 0, 0)), ((0, -9, 0), (0, -1, 0)),
                                               ((0, 0, -9), (0, 0, -1))]]
    # TestPolycone.cpp:190-199
    + [("polycone", PSIMPLE, p, n, None) for p, n in [((70, 0, -5), (1, 0, 0)), ((-70, 0, -5), (-1, 0, 0)),
                                                      ((0, 70, -5), (0, 1, 0)), ((0, 0, 10), (0, 0, 1)),
                                                      ((0, 0, -10), (0, 0, -1))]]
)
# TestPolycone.cpp:171-186: away from the surface Normal is invalid, with a unit vector
NORMAL_OFF_SURFACE = [("polycone", PSIMPLE, p) for p in [(70, 70, -5), (70, 70, -10), (0, 0, 100), (0, 0, -100)]]


# ---- Sphere (TestSphere.cpp:43-320): (shape, method, point, direction, expectation) ----
_SB4 = (6.0, 8.0, 0.0, 2 * PI, 0.0, PI)
_SB5 = (6.0, 8.0, 0.0, PI, 0.0, PI)
_SB6 = (6.0, 8.0, PI / 6, PI / 3, 0.0, PI)
_SB7 = (6.0, 8.0, PI / 6, PI / 6, 0.0, PI / 6)
_SB8 = (6.0, 8.0, PI / 6, PI / 6, 0.0, PI)
_SB9 = (6.0, 8.0, PI / 6, PI / 6, PI / 6, PI / 6)
_BIG = [(100, 0, 0), (0, 100, 0), (0, 0, 100), (-100, 0, 0), (0, -100, 0), (0, 0, -100)]
_PO_31_45 = (3.038887886090341, 3.240358815458723, 5.409735221140894)
SPHERE_KATS = [
    (_SB4, "Inside", (0.337317535387103, 6.265557026812151, 3.103064940359008), None, ("inside", 1)),
    (_SB4, "Inside", (0.162722224118177, -8.624868144449298, -2.566158796458579), None, ("inside", 3)),
    (_SB4, "Inside", (-2.393620784732096, -4.753290873255242, 5.973006396542139), None, ("inside", 2)),
    (_SB4, "Inside", (-4.231374871926247, -4.183095200116632, -0.772775025466741), None, ("inside", 2)),
    (_SB4, "SafetyToIn", (100, 0, 0), None, ("eq", 92.0)),
    (_SB4, "SafetyToOut", (0, 0, 0), None, ("le", 0.0)),
    (_SB4, "SafetyToOut", (0.162722224118177, -8.624868144449298, -2.566158796458579), None, ("le", 0.0)),
    (_SB4, "SafetyToOut", (9, 0, 0), None, ("le", 0.0)),
    (_SB4, "SafetyToOut", (0, 9, 0), None, ("le", 0.0)),
    (_SB4, "SafetyToOut", (0, 0, 9), None, ("le", 0.0)),
] + [(_SB4, "SafetyToOut", p, None, ("eq", 0.5)) for p in [(6.5, 0, 0), (0, 6.5, 0), (0, 0, 6.5), (7.5, 0, 0),
                                                           (0, 7.5, 0), (0, 0, 7.5)]] + [
    (_SB4, "SafetyToIn", p, None, ("le", 0.0)) for p in [(7.5, 0, 0), (0, 7.5, 0), (0, 0, 7.5),
                                                         (3.796560684305335, -6.207283535497058, 2.519078815824183)]
] + [
    (_SB5, "Inside", (-6.354024677915919, 2.724956501738179, -1.095893451057325), None, ("inside", 1)),
    (_SB5, "Inside", (-2.523183636323228, -3.188902592924771, -5.697757856405303), None, ("inside", 3)),
    (_SB5, "Inside", (-6.975035234148935, 0.000000000583791, -0.590663594934467), None, ("inside", 2)),
    (_SB6, "Inside", (5.343263338886987, 3.536603816131425, -2.818150162612158), None, ("inside", 1)),
    (_SB6, "Inside", (6.043467293054153, 0.461937628581358, 3.501873313683026), None, ("inside", 3)),
    (_SB6, "Inside", (0.139868946641446, 0.080753374000628, 6.998136578429498), None, ("inside", 2)),
    (_SB6, "Inside", (4.589551141500899, 2.649778586424563, 4.573258549707597), None, ("inside", 2)),
    (_SB8, "SafetyToIn", (0, 5, 0), None, ("eq", 2.5)),
] + [(_SB8, "SafetyToIn", p, None, ("eq", 92.0)) for p in _BIG] + [
    (_SB8, "SafetyToOut", p, None, ("le", 0.0)) for p in _BIG] + [
    (_SB8, "SafetyToOut", (0.041433182037376, 3.902131470711286, -6.637895244481554), None, ("le", 0.0)),
    (_SB8, "SafetyToIn", (0.041433182037376, 3.902131470711286, -6.637895244481554), None, ("eq", 1.91518354715165)),
    (_SB7, "Inside", (2.051950506806283, 1.301790637821730, 6.564666042754737), None, ("inside", 1)),
    (_SB7, "Inside", _PO_31_45, None, ("inside", 3)),
    (_SB7, "Inside", (2.421058285775014, 2.527543625149363, 6.062177826478448), None, ("inside", 2)),
    (_SB7, "Inside", (2.983684745081379, 1.829651698111631, 6.062177826950138), None, ("inside", 2)),
] + [(_SB9, "SafetyToIn", p, None, ("eq", 92.0)) for p in _BIG] + [
    (_SB9, "SafetyToOut", p, None, ("le", 0.0)) for p in _BIG] + [
    (_SB7, "SafetyToIn", (2.051950506806283, 1.301790637821730, 6.564666042754737), None, ("le", 0.0)),
    (_SB7, "SafetyToIn", _PO_31_45, None, ("eq", 1.142348737370068)),
]
# TestSphere.cpp:188-205, 290-293, 313-314 (the last two "verified with Geant4", to 6 digits)
SPHERE_NORMAL_KATS = [
    (_SB4, (8, 0, 0), (1, 0, 0), True), (_SB4, (0, 8, 0), (0, 1, 0), True), (_SB4, (0, 0, 8), (0, 0, 1), True),
    (_SB4, (-8, 0, 0), (-1, 0, 0), True), (_SB4, (0, -8, 0), (0, -1, 0), True),
    (_SB4, (0, 0, -8), (0, 0, -1), True),
]
SPHERE_NORMAL_APPROX = [
    ((6.0, 8.0, PI / 3, PI / 3, 0.0, PI), (2.821923096621196, 6.085843652625694, 5.999938089059868),
     (0.866025, -0.5, 0.0)),
    (_SB7, _PO_31_45, (-0.434127, -0.462908, -0.772819)),
]

# TestSphere.cpp:460-573 (the "Solid Sphere_t" family)
_S1 = (0.0, 50.0, 0.0, 2 * PI, 0.0, PI)
_SN1 = (0.0, 50.0, 0.5 * PI, 1.5 * PI, 0.0, PI)
_SN11 = (0.0, 50.0, 0.0, 2 * PI, 0.0, 0.5 * PI)
_SN12 = (0.0, 50.0, 0.0, 2 * PI, 0.0, 0.25 * PI)
_SN13 = (0.0, 50.0, 0.0, 2 * PI, 0.75 * PI, 0.25 * PI)
_S2 = (45.0, 50.0, 0.0, 2 * PI, 0.0, PI)
_SN2 = (45.0, 50.0, 0.5 * PI, 0.5 * PI, 0.0, PI)
_S3 = (45.0, 50.0, 0.0, 2 * PI, PI / 4, 0.5 * PI)
_S4 = (45.0, 50.0, -PI / 4, 0.5 * PI, 0.0, 2 * PI)
_S41 = (5.0, 50.0, -PI, 1.5 * PI, 0.0, 2 * PI)
_S42 = (5.0, 50.0, -PI / 2, 1.5 * PI, 0.0, 2 * PI)
_S5 = (45.0, 50.0, -PI / 4, 0.5 * PI, PI / 4, 0.5 * PI)
_S9 = (0.0, 410.0, 0.0, 2 * PI, 0.5 * PI, 0.5 * PI)
_R2 = math.sqrt(2.0)
_PONRMIN2, _PONRMAX2 = (45 / _R2, 45 / _R2, 0.0), (50 / _R2, 50 / _R2, 0.0)
_PONTHETA1 = (0.0, 48 * math.sin(PI / 4), 48 * math.cos(PI / 4))
_PONTHETA2 = (0.0, 48 * math.sin(PI / 4), -48 * math.cos(PI / 4))
_PONPHI1, _POVERPHI, _PINPHI = (48 / _R2, -48 / _R2, 0.0), (-48 / _R2, 48 / _R2, 0.0), (48 * 0.866, -24.0, 0.0)
SPHERE_KATS += [
    (_S1, "Inside", (0, 0, 0), None, ("inside", 1)), (_S2, "Inside", (0, 0, 0), None, ("inside", 3)),
    (_S2, "Inside", _PONRMIN2, None, ("inside", 2)), (_S2, "Inside", _PONRMAX2, None, ("inside", 2)),
    (_S3, "Inside", _PONTHETA1, None, ("inside", 2)), (_S3, "Inside", _PONTHETA2, None, ("inside", 2)),
    (_S4, "Inside", _PONPHI1, None, ("inside", 2)), (_S4, "Inside", _POVERPHI, None, ("inside", 3)),
    (_S4, "Inside", _PINPHI, None, ("inside", 1)), (_S5, "Inside", (0, 0, 100), None, ("inside", 3)),
    (_S41, "Inside", (-30, 0, 0), None, ("inside", 2)), (_S42, "Inside", (-30, 0, 0), None, ("inside", 2)),
    (_SN11, "Inside", (0, 0, 0), None, ("inside", 2)), (_SN12, "Inside", (0, 0, 0), None, ("inside", 2)),
    (_SN13, "Inside", (0, 0, 0), None, ("inside", 2)), (_SN13, "Inside", (0, 0, -30), None, ("inside", 1)),
    (_SN12, "Inside", (0, 0, 30), None, ("inside", 1)), (_SN12, "Inside", (30, 0, 0), None, ("inside", 3)),
    (_SN11, "Inside", (30, 0, 0), None, ("inside", 2)),
    (_S1, "SafetyToOut", (0, 0, 0), None, ("eq", 50.0)), (_S1, "SafetyToOut", (50, 0, 0), None, ("eq", 0.0)),
    (_S1, "SafetyToOut", (30, 0, 0), None, ("eq", 20.0)), (_S2, "SafetyToOut", (0, 0, 0), None, ("le", 0.0)),
    (_S2, "SafetyToOut", (50, 0, 0), None, ("eq", 0.0)), (_S2, "SafetyToOut", (45, 0, 0), None, ("eq", 0.0)),
    (_SN11, "SafetyToOut", (0, 0, 0), None, ("eq", 0.0)), (_SN11, "SafetyToOut", (30, 0, 0), None, ("eq", 0.0)),
    (_SN12, "SafetyToOut", (0, 0, 0), None, ("eq", 0.0)), (_SN12, "SafetyToOut", (0, 0, 30), None, ("eq", 20.0)),
    (_SN13, "SafetyToOut", (0, 0, -30), None, ("eq", 20.0)), (_S3, "SafetyToOut", (0, 0, 0), None, ("le", 0.0)),
    (_S9, "SafetyToOut", (0, 0, 0), None, ("eq", 0.0)), (_S9, "SafetyToOut", (30, 0, 0), None, ("eq", 0.0)),
    (_S41, "SafetyToOut", (0, 0, 0), None, ("le", 0.0)), (_S41, "SafetyToOut", (50, 0, 0), None, ("eq", 0.0)),
    (_S1, "DistanceToOut", (0, 0, 30), VZ, ("eq", 20.0)), (_S1, "DistanceToOut", (50, 0, 0), VX, ("eq", 0.0)),
    (_S2, "DistanceToOut", (45, 0, 0), VX, ("eq", 5.0)), (_S2, "DistanceToOut", _PONRMAX2, (1 / _R2, 1 / _R2, 0.0),
                                                          ("eq", 0.0)),
    (_S1, "DistanceToOut", (0, 0, 0), VX, ("eq", 50.0)), (_S1, "DistanceToOut", (0, 0, 0), VMX, ("eq", 50.0)),
    (_S1, "DistanceToOut", (0, 0, 0), VY, ("eq", 50.0)), (_S1, "DistanceToOut", (0, 0, 0), VMY, ("eq", 50.0)),
]
_P3S = 1 / math.sqrt(3.0)
SPHERE_NORMAL_KATS += [
    (_SN1, (0.0, 0.0, 50.0), (_P3S, _P3S, _P3S), None), (_SN1, (0.0, 0.0, 0.0), (1 / _R2, 1 / _R2, 0.0), None),
    (_SN11, (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), None), (_SN12, (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), None),
    (_SN13, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), None), (_SN2, (-45.0, 0.0, 0.0), (1 / _R2, -1 / _R2, 0.0), None),
    (_S1, (50.0, 0.0, 0.0), (1.0, 0.0, 0.0), None),
]

# TestTorus2.cpp:17-255 (rmin, rmax, rtor, sphi, dphi); Rtor 100, Rmax 90, Rmin 10
_T1 = (0.0, 90.0, 100.0, 0.0, 2 * PI)
_T2 = (10.0, 90.0, 100.0, 0.0, PI / 2)
_TN2 = (10.0, 90.0, 100.0, PI / 2, PI / 2)
_TN3 = (10.0, 90.0, 100.0, PI / 2, 3 * PI / 2)
_T3 = (40.0, 90.0, 100.0, PI / 2 - PI / 24, PI / 12)
_T6 = (100.0, 150.0, 200.0, 0.0, PI / 3)
_CMS_EE_COOL3 = (0.0, 6.0, 1640.0, 0.09599, 1.4312)
_TBX, _TBY, _TBZ = (240.0, 0.0, 0.0), (0.0, 240.0, 0.0), (0.0, 0.0, 240.0)
_TBMX, _TBMY, _TBMZ = (-240.0, 0.0, 0.0), (0.0, -240.0, 0.0), (0.0, 0.0, -240.0)
_TONRMAX, _TONRMIN, _TONRTOR = (190.0, 0.0, 0.0), (0.0, 110.0, 0.0), (0.0, 100.0, 0.0)
_TONPHI1, _TONPHI2 = (100 / _R2, 100 / _R2, 0.0), (-100 / _R2, 100 / _R2, 0.0)
_TONPHI12 = (190 / _R2, 190 / _R2, 0.0)
_VXY = (1 / _R2, 1 / _R2, 0.0)
TORUS_KATS = [
    (_CMS_EE_COOL3, "Inside", (1230.993171, 1078.075896, -2.947036), None, ("inside", 1)),
    (_T1, "Inside", O, None, ("inside", 3)), (_T1, "Inside", _TBX, None, ("inside", 3)),
    (_T1, "Inside", _TONRMAX, None, ("inside", 2)), (_T2, "Inside", _TONRMIN, None, ("inside", 2)),
    (_T2, "Inside", _TBX, None, ("inside", 3)), (_T2, "Inside", _TBY, None, ("inside", 3)),
    (_T2, "Inside", _TONPHI1, None, ("inside", 3)), (_T2, "Inside", _TONPHI2, None, ("inside", 3)),
    (_T2, "Inside", (20.0, 0.0, 0.0), None, ("inside", 2)), (_T2, "Inside", (0.0, 20.0, 0.0), None, ("inside", 2)),
    (_T6, "Inside", (70.75950555416668, -3.552713678800501e-15, 22.37458414788935), None, ("inside", 2)),
    (_T1, "SafetyToOut", _TONRMIN, None, ("eq", 80.0)), (_T1, "SafetyToOut", _TONRMAX, None, ("eq", 0.0)),
    (_T2, "SafetyToOut", (20.0, 0.0, 0.0), None, ("eq", 0.0)), (_T2, "SafetyToOut", (0.0, 20.0, 0.0), None, ("eq", 0.0)),
    (_T1, "DistanceToOut", _TONRMAX, VX, ("eq", 0.0)), (_T1, "DistanceToOut", (130.0, 0.0, 0.0), VX, ("eq", 60.0)),
    (_T1, "DistanceToOut", _TONRMIN, VY, ("eq", 80.0)), (_T1, "DistanceToOut", _TONRMIN, VMY, ("eq", 100.0)),
    (_T1, "DistanceToOut", (100.0, 0.0, 0.0), unit((0.00001, 0.000001, 1.0)), ("eq", 90.0)),
    (_T2, "DistanceToOut", (7.07106781186547524400844362, 7.07106781186547524400844362, 0.0), VMX, ("eq", 0.0)),
] + [(_T1, "SafetyToIn", p, None, ("eq", 50.0)) for p in (_TBX, _TBMX, _TBY, _TBMY)] + [
    (_T1, "DistanceToIn", _TBX, VMX, ("eq", 50.0)), (_T1, "DistanceToIn", _TBMX, VX, ("eq", 50.0)),
    (_T1, "DistanceToIn", _TBY, VMY, ("eq", 50.0)), (_T1, "DistanceToIn", _TBMY, VY, ("eq", 50.0)),
    (_T1, "DistanceToIn", _TBZ, VMZ, ("inf",)), (_T1, "DistanceToIn", _TBMZ, VZ, ("inf",)),
    (_T1, "DistanceToIn", _TBX, _VXY, ("inf",)),
    (_T2, "DistanceToIn", O, VY, ("eq", 10.0)), (_T2, "DistanceToIn", _TONPHI12, VY, ("inf",)),
    (_T2, "DistanceToIn", _TONPHI12, VMY, ("eq", 0.0)),
    (_T2, "DistanceToIn", _TONPHI1, VY, ("eq", 13.550819613108856743)),
    (_T2, "DistanceToIn", _TONRMIN, VY, ("inf",)), (_T2, "DistanceToIn", _TONRMIN, VMY, ("inf",)),
] + [(_T3, "DistanceToIn", _TONRTOR, v, ("eq", 40.0)) for v in (VY, VMY, VZ, VMZ)] + [
    (_T3, "DistanceToIn", _TONRTOR, v, ("inf",)) for v in (VX, VMX)]
# TestTorus2.cpp:108-139: (params, point, normal)
TORUS_NORMAL_KATS = [
    (_T1, _TONRMAX, VX), (_T1, (0.0, 190.0, 0.0), VY),
    (_TN2, (0.0, 190.0, 0.0), (1 / _R2, 1 / _R2, 0.0)), (_TN2, (0.0, 110.0, 0.0), (1 / _R2, -1 / _R2, 0.0)),
    (_TN2, (0.0, 90.0, 0.0), (1 / _R2, 1 / _R2, 0.0)), (_TN2, (0.0, 10.0, 0.0), (1 / _R2, -1 / _R2, 0.0)),
    (_TN3, (100.0, 0.0, 90.0), (0.0, 1 / _R2, 1 / _R2)), (_TN3, (0.0, 100.0, 90.0), (1 / _R2, 0.0, 1 / _R2)),
    (_T2, _TONRMIN, (-1 / _R2, -1 / _R2, 0.0)), (_T2, (20.0, 0.0, 0.0), VMY), (_T2, (0.0, 20.0, 0.0), VMX),
]

# test/core/PhiWedgeTest.cpp and test/core/ThetaConeTest.cpp, through a
# sphere whose radial range holds the points (rmin 0, rmax 1000): its Inside
# and Contains are then those of its phi wedge (Wedge_Evolution.h, the same
# kernels as the Wedge these tests construct) or its theta cone
# (ThetaCone_Evolution.h, the same kernels as ThetaCone).  Sphere parameters
# (rmin, rmax, sphi, dphi, stheta, dtheta).
_KT = 1e-9  # kTolerance
_KA = 1e-9  # kAngTolerance


def _wedge(angle, zero=0.0):
    return (0.0, 1000.0, zero, angle, 0.0, PI)


def _theta(st, dt):
    return (0.0, 1000.0, 0.0, 2 * PI, st, dt)


def _cs(a):
    return (math.cos(a), math.sin(a), 0.0)


_W1 = _wedge(1.5 * PI, PI / 6)
_W2 = _wedge(PI / 3)
_A3 = PI + PI / 3
_W3 = _wedge(_A3)
_IN, _SURF, _OUT = ("inside", 1), ("inside", 2), ("inside", 3)
_YES, _NO = ("inside", 1), ("inside", 0)
WEDGE_THETA_KATS = [
    (_W1, "Inside", (4.589551141500899, 2.649778586424563, 4.573258549707597), None, _SURF),
    (_W1, "Contains", (-4.589551141500899, -2.649778586424563, 4.573258549707597), None, _YES),
    (_W1, "Inside", (-4.589551141500899, -2.649778586424563, 4.573258549707597), None, _IN),
    (_W1, "Inside", (0.0, 0.0, 0.0), None, _SURF),
] + [(_W2, "Contains", p, None, e) for p, e in [
    ((1.0, -2 * _KT, 0.0), _NO), ((1.0, 0.0, 0.0), _YES), ((1.0, -_KT, 0.0), _NO), ((-1.0, -1.0, 0.0), _NO),
    (_cs(PI / 3), _YES), (_cs(PI / 3 + _KA), _NO), (_cs(PI / 6), _YES), (_cs(PI / 2), _NO), ((-1.0, 0.0, 0.0), _NO)]
] + [(_W2, "Inside", p, None, e) for p, e in [
    ((1.0, 0.0, 0.0), _SURF), ((1.0, -0.5 * _KT, 0.0), _SURF), ((1.0, 0.5 * _KT, 0.0), _SURF), (_cs(PI / 3), _SURF),
    (_cs(PI / 3 + _KA), _OUT), (_cs(PI / 6), _IN), (_cs(PI / 2), _OUT)]
] + [(_W3, "Contains", p, None, e) for p, e in [
    ((1.0, -2 * _KT, 0.0), _NO), ((1.0, 0.0, 0.0), _YES), ((1.0, -_KT, 0.0), _NO), (_cs(_A3 / 3), _YES),
    (_cs(_A3 + _KA), _NO), (_cs(_A3 / 6), _YES), (_cs(_A3 + 0.1), _NO)]
] + [(_W3, "Inside", p, None, e) for p, e in [
    ((1.0, 0.0, 0.0), _SURF), ((1.0, -0.5 * _KT, 0.0), _SURF), ((1.0, 0.5 * _KT, 0.0), _SURF), (_cs(_A3), _SURF),
    (_cs(_A3 + _KA), _OUT), (_cs(_A3 / 6), _IN), (_cs(_A3 + 0.1), _OUT)]
]
_R3 = math.sqrt(3.0)
_TC1, _TC2, _TC3 = _theta(0.0, PI / 3), _theta(2 * PI / 3, PI / 3), _theta(0.0, 2 * PI / 3)
_TC4, _TC5, _TC6 = _theta(PI / 3, 2 * PI / 3), _theta(0.0, 0.5 * PI), _theta(0.5 * PI, 0.5 * PI)
_PX, _PY, _PZ = (30.0, 0.0, 0.0), (0.0, 30.0, 0.0), (0.0, 0.0, 30.0)


def _neg(p):
    return tuple(-c for c in p)


WEDGE_THETA_KATS += [(_TC1, "Inside", p, None, e) for p, e in [
    (VZ, _IN), (VMZ, _OUT), (VX, _OUT), (VMX, _OUT), (VY, _OUT), (VMY, _OUT), ((1.0, 0.0, 1.0), _IN),
    ((0.0, 1.0, 1.0), _IN), ((0.0, _R3, 1.0), _SURF), ((-_R3, 0.0, 1.0), _SURF), ((-_R3, 0.0, -1.0), _OUT)]
] + [(_TC1, "Contains", p, None, e) for p, e in [
    (VX, _NO), (VMX, _NO), (VY, _NO), (VMY, _NO), ((1.0, 0.0, 1.0), _YES), ((-1.0, 0.0, 1.0), _YES),
    ((0.0, 2.0, 1.0), _NO), ((0.0, -1.0, 1.0), _YES)]
] + [(_TC2, "Inside", p, None, e) for p, e in [
    (VZ, _OUT), (VMZ, _IN), (VX, _OUT), (VMX, _OUT), (VY, _OUT), (VMY, _OUT), ((1.0, 0.0, -1.0), _IN),
    ((0.0, 1.0, -1.0), _IN), ((0.0, _R3, -1.0), _SURF), ((-_R3, 0.0, -1.0), _SURF), ((-_R3, 0.0, 1.0), _OUT)]
] + [(_TC3, "Inside", p, None, e) for p, e in [(VZ, _IN), (VX, _IN), (VY, _IN), (VMZ, _OUT)]
] + [(_TC4, "Inside", p, None, e) for p, e in [(VZ, _OUT), (VX, _IN), (VY, _IN), (VMZ, _IN)]
] + [(_TC5, "Inside", p, None, e) for p, e in [
    (VZ, _IN), (VMZ, _OUT), (VX, _SURF), (VMX, _SURF), (VY, _SURF), (VMY, _SURF), (_PZ, _IN), (_neg(_PZ), _OUT),
    (_PX, _SURF), (_neg(_PX), _SURF), (_PY, _SURF), (_neg(_PY), _SURF)]
] + [(_TC6, "Inside", p, None, e) for p, e in [
    (VZ, _OUT), (VMZ, _IN), (VX, _SURF), (VMX, _SURF), (VY, _SURF), (VMY, _SURF), (_PZ, _OUT), (_neg(_PZ), _IN),
    (_PX, _SURF), (_neg(_PX), _SURF), (_PY, _SURF), (_neg(_PY), _SURF)]
]
