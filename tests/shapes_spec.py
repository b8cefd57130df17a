"""Shape specifications shared by the oracle and GPU parity tests.

Each spec is (builder name, constructor arguments); both vecgeom_amd.Context and
oracle.Geometry expose builders with the same names and argument order.
"""
import math

import numpy as np

# Rotation used to exercise the generic Transformation3D path: rotation about
# an oblique axis, stored in the reference's 9-element order.


def rotation_matrix(axis, angle):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    x, y, z = a
    c, s = math.cos(angle), math.sin(angle)
    C = 1 - c
    return np.array([
        [c + x * x * C, x * y * C - z * s, x * z * C + y * s],
        [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
        [z * x * C - y * s, z * y * C + x * s, c + z * z * C],
    ])


def transform12(translation, axis=(1.0, 2.0, 3.0), angle=0.7):
    return np.concatenate([np.asarray(translation, dtype=np.float64), rotation_matrix(axis, angle).reshape(-1)])


SHAPES = {
    "box_c1": ("box", (10.0, 20.0, 30.0)),
    "box_thin": ("box", (0.15, 24.707, 22.7)),
}
