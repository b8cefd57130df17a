"""Seeded synthetic inputs (points and directions as SOA3D arrays).

A counter-based generator keeps inputs reproducible and independent of the
platform: stream `s` of seed `k` is the SplitMix64 sequence started from
key = mix(seed << 32 | s).  Uniform doubles take the top 53 bits.
Directions follow volumeUtilities::SampleDirection
(VecGeom/volumes/utilities/VolumeUtilities.h:113-125): normalise
(1-2u1, 1-2u2, 1-2u3).
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def key(seed: int, stream: int) -> np.uint64:
    return _mix(np.array([(int(seed) << 32) | (int(stream) & 0xFFFFFFFF)], dtype=np.uint64))[0]


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n doubles in [0, 1) from stream `stream` of `seed`, starting at `offset`."""
    k = key(seed, stream)
    ctr = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = _mix(k + ctr * _GOLDEN)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def points_in_box(seed: int, stream: int, n: int, half, scale: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Uniform points in the box |x_i - c_i| <= scale*half_i; returns (x, y, z)."""
    out = []
    for k in range(3):
        u = uniform(seed, stream * 8 + k, n)
        out.append(np.ascontiguousarray(center[k] + (2.0 * u - 1.0) * scale * half[k]))
    return tuple(out)


def isotropic_directions(seed: int, stream: int, n: int):
    """Unit directions: normalised (1-2u1, 1-2u2, 1-2u3); returns (x, y, z)."""
    u = [1.0 - 2.0 * uniform(seed, stream * 8 + 4 + k, n) for k in range(3)]
    mag = np.sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2])
    mag[mag == 0] = 1.0
    return tuple(np.ascontiguousarray(c / mag) for c in u)
