"""GDML reader for the solids this library implements, with the reference's
reading rules (vgdml Middleware, persistency/gdml/source/src/Middleware.cpp).

It builds the geometry through the constructor calls of a geometry object:
a library `Context` (GPU queries and navigation) or the CPU oracle's
`Geometry`, which expose the same builder names.  Materials and auxiliary
information are read past; the hot path does not use them.

Reading rules kept from the reference:
  - A numeric attribute is read as `std::stod` reads it: the longest numeric
    prefix, so "2*pi" reads 2.  When no prefix parses, the whole string names
    a `<constant>`; an unknown name reads 0, as a default-inserted map entry
    does (Middleware::GetDoubleAttribute, :141-150).  Constants are read the
    same way, without the lookup.  `<variable>` and `<quantity>` are not
    definitions to the reference and are ignored here too.
  - Lengths are scaled by the length unit (`lunit`; `unit` on positions):
    mm -> mm_unit, cm, m, km, um, nm relative to it, anything else mm
    (:665-685).  The reference's millimetre is GeoManager's unit, 0.1 by
    default (lengths in cm); `mm_unit` sets it.  Angles: `aunit` (`unit` on
    rotations) "deg" -> pi/180, else radians (:687-695).
  - box, tube, cone, trd and trap take full lengths and are built with their
    halves (:697-990); polycone and polyhedra read every child element as a
    z plane (:794-928).
  - A tessellated solid takes triangular and quadrangular facets whose
    vertices name `<position>`s; RELATIVE vertices accumulate from the first
    (TessellatedStruct::AddTriangularFacet, VecGeom/volumes/TessellatedStruct.h:596-611),
    and a quadrangle is the triangles (0,1,2) and (0,2,3), the second added
    only when the first is valid (:624-658).
  - A physvol is placed with its position and the rotation matrix of its
    Cartesian angles (makeRotationMatrixFromCartesianAngles, :111-129;
    processPhysicalVolume, :1305-1389); the world volume is placed with the
    identity (:1446-1464).

Solids outside the implemented families, scaled solids and scaled physvols
raise GdmlError naming what was found.
"""
from __future__ import annotations

import math
import re
import xml.etree.ElementTree as ET
from typing import Dict, List, Tuple

__all__ = ["GdmlError", "load", "SUPPORTED_SOLIDS"]

SUPPORTED_SOLIDS = ("box", "tube", "cutTube", "cone", "trd", "trap", "polycone", "polyhedra", "tessellated", "orb",
                    "sphere", "torus", "subtraction", "intersection", "union")

# everything the reference reads as a solid (Middleware.cpp:64-103); the
# others are reported as unsupported rather than skipped
_REFERENCE_SOLIDS = SUPPORTED_SOLIDS + (
    "torus", "para", "arb8", "ellipsoid", "elcone", "eltube", "xtru",
    "genericPolycone", "genericPolyhedra", "hype", "paraboloid", "tet", "twistedbox", "twistedtrap",
    "twistedtrd", "twistedtubs", "scaledSolid", "multiUnion")

_STOD = re.compile(r"\s*[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan)", re.IGNORECASE)
_KTOL = 1e-9  # kTolerance, the facet validity threshold


class GdmlError(RuntimeError):
    pass


def _stod(s: str) -> float:
    """std::stod: the longest numeric prefix after leading white space."""
    m = _STOD.match(s)
    if not m:
        raise ValueError(s)
    return float(m.group(0))


def _rotation(x: float, y: float, z: float) -> List[float]:
    """makeRotationMatrixFromCartesianAngles (Middleware.cpp:111-129), row-major."""
    s1, c1 = -math.sin(x), math.cos(x)
    s2, c2 = -math.sin(y), math.cos(y)
    s3, c3 = -math.sin(z), math.cos(z)
    return [c2 * c3, -c2 * s3, s2,
            c1 * s3 + c3 * s1 * s2, c1 * c3 - s1 * s2 * s3, -c2 * s1,
            s1 * s3 - c1 * c3 * s2, c3 * s1 + c1 * s2 * s3, c1 * c2]


def _valid_triangle(v) -> bool:
    """The facet test of the tessellated builder: three distinct vertices and
    a non-degenerate normal."""
    def sub(a, b):
        return (a[0] - b[0], a[1] - b[1], a[2] - b[2])

    def mag2(a):
        return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]

    def cross(a, b):
        return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])

    if sum(mag2(sub(v[(i + 1) % 3], v[i])) < _KTOL for i in range(3)):
        return False
    for i in range(2):
        e1 = sub(v[i + 1], v[i])
        if mag2(e1) < _KTOL:
            continue
        for j in range(i + 1, 3):
            e2 = sub(v[(j + 1) % 3], v[j])
            if mag2(e2) < _KTOL:
                continue
            if mag2(cross(e1, e2)) >= _KTOL:
                return True
    return False


class _Reader:
    def __init__(self, geom, mm_unit: float):
        self.geom = geom
        self.mm = mm_unit
        self.constants: Dict[str, float] = {}
        self.positions: Dict[str, Tuple[float, float, float]] = {}
        self.rotations: Dict[str, Tuple[float, float, float]] = {}
        self.solids: Dict[str, int] = {}
        self.volumes: Dict[str, int] = {}
        self.world = None

    # ---- attributes ----------------------------------------------------------
    def num(self, el, name: str) -> float:
        s = el.get(name, "")
        try:
            return _stod(s)
        except ValueError:
            return self.constants.get(s, 0.0)

    def integer(self, el, name: str) -> int:
        s = el.get(name, "")
        return int(_stod(s)) if s else 0

    def length_unit(self, el) -> float:
        unit = el.get("unit" if el.tag == "position" else "lunit", "")
        mm = self.mm
        return {"mm": mm, "m": 1e3 * mm, "km": 1e6 * mm, "um": 1e-3 * mm, "nm": 1e-6 * mm, "cm": 10 * mm}.get(unit, mm)

    def angle_unit(self, el) -> float:
        unit = el.get("unit" if el.tag == "rotation" else "aunit", "")
        return math.pi / 180.0 if unit == "deg" else 1.0

    def lengths(self, el, *names):
        lm = self.length_unit(el)
        return [lm * self.num(el, n) for n in names]

    def angles(self, el, *names):
        am = self.angle_unit(el)
        return [am * self.num(el, n) for n in names]

    # ---- document ------------------------------------------------------------
    def node(self, el):
        tag = el.tag
        if tag in _REFERENCE_SOLIDS:
            self.solid(el)
        elif tag == "constant":
            try:
                self.constants.setdefault(el.get("name", ""), _stod(el.get("value", "")))
            except ValueError:
                self.constants.setdefault(el.get("name", ""), float("nan"))
        elif tag == "position":
            self.position(el)
        elif tag == "rotation":
            self.rotation(el)
        elif tag == "scale":
            pass  # only used by scaled solids and physvols, which are refused where used
        elif tag in ("isotope", "element", "material", "userinfo"):
            pass
        elif tag == "volume":
            self.volume(el)
        elif tag == "world":
            self.world_volume(el)
        else:
            for child in el:
                self.node(child)

    def position(self, el):
        x, y, z = self.lengths(el, "x", "y", "z")
        self.positions.setdefault(el.get("name", ""), (x, y, z))

    def rotation(self, el):
        x, y, z = self.angles(el, "x", "y", "z")
        self.rotations.setdefault(el.get("name", ""), (x, y, z))

    # ---- solids --------------------------------------------------------------
    def solid(self, el):
        tag, name, g = el.tag, el.get("name", ""), self.geom
        if tag not in SUPPORTED_SOLIDS:
            raise GdmlError(f"solid '{name}': <{tag}> is not one of the implemented families "
                            f"({', '.join(SUPPORTED_SOLIDS)})")
        if tag == "box":
            x, y, z = self.lengths(el, "x", "y", "z")
            s = g.box(x / 2.0, y / 2.0, z / 2.0)
        elif tag == "tube":
            z, rmin, rmax = self.lengths(el, "z", "rmin", "rmax")
            sphi, dphi = self.angles(el, "startphi", "deltaphi")
            s = g.tube(rmin, rmax, z / 2.0, sphi, dphi)
        elif tag in ("subtraction", "intersection", "union"):
            # Middleware.cpp:415-499, 622-623: the first solid placed as it is,
            # the second by the position and rotation (inline or by reference)
            first = second = None
            pos, rot = (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)
            for c in el:
                if c.tag in ("first", "second"):
                    ref = c.get("ref", "")
                    if ref not in self.solids:
                        raise GdmlError(f"solid '{name}': unknown {c.tag} solid '{ref}'")
                    if c.tag == "first":
                        first = self.solids[ref]
                    else:
                        second = self.solids[ref]
                elif c.tag == "positionref":
                    pos = self.positions.get(c.get("ref", ""), (0.0, 0.0, 0.0))
                elif c.tag == "rotationref":
                    rot = self.rotations.get(c.get("ref", ""), (0.0, 0.0, 0.0))
                elif c.tag == "position":
                    pos = tuple(self.lengths(c, "x", "y", "z"))
                elif c.tag == "rotation":
                    rot = tuple(self.angles(c, "x", "y", "z"))
            if first is None or second is None:
                raise GdmlError(f"solid '{name}': <{tag}> needs <first> and <second>")
            s = g.boolean(tag, first, second, None, [pos[0], pos[1], pos[2]] + _rotation(*rot))
        elif tag == "sphere":
            # Middleware.cpp:885-904: lengths and angles as given (no halving)
            rmin, rmax = self.lengths(el, "rmin", "rmax")
            sphi, dphi, sth, dth = self.angles(el, "startphi", "deltaphi", "starttheta", "deltatheta")
            s = g.sphere(rmin, rmax, sphi, dphi, sth, dth)
        elif tag == "torus":
            # Middleware.cpp:865-883: lengths and angles as given
            rmin, rmax, rtor = self.lengths(el, "rmin", "rmax", "rtor")
            sphi, dphi = self.angles(el, "startphi", "deltaphi")
            s = g.torus(rmin, rmax, rtor, sphi, dphi)
        elif tag == "cutTube":
            # the normals read as lengths; z halved (Middleware.cpp:738-768)
            z, rmin, rmax, lx, ly, lz, hx, hy, hz = self.lengths(el, "z", "rmin", "rmax", "lowX", "lowY", "lowZ",
                                                                  "highX", "highY", "highZ")
            sphi, dphi = self.angles(el, "startphi", "deltaphi")
            s = g.cut_tube(rmin, rmax, z / 2.0, sphi, dphi, (lx, ly, lz), (hx, hy, hz))
        elif tag == "cone":
            z, rmin1, rmin2, rmax1, rmax2 = self.lengths(el, "z", "rmin1", "rmin2", "rmax1", "rmax2")
            sphi, dphi = self.angles(el, "startphi", "deltaphi")
            s = g.cone(rmin1, rmax1, rmin2, rmax2, z / 2.0, sphi, dphi)
        elif tag == "orb":
            # the radius as written, not halved (Middleware.cpp:650-662)
            r, = self.lengths(el, "r")
            s = g.orb(r)
        elif tag == "trd":
            x1, x2, y1, y2, z = self.lengths(el, "x1", "x2", "y1", "y2", "z")
            s = g.trd(x1 / 2.0, x2 / 2.0, y1 / 2.0, y2 / 2.0, z / 2.0)
        elif tag == "trap":
            z, y1, x1, x2, y2, x3, x4 = self.lengths(el, "z", "y1", "x1", "x2", "y2", "x3", "x4")
            theta, phi, alpha1, alpha2 = self.angles(el, "theta", "phi", "alpha1", "alpha2")
            s = g.trapezoid(z / 2.0, theta, phi, y1 / 2.0, x1 / 2.0, x2 / 2.0, alpha1, y2 / 2.0, x3 / 2.0, x4 / 2.0,
                            alpha2)
        elif tag in ("polycone", "polyhedra"):
            sphi, dphi = self.angles(el, "startphi", "deltaphi")
            lm = self.length_unit(el)
            zs, rmins, rmaxs = [], [], []
            for zp in el:
                rmaxs.append(lm * self.num(zp, "rmax"))
                rmins.append(lm * self.num(zp, "rmin"))
                zs.append(lm * self.num(zp, "z"))
            if tag == "polycone":
                s = g.polycone(sphi, dphi, zs, rmins, rmaxs)
            else:
                s = g.polyhedron(sphi, dphi, self.integer(el, "numsides"), zs, rmins, rmaxs)
        else:
            s = g.tessellated(self.facets(el, name))
        self.solids.setdefault(name, s)

    def facets(self, el, name):
        tris = []
        zero = (0.0, 0.0, 0.0)

        def add(a, b):
            return (a[0] + b[0], a[1] + b[1], a[2] + b[2])

        for f in el:
            absolute = f.get("type", "") == "ABSOLUTE"
            nv = 3 if f.tag == "triangular" else 4
            v = [self.positions.get(f.get(f"vertex{k + 1}", ""), zero) for k in range(nv)]
            if nv == 3:
                t = v if absolute else [v[0], add(v[0], v[1]), add(add(v[0], v[1]), v[2])]
                if _valid_triangle(t):
                    tris.append(t)
                continue
            t1 = [v[0], v[1], v[2]] if absolute else [v[0], add(v[0], v[1]), add(add(v[0], v[1]), v[2])]
            t2 = [v[0], v[2], v[3]] if absolute else [v[0], add(v[0], v[2]), add(add(v[0], v[2]), v[3])]
            if not _valid_triangle(t1):
                continue
            tris.append(t1)
            if _valid_triangle(t2):
                tris.append(t2)
        if not tris:
            raise GdmlError(f"tessellated solid '{name}' has no valid facet")
        return tris

    # ---- structure -----------------------------------------------------------
    def volume(self, el):
        name = el.get("name", "")
        lv = None
        for child in el:
            if child.tag == "solidref":
                ref = child.get("ref", "")
                if ref not in self.solids:
                    raise GdmlError(f"volume '{name}': unknown solid '{ref}'")
                lv = self.geom.logical_volume(self.solids[ref])
                self.volumes.setdefault(name, lv)
            elif child.tag == "physvol":
                if lv is None:
                    raise GdmlError(f"volume '{name}': physvol before its solid")
                self.physvol(child, lv)

    def physvol(self, el, mother_lv):
        lv = None
        pos, rot = (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)
        for child in el:
            tag = child.tag
            if tag == "volumeref":
                ref = child.get("ref", "")
                if ref not in self.volumes:
                    raise GdmlError(f"physvol: unknown volume '{ref}'")
                lv = self.volumes[ref]
            elif tag == "position":
                self.position(child)
                pos = self.positions[child.get("name", "")]
            elif tag == "positionref":
                pos = self.positions.get(child.get("ref", ""), (0.0, 0.0, 0.0))
            elif tag == "rotation":
                self.rotation(child)
                rot = self.rotations[child.get("name", "")]
            elif tag == "rotationref":
                rot = self.rotations.get(child.get("ref", ""), (0.0, 0.0, 0.0))
            elif tag in ("scale", "scaleref"):
                raise GdmlError("physvol with a scale (reflection factory) is not supported")
        if lv is None:
            raise GdmlError("physvol without a volumeref")
        transform = [pos[0], pos[1], pos[2]] + _rotation(*rot)
        self.geom.add_daughter(mother_lv, self.geom.place(lv, transform))

    def world_volume(self, el):
        ref = el.get("ref", "")
        if ref not in self.volumes:
            raise GdmlError(f"world: unknown volume '{ref}'")
        self.world = self.geom.place(self.volumes[ref])


def load(path_or_text: str, geom, mm_unit: float = 0.1):
    """Reads a GDML file (or its text) into `geom`.

    Returns (world placed volume, {solid name: shape id}, {volume name:
    logical volume id}).  `mm_unit` is the length of one millimetre in the
    geometry's unit (the reference's default 0.1: centimetres).
    """
    text = path_or_text
    if not path_or_text.lstrip().startswith("<"):
        with open(path_or_text, "r", encoding="utf-8") as fh:
            text = fh.read()
    root = ET.fromstring(text)
    r = _Reader(geom, mm_unit)
    r.node(root)
    if r.world is None:
        raise GdmlError("no <world> in <setup>")
    return r.world, dict(r.solids), dict(r.volumes)
