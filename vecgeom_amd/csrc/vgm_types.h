// Plain-old-data geometry records shared by the host geometry builder and the
// gfx950 kernels.  Everything the device reads lives in a handful of flat
// arrays (one upload, no pointers inside records): shape records index into a
// double-precision parameter pool, placed volumes carry their transformation
// inline, logical volumes index a daughter list and a BVH.
#pragma once
#include <stdint.h>

namespace vgm {

enum ShapeType : int32_t {
  kBox = 0,
  kTube = 1,
  kCone = 2,
  kTrd = 3,
  kTrapezoid = 4,
  kPolycone = 5,
  kPolyhedron = 6,
  kTessellated = 7,
  kOrb = 8,
  kCutTube = 9,
  kBoolean = 10,
  kSphere = 11,
  kTorus = 12,
  kNumShapeTypes = 13
};

enum Method : int32_t {
  kContains = 0,
  kInside = 1,
  kDistanceToIn = 2,
  kDistanceToOut = 3,
  kSafetyToIn = 4,
  kSafetyToOut = 5,
  kNumMethods = 6
};

// Parameter blocks up to this size are staged in LDS by the kernels of the
// families whose lanes index their tables divergently (polycone sections,
// polyhedron quadrilaterals); larger ones are read from global memory.
constexpr int kMaxStageBytes = 64 * 1024;

// Inside_t values (VecGeom/base/Global.h:71-88).
enum : int32_t { kInsideV = 1, kSurfaceV = 2, kOutsideV = 3 };

// Transformation3D: fTranslation[3] + fRotation[9]
// (VecGeom/base/Transformation3D.h:54-63).  local = R^T (master - t) with the
// row/column convention of the reference: local_i = sum_j m_j * rot[3*j + i].
struct Transform {
  double t[3];
  double r[9];
};

struct ShapeRec {
  int32_t type;  // ShapeType
  int32_t off;   // offset of the type's parameter block in the pool (doubles)
  int32_t n0;    // type-specific counts (sections, segments, facets ...)
  int32_t n1;
  int32_t nd;    // size of the parameter block (doubles)
  int32_t pad;
  double bmin[3];  // local bounding box (Extent)
  double bmax[3];
};

struct LogicalRec {
  int32_t shape;
  int32_t dbegin;  // first entry in the daughter list
  int32_t dcount;  // number of daughters
  int32_t bvh;     // index into BVH headers, -1 if no daughters
};

struct PlacedRec {
  int32_t lv;
  int32_t id;
  Transform tr;
};

}  // namespace vgm
