// Boolean solid parameter block (BooleanStruct, VecGeom/volumes/BooleanStruct.h:19-31):
// the operation, the two operand solids as placed volumes (type, placement,
// and a copy of the operand's own parameter block appended to this one).
#pragma once
#include <stdint.h>
#include "../vgm_types.h"

namespace vgm {

enum BooleanOp : int32_t { kBoolUnion = 0, kBoolIntersection = 1, kBoolSubtraction = 2 };  // BooleanOperation

// Shape families an operand may be: the primitive solids whose queries the
// boolean kernels dispatch to (not booleans, tessellated solids or tori).
constexpr unsigned kBoolOperandMask = (1u << kBox) | (1u << kTube) | (1u << kCone) | (1u << kTrd) |
                                      (1u << kTrapezoid) | (1u << kPolycone) | (1u << kPolyhedron) | (1u << kOrb) |
                                      (1u << kCutTube) | (1u << kSphere);

struct BooleanData {
  int32_t op;            // BooleanOp
  int32_t ltype, rtype;  // operand shape types
  int32_t loff, roff;    // operand parameter blocks: offsets in doubles from the start of this block
  int32_t pad;
  Transform ltr, rtr;    // operand placements in the boolean's frame
};

}  // namespace vgm
