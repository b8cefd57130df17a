// Boolean solid parameter block (BooleanStruct, VecGeom/volumes/BooleanStruct.h:19-31):
// the operation, the two operand solids as placed volumes (type, placement,
// and a copy of the operand's own parameter block appended to this one).
#pragma once
#include <stdint.h>
#include "../vgm_types.h"

namespace vgm {

enum BooleanOp : int32_t { kBoolUnion = 0, kBoolIntersection = 1, kBoolSubtraction = 2 };  // BooleanOperation

struct BooleanData {
  int32_t op;            // BooleanOp
  int32_t ltype, rtype;  // operand shape types
  int32_t loff, roff;    // operand parameter blocks: offsets in doubles from the start of this block
  int32_t pad;
  Transform ltr, rtr;    // operand placements in the boolean's frame
};

}  // namespace vgm
