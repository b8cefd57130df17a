// Box parameter block (BoxStruct::fDimensions, VecGeom/volumes/BoxStruct.h:19-27).
#pragma once

namespace vgm {

struct BoxData {
  double dim[3];  // half-lengths
};

}  // namespace vgm
