// Orb parameter block (OrbStruct::fR, VecGeom/volumes/OrbStruct.h).
#pragma once

namespace vgm {

struct OrbData {
  double r;  // radius
};

}  // namespace vgm
