// Cut tube parameter block (CutTubeStruct, VecGeom/volumes/CutTubeStruct.h:21-53):
// the tube with an unbounded z half-length and the two cut planes.
#pragma once
#include "tube_data.h"

namespace vgm {

struct CutTubeData {
  TubeData tube;                 // TubeStruct(rmin, rmax, kInfLength, sphi, dphi)
  double n0x, n0y, n0z, d0;      // bottom plane: unit normal and distance from the origin
  double n1x, n1y, n1z, d1;      // top plane
  double dz;                     // z half-length of the uncut tube
  double maxval;                 // Max(rmax, dz): the far-point pre-step scale
};

}  // namespace vgm
