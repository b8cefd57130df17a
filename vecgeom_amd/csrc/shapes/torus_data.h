// Torus parameter block (TorusStruct2, VecGeom/volumes/TorusStruct2.h:25-43):
// the radii and their squares, the phi wedge of the constructor's angles
// (Wedge_Evolution.h) and the bounding tube the entry distance starts from.
#pragma once
#include <stdint.h>
#include "tube_data.h"

namespace vgm {

struct TorusData {
  double rmin, rmax, rtor, sphi, dphi;
  double rmin2, rmax2, rtor2;
  // Wedge(dphi, sphi): along and normal vectors of its two planes
  double w_a1x, w_a1y, w_a2x, w_a2y;
  double w_n1x, w_n1y, w_n2x, w_n2y;
  // Normal's surface band: max(10 kTolerance, 1e-8 (rtor + rmax))
  double normal_delta;
  int32_t has_phi;  // dphi < 2 pi
  int32_t w_lt_pi;  // dphi < pi (the wedge's convexity test)
  // GenericUnplacedTube(rtor - rmax - kTolerance, rtor + rmax + kTolerance, rmax, sphi, dphi)
  TubeData tube;
};
static_assert(sizeof(TorusData) % sizeof(double) == 0, "torus block is whole doubles");

}  // namespace vgm
