// Tube parameter block: TubeStruct cached values (VecGeom/volumes/TubeStruct.h:15-176)
// plus the phi-wedge vectors (source/Wedge_Evolution.cpp).
#pragma once
#include <stdint.h>

namespace vgm {

struct TubeData {
  double rmin, rmax, z, sphi, dphi;
  double rmin2, rmax2;
  double along1x, along1y, along2x, along2y;  // (cos, sin) of sphi and sphi+dphi
  double normal1x, normal1y, normal2x, normal2y;  // wedge plane normals
  double tolIrmin2, tolOrmin2, tolIrmax2, tolOrmax2, tolIz, tolOz;
  double maxval;
  int32_t phi_treat;   // dphi < 2pi
  int32_t rmin_treat;  // rmin > 0
  int32_t le_pi;       // dphi <= pi  (PointInCyclicalSector)
  int32_t lt_pi;       // dphi <  pi  (Wedge)
  int32_t gt_pi;       // dphi >  pi  (PhiPlaneSafety)
  int32_t pad;
};

}  // namespace vgm
