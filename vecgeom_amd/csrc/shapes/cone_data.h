// Cone parameter block: ConeStruct cached values (VecGeom/volumes/ConeStruct.h)
// plus the phi-wedge vectors built from the constructor's raw angles.
#pragma once
#include <stdint.h>

namespace vgm {

struct ConeData {
  double rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi;  // rmax may carry the +kConeTolerance of rmin == rmax
  double orig_rmax1, orig_rmax2;
  double along1x, along1y, along2x, along2y;              // from the normalised sphi/dphi
  double w_along1x, w_along1y, w_along2x, w_along2y;      // wedge vectors (raw angles)
  double w_normal1x, w_normal1y, w_normal2x, w_normal2y;
  double inner_slope, outer_slope, inner_offset, outer_offset, inner_tol, outer_tol;
  double sq_rmin1, sq_rmin2, sq_rmax1, sq_rmax2, tol_iz, tol_oz;
  double inner_apex, tan_inner_apex, outer_apex, tan_outer_apex;
  double sec_rmin, sec_rmax, inv_sec_rmin, inv_sec_rmax, tan_rmin, tan_rmax, znorm_inner, znorm_outer;
  int32_t phi_treat;   // dphi < 2pi
  int32_t rmin_treat;  // rmin1 > 0 || rmin2 > 0
  int32_t le_pi;       // dphi <= pi (PointInCyclicalSector)
  int32_t lt_pi;       // dphi <  pi (Wedge)
};

}  // namespace vgm
