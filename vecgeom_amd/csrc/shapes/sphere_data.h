// Sphere parameter block (SphereStruct, VecGeom/volumes/SphereStruct.h): the
// normalised angles of the struct, the phi wedge and the theta cone (both
// built from the constructor's raw angles, as the reference builds them).
#pragma once
#include <stdint.h>

namespace vgm {

struct SphereData {
  double rmin, rmax;
  double sphi, dphi;             // after CheckPhiAngles
  double stheta, dtheta, etheta; // after CheckThetaAngles; etheta = stheta + dtheta
  int32_t full_phi, full_theta, full_sphere, pad;
  // phi wedge (Wedge_Evolution.h) of the raw angles: along and normal
  // vectors of its two planes, and its opening angle
  double w_a1x, w_a1y, w_a2x, w_a2y;
  double w_n1x, w_n1y, w_n2x, w_n2y;
  double w_dphi;
  // theta cone (ThetaCone_Evolution.h) of the raw angles
  double tc_s, tc_e;                 // fSTheta, fETheta
  double tan_s, tan_e, tan_bis;      // tanSTheta, tanETheta, tanBisector
  double slope1, slope2;
  double tan_s2, tan_e2;
};
static_assert(sizeof(SphereData) % sizeof(double) == 0, "sphere block is whole doubles");

}  // namespace vgm
