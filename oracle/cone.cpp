// Oracle: cone factory (the restatement itself is in cone.h).
#include "cone.h"

namespace orc {

Shape *make_cone(double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi, double dphi)
{
  return new Cone(rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi);
}

}  // namespace orc
