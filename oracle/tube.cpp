// Oracle: tube factory (the restatement itself is in tube.h).
#include "tube.h"

namespace orc {

Shape *make_tube(double rmin, double rmax, double z, double sphi, double dphi)
{
  return new Tube(rmin, rmax, z, sphi, dphi);
}

}  // namespace orc
