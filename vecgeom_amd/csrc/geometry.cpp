// Host geometry tables and shape parameter derivation.
#include "geometry.h"
#include <cstring>
#include "shapes/box_data.h"

namespace vgm {

int HostGeometry::add_shape(int type, const void *block, size_t ndoubles, int n0, int n1, const double bmin[3],
                            const double bmax[3])
{
  ShapeRec r;
  r.type = type;
  r.off = int32_t(pool.size());
  r.n0 = n0;
  r.n1 = n1;
  for (int k = 0; k < 3; ++k) {
    r.bmin[k] = bmin[k];
    r.bmax[k] = bmax[k];
  }
  const size_t at = pool.size();
  pool.resize(at + ndoubles);
  std::memcpy(pool.data() + at, block, ndoubles * sizeof(double));
  shapes.push_back(r);
  ++version;
  return int(shapes.size()) - 1;
}

// UnplacedBox(dx, dy, dz) (VecGeom/volumes/UnplacedBox.h, BoxStruct.h:19-27).
int build_box(HostGeometry &g, double dx, double dy, double dz, std::string &err)
{
  if (!(dx > 0 && dy > 0 && dz > 0)) {
    err = "box half-lengths must be positive";
    return -1;
  }
  BoxData b{{dx, dy, dz}};
  const double lo[3] = {-dx, -dy, -dz}, hi[3] = {dx, dy, dz};
  return g.add_shape(kBox, &b, sizeof(b) / sizeof(double), 0, 0, lo, hi);
}

}  // namespace vgm
