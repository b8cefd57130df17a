// Polycone builder: PolyconeStruct::Init (VecGeom/volumes/PolyconeStruct.h):
// z planes are put in increasing order, a cone section is created between
// consecutive planes more than kTolerance apart, with the constructor's raw
// phi angles, and shifted to the middle of its z range.
#include <vector>
#include "geometry.h"
#include "shapes/polycone_data.h"

namespace vgm {

int build_polycone(HostGeometry &g, double sphi, double dphi, int nz, const double *zPlane, const double *rInner,
                   const double *rOuter, std::string &err)
{
  if (nz < 2 || !zPlane || !rInner || !rOuter) {
    err = "polycone needs at least two z planes";
    return -1;
  }
  std::vector<double> z(nz), rmin(nz), rmax(nz);
  const bool flip = zPlane[0] > zPlane[nz - 1];
  for (int i = 0; i < nz; ++i) {
    const int j = flip ? nz - 1 - i : i;
    z[i] = zPlane[j];
    rmin[i] = rInner[j];
    rmax[i] = rOuter[j];
  }
  double rmax_ext = rmax[0];
  for (int i = 1; i < nz; ++i) rmax_ext = rmax[i] > rmax_ext ? rmax[i] : rmax_ext;
  std::vector<double> zs, shifts;
  std::vector<ConeData> secs;
  const double kTol = 1e-9;
  double prevZ = z[0], prevRmax = 0, prevRmin = 0;
  const int dirZ = (z[1] < z[0]) ? -1 : 1;
  for (int i = 0; i < nz; ++i) {
    if (i == 0) {
      zs.push_back(z[i]);
    } else if (((z[i] > prevZ + kTol) && (dirZ > 0)) || ((z[i] < prevZ - kTol) && (dirZ < 0))) {
      ConeData c;
      if (!make_cone_data(c, prevRmin, prevRmax, rmin[i], rmax[i], (z[i] - prevZ) / 2, sphi, dphi, err)) return -1;
      zs.push_back(z[i]);
      const size_t k = zs.size() - 1;
      shifts.push_back(zs[k - 1] + 0.5 * (zs[k] - zs[k - 1]));
      secs.push_back(c);
    }
    prevZ = z[i];
    prevRmin = rmin[i];
    prevRmax = rmax[i];
  }
  const int nsec = int(secs.size());
  if (nsec < 1) {
    err = "polycone has no section of non-zero length";
    return -1;
  }
  std::vector<double> block(PolyconeData::bytes(nsec) / sizeof(double), 0.0);
  PolyconeData *h = reinterpret_cast<PolyconeData *>(block.data());
  h->nsec = nsec;
  h->pad = 0;
  double *zz = const_cast<double *>(h->zs());
  for (int i = 0; i <= nsec; ++i) zz[i] = zs[i];
  double *sh = const_cast<double *>(h->shift());
  for (int i = 0; i < nsec; ++i) sh[i] = shifts[i];
  ConeData *cs = const_cast<ConeData *>(h->sec());
  for (int i = 0; i < nsec; ++i) cs[i] = secs[i];
  const double lo[3] = {-rmax_ext, -rmax_ext, zs.front()}, hi[3] = {rmax_ext, rmax_ext, zs.back()};
  return g.add_shape(kPolycone, block.data(), block.size(), nsec, 0, lo, hi);
}

}  // namespace vgm
