// Host geometry tables and shape parameter derivation.
#include "geometry.h"
#include <cstring>
#include <cmath>
#include <cfloat>
#include "shapes/box_data.h"
#include "shapes/orb_data.h"
#include "shapes/cut_tube_data.h"
#include "shapes/boolean_data.h"
#include "shapes/sphere_data.h"
#include "shapes/torus_data.h"
#include "transform.hpp"
#include "shapes/tube_data.h"
#include "shapes/cone_data.h"
#include "shapes/trd_data.h"
#include "shapes/trapezoid_data.h"

namespace vgm {

int HostGeometry::add_shape(int type, const void *block, size_t ndoubles, int n0, int n1, const double bmin[3],
                            const double bmax[3])
{
  ShapeRec r;
  r.type = type;
  r.off = int32_t(pool.size());
  r.n0 = n0;
  r.n1 = n1;
  r.nd = int32_t(ndoubles);
  r.pad = 0;
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

// UnplacedTube(rmin, rmax, z, sphi, dphi) with TubeStruct::CalculateCached
// (VecGeom/volumes/TubeStruct.h:110-176) and the wedge vectors of
// source/Wedge_Evolution.cpp.  The bounding box is the full-phi one.
void make_tube_data(TubeData &t, double rmin, double rmax, double z, double sphi, double dphi)
{
  const double kHalfTol = 0.5e-9;
  t = TubeData{};
  t.rmin = rmin < 0.0 ? 0.0 : rmin;
  t.rmax = rmax;
  t.z = z;
  t.sphi = sphi;
  t.dphi = dphi;
  t.tolIz = t.z - kHalfTol;
  t.tolOz = t.z + kHalfTol;
  t.rmin2 = t.rmin * t.rmin;
  const double tolOrmin = t.rmin - kHalfTol, tolIrmin = t.rmin + kHalfTol;
  t.tolOrmin2 = tolOrmin * tolOrmin;
  t.tolIrmin2 = tolIrmin * tolIrmin;
  t.rmax2 = t.rmax * t.rmax;
  const double tolOrmax = t.rmax + kHalfTol, tolIrmax = t.rmax - kHalfTol;
  t.tolOrmax2 = tolOrmax * tolOrmax;
  t.tolIrmax2 = tolIrmax * tolIrmax;
  t.along1x = std::cos(sphi);
  t.along1y = std::sin(sphi);
  t.along2x = std::cos(sphi + dphi);
  t.along2y = std::sin(sphi + dphi);
  t.normal1x = -std::sin(sphi);
  t.normal1y = std::cos(sphi);
  t.normal2x = std::sin(sphi + dphi);
  t.normal2y = -std::cos(sphi + dphi);
  t.maxval = t.rmax > t.z ? t.rmax : t.z;
  const double kPi = 3.14159265358979323846;
  t.phi_treat = dphi < 2. * kPi;
  t.rmin_treat = t.rmin > 0;
  t.le_pi = dphi <= M_PI;
  t.lt_pi = dphi < kPi;
  t.gt_pi = dphi > M_PI;
}

// UnplacedOrb(r): extent +-r (OrbStruct.h)
int build_orb(HostGeometry &g, double r, std::string &err)
{
  if (!(r > 0)) {
    err = "orb radius must be positive";
    return -1;
  }
  OrbData o{r};
  const double lo[3] = {-r, -r, -r}, hi[3] = {r, r, r};
  return g.add_shape(kOrb, &o, sizeof(o) / sizeof(double), 0, 0, lo, hi);
}

// UnplacedSphere(rmin, rmax, sphi, dphi, stheta, dtheta): SphereStruct's
// angle checks (VecGeom/volumes/SphereStruct.h:120-210), the wedge and theta
// cone of the raw angles (Wedge_Evolution.h, ThetaCone_Evolution.h:40-60),
// and the Extent of source/UnplacedSphere.cpp:76-164
int build_sphere(HostGeometry &g, double rmin, double rmax, double sphi, double dphi, double stheta, double dtheta,
                 std::string &err)
{
  const double kPi = 3.14159265358979323846, kAngTol = 1e-9;
  if (!(rmin >= 0 && rmin < rmax && dphi > 0 && stheta >= 0 && stheta <= kPi && dtheta > 0)) {
    err = "sphere needs 0 <= rmin < rmax, dphi > 0, 0 <= stheta <= pi, dtheta > 0";
    return -1;
  }
  SphereData s{};
  s.rmin = rmin;
  s.rmax = rmax;
  s.sphi = sphi;
  s.dphi = dphi;
  s.stheta = stheta;
  s.dtheta = dtheta;
  bool fullPhi = true;
  if (dphi >= 2 * kPi - kAngTol * 0.5) {
    s.dphi = 2 * kPi;
    s.sphi = 0;
  } else {
    fullPhi = false;
    s.sphi = sphi < 0 ? 2 * kPi - std::fmod(std::fabs(sphi), 2 * kPi) : std::fmod(sphi, 2 * kPi);
    if (s.sphi + s.dphi > 2 * kPi) s.sphi -= 2 * kPi;
  }
  if (dtheta + stheta >= kPi) s.dtheta = kPi - stheta;
  const bool fullTheta = !(s.dtheta - s.stheta < kPi);
  s.etheta = s.stheta + s.dtheta;
  s.full_phi = fullPhi;
  s.full_theta = fullTheta;
  s.full_sphere = fullPhi && fullTheta;
  // Wedge(dphi, sphi) of the raw angles
  s.w_a1x = std::cos(sphi);
  s.w_a1y = std::sin(sphi);
  s.w_a2x = std::cos(sphi + dphi);
  s.w_a2y = std::sin(sphi + dphi);
  s.w_n1x = -std::sin(sphi);
  s.w_n1y = std::cos(sphi);
  s.w_n2x = std::sin(sphi + dphi);
  s.w_n2y = -std::cos(sphi + dphi);
  s.w_dphi = dphi;
  // ThetaCone(stheta, dtheta) of the raw angles
  s.tc_s = stheta;
  s.tc_e = stheta + dtheta;
  double ts = s.tc_s, te = s.tc_e;
  if (s.tc_s > kPi / 2) ts = kPi - s.tc_s;
  if (s.tc_e > kPi / 2) te = kPi - s.tc_e;
  s.tan_s = std::tan(ts);
  s.tan_s2 = s.tan_s * s.tan_s;
  s.tan_e = std::tan(te);
  s.tan_e2 = s.tan_e * s.tan_e;
  s.tan_bis = std::tan(ts + (dtheta / 2));
  if (s.tc_s > kPi / 2 && s.tc_e > kPi / 2) s.tan_bis = std::tan(ts - (dtheta / 2));
  s.slope1 = std::tan(kPi / 2 - s.tc_s);
  s.slope2 = std::tan(kPi / 2 - s.tc_e);
  // Extent
  double lo[3] = {-rmax, -rmax, -rmax}, hi[3] = {rmax, rmax, rmax};
  const double eT = s.stheta + s.dtheta;
  if (!s.full_sphere) {
    double st1 = 0., st2 = 0.;
    if (!fullTheta) {
      hi[2] = ((s.stheta <= kPi / 2.) ? rmax : rmin) * std::cos(s.stheta);
      lo[2] = ((eT <= kPi / 2.) ? rmin : rmax) * std::cos(eT);
      st1 = std::sin(s.stheta);
      st2 = std::sin(s.etheta);
      if (s.stheta <= kPi / 2.) {
        if (eT < kPi / 2.) {
          hi[0] = rmax * st2;
          lo[0] = -rmax * st2;
          hi[1] = rmax * st2;
          lo[1] = -rmax * st2;
        }
      } else {
        hi[0] = rmax * st1;
        lo[0] = -rmax * st1;
        hi[1] = rmax * st1;
        lo[1] = -rmax * st1;
      }
    }
    if (!fullPhi) {
      double Rmax = rmax;
      if (s.stheta > kPi / 2.) Rmax *= st1;
      if (eT < kPi / 2.) Rmax *= st2;
      const double Rmin = rmin * (st1 < st2 ? st1 : st2);
      const double Rin = 0.5 * (Rmax + Rmin);
      auto wedge_has = [&](double x, double y) {
        const double startCheck = (-x * s.w_a1y + y * s.w_a1x);
        const double endCheck = (-s.w_a2x * y + s.w_a2y * x);
        bool out = startCheck < 0.;
        if (s.w_dphi < kPi)
          out |= endCheck < 0.;
        else
          out &= endCheck < 0.;
        return !out;
      };
      const bool phi0out = !wedge_has(Rin, 0), phi90out = !wedge_has(0, Rin);
      const bool phi180out = !wedge_has(-Rin, 0), phi270out = !wedge_has(0, -Rin);
      if (phi0out || phi90out || phi180out || phi270out) {
        const double Cx = Rmax * std::cos(s.sphi), Dx = Rmax * std::cos(s.sphi + s.dphi);
        const double Cy = Rmax * std::sin(s.sphi), Dy = Rmax * std::sin(s.sphi + s.dphi);
        if (phi0out) hi[0] = Cx > Dx ? Cx : Dx;
        if (phi90out) hi[1] = Cy > Dy ? Cy : Dy;
        if (phi180out) lo[0] = Cx < Dx ? Cx : Dx;
        if (phi270out) lo[1] = Cy < Dy ? Cy : Dy;
        if (s.dphi < kPi) {
          const double Ax = Rmin * std::cos(s.sphi), Bx = Rmin * std::cos(s.sphi + s.dphi);
          const double Ay = Rmin * std::sin(s.sphi), By = Rmin * std::sin(s.sphi + s.dphi);
          double t = Ax > Bx ? Ax : Bx;
          hi[0] = t > hi[0] ? t : hi[0];
          t = Ay > By ? Ay : By;
          hi[1] = t > hi[1] ? t : hi[1];
          t = Ax < Bx ? Ax : Bx;
          lo[0] = t < lo[0] ? t : lo[0];
          t = Ay < By ? Ay : By;
          lo[1] = t < lo[1] ? t : lo[1];
        }
      }
    }
  }
  return g.add_shape(kSphere, &s, sizeof(s) / sizeof(double), 0, 0, lo, hi);
}

// UnplacedTorus2(rmin, rmax, rtor, sphi, dphi): TorusStruct2's cached squares,
// Wedge(dphi, sphi) and the bounding tube (VecGeom/volumes/TorusStruct2.h:37-43)
// of the constructor's angles as given.  The extent is the bounding tube's
// full box (UnplacedTorus2.h:210 asks the tube, which trims it by phi; the
// full box contains the trimmed one).
int build_torus(HostGeometry &g, double rmin, double rmax, double rtor, double sphi, double dphi, std::string &err)
{
  const double kPi = 3.14159265358979323846, kTol = 1e-9;
  if (!(rmin >= 0 && rmin < rmax && rtor > 0 && dphi > 0)) {
    err = "torus needs 0 <= rmin < rmax, rtor > 0, dphi > 0";
    return -1;
  }
  TorusData t{};
  t.rmin = rmin;
  t.rmax = rmax;
  t.rtor = rtor;
  t.sphi = sphi;
  t.dphi = dphi;
  t.rmin2 = rmin * rmin;
  t.rmax2 = rmax * rmax;
  t.rtor2 = rtor * rtor;
  t.w_a1x = std::cos(sphi);
  t.w_a1y = std::sin(sphi);
  t.w_a2x = std::cos(sphi + dphi);
  t.w_a2y = std::sin(sphi + dphi);
  t.w_n1x = -std::sin(sphi);
  t.w_n1y = std::cos(sphi);
  t.w_n2x = std::sin(sphi + dphi);
  t.w_n2y = -std::cos(sphi + dphi);
  const double band = 1.0e-8 * (rtor + rmax);
  t.normal_delta = 10.0 * kTol > band ? 10.0 * kTol : band;
  t.has_phi = dphi < 2 * kPi;
  t.w_lt_pi = dphi < kPi;
  make_tube_data(t.tube, rtor - rmax - kTol, rtor + rmax + kTol, rmax, sphi, dphi);
  const double r = t.tube.rmax;
  const double lo[3] = {-r, -r, -rmax}, hi[3] = {r, r, rmax};
  return g.add_shape(kTorus, &t, sizeof(t) / sizeof(double), 0, 0, lo, hi);
}

// Plane::Set(normal, origin) (VecGeom/volumes/Plane.h:30-37) of the unit
// normal: components below kTolerance are zeroed, the normal renormalised and
// the distance taken from the origin point (0, 0, oz)
static void set_cut_plane(double n[3], double &dd, const double in[3], double oz)
{
  const double kTolerance = 1e-9;
  // Vector3D::Unit (VecGeom/base/Vector3D.h:284-290): divided by sqrt(mag2 + kMinimum)
  const double mag2 = in[0] * in[0] + in[1] * in[1] + in[2] * in[2];
  const double s = std::sqrt(mag2 + DBL_MIN);
  double f[3] = {in[0] / s, in[1] / s, in[2] / s};
  for (double &c : f)
    if (std::fabs(c) < kTolerance) c = 0.0;  // FixZeroes
  const double inverseLength = 1. / std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (int i = 0; i < 3; ++i) n[i] = inverseLength * f[i];
  dd = inverseLength * -(f[0] * 0.0 + f[1] * 0.0 + f[2] * oz);
}

// Wedge::Contains (VecGeom/volumes/Wedge_Evolution.h:296-338)
static bool wedge_contains(const TubeData &t, double x, double y)
{
  const double startCheck = (-x * t.along1y + y * t.along1x);
  const double endCheck = (-t.along2x * y + t.along2y * x);
  bool out = startCheck < 0.;
  if (t.lt_pi)
    out = out || endCheck < 0.;
  else
    out = out && endCheck < 0.;
  return !out;
}

// UnplacedCutTube(rmin, rmax, z, sphi, dphi, bx, by, bz, tx, ty, tz)
// (VecGeom/volumes/UnplacedCutTube.h:54-70, CutTubeStruct.h:38-49) and its
// Extent (source/UnplacedCutTube.cpp:128-172)
int build_cut_tube(HostGeometry &g, double rmin, double rmax, double z, double sphi, double dphi, const double bot[3],
                   const double top[3], std::string &err)
{
  if (!(rmin >= 0 && rmin < rmax && z > 0 && dphi > 0)) {
    err = "cut tube needs 0 <= rmin < rmax, z > 0, dphi > 0";
    return -1;
  }
  if (bot[2] >= 0 || top[2] <= 0) {
    err = "illegal normal direction for cut planes (bottom z < 0 and top z > 0 required)";
    return -1;
  }
  CutTubeData c{};
  const double kPi = 3.14159265358979323846;
  make_tube_data(c.tube, rmin, rmax, DBL_MAX, sphi, dphi);  // kInfLength
  double n0[3], n1[3];
  set_cut_plane(n0, c.d0, bot, -z);
  set_cut_plane(n1, c.d1, top, z);
  c.n0x = n0[0], c.n0y = n0[1], c.n0z = n0[2];
  c.n1x = n1[0], c.n1y = n1[1], c.n1z = n1[2];
  c.dz = z;
  c.maxval = rmax > z ? rmax : z;
  const double rmn = c.tube.rmin, rmx = rmax;
  const double dztop = rmx * std::sqrt(1. - n1[2] * n1[2]) / n1[2];
  const double dzbottom = -rmx * std::sqrt(1. - n0[2] * n0[2]) / n0[2];
  double lo[3] = {-rmx, -rmx, -z - dzbottom}, hi[3] = {rmx, rmx, z + dztop};
  if (!(dphi == 2. * kPi)) {
    auto zt = [&](double r, double phi) { return z - (r / n1[2]) * (n1[0] * std::cos(phi) + n1[1] * std::sin(phi)); };
    auto zb = [&](double r, double phi) { return -z - (r / n0[2]) * (n0[0] * std::cos(phi) + n0[1] * std::sin(phi)); };
    auto mx = [](double a, double b) { return a > b ? a : b; };
    auto mn = [](double a, double b) { return a < b ? a : b; };
    const double s1 = sphi, s2 = sphi + dphi;
    const double c1 = std::cos(sphi), si1 = std::sin(sphi), c2 = std::cos(sphi + dphi), si2 = std::sin(sphi + dphi);
    if (!wedge_contains(c.tube, -n1[0], -n1[1])) {
      hi[2] = mx(zt(rmx, s1), zt(rmx, s2));
      hi[2] = mx(hi[2], mx(zt(rmn, s1), zt(rmn, s2)));
    }
    if (!wedge_contains(c.tube, -n0[0], -n0[1])) {
      lo[2] = mn(zb(rmx, s1), zb(rmx, s2));
      lo[2] = mn(lo[2], mn(zb(rmn, s1), zb(rmn, s2)));
    }
    if (!wedge_contains(c.tube, 1., 0.)) hi[0] = mx(mx(rmx * c1, rmx * c2), mx(rmn * c1, rmn * c2));
    if (!wedge_contains(c.tube, -1., 0.)) lo[0] = mn(mn(rmx * c1, rmx * c2), mn(rmn * c1, rmn * c2));
    if (!wedge_contains(c.tube, 0., 1.)) hi[1] = mx(mx(rmx * si1, rmx * si2), mx(rmn * si1, rmn * si2));
    if (!wedge_contains(c.tube, 0., -1.)) lo[1] = mn(mn(rmx * si1, rmx * si2), mn(rmn * si1, rmn * si2));
  }
  return g.add_shape(kCutTube, &c, sizeof(c) / sizeof(double), 0, 0, lo, hi);
}

// TransformedExtent (source/UnplacedBooleanVolume.cpp:403-442): the box of
// the eight corners of an operand's extent taken to the boolean's frame
static void transformed_extent(const ShapeRec &s, const Transform &T, double mn[3], double mx[3])
{
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) mn[k] = DBL_MAX, mx[k] = -DBL_MAX, lo[k] = s.bmin[k], hi[k] = s.bmax[k];
  const double delta[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  for (int x = 0; x <= 1; ++x)
    for (int y = 0; y <= 1; ++y)
      for (int z = 0; z <= 1; ++z) {
        const V3<double> c{lo[0] + x * delta[0], lo[1] + y * delta[1], lo[2] + z * delta[2]};
        const V3<double> m = to_master_point<double>(T, c);
        const double v[3] = {m.x, m.y, m.z};
        for (int k = 0; k < 3; ++k) {
          mn[k] = v[k] < mn[k] ? v[k] : mn[k];
          mx[k] = mx[k] < v[k] ? v[k] : mx[k];
        }
      }
}

// UnplacedBooleanVolume<kUnion / kIntersection / kSubtraction>(left, right)
// (VecGeom/volumes/UnplacedBooleanVolume.h, BooleanStruct.h:19-31) and its
// Extent (source/UnplacedBooleanVolume.cpp:444-482): the operand blocks are
// copied after the header so that the boolean's block is self-contained
int build_boolean(HostGeometry &g, int op, int left, const double *ltr, int right, const double *rtr,
                  std::string &err)
{
  if (op != kBoolUnion && op != kBoolIntersection && op != kBoolSubtraction) {
    err = "boolean operation must be union (0), intersection (1) or subtraction (2)";
    return -1;
  }
  const int n = int(g.shapes.size());
  if (left < 0 || left >= n || right < 0 || right >= n) {
    err = "boolean operand is not a shape of this context";
    return -1;
  }
  const ShapeRec L = g.shapes[left], R = g.shapes[right];
  for (const ShapeRec *s : {&L, &R})
    if (!(kBoolOperandMask & (1u << s->type))) {
      err = "boolean operands must be primitive solids (no booleans, no tessellated solids, no tori)";
      return -1;
    }
  BooleanData b{};
  b.op = op;
  b.ltype = L.type;
  b.rtype = R.type;
  auto set_tr = [](Transform &T, const double *tr) {
    for (int k = 0; k < 3; ++k) T.t[k] = tr ? tr[k] : 0.0;
    for (int k = 0; k < 9; ++k) T.r[k] = tr ? tr[3 + k] : ((k % 4 == 0) ? 1.0 : 0.0);
  };
  set_tr(b.ltr, ltr);
  set_tr(b.rtr, rtr);
  const int hdr = int((sizeof(BooleanData) + sizeof(double) - 1) / sizeof(double));
  b.loff = hdr;
  b.roff = hdr + L.nd;
  std::vector<double> block(size_t(hdr + L.nd + R.nd), 0.0);
  std::memcpy(block.data(), &b, sizeof(b));
  std::memcpy(block.data() + b.loff, g.pool.data() + L.off, size_t(L.nd) * sizeof(double));
  std::memcpy(block.data() + b.roff, g.pool.data() + R.off, size_t(R.nd) * sizeof(double));
  double mnL[3], mxL[3], mnR[3], mxR[3], lo[3], hi[3];
  transformed_extent(L, b.ltr, mnL, mxL);
  transformed_extent(R, b.rtr, mnR, mxR);
  for (int k = 0; k < 3; ++k) {
    if (op == kBoolSubtraction) {
      lo[k] = mnL[k], hi[k] = mxL[k];
    } else if (op == kBoolUnion) {
      lo[k] = mnL[k] < mnR[k] ? mnL[k] : mnR[k];
      hi[k] = mxL[k] > mxR[k] ? mxL[k] : mxR[k];
    } else {
      lo[k] = mnL[k] > mnR[k] ? mnL[k] : mnR[k];
      hi[k] = mxL[k] < mxR[k] ? mxL[k] : mxR[k];
    }
  }
  return g.add_shape(kBoolean, block.data(), block.size(), 0, 0, lo, hi);
}

int build_tube(HostGeometry &g, double rmin, double rmax, double z, double sphi, double dphi, std::string &err)
{
  if (!(rmax > 0 && z > 0 && dphi > 0 && rmin < rmax)) {
    err = "tube needs 0 <= rmin < rmax, z > 0, dphi > 0";
    return -1;
  }
  TubeData t;
  make_tube_data(t, rmin, rmax, z, sphi, dphi);
  const double lo[3] = {-rmax, -rmax, -z}, hi[3] = {rmax, rmax, z};
  return g.add_shape(kTube, &t, sizeof(t) / sizeof(double), 0, 0, lo, hi);
}

static double nonzero_h(double x) { return x + std::copysign(1.0e-30, x); }

// UnplacedCone(rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi): ConeStruct
// constructor order (VecGeom/volumes/ConeStruct.h): SetAndCheckDPhiAngle,
// SetAndCheckSPhiAngle, CalculateCached.  The wedge keeps the raw angles.
bool make_cone_data(ConeData &c, double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi,
                    double dphi, std::string &err)
{
  if (!(dz > 0 && dphi > 0 && rmin1 >= 0 && rmin2 >= 0 && rmax1 >= rmin1 && rmax2 >= rmin2 &&
        (rmax1 > 0 || rmax2 > 0))) {
    err = "cone needs 0 <= rmin <= rmax, dz > 0, dphi > 0";
    return false;
  }
  const double kPi = 3.14159265358979323846, kTwoPi = 2. * kPi, kConeTol = 1e-7, kHalfTol = 0.5e-9;
  c = ConeData{};
  c.rmin1 = rmin1;
  c.rmax1 = rmax1;
  c.rmin2 = rmin2;
  c.rmax2 = rmax2;
  c.dz = dz;
  double fS = sphi, fD = dphi;
  if (dphi >= kTwoPi - 0.5 * 1e-9) {
    fD = kTwoPi;
    fS = 0;
  }
  if (sphi < 0)
    fS = kTwoPi - std::fmod(std::fabs(sphi), kTwoPi);
  else
    fS = std::fmod(sphi, kTwoPi);
  if (fS + fD > kTwoPi) fS -= kTwoPi;
  c.sphi = fS;
  c.dphi = fD;
  c.along1x = std::cos(fS);
  c.along1y = std::sin(fS);
  c.along2x = std::cos(fS + fD);
  c.along2y = std::sin(fS + fD);
  c.w_along1x = std::cos(sphi);
  c.w_along1y = std::sin(sphi);
  c.w_along2x = std::cos(sphi + dphi);
  c.w_along2y = std::sin(sphi + dphi);
  c.w_normal1x = -std::sin(sphi);
  c.w_normal1y = std::cos(sphi);
  c.w_normal2x = std::sin(sphi + dphi);
  c.w_normal2y = -std::cos(sphi + dphi);
  // CalculateCached
  c.orig_rmax1 = c.rmax1;
  c.orig_rmax2 = c.rmax2;
  if (c.rmin1 == c.rmax1) c.rmax1 += kConeTol;
  if (c.rmin2 == c.rmax2) c.rmax2 += kConeTol;
  c.sq_rmin1 = c.rmin1 * c.rmin1;
  c.sq_rmax1 = c.rmax1 * c.rmax1;
  c.sq_rmin2 = c.rmin2 * c.rmin2;
  c.sq_rmax2 = c.rmax2 * c.rmax2;
  c.tan_rmin = (c.rmin2 - c.rmin1) * 0.5 / c.dz;
  c.sec_rmin = std::sqrt(1.0 + c.tan_rmin * c.tan_rmin);
  c.inv_sec_rmin = 1. / nonzero_h(c.sec_rmin);
  c.tan_rmax = (c.rmax2 - c.rmax1) * 0.5 / c.dz;
  c.sec_rmax = std::sqrt(1.0 + c.tan_rmax * c.tan_rmax);
  c.inv_sec_rmax = 1. / nonzero_h(c.sec_rmax);
  c.inner_slope = -(c.rmin1 - c.rmin2) / (2. * c.dz);
  c.outer_slope = -(c.rmax1 - c.rmax2) / (2. * c.dz);
  c.inner_offset = c.rmin2 - c.inner_slope * c.dz;
  c.outer_offset = c.rmax2 - c.outer_slope * c.dz;
  c.inner_tol = kConeTol * c.sec_rmin;
  c.outer_tol = kConeTol * c.sec_rmax;
  if (c.rmin2 > c.rmin1) {
    c.inner_apex = 2 * c.dz * c.rmin1 / (c.rmin2 - c.rmin1);
    c.tan_inner_apex = c.rmin2 / (2 * c.dz + c.inner_apex);
  } else {
    c.inner_apex = 2 * c.dz * c.rmin2 / nonzero_h(c.rmin1 - c.rmin2);
    c.tan_inner_apex = c.rmin1 / (2 * c.dz + c.inner_apex);
  }
  if (c.rmin1 == 0. || c.rmin2 == 0.) c.inner_apex = 0.;
  if (c.rmin1 == 0.) c.tan_inner_apex = c.rmin2 / (2 * c.dz);
  if (c.rmin2 == 0.) c.tan_inner_apex = c.rmin1 / (2 * c.dz);
  if (c.rmax2 > c.rmax1) {
    c.outer_apex = 2 * c.dz * c.rmax1 / (c.rmax2 - c.rmax1);
    c.tan_outer_apex = c.rmax2 / (2 * c.dz + c.outer_apex);
  } else {
    c.outer_apex = 2 * c.dz * c.rmax2 / nonzero_h(c.rmax1 - c.rmax2);
    c.tan_outer_apex = c.rmax1 / (2 * c.dz + c.outer_apex);
  }
  if (c.rmax1 == 0. || c.rmax2 == 0.) c.outer_apex = 0.;
  if (c.rmax1 == 0.) c.tan_outer_apex = c.rmax2 / (2 * c.dz);
  if (c.rmax2 == 0.) c.tan_outer_apex = c.rmax1 / (2 * c.dz);
  c.znorm_inner = c.tan_rmin / nonzero_h(c.sec_rmin);
  c.znorm_outer = -c.tan_rmax / nonzero_h(c.sec_rmax);
  c.tol_iz = c.dz - kHalfTol;
  c.tol_oz = c.dz + kHalfTol;
  c.phi_treat = fD < kTwoPi;
  c.rmin_treat = c.rmin1 > 0 || c.rmin2 > 0;
  c.le_pi = fD <= M_PI;
  c.lt_pi = fD < kPi;
  return true;
}

int build_cone(HostGeometry &g, double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi,
               double dphi, std::string &err)
{
  ConeData c;
  if (!make_cone_data(c, rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi, err)) return -1;
  const double r = c.orig_rmax1 > c.orig_rmax2 ? c.orig_rmax1 : c.orig_rmax2;
  const double lo[3] = {-r, -r, -dz}, hi[3] = {r, r, dz};
  return g.add_shape(kCone, &c, sizeof(c) / sizeof(double), 0, 0, lo, hi);
}

// UnplacedTrd(x1, x2, y1, y2, z) with TrdStruct::CalculateCached
// (VecGeom/volumes/TrdStruct.h); fToleranceY uses the x extents as there.
int build_trd(HostGeometry &g, double x1, double x2, double y1, double y2, double z, std::string &err)
{
  if (!(x1 >= 0 && x2 >= 0 && y1 >= 0 && y2 >= 0 && z > 0 && (x1 > 0 || x2 > 0) && (y1 > 0 || y2 > 0))) {
    err = "trd needs non-negative half-lengths and z > 0";
    return -1;
  }
  TrdData t{};
  t.dx1 = x1;
  t.dx2 = x2;
  t.dy1 = y1;
  t.dy2 = y2;
  t.dz = z;
  t.x2minusx1 = t.dx2 - t.dx1;
  t.y2minusy1 = t.dy2 - t.dy1;
  t.half_x1plusx2 = 0.5 * (t.dx1 + t.dx2);
  t.half_y1plusy2 = 0.5 * (t.dy1 + t.dy2);
  t.fx = 0.5 * (t.dx1 - t.dx2) / t.dz;
  t.fy = 0.5 * (t.dy1 - t.dy2) / t.dz;
  t.calfx = 1. / std::sqrt(1.0 + t.fx * t.fx);
  t.calfy = 1. / std::sqrt(1.0 + t.fy * t.fy);
  t.tolx = 1e-9 * std::sqrt(t.x2minusx1 * t.x2minusx1 + 4 * t.dz * t.dz);
  t.toly = 1e-9 * std::sqrt(t.x2minusx1 * t.x2minusx1 + 4 * t.dz * t.dz);
  t.vary_y = t.dy1 != t.dy2;
  const double x = x1 > x2 ? x1 : x2, y = y1 > y2 ? y1 : y2;
  const double lo[3] = {-x, -y, -z}, hi[3] = {x, y, z};
  return g.add_shape(kTrd, &t, sizeof(t) / sizeof(double), 0, 0, lo, hi);
}

// UnplacedTrapezoid(dz, theta, phi, dy1, dx1, dx2, alpha1, dy2, dx3, dx4, alpha2):
// eight corners (source/UnplacedTrapezoid.cpp:447-465) and the four side
// planes through them (MakeAPlane/MakePlanes, :560-660).
int build_trapezoid(HostGeometry &g, const double p[11], std::string &err)
{
  const double dz = p[0], theta = p[1], phi = p[2], dy1 = p[3], dx1 = p[4], dx2 = p[5], dy2 = p[7], dx3 = p[8],
               dx4 = p[9];
  if (!(dz > 0 && dy1 > 0 && dx1 > 0 && dx2 > 0 && dy2 > 0 && dx3 > 0 && dx4 > 0)) {
    err = "trapezoid needs positive half-lengths";
    return -1;
  }
  const double tanA1 = std::tan(p[6]), tanA2 = std::tan(p[10]);
  const double tthetaCphi = std::tan(theta) * std::cos(phi);
  const double tthetaSphi = std::tan(theta) * std::sin(phi);
  const double dxdyDy1 = tanA1 * dy1, dxdyDy2 = tanA2 * dy2;
  const double dxdzDz = tthetaCphi * dz, dydzDz = tthetaSphi * dz;
  double pt[8][3] = {{-dxdzDz - dxdyDy1 - dx1, -dydzDz - dy1, -dz}, {-dxdzDz - dxdyDy1 + dx1, -dydzDz - dy1, -dz},
                     {-dxdzDz + dxdyDy1 - dx2, -dydzDz + dy1, -dz}, {-dxdzDz + dxdyDy1 + dx2, -dydzDz + dy1, -dz},
                     {+dxdzDz - dxdyDy2 - dx3, +dydzDz - dy2, +dz}, {+dxdzDz - dxdyDy2 + dx3, +dydzDz - dy2, +dz},
                     {+dxdzDz + dxdyDy2 - dx4, +dydzDz + dy2, +dz}, {+dxdzDz + dxdyDy2 + dx4, +dydzDz + dy2, +dz}};
  TrapezoidData t{};
  t.dz = dz;
  const int quads[4][4] = {{0, 1, 5, 4}, {2, 6, 7, 3}, {0, 4, 6, 2}, {1, 3, 7, 5}};
  for (int i = 0; i < 4; ++i) {
    const double *p1 = pt[quads[i][0]], *p2 = pt[quads[i][1]], *p3 = pt[quads[i][2]], *p4 = pt[quads[i][3]];
    const double u[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
    const double v[3] = {p4[0] - p2[0], p4[1] - p2[1], p4[2] - p2[2]};
    double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    const double inv = 1. / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; ++k) n[k] = n[k] * inv;
    double cen[3];
    for (int k = 0; k < 3; ++k) cen[k] = 0.25 * (p1[k] + p2[k] + p3[k] + p4[k]);
    t.a[i] = n[0];
    t.b[i] = n[1];
    t.c[i] = n[2];
    t.d[i] = -(n[0] * cen[0] + n[1] * cen[1] + n[2] * cen[2]);
  }
  double lo[3] = {pt[0][0], pt[0][1], pt[0][2]}, hi[3] = {pt[0][0], pt[0][1], pt[0][2]};
  for (int i = 1; i < 8; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = pt[i][k] < lo[k] ? pt[i][k] : lo[k];
      hi[k] = pt[i][k] > hi[k] ? pt[i][k] : hi[k];
    }
  return g.add_shape(kTrapezoid, &t, sizeof(t) / sizeof(double), 0, 0, lo, hi);
}

}  // namespace vgm
