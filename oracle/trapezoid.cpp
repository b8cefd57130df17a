// Oracle: trapezoid with the PlaneShell<4> side planes (VECGEOM_PLANESHELL is
// ON by default), scalar backend.
//   constructor/planes: VecGeom/volumes/UnplacedTrapezoid.h:56-66,
//     source/UnplacedTrapezoid.cpp:447-465 (corners), :560-660 (MakeAPlane, MakePlanes),
//     VecGeom/volumes/TrapezoidStruct.h (CalculateCached)
//   kernels: VecGeom/volumes/kernel/TrapezoidImplementation.h:98-395,
//     VecGeom/base/PlaneShell.h:130-290
#include "core.h"

namespace orc {

struct Trapezoid final : Shape {
  double fDz, fTheta, fPhi, fDy1, fDx1, fDx2, fTanAlpha1, fDy2, fDx3, fDx4, fTanAlpha2, fTthetaCphi, fTthetaSphi;
  double fA[4], fB[4], fC[4], fD[4];
  Vec3 corners[8];

  Trapezoid(double dz, double theta, double phi, double dy1, double dx1, double dx2, double alpha1, double dy2,
            double dx3, double dx4, double alpha2)
      : fDz(dz), fTheta(theta), fPhi(phi), fDy1(dy1), fDx1(dx1), fDx2(dx2), fTanAlpha1(std::tan(alpha1)), fDy2(dy2),
        fDx3(dx3), fDx4(dx4), fTanAlpha2(std::tan(alpha2))
  {
    fTthetaCphi = std::tan(fTheta) * std::cos(fPhi);
    fTthetaSphi = std::tan(fTheta) * std::sin(fPhi);
    const double dxdyDy1 = fTanAlpha1 * fDy1;
    const double dxdyDy2 = fTanAlpha2 * fDy2;
    const double dxdzDz = fTthetaCphi * fDz;
    const double dydzDz = fTthetaSphi * fDz;
    Vec3 *pt = corners;
    pt[0] = Vec3(-dxdzDz - dxdyDy1 - fDx1, -dydzDz - fDy1, -fDz);
    pt[1] = Vec3(-dxdzDz - dxdyDy1 + fDx1, -dydzDz - fDy1, -fDz);
    pt[2] = Vec3(-dxdzDz + dxdyDy1 - fDx2, -dydzDz + fDy1, -fDz);
    pt[3] = Vec3(-dxdzDz + dxdyDy1 + fDx2, -dydzDz + fDy1, -fDz);
    pt[4] = Vec3(+dxdzDz - dxdyDy2 - fDx3, +dydzDz - fDy2, +fDz);
    pt[5] = Vec3(+dxdzDz - dxdyDy2 + fDx3, +dydzDz - fDy2, +fDz);
    pt[6] = Vec3(+dxdzDz + dxdyDy2 - fDx4, +dydzDz + fDy2, +fDz);
    pt[7] = Vec3(+dxdzDz + dxdyDy2 + fDx4, +dydzDz + fDy2, +fDz);
    MakeAPlane(pt[0], pt[1], pt[5], pt[4], 0);
    MakeAPlane(pt[2], pt[6], pt[7], pt[3], 1);
    MakeAPlane(pt[0], pt[4], pt[6], pt[2], 2);
    MakeAPlane(pt[1], pt[3], pt[7], pt[5], 3);
  }

  void MakeAPlane(const Vec3 &p1, const Vec3 &p2, const Vec3 &p3, const Vec3 &p4, int i)
  {
    Vec3 n = Cross(p3 - p1, p4 - p2);
    const double inv = 1. / n.Mag();
    n = n * inv;
    Vec3 s = p1 + p2 + p3 + p4;
    Vec3 centr(0.25 * s.x, 0.25 * s.y, 0.25 * s.z);
    double d = -Dot(n, centr);
    fA[i] = n.x;
    fB[i] = n.y;
    fC[i] = n.z;
    fD[i] = d;
  }

  double PlaneDist(int i, const Vec3 &p) const { return fA[i] * p.x + fB[i] * p.y + fC[i] * p.z + fD[i]; }
  double PlaneProj(int i, const Vec3 &d) const { return fA[i] * d.x + fB[i] * d.y + fC[i] * d.z; }

  void Generic(const Vec3 &p, bool forInside, bool &cin, bool &cout) const
  {
    cout = Abs(p.z) > (forInside ? fDz + kHalfTolerance : fDz);
    if (forInside) cin = Abs(p.z) < fDz - kHalfTolerance;
    double dist[4];
    for (int i = 0; i < 4; ++i) dist[i] = PlaneDist(i, p);
    for (int i = 0; i < 4; ++i) {
      cout = cout || (dist[i] > (forInside ? 0. + kHalfTolerance : 0.));
      if (forInside) cin = cin && (dist[i] < 0. - kHalfTolerance);
    }
  }

  bool Contains(const Vec3 &p) const override
  {
    bool ci = false, co = false;
    Generic(p, false, ci, co);
    return !co;
  }

  int Inside(const Vec3 &p) const override
  {
    bool ci = false, co = false;
    Generic(p, true, ci, co);
    int r = kSurface;
    if (co) r = kOutside;
    if (ci) r = kInside;
    return r;
  }

  double DistanceToIn(const Vec3 &point, const Vec3 &dir, double) const override
  {
    double signZdir = Sign(dir.z);
    double max = signZdir * fDz - point.z;
    if (signZdir * max < 0. + kHalfTolerance) return kInfLength;
    double invdir = 1.0 / NonZero(dir.z);
    double smax = max * invdir;
    double smin = -(signZdir * fDz + point.z) * invdir;
    // PlaneShell::DistanceToIn
    double pdist[4], proj[4], vdist[4];
    for (int i = 0; i < 4; ++i) {
      pdist[i] = PlaneDist(i, point);
      proj[i] = PlaneProj(i, dir);
      vdist[i] = -pdist[i] / NonZero(proj[i]);
    }
    bool done = false;
    for (int i = 0; i < 4; ++i) {
      done = done || (pdist[i] > 0. + kHalfTolerance && proj[i] >= 0.);
      done = done || (pdist[i] > 0. - kHalfTolerance && proj[i] > 0.);
    }
    if (done) return kInfLength;
    for (int i = 0; i < 4; ++i) {
      bool posPoint = pdist[i] > 0. - kHalfTolerance;
      bool posDir = proj[i] > 0;
      bool fromIn = (!posPoint && posDir);
      done = done || (fromIn && vdist[i] < smin);
      bool fromOut = (posPoint && !posDir);
      done = done || (fromOut && vdist[i] > smax);
      if (done) return kInfLength;
      if (fromOut && vdist[i] > smin) smin = vdist[i];
      if (fromIn && vdist[i] < smax) smax = vdist[i];
    }
    return (smin <= smax) ? smin : kInfLength;
  }

  double DistanceToOut(const Vec3 &point, const Vec3 &dir, double) const override
  {
    if (Abs(point.z) > fDz + kHalfTolerance) return -1.0;
    double distance = kInfLength;
    double distz = (Sign(dir.z) * fDz - point.z) / NonZero(dir.z);
    if (dir.z != 0.) distance = distz;
    double distOut = kInfLength;
    for (int i = 0; i < 4; ++i) {
      double pdist = PlaneDist(i, point);
      double proj = PlaneProj(i, dir);
      double vdist = -pdist / NonZero(proj);
      if (pdist > kHalfTolerance) distOut = -1.;
      if (proj > 0. && vdist < distOut) distOut = vdist;
    }
    if (distOut < distance) distance = distOut;
    return distance;
  }

  double SafetyToIn(const Vec3 &point) const override
  {
    double safety = Abs(point.z) - fDz;
    for (int i = 0; i < 4; ++i) {
      double d = PlaneDist(i, point);
      if (d > safety) safety = d;
    }
    return safety;
  }

  double SafetyToOut(const Vec3 &point) const override
  {
    double safety = fDz - Abs(point.z);
    for (int i = 0; i < 4; ++i) {
      double d = -(PlaneDist(i, point));
      if (d < safety) safety = d;
    }
    return safety;
  }

  // NormalKernel (TrapezoidImplementation.h:406-462) over
  // PlaneShell::NormalKernel (VecGeom/base/PlaneShell.h:342-370)
  bool Normal(const Vec3 &point, Vec3 &normal) const override
  {
    const double delta = 1000. * kTolerance;
    normal = Vec3(0., 0., 0.);
    double safety = -kInfLength;
    double dist[4];
    for (int i = 0; i < 4; ++i) dist[i] = PlaneDist(i, point);
    for (int i = 0; i < 4; ++i) {
      const double saf_i = dist[i] - safety;
      if (Abs(saf_i) < kHalfTolerance && dist[i] >= -kHalfTolerance) normal = normal + Vec3(fA[i], fB[i], fC[i]);
      if (saf_i > 0.) normal = Vec3(fA[i], fB[i], fC[i]);
      if (saf_i > 0.) safety = dist[i];
    }
    const double safz = Sign(point[2]) * point[2] - fDz;
    if (Abs(safz - safety) < kHalfTolerance && safz >= -kHalfTolerance)
      normal = normal + Vec3(0, 0, Sign(point.z));
    if (safz > safety && safz >= -kHalfTolerance) normal = Vec3(0, 0, Sign(point.z));
    if (safz > safety) safety = safz;
    const bool valid = Abs(safety) <= delta;
    if (normal.Mag2() > 1.0) normal = normal.Normalized();
    return valid;
  }
  bool HasNormal() const override { return true; }

  void Extent(Vec3 &mn, Vec3 &mx) const override
  {
    mn = corners[0];
    mx = corners[0];
    for (int i = 1; i < 8; ++i)
      for (int k = 0; k < 3; ++k) {
        mn[k] = Min(mn[k], corners[i][k]);
        mx[k] = Max(mx[k], corners[i][k]);
      }
  }
};

Shape *make_trapezoid(double dz, double theta, double phi, double dy1, double dx1, double dx2, double alpha1,
                      double dy2, double dx3, double dx4, double alpha2)
{
  return new Trapezoid(dz, theta, phi, dy1, dx1, dx2, alpha1, dy2, dx3, dx4, alpha2);
}

}  // namespace orc
