// Oracle: tube with the UniversalTube runtime checks (VECGEOM_NO_SPECIALIZATION
// default): VecGeom/volumes/kernel/TubeImplementation.h (TUBE_SAFETY_OLD
// defined at :14), TubeStruct.h:15-176, Wedge_Evolution.h:43-420,
// source/Wedge_Evolution.cpp (normals), shapetypes/TubeTypes.h
// (checkPhiTreatment: dphi < 2pi; checkRminTreatment: rmin > 0;
// SectorType = kUnknownAngle).  Scalar backend: EarlyReturnAllowed() is true,
// so every `if (EarlyReturnAllowed() && MaskFull(done)) return` returns.
#pragma once
#include "core.h"
#include "wedge.h"

namespace orc {

struct Tube final : Shape {
  double fRmin, fRmax, fZ, fSphi, fDphi;
  double fRmin2, fRmax2, fAlongPhi1x, fAlongPhi1y, fAlongPhi2x, fAlongPhi2y;
  double fTolIrmin2, fTolOrmin2, fTolIrmax2, fTolOrmax2, fTolIz, fTolOz;
  double fTolIrmin, fTolOrmin, fTolIrmax, fTolOrmax, fMaxVal;
  Wedge fPhiWedge;

  Tube(double rmin, double rmax, double z, double sphi, double dphi)
      : fRmin(rmin < 0.0 ? 0.0 : rmin), fRmax(rmax), fZ(z), fSphi(sphi), fDphi(dphi), fPhiWedge(dphi, sphi)
  {
    // TubeStruct::CalculateCached (TubeStruct.h:153-163)
    fTolIz = fZ - kHalfTolerance;
    fTolOz = fZ + kHalfTolerance;
    fRmin2 = fRmin * fRmin;
    fTolOrmin = (fRmin - kHalfTolerance);
    fTolIrmin = (fRmin + kHalfTolerance);
    fTolOrmin2 = fTolOrmin * fTolOrmin;
    fTolIrmin2 = fTolIrmin * fTolIrmin;
    fRmax2 = fRmax * fRmax;
    fTolOrmax = (fRmax + kHalfTolerance);
    fTolIrmax = (fRmax - kHalfTolerance);
    fTolOrmax2 = fTolOrmax * fTolOrmax;
    fTolIrmax2 = fTolIrmax * fTolIrmax;
    fAlongPhi1x = std::cos(fSphi);
    fAlongPhi1y = std::sin(fSphi);
    fAlongPhi2x = std::cos(fSphi + fDphi);
    fAlongPhi2y = std::sin(fSphi + fDphi);
    fMaxVal = Max(fRmax, fZ);
  }

  bool phiT() const { return fDphi < kTwoPi; }
  bool rminT() const { return fRmin > 0; }

  // TubeUtilities::PointInCyclicalSector (:62-101), onSurfaceT = false
  bool InSector(double x, double y, bool includeSurface) const
  {
    const bool smallerthanpi = fDphi <= M_PI;
    double startCheck = (-x * fAlongPhi1y + y * fAlongPhi1x);
    double endCheck = (-fAlongPhi2x * y + fAlongPhi2y * x);
    if (smallerthanpi) {
      if (includeSurface) return (startCheck >= -kHalfTolerance) & (endCheck >= -kHalfTolerance);
      return (startCheck >= kHalfTolerance) & (endCheck >= kHalfTolerance);
    }
    if (includeSurface) return (startCheck >= -kHalfTolerance) | (endCheck >= -kHalfTolerance);
    return (startCheck >= kHalfTolerance) | (endCheck >= kHalfTolerance);
  }

  // CircleTrajectoryIntersection (:106-162)
  bool Circle(double b, double c, const Vec3 &pos, const Vec3 &dir, bool largest, bool insectorCheck,
              double &dist) const
  {
    double delta = b * b - c;
    bool ok = delta > 0.;
    if (largest) ok |= delta == 0.;
    if (!ok) delta = 0.;
    delta = Sqrt(delta);
    if (!largest) delta = -delta;
    dist = -b + delta;
    ok &= dist >= -2 * kTolerance;
    if (!ok) return false;
    if (insectorCheck) {
      double hitz = pos.z + dist * dir.z;
      ok &= (Abs(hitz) <= fZ);
      if (!ok) return false;
      if (phiT()) {
        double hitx = pos.x + dist * dir.x;
        double hity = pos.y + dist * dir.y;
        ok &= InSector(hitx, hity, true);
      }
    }
    return ok;
  }

  // PhiPlaneTrajectoryIntersection (:264-308), PositiveDirectionOfPhiVector = true
  bool PhiPlane(double alongX, double alongY, double normalX, double normalY, const Vec3 &pos, const Vec3 &dir,
                bool insectorCheck, double &dist) const
  {
    dist = kInfLength;
    double dirDotNorm = dir.x * normalX + dir.y * normalY;
    bool ok = insectorCheck ? (dirDotNorm > 0.) : (dirDotNorm < 0.);
    double dirDotXY = (dir.y * alongX - dir.x * alongY);
    dist = (alongY * pos.x - alongX * pos.y) / NonZero(dirDotXY);
    ok &= (dist * Abs(dirDotNorm)) > -kHalfTolerance;
    if (insectorCheck) {
      double hitx = pos.x + dist * dir.x;
      double hity = pos.y + dist * dir.y;
      double hitz = pos.z + dist * dir.z;
      double r2 = hitx * hitx + hity * hity;
      ok &= Abs(hitz) <= fTolOz && (r2 >= fTolOrmin2) && (r2 <= fTolOrmax2);
      ok = ok && (hitx * alongX + hity * alongY) > 0.;
    } else {
      double hitx = pos.x + dist * dir.x;
      double hity = pos.y + dist * dir.y;
      ok = ok && (hitx * alongX + hity * alongY) >= 0.;
    }
    return ok;
  }

  // GenericKernelForContainsAndInside (:392-447)
  void Generic(const Vec3 &p, bool forInside, bool &cin, bool &cout) const
  {
    double absz = Abs(p.z);
    cout = absz > (forInside ? fZ + kHalfTolerance : fZ);
    if (forInside) cin = absz < fZ - kHalfTolerance;
    if (cout) return;
    double r2 = p.x * p.x + p.y * p.y;
    cout |= r2 > (forInside ? fRmax * (fRmax + 2.0 * kTolerance) : fRmax * fRmax);
    if (forInside) cin &= r2 < Abs(fRmax * (fRmax - 2.0 * kTolerance));
    if (cout) return;
    if (rminT()) {
      cout |= r2 <= (forInside ? Abs(fRmin * (fRmin - 2.0 * kTolerance)) : fRmin * fRmin);
      if (forInside) cin &= r2 > fRmin * (fRmin + 2.0 * kTolerance);
      if (cout) return;
    }
    if (phiT()) {
      bool ci = false, co = false;
      fPhiWedge.GenericKernel(p, forInside, ci, co);
      cout |= co;
      if (forInside) cin &= ci;
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
    int in = kSurface;
    if (co) in = kOutside;
    if (ci) in = kInside;
    return in;
  }

  // IsMovingInsideTubeSurface (:309-343)
  bool MovingInsideSurface(const Vec3 &p, const Vec3 &d, bool inner) const
  {
    const double rho = p.x * p.x + p.y * p.y;
    bool on = inner ? ((rho >= fTolOrmin2) && (rho <= fTolIrmin2) && (Abs(p.z) < (fZ + kTolerance)))
                    : ((rho >= fTolIrmax2) && (rho <= fTolOrmax2) && (Abs(p.z) < (fZ + kTolerance)));
    Vec3 n = inner ? Vec3(-p.x, -p.y, 0.) : Vec3(p.x, p.y, 0.);
    return on && (Dot(d, n) <= 0.);
  }

  double DistanceToIn(const Vec3 &pt, const Vec3 &dir, double stepMax) const override  // :478-502
  {
    Vec3 point = pt;
    double ptDist = point.Mag();
    double distToMove = 0.;
    const double order = 100.;
    if (ptDist > order * fMaxVal) {
      distToMove = (ptDist - order * fMaxVal);
      point = point + dir * distToMove;
    }
    double distance = DistanceToInKernel(point, dir);
    distance += distToMove;
    return distance;
  }

  double DistanceToInKernel(const Vec3 &point, const Vec3 &dir) const  // :507-681
  {
    bool done = false;
    double distance = kInfLength;
    double distz = Abs(point.z) - fZ;
    done |= distz > kHalfTolerance && point.z * dir.z >= 0;
    double rsq = point.x * point.x + point.y * point.y;
    double rdotn = point.x * dir.x + point.y * dir.y;
    done |= rsq > fTolIrmax2 && rdotn >= 0;
    if (done) return distance;
    distance = -1.0;
    bool inside = distz < -kHalfTolerance;
    inside &= rsq < fTolIrmax2;
    if (rminT()) inside &= rsq > fTolIrmin2;
    if (phiT() && inside) inside &= InSector(point.x, point.y, false);
    done |= inside;
    if (done) return distance;
    distance = kInfLength;
    distz /= NonZeroAbs(dir.z);
    double hitx = point.x + distz * dir.x;
    double hity = point.y + distz * dir.y;
    double r2 = hitx * hitx + hity * hity;
    bool okz = distz > -kHalfTolerance && (point.z * dir.z < 0);
    okz &= (r2 <= fRmax2);
    if (rminT()) okz &= (fRmin2 <= r2);
    if (phiT() && okz) okz &= InSector(hitx, hity, true);
    if (!done && okz) distance = distz;
    done |= okz;
    bool onSurfIn = MovingInsideSurface(point, dir, false);
    if (rminT()) onSurfIn |= MovingInsideSurface(point, dir, true);
    if (!phiT()) {
      if (!done && onSurfIn) distance = 0.;
      done |= onSurfIn;
      if (done) return distance;
    } else {
      bool insector = InSector(point.x, point.y, true);
      if (!done && insector && onSurfIn) distance = 0.;
      done |= (insector && onSurfIn);
      if (done) return distance;
    }
    double invnsq = 1. / NonZero(1. - dir.z * dir.z);
    double b = invnsq * rdotn;
    double crmax = invnsq * (rsq - fRmax2);
    double dist_rmax;
    bool ok_rmax = Circle(b, crmax, point, dir, false, true, dist_rmax);
    ok_rmax &= dist_rmax < distance;
    if (!done && ok_rmax) distance = dist_rmax;
    done |= ok_rmax;
    if (done) return distance;
    if (rminT()) {
      double crmin = invnsq * (rsq - fRmin2);
      double dist_rmin = -kInfLength;
      bool ok_rmin = Circle(b, crmin, point, dir, true, true, dist_rmin);
      ok_rmin &= dist_rmin < distance;
      if (!done && ok_rmin) distance = dist_rmin;
    }
    if (phiT()) {
      double dist_phi;
      const Vec3 n1 = fPhiWedge.GetNormal1(), n2 = fPhiWedge.GetNormal2();
      bool ok_phi = PhiPlane(fAlongPhi1x, fAlongPhi1y, n1.x, n1.y, point, dir, true, dist_phi);
      ok_phi &= dist_phi < distance;
      if (!done && ok_phi) distance = dist_phi;
      done |= ok_phi;
      ok_phi = PhiPlane(fAlongPhi2x, fAlongPhi2y, n2.x, n2.y, point, dir, true, dist_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
    }
    return distance;
  }

  double DistanceToOut(const Vec3 &point, const Vec3 &dir, double) const override  // :686-821
  {
    double distance = -1.;
    bool done = false;
    double distz = fZ - Abs(point.z);
    done |= distz < -kHalfTolerance;
    if (done) return distance;
    double rsq = point.x * point.x + point.y * point.y;
    double rdotn = dir.x * point.x + dir.y * point.y;
    double crmax = rsq - fRmax2;
    double crmin = rsq;
    done |= crmax > 2.0 * kTolerance * fRmax;
    if (done) return distance;
    if (rminT()) {
      crmin -= fRmin2;
      done |= crmin < -2.0 * kTolerance * fRmin;
      if (done) return distance;
    }
    if (phiT()) {
      bool ci = false, co = false;
      fPhiWedge.GenericKernel(point, true, ci, co);
      done |= co;
      if (done) return distance;
    }
    distance = kInfLength;
    double invdirz = 1. / NonZero(dir.z);
    distz = (fZ - point.z) * invdirz;
    if (dir.z < 0) distz = (-fZ - point.z) * invdirz;
    if (!done && dir.z != 0. && distz < distance) distance = distz;
    double invnsq = 1. / NonZero(1. - dir.z * dir.z);
    double b = invnsq * rdotn;
    if (rminT()) {
      double dist_rmin = kInfLength;
      crmin *= invnsq;
      bool ok_rmin = Circle(b, crmin, point, dir, false, false, dist_rmin);
      if (ok_rmin && dist_rmin < distance) distance = dist_rmin;
    }
    double dist_rmax = kInfLength;
    crmax *= invnsq;
    bool ok_rmax = Circle(b, crmax, point, dir, true, false, dist_rmax);
    if (ok_rmax && dist_rmax < distance) distance = dist_rmax;
    if (phiT()) {
      double dist_phi = kInfLength;
      const Vec3 n1 = fPhiWedge.GetNormal1(), n2 = fPhiWedge.GetNormal2();
      bool ok_phi = PhiPlane(fAlongPhi1x, fAlongPhi1y, n1.x, n1.y, point, dir, false, dist_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
      ok_phi = PhiPlane(fAlongPhi2x, fAlongPhi2y, n2.x, n2.y, point, dir, false, dist_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
    }
    return distance;
  }

  // PhiPlaneSafety<inside = false> (:209-243), kUnknownAngle
  double PhiSafety(const Vec3 &pos) const
  {
    double safety = (fDphi > M_PI) ? Sqrt(pos.x * pos.x + pos.y * pos.y) : kInfLength;
    double phi1 = pos.x * fAlongPhi1y - pos.y * fAlongPhi1x;
    if (phi1 > -kHalfTolerance && phi1 < safety) safety = phi1;
    double phi2 = pos.x * fAlongPhi2y - pos.y * fAlongPhi2x;
    phi2 *= -1;
    if (phi2 > -kHalfTolerance && phi2 < safety) safety = phi2;
    return safety;
  }

  double SafetyToIn(const Vec3 &point) const override  // SafetyToInOld :912-939
  {
    double safety = Abs(point.z) - fZ;
    double r = Sqrt(point.x * point.x + point.y * point.y);
    double safermax = r - fRmax;
    if (safermax > safety) safety = safermax;
    if (rminT()) {
      double safermin = fRmin - r;
      if (safermin > safety) safety = safermin;
    }
    if (phiT()) {
      bool insector = InSector(point.x, point.y, false);
      if (insector) return safety;
      double safephi = PhiSafety(point);
      if (!insector && safephi < kInfLength && safephi > safety) safety = safephi;
    }
    return safety;
  }

  double SafetyToOut(const Vec3 &point) const override  // SafetyToOutOld :944-964
  {
    double safety = fZ - Abs(point.z);
    double r = Sqrt(point.x * point.x + point.y * point.y);
    double safermax = fRmax - r;
    if (safermax < safety) safety = safermax;
    if (rminT()) {
      double safermin = r - fRmin;
      if (safermin < safety) safety = safermin;
    }
    if (phiT()) {
      double safephi = fPhiWedge.SafetyToOut(point);
      if (safephi < safety) safety = safephi;
    }
    return safety;
  }

  // ApproxSurfaceNormalKernel (:969-1015): the nearest surface of a point
  // that is not on one
  Vec3 ApproxSurfaceNormal(const Vec3 &p) const
  {
    Vec3 norm(0., 0., 0.);
    const double radius = p.Perp();
    double distRMax = Abs(radius - fRmax);
    double distRMin = kInfLength;
    if (distRMax < 0.) distRMax = kInfLength;
    if (fRmin != 0.) {
      distRMin = Abs(fRmin - radius);
      if (distRMin < 0.) distRMin = kInfLength;
    }
    double distMin = Min(distRMin, distRMax);
    double distPhi1 = kInfLength, distPhi2 = kInfLength;
    const Vec3 n1 = fPhiWedge.GetNormal1(), n2 = fPhiWedge.GetNormal2();
    if (fDphi != kTwoPi) {
      distPhi1 = p.x * n1.x + p.y * n1.y;
      distPhi2 = p.x * n2.x + p.y * n2.y;
      if (distPhi1 < 0.) distPhi1 = kInfLength;
      if (distPhi2 < 0.) distPhi2 = kInfLength;
      distMin = Min(distMin, Min(distPhi1, distPhi2));
    }
    double distZ = kInfLength;
    if (p.z < 0.) distZ = Abs(p.z + fZ);
    if (p.z >= 0.) distZ = Abs(p.z - fZ);
    distMin = Min(distMin, distZ);
    if (fDphi != 0.) {
      if (distMin == distPhi1) norm = -n1;
      if (distMin == distPhi2) norm = -n2;
    }
    if ((distMin == distZ) && (p.z < 0.)) norm = Vec3(0., 0., -1.);
    if ((distMin == distZ) && (p.z >= 0.)) norm = Vec3(0., 0., 1.);
    if (Abs(p.z) < (fZ + kTolerance)) {
      const Vec3 temp(p.x, p.y, 0.);
      if (distMin == distRMax) norm = temp.Unit();
      if (fRmin != 0.)
        if (distMin == distRMin) norm = -temp.Unit();
    }
    return norm;
  }

  // NormalKernel (:1019-1100): the sum of the normals of the surfaces the
  // point is on, over the square root of their number
  bool Normal(const Vec3 &p, Vec3 &norm) const override
  {
    bool ci = false, co = false;
    Generic(p, true, ci, co);
    if (ci || co) {
      norm = ApproxSurfaceNormal(p);
      return false;
    }
    norm = Vec3();
    int nosurface = 0;
    const double x2y2 = Sqrt(p.x * p.x + p.y * p.y);
    const bool inZ = ((p.z < fZ + kTolerance) && (p.z > -fZ - kTolerance));
    const bool inR = ((x2y2 >= fRmin - kTolerance) && (x2y2 <= fRmax + kTolerance));
    if (inR && (Abs(p.z - fZ) <= kTolerance)) {
      norm = Vec3(0., 0., 1.);
      nosurface++;
    }
    if (inR && (Abs(p.z + fZ) <= kTolerance)) {
      if (nosurface > 0)
        norm.z += -1.;
      else
        norm = Vec3(0., 0., -1.);
      nosurface++;
    }
    if (fRmin > 0.) {
      if (inZ && (Abs(x2y2 - fRmin) <= kTolerance)) {
        const double inv = 1. / x2y2;
        if (nosurface == 0) {
          norm.x = -p.x * inv;
          norm.y = -p.y * inv;
          norm.z = 0.;
        } else {
          norm.x += -p.x * inv;
          norm.y += -p.y * inv;
        }
        nosurface++;
      }
    }
    if (inZ && (Abs(x2y2 - fRmax) <= kTolerance)) {
      const double inv = 1. / x2y2;
      if (nosurface > 0) {
        norm.x += p.x * inv;
        norm.y += p.y * inv;
      } else {
        norm.x = p.x * inv;
        norm.y = p.y * inv;
        norm.z = 0.;
      }
      nosurface++;
    }
    if (fDphi < kTwoPi) {
      const Vec3 n1 = fPhiWedge.GetNormal1(), n2 = fPhiWedge.GetNormal2();
      if (inR && Wedge::IsOnSurfaceGeneric(fPhiWedge.fAlong1, n1, p)) {
        if (nosurface == 0)
          norm = -n1;
        else
          norm += -n1;
        nosurface++;
      }
      if (inR && Wedge::IsOnSurfaceGeneric(fPhiWedge.fAlong2, n2, p)) {
        if (nosurface == 0)
          norm = -n2;
        else
          norm += -n2;
        nosurface++;
      }
    }
    if (nosurface > 1) norm = norm / std::sqrt(1. * nosurface);
    return nosurface != 0;
  }
  bool HasNormal() const override { return true; }

  void Extent(Vec3 &mn, Vec3 &mx) const override
  {
    mn = Vec3(-fRmax, -fRmax, -fZ);
    mx = Vec3(fRmax, fRmax, fZ);
  }
};

}  // namespace orc
