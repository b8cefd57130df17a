// Oracle: cone, UniversalCone runtime checks, scalar backend.
//   ConeStruct::CalculateCached, SetAndCheck{S,D}PhiAngle: VecGeom/volumes/ConeStruct.h
//   kernels: VecGeom/volumes/kernel/ConeImplementation.h:99-517
//   helpers: VecGeom/volumes/ConeUtilities.h (ConeUtilities namespace and the
//            scalar ConeHelpers<Precision, coneTypeT> specialisation :560-830)
//   checkRminTreatment = rmin1 > 0 || rmin2 > 0, checkPhiTreatment = dphi < 2pi,
//   SectorType = kUnknownAngle (shapetypes/ConeTypes.h:101-150).
// EDGE_POINTS is not defined, so the block at ConeImplementation.h:175-186 is absent.
#pragma once
#include "core.h"
#include "wedge.h"

namespace orc {

struct Cone final : Shape {
  double fRmin1, fRmax1, fRmin2, fRmax2, fDz, fSPhi, fDPhi;
  Wedge fPhiWedge;
  double fAlongPhi1x, fAlongPhi1y, fAlongPhi2x, fAlongPhi2y;
  double fInnerSlope, fOuterSlope, fInnerOffset, fOuterOffset, fInnerTolerance, fOuterTolerance;
  double fSqRmin1, fSqRmin2, fSqRmax1, fSqRmax2, fTolIz, fTolOz;
  double fInnerConeApex, fTanInnerApexAngle, fOuterConeApex, fTanOuterApexAngle;
  double fSecRMin, fSecRMax, fInvSecRMin, fInvSecRMax, fTanRMin, fTanRMax, fZNormInner, fZNormOuter;
  double fOriginalRmax1, fOriginalRmax2;

  Cone(double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi, double dphi)
      : fRmin1(rmin1), fRmax1(rmax1), fRmin2(rmin2), fRmax2(rmax2), fDz(dz), fSPhi(sphi), fDPhi(dphi),
        fPhiWedge(dphi, sphi)
  {
    // SetAndCheckDPhiAngle(dphi)
    if (dphi >= kTwoPi - 0.5 * kAngTolerance) {
      fDPhi = kTwoPi;
      fSPhi = 0;
    } else if (dphi > 0) {
      fDPhi = dphi;
    }
    fPhiWedge.fDPhi = fDPhi;
    // SetAndCheckSPhiAngle(sphi)
    if (sphi < 0)
      fSPhi = kTwoPi - std::fmod(std::fabs(sphi), kTwoPi);
    else
      fSPhi = std::fmod(sphi, kTwoPi);
    if (fSPhi + fDPhi > kTwoPi) fSPhi -= kTwoPi;
    fPhiWedge.fSPhi = fSPhi;
    fAlongPhi1x = std::cos(fSPhi);
    fAlongPhi1y = std::sin(fSPhi);
    fAlongPhi2x = std::cos(fSPhi + fDPhi);
    fAlongPhi2y = std::sin(fSPhi + fDPhi);
    // CalculateCached
    fOriginalRmax1 = fRmax1;
    fOriginalRmax2 = fRmax2;
    if (fRmin1 == fRmax1) fRmax1 += kConeTolerance;
    if (fRmin2 == fRmax2) fRmax2 += kConeTolerance;
    fSqRmin1 = fRmin1 * fRmin1;
    fSqRmax1 = fRmax1 * fRmax1;
    fSqRmin2 = fRmin2 * fRmin2;
    fSqRmax2 = fRmax2 * fRmax2;
    fTanRMin = (fRmin2 - fRmin1) * 0.5 / fDz;
    fSecRMin = std::sqrt(1.0 + fTanRMin * fTanRMin);
    fInvSecRMin = 1. / NonZero(fSecRMin);
    fTanRMax = (fRmax2 - fRmax1) * 0.5 / fDz;
    fSecRMax = std::sqrt(1.0 + fTanRMax * fTanRMax);
    fInvSecRMax = 1. / NonZero(fSecRMax);
    fInnerSlope = -(fRmin1 - fRmin2) / (2. * fDz);
    fOuterSlope = -(fRmax1 - fRmax2) / (2. * fDz);
    fInnerOffset = fRmin2 - fInnerSlope * fDz;
    fOuterOffset = fRmax2 - fOuterSlope * fDz;
    fInnerTolerance = kConeTolerance * fSecRMin;
    fOuterTolerance = kConeTolerance * fSecRMax;
    if (fRmin2 > fRmin1) {
      fInnerConeApex = 2 * fDz * fRmin1 / (fRmin2 - fRmin1);
      fTanInnerApexAngle = fRmin2 / (2 * fDz + fInnerConeApex);
    } else {
      fInnerConeApex = 2 * fDz * fRmin2 / NonZero(fRmin1 - fRmin2);
      fTanInnerApexAngle = fRmin1 / (2 * fDz + fInnerConeApex);
    }
    if (fRmin1 == 0. || fRmin2 == 0.) fInnerConeApex = 0.;
    if (fRmin1 == 0.) fTanInnerApexAngle = fRmin2 / (2 * fDz);
    if (fRmin2 == 0.) fTanInnerApexAngle = fRmin1 / (2 * fDz);
    if (fRmax2 > fRmax1) {
      fOuterConeApex = 2 * fDz * fRmax1 / (fRmax2 - fRmax1);
      fTanOuterApexAngle = fRmax2 / (2 * fDz + fOuterConeApex);
    } else {
      fOuterConeApex = 2 * fDz * fRmax2 / NonZero(fRmax1 - fRmax2);
      fTanOuterApexAngle = fRmax1 / (2 * fDz + fOuterConeApex);
    }
    if (fRmax1 == 0. || fRmax2 == 0.) fOuterConeApex = 0.;
    if (fRmax1 == 0.) fTanOuterApexAngle = fRmax2 / (2 * fDz);
    if (fRmax2 == 0.) fTanOuterApexAngle = fRmax1 / (2 * fDz);
    fZNormInner = fTanRMin / NonZero(fSecRMin);
    fZNormOuter = -fTanRMax / NonZero(fSecRMax);
    fTolIz = fDz - kHalfTolerance;
    fTolOz = fDz + kHalfTolerance;
  }

  bool rminT() const { return fRmin1 > 0 || fRmin2 > 0; }
  bool phiT() const { return fDPhi < kTwoPi; }

  static double PlusTolSq(double x, double tol) { return x * (x + 2.0 * tol); }
  static double MinusTolSq(double x, double tol) { return Abs(x * (x - 2.0 * tol)); }

  // ConeUtilities::PointInCyclicalSector (kConeTolerance bands)
  bool InSector(double x, double y, bool includeSurface) const
  {
    const bool smallerthanpi = fDPhi <= M_PI;
    double startCheck = (-x * fAlongPhi1y) + (y * fAlongPhi1x);
    double endCheck = (-fAlongPhi2x * y) + (fAlongPhi2y * x);
    if (smallerthanpi) {
      if (includeSurface) return (startCheck >= -kConeTolerance) & (endCheck >= -kConeTolerance);
      return (startCheck >= kConeTolerance) & (endCheck >= kConeTolerance);
    }
    if (includeSurface) return (startCheck >= -kConeTolerance) | (endCheck >= -kConeTolerance);
    return (startCheck >= kConeTolerance) | (endCheck >= kConeTolerance);
  }

  // ConeUtilities::GetRadiusOfConeAtPoint
  double RadiusAt(double z, bool inner) const
  {
    if (inner) return (fRmin1 == fRmin2) ? fRmin1 : fInnerSlope * z + fInnerOffset;
    return (fOriginalRmax1 == fOriginalRmax2) ? fOriginalRmax1 : fOuterSlope * z + fOuterOffset;
  }

  // ConeUtilities::GetNormal
  Vec3 Normal(const Vec3 &p, bool inner) const
  {
    double rho = p.Perp();
    if (inner) {
      if ((fRmin1 == fRmin2) && (fRmin1 != 0.)) return Vec3(-p.x, -p.y, 0.);
      return Vec3(-p.x, -p.y, fZNormInner * (rho * fSecRMin));
    }
    if ((fRmax1 == fRmax2) && (fRmax1 != 0.)) return Vec3(p.x, p.y, 0.);
    return Vec3(p.x, p.y, fZNormOuter * (rho * fSecRMax));
  }

  bool OnConicalSurface(const Vec3 &p, bool inner) const
  {
    const double rho = p.Perp2();
    const double coneRad = RadiusAt(p.z, inner);
    const double coneRad2 = coneRad * coneRad;
    const double tol = inner ? fInnerTolerance : fOuterTolerance;
    return (rho >= (coneRad2 - tol * coneRad)) && (rho <= (coneRad2 + tol * coneRad)) &&
           (Abs(p.z) < (fDz + kConeTolerance));
  }

  // ConeUtilities::PhiPlaneTrajectoryIntersection, PositiveDirectionOfPhiVector = true
  bool PhiPlane(double alongX, double alongY, double normalX, double normalY, const Vec3 &pos, const Vec3 &dir,
                bool insectorCheck, double &dist) const
  {
    dist = kInfLength;
    bool ok = insectorCheck ? ((dir.x * normalX) + (dir.y * normalY) > 0.)
                            : ((dir.x * normalX) + (dir.y * normalY) < 0.);
    double dirDotXY = (dir.y * alongX) - (dir.x * alongY);
    if (dirDotXY != 0) dist = ((alongY * pos.x) - (alongX * pos.y)) / NonZero(dirDotXY);
    ok &= dist > -kConeTolerance;
    if (insectorCheck) {
      double hitx = pos.x + dist * dir.x;
      double hity = pos.y + dist * dir.y;
      double hitz = pos.z + dist * dir.z;
      double r2 = (hitx * hitx) + (hity * hity);
      double innerRadIrTol = RadiusAt(hitz, true) + kTolerance;
      double outerRadIrTol = RadiusAt(hitz, false) - kTolerance;
      ok &= Abs(hitz) <= fTolIz && (r2 >= innerRadIrTol * innerRadIrTol) && (r2 <= outerRadIrTol * outerRadIrTol);
      ok = ok && ((hitx * alongX) + (hity * alongY)) > 0.;
    } else {
      double hitx = pos.x + dist * dir.x;
      double hity = pos.y + dist * dir.y;
      ok = ok && ((hitx * alongX) + (hity * alongY)) >= 0.;
    }
    return ok;
  }

  // ConeHelpers<Precision>::DetectIntersectionAndCalculateDistanceToConicalSurface
  bool Conical(const Vec3 &point, const Vec3 &direction, bool forIn, bool inner, double &distance) const
  {
    distance = kInfLength;
    bool onConicalSurface = OnConicalSurface(point, inner);
    Vec3 normal = Normal(point, inner);
    bool tangent = Abs(Dot(direction, normal)) == 0.;
    if (onConicalSurface && tangent) return false;
    if (onConicalSurface) {
      bool moving = forIn ? (Dot(direction, normal) <= 0.) : (Dot(direction, normal) >= 0.);
      if (!phiT()) {
        if (moving) {
          distance = 0.;
          return true;
        }
      } else {
        bool insector = InSector(point.x, point.y, true);
        if (insector && moving) {
          distance = 0.;
          return true;
        }
      }
    }
    bool ok = false;
    double pDotV2D = point.x * direction.x + point.y * direction.y;
    double a, b, c;
    if (inner) {
      if (fRmin1 == fRmin2) {
        b = pDotV2D;
        a = direction.Perp2();
        c = point.Perp2() - fRmin2 * fRmin2;
      } else {
        double newPz;
        if (fRmin2 > fRmin1)
          newPz = (point.z + fDz + fInnerConeApex) * fTanInnerApexAngle;
        else
          newPz = (point.z - fDz - fInnerConeApex) * fTanInnerApexAngle;
        double dirT = direction.z * fTanInnerApexAngle;
        a = (direction.x * direction.x) + (direction.y * direction.y) - dirT * dirT;
        b = pDotV2D - (newPz * dirT);
        c = point.Perp2() - (newPz * newPz);
      }
      double b2 = b * b;
      double ac = a * c;
      if (b2 < ac) return false;
      double d2 = b2 - ac;
      double delta = Sqrt(d2);
      if (forIn) {
        if (b >= 0.)
          distance = (c / NonZero(-b - delta));
        else
          distance = (-b + delta) / NonZero(a);
      } else {
        if (b == 0. && delta == 0.) return false;
        if (b >= 0.)
          distance = (-b - delta) / NonZero(a);
        else
          distance = (c / NonZero(-b + delta));
      }
      if (distance < 0.) return false;
      double newZ = point.z + (direction.z * distance);
      ok = (Abs(newZ) < fDz);
    } else {
      if (fOriginalRmax1 == fOriginalRmax2) {
        a = direction.Perp2();
        b = pDotV2D;
        c = (point.Perp2() - fOriginalRmax2 * fOriginalRmax2);
      } else {
        double newPz;
        if (fOriginalRmax2 > fOriginalRmax1)
          newPz = (point.z + fDz + fOuterConeApex) * fTanOuterApexAngle;
        else
          newPz = (point.z - fDz - fOuterConeApex) * fTanOuterApexAngle;
        double dirT = direction.z * fTanOuterApexAngle;
        a = direction.x * direction.x + direction.y * direction.y - dirT * dirT;
        b = (pDotV2D - (newPz * dirT));
        c = (point.Perp2() - (newPz * newPz));
      }
      double b2 = b * b;
      double ac = a * c;
      double d2 = b2 - ac;
      if (d2 < 0) return false;
      double delta = Sqrt(d2);
      if (forIn) {
        if (b == 0. && delta == 0.) return false;
        if (b > 0.)
          distance = (-b - delta) / NonZero(a);
        else
          distance = (c / NonZero(-b + delta));
        double newZ = point.z + (direction.z * distance);
        ok = (Abs(newZ) < fDz + kHalfTolerance);
      } else {
        if (b < 0.)
          distance = (-b + delta) / NonZero(a);
        else
          distance = (c / NonZero(-b - delta));
        ok = distance > 0.;
      }
      if (distance < 0.) return false;
    }
    if (phiT()) {
      double hitx = 0, hity = 0;
      if (distance < kInfLength) {
        hitx = point.x + distance * direction.x;
        hity = point.y + distance * direction.y;
      }
      bool insector = InSector(hitx, hity, true);
      ok &= ((insector) && (distance < kInfLength));
    }
    return ok;
  }

  void Generic(const Vec3 &p, bool forInside, bool &cin, bool &cout) const
  {
    double absz = Abs(p.z);
    cout = absz > (forInside ? fDz + kConeTolerance : fDz);
    if (forInside) cin = absz < fDz - kConeTolerance;
    if (cout) return;
    double r2 = p.x * p.x + p.y * p.y;
    double rmax = (fOriginalRmax1 == fOriginalRmax2) ? fOriginalRmax1 : fOuterSlope * p.z + fOuterOffset;
    cout |= r2 > (forInside ? PlusTolSq(rmax, fOuterTolerance) : rmax * rmax);
    if (forInside) cin &= r2 < MinusTolSq(rmax, fOuterTolerance);
    if (cout) return;
    if (rminT()) {
      double rmin = fInnerSlope * p.z + fInnerOffset;
      cout |= r2 <= (forInside ? MinusTolSq(rmin, fInnerTolerance) : rmin * rmin);
      if (forInside) cin &= r2 > PlusTolSq(rmin, fInnerTolerance);
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

  double DistanceToIn(const Vec3 &point, const Vec3 &dir, double) const override  // :120-255
  {
    double distance = kInfLength;
    double distz = Abs(point.z) - fDz;
    bool outZAndGoingOut = (distz > kConeTolerance && (point.z * dir.z) >= 0.) ||
                           (Abs(distz) < kConeTolerance && (point.z * dir.z) > 0.);
    if (outZAndGoingOut) return distance;
    double outerRad = RadiusAt(point.z, false);
    double rsq = point.Perp2();
    if ((rsq > MinusTolSq(outerRad, fOuterTolerance)) && (Dot(dir, Normal(point, false)) >= 0.)) return distance;
    distance = -1.0;
    bool inside = distz < -kConeTolerance;
    inside &= rsq < MinusTolSq(outerRad, fOuterTolerance);
    if (rminT()) {
      double innerRad = RadiusAt(point.z, true);
      inside &= rsq > PlusTolSq(innerRad, fInnerTolerance);
    }
    if (phiT()) inside &= InSector(point.x, point.y, false);
    if (inside) return distance;
    distance = kInfLength;
    distz /= NonZero(Abs(dir.z));
    double hitx = point.x + distz * dir.x;
    double hity = point.y + distz * dir.y;
    double r2 = (hitx * hitx) + (hity * hity);
    double innerZTol = fTolIz;
    bool top = (point.z >= innerZTol) && (r2 <= fSqRmax2 + kTolerance);
    bool bottom = (point.z <= -innerZTol) && (r2 <= fSqRmax1 + kTolerance);
    bool okz = (top || bottom);
    if (rminT()) {
      top &= (r2 >= fSqRmin2 - kTolerance);
      bottom &= (r2 >= fSqRmin1 - kTolerance);
      okz &= (top || bottom);
    }
    if (phiT()) okz &= InSector(hitx, hity, true);
    if (okz) return distz;
    double dist_rOuter = kInfLength;
    bool ok_outer = Conical(point, dir, true, false, dist_rOuter);
    ok_outer &= dist_rOuter < distance;
    if (ok_outer) return dist_rOuter;
    if (rminT()) {
      double dist_rInner = kInfLength;
      bool ok_inner = Conical(point, dir, true, true, dist_rInner);
      ok_inner &= dist_rInner < distance;
      if (ok_inner) distance = dist_rInner;
    }
    if (phiT()) {
      double dist_phi;
      const Vec3 n1 = fPhiWedge.GetNormal1(), n2 = fPhiWedge.GetNormal2();
      bool ok_phi = PhiPlane(fAlongPhi1x, fAlongPhi1y, n1.x, n1.y, point, dir, true, dist_phi);
      ok_phi &= dist_phi < distance;
      if (ok_phi) distance = dist_phi;
      ok_phi = PhiPlane(fAlongPhi2x, fAlongPhi2y, n2.x, n2.y, point, dir, true, dist_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
    }
    return distance;
  }

  double DistanceToOut(const Vec3 &point, const Vec3 &direction, double) const override  // :260-366
  {
    double distance = -1.0;
    double distz = Abs(point.z) - fDz;
    bool outside = distz > kConeTolerance;
    double rsq = point.Perp2();
    double outerRad = RadiusAt(point.z, false);
    outside |= rsq > PlusTolSq(outerRad, fOuterTolerance);
    if (outside) return distance;
    if (rminT()) {
      double innerRad = RadiusAt(point.z, true);
      outside |= rsq < MinusTolSq(innerRad, fInnerTolerance);
      if (outside) return distance;
    }
    // the phi test at :285-289 runs only when `outside` has a set lane: never
    // for a scalar that reached this point.
    bool isGoingUp = direction.z > 0.;
    bool isGoingDown = direction.z < 0.;
    bool onZOut = ((isGoingUp && point.z > 0. && Abs(distz) < kConeTolerance) ||
                   (isGoingDown && point.z < 0. && Abs(distz) < kConeTolerance));
    if (onZOut) return distz;
    distance = kInfLength;
    double dirZInv = 1. / NonZero(direction.z);
    if (isGoingUp) distance = (fDz - point.z) * dirZInv;
    if (isGoingDown) distance = (-fDz - point.z) * dirZInv;
    double dist_rOuter = kInfLength;
    bool ok_outer = Conical(point, direction, false, false, dist_rOuter);
    if (ok_outer && dist_rOuter < distance) distance = dist_rOuter;
    if (rminT()) {
      double dist_rInner = kInfLength;
      bool ok_inner = Conical(point, direction, false, true, dist_rInner);
      if (ok_inner && dist_rInner < distance) distance = dist_rInner;
    }
    if (phiT()) {
      const Vec3 n1 = fPhiWedge.GetNormal1(), n2 = fPhiWedge.GetNormal2();
      bool onStart = Wedge::IsOnSurfaceGeneric(fPhiWedge.fAlong1, n1, point);
      bool onEnd = Wedge::IsOnSurfaceGeneric(fPhiWedge.fAlong2, n2, point);
      bool cond = (onStart && Dot(direction, -n1) > 0.) || (onEnd && Dot(direction, -n2) > 0.);
      if (cond) return 0.;
      double dist_phi;
      bool ok_phi = PhiPlane(fAlongPhi1x, fAlongPhi1y, n1.x, n1.y, point, direction, false, dist_phi);
      ok_phi &= dist_phi < distance;
      if (ok_phi) distance = dist_phi;
      ok_phi = PhiPlane(fAlongPhi2x, fAlongPhi2y, n2.x, n2.y, point, direction, false, dist_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
    }
    if (distance < 0. && Abs(distance) < kTolerance) distance = 0.;
    return distance;
  }

  // ConeImplementation::SafeDistanceToConicalSurface (:504-517)
  double SafeConical(const Vec3 &p, bool inner) const
  {
    double rho = p.Perp();
    if (inner) {
      double pRMin = fTanRMin * p.z + (fRmin1 + fRmin2) * 0.5;
      return (rho - pRMin) * fInvSecRMin;
    }
    double pRMax = fTanRMax * p.z + (fRmax1 + fRmax2) * 0.5;
    return (pRMax - rho) * fInvSecRMax;
  }

  double SafetyToIn(const Vec3 &point) const override  // :371-435
  {
    double safety = -1.0;
    double distz = Abs(point.z) - fDz;
    bool inside = distz < -kConeTolerance;
    double outerRad = RadiusAt(point.z, false);
    double rsq = point.Perp2();
    inside &= rsq < MinusTolSq(outerRad, fOuterTolerance);
    if (rminT()) {
      double innerRad = RadiusAt(point.z, true);
      inside &= rsq > PlusTolSq(innerRad, fInnerTolerance);
    }
    if (phiT() && inside) inside &= InSector(point.x, point.y, false);
    if (inside) return safety;
    double safeZ = Abs(point.z) - fDz;
    double safeOuter = -SafeConical(point, false);
    double safeInner = -kInfLength;
    if (rminT()) safeInner = -SafeConical(point, true);
    safety = Max(safeZ, Max(safeOuter, safeInner));
    if (phiT()) {
      double safetyPhi = fPhiWedge.SafetyToIn(point);
      safety = Max(safetyPhi, safety);
    }
    if (Abs(safety) < kTolerance) safety = 0.;
    return safety;
  }

  double SafetyToOut(const Vec3 &point) const override  // :437-500
  {
    double safety = -1.0;
    double distz = Abs(point.z) - fDz;
    double rsq = point.Perp2();
    bool outside = distz > kConeTolerance;
    double outerRad = RadiusAt(point.z, false);
    outside |= rsq > PlusTolSq(outerRad, fOuterTolerance);
    if (rminT()) {
      double innerRad = RadiusAt(point.z, true);
      outside |= rsq < MinusTolSq(innerRad, fInnerTolerance);
    }
    if (outside) return safety;
    double safeZ = fDz - Abs(point.z);
    double safeOuter = SafeConical(point, false);
    double safeInner = kInfLength;
    if (rminT()) safeInner = SafeConical(point, true);
    safety = Min(safeZ, Min(safeOuter, safeInner));
    if (phiT()) {
      double safetyPhi = fPhiWedge.SafetyToOut(point);
      safety = Min(safetyPhi, safety);
    }
    if (Abs(safety) < kTolerance) safety = 0.;
    return safety;
  }

  // ConeStruct::Normal (ConeStruct.h:177-306).  _frmin/_frmax are the
  // constructor's radii: rmin unchanged, rmax before the kConeTolerance bump.
  bool Normal(const Vec3 &p, Vec3 &norm) const override
  {
    const double _frmin1 = fRmin1, _frmin2 = fRmin2, _frmax1 = fOriginalRmax1, _frmax2 = fOriginalRmax2;
    int noSurfaces = 0;
    double distSPhi = kInfLength, distEPhi = kInfLength;
    Vec3 sumnorm(0., 0., 0.), nZ(0., 0., 1.);
    Vec3 nR, nr(0., 0., 0.), nPs, nPe;
    norm = sumnorm;
    double distZ = Abs(p.z) - fDz;
    const double rho = Sqrt(p.x * p.x + p.y * p.y);
    double distRMin, distRMax;
    const double pRMin = rho - p.z * fTanRMin;
    const double widRMin = fRmin2 - fDz * fTanRMin;
    if (Abs(_frmin1 - _frmin2) < fInnerTolerance)
      distRMin = (rho - _frmin2);
    else
      distRMin = (pRMin - widRMin) / fSecRMin;
    const double pRMax = rho - p.z * fTanRMax;
    const double widRMax = fRmax2 - fDz * fTanRMax;
    if (Abs(_frmax1 - _frmax2) < fOuterTolerance)
      distRMax = (rho - _frmax2);
    else
      distRMax = (pRMax - widRMax) / fSecRMax;
    bool inside = distZ < kTolerance && distRMax < fOuterTolerance;
    if (fRmin1 != 0. || fRmin2 != 0.) inside &= distRMin > -fInnerTolerance;
    distZ = std::fabs(distZ);
    distRMax = std::fabs(distRMax);
    distRMin = std::fabs(distRMin);
    double distNearest = distZ;
    Vec3 normNearest = nZ;
    if (p.z < 0.) normNearest = Vec3(0, 0, -1.);
    const bool fullPhi = fDPhi == kTwoPi;
    if (!fullPhi) {
      if (rho != 0.) {
        double pPhi = std::atan2(p.y, p.x);
        if (pPhi < fSPhi - kHalfTolerance)
          pPhi += 2 * kPi;
        else if (pPhi > fSPhi + fDPhi + kHalfTolerance)
          pPhi -= 2 * kPi;
        distSPhi = rho * (pPhi - fSPhi);
        distEPhi = rho * (pPhi - fSPhi - fDPhi);
        inside = inside && (distSPhi > -kTolerance) && (distEPhi < kTolerance);
        distSPhi = Abs(distSPhi);
        distEPhi = Abs(distEPhi);
      } else if (!(fRmin1 != 0.) || !(fRmin2 != 0.)) {
        distSPhi = 0.;
        distEPhi = 0.;
      }
      nPs = Vec3(std::sin(fSPhi), -std::cos(fSPhi), 0);
      nPe = Vec3(-std::sin(fSPhi + fDPhi), std::cos(fSPhi + fDPhi), 0);
    }
    if (rho > kHalfTolerance) {
      nR = Vec3(p.x / rho / fSecRMax, p.y / rho / fSecRMax, -fTanRMax / fSecRMax);
      if (fRmin1 != 0. || fRmin2 != 0.) nr = Vec3(-p.x / rho / fSecRMin, -p.y / rho / fSecRMin, fTanRMin / fSecRMin);
    }
    if (inside && distZ <= kHalfTolerance) {
      noSurfaces++;
      if (p.z >= 0.)
        sumnorm += nZ;
      else
        sumnorm = Vec3(0, 0, -1.);
    }
    if (inside && distRMax <= fOuterTolerance) {
      noSurfaces++;
      sumnorm += nR;
    } else if (noSurfaces == 0 && distRMax < distNearest) {
      distNearest = distRMax;
      normNearest = nR;
    }
    if (fRmin1 != 0. || fRmin2 != 0.) {
      if (inside && distRMin <= fInnerTolerance) {
        noSurfaces++;
        sumnorm += nr;
      } else if (noSurfaces == 0 && distRMin < distNearest) {
        distNearest = distRMin;
        normNearest = nr;
      }
    }
    if (!fullPhi) {
      if (inside && distSPhi <= kHalfTolerance) {
        noSurfaces++;
        sumnorm += nPs;
      } else if (noSurfaces == 0 && distSPhi < distNearest) {
        distNearest = distSPhi;
        normNearest = nPs;
      }
      if (inside && distEPhi <= kHalfTolerance) {
        noSurfaces++;
        sumnorm += nPe;
      } else if (noSurfaces == 0 && distEPhi < distNearest) {
        normNearest = nPe;
      }
    }
    if (noSurfaces == 0)
      norm = normNearest;
    else if (noSurfaces == 1)
      norm = sumnorm;
    else
      norm = sumnorm.Unit();
    bool valid = noSurfaces != 0;
    if (noSurfaces > 2) valid = false;
    return valid;
  }
  bool HasNormal() const override { return true; }

  void Extent(Vec3 &mn, Vec3 &mx) const override
  {
    double r = Max(fOriginalRmax1, fOriginalRmax2);
    mn = Vec3(-r, -r, -fDz);
    mx = Vec3(r, r, fDz);
  }
};

}  // namespace orc
