// Oracle: sphere with phi and theta cuts.  SphereStruct
// (VecGeom/volumes/SphereStruct.h), kernels
// (VecGeom/volumes/kernel/SphereImplementation.h), SphereUtilities.h, the
// theta cone (VecGeom/volumes/ThetaCone_Evolution.h), the phi wedge
// (Wedge_Evolution.h) and the Extent (source/UnplacedSphere.cpp:76-164).
// The wedge and the theta cone are built from the constructor's raw angles,
// the struct's own checks from the normalised ones, as in the reference.
// TEST ONLY.
#include "core.h"
#include "wedge.h"

namespace orc {

namespace {

double PlusTolSq(double x) { return x * (x + 2.0 * kTolerance); }              // MakePlusTolerantSquare<true>
double MinusTolSq(double x) { return Abs(x * (x - 2.0 * kTolerance)); }        // MakeMinusTolerantSquare<true>

// ThetaCone (ThetaCone_Evolution.h): the region between the cones of polar
// angles sTheta and sTheta + dTheta.  kHalfPi is the reference's macro
// 0.5 * kPi.
struct ThetaCone {
  double fSTheta, fDTheta, kAngTol, halfAngTolerance, fETheta;
  double tanSTheta, tanETheta, tanBisector, slope1, slope2, tanSTheta2, tanETheta2;

  ThetaCone(double sTheta, double dTheta) : fSTheta(sTheta), fDTheta(dTheta), kAngTol(kTolerance)
  {
    fETheta = fSTheta + fDTheta;
    halfAngTolerance = (0.5 * kAngTol);
    double tempfSTheta = fSTheta, tempfETheta = fETheta;
    if (fSTheta > kPi / 2) tempfSTheta = kPi - fSTheta;
    if (fETheta > kPi / 2) tempfETheta = kPi - fETheta;
    tanSTheta = std::tan(tempfSTheta);
    tanSTheta2 = tanSTheta * tanSTheta;
    tanETheta = std::tan(tempfETheta);
    tanETheta2 = tanETheta * tanETheta;
    tanBisector = std::tan(tempfSTheta + (fDTheta / 2));
    if (fSTheta > kPi / 2 && fETheta > kPi / 2) tanBisector = std::tan(tempfSTheta - (fDTheta / 2));
    slope1 = std::tan(kPi / 2 - fSTheta);
    slope2 = std::tan(kPi / 2 - fETheta);
  }

  Vec3 GetNormal1(const Vec3 &p) const
  {
    const Vec3 n(2. * p.x, 2. * p.y, -2. * tanSTheta2 * p.z);
    return (fSTheta <= kPi / 2.) ? -n : n;
  }
  Vec3 GetNormal2(const Vec3 &p) const
  {
    const Vec3 n(2 * p.x, 2 * p.y, -2 * tanETheta2 * p.z);
    return (fETheta <= kPi / 2.) ? n : -n;
  }
  Vec3 GetNormal(const Vec3 &p, bool start) const
  {
    if (start) {
      const Vec3 n(p.x, p.y, -tanSTheta2 * p.z);
      return (fSTheta <= 0.5 * kPi) ? -n : n;
    }
    const Vec3 n(p.x, p.y, -tanETheta2 * p.z);
    return (fETheta <= 0.5 * kPi) ? n : -n;
  }
  bool IsOnSurfaceGeneric(const Vec3 &p, bool start) const
  {
    const double rhs = start ? Abs(tanSTheta * p.z) : Abs(tanETheta * p.z);
    const double rho2 = p.Perp2();
    return rho2 >= MinusTolSq(rhs) && rho2 <= PlusTolSq(rhs);
  }
  bool IsPointOnSurfaceAndMovingOut(const Vec3 &p, const Vec3 &d, bool start, bool movingOut) const
  {
    if (!IsOnSurfaceGeneric(p, start)) return false;
    const double dn = Dot(d, GetNormal(p, start));
    return movingOut ? dn > 0. : dn < 0.;
  }
  bool Contains(const Vec3 &p) const { return !IsCompletelyOutside(p); }

  double DistanceToLine(double slope, double x, double y) const
  {
    const double dist = (y - slope * x) / Sqrt(1. + slope * slope);
    return Abs(dist);
  }

  double SafetyToIn(const Vec3 &p) const
  {
    const double pointRad = Sqrt(p.x * p.x + p.y * p.y);
    const double sfTh1 = DistanceToLine(slope1, pointRad, p.z);
    const double sfTh2 = DistanceToLine(slope2, pointRad, p.z);
    double safeTheta = Min(sfTh1, sfTh2);
    if (Contains(p)) return 0.;
    if (fSTheta < kPi / 2 + halfAngTolerance) {
      if (fETheta < kPi / 2 + halfAngTolerance) {
        if (fSTheta < fETheta && p.z < 0.) safeTheta = sfTh2;
      }
      if (fETheta > kPi / 2 + halfAngTolerance) {
        if (fSTheta < fETheta) {
          if (p.z > 0.) safeTheta = sfTh1;
          if (p.z < 0.) safeTheta = sfTh2;
        }
      }
    }
    if (fETheta > kPi / 2 + halfAngTolerance) {
      if (fSTheta > kPi / 2 + halfAngTolerance) {
        if (fSTheta < fETheta && p.z > 0.) safeTheta = sfTh1;
      }
    }
    return safeTheta;
  }

  double SafetyToOut(const Vec3 &p) const
  {
    const double pointRad = Sqrt(p.x * p.x + p.y * p.y);
    const double bisectorRad = Abs(p.z * tanBisector);
    bool condition = false;
    const double sfTh1 = DistanceToLine(slope1, pointRad, p.z);
    const double sfTh2 = DistanceToLine(slope2, pointRad, p.z);
    if (fSTheta < kPi / 2 + halfAngTolerance) {
      if (fETheta < kPi / 2 + halfAngTolerance) {
        if (fSTheta < fETheta) condition = (pointRad < bisectorRad) && (fSTheta != 0.);
      }
      if (fETheta > kPi / 2 + halfAngTolerance) {
        if (fSTheta < fETheta) condition = sfTh1 < sfTh2;
      }
    }
    if (fETheta > kPi / 2 + halfAngTolerance) {
      if (fSTheta > kPi / 2 + halfAngTolerance) {
        if (fSTheta < fETheta) condition = !((pointRad < bisectorRad) && (fETheta != kPi));
      }
    }
    return condition ? sfTh1 : sfTh2;
  }

  // the "done" mask of the root cleanup: |root| < 3 kTolerance snaps to 0
  static void Snap(double &root, bool &done)
  {
    done |= (Abs(root) < 3. * kTolerance);
    if (Abs(root) < 3. * kTolerance) root = 0.;
  }

  void DistanceToIn(const Vec3 &p, const Vec3 &d, double &dist1, double &dist2, bool &int1, bool &int2) const
  {
    bool done = false;
    double firstRoot = kInfLength, secondRoot = kInfLength;
    const double pDotV2d = p.x * d.x + p.y * d.y;
    const double rho2 = p.x * p.x + p.y * p.y;
    const double dirRho2 = d.Perp2();
    const double b = pDotV2d - p.z * d.z * tanSTheta2;
    const double a = dirRho2 - d.z * d.z * tanSTheta2;
    const double c = rho2 - p.z * p.z * tanSTheta2;
    const double d2 = b * b - a * c;
    const double aInv = 1. / NonZero(a);
    if (d2 > 0.) firstRoot = (-b + Sqrt(Abs(d2))) * aInv;
    done |= (Abs(firstRoot) < 3. * kTolerance);
    if (Abs(firstRoot) < 3. * kTolerance) firstRoot = 0.;
    if (!done && (firstRoot < 0.)) firstRoot = kInfLength;
    const double b2 = pDotV2d - p.z * d.z * tanETheta2;
    const double a2 = dirRho2 - d.z * d.z * tanETheta2;
    const double c2 = rho2 - p.z * p.z * tanETheta2;
    const double d22 = b2 * b2 - a2 * c2;
    const double a2Inv = 1. / NonZero(a2);
    if (d22 > 0.) secondRoot = (-b2 - Sqrt(Abs(d22))) * a2Inv;
    if (!done && (Abs(secondRoot) < 3. * kTolerance)) secondRoot = 0.;
    done |= (Abs(secondRoot) < 3. * kTolerance);
    if (!done && (secondRoot < 0.)) secondRoot = kInfLength;

    if (fSTheta < 0.5 * kPi + halfAngTolerance) {
      if (fETheta < 0.5 * kPi + halfAngTolerance) {
        if (fSTheta < fETheta) {
          dist1 = firstRoot;
          dist2 = secondRoot;
          const double z1 = (p.z + dist1 * d.z), z2 = (p.z + dist2 * d.z);
          int1 = ((d2 > 0.) && (z1 > 0.));
          int2 = ((d22 > 0.) && (z2 > 0.));
        }
      }
      if (fETheta >= 0.5 * kPi - halfAngTolerance && fETheta <= 0.5 * kPi + halfAngTolerance) {
        if (d.z > 0.) dist2 = -p.z / d.z;
        const double z2 = (p.z + dist2 * d.z);
        int2 = ((dist2 != kInfLength) && (Abs(z2) < halfAngTolerance));
      }
      if (fETheta > 0.5 * kPi + halfAngTolerance) {
        if (fSTheta < fETheta) {
          dist1 = firstRoot;
          if (d22 > 0.) secondRoot = (-b2 + Sqrt(Abs(d22))) * a2Inv;
          done = false;
          Snap(secondRoot, done);
          if (!done && (secondRoot < 0.)) secondRoot = kInfLength;
          dist2 = secondRoot;
          const double z1 = (p.z + dist1 * d.z), z2 = (p.z + dist2 * d.z);
          int1 = ((d2 > 0) && (dist1 != kInfLength) && (z1 > 0.));
          int2 = ((d22 > 0) && (dist2 != kInfLength) && (z2 < 0.));
        }
      }
    }
    if (fSTheta >= 0.5 * kPi - halfAngTolerance) {
      if (fETheta > 0.5 * kPi + halfAngTolerance) {
        if (fSTheta < fETheta) {
          if (d2 > 0.) firstRoot = (-b - Sqrt(Abs(d2))) * aInv;
          done = false;
          Snap(firstRoot, done);
          if (!done && (firstRoot < 0.)) firstRoot = kInfLength;
          dist1 = firstRoot;
          if (d22 > 0.) secondRoot = (-b2 + Sqrt(Abs(d22))) * a2Inv;
          done = false;
          Snap(secondRoot, done);
          if (!done && (secondRoot < 0.)) secondRoot = kInfLength;
          dist2 = secondRoot;
          const double z1 = (p.z + dist1 * d.z), z2 = (p.z + dist2 * d.z);
          int1 = ((d2 > 0) && (dist1 != kInfLength) && (z1 < 0.));
          int2 = ((d22 > 0) && (dist2 != kInfLength) && (z2 < 0.));
        }
      }
    }
    if (fSTheta >= 0.5 * kPi - halfAngTolerance && fSTheta <= 0.5 * kPi + halfAngTolerance) {
      if (d.z < 0.) dist1 = -p.z / d.z;
      const double z1 = (p.z + dist1 * d.z);
      int1 = ((dist1 != kInfLength) && (Abs(z1) < halfAngTolerance));
    }
  }

  void DistanceToOut(const Vec3 &p, const Vec3 &d, double &dist1, double &dist2, bool &int1, bool &int2) const
  {
    const double inf = kInfLength;
    double firstRoot = kInfLength, secondRoot = kInfLength;
    const double pDotV2d = p.x * d.x + p.y * d.y;
    const double rho2 = p.x * p.x + p.y * p.y;
    const double b = pDotV2d - p.z * d.z * tanSTheta2;
    const double a = d.x * d.x + d.y * d.y - d.z * d.z * tanSTheta2;
    const double c = rho2 - p.z * p.z * tanSTheta2;
    double d2 = b * b - a * c;
    if (d2 < 0. && Abs(d2) < kHalfTolerance) d2 = 0.;
    if ((d2 >= 0.) && b >= 0. && a != 0.) firstRoot = ((-b - Sqrt(Abs(d2))) / NonZero(a));
    if ((d2 >= 0.) && b < 0.) firstRoot = ((c) / NonZero(-b + Sqrt(Abs(d2))));
    if (firstRoot < 0.) firstRoot = kInfLength;
    const double b2 = p.x * d.x + p.y * d.y - p.z * d.z * tanETheta2;
    const double a2 = d.x * d.x + d.y * d.y - d.z * d.z * tanETheta2;
    const double c2 = p.x * p.x + p.y * p.y - p.z * p.z * tanETheta2;
    double d22 = (b2 * b2) - (a2 * c2);
    if (d22 < 0. && Abs(d22) < kHalfTolerance) d22 = 0.;
    if ((d22 >= 0.) && b2 >= 0.) secondRoot = ((c2) / NonZero(-b2 - Sqrt(Abs(d22))));
    if ((d22 >= 0.) && b2 < 0. && a2 != 0.) secondRoot = ((-b2 + Sqrt(Abs(d22))) / NonZero(a2));
    if (secondRoot < 0. && Abs(secondRoot) > kTolerance) secondRoot = kInfLength;
    if (Abs(secondRoot) < kTolerance) secondRoot = 0.;

    if (fSTheta < kPi / 2 + halfAngTolerance) {
      if (fETheta < kPi / 2 + halfAngTolerance) {
        if (fSTheta < fETheta) {
          dist1 = firstRoot;
          dist2 = secondRoot;
          const double z1 = (p.z + dist1 * d.z), z2 = (p.z + dist2 * d.z);
          int1 = ((d2 > 0) && (dist1 != kInfLength) && ((z1) > -kHalfTolerance));
          int2 = ((d22 > 0) && (dist2 != kInfLength) && ((z2) > -kHalfTolerance));
          const double dirRho2 = d.x * d.x + d.y * d.y;
          double zs = kInfLength;
          if (fSTheta != 0.) zs = dirRho2 / tanSTheta;
          double ze = kInfLength;
          if (fETheta != 0.) ze = dirRho2 / tanETheta;
          const bool cond = (p.x == 0. && p.y == 0. && p.z == 0. && d.z < zs && d.z < ze);
          if (cond) {
            dist1 = 0.;
            dist2 = 0.;
          }
          int1 |= cond;
          int2 |= cond;
        }
      }
      if (fETheta >= kPi / 2 - halfAngTolerance && fETheta <= kPi / 2 + halfAngTolerance) {
        dist1 = firstRoot;
        dist2 = inf;
        if (d.z < 0.) dist2 = -1. * p.z / d.z;
        const double z2 = (p.z + dist2 * d.z), z1 = (p.z + dist1 * d.z);
        int2 = ((dist2 != kInfLength) && (Abs(z2) < kHalfTolerance) && !(d.z == 0.));
        int1 = ((d2 >= 0) && (dist1 != kInfLength) && (Abs(z1) < kHalfTolerance) && !(d.z == 0.));
      }
      if (fETheta > kPi / 2 + halfAngTolerance) {
        if (fSTheta < fETheta) {
          dist1 = firstRoot;
          if ((d22 >= 0.) && b2 > 0. && a2 != 0.) secondRoot = ((-b2 - Sqrt(Abs(d22))) / NonZero(a2));
          if ((d22 >= 0.) && b2 <= 0.) secondRoot = ((c2) / NonZero(-b2 + Sqrt(Abs(d22))));
          if (secondRoot < 0.) secondRoot = kInfLength;
          dist2 = secondRoot;
          const double z1 = (p.z + dist1 * d.z), z2 = (p.z + dist2 * d.z);
          int1 = ((d2 >= 0) && (dist1 != kInfLength) && ((z1) > -kHalfTolerance));
          int2 = ((d22 >= 0) && (dist2 != kInfLength) && ((z2) < kHalfTolerance));
        }
      }
    }
    if (fETheta > kPi / 2 + halfAngTolerance) {
      if (fSTheta < fETheta) {
        secondRoot = kInfLength;
        if ((d22 >= 0.) && b2 > 0. && a2 != 0.) secondRoot = ((-b2 - Sqrt(Abs(d22))) / NonZero(a2));
        if ((d22 >= 0.) && b2 <= 0.) secondRoot = ((c2) / NonZero(-b2 + Sqrt(Abs(d22))));
        dist2 = secondRoot;
        if (fSTheta > kPi / 2 + halfAngTolerance) {
          if ((d2 >= 0.) && b > 0.) firstRoot = ((c) / NonZero(-b - Sqrt(Abs(d2))));
          if ((d2 >= 0.) && b <= 0. && a != 0.) firstRoot = ((-b + Sqrt(Abs(d2))) / NonZero(a));
          if (firstRoot < 0.) firstRoot = kInfLength;
          dist1 = firstRoot;
          const double z1 = (p.z + dist1 * d.z);
          int1 = ((d2 > 0) && (dist1 != kInfLength) && ((z1) < kHalfTolerance));
          const double z2 = (p.z + dist2 * d.z);
          int2 = ((d22 > 0) && (dist2 != kInfLength) && ((z2) < kHalfTolerance));
          const double dirRho2 = d.x * d.x + d.y * d.y;
          double zs = -kInfLength;
          if (tanSTheta != 0.) zs = -dirRho2 / tanSTheta;
          double ze = -kInfLength;
          if (tanETheta != 0.) ze = -dirRho2 / tanETheta;
          const bool cond = (p.x == 0. && p.y == 0. && p.z == 0. && d.z > zs && d.z > ze);
          if (cond) {
            dist1 = 0.;
            dist2 = 0.;
          }
          int1 |= cond;
          int2 |= cond;
        }
      }
    }
    if (fSTheta >= kPi / 2 - halfAngTolerance && fSTheta <= kPi / 2 + halfAngTolerance) {
      dist2 = secondRoot;
      dist1 = kInfLength;
      if (d.z > 0.) dist1 = -1. * p.z / NonZero(d.z);
      const double z1 = (p.z + dist1 * d.z), z2 = (p.z + dist2 * d.z);
      int1 = ((dist1 != kInfLength) && (Abs(z1) < kHalfTolerance) && (d.z != 0.));
      int2 = ((d22 >= 0) && (dist2 != kInfLength) && (Abs(z2) < kHalfTolerance) && (d.z != 0.));
    }
  }

  bool IsCompletelyInside(const Vec3 &p) const
  {
    const double rho = Sqrt(p.Mag2() - (p.z * p.z));
    const double cone1Radius = Abs(p.z * tanSTheta);
    const double cone2Radius = Abs(p.z * tanETheta);
    const bool isPointOnZAxis = p.z != 0. && p.x == 0. && p.y == 0.;
    const bool isPointOnXYPlane = p.z == 0. && (p.x != 0. || p.y != 0.);
    const double startTheta = fSTheta, endTheta = fETheta;
    bool ci = (isPointOnZAxis && ((startTheta == 0. && endTheta == kPi) || (p.z > 0. && startTheta == 0.) ||
                                  (p.z < 0. && endTheta == kPi)));
    ci |= (!ci && (isPointOnXYPlane && (startTheta < 0.5 * kPi && endTheta > 0.5 * kPi &&
                                         (0.5 * kPi - startTheta) > kAngTol && (endTheta - 0.5 * kPi) > kTolerance)));
    if (fSTheta < 0.5 * kPi + halfAngTolerance) {
      if (fETheta < 0.5 * kPi + halfAngTolerance) {
        if (fSTheta < fETheta) {
          const double tolAngMin = cone1Radius + 2 * kAngTol * 10.;
          const double tolAngMax = cone2Radius - 2 * kAngTol * 10.;
          ci |= (!ci && (((rho <= tolAngMax) && (rho >= tolAngMin) && (p.z > 0.) && (fSTheta != 0.)) ||
                         ((rho <= tolAngMax) && (fSTheta == 0.) && (p.z > 0.))));
        }
      }
      if (fETheta > 0.5 * kPi + halfAngTolerance) {
        if (fSTheta < fETheta) {
          const double tolAngMin = cone1Radius + 2 * kAngTol * 10.;
          const double tolAngMax = cone2Radius + 2 * kAngTol * 10.;
          ci |= (!ci && (((rho >= tolAngMin) && (p.z > 0.)) || ((rho >= tolAngMax) && (p.z < 0.))));
        }
      }
      if (fETheta >= 0.5 * kPi - halfAngTolerance && fETheta <= 0.5 * kPi + halfAngTolerance)
        ci &= !(Abs(p.z) < halfAngTolerance);
    }
    if (fETheta > 0.5 * kPi + halfAngTolerance) {
      if (fSTheta >= 0.5 * kPi - halfAngTolerance && fSTheta <= 0.5 * kPi + halfAngTolerance)
        ci &= !(Abs(p.z) < halfAngTolerance);
      if (fSTheta > 0.5 * kPi + halfAngTolerance) {
        if (fSTheta < fETheta) {
          const double tolAngMin = cone1Radius - 2 * kAngTol * 10.;
          const double tolAngMax = cone2Radius + 2 * kAngTol * 10.;
          ci |= (!ci && (((rho <= tolAngMin) && (rho >= tolAngMax) && (p.z < 0.) && (fETheta != kPi)) ||
                         ((rho <= tolAngMin) && (p.z < 0.) && (fETheta == kPi))));
        }
      }
    }
    return ci;
  }

  bool IsCompletelyOutside(const Vec3 &p) const
  {
    const double diff = p.Perp2();
    const double rho = Sqrt(diff);
    const double cone1Radius = Abs(p.z * tanSTheta);
    const double cone2Radius = Abs(p.z * tanETheta);
    const bool isPointOnZAxis = p.z != 0. && p.x == 0. && p.y == 0.;
    const bool isPointOnXYPlane = p.z == 0. && (p.x != 0. || p.y != 0.);
    bool co = (isPointOnZAxis &&
               ((fSTheta != 0. && fETheta != kPi) || (p.z > 0. && fSTheta != 0.) || (p.z < 0. && fETheta != kPi)));
    co |= (!co && (isPointOnXYPlane &&
                   (((fSTheta < 0.5 * kPi) && (fETheta < 0.5 * kPi) && ((0.5 * kPi - fSTheta) > kAngTol) &&
                     ((0.5 * kPi - fETheta) > kTolerance)) ||
                    ((fSTheta > 0.5 * kPi && fETheta > 0.5 * kPi) && ((fSTheta - 0.5 * kPi) > kAngTol) &&
                     ((fETheta - 0.5 * kPi) > kTolerance)))));
    if (fSTheta < 0.5 * kPi + halfAngTolerance) {
      if (fETheta < 0.5 * kPi + halfAngTolerance) {
        if (fSTheta < fETheta) {
          const double tolAngMin2 = cone1Radius - 2 * kAngTol * 10.;
          const double tolAngMax2 = cone2Radius + 2 * kAngTol * 10.;
          co |= (!co && ((rho < tolAngMin2) || (rho > tolAngMax2) || (p.z < 0.)));
        }
      }
      if (fETheta > 0.5 * kPi + halfAngTolerance) {
        if (fSTheta < fETheta) {
          const double tolAngMin2 = cone1Radius - 2 * kAngTol * 10.;
          const double tolAngMax2 = cone2Radius - 2 * kAngTol * 10.;
          co |= (!co && (((rho < tolAngMin2) && (p.z > 0.)) || ((rho < tolAngMax2) && (p.z < 0.))));
        }
      }
      if (fETheta >= 0.5 * kPi - halfAngTolerance && fETheta <= 0.5 * kPi + halfAngTolerance)
        co &= !(Abs(p.z) < halfAngTolerance);
    }
    if (fETheta > 0.5 * kPi + halfAngTolerance) {
      if (fSTheta >= 0.5 * kPi - halfAngTolerance && fSTheta <= 0.5 * kPi + halfAngTolerance)
        co &= !(Abs(p.z) < halfAngTolerance);
      if (fSTheta > 0.5 * kPi + halfAngTolerance) {
        if (fSTheta < fETheta) {
          const double tolAngMin2 = cone1Radius + 2 * kAngTol * 10.;
          const double tolAngMax2 = cone2Radius - 2 * kAngTol * 10.;
          co |= (!co && ((rho < tolAngMax2) || (rho > tolAngMin2) || (p.z > 0.)));
        }
      }
    }
    return co;
  }
};

}  // namespace

struct Sphere final : Shape {
  double fRmin, fRmax, fSPhi, fDPhi, fSTheta, fDTheta, eTheta;
  bool fFullPhiSphere = true, fFullThetaSphere = true, fFullSphere = true;
  Wedge fPhiWedge;
  ThetaCone fThetaCone;

  // SphereStruct(pRmin, pRmax, pSPhi, pDPhi, pSTheta, pDTheta) with
  // CheckPhiAngles and CheckThetaAngles (SphereStruct.h:120-210)
  Sphere(double rmin, double rmax, double sphi, double dphi, double stheta, double dtheta)
      : fRmin(rmin), fRmax(rmax), fSPhi(sphi), fDPhi(dphi), fSTheta(stheta), fDTheta(dtheta),
        fPhiWedge(dphi, sphi), fThetaCone(stheta, dtheta)
  {
    if (dphi >= 2 * kPi - kAngTolerance * 0.5) {
      fDPhi = 2 * kPi;
      fSPhi = 0;
      fFullPhiSphere = true;
    } else {
      fFullPhiSphere = false;
      if (dphi > 0) fDPhi = dphi;
      if (sphi < 0)
        fSPhi = 2 * kPi - std::fmod(std::fabs(sphi), 2 * kPi);
      else
        fSPhi = std::fmod(sphi, 2 * kPi);
      if (fSPhi + fDPhi > 2 * kPi) fSPhi -= 2 * kPi;
    }
    fFullSphere = fFullPhiSphere && fFullThetaSphere;
    if (!((stheta < 0) || (stheta > kPi))) fSTheta = stheta;
    if (dtheta + stheta >= kPi)
      fDTheta = kPi - stheta;
    else if (dtheta > 0)
      fDTheta = dtheta;
    fFullThetaSphere = !(fDTheta - fSTheta < kPi);
    fFullSphere = fFullPhiSphere && fFullThetaSphere;
    eTheta = fSTheta + fDTheta;
  }

  // GenericKernelForContainsAndInside (SphereImplementation.h:78-120)
  void Generic(const Vec3 &p, bool forInside, bool &cin, bool &cout) const
  {
    const double rad2 = p.Mag2();
    if (fRmin != 0.)
      cin = rad2 <= MinusTolSq(fRmax) && rad2 >= PlusTolSq(fRmin);
    else
      cin = rad2 <= MinusTolSq(fRmax);
    if (fRmin != 0.)
      cout = rad2 >= PlusTolSq(fRmax) || rad2 <= MinusTolSq(fRmin);
    else
      cout = rad2 >= PlusTolSq(fRmax);
    if (!fFullPhiSphere) {
      bool ci = false, co = false;
      fPhiWedge.GenericKernel(p, forInside, ci, co);
      cout |= co;
      if (forInside) cin &= ci;
    }
    if (!fFullThetaSphere) {
      const bool ci = forInside ? fThetaCone.IsCompletelyInside(p) : false;
      const bool co = fThetaCone.IsCompletelyOutside(p);
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

  // SphereUtilities::IsCompletelyOutside / IsCompletelyInside (:112-140)
  bool UtilCompletelyOutside(const Vec3 &p) const
  {
    const double rad = p.Mag();
    const bool outsideRadiusRange = (rad > (fRmax + kTolerance)) || (rad < (fRmin - kTolerance));
    bool inPhi = false, outPhi = false;
    fPhiWedge.GenericKernel(p, true, inPhi, outPhi);
    const bool outsideThetaRange = fThetaCone.IsCompletelyOutside(p);
    return outsideRadiusRange || outPhi || outsideThetaRange;
  }
  bool UtilCompletelyInside(const Vec3 &p) const
  {
    const double rad = p.Mag();
    const bool insideRadiusRange = (rad < (fRmax - kTolerance)) && (rad > (fRmin + kTolerance));
    bool inPhi = false, outPhi = false;
    fPhiWedge.GenericKernel(p, true, inPhi, outPhi);
    const bool insideThetaRange = fThetaCone.IsCompletelyInside(p);
    return insideRadiusRange && inPhi && insideThetaRange;
  }

  bool OnInnerRadius(const Vec3 &p) const
  {
    const double m2 = p.Mag2();
    return m2 <= PlusTolSq(fRmin) && m2 >= MinusTolSq(fRmin);
  }
  bool OnOuterRadius(const Vec3 &p) const
  {
    const double m2 = p.Mag2();
    return m2 <= PlusTolSq(fRmax) && m2 >= MinusTolSq(fRmax);
  }
  bool WedgeOnSurfaceMovingOut(const Vec3 &p, const Vec3 &d, bool start, bool movingOut) const
  {
    const Vec3 a = start ? fPhiWedge.fAlong1 : fPhiWedge.fAlong2;
    const Vec3 n = start ? fPhiWedge.fNormal1 : fPhiWedge.fNormal2;
    if (!Wedge::IsOnSurfaceGeneric(a, n, p)) return false;
    const double dn = Dot(d, -n);
    return movingOut ? dn > 0.005 * kHalfTolerance : dn < 0.005 * kHalfTolerance;
  }
  // SphereUtilities::IsPointOnSurfaceAndMovingOut (:164-199)
  bool OnSurfaceAndMovingOut(const Vec3 &p, const Vec3 &d, bool movingOut) const
  {
    const double kSqrtTolerance = 3.1622777e-5;  // VecGeom/base/Math.h:29
    auto radial = [&](bool inner) {
      if (inner)
        return OnInnerRadius(p) && (movingOut ? Dot(d, -p) > kSqrtTolerance : Dot(d, -p) < -kSqrtTolerance);
      return OnOuterRadius(p) && (movingOut ? Dot(d, p) > kSqrtTolerance : Dot(d, p) < -kSqrtTolerance);
    };
    const bool tempOuterRad = radial(false);
    bool tempInnerRad = false, tempStartPhi = false, tempEndPhi = false, tempStartTheta = false, tempEndTheta = false;
    if (fRmin != 0.) tempInnerRad = radial(true);
    if (fDPhi < (kTwoPi - kHalfTolerance)) {
      tempStartPhi = WedgeOnSurfaceMovingOut(p, d, true, movingOut);
      tempEndPhi = WedgeOnSurfaceMovingOut(p, d, false, movingOut);
    }
    if (fDTheta < (kPi - kHalfTolerance)) {
      tempStartTheta = fThetaCone.IsPointOnSurfaceAndMovingOut(p, d, true, movingOut);
      tempEndTheta = fThetaCone.IsPointOnSurfaceAndMovingOut(p, d, false, movingOut);
    }
    const double m2 = p.Mag2();
    return ((tempOuterRad || tempInnerRad) && fPhiWedge.Contains(p) && fThetaCone.Contains(p)) ||
           ((tempStartPhi || tempEndPhi) && (m2 >= fRmin * fRmin) && (m2 <= fRmax * fRmax) &&
            fThetaCone.Contains(p)) ||
           ((tempStartTheta || tempEndTheta) && (m2 >= fRmin * fRmin) && (m2 <= fRmax * fRmax) &&
            fPhiWedge.Contains(p));
  }

  // GetMinDistFromPhi<DistToIn> (SphereImplementation.h:285-330)
  void MinDistFromPhi(const Vec3 &p, const Vec3 &d, bool toIn, double &distance) const
  {
    double distPhi1 = kInfLength, distPhi2 = kInfLength;
    if (toIn)
      fPhiWedge.DistanceToIn(p, d, distPhi1, distPhi2);
    else
      fPhiWedge.DistanceToOut(p, d, distPhi1, distPhi2);
    auto check = [&](double dist) {
      const Vec3 tmpPt = p + d * dist;
      const double rad2 = tmpPt.Mag2();
      const bool tempCond =
          ((dist == distPhi1) && Wedge::IsOnSurfaceGeneric(fPhiWedge.fAlong1, fPhiWedge.fNormal1, tmpPt)) ||
          ((dist == distPhi2) && Wedge::IsOnSurfaceGeneric(fPhiWedge.fAlong2, fPhiWedge.fNormal2, tmpPt));
      return tempCond && (rad2 > fRmin * fRmin) && (rad2 < fRmax * fRmax) && fThetaCone.Contains(tmpPt);
    };
    double dist = Min(distPhi1, distPhi2);
    const bool containsCond1 = check(dist);
    if (containsCond1) distance = Min(dist, distance);
    dist = Max(distPhi1, distPhi2);
    const bool containsCond2 = check(dist);
    if (!containsCond1 && containsCond2) distance = Min(dist, distance);
  }

  double DistanceToIn(const Vec3 &point, const Vec3 &direction, double) const override  // :128-222
  {
    double distance = kInfLength;
    const double rad2 = point.Mag2();
    const double pDotV3d = Dot(point, direction);
    double c = rad2 - fRmax * fRmax;
    if (UtilCompletelyInside(point)) return -1.0;
    if (OnSurfaceAndMovingOut(point, direction, false)) return 0.0;
    double sd1 = kInfLength, sd2 = kInfLength;
    double d2 = (pDotV3d * pDotV3d - c);
    if (d2 < 0. || ((c > 0.) && (pDotV3d > 0.))) return distance;
    if (d2 >= 0.) sd1 = (-pDotV3d - Sqrt(Abs(d2)));
    double outerDist = kInfLength, innerDist = kInfLength;
    if (fFullSphere) {
      if (sd1 >= 0.) outerDist = sd1;
    } else {
      const Vec3 tmpPt = point + direction * sd1;
      if (fPhiWedge.Contains(tmpPt) && fThetaCone.Contains(tmpPt) && (sd1 >= 0.)) outerDist = sd1;
    }
    if (fRmin != 0.) {
      c = rad2 - fRmin * fRmin;
      d2 = pDotV3d * pDotV3d - c;
      if (d2 >= 0.) sd2 = (-pDotV3d + Sqrt(Abs(d2)));
      if (fFullSphere) {
        if (sd2 >= 0.) innerDist = sd2;
      } else {
        const Vec3 tmpPt = point + direction * sd2;
        if ((sd2 >= 0.) && fPhiWedge.Contains(tmpPt) && fThetaCone.Contains(tmpPt)) innerDist = sd2;
      }
    }
    distance = Min(outerDist, innerDist);
    if (!fFullPhiSphere) MinDistFromPhi(point, direction, true, distance);
    double distThetaMin = kInfLength;
    if (!fFullThetaSphere) {
      bool intsect1 = false, intsect2 = false;
      double distTheta1 = kInfLength, distTheta2 = kInfLength;
      fThetaCone.DistanceToIn(point, direction, distTheta1, distTheta2, intsect1, intsect2);
      const Vec3 coneIntSecPt1 = point + direction * distTheta1;
      const double distCone1 = coneIntSecPt1.Mag2();
      const Vec3 coneIntSecPt2 = point + direction * distTheta2;
      const double distCone2 = coneIntSecPt2.Mag2();
      bool isValidCone1 = (distCone1 >= fRmin * fRmin && distCone1 <= fRmax * fRmax) && intsect1;
      bool isValidCone2 = (distCone2 >= fRmin * fRmin && distCone2 <= fRmax * fRmax) && intsect2;
      if (!fFullPhiSphere) {
        isValidCone1 &= fPhiWedge.Contains(coneIntSecPt1);
        isValidCone2 &= fPhiWedge.Contains(coneIntSecPt2);
      }
      if (isValidCone2 && !isValidCone1) distThetaMin = distTheta2;
      if (isValidCone1 && !isValidCone2) distThetaMin = distTheta1;
      if (isValidCone1 && isValidCone2) distThetaMin = Min(distTheta1, distTheta2);
    }
    distance = Min(distThetaMin, distance);
    const Vec3 directDir = Vec3(0., 0., 0.) - point;
    const double newDist = directDir.Mag();
    const Vec3 u = directDir.Unit();
    if ((fSTheta > kHalfTolerance || eTheta < (kPi - kHalfTolerance)) && (Abs(u.x - direction.x) < kHalfTolerance) &&
        (Abs(u.y - direction.y) < kHalfTolerance) && (Abs(u.z - direction.z) < kHalfTolerance))
      distance = Min(distance, newDist);
    return distance;
  }

  double DistanceToOut(const Vec3 &point, const Vec3 &direction, double) const override  // :227-280
  {
    double distance = kInfLength;
    const double pDotV3d = Dot(point, direction);
    const double rad2 = point.Mag2();
    double c = rad2 - fRmax * fRmax;
    double sd1 = kInfLength, sd2 = kInfLength;
    if (UtilCompletelyOutside(point)) return -1.0;
    if (OnSurfaceAndMovingOut(point, direction, true)) return 0.0;
    double d2 = (pDotV3d * pDotV3d - c);
    if (d2 >= 0.) sd1 = (-pDotV3d + Sqrt(Abs(d2)));
    if (fRmin != 0.) {
      c = rad2 - fRmin * fRmin;
      d2 = (pDotV3d * pDotV3d - c);
      if ((d2 >= 0.) && (pDotV3d < 0.)) sd2 = (-pDotV3d - Sqrt(Abs(d2)));
    }
    const double snxt = Min(sd1, sd2);
    const bool condSemi = ((fSTheta == 0. && eTheta == kPi / 2.) && direction.z >= 0.) ||
                          ((fSTheta == kPi / 2. && eTheta == kPi) && direction.z <= 0.);
    if (condSemi) return snxt;
    double distThetaMin = kInfLength;
    if (!fFullThetaSphere) {
      bool intsect1 = false, intsect2 = false;
      double distTheta1 = kInfLength, distTheta2 = kInfLength;
      fThetaCone.DistanceToOut(point, direction, distTheta1, distTheta2, intsect1, intsect2);
      if (intsect2 && !intsect1) distThetaMin = distTheta2;
      if (!intsect2 && intsect1) distThetaMin = distTheta1;
      if (intsect2 && intsect1) distThetaMin = Min(distTheta1, distTheta2);
    }
    distance = Min(distThetaMin, snxt);
    if (!fFullPhiSphere) {
      if (fDPhi <= kPi) {
        double distPhi1, distPhi2;
        fPhiWedge.DistanceToOut(point, direction, distPhi1, distPhi2);
        const double distPhiMin = Min(distPhi1, distPhi2);
        distance = Min(distPhiMin, distance);
      } else {
        MinDistFromPhi(point, direction, false, distance);
      }
    }
    return distance;
  }

  double SafetyToIn(const Vec3 &point) const override  // :335-372
  {
    const double rad = point.Mag();
    bool ci = false, co = false;
    Generic(point, true, ci, co);
    if (ci) return -1.0;
    if (!ci && !co) return 0.0;
    double safety;
    if (fRmin != 0.) {
      const double safeRMin = fRmin - rad, safeRMax = rad - fRmax;
      safety = (safeRMin > safeRMax) ? safeRMin : safeRMax;
    } else {
      safety = (rad - fRmax);
    }
    if (!fFullPhiSphere) safety = Max(fPhiWedge.SafetyToIn(point), safety);
    if (!fFullThetaSphere) safety = Max(fThetaCone.SafetyToIn(point), safety);
    return safety;
  }

  double SafetyToOut(const Vec3 &point) const override  // :377-413
  {
    const double rad = point.Mag();
    bool ci = false, co = false;
    Generic(point, true, ci, co);
    if (co) return -1.0;
    if (!ci && !co) return 0.0;
    double safety;
    if (fRmin != 0.) {
      const double safeRMin = (rad - fRmin), safeRMax = (fRmax - rad);
      safety = (safeRMin < safeRMax) ? safeRMin : safeRMax;
    } else {
      safety = (fRmax - rad);
    }
    if (!fFullPhiSphere) safety = Min(fPhiWedge.SafetyToOut(point), safety);
    if (!fFullThetaSphere) safety = Min(fThetaCone.SafetyToOut(point), safety);
    return safety;
  }

  // ApproxSurfaceNormalKernel (:423-465)
  Vec3 ApproxNormal(const Vec3 &point) const
  {
    Vec3 norm(0., 0., 0.);
    const double radius = point.Mag();
    const double distRMax = Abs(radius - fRmax);
    double distRMin = kInfLength, distPhi1 = kInfLength, distPhi2 = kInfLength, distTheta1 = kInfLength,
           distTheta2 = kInfLength;
    double distMin = distRMax;
    if (fRmin > 0.) {
      distRMin = Abs(fRmin - radius);
      distMin = Min(distRMin, distRMax);
    }
    const Vec3 n1 = fPhiWedge.GetNormal1(), n2 = fPhiWedge.GetNormal2();
    if (!fFullPhiSphere) {
      distPhi1 = Abs(point.x * n1.x + point.y * n1.y);
      distPhi2 = Abs(point.x * n2.x + point.y * n2.y);
      distMin = Min(Min(distMin, distPhi1), distPhi2);
    }
    if (!fFullThetaSphere) {
      const double rho = point.Perp();
      distTheta1 = fThetaCone.DistanceToLine(fThetaCone.slope1, rho, point.z);
      distTheta2 = fThetaCone.DistanceToLine(fThetaCone.slope2, rho, point.z);
      distMin = Min(Min(distMin, distTheta1), distTheta2);
    }
    if (distMin == distRMax) norm = point.Unit();
    if (distMin == distRMin) norm = -point.Unit();
    if (distMin == distPhi1) norm = -n1;
    if (distMin == distPhi2) norm = -n2;
    if (distMin == distTheta1) norm = norm + fThetaCone.GetNormal1(point);
    if (distMin == distTheta2) norm = norm + fThetaCone.GetNormal2(point);
    return norm;
  }

  // Normal (:470-545): the sum of the normals of the surfaces the point is
  // on (starting from (1e-30, 1e-30, 1e-30)), normalised
  bool Normal(const Vec3 &point, Vec3 &out) const override
  {
    Vec3 normal(1e-30, 1e-30, 1e-30);
    bool isPointInside = false, isPointOutside = false;
    Generic(point, true, isPointInside, isPointOutside);
    if (isPointInside || isPointOutside) normal = ApproxNormal(point);
    double noSurfaces = 0.;
    if (OnOuterRadius(point)) {
      noSurfaces = noSurfaces + 1;
      if (!isPointOutside) normal = normal + point.Unit();
    }
    if (fRmin != 0.) {
      if (OnInnerRadius(point)) {
        noSurfaces = noSurfaces + 1;
        if (!isPointOutside) normal = normal - point.Unit();
      }
    }
    if (!fFullPhiSphere) {
      const bool onStart = Wedge::IsOnSurfaceGeneric(fPhiWedge.fAlong1, fPhiWedge.fNormal1, point);
      const bool onEnd = Wedge::IsOnSurfaceGeneric(fPhiWedge.fAlong2, fPhiWedge.fNormal2, point);
      if (onStart) noSurfaces = noSurfaces + 1;
      if (onEnd) noSurfaces = noSurfaces + 1;
      if (!isPointOutside && onStart) normal = normal - fPhiWedge.GetNormal1();
      if (!isPointOutside && onEnd) normal = normal - fPhiWedge.GetNormal2();
    }
    if (!fFullThetaSphere) {
      const bool onStartTheta = fThetaCone.IsOnSurfaceGeneric(point, true);
      const bool onEndTheta = fThetaCone.IsOnSurfaceGeneric(point, false);
      if (onStartTheta) noSurfaces = noSurfaces + 1;
      if (!isPointOutside && onStartTheta) normal = normal + fThetaCone.GetNormal1(point);
      if (onEndTheta) noSurfaces = noSurfaces + 1;
      if (!isPointOutside && onEndTheta) normal = normal + fThetaCone.GetNormal2(point);
      if (!isPointOutside && onStartTheta && onEndTheta && (eTheta <= kPi / 2.)) normal = Vec3(0., 0., -1.);
      if (!isPointOutside && onStartTheta && onEndTheta && (fSTheta >= kPi / 2.)) normal = Vec3(0., 0., 1.);
    }
    out = normal.Normalized();
    return noSurfaces > 0.;
  }
  bool HasNormal() const override { return true; }

  // UnplacedSphere::Extent (source/UnplacedSphere.cpp:76-164)
  void Extent(Vec3 &aMin, Vec3 &aMax) const override
  {
    aMin = Vec3(-fRmax, -fRmax, -fRmax);
    aMax = Vec3(fRmax, fRmax, fRmax);
    const double eT = fSTheta + fDTheta;
    if (fFullSphere) return;
    double st1 = 0., st2 = 0.;
    if (!fFullThetaSphere) {
      aMax.z = ((fSTheta <= kPi / 2.) ? fRmax : fRmin) * std::cos(fSTheta);
      aMin.z = ((eT <= kPi / 2.) ? fRmin : fRmax) * std::cos(eT);
      st1 = std::sin(fSTheta);
      st2 = std::sin(eTheta);
      if (fSTheta <= kPi / 2.) {
        if ((eT) < kPi / 2.) {
          aMax.x = fRmax * st2;
          aMin.x = -fRmax * st2;
          aMax.y = fRmax * st2;
          aMin.y = -fRmax * st2;
        }
      } else {
        aMax.x = fRmax * st1;
        aMin.x = -fRmax * st1;
        aMax.y = fRmax * st1;
        aMin.y = -fRmax * st1;
      }
    }
    if (!fFullPhiSphere) {
      double Rmax = fRmax;
      if (fSTheta > kPi / 2.) Rmax *= st1;
      if (eT < kPi / 2.) Rmax *= st2;
      const double Rmin = fRmin * Min(st1, st2);
      const double Rin = 0.5 * (Rmax + Rmin);
      const bool phi0out = !fPhiWedge.Contains(Vec3(Rin, 0, 0));
      const bool phi90out = !fPhiWedge.Contains(Vec3(0, Rin, 0));
      const bool phi180out = !fPhiWedge.Contains(Vec3(-Rin, 0, 0));
      const bool phi270out = !fPhiWedge.Contains(Vec3(0, -Rin, 0));
      if (!(phi0out || phi90out || phi180out || phi270out)) return;
      const double Cx = Rmax * std::cos(fSPhi), Dx = Rmax * std::cos(fSPhi + fDPhi);
      const double Cy = Rmax * std::sin(fSPhi), Dy = Rmax * std::sin(fSPhi + fDPhi);
      if (phi0out) aMax.x = Max(Cx, Dx);
      if (phi90out) aMax.y = Max(Cy, Dy);
      if (phi180out) aMin.x = Min(Cx, Dx);
      if (phi270out) aMin.y = Min(Cy, Dy);
      if (fDPhi >= kPi) return;
      const double Ax = Rmin * std::cos(fSPhi), Bx = Rmin * std::cos(fSPhi + fDPhi);
      const double Ay = Rmin * std::sin(fSPhi), By = Rmin * std::sin(fSPhi + fDPhi);
      double temp = Max(Ax, Bx);
      aMax.x = temp > aMax.x ? temp : aMax.x;
      temp = Max(Ay, By);
      aMax.y = temp > aMax.y ? temp : aMax.y;
      temp = Min(Ax, Bx);
      aMin.x = temp < aMin.x ? temp : aMin.x;
      temp = Min(Ay, By);
      aMin.y = temp < aMin.y ? temp : aMin.y;
    }
  }
};

Shape *make_sphere(double rmin, double rmax, double sphi, double dphi, double stheta, double dtheta)
{
  return new Sphere(rmin, rmax, sphi, dphi, stheta, dtheta);
}

}  // namespace orc
