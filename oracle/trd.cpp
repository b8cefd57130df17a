// Oracle: trd, UniversalTrd (runtime checkVaryingY = dy1 != dy2; HasVaryingY =
// kUnknown so UnplacedInside always takes the line-orientation test for y),
// scalar backend.  VecGeom/volumes/TrdStruct.h (CalculateCached; note that
// fToleranceY is computed from the x extents, as in the reference) and
// VecGeom/volumes/kernel/TrdImplementation.h:60-522.
#include "core.h"

namespace orc {

struct Trd final : Shape {
  double fDX1, fDX2, fDY1, fDY2, fDZ;
  double fX2minusX1, fY2minusY1, fHalfX1plusX2, fHalfY1plusY2, fCalfX, fCalfY, fSecxz, fSecyz;
  double fToleranceX, fToleranceY, fFx, fFy;

  Trd(double x1, double x2, double y1, double y2, double z) : fDX1(x1), fDX2(x2), fDY1(y1), fDY2(y2), fDZ(z)
  {
    fX2minusX1 = fDX2 - fDX1;
    fY2minusY1 = fDY2 - fDY1;
    fHalfX1plusX2 = 0.5 * (fDX1 + fDX2);
    fHalfY1plusY2 = 0.5 * (fDY1 + fDY2);
    fFx = 0.5 * (fDX1 - fDX2) / fDZ;
    fFy = 0.5 * (fDY1 - fDY2) / fDZ;
    fSecxz = std::sqrt(1 + fFx * fFx);
    fSecyz = std::sqrt(1 + fFy * fFy);
    fCalfX = 1. / Sqrt(1.0 + fFx * fFx);
    fCalfY = 1. / Sqrt(1.0 + fFy * fFy);
    fToleranceX = kTolerance * Sqrt(fX2minusX1 * fX2minusX1 + 4 * fDZ * fDZ);
    fToleranceY = kTolerance * Sqrt(fX2minusX1 * fX2minusX1 + 4 * fDZ * fDZ);
  }

  bool varY() const { return fDY1 != fDY2; }

  // TrdUtilities::FaceTrajectoryIntersection (:83-138)
  bool Face(const Vec3 &pos, const Vec3 &dir, bool forY, bool mirrored, bool toInside, double &dist) const
  {
    double alongV, posV, dirV, posK, dirK, fV, fK, halfKplus, v1;
    if (forY) {
      alongV = fY2minusY1;
      v1 = fDY1;
      posV = pos.y;
      posK = pos.x;
      dirV = dir.y;
      dirK = dir.x;
      fK = fFx;
      fV = fFy;
      halfKplus = fHalfX1plusX2;
    } else {
      alongV = fX2minusX1;
      v1 = fDX1;
      posV = pos.x;
      posK = pos.y;
      dirV = dir.x;
      dirK = dir.y;
      fK = fFy;
      fV = fFx;
      halfKplus = fHalfY1plusY2;
    }
    if (mirrored) {
      posV *= -1.;
      dirV *= -1.;
    }
    double ndotv = dirV + fV * dir.z;
    bool ok = toInside ? ndotv < 0. : ndotv > 0.;
    if (!ok) return false;
    double alongZ = 2.0 * fDZ;
    dist = (alongZ * (posV - v1) - alongV * (pos.z + fDZ)) / (dir.z * alongV - dirV * alongZ + 1e-30);
    ok &= dist > (0. - kHalfTolerance);
    if (ok) {
      double hitz = pos.z + dist * dir.z;
      ok &= Abs(hitz) <= fDZ;
      double hitk = posK + dist * dirK;
      double dK = halfKplus - fK * hitz;
      ok &= Abs(hitk) <= dK;
      if (ok & (Abs(dist) < kHalfTolerance)) dist = 0.0;
    }
    return ok;
  }

  // TrdUtilities::Safety (:143-170)
  double Safety(const Vec3 &pos, bool inside) const
  {
    double dist = fDZ - Abs(pos.z);
    double distx = fHalfX1plusX2 - fFx * pos.z;
    double safx = (distx - Abs(pos.x)) * fCalfX;
    if (distx >= 0 && safx < dist) dist = safx;
    if (varY()) {
      double disty = fHalfY1plusY2 - fFy * pos.z;
      double safy = (disty - Abs(pos.y)) * fCalfY;
      if (disty >= 0 && safy < dist) dist = safy;
    } else {
      double safy = fDY1 - Abs(pos.y);
      if (safy < dist) dist = safy;
    }
    if (!inside) dist = -dist;
    return dist;
  }

  // TrdUtilities::UnplacedInside (:173-215)
  void UInside(const Vec3 &p, bool surfaceT, bool &cin, bool &cout) const
  {
    double pzPlusDz = p.z + fDZ;
    cout = Abs(p.z) > (surfaceT ? fDZ + kHalfTolerance : fDZ);
    if (surfaceT) cin = Abs(p.z) < fDZ - kHalfTolerance;
    double cross = fX2minusX1 * pzPlusDz - (2.0 * fDZ) * (Abs(p.x) - fDX1);
    if (surfaceT) {
      cout |= cross < -fToleranceX;
      cin &= cross > fToleranceX;
    } else {
      cout |= cross < 0;
    }
    cross = fY2minusY1 * pzPlusDz - (2.0 * fDZ) * (Abs(p.y) - fDY1);
    if (surfaceT) {
      cout |= cross < -fToleranceY;
      cin &= cross > fToleranceY;
    } else {
      cout |= cross < 0;
    }
  }

  bool Contains(const Vec3 &p) const override
  {
    bool ci = false, co = false;
    UInside(p, false, ci, co);
    return !co;
  }

  int Inside(const Vec3 &p) const override
  {
    bool ci = false, co = false;
    UInside(p, true, ci, co);
    int r = kSurface;
    if (ci) r = kInside;
    if (co) r = kOutside;
    return r;
  }

  static double Clamp0(double d) { return Abs(d) < kHalfTolerance ? 0.0 : d; }

  double DistanceToIn(const Vec3 &point, const Vec3 &direction, double) const override  // :303-394
  {
    double distance = kInfLength;
    bool inz = Abs(point.z) < fDZ - kHalfTolerance;
    double distx = fHalfX1plusX2 - fFx * point.z;
    bool inx = (distx - Abs(point.x)) * fCalfX > (0. + kHalfTolerance);
    double disty;
    bool iny;
    if (varY()) {
      disty = fHalfY1plusY2 - fFy * point.z;
      iny = (disty - Abs(point.y)) * fCalfY > (0. + kHalfTolerance);
    } else {
      disty = Abs(point.y) - fDY1;
      iny = disty < (0. - kHalfTolerance);
    }
    bool inside = inx & iny & inz;
    if (inside) distance = -1.;
    bool done = inside;
    bool okz = point.z * direction.z < 0.;
    okz &= !inz;
    if (okz) {
      double distz = (Abs(point.z) - fDZ) / Abs(direction.z);
      double hitx = Abs(point.x + distz * direction.x);
      double hity = Abs(point.y + distz * direction.y);
      bool okzt = point.z > (fDZ - kHalfTolerance) && hitx <= fDX2 && hity <= fDY2;
      bool okzb = point.z < (-fDZ + kHalfTolerance) && hitx <= fDX1 && hity <= fDY1;
      okz &= (okzt | okzb);
      if (okz) distance = distz;
    }
    done |= okz;
    if (done) return Clamp0(distance);
    bool okx = false;
    if (!inx) {
      okx = Face(point, direction, false, false, true, distx);
      if (okx) distance = distx;
      okx = Face(point, direction, false, true, true, distx);
      if (okx) distance = distx;
    }
    done |= okx;
    if (done) return Clamp0(distance);
    if (varY()) {
      if (!iny) {
        bool oky = Face(point, direction, true, false, true, disty);
        if (oky) distance = disty;
        oky = Face(point, direction, true, true, true, disty);
        if (oky) distance = disty;
      }
    } else {
      if (!iny) {
        disty /= Abs(direction.y);
        double zhit = point.z + disty * direction.z;
        double xhit = point.x + disty * direction.x;
        double dx = fHalfX1plusX2 - fFx * zhit;
        bool oky = point.y * direction.y < 0 && disty > -kHalfTolerance && Abs(xhit) < dx && Abs(zhit) < fDZ;
        if (oky) distance = disty;
      }
    }
    return Clamp0(distance);
  }

  double DistanceToOut(const Vec3 &point, const Vec3 &dir, double) const override  // :399-500
  {
    double distance = 0.0;
    double invdir = 1. / Abs(dir.z + 1e-30);
    double safz = fDZ - Abs(point.z);
    bool out = safz < (0. - kHalfTolerance);
    double distx = fHalfX1plusX2 - fFx * point.z;
    out |= (distx - Abs(point.x)) * fCalfX < (0. - kHalfTolerance);
    double disty;
    if (varY()) {
      disty = fHalfY1plusY2 - fFy * point.z;
      out |= (disty - Abs(point.y)) * fCalfY < (0. - kHalfTolerance);
    } else {
      disty = fDY1 - Abs(point.y);
      out |= disty < (0. - kHalfTolerance);
    }
    if (out) return -1.;
    if (dir.z > 0.) {
      double distz = (fDZ - point.z) * invdir;
      double hitx = Abs(point.x + distz * dir.x);
      double hity = Abs(point.y + distz * dir.y);
      if (hitx <= fDX2 && hity <= fDY2) return Clamp0(distz);
    }
    if (dir.z < 0.) {
      double distz = (point.z + fDZ) * invdir;
      double hitx = Abs(point.x + distz * dir.x);
      double hity = Abs(point.y + distz * dir.y);
      if (hitx <= fDX1 && hity <= fDY1) return Clamp0(distz);
    }
    if (Face(point, dir, false, false, false, distx)) return Clamp0(distx);
    if (Face(point, dir, false, true, false, distx)) return Clamp0(distx);
    if (varY()) {
      if (Face(point, dir, true, false, false, disty)) return Clamp0(disty);
      if (Face(point, dir, true, true, false, disty)) distance = disty;
    } else {
      double plane = fDY1;
      if (dir.y < 0.) plane = -fDY1;
      disty = (plane - point.y) / dir.y;
      double zhit = point.z + disty * dir.z;
      double xhit = point.x + disty * dir.x;
      double dx = fHalfX1plusX2 - fFx * zhit;
      if (Abs(xhit) < dx && Abs(zhit) < fDZ) distance = disty;
    }
    return Clamp0(distance);
  }

  double SafetyToIn(const Vec3 &p) const override { return Safety(p, false); }
  double SafetyToOut(const Vec3 &p) const override { return Safety(p, true); }

  // UnplacedTrd::Normal (source/UnplacedTrd.cpp:277-352)
  bool Normal(const Vec3 &point, Vec3 &norm) const override
  {
    int noSurfaces = 0;
    Vec3 sumnorm(0., 0., 0.), vecnorm(0., 0., 0.);
    const double distz = Abs(Abs(point[2]) - fDZ);
    const double xnorm = 1.0 / std::sqrt(4 * fDZ * fDZ + (fDX2 - fDX1) * (fDX2 - fDX1));
    const double ynorm = 1.0 / std::sqrt(4 * fDZ * fDZ + (fDY2 - fDY1) * (fDY2 - fDY1));
    double distmx = -2.0 * fDZ * point[0] - (fDX2 - fDX1) * point[2] - fDZ * (fDX1 + fDX2);
    distmx *= xnorm;
    double distpx = 2.0 * fDZ * point[0] - (fDX2 - fDX1) * point[2] - fDZ * (fDX1 + fDX2);
    distpx *= xnorm;
    double distmy = -2.0 * fDZ * point[1] - (fDY2 - fDY1) * point[2] - fDZ * (fDY1 + fDY2);
    distmy *= ynorm;
    double distpy = 2.0 * fDZ * point[1] - (fDY2 - fDY1) * point[2] - fDZ * (fDY1 + fDY2);
    distpy *= ynorm;
    double safmin = Min(Abs(distmx), Abs(distpx));
    safmin = Min(safmin, Min(Abs(distmy), Abs(distpy)));
    safmin = Min(safmin, Abs(distz));
    const bool valid = safmin <= kHalfTolerance;
    if (Abs(distmx) - safmin <= kHalfTolerance) {
      noSurfaces++;
      sumnorm += Vec3(-2.0 * fDZ, 0.0, -(fDX2 - fDX1)) * xnorm;
    }
    if (Abs(distpx) - safmin <= kHalfTolerance) {
      noSurfaces++;
      sumnorm += Vec3(2.0 * fDZ, 0.0, -(fDX2 - fDX1)) * xnorm;
    }
    if (Abs(distpy) - safmin <= kHalfTolerance) {
      noSurfaces++;
      sumnorm += Vec3(0.0, 2.0 * fDZ, -(fDY2 - fDY1)) * ynorm;
    }
    if (Abs(distmy) - safmin <= kHalfTolerance) {
      noSurfaces++;
      sumnorm += Vec3(0.0, -2.0 * fDZ, -(fDY2 - fDY1)) * ynorm;
    }
    if (Abs(distz) - safmin <= kHalfTolerance) {
      noSurfaces++;
      if (point[2] >= 0.)
        sumnorm += Vec3(0., 0., 1.);
      else
        sumnorm += Vec3(0., 0., -1.);
    }
    if (noSurfaces == 0)
      vecnorm = Vec3(0., 0., 1.);
    else if (noSurfaces == 1)
      vecnorm = sumnorm;
    else
      vecnorm = sumnorm.Unit();
    norm = vecnorm;
    return valid;
  }
  bool HasNormal() const override { return true; }

  void Extent(Vec3 &mn, Vec3 &mx) const override
  {
    double x = Max(fDX1, fDX2), y = Max(fDY1, fDY2);
    mn = Vec3(-x, -y, -fDZ);
    mx = Vec3(x, y, fDZ);
  }
};

Shape *make_trd(double x1, double x2, double y1, double y2, double z) { return new Trd(x1, x2, y1, y2, z); }

}  // namespace orc
