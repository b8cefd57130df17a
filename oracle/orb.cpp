// Oracle: orb, a full sphere of radius r (VecGeom/volumes/kernel/OrbImplementation.h:64-225).
#include "core.h"

namespace orc {

struct Orb final : Shape {
  double r;
  explicit Orb(double radius) : r(radius) {}

  static double mag2(const Vec3 &p) { return p.x * p.x + p.y * p.y + p.z * p.z; }
  static double dot(const Vec3 &p, const Vec3 &d) { return p.x * d.x + p.y * d.y + p.z * d.z; }

  // GenericKernelForContainsAndInside (:103-113)
  bool Contains(const Vec3 &p) const override  // :64-70
  {
    const double tolR = r + kTolerance;
    const bool completelyoutside = mag2(p) >= tolR * tolR;
    return !completelyoutside;
  }

  int Inside(const Vec3 &p) const override  // :72-84
  {
    const double rad2 = mag2(p);
    const double tolIn = r - kTolerance, tolOut = r + kTolerance;
    if (rad2 <= tolIn * tolIn) return kInside;
    if (rad2 >= tolOut * tolOut) return kOutside;
    return kSurface;
  }

  // DetectIntersectionAndCalculateDistance (:190-208)
  bool Intersect(const Vec3 &p, const Vec3 &d, bool forIn, double &distance) const
  {
    const double pDotV = dot(p, d);
    const double d2 = pDotV * pDotV - (mag2(p) - r * r);
    if (forIn) {
      if (d2 >= 0.0 && pDotV <= 0.0) {
        distance = -pDotV - std::sqrt(Abs(d2));
        return true;
      }
      return false;
    }
    if (d2 >= 0.0) {
      distance = -pDotV + std::sqrt(Abs(d2));
      return true;
    }
    return false;
  }

  double DistanceToIn(const Vec3 &p, const Vec3 &d, double) const override  // :122-145
  {
    const double rad = std::sqrt(mag2(p));
    if (rad < r - kTolerance) return -1.0;
    if (rad >= r - kTolerance && rad <= r + kTolerance && dot(p, d) < 0.0) return 0.0;
    double dist = kInfLength;
    return Intersect(p, d, true, dist) ? dist : kInfLength;
  }

  double DistanceToOut(const Vec3 &p, const Vec3 &d, double) const override  // :147-171
  {
    const double rad = std::sqrt(mag2(p));
    if (rad > r + kTolerance) return -1.0;
    if (rad >= r - kTolerance && rad <= r + kTolerance && dot(p, d) > 0.0) return 0.0;
    double dist = kInfLength;
    return Intersect(p, d, false, dist) ? dist : kInfLength;
  }

  double SafetyToIn(const Vec3 &p) const override  // :173-184
  {
    const double rad = std::sqrt(mag2(p));
    if (rad < r - kTolerance) return -1.0;
    if (rad > r - kTolerance && rad < r + kTolerance) return 0.0;
    return rad - r;
  }

  double SafetyToOut(const Vec3 &p) const override  // :186-197
  {
    const double rad = std::sqrt(mag2(p));
    if (rad > r + kTolerance) return -1.0;
    if (rad > r - kTolerance && rad < r + kTolerance) return 0.0;
    return r - rad;
  }

  // NormalKernel (:227-240)
  bool Normal(const Vec3 &p, Vec3 &n) const override
  {
    const double rad2 = mag2(p);
    const double invRadius = 1. / Sqrt(rad2);
    n = Vec3(p.x * invRadius, p.y * invRadius, p.z * invRadius);
    const double tolRMaxO = r + kTolerance, tolRMaxI = r - kTolerance;
    return (rad2 <= tolRMaxO * tolRMaxO) && (rad2 >= tolRMaxI * tolRMaxI);
  }
  bool HasNormal() const override { return true; }

  void Extent(Vec3 &mn, Vec3 &mx) const override
  {
    mn = Vec3(-r, -r, -r);
    mx = Vec3(r, r, r);
  }
};

Shape *make_orb(double r) { return new Orb(r); }

}  // namespace orc
