// Oracle: box (VecGeom/volumes/kernel/BoxImplementation.h:76-167).
#include "core.h"

namespace orc {

struct Box final : Shape {
  Vec3 h;
  explicit Box(double dx, double dy, double dz) : h(dx, dy, dz) {}

  bool Contains(const Vec3 &p) const override { return (p.Abs() - h).MaxC() < 0.0; }  // :76-80

  int Inside(const Vec3 &p) const override  // :84-90
  {
    double dist = (p.Abs() - h).MaxC();
    int in = dist < 0.0 ? kInside : kOutside;
    if (Abs(dist) < kHalfTolerance) in = kSurface;
    return in;
  }

  double DistanceToIn(const Vec3 &p, const Vec3 &d, double) const override  // :109-131
  {
    const Vec3 invDir(1.0 / NonZero(d.x), 1.0 / NonZero(d.y), 1.0 / NonZero(d.z));
    const Vec3 signDir(Sign(d.x), Sign(d.y), Sign(d.z));
    const Vec3 tempIn(-signDir.x * h.x - p.x, -signDir.y * h.y - p.y, -signDir.z * h.z - p.z);
    const Vec3 tempOut(signDir.x * h.x - p.x, signDir.y * h.y - p.y, signDir.z * h.z - p.z);
    const double absOrthogOut =
        Abs(Vec3(signDir.x * tempOut.x, signDir.y * tempOut.y, signDir.z * tempOut.z).MinC());
    const double distOut = Vec3(tempOut.x * invDir.x, tempOut.y * invDir.y, tempOut.z * invDir.z).MinC();
    double distance = Vec3(tempIn.x * invDir.x, tempIn.y * invDir.y, tempIn.z * invDir.z).MaxC();
    if (distance >= distOut || distOut <= kHalfTolerance || absOrthogOut <= kHalfTolerance) distance = kInfLength;
    return distance;
  }

  double DistanceToOut(const Vec3 &p, const Vec3 &d, double) const override  // :136-151
  {
    const Vec3 invDir(1.0 / NonZero(d.x), 1.0 / NonZero(d.y), 1.0 / NonZero(d.z));
    const Vec3 signDir(Sign(d.x), Sign(d.y), Sign(d.z));
    const double safetyIn = (p.Abs() - h).MaxC();
    const Vec3 tempOut(signDir.x * h.x - p.x, signDir.y * h.y - p.y, signDir.z * h.z - p.z);
    double distance = Vec3(tempOut.x * invDir.x, tempOut.y * invDir.y, tempOut.z * invDir.z).MinC();
    if (safetyIn > kHalfTolerance) distance = -1.0;
    return distance;
  }

  double SafetyToIn(const Vec3 &p) const override { return (p.Abs() - h).MaxC(); }    // :156-159
  double SafetyToOut(const Vec3 &p) const override { return (h - p.Abs()).MinC(); }   // :164-167

  void Extent(Vec3 &mn, Vec3 &mx) const override
  {
    mn = -h;
    mx = h;
  }

  // NormalKernel (:172-189): the faces within tolerance of the nearest one,
  // averaged on edges and corners
  bool Normal(const Vec3 &p, Vec3 &n) const override
  {
    const Vec3 safety = (p.Abs() - h).Abs();
    const double safmin = safety.MinC();
    const bool valid = safmin < kHalfTolerance;
    Vec3 normal(0., 0., 0.);
    if (safety.x - safmin < kHalfTolerance) normal.x = Sign(p.x);
    if (safety.y - safmin < kHalfTolerance) normal.y = Sign(p.y);
    if (safety.z - safmin < kHalfTolerance) normal.z = Sign(p.z);
    if (normal.Mag2() > 1.0) normal = normal.Normalized();
    n = normal;
    return valid;
  }
  bool HasNormal() const override { return true; }
};

Shape *make_box(double dx, double dy, double dz) { return new Box(dx, dy, dz); }

}  // namespace orc
