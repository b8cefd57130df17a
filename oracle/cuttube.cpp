// Oracle: cut tube, a tube of unbounded length between two cut planes.
// VecGeom/volumes/kernel/CutTubeImplementation.h:106-321 (Contains, Inside,
// DistanceToIn with DistanceToInKernel under USE_CONV_WRONG_SIDE,
// DistanceToOut, SafetyToIn, SafetyToOut), VecGeom/volumes/CutTubeStruct.h:38-49,
// VecGeom/volumes/CutPlanes.h:100-190, VecGeom/volumes/Plane.h:30-150,
// source/UnplacedCutTube.cpp:128-172 (Extent).  Scalar backend:
// EarlyReturnAllowed() is true, so every early return returns.  TEST ONLY.
#include <cfloat>

#include "tube.h"

namespace orc {

// Plane (VecGeom/volumes/Plane.h)
struct Plane {
  Vec3 fNormal;
  double fDistance = 0;

  void Set(const Vec3 &normal, const Vec3 &origin)  // :30-37
  {
    Vec3 fixedNormal(normal);
    for (int i = 0; i < 3; ++i)  // Vector3D::FixZeroes
      if (Abs(fixedNormal[i]) < kTolerance) fixedNormal[i] = 0.0;
    const double inverseLength = 1. / fixedNormal.Mag();
    fNormal = Vec3(inverseLength * fixedNormal.x, inverseLength * fixedNormal.y, inverseLength * fixedNormal.z);
    fDistance = inverseLength * -(fixedNormal.x * origin.x + fixedNormal.y * origin.y + fixedNormal.z * origin.z);
  }
  static double Dot(const Vec3 &a, const Vec3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
  double DistPlane(const Vec3 &p) const { return Dot(p, fNormal) + fDistance; }  // :85-90
  double DistanceToIn(const Vec3 &p, const Vec3 &d) const  // :104-116
  {
    const double ndd = NonZero(Dot(d, fNormal));
    const double saf = DistPlane(p);
    if (ndd < 0. && saf > -kTolerance) return -saf / ndd;
    return -kInfLength;
  }
  double DistanceToOut(const Vec3 &p, const Vec3 &d) const  // :118-131
  {
    double distance = kInfLength;
    const double ndd = NonZero(Dot(d, fNormal));
    const double saf = DistPlane(p);
    const bool valid = ndd > 0. && saf < kTolerance;
    if (saf > kTolerance) distance = -kInfLength;
    if (!valid) return distance;
    return -saf / ndd;
  }
};

// Vector3D::Unit (VecGeom/base/Vector3D.h:284-290)
static Vec3 Unit(const Vec3 &v)
{
  const double s = std::sqrt(v.Mag2() + DBL_MIN);
  return Vec3(v.x / s, v.y / s, v.z / s);
}

struct CutTube final : Shape {
  double fDz, fMaxVal, fRmin, fRmax, fSphi, fDphi;
  Tube fTube;  // TubeStruct(rmin, rmax, kInfLength, sphi, dphi)
  Plane fCut[2];

  CutTube(double rmin, double rmax, double z, double sphi, double dphi, const Vec3 &bot, const Vec3 &top)
      : fDz(z), fMaxVal(Max(rmax, z)), fRmin(rmin), fRmax(rmax), fSphi(sphi), fDphi(dphi),
        fTube(rmin, rmax, kInfLength, sphi, dphi)
  {
    fCut[0].Set(Unit(bot), Vec3(0., 0., -z));
    fCut[1].Set(Unit(top), Vec3(0., 0., z));
  }

  // CutPlanes::Inside (CutPlanes.h:112-121)
  int PlanesInside(const Vec3 &p) const
  {
    const double d0 = fCut[0].DistPlane(p), d1 = fCut[1].DistPlane(p);
    int inside = (d0 < 0.0 && d1 < 0.0) ? kInside : kOutside;
    if (Abs(d0) < kTolerance || Abs(d1) < kTolerance) inside = kSurface;
    return inside;
  }
  // CutPlanes::DistanceToIn (:123-137)
  double PlanesDistanceToIn(const Vec3 &p, const Vec3 &d) const
  {
    double d0 = fCut[0].DistanceToIn(p, d);
    if (Plane::Dot(d, fCut[0].fNormal) > 0.) d0 = -kInfLength;
    double d1 = fCut[1].DistanceToIn(p, d);
    if (Plane::Dot(d, fCut[1].fNormal) > 0.) d1 = -kInfLength;
    return Max(d0, d1);
  }
  double PlanesDistanceToOut(const Vec3 &p, const Vec3 &d) const  // :139-147
  {
    return Min(fCut[0].DistanceToOut(p, d), fCut[1].DistanceToOut(p, d));
  }
  double PlanesSafetyToIn(const Vec3 &p) const { return Max(fCut[0].DistPlane(p), fCut[1].DistPlane(p)); }
  double PlanesSafetyToOut(const Vec3 &p) const { return Min(-fCut[0].DistPlane(p), -fCut[1].DistPlane(p)); }

  bool Contains(const Vec3 &p) const override  // :106-121
  {
    const bool inside_cutplanes = fCut[0].DistPlane(p) < 0. && fCut[1].DistPlane(p) < 0.;
    if (!inside_cutplanes) return false;
    return fTube.Contains(p) && inside_cutplanes;
  }

  int Inside(const Vec3 &p) const override  // :123-141
  {
    const int inside_cutplanes = PlanesInside(p);
    if (inside_cutplanes == kOutside) return kOutside;
    int inside = fTube.Inside(p);
    if (inside_cutplanes == kOutside || (inside_cutplanes == kSurface && inside != kOutside)) inside = inside_cutplanes;
    return inside;
  }

  double DistanceToIn(const Vec3 &pointT, const Vec3 &dir, double stepMax) const override  // :143-169
  {
    Vec3 point = pointT;
    const double ptDist = point.Mag();
    double distToMove = 0.;
    const double order = 100.;
    if (ptDist > order * fMaxVal) {
      distToMove = ptDist - order * fMaxVal;
      point = point + dir * distToMove;
    }
    double distance = DistanceToInKernel(point, dir, stepMax);
    distance += distToMove;
    return distance;
  }

  double DistanceToInKernel(const Vec3 &point, const Vec3 &direction, double stepMax) const  // :171-268
  {
    double distance = kInfLength;
    int inside_cutplanes = PlanesInside(point);
    int instart = fTube.Inside(point);
    if (inside_cutplanes == kOutside || (inside_cutplanes == kSurface && instart != kOutside))
      instart = inside_cutplanes;
    if (instart == kInside) distance = -1.;
    double dplanes = PlanesDistanceToIn(point, direction);
    const double safplanes = PlanesSafetyToIn(point);
    const bool hitplanes = Abs(dplanes) < stepMax && safplanes > -kTolerance;
    if (inside_cutplanes != kInside && !hitplanes) return distance;
    Vec3 propagated = point;
    if (!hitplanes) dplanes = 0.;
    if (hitplanes) {
      propagated = propagated + direction * dplanes;
      inside_cutplanes = PlanesInside(propagated);
    }
    bool done = inside_cutplanes == kOutside;
    done |= instart == kInside;
    if (done) return distance;
    instart = fTube.Inside(propagated);
    const bool inside = hitplanes && (instart != kOutside);
    if (inside && !done) distance = dplanes;
    done |= inside;
    if (done) {
      if (Abs(distance) < kTolerance) distance = 0.;
      return distance;
    }
    const double dexit = PlanesDistanceToOut(propagated, direction);
    double dtube = fTube.DistanceToInKernel(propagated, direction);
    if (dtube < 0.) dtube = 0.;
    if (dexit < dtube) dtube = kInfLength;
    if (!done && (dtube + dplanes) < stepMax) distance = dtube + dplanes;
    return distance;
  }

  double DistanceToOut(const Vec3 &point, const Vec3 &direction, double stepMax) const override  // :270-289
  {
    double distance = PlanesDistanceToOut(point, direction);
    const double dtube = fTube.DistanceToOut(point, direction, stepMax);
    if (dtube < distance) distance = dtube;
    return distance;
  }

  double SafetyToIn(const Vec3 &point) const override  // :291-305
  {
    double safety = PlanesSafetyToIn(point);
    const double saftube = fTube.SafetyToIn(point);
    if (saftube > safety) safety = saftube;
    return safety;
  }

  double SafetyToOut(const Vec3 &point) const override  // :307-321
  {
    double safety = PlanesSafetyToOut(point);
    const double saftube = fTube.SafetyToOut(point);
    if (saftube < safety) safety = saftube;
    return safety;
  }

  // ZlimitTop / ZlimitBottom (VecGeom/volumes/UnplacedCutTube.h:135-147)
  double ZlimitBottom(double r, double phi) const
  {
    const Vec3 &n = fCut[0].fNormal;
    return (-fDz - (r / n.z) * (n.x * std::cos(phi) + n.y * std::sin(phi)));
  }
  double ZlimitTop(double r, double phi) const
  {
    const Vec3 &n = fCut[1].fNormal;
    return (fDz - (r / n.z) * (n.x * std::cos(phi) + n.y * std::sin(phi)));
  }

  // NormalKernel (CutTubeImplementation.h:340-360): the tube's normal where
  // the tube is nearer than the cut planes, else the cut plane of the point's
  // half
  bool Normal(const Vec3 &point, Vec3 &normal) const override
  {
    const double safcut = PlanesSafetyToOut(point);
    const double saftube = fTube.SafetyToOut(point);
    if (Abs(saftube) < Abs(safcut)) return fTube.Normal(point, normal);
    normal = point.z < 0 ? fCut[0].fNormal : fCut[1].fNormal;
    return true;
  }
  bool HasNormal() const override { return true; }

  void Extent(Vec3 &aMin, Vec3 &aMax) const override  // source/UnplacedCutTube.cpp:128-172
  {
    const Vec3 &tn = fCut[1].fNormal, &bn = fCut[0].fNormal;
    const double rmax = fRmax, rmin = fTube.fRmin, sphi = fSphi, dphi = fDphi;
    const double dztop = rmax * std::sqrt(1. - tn.z * tn.z) / tn.z;
    const double dzbottom = -rmax * std::sqrt(1. - bn.z * bn.z) / bn.z;
    aMin = Vec3(-rmax, -rmax, -fDz - dzbottom);
    aMax = Vec3(rmax, rmax, fDz + dztop);
    if (dphi == kTwoPi) return;
    const Wedge &w = fTube.fPhiWedge;
    const double c1 = std::cos(sphi), s1 = std::sin(sphi), c2 = std::cos(sphi + dphi), s2 = std::sin(sphi + dphi);
    if (!w.Contains(Vec3(-tn.x, -tn.y, 0.))) {
      aMax.z = Max(ZlimitTop(rmax, sphi), ZlimitTop(rmax, sphi + dphi));
      aMax.z = Max(aMax.z, Max(ZlimitTop(rmin, sphi), ZlimitTop(rmin, sphi + dphi)));
    }
    if (!w.Contains(Vec3(-bn.x, -bn.y, 0.))) {
      aMin.z = Min(ZlimitBottom(rmax, sphi), ZlimitBottom(rmax, sphi + dphi));
      aMin.z = Min(aMin.z, Min(ZlimitBottom(rmin, sphi), ZlimitBottom(rmin, sphi + dphi)));
    }
    if (!w.Contains(Vec3(1., 0., 0.))) {
      aMax.x = Max(rmax * c1, rmax * c2);
      aMax.x = Max(aMax.x, Max(rmin * c1, rmin * c2));
    }
    if (!w.Contains(Vec3(-1., 0., 0.))) {
      aMin.x = Min(rmax * c1, rmax * c2);
      aMin.x = Min(aMin.x, Min(rmin * c1, rmin * c2));
    }
    if (!w.Contains(Vec3(0., 1., 0.))) {
      aMax.y = Max(rmax * s1, rmax * s2);
      aMax.y = Max(aMax.y, Max(rmin * s1, rmin * s2));
    }
    if (!w.Contains(Vec3(0., -1., 0.))) {
      aMin.y = Min(rmax * s1, rmax * s2);
      aMin.y = Min(aMin.y, Min(rmin * s1, rmin * s2));
    }
  }
};

Shape *make_cut_tube(double rmin, double rmax, double z, double sphi, double dphi, const double bot[3],
                     const double top[3])
{
  return new CutTube(rmin, rmax, z, sphi, dphi, Vec3(bot[0], bot[1], bot[2]), Vec3(top[0], top[1], top[2]));
}

}  // namespace orc
