// Cut tube kernels: VecGeom/volumes/kernel/CutTubeImplementation.h:106-330, a
// tube of unbounded length (the UniversalTube kernels of tube.hpp) between two
// cut planes (VecGeom/volumes/CutPlanes.h:100-190 over Plane.h:85-150).
// Scalar semantics: EarlyReturnAllowed() is true, so every early return of
// the reference returns, together with the rounding to zero it does there.
#pragma once
#include "../vgm_math.hpp"
#include "cut_tube_data.h"
#include "tube.hpp"

namespace vgm {

struct CutTubeShape {
  using Data = CutTubeData;
  static constexpr bool kStage = false;  // parameter block staged in LDS
  // waves per SIMD the transport kernel asks the compiler for (1: no request)
  static constexpr int kWavesPerEU = 1;
  // safety is the distance to the surface (1-Lipschitz): carried bounds are valid
  static constexpr bool kSafetyIsDistance = false;
  // the distance-first step is not proven for this family: it takes the FP64 step
  static constexpr bool kSafetyBelowDistance = false;
  // mixed precision: Inside/Contains classified in float away from the surface
  static constexpr bool kFloatClassify = true;
  // mixed precision: the float safety less the guard band bounds the FP64 safety
  static constexpr bool kFloatSafety = true;
  // mixed precision: distance_to_in_mixed / distance_to_out_mixed (float selects
  // the surface, FP64 evaluates it)
  static constexpr bool kMixedDistance = false;

  // Plane::DistPlane (Plane.h:85-90): point . normal + distance
  template <typename R>
  VGM_DEV static R dist_plane(double nx, double ny, double nz, double dd, V3<R> p)
  {
    return (p.x * R(nx) + p.y * R(ny) + p.z * R(nz)) + R(dd);
  }
  template <typename R>
  VGM_DEV static R dist0(const Data &c, V3<R> p)
  {
    return dist_plane<R>(c.n0x, c.n0y, c.n0z, c.d0, p);
  }
  template <typename R>
  VGM_DEV static R dist1(const Data &c, V3<R> p)
  {
    return dist_plane<R>(c.n1x, c.n1y, c.n1z, c.d1, p);
  }

  // CutPlanes::Inside (CutPlanes.h:112-121)
  template <typename R>
  VGM_DEV static int planes_inside(const Data &c, V3<R> p)
  {
    const R d0 = dist0<R>(c, p), d1 = dist1<R>(c, p);
    int in = (d0 < R(0) && d1 < R(0)) ? kInsideV : kOutsideV;
    if (vabs(d0) < R(kTolerance) || vabs(d1) < R(kTolerance)) in = kSurfaceV;
    return in;
  }

  // Plane::DistanceToIn (Plane.h:104-116): -inf unless moving against the
  // normal from the outer side
  template <typename R>
  VGM_DEV static R plane_in(double nx, double ny, double nz, double dd, V3<R> p, V3<R> d)
  {
    R distance = -inf_length<R>();
    const R ndd = nonzero(d.x * R(nx) + d.y * R(ny) + d.z * R(nz));
    const R saf = dist_plane<R>(nx, ny, nz, dd, p);
    if (ndd < R(0) && saf > R(-kTolerance)) distance = -saf / ndd;
    return distance;
  }

  // Plane::DistanceToOut (Plane.h:118-131)
  template <typename R>
  VGM_DEV static R plane_out(double nx, double ny, double nz, double dd, V3<R> p, V3<R> d)
  {
    R distance = inf_length<R>();
    const R ndd = nonzero(d.x * R(nx) + d.y * R(ny) + d.z * R(nz));
    const R saf = dist_plane<R>(nx, ny, nz, dd, p);
    if (saf > R(kTolerance)) distance = -inf_length<R>();
    if (ndd > R(0) && saf < R(kTolerance)) distance = -saf / ndd;
    return distance;
  }

  // CutPlanes::DistanceToIn (CutPlanes.h:123-137): a plane the ray leaves
  // through counts -kInfLength; the larger of the two
  template <typename R>
  VGM_DEV static R planes_distance_to_in(const Data &c, V3<R> p, V3<R> d)
  {
    R d0 = plane_in<R>(c.n0x, c.n0y, c.n0z, c.d0, p, d);
    if (d.x * R(c.n0x) + d.y * R(c.n0y) + d.z * R(c.n0z) > R(0)) d0 = -inf_length<R>();
    R d1 = plane_in<R>(c.n1x, c.n1y, c.n1z, c.d1, p, d);
    if (d.x * R(c.n1x) + d.y * R(c.n1y) + d.z * R(c.n1z) > R(0)) d1 = -inf_length<R>();
    return vmax(d0, d1);
  }

  // CutPlanes::DistanceToOut (:139-147)
  template <typename R>
  VGM_DEV static R planes_distance_to_out(const Data &c, V3<R> p, V3<R> d)
  {
    const R d0 = plane_out<R>(c.n0x, c.n0y, c.n0z, c.d0, p, d);
    const R d1 = plane_out<R>(c.n1x, c.n1y, c.n1z, c.d1, p, d);
    return vmin(d0, d1);
  }

  // Contains (CutTubeImplementation.h:106-121)
  template <typename R>
  VGM_DEV static bool contains(const Data &c, V3<R> p)
  {
    if (!(dist0<R>(c, p) < R(0) && dist1<R>(c, p) < R(0))) return false;
    return TubeShape::contains<R>(c.tube, p);
  }

  // Inside (:123-141): the plane classification wins when it is outside, or
  // surface unless the tube says outside
  template <typename R>
  VGM_DEV static int inside(const Data &c, V3<R> p)
  {
    const int icp = planes_inside<R>(c, p);
    if (icp == kOutsideV) return kOutsideV;
    int in = TubeShape::inside<R>(c.tube, p);
    if (icp == kSurfaceV && in != kOutsideV) in = icp;
    return in;
  }

  // DistanceToIn (:143-169): far points first move to 100 x Max(rmax, dz)
  template <typename R>
  VGM_DEV static R distance_to_in(const Data &c, V3<R> pt, V3<R> dir, R step_max)
  {
    V3<R> point = pt;
    const R ptDist = vsqrt(point.x * point.x + point.y * point.y + point.z * point.z);
    R distToMove = R(0);
    if (ptDist > R(100. * c.maxval)) {
      distToMove = ptDist - R(100. * c.maxval);
      point.x = point.x + distToMove * dir.x;
      point.y = point.y + distToMove * dir.y;
      point.z = point.z + distToMove * dir.z;
    }
    R distance = distance_to_in_kernel<R>(c, point, dir, step_max);
    distance += distToMove;
    return distance;
  }

  // DistanceToInKernel (:171-268), USE_CONV_WRONG_SIDE
  template <typename R>
  VGM_DEV static R distance_to_in_kernel(const Data &c, V3<R> point, V3<R> dir, R step_max)
  {
    R distance = inf_length<R>();
    int icp = planes_inside<R>(c, point);
    int instart = TubeShape::inside<R>(c.tube, point);
    if (icp == kOutsideV || (icp == kSurfaceV && instart != kOutsideV)) instart = icp;
    if (instart == kInsideV) distance = R(-1.);
    R dplanes = planes_distance_to_in<R>(c, point, dir);
    // CutPlanes::SafetyToIn: the larger plane distance
    const R safplanes = vmax(dist0<R>(c, point), dist1<R>(c, point));
    const bool hitplanes = vabs(dplanes) < step_max && safplanes > R(-kTolerance);
    if (icp != kInsideV && !hitplanes) return distance;
    // propagate the tracks that hit a plane onto it
    V3<R> propagated = point;
    if (!hitplanes) dplanes = R(0);
    if (hitplanes) {
      propagated.x += dplanes * dir.x;
      propagated.y += dplanes * dir.y;
      propagated.z += dplanes * dir.z;
      icp = planes_inside<R>(c, propagated);
    }
    if (icp == kOutsideV || instart == kInsideV) return distance;
    // between the planes now: entering through a cut face?
    instart = TubeShape::inside<R>(c.tube, propagated);
    if (hitplanes && instart != kOutsideV) {
      distance = dplanes;
      if (vabs(distance) < R(kTolerance)) distance = R(0);
      return distance;
    }
    // the tube hit must come before the ray leaves the planes
    const R dexit = planes_distance_to_out<R>(c, propagated, dir);
    R dtube = TubeShape::distance_to_in_kernel<R>(c.tube, propagated, dir);
    if (dtube < R(0)) dtube = R(0);
    if (dexit < dtube) dtube = inf_length<R>();
    if ((dtube + dplanes) < step_max) distance = dtube + dplanes;
    return distance;
  }

  // DistanceToOut (:270-289): the nearer of the planes and the tube
  template <typename R>
  VGM_DEV static R distance_to_out(const Data &c, V3<R> point, V3<R> dir, R step_max)
  {
    R distance = planes_distance_to_out<R>(c, point, dir);
    const R dtube = TubeShape::distance_to_out<R>(c.tube, point, dir, step_max);
    if (dtube < distance) distance = dtube;
    return distance;
  }

  // SafetyToIn (:291-305): the larger of the plane and tube safeties
  template <typename R>
  VGM_DEV static R safety_to_in(const Data &c, V3<R> point)
  {
    R safety = vmax(dist0<R>(c, point), dist1<R>(c, point));
    const R saftube = TubeShape::safety_to_in<R>(c.tube, point);
    if (saftube > safety) safety = saftube;
    return safety;
  }

  // SafetyToOut (:307-321): the smaller
  template <typename R>
  VGM_DEV static R safety_to_out(const Data &c, V3<R> point)
  {
    R safety = vmin(-dist0<R>(c, point), -dist1<R>(c, point));
    const R saftube = TubeShape::safety_to_out<R>(c.tube, point);
    if (saftube < safety) safety = saftube;
    return safety;
  }

  // NormalKernel (CutTubeImplementation.h:340-360): the tube's normal where
  // the tube is nearer than the cut planes, else the cut plane of the
  // point's half
  VGM_DEV static bool normal(const Data &c, V3<double> p, V3<double> &n)
  {
    const double safcut = vmin(-dist0<double>(c, p), -dist1<double>(c, p));
    const double saftube = TubeShape::safety_to_out<double>(c.tube, p);
    if (vabs(saftube) < vabs(safcut)) return TubeShape::normal(c.tube, p, n);
    n = p.z < 0 ? V3<double>{c.n0x, c.n0y, c.n0z} : V3<double>{c.n1x, c.n1y, c.n1z};
    return true;
  }
};

}  // namespace vgm
