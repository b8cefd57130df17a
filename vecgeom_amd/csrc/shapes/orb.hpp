// Orb kernels: VecGeom/volumes/kernel/OrbImplementation.h:64-225, a full
// sphere of radius r about the origin.
#pragma once
#include "../vgm_math.hpp"
#include "orb_data.h"

namespace vgm {

struct OrbShape {
  using Data = OrbData;
  static constexpr bool kStage = false;  // parameter block staged in LDS
  // waves per SIMD the transport kernel asks the compiler for (1: no request)
  static constexpr int kWavesPerEU = 1;
  // safety is the distance to the surface (1-Lipschitz): carried bounds are valid
  static constexpr bool kSafetyIsDistance = false;
  // SafetyToIn/Out never exceed DistanceToIn/Out beyond the surface tolerance:
  // the distance-first step takes the FP64 step's decisions
  static constexpr bool kSafetyBelowDistance = true;
  // mixed precision: Inside/Contains classified in float away from the surface
  static constexpr bool kFloatClassify = true;
  // mixed precision: the float safety less the guard band bounds the FP64 safety
  static constexpr bool kFloatSafety = true;
  // mixed precision: distance_to_in_mixed / distance_to_out_mixed (float selects
  // the surface, FP64 evaluates it)
  static constexpr bool kMixedDistance = false;

  template <typename R>
  VGM_DEV static R mag2(V3<R> p)
  {
    return p.x * p.x + p.y * p.y + p.z * p.z;
  }

  // GenericKernelForContainsAndInside (:103-113): inside below (r - tol)^2,
  // outside from (r + tol)^2
  template <typename R>
  VGM_DEV static bool contains(const Data &o, V3<R> p)
  {
    const R tolR = R(o.r) + R(kTolerance);
    return !(mag2(p) >= tolR * tolR);
  }

  template <typename R>
  VGM_DEV static int inside(const Data &o, V3<R> p)
  {
    const R rad2 = mag2(p);
    R tolR = R(o.r) - R(kTolerance);
    const bool completelyinside = rad2 <= tolR * tolR;
    tolR = R(o.r) + R(kTolerance);
    const bool completelyoutside = rad2 >= tolR * tolR;
    int in = kSurfaceV;
    if (completelyoutside) in = kOutsideV;
    if (completelyinside) in = kInsideV;
    return in;
  }

  // DetectIntersectionAndCalculateDistance (:190-208): the entering root
  // -p.d - sqrt(|d2|) for a ray heading in, the exiting root -p.d + sqrt(|d2|)
  template <typename R>
  VGM_DEV static bool intersect(const Data &o, V3<R> p, V3<R> d, bool to_in, R &distance)
  {
    const R rad2 = mag2(p);
    const R pDotV3D = p.x * d.x + p.y * d.y + p.z * d.z;
    const R fR = R(o.r);
    const R c = rad2 - fR * fR;
    const R d2 = pDotV3D * pDotV3D - c;
    if (to_in) {
      const bool cond = (d2 >= R(0)) && (pDotV3D <= R(0));
      if (cond) distance = -pDotV3D - vsqrt(vabs(d2));
      return cond;
    }
    if (d2 >= R(0)) distance = -pDotV3D + vsqrt(vabs(d2));
    return d2 >= R(0);
  }

  // DistanceToIn (:122-145)
  template <typename R>
  VGM_DEV static R distance_to_in(const Data &o, V3<R> p, V3<R> d, R /*step_max*/)
  {
    const R rad = vsqrt(mag2(p));
    if (rad < R(o.r) - R(kTolerance)) return R(-1);
    const R pDotV3D = p.x * d.x + p.y * d.y + p.z * d.z;
    const bool onSurface = (rad >= R(o.r) - R(kTolerance)) && (rad <= R(o.r) + R(kTolerance));
    if (onSurface && pDotV3D < R(0)) return R(0);
    R dist = inf_length<R>();
    R distance = inf_length<R>();
    if (intersect<R>(o, p, d, true, dist)) distance = dist;
    return distance;
  }

  // DistanceToOut (:147-171)
  template <typename R>
  VGM_DEV static R distance_to_out(const Data &o, V3<R> p, V3<R> d, R /*step_max*/)
  {
    const R rad = vsqrt(mag2(p));
    if (rad > R(o.r) + R(kTolerance)) return R(-1);
    const R pDotV3D = p.x * d.x + p.y * d.y + p.z * d.z;
    const bool onSurface = (rad >= R(o.r) - R(kTolerance)) && (rad <= R(o.r) + R(kTolerance));
    if (onSurface && pDotV3D > R(0)) return R(0);
    R dist = inf_length<R>();
    R distance = inf_length<R>();
    if (intersect<R>(o, p, d, false, dist)) distance = dist;
    return distance;
  }

  // SafetyToIn (:173-184)
  template <typename R>
  VGM_DEV static R safety_to_in(const Data &o, V3<R> p)
  {
    const R rad = vsqrt(mag2(p));
    R safety = rad - R(o.r);
    if (rad < R(o.r) - R(kTolerance)) return R(-1);
    if ((rad > R(o.r) - R(kTolerance)) && (rad < R(o.r) + R(kTolerance))) safety = R(0);
    return safety;
  }

  // SafetyToOut (:186-197)
  template <typename R>
  VGM_DEV static R safety_to_out(const Data &o, V3<R> p)
  {
    const R rad = vsqrt(mag2(p));
    R safety = R(o.r) - rad;
    if (rad > R(o.r) + R(kTolerance)) return R(-1);
    if ((rad > R(o.r) - R(kTolerance)) && (rad < R(o.r) + R(kTolerance))) safety = R(0);
    return safety;
  }

  // NormalKernel (OrbImplementation.h:227-240): radial
  VGM_DEV static bool normal(const Data &o, V3<double> p, V3<double> &n)
  {
    const double rad2 = p.x * p.x + p.y * p.y + p.z * p.z;
    const double invRadius = 1. / __builtin_sqrt(rad2);
    n = V3<double>{p.x * invRadius, p.y * invRadius, p.z * invRadius};
    const double tolRMaxO = o.r + kTolerance, tolRMaxI = o.r - kTolerance;
    return (rad2 <= tolRMaxO * tolRMaxO) && (rad2 >= tolRMaxI * tolRMaxI);
  }
};

}  // namespace vgm
