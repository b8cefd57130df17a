// Box kernels: VecGeom/volumes/kernel/BoxImplementation.h:76-167.
// Parameter block: half-lengths (BoxStruct::fDimensions).
#pragma once
#include "../vgm_math.hpp"
#include "box_data.h"

namespace vgm {

struct BoxShape {
  using Data = BoxData;
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

  // max(|p| - h) (BoxImplementation.h:76-90)
  template <typename R>
  VGM_DEV static R excess(const Data &b, V3<R> p)
  {
    return vmax3(vabs(p.x) - R(b.dim[0]), vabs(p.y) - R(b.dim[1]), vabs(p.z) - R(b.dim[2]));
  }

  template <typename R>
  VGM_DEV static bool contains(const Data &b, V3<R> p)
  {
    return excess(b, p) < R(0);
  }

  template <typename R>
  VGM_DEV static int inside(const Data &b, V3<R> p)
  {
    R dist = excess(b, p);
    int in = dist < R(0) ? kInsideV : kOutsideV;
    if (vabs(dist) < R(kHalfTolerance)) in = kSurfaceV;
    return in;
  }

  // Slab method (BoxImplementation.h:109-131).
  template <typename R>
  VGM_DEV static R distance_to_in(const Data &b, V3<R> p, V3<R> d, R /*step_max*/)
  {
    const R ix = R(1) / nonzero(d.x), iy = R(1) / nonzero(d.y), iz = R(1) / nonzero(d.z);
    const R sx = vsign(d.x), sy = vsign(d.y), sz = vsign(d.z);
    const R hx = R(b.dim[0]), hy = R(b.dim[1]), hz = R(b.dim[2]);
    const R inx = -sx * hx - p.x, iny = -sy * hy - p.y, inz = -sz * hz - p.z;
    const R outx = sx * hx - p.x, outy = sy * hy - p.y, outz = sz * hz - p.z;
    const R absOrthogOut = vabs(vmin3(sx * outx, sy * outy, sz * outz));
    const R distOut = vmin3(outx * ix, outy * iy, outz * iz);
    R distance = vmax3(inx * ix, iny * iy, inz * iz);
    if (distance >= distOut || distOut <= R(kHalfTolerance) || absOrthogOut <= R(kHalfTolerance))
      distance = inf_length<R>();
    return distance;
  }

  // BoxImplementation.h:136-151
  template <typename R>
  VGM_DEV static R distance_to_out(const Data &b, V3<R> p, V3<R> d, R /*step_max*/)
  {
    const R ix = R(1) / nonzero(d.x), iy = R(1) / nonzero(d.y), iz = R(1) / nonzero(d.z);
    const R sx = vsign(d.x), sy = vsign(d.y), sz = vsign(d.z);
    const R safetyIn = excess(b, p);
    const R outx = sx * R(b.dim[0]) - p.x, outy = sy * R(b.dim[1]) - p.y, outz = sz * R(b.dim[2]) - p.z;
    R distance = vmin3(outx * ix, outy * iy, outz * iz);
    if (safetyIn > R(kHalfTolerance)) distance = R(-1);
    return distance;
  }

  template <typename R>
  VGM_DEV static R safety_to_in(const Data &b, V3<R> p)
  {
    return excess(b, p);
  }

  template <typename R>
  VGM_DEV static R safety_to_out(const Data &b, V3<R> p)
  {
    return vmin3(R(b.dim[0]) - vabs(p.x), R(b.dim[1]) - vabs(p.y), R(b.dim[2]) - vabs(p.z));
  }

  // NormalKernel (BoxImplementation.h:172-189): the faces within tolerance of
  // the nearest one, averaged on edges and corners
  VGM_DEV static bool normal(const Data &b, V3<double> p, V3<double> &n)
  {
    const double sx = vabs(vabs(p.x) - b.dim[0]), sy = vabs(vabs(p.y) - b.dim[1]), sz = vabs(vabs(p.z) - b.dim[2]);
    const double safmin = vmin3(sx, sy, sz);
    const bool valid = safmin < kHalfTolerance;
    V3<double> v{0.0, 0.0, 0.0};
    if (sx - safmin < kHalfTolerance) v.x = vsign(p.x);
    if (sy - safmin < kHalfTolerance) v.y = vsign(p.y);
    if (sz - safmin < kHalfTolerance) v.z = vsign(p.z);
    if (v3_mag2(v) > 1.0) v = v3_normalized(v);
    n = v;
    return valid;
  }
};

}  // namespace vgm
