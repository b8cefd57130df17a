// Trapezoid kernels with PlaneShell<4> side planes, scalar semantics
// (VecGeom/volumes/kernel/TrapezoidImplementation.h:98-395,
//  VecGeom/base/PlaneShell.h:130-290).
#pragma once
#include "../vgm_math.hpp"
#include "trapezoid_data.h"

namespace vgm {

struct TrapezoidShape {
  using Data = TrapezoidData;
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
  VGM_DEV static R pdist(const Data &t, int i, V3<R> p)
  {
    return R(t.a[i]) * p.x + R(t.b[i]) * p.y + R(t.c[i]) * p.z + R(t.d[i]);
  }
  template <typename R>
  VGM_DEV static R proj(const Data &t, int i, V3<R> d)
  {
    return R(t.a[i]) * d.x + R(t.b[i]) * d.y + R(t.c[i]) * d.z;
  }

  template <typename R, bool ForInside>
  VGM_DEV static void generic(const Data &t, V3<R> p, bool &cin, bool &cout)
  {
    cout = vabs(p.z) > (ForInside ? R(t.dz) + R(kHalfTolerance) : R(t.dz));
    if (ForInside) cin = vabs(p.z) < R(t.dz) - R(kHalfTolerance);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const R dist = pdist<R>(t, i, p);
      cout = cout || (dist > (ForInside ? R(kHalfTolerance) : R(0)));
      if (ForInside) cin = cin && (dist < R(-kHalfTolerance));
    }
  }

  template <typename R>
  VGM_DEV static bool contains(const Data &t, V3<R> p)
  {
    bool ci = false, co = false;
    generic<R, false>(t, p, ci, co);
    return !co;
  }

  template <typename R>
  VGM_DEV static int inside(const Data &t, V3<R> p)
  {
    bool ci = false, co = false;
    generic<R, true>(t, p, ci, co);
    return ci ? kInsideV : (co ? kOutsideV : kSurfaceV);
  }

  template <typename R>
  VGM_DEV static R distance_to_in(const Data &t, V3<R> point, V3<R> dir, R /*step_max*/)
  {
    const R signZdir = vsign(dir.z);
    const R max = signZdir * R(t.dz) - point.z;
    if (signZdir * max < R(kHalfTolerance)) return inf_length<R>();
    const R invdir = R(1.0) / nonzero(dir.z);
    R smax = max * invdir;
    R smin = -(signZdir * R(t.dz) + point.z) * invdir;
    R pd[4], pj[4], vd[4];
    bool done = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      pd[i] = pdist<R>(t, i, point);
      pj[i] = proj<R>(t, i, dir);
      vd[i] = -pd[i] / nonzero(pj[i]);
      done = done || (pd[i] > R(kHalfTolerance) && pj[i] >= R(0));
      done = done || (pd[i] > R(-kHalfTolerance) && pj[i] > R(0));
    }
    if (done) return inf_length<R>();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool posPoint = pd[i] > R(-kHalfTolerance);
      const bool posDir = pj[i] > R(0);
      const bool fromIn = !posPoint && posDir;
      const bool fromOut = posPoint && !posDir;
      if ((fromIn && vd[i] < smin) || (fromOut && vd[i] > smax)) return inf_length<R>();
      if (fromOut && vd[i] > smin) smin = vd[i];
      if (fromIn && vd[i] < smax) smax = vd[i];
    }
    return (smin <= smax) ? smin : inf_length<R>();
  }

  template <typename R>
  VGM_DEV static R distance_to_out(const Data &t, V3<R> point, V3<R> dir, R /*step_max*/)
  {
    if (vabs(point.z) > R(t.dz) + R(kHalfTolerance)) return R(-1.0);
    R distance = inf_length<R>();
    const R distz = (vsign(dir.z) * R(t.dz) - point.z) / nonzero(dir.z);
    if (dir.z != R(0)) distance = distz;
    R distOut = inf_length<R>();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const R pd = pdist<R>(t, i, point);
      const R pj = proj<R>(t, i, dir);
      const R vd = -pd / nonzero(pj);
      if (pd > R(kHalfTolerance)) distOut = R(-1.);
      if (pj > R(0) && vd < distOut) distOut = vd;
    }
    if (distOut < distance) distance = distOut;
    return distance;
  }

  template <typename R>
  VGM_DEV static R safety_to_in(const Data &t, V3<R> point)
  {
    R safety = vabs(point.z) - R(t.dz);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const R d = pdist<R>(t, i, point);
      if (d > safety) safety = d;
    }
    return safety;
  }

  template <typename R>
  VGM_DEV static R safety_to_out(const Data &t, V3<R> point)
  {
    R safety = R(t.dz) - vabs(point.z);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const R d = -(pdist<R>(t, i, point));
      if (d < safety) safety = d;
    }
    return safety;
  }

  // NormalKernel (TrapezoidImplementation.h:406-462) over
  // PlaneShell::NormalKernel (VecGeom/base/PlaneShell.h:342-370)
  VGM_DEV static bool normal(const Data &t, V3<double> p, V3<double> &n)
  {
    const double delta = 1000. * kTolerance;
    V3<double> normal{0., 0., 0.};
    double safety = -kInfLength;
    double dist[4];
    for (int i = 0; i < 4; ++i) dist[i] = pdist<double>(t, i, p);
    for (int i = 0; i < 4; ++i) {
      const double saf_i = dist[i] - safety;
      if (vabs(saf_i) < kHalfTolerance && dist[i] >= -kHalfTolerance)
        normal = V3<double>{normal.x + t.a[i], normal.y + t.b[i], normal.z + t.c[i]};
      if (saf_i > 0.) normal = V3<double>{t.a[i], t.b[i], t.c[i]};
      if (saf_i > 0.) safety = dist[i];
    }
    const double sz = vsign(p.z);
    const double safz = sz * p.z - t.dz;
    if (vabs(safz - safety) < kHalfTolerance && safz >= -kHalfTolerance)
      normal = V3<double>{normal.x + 0, normal.y + 0, normal.z + sz};
    if (safz > safety && safz >= -kHalfTolerance) normal = V3<double>{0, 0, sz};
    if (safz > safety) safety = safz;
    const bool valid = vabs(safety) <= delta;
    if (v3_mag2(normal) > 1.0) normal = v3_normalized(normal);
    n = normal;
    return valid;
  }
};

}  // namespace vgm
