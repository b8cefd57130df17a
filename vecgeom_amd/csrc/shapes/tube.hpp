// Tube kernels, UniversalTube runtime-checked variant
// (VecGeom/volumes/kernel/TubeImplementation.h, TUBE_SAFETY_OLD; shape type
// traits VecGeom/volumes/kernel/shapetypes/TubeTypes.h).  One code path per
// thread; the rmin/phi branches are wave-uniform because the parameters are.
#pragma once
#include "../vgm_math.hpp"
#include "tube_data.h"
#include "wedge.hpp"

namespace vgm {

struct TubeShape {
  using Data = TubeData;
  static constexpr bool kStage = false;  // parameter block staged in LDS
  // waves per SIMD the transport kernel asks the compiler for (1: no request)
  static constexpr int kWavesPerEU = 1;
  // safety is the distance to the surface (1-Lipschitz): carried bounds are valid
  static constexpr bool kSafetyIsDistance = false;
  // SafetyToIn/Out never exceed DistanceToIn/Out beyond the surface tolerance:
  // the distance-first step takes the FP64 step's decisions
  static constexpr bool kSafetyBelowDistance = true;
  // mixed-precision DistanceToOut may be evaluated in float
  static constexpr bool kFloatDistanceToOut = true;
  // mixed-precision DistanceToIn may be evaluated in float
  static constexpr bool kFloatDistanceToIn = true;
  // mixed precision may evaluate the queries in float at all
  static constexpr bool kFloatQueries = true;

  // PointInCyclicalSector (TubeImplementation.h:62-101)
  template <typename R>
  VGM_DEV static bool in_sector(const Data &t, R x, R y, bool includeSurface)
  {
    const R startCheck = (-x * R(t.along1y) + y * R(t.along1x));
    const R endCheck = (-R(t.along2x) * y + R(t.along2y) * x);
    const R tol = includeSurface ? R(-kHalfTolerance) : R(kHalfTolerance);
    if (t.le_pi) return (startCheck >= tol) && (endCheck >= tol);
    return (startCheck >= tol) || (endCheck >= tol);
  }

  // CircleTrajectoryIntersection (:106-162)
  template <typename R, bool Largest, bool InsectorCheck>
  VGM_DEV static bool circle(const Data &t, R b, R c, V3<R> pos, V3<R> dir, R &dist)
  {
    R delta = b * b - c;
    bool ok = delta > R(0);
    if (Largest) ok = ok || (delta == R(0));
    if (!ok) delta = R(0);
    delta = vsqrt(delta);
    if (!Largest) delta = -delta;
    dist = -b + delta;
    ok = ok && (dist >= R(-2 * kTolerance));
    if (!ok) return false;
    if (InsectorCheck) {
      const R hitz = pos.z + dist * dir.z;
      if (!(vabs(hitz) <= R(t.z))) return false;
      if (t.phi_treat) {
        const R hitx = pos.x + dist * dir.x;
        const R hity = pos.y + dist * dir.y;
        return in_sector<R>(t, hitx, hity, true);
      }
    }
    return true;
  }

  // PhiPlaneTrajectoryIntersection (:264-308) with PositiveDirectionOfPhiVector
  template <typename R, bool InsectorCheck>
  VGM_DEV static bool phi_plane(const Data &t, double ax, double ay, double nx, double ny, V3<R> pos, V3<R> dir,
                                R &dist)
  {
    const R dirDotNorm = dir.x * R(nx) + dir.y * R(ny);
    bool ok = InsectorCheck ? (dirDotNorm > R(0)) : (dirDotNorm < R(0));
    const R dirDotXY = (dir.y * R(ax) - dir.x * R(ay));
    dist = (R(ay) * pos.x - R(ax) * pos.y) / nonzero(dirDotXY);
    ok = ok && ((dist * vabs(dirDotNorm)) > R(-kHalfTolerance));
    const R hitx = pos.x + dist * dir.x;
    const R hity = pos.y + dist * dir.y;
    if (InsectorCheck) {
      const R hitz = pos.z + dist * dir.z;
      const R r2 = hitx * hitx + hity * hity;
      ok = ok && vabs(hitz) <= R(t.tolOz) && (r2 >= R(t.tolOrmin2)) && (r2 <= R(t.tolOrmax2));
      ok = ok && (hitx * R(ax) + hity * R(ay)) > R(0);
    } else {
      ok = ok && (hitx * R(ax) + hity * R(ay)) >= R(0);
    }
    return ok;
  }

  // GenericKernelForContainsAndInside (:392-447)
  template <typename R, bool ForInside>
  VGM_DEV static void generic(const Data &t, V3<R> p, bool &cin, bool &cout)
  {
    const R absz = vabs(p.z);
    cout = absz > (ForInside ? R(t.z) + R(kHalfTolerance) : R(t.z));
    if (ForInside) cin = absz < R(t.z) - R(kHalfTolerance);
    if (cout) return;
    const R r2 = p.x * p.x + p.y * p.y;
    const R rmax = R(t.rmax);
    cout = cout || (r2 > (ForInside ? rmax * (rmax + R(2.0 * kTolerance)) : rmax * rmax));
    if (ForInside) cin = cin && (r2 < vabs(rmax * (rmax - R(2.0 * kTolerance))));
    if (cout) return;
    if (t.rmin_treat) {
      const R rmin = R(t.rmin);
      cout = cout || (r2 <= (ForInside ? vabs(rmin * (rmin - R(2.0 * kTolerance))) : rmin * rmin));
      if (ForInside) cin = cin && (r2 > rmin * (rmin + R(2.0 * kTolerance)));
      if (cout) return;
    }
    if (t.phi_treat) {
      bool ci = false, co = false;
      wedge_generic<R, ForInside>(t.along1x, t.along1y, t.along2x, t.along2y, t.normal1x, t.normal1y, t.normal2x,
                                  t.normal2y, t.lt_pi, p, ci, co);
      cout = cout || co;
      if (ForInside) cin = cin && ci;
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

  // IsMovingInsideTubeSurface (:309-343)
  template <typename R, bool Inner>
  VGM_DEV static bool moving_inside_surface(const Data &t, V3<R> p, V3<R> d)
  {
    const R rho = p.x * p.x + p.y * p.y;
    const bool on = Inner ? ((rho >= R(t.tolOrmin2)) && (rho <= R(t.tolIrmin2)))
                          : ((rho >= R(t.tolIrmax2)) && (rho <= R(t.tolOrmax2)));
    if (!(on && (vabs(p.z) < (R(t.z) + R(kTolerance))))) return false;
    const R dot = Inner ? (d.x * (-p.x) + d.y * (-p.y) + d.z * R(0)) : (d.x * p.x + d.y * p.y + d.z * R(0));
    return dot <= R(0);
  }

  template <typename R>
  VGM_DEV static R distance_to_in(const Data &t, V3<R> pt, V3<R> dir, R /*step_max*/)
  {
    // far-point pre-step (TubeImplementation.h:478-502)
    V3<R> point = pt;
    const R ptDist = vsqrt(point.x * point.x + point.y * point.y + point.z * point.z);
    R distToMove = R(0);
    if (ptDist > R(100.) * R(t.maxval)) {
      distToMove = ptDist - R(100.) * R(t.maxval);
      point.x = point.x + distToMove * dir.x;
      point.y = point.y + distToMove * dir.y;
      point.z = point.z + distToMove * dir.z;
    }
    R distance = distance_to_in_kernel<R>(t, point, dir);
    distance += distToMove;
    return distance;
  }

  // DistanceToInKernel (:507-681), scalar early returns
  template <typename R>
  VGM_DEV static R distance_to_in_kernel(const Data &t, V3<R> point, V3<R> dir)
  {
    bool done = false;
    R distance = inf_length<R>();
    R distz = vabs(point.z) - R(t.z);
    done = done || (distz > R(kHalfTolerance) && point.z * dir.z >= R(0));
    const R rsq = point.x * point.x + point.y * point.y;
    const R rdotn = point.x * dir.x + point.y * dir.y;
    done = done || (rsq > R(t.tolIrmax2) && rdotn >= R(0));
    if (done) return distance;
    distance = R(-1.0);
    bool inside = distz < R(-kHalfTolerance);
    inside = inside && (rsq < R(t.tolIrmax2));
    if (t.rmin_treat) inside = inside && (rsq > R(t.tolIrmin2));
    if (t.phi_treat && inside) inside = in_sector<R>(t, point.x, point.y, false);
    if (inside) return distance;
    distance = inf_length<R>();
    distz /= nonzero_abs(dir.z);
    const R hitx = point.x + distz * dir.x;
    const R hity = point.y + distz * dir.y;
    const R r2 = hitx * hitx + hity * hity;
    bool okz = distz > R(-kHalfTolerance) && (point.z * dir.z < R(0));
    okz = okz && (r2 <= R(t.rmax2));
    if (t.rmin_treat) okz = okz && (R(t.rmin2) <= r2);
    if (t.phi_treat && okz) okz = in_sector<R>(t, hitx, hity, true);
    if (okz) distance = distz;
    done = okz;
    bool onSurfIn = moving_inside_surface<R, false>(t, point, dir);
    if (t.rmin_treat) onSurfIn = onSurfIn || moving_inside_surface<R, true>(t, point, dir);
    if (!t.phi_treat) {
      if (!done && onSurfIn) distance = R(0);
      done = done || onSurfIn;
    } else {
      const bool insector = in_sector<R>(t, point.x, point.y, true);
      if (!done && insector && onSurfIn) distance = R(0);
      done = done || (insector && onSurfIn);
    }
    if (done) return distance;
    const R invnsq = R(1.) / nonzero(R(1.) - dir.z * dir.z);
    const R b = invnsq * rdotn;
    const R crmax = invnsq * (rsq - R(t.rmax2));
    R dist_rmax;
    bool ok_rmax = circle<R, false, true>(t, b, crmax, point, dir, dist_rmax);
    ok_rmax = ok_rmax && (dist_rmax < distance);
    if (ok_rmax) return dist_rmax;
    if (t.rmin_treat) {
      const R crmin = invnsq * (rsq - R(t.rmin2));
      R dist_rmin = -inf_length<R>();
      bool ok_rmin = circle<R, true, true>(t, b, crmin, point, dir, dist_rmin);
      if (ok_rmin && dist_rmin < distance) distance = dist_rmin;
    }
    if (t.phi_treat) {
      R dist_phi;
      bool ok_phi = phi_plane<R, true>(t, t.along1x, t.along1y, t.normal1x, t.normal1y, point, dir, dist_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
      ok_phi = phi_plane<R, true>(t, t.along2x, t.along2y, t.normal2x, t.normal2y, point, dir, dist_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
    }
    return distance;
  }

  // DistanceToOut (:686-821); UniversalTube takes the last phi branch
  template <typename R>
  VGM_DEV static R distance_to_out(const Data &t, V3<R> point, V3<R> dir, R /*step_max*/)
  {
    R distz = R(t.z) - vabs(point.z);
    if (distz < R(-kHalfTolerance)) return R(-1.);
    const R rsq = point.x * point.x + point.y * point.y;
    const R rdotn = dir.x * point.x + dir.y * point.y;
    R crmax = rsq - R(t.rmax2);
    R crmin = rsq;
    if (crmax > R(2.0 * kTolerance) * R(t.rmax)) return R(-1.);
    if (t.rmin_treat) {
      crmin -= R(t.rmin2);
      if (crmin < R(-2.0 * kTolerance) * R(t.rmin)) return R(-1.);
    }
    if (t.phi_treat) {
      bool ci = false, co = false;
      wedge_generic<R, true>(t.along1x, t.along1y, t.along2x, t.along2y, t.normal1x, t.normal1y, t.normal2x,
                             t.normal2y, t.lt_pi, point, ci, co);
      if (co) return R(-1.);
    }
    R distance = inf_length<R>();
    const R invdirz = R(1.) / nonzero(dir.z);
    distz = (R(t.z) - point.z) * invdirz;
    if (dir.z < R(0)) distz = (-R(t.z) - point.z) * invdirz;
    if (dir.z != R(0) && distz < distance) distance = distz;
    const R invnsq = R(1.) / nonzero(R(1.) - dir.z * dir.z);
    const R b = invnsq * rdotn;
    if (t.rmin_treat) {
      R dist_rmin = inf_length<R>();
      crmin *= invnsq;
      bool ok_rmin = circle<R, false, false>(t, b, crmin, point, dir, dist_rmin);
      if (ok_rmin && dist_rmin < distance) distance = dist_rmin;
    }
    R dist_rmax = inf_length<R>();
    crmax *= invnsq;
    bool ok_rmax = circle<R, true, false>(t, b, crmax, point, dir, dist_rmax);
    if (ok_rmax && dist_rmax < distance) distance = dist_rmax;
    if (t.phi_treat) {
      R dist_phi;
      bool ok_phi = phi_plane<R, false>(t, t.along1x, t.along1y, t.normal1x, t.normal1y, point, dir, dist_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
      ok_phi = phi_plane<R, false>(t, t.along2x, t.along2y, t.normal2x, t.normal2y, point, dir, dist_phi);
      if (ok_phi && dist_phi < distance) distance = dist_phi;
    }
    return distance;
  }

  // PhiPlaneSafety<inside=false> (:209-243) with kUnknownAngle
  template <typename R>
  VGM_DEV static R phi_safety(const Data &t, V3<R> pos)
  {
    R safety = t.gt_pi ? vsqrt(pos.x * pos.x + pos.y * pos.y) : inf_length<R>();
    const R phi1 = pos.x * R(t.along1y) - pos.y * R(t.along1x);
    if (phi1 > R(-kHalfTolerance) && phi1 < safety) safety = phi1;
    R phi2 = pos.x * R(t.along2y) - pos.y * R(t.along2x);
    phi2 *= R(-1);
    if (phi2 > R(-kHalfTolerance) && phi2 < safety) safety = phi2;
    return safety;
  }

  // SafetyToInOld (:912-939)
  template <typename R>
  VGM_DEV static R safety_to_in(const Data &t, V3<R> point)
  {
    R safety = vabs(point.z) - R(t.z);
    const R r = vsqrt(point.x * point.x + point.y * point.y);
    const R safermax = r - R(t.rmax);
    if (safermax > safety) safety = safermax;
    if (t.rmin_treat) {
      const R safermin = R(t.rmin) - r;
      if (safermin > safety) safety = safermin;
    }
    if (t.phi_treat) {
      if (in_sector<R>(t, point.x, point.y, false)) return safety;
      const R safephi = phi_safety<R>(t, point);
      if (safephi < inf_length<R>() && safephi > safety) safety = safephi;
    }
    return safety;
  }

  // SafetyToOutOld (:944-964)
  template <typename R>
  VGM_DEV static R safety_to_out(const Data &t, V3<R> point)
  {
    R safety = R(t.z) - vabs(point.z);
    const R r = vsqrt(point.x * point.x + point.y * point.y);
    const R safermax = R(t.rmax) - r;
    if (safermax < safety) safety = safermax;
    if (t.rmin_treat) {
      const R safermin = r - R(t.rmin);
      if (safermin < safety) safety = safermin;
    }
    if (t.phi_treat) {
      const R safephi = wedge_safety_to_out<R>(t.normal1x, t.normal1y, t.normal2x, t.normal2y, t.lt_pi, point);
      if (safephi < safety) safety = safephi;
    }
    return safety;
  }
};

}  // namespace vgm
