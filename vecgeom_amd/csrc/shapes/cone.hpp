// Cone kernels, UniversalCone runtime-checked variant, scalar semantics
// (VecGeom/volumes/kernel/ConeImplementation.h:99-517 and the scalar
// ConeHelpers of VecGeom/volumes/ConeUtilities.h).
#pragma once
#include "../vgm_math.hpp"
#include "cone_data.h"
#include "wedge.hpp"
#include "mixed.hpp"

namespace vgm {

struct ConeShape {
  using Data = ConeData;
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
  static constexpr bool kMixedDistance = true;

  template <typename R>
  VGM_DEV static R plus_tol_sq(R x, double tol) { return x * (x + R(2.0 * tol)); }
  template <typename R>
  VGM_DEV static R minus_tol_sq(R x, double tol) { return vabs(x * (x - R(2.0 * tol))); }

  // ConeUtilities::PointInCyclicalSector with kConeTolerance bands
  template <typename R>
  VGM_DEV static bool in_sector(const Data &c, R x, R y, bool includeSurface)
  {
    const R startCheck = (-x * R(c.along1y)) + (y * R(c.along1x));
    const R endCheck = (-R(c.along2x) * y) + (R(c.along2y) * x);
    const R tol = includeSurface ? R(-kConeTolerance) : R(kConeTolerance);
    if (c.le_pi) return (startCheck >= tol) && (endCheck >= tol);
    return (startCheck >= tol) || (endCheck >= tol);
  }

  template <typename R>
  VGM_DEV static R radius_at(const Data &c, R z, bool inner)
  {
    if (inner) return (c.rmin1 == c.rmin2) ? R(c.rmin1) : R(c.inner_slope) * z + R(c.inner_offset);
    return (c.orig_rmax1 == c.orig_rmax2) ? R(c.orig_rmax1) : R(c.outer_slope) * z + R(c.outer_offset);
  }

  // n . d with the (unnormalised) surface normal of ConeUtilities::GetNormal
  template <typename R>
  VGM_DEV static R normal_dot(const Data &c, V3<R> p, V3<R> d, bool inner)
  {
    const R rho = vsqrt(p.x * p.x + p.y * p.y);
    if (inner) {
      const R nz = ((c.rmin1 == c.rmin2) && (c.rmin1 != 0.)) ? R(0) : R(c.znorm_inner) * (rho * R(c.sec_rmin));
      return d.x * (-p.x) + d.y * (-p.y) + d.z * nz;
    }
    const R nz = ((c.rmax1 == c.rmax2) && (c.rmax1 != 0.)) ? R(0) : R(c.znorm_outer) * (rho * R(c.sec_rmax));
    return d.x * p.x + d.y * p.y + d.z * nz;
  }

  template <typename R>
  VGM_DEV static bool on_conical_surface(const Data &c, V3<R> p, bool inner)
  {
    const R rho = p.x * p.x + p.y * p.y;
    const R coneRad = radius_at<R>(c, p.z, inner);
    const R coneRad2 = coneRad * coneRad;
    const R tol = inner ? R(c.inner_tol) : R(c.outer_tol);
    return (rho >= (coneRad2 - tol * coneRad)) && (rho <= (coneRad2 + tol * coneRad)) &&
           (vabs(p.z) < (R(c.dz) + R(kConeTolerance)));
  }

  // ConeUtilities::PhiPlaneTrajectoryIntersection, PositiveDirectionOfPhiVector
  template <typename R, bool InsectorCheck>
  VGM_DEV static bool phi_plane(const Data &c, double ax, double ay, double nx, double ny, V3<R> pos, V3<R> dir,
                                R &dist)
  {
    dist = inf_length<R>();
    bool ok = InsectorCheck ? ((dir.x * R(nx)) + (dir.y * R(ny)) > R(0)) : ((dir.x * R(nx)) + (dir.y * R(ny)) < R(0));
    const R dirDotXY = (dir.y * R(ax)) - (dir.x * R(ay));
    if (dirDotXY != R(0)) dist = ((R(ay) * pos.x) - (R(ax) * pos.y)) / nonzero(dirDotXY);
    ok = ok && (dist > R(-kConeTolerance));
    const R hitx = pos.x + dist * dir.x;
    const R hity = pos.y + dist * dir.y;
    if (InsectorCheck) {
      const R hitz = pos.z + dist * dir.z;
      const R r2 = (hitx * hitx) + (hity * hity);
      const R innerRadIrTol = radius_at<R>(c, hitz, true) + R(kTolerance);
      const R outerRadIrTol = radius_at<R>(c, hitz, false) - R(kTolerance);
      ok = ok && vabs(hitz) <= R(c.tol_iz) && (r2 >= innerRadIrTol * innerRadIrTol) &&
           (r2 <= outerRadIrTol * outerRadIrTol);
      ok = ok && ((hitx * R(ax)) + (hity * R(ay))) > R(0);
    } else {
      ok = ok && ((hitx * R(ax)) + (hity * R(ay))) >= R(0);
    }
    return ok;
  }

  // Scalar DetectIntersectionAndCalculateDistanceToConicalSurface
  template <typename R, bool ForIn, bool Inner>
  VGM_DEV static bool conical(const Data &c, V3<R> point, V3<R> direction, R &distance)
  {
    distance = inf_length<R>();
    const bool onSurf = on_conical_surface<R>(c, point, Inner);
    const R ndot = normal_dot<R>(c, point, direction, Inner);
    if (onSurf && vabs(ndot) == R(0)) return false;
    if (onSurf) {
      const bool moving = ForIn ? (ndot <= R(0)) : (ndot >= R(0));
      const bool insector = c.phi_treat ? in_sector<R>(c, point.x, point.y, true) : true;
      if (insector && moving) {
        distance = R(0);
        return true;
      }
    }
    bool ok = false;
    const R pDotV2D = point.x * direction.x + point.y * direction.y;
    R a, b, cc;
    if (Inner) {
      if (c.rmin1 == c.rmin2) {
        b = pDotV2D;
        a = direction.x * direction.x + direction.y * direction.y;
        cc = (point.x * point.x + point.y * point.y) - R(c.rmin2) * R(c.rmin2);
      } else {
        const R t = R(c.tan_inner_apex);
        const R newPz = (c.rmin2 > c.rmin1) ? (point.z + R(c.dz) + R(c.inner_apex)) * t
                                            : (point.z - R(c.dz) - R(c.inner_apex)) * t;
        const R dirT = direction.z * t;
        a = (direction.x * direction.x) + (direction.y * direction.y) - dirT * dirT;
        b = pDotV2D - (newPz * dirT);
        cc = (point.x * point.x + point.y * point.y) - (newPz * newPz);
      }
      const R b2 = b * b;
      const R ac = a * cc;
      if (b2 < ac) return false;
      const R d2 = b2 - ac;
      const R delta = vsqrt(d2);
      if (ForIn) {
        distance = (b >= R(0)) ? (cc / nonzero(-b - delta)) : (-b + delta) / nonzero(a);
      } else {
        if (b == R(0) && delta == R(0)) return false;
        distance = (b >= R(0)) ? (-b - delta) / nonzero(a) : (cc / nonzero(-b + delta));
      }
      if (distance < R(0)) return false;
      const R newZ = point.z + (direction.z * distance);
      ok = (vabs(newZ) < R(c.dz));
    } else {
      if (c.orig_rmax1 == c.orig_rmax2) {
        a = direction.x * direction.x + direction.y * direction.y;
        b = pDotV2D;
        cc = ((point.x * point.x + point.y * point.y) - R(c.orig_rmax2) * R(c.orig_rmax2));
      } else {
        const R t = R(c.tan_outer_apex);
        const R newPz = (c.orig_rmax2 > c.orig_rmax1) ? (point.z + R(c.dz) + R(c.outer_apex)) * t
                                                      : (point.z - R(c.dz) - R(c.outer_apex)) * t;
        const R dirT = direction.z * t;
        a = direction.x * direction.x + direction.y * direction.y - dirT * dirT;
        b = (pDotV2D - (newPz * dirT));
        cc = ((point.x * point.x + point.y * point.y) - (newPz * newPz));
      }
      const R b2 = b * b;
      const R ac = a * cc;
      const R d2 = b2 - ac;
      if (d2 < R(0)) return false;
      const R delta = vsqrt(d2);
      if (ForIn) {
        if (b == R(0) && delta == R(0)) return false;
        distance = (b > R(0)) ? (-b - delta) / nonzero(a) : (cc / nonzero(-b + delta));
        const R newZ = point.z + (direction.z * distance);
        ok = (vabs(newZ) < R(c.dz) + R(kHalfTolerance));
      } else {
        distance = (b < R(0)) ? (-b + delta) / nonzero(a) : (cc / nonzero(-b - delta));
        ok = distance > R(0);
      }
      if (distance < R(0)) return false;
    }
    if (c.phi_treat) {
      R hitx = R(0), hity = R(0);
      if (distance < inf_length<R>()) {
        hitx = point.x + distance * direction.x;
        hity = point.y + distance * direction.y;
      }
      ok = ok && in_sector<R>(c, hitx, hity, true) && (distance < inf_length<R>());
    }
    return ok;
  }

  template <typename R, bool ForInside>
  VGM_DEV static void generic(const Data &c, V3<R> p, bool &cin, bool &cout)
  {
    const R absz = vabs(p.z);
    cout = absz > (ForInside ? R(c.dz) + R(kConeTolerance) : R(c.dz));
    if (ForInside) cin = absz < R(c.dz) - R(kConeTolerance);
    if (cout) return;
    const R r2 = p.x * p.x + p.y * p.y;
    const R rmax = (c.orig_rmax1 == c.orig_rmax2) ? R(c.orig_rmax1) : R(c.outer_slope) * p.z + R(c.outer_offset);
    cout = cout || (r2 > (ForInside ? plus_tol_sq<R>(rmax, c.outer_tol) : rmax * rmax));
    if (ForInside) cin = cin && (r2 < minus_tol_sq<R>(rmax, c.outer_tol));
    if (cout) return;
    if (c.rmin_treat) {
      const R rmin = R(c.inner_slope) * p.z + R(c.inner_offset);
      cout = cout || (r2 <= (ForInside ? minus_tol_sq<R>(rmin, c.inner_tol) : rmin * rmin));
      if (ForInside) cin = cin && (r2 > plus_tol_sq<R>(rmin, c.inner_tol));
      if (cout) return;
    }
    if (c.phi_treat) {
      bool ci = false, co = false;
      wedge_generic<R, ForInside>(c.w_along1x, c.w_along1y, c.w_along2x, c.w_along2y, c.w_normal1x, c.w_normal1y,
                                  c.w_normal2x, c.w_normal2y, c.lt_pi, p, ci, co);
      cout = cout || co;
      if (ForInside) cin = cin && ci;
    }
  }

  template <typename R>
  VGM_DEV static bool contains(const Data &c, V3<R> p)
  {
    bool ci = false, co = false;
    generic<R, false>(c, p, ci, co);
    return !co;
  }

  template <typename R>
  VGM_DEV static int inside(const Data &c, V3<R> p)
  {
    bool ci = false, co = false;
    generic<R, true>(c, p, ci, co);
    return ci ? kInsideV : (co ? kOutsideV : kSurfaceV);
  }

  // DistanceToIn (ConeImplementation.h:120-255): the tolerance decisions up
  // to and including the z-plane entry (true: `distance` is the result)...
  template <typename R>
  VGM_DEV static bool din_prologue(const Data &c, V3<R> point, V3<R> dir, R &distance)
  {
    distance = inf_length<R>();
    R distz = vabs(point.z) - R(c.dz);
    const bool outZ = (distz > R(kConeTolerance) && (point.z * dir.z) >= R(0)) ||
                      (vabs(distz) < R(kConeTolerance) && (point.z * dir.z) > R(0));
    if (outZ) return true;
    const R outerRad = radius_at<R>(c, point.z, false);
    const R rsq = point.x * point.x + point.y * point.y;
    if ((rsq > minus_tol_sq<R>(outerRad, c.outer_tol)) && (normal_dot<R>(c, point, dir, false) >= R(0)))
      return true;
    bool inside = distz < R(-kConeTolerance);
    inside = inside && (rsq < minus_tol_sq<R>(outerRad, c.outer_tol));
    if (c.rmin_treat) inside = inside && (rsq > plus_tol_sq<R>(radius_at<R>(c, point.z, true), c.inner_tol));
    if (c.phi_treat) inside = inside && in_sector<R>(c, point.x, point.y, false);
    if (inside) {
      distance = R(-1.0);
      return true;
    }
    distance = inf_length<R>();
    distz /= nonzero(vabs(dir.z));
    const R hitx = point.x + distz * dir.x;
    const R hity = point.y + distz * dir.y;
    const R r2 = (hitx * hitx) + (hity * hity);
    bool top = (point.z >= R(c.tol_iz)) && (r2 <= R(c.sq_rmax2) + R(kTolerance));
    bool bottom = (point.z <= -R(c.tol_iz)) && (r2 <= R(c.sq_rmax1) + R(kTolerance));
    bool okz = (top || bottom);
    if (c.rmin_treat) {
      top = top && (r2 >= R(c.sq_rmin2) - R(kTolerance));
      bottom = bottom && (r2 >= R(c.sq_rmin1) - R(kTolerance));
      okz = okz && (top || bottom);
    }
    if (c.phi_treat) okz = okz && in_sector<R>(c, hitx, hity, true);
    if (okz) {
      distance = distz;
      return true;
    }
    return false;
  }

  // ...then the outer cone (returned at once when hit), the inner cone and
  // the phi planes; `which` >= 0 evaluates that candidate alone (0 outer,
  // 1 inner, 2/3 phi) and returns it, kInfLength when it is rejected
  template <typename R>
  VGM_DEV static R din_surfaces(const Data &c, V3<R> point, V3<R> dir, int which = -1)
  {
    R distance = inf_length<R>();
    if (which < 0 || which == 0) {
      R dist_rOuter;
      if (conical<R, true, false>(c, point, dir, dist_rOuter) && dist_rOuter < distance) return dist_rOuter;
      if (which == 0) return distance;
    }
    if (c.rmin_treat && (which < 0 || which == 1)) {
      R dist_rInner;
      if (conical<R, true, true>(c, point, dir, dist_rInner) && dist_rInner < distance) distance = dist_rInner;
    }
    if (c.phi_treat) {
      R dist_phi;
      if (which < 0 || which == 2) {
        bool ok_phi = phi_plane<R, true>(c, c.along1x, c.along1y, c.w_normal1x, c.w_normal1y, point, dir, dist_phi);
        if (ok_phi && dist_phi < distance) distance = dist_phi;
      }
      if (which < 0 || which == 3) {
        bool ok_phi = phi_plane<R, true>(c, c.along2x, c.along2y, c.w_normal2x, c.w_normal2y, point, dir, dist_phi);
        if (ok_phi && dist_phi < distance) distance = dist_phi;
      }
    }
    return distance;
  }

  template <typename R>
  VGM_DEV static R distance_to_in(const Data &c, V3<R> point, V3<R> dir, R /*step_max*/)
  {
    R distance;
    if (din_prologue<R>(c, point, dir, distance)) return distance;
    return din_surfaces<R>(c, point, dir);
  }

  // DistanceToOut (ConeImplementation.h:260-366): the wrong-side and
  // z-surface returns (true: `distance` is the result)...
  template <typename R>
  VGM_DEV static bool dout_prologue(const Data &c, V3<R> point, V3<R> direction, R &distance)
  {
    const R distz = vabs(point.z) - R(c.dz);
    const R rsq = point.x * point.x + point.y * point.y;
    distance = R(-1.0);
    if (distz > R(kConeTolerance)) return true;
    if (rsq > plus_tol_sq<R>(radius_at<R>(c, point.z, false), c.outer_tol)) return true;
    if (c.rmin_treat && rsq < minus_tol_sq<R>(radius_at<R>(c, point.z, true), c.inner_tol)) return true;
    const bool up = direction.z > R(0);
    const bool down = direction.z < R(0);
    if ((up && point.z > R(0) && vabs(distz) < R(kConeTolerance)) ||
        (down && point.z < R(0) && vabs(distz) < R(kConeTolerance))) {
      distance = distz;
      return true;
    }
    return false;
  }

  // a point on a phi plane moving out leaves at once (distance 0, whatever
  // the other surfaces give)
  template <typename R>
  VGM_DEV static bool dout_phi_exit(const Data &c, V3<R> point, V3<R> direction)
  {
    const bool onStart = wedge_on_surface<R>(c.w_along1x, c.w_along1y, c.w_normal1x, c.w_normal1y, point);
    const bool onEnd = wedge_on_surface<R>(c.w_along2x, c.w_along2y, c.w_normal2x, c.w_normal2y, point);
    const R dn1 = direction.x * R(-c.w_normal1x) + direction.y * R(-c.w_normal1y) + direction.z * R(-0.0);
    const R dn2 = direction.x * R(-c.w_normal2x) + direction.y * R(-c.w_normal2y) + direction.z * R(-0.0);
    return (onStart && dn1 > R(0)) || (onEnd && dn2 > R(0));
  }

  // ...then the smallest of the z plane (0), outer cone (1), inner cone (2)
  // and phi planes (3, 4); `which` >= 0 evaluates that candidate alone
  template <typename R>
  VGM_DEV static R dout_surfaces(const Data &c, V3<R> point, V3<R> direction, int which = -1)
  {
    const bool up = direction.z > R(0);
    const bool down = direction.z < R(0);
    R distance = inf_length<R>();
    if (which < 0 || which == 0) {
      const R dirZInv = R(1.) / nonzero(direction.z);
      if (up) distance = (R(c.dz) - point.z) * dirZInv;
      if (down) distance = (-R(c.dz) - point.z) * dirZInv;
    }
    if (which < 0 || which == 1) {
      R dist_rOuter;
      if (conical<R, false, false>(c, point, direction, dist_rOuter) && dist_rOuter < distance)
        distance = dist_rOuter;
    }
    if (c.rmin_treat && (which < 0 || which == 2)) {
      R dist_rInner;
      if (conical<R, false, true>(c, point, direction, dist_rInner) && dist_rInner < distance)
        distance = dist_rInner;
    }
    if (c.phi_treat) {
      if (which < 0 && dout_phi_exit<R>(c, point, direction)) return R(0);
      R dist_phi;
      if (which < 0 || which == 3) {
        bool ok_phi =
            phi_plane<R, false>(c, c.along1x, c.along1y, c.w_normal1x, c.w_normal1y, point, direction, dist_phi);
        if (ok_phi && dist_phi < distance) distance = dist_phi;
      }
      if (which < 0 || which == 4) {
        bool ok_phi =
            phi_plane<R, false>(c, c.along2x, c.along2y, c.w_normal2x, c.w_normal2y, point, direction, dist_phi);
        if (ok_phi && dist_phi < distance) distance = dist_phi;
      }
    }
    if (distance < R(0) && vabs(distance) < R(kTolerance)) distance = R(0);
    return distance;
  }

  template <typename R>
  VGM_DEV static R distance_to_out(const Data &c, V3<R> point, V3<R> direction, R /*step_max*/)
  {
    R distance;
    if (dout_prologue<R>(c, point, direction, distance)) return distance;
    return dout_surfaces<R>(c, point, direction);
  }

  // ---- mixed precision: float selects the surface, FP64 evaluates it -------
  // PointInCyclicalSector (kConeTolerance bands) at a point known to within e
  VGM_MIX static int in_sector_cert(const Data &c, float x, float y, float e, bool includeSurface)
  {
    const float startCheck = (-x * float(c.along1y)) + (y * float(c.along1x));
    const float endCheck = (-float(c.along2x) * y) + (float(c.along2y) * x);
    const float tol = includeSurface ? float(-kConeTolerance) : float(kConeTolerance);
    const float m = 2.0f * e + kMixEps * (vabs(x) + vabs(y));
    const int a = cert_ge(startCheck, tol, m), b = cert_ge(endCheck, tol, m);
    return c.le_pi ? cert_and(a, b) : cert_or(a, b);
  }

  // radius of the inner or outer surface at z (radius_at) in float, and a
  // bound on its error for z known to within e
  VGM_MIX static float radius_at_f(const Data &c, float z, bool inner, float e, float &err)
  {
    const bool flat = inner ? (c.rmin1 == c.rmin2) : (c.orig_rmax1 == c.orig_rmax2);
    if (flat) {
      const float r = inner ? float(c.rmin1) : float(c.orig_rmax1);
      err = kMixEps * r;
      return r;
    }
    const float k = inner ? float(c.inner_slope) : float(c.outer_slope);
    const float o = inner ? float(c.inner_offset) : float(c.outer_offset);
    err = vabs(k) * e + kMixEps * (vabs(k * z) + vabs(o));
    return k * z + o;
  }

  // DetectIntersectionAndCalculateDistanceToConicalSurface in float, with the
  // tests certified.  The quadratic is a s^2 + 2 b s + cc = 0; a root moves by
  // at most (|da| s^2 + 2 |db| s + |dcc|) / (2 sqrt(b^2 - a cc)) under
  // perturbations da, db, dcc, which bounds its float error.
  template <bool ForIn, bool Inner>
  VGM_MIX static FCand conical_f(const Data &c, V3<double> pd, V3<float> point, V3<float> direction, float m)
  {
    // a point on the conical surface takes the FP64 kernel's tolerance branch
    if (on_conical_surface<double>(c, pd, Inner)) return cand_unsure();
    const float pxy2 = point.x * point.x + point.y * point.y;
    const float dxy2 = direction.x * direction.x + direction.y * direction.y;
    const float pDotV2D = point.x * direction.x + point.y * direction.y;
    const float pdv = vabs(point.x * direction.x) + vabs(point.y * direction.y);
    float a, b, cc, A, B, C;
    const bool cyl = Inner ? (c.rmin1 == c.rmin2) : (c.orig_rmax1 == c.orig_rmax2);
    if (cyl) {
      const float R = Inner ? float(c.rmin2) : float(c.orig_rmax2);
      a = dxy2, b = pDotV2D, cc = pxy2 - R * R;
      A = dxy2, B = pdv, C = pxy2 + R * R;
    } else {
      const float t = Inner ? float(c.tan_inner_apex) : float(c.tan_outer_apex);
      const float apex = Inner ? float(c.inner_apex) : float(c.outer_apex);
      const bool grows = Inner ? (c.rmin2 > c.rmin1) : (c.orig_rmax2 > c.orig_rmax1);
      const float dz = float(c.dz);
      const float newPz = grows ? (point.z + dz + apex) * t : (point.z - dz - apex) * t;
      const float ez = (vabs(point.z) + dz + vabs(apex)) * vabs(t);  // rounding scale of newPz
      const float dirT = direction.z * t;
      a = dxy2 - dirT * dirT;
      b = pDotV2D - newPz * dirT;
      cc = pxy2 - newPz * newPz;
      A = dxy2 + dirT * dirT;
      B = pdv + vabs(newPz * dirT) + ez * vabs(dirT);
      C = pxy2 + newPz * newPz + 2.0f * vabs(newPz) * ez;
    }
    const float d2 = b * b - a * cc;
    const float ed = kMixEps * (2.0f * vabs(b) * B + vabs(a) * C + vabs(cc) * A + b * b + vabs(a * cc));
    if (d2 < -ed) return cand_no();  // the FP64 discriminant test fails
    if (d2 <= ed) return cand_unsure();
    const float delta = mix_sqrt(d2);
    float dist;
    if (Inner) {
      if (ForIn) dist = (b >= 0.0f) ? mix_div(cc, -b - delta) : mix_div(-b + delta, a);
      else dist = (b >= 0.0f) ? mix_div(-b - delta, a) : mix_div(cc, -b + delta);
    } else {
      if (ForIn) dist = (b > 0.0f) ? mix_div(-b - delta, a) : mix_div(cc, -b + delta);
      else dist = (b < 0.0f) ? mix_div(-b + delta, a) : mix_div(cc, -b - delta);
    }
    if (!(vabs(dist) < 1.0e30f)) return cand_unsure();
    const float err = kMixEps * (mix_div(A * dist * dist + 2.0f * B * vabs(dist) + C, 2.0f * delta) + vabs(dist)) + m;
    int acc;
    const float newZ = point.z + direction.z * dist;
    if (Inner) {
      acc = cert_and(cert_ge(dist, 0.0f, err), cert_le(vabs(newZ), float(c.dz), err));
    } else if (ForIn) {
      acc = cert_and(cert_le(vabs(newZ), float(c.dz + kHalfTolerance), err), cert_ge(dist, 0.0f, err));
    } else {
      acc = cert_ge(dist, 0.0f, err);
    }
    if (c.phi_treat && acc != kCertNo)
      acc = cert_and(acc, in_sector_cert(c, point.x + dist * direction.x, point.y + dist * direction.y, err, true));
    return FCand{acc, dist, err};
  }

  // ConeUtilities::PhiPlaneTrajectoryIntersection in float, certified
  template <bool InsectorCheck>
  VGM_MIX static FCand phi_plane_f(const Data &c, double ax_, double ay_, double nx_, double ny_, V3<float> pos,
                                   V3<float> dir, float m)
  {
    const float ax = float(ax_), ay = float(ay_), nx = float(nx_), ny = float(ny_);
    const float dn = (dir.x * nx) + (dir.y * ny);
    int acc = InsectorCheck ? cert_ge(dn, 0.0f, kMixEps) : cert_le(dn, 0.0f, kMixEps);
    if (acc == kCertNo) return cand_no();
    const float dirDotXY = (dir.y * ax) - (dir.x * ay);
    if (vabs(dirDotXY) < 1.0e-3f) return cand_unsure();
    const float rxy = __builtin_amdgcn_rcpf(dirDotXY);
    const float dist = ((ay * pos.x) - (ax * pos.y)) * rxy;
    const float err = (kMixEps * (vabs(pos.x) + vabs(pos.y)) + m) * vabs(rxy) + kMixEps * vabs(dist);
    acc = cert_and(acc, cert_ge(dist, float(-kConeTolerance), err));
    if (acc == kCertNo) return cand_no();
    const float hitx = pos.x + dist * dir.x;
    const float hity = pos.y + dist * dir.y;
    const float eh = err + kMixEps * vabs(dist);
    if (InsectorCheck) {
      const float hitz = pos.z + dist * dir.z;
      const float r2 = (hitx * hitx) + (hity * hity);
      const float r = mix_sqrt(r2);
      float ei, eo;
      const float ri = radius_at_f(c, hitz, true, eh, ei) + float(kTolerance);
      const float ro = radius_at_f(c, hitz, false, eh, eo) - float(kTolerance);
      const float er_i = 2.0f * (r + eh) * eh + 2.0f * vabs(ri) * ei + ei * ei + kMixEps * (r2 + ri * ri);
      const float er_o = 2.0f * (r + eh) * eh + 2.0f * vabs(ro) * eo + eo * eo + kMixEps * (r2 + ro * ro);
      acc = cert_and(acc, cert_le(vabs(hitz), float(c.tol_iz), eh));
      acc = cert_and(acc, cert_ge(r2, ri * ri, er_i));
      acc = cert_and(acc, cert_le(r2, ro * ro, er_o));
      acc = cert_and(acc, cert_ge((hitx * ax) + (hity * ay), 0.0f, 2.0f * eh));
    } else {
      acc = cert_and(acc, cert_ge((hitx * ax) + (hity * ay), 0.0f, 2.0f * eh));
    }
    return FCand{acc, dist, err};
  }

  // DistanceToIn, mixed: the FP64 kernel's tolerance decisions (outside in z,
  // moving away, inside, z-plane entry) in FP64, the conical and phi surfaces
  // selected in float, the selected one evaluated by the FP64 kernel's
  // expressions.  Equal to distance_to_in<double> bit for bit.
  VGM_MIX static double distance_to_in_mixed(const Data &c, V3<double> pt, V3<double> dir, float scale)
  {
    double distance;
    if (din_prologue<double>(c, pt, dir, distance)) return distance;
    const V3<float> pf{float(pt.x), float(pt.y), float(pt.z)};
    const V3<float> df{float(dir.x), float(dir.y), float(dir.z)};
    const float m = kMixLen * (vabs(pf.x) + vabs(pf.y) + vabs(pf.z) + scale);
    int which;
    const FCand co = conical_f<true, false>(c, pt, pf, df, m);
    if (co.acc == kCertUnsure) return din_surfaces<double>(c, pt, dir);
    if (co.acc == kCertYes) {
      which = 0;  // an outer hit is returned before the other surfaces
    } else {
      FSelect sel;
      if (c.rmin_treat) sel.add(1, conical_f<true, true>(c, pt, pf, df, m));
      if (c.phi_treat) {
        sel.add(2, phi_plane_f<true>(c, c.along1x, c.along1y, c.w_normal1x, c.w_normal1y, pf, df, m));
        sel.add(3, phi_plane_f<true>(c, c.along2x, c.along2y, c.w_normal2x, c.w_normal2y, pf, df, m));
      }
      if (sel.unsure) return din_surfaces<double>(c, pt, dir);
      if (sel.best < 0) return kInfLength;  // certified miss
      which = sel.best;
    }
    const double r = din_surfaces<double>(c, pt, dir, which);
    if (!(r < kInfLength)) return din_surfaces<double>(c, pt, dir);
    return r;
  }

  // DistanceToOut, mixed: wrong side, z-surface and phi-surface exits in
  // FP64, the smallest surface selected in float and evaluated by the FP64
  // kernel's expressions.
  VGM_MIX static double distance_to_out_mixed(const Data &c, V3<double> pt, V3<double> dir, float scale)
  {
    double distance;
    if (dout_prologue<double>(c, pt, dir, distance)) return distance;
    if (c.phi_treat && dout_phi_exit<double>(c, pt, dir)) return 0.0;
    const V3<float> pf{float(pt.x), float(pt.y), float(pt.z)};
    const V3<float> df{float(dir.x), float(dir.y), float(dir.z)};
    const float m = kMixLen * (vabs(pf.x) + vabs(pf.y) + vabs(pf.z) + scale);
    FSelect sel;
    if (dir.z != 0.0) {  // the z plane is taken exactly when dir.z != 0
      const float zt = dir.z > 0.0 ? float(c.dz) : -float(c.dz);
      sel.add(0, z_plane_cand(zt, pf.z, df.z, m));
    }
    sel.add(1, conical_f<false, false>(c, pt, pf, df, m));
    if (c.rmin_treat) sel.add(2, conical_f<false, true>(c, pt, pf, df, m));
    if (c.phi_treat) {
      sel.add(3, phi_plane_f<false>(c, c.along1x, c.along1y, c.w_normal1x, c.w_normal1y, pf, df, m));
      sel.add(4, phi_plane_f<false>(c, c.along2x, c.along2y, c.w_normal2x, c.w_normal2y, pf, df, m));
    }
    if (sel.unsure || sel.best < 0) return dout_surfaces<double>(c, pt, dir);
    const double r = dout_surfaces<double>(c, pt, dir, sel.best);
    if (!(r < kInfLength)) return dout_surfaces<double>(c, pt, dir);
    return r;
  }

  template <typename R>
  VGM_DEV static R safe_conical(const Data &c, V3<R> p, bool inner)
  {
    const R rho = vsqrt(p.x * p.x + p.y * p.y);
    if (inner) {
      const R pRMin = R(c.tan_rmin) * p.z + (R(c.rmin1) + R(c.rmin2)) * R(0.5);
      return (rho - pRMin) * R(c.inv_sec_rmin);
    }
    const R pRMax = R(c.tan_rmax) * p.z + (R(c.rmax1) + R(c.rmax2)) * R(0.5);
    return (pRMax - rho) * R(c.inv_sec_rmax);
  }

  template <typename R>
  VGM_DEV static R safety_to_in(const Data &c, V3<R> point)
  {
    const R distz = vabs(point.z) - R(c.dz);
    bool inside = distz < R(-kConeTolerance);
    const R rsq = point.x * point.x + point.y * point.y;
    inside = inside && (rsq < minus_tol_sq<R>(radius_at<R>(c, point.z, false), c.outer_tol));
    if (c.rmin_treat) inside = inside && (rsq > plus_tol_sq<R>(radius_at<R>(c, point.z, true), c.inner_tol));
    if (c.phi_treat && inside) inside = in_sector<R>(c, point.x, point.y, false);
    if (inside) return R(-1.0);
    const R safeZ = vabs(point.z) - R(c.dz);
    const R safeOuter = -safe_conical<R>(c, point, false);
    const R safeInner = c.rmin_treat ? -safe_conical<R>(c, point, true) : -inf_length<R>();
    R safety = vmax(safeZ, vmax(safeOuter, safeInner));
    if (c.phi_treat) {
      const R d1 = point.x * R(c.w_normal1x) + point.y * R(c.w_normal1y);
      const R d2 = point.x * R(c.w_normal2x) + point.y * R(c.w_normal2y);
      const R safetyPhi = c.lt_pi ? vmax(-d1, -d2) : vmin(-d1, -d2);
      safety = vmax(safetyPhi, safety);
    }
    if (vabs(safety) < R(kTolerance)) safety = R(0);
    return safety;
  }

  template <typename R>
  VGM_DEV static R safety_to_out(const Data &c, V3<R> point)
  {
    const R distz = vabs(point.z) - R(c.dz);
    const R rsq = point.x * point.x + point.y * point.y;
    bool outside = distz > R(kConeTolerance);
    outside = outside || (rsq > plus_tol_sq<R>(radius_at<R>(c, point.z, false), c.outer_tol));
    if (c.rmin_treat) outside = outside || (rsq < minus_tol_sq<R>(radius_at<R>(c, point.z, true), c.inner_tol));
    if (outside) return R(-1.0);
    const R safeZ = R(c.dz) - vabs(point.z);
    const R safeOuter = safe_conical<R>(c, point, false);
    const R safeInner = c.rmin_treat ? safe_conical<R>(c, point, true) : inf_length<R>();
    R safety = vmin(safeZ, vmin(safeOuter, safeInner));
    if (c.phi_treat) {
      const R safetyPhi =
          wedge_safety_to_out<R>(c.w_normal1x, c.w_normal1y, c.w_normal2x, c.w_normal2y, c.lt_pi, point);
      safety = vmin(safetyPhi, safety);
    }
    if (vabs(safety) < R(kTolerance)) safety = R(0);
    return safety;
  }

  // ConeStruct::Normal (ConeStruct.h:177-306).  The reference's _frmin /
  // _frmax are the constructor's radii: rmin as given, rmax before the
  // kConeTolerance bump.  The phi-plane normals come from the normalised
  // start and end angles (along1/along2 hold their cosines and sines).
  VGM_DEV static bool normal(const Data &c, V3<double> p, V3<double> &norm)
  {
    const double frmin1 = c.rmin1, frmin2 = c.rmin2, frmax2 = c.orig_rmax2, frmax1 = c.orig_rmax1;
    int noSurfaces = 0;
    double distSPhi = kInfLength, distEPhi = kInfLength;
    V3<double> sumnorm{0., 0., 0.};
    V3<double> nR{0., 0., 0.}, nr{0., 0., 0.}, nPs{0., 0., 0.}, nPe{0., 0., 0.};
    norm = sumnorm;
    double distZ = vabs(p.z) - c.dz;
    const double rho = __builtin_sqrt(p.x * p.x + p.y * p.y);
    double distRMin, distRMax;
    const double pRMin = rho - p.z * c.tan_rmin;
    const double widRMin = c.rmin2 - c.dz * c.tan_rmin;
    if (vabs(frmin1 - frmin2) < c.inner_tol)
      distRMin = (rho - frmin2);
    else
      distRMin = (pRMin - widRMin) / c.sec_rmin;
    const double pRMax = rho - p.z * c.tan_rmax;
    const double widRMax = c.rmax2 - c.dz * c.tan_rmax;
    if (vabs(frmax1 - frmax2) < c.outer_tol)
      distRMax = (rho - frmax2);
    else
      distRMax = (pRMax - widRMax) / c.sec_rmax;
    const bool hasRmin = c.rmin1 != 0. || c.rmin2 != 0.;
    bool inside = distZ < kTolerance && distRMax < c.outer_tol;
    if (hasRmin) inside &= distRMin > -c.inner_tol;
    distZ = vabs(distZ);
    distRMax = vabs(distRMax);
    distRMin = vabs(distRMin);
    double distNearest = distZ;
    V3<double> normNearest{0., 0., 1.};
    if (p.z < 0.) normNearest = V3<double>{0., 0., -1.};
    const bool fullPhi = c.dphi == kTwoPi;
    if (!fullPhi) {
      if (rho != 0.) {
        double pPhi = atan2(p.y, p.x);
        if (pPhi < c.sphi - kHalfTolerance)
          pPhi += 2 * kPi;
        else if (pPhi > c.sphi + c.dphi + kHalfTolerance)
          pPhi -= 2 * kPi;
        distSPhi = rho * (pPhi - c.sphi);
        distEPhi = rho * (pPhi - c.sphi - c.dphi);
        inside = inside && (distSPhi > -kTolerance) && (distEPhi < kTolerance);
        distSPhi = vabs(distSPhi);
        distEPhi = vabs(distEPhi);
      } else if (!(c.rmin1 != 0.) || !(c.rmin2 != 0.)) {
        distSPhi = 0.;
        distEPhi = 0.;
      }
      nPs = V3<double>{c.along1y, -c.along1x, 0.};
      nPe = V3<double>{-c.along2y, c.along2x, 0.};
    }
    if (rho > kHalfTolerance) {
      nR = V3<double>{p.x / rho / c.sec_rmax, p.y / rho / c.sec_rmax, -c.tan_rmax / c.sec_rmax};
      if (hasRmin) nr = V3<double>{-p.x / rho / c.sec_rmin, -p.y / rho / c.sec_rmin, c.tan_rmin / c.sec_rmin};
    }
    auto acc = [&](const V3<double> &v) {
      sumnorm.x += v.x;
      sumnorm.y += v.y;
      sumnorm.z += v.z;
    };
    if (inside && distZ <= kHalfTolerance) {
      noSurfaces++;
      if (p.z >= 0.)
        acc(V3<double>{0., 0., 1.});
      else
        sumnorm = V3<double>{0., 0., -1.};
    }
    if (inside && distRMax <= c.outer_tol) {
      noSurfaces++;
      acc(nR);
    } else if (noSurfaces == 0 && distRMax < distNearest) {
      distNearest = distRMax;
      normNearest = nR;
    }
    if (hasRmin) {
      if (inside && distRMin <= c.inner_tol) {
        noSurfaces++;
        acc(nr);
      } else if (noSurfaces == 0 && distRMin < distNearest) {
        distNearest = distRMin;
        normNearest = nr;
      }
    }
    if (!fullPhi) {
      if (inside && distSPhi <= kHalfTolerance) {
        noSurfaces++;
        acc(nPs);
      } else if (noSurfaces == 0 && distSPhi < distNearest) {
        distNearest = distSPhi;
        normNearest = nPs;
      }
      if (inside && distEPhi <= kHalfTolerance) {
        noSurfaces++;
        acc(nPe);
      } else if (noSurfaces == 0 && distEPhi < distNearest) {
        normNearest = nPe;
      }
    }
    if (noSurfaces == 0)
      norm = normNearest;
    else if (noSurfaces == 1)
      norm = sumnorm;
    else
      norm = v3_unit(sumnorm);
    bool valid = noSurfaces != 0;
    if (noSurfaces > 2) valid = false;
    return valid;
  }
};

}  // namespace vgm
