// Trd kernels, UniversalTrd variant with scalar semantics
// (VecGeom/volumes/kernel/TrdImplementation.h:60-522).
#pragma once
#include "../vgm_math.hpp"
#include "trd_data.h"

namespace vgm {

struct TrdShape {
  using Data = TrdData;
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
  VGM_DEV static R clamp0(R d) { return vabs(d) < R(kHalfTolerance) ? R(0) : d; }

  // FaceTrajectoryIntersection (:83-138)
  template <typename R, bool ForY, bool Mirrored, bool ToInside>
  VGM_DEV static bool face(const Data &t, V3<R> pos, V3<R> dir, R &dist)
  {
    const R alongV = ForY ? R(t.y2minusy1) : R(t.x2minusx1);
    const R v1 = ForY ? R(t.dy1) : R(t.dx1);
    R posV = ForY ? pos.y : pos.x;
    const R posK = ForY ? pos.x : pos.y;
    R dirV = ForY ? dir.y : dir.x;
    const R dirK = ForY ? dir.x : dir.y;
    const R fK = ForY ? R(t.fx) : R(t.fy);
    const R fV = ForY ? R(t.fy) : R(t.fx);
    const R halfKplus = ForY ? R(t.half_x1plusx2) : R(t.half_y1plusy2);
    if (Mirrored) {
      posV *= R(-1.);
      dirV *= R(-1.);
    }
    const R ndotv = dirV + fV * dir.z;
    if (!(ToInside ? ndotv < R(0) : ndotv > R(0))) return false;
    const R alongZ = R(2.0) * R(t.dz);
    dist = (alongZ * (posV - v1) - alongV * (pos.z + R(t.dz))) / (dir.z * alongV - dirV * alongZ + R(1e-30));
    if (!(dist > R(-kHalfTolerance))) return false;
    const R hitz = pos.z + dist * dir.z;
    bool ok = vabs(hitz) <= R(t.dz);
    const R hitk = posK + dist * dirK;
    const R dK = halfKplus - fK * hitz;
    ok = ok && (vabs(hitk) <= dK);
    if (ok && (vabs(dist) < R(kHalfTolerance))) dist = R(0);
    return ok;
  }

  template <typename R, bool Inside>
  VGM_DEV static R safety(const Data &t, V3<R> pos)
  {
    R dist = R(t.dz) - vabs(pos.z);
    const R distx = R(t.half_x1plusx2) - R(t.fx) * pos.z;
    const R safx = (distx - vabs(pos.x)) * R(t.calfx);
    if (distx >= R(0) && safx < dist) dist = safx;
    if (t.vary_y) {
      const R disty = R(t.half_y1plusy2) - R(t.fy) * pos.z;
      const R safy = (disty - vabs(pos.y)) * R(t.calfy);
      if (disty >= R(0) && safy < dist) dist = safy;
    } else {
      const R safy = R(t.dy1) - vabs(pos.y);
      if (safy < dist) dist = safy;
    }
    return Inside ? dist : -dist;
  }

  template <typename R, bool SurfaceT>
  VGM_DEV static void uinside(const Data &t, V3<R> p, bool &cin, bool &cout)
  {
    const R pzPlusDz = p.z + R(t.dz);
    cout = vabs(p.z) > (SurfaceT ? R(t.dz) + R(kHalfTolerance) : R(t.dz));
    if (SurfaceT) cin = vabs(p.z) < R(t.dz) - R(kHalfTolerance);
    R cross = R(t.x2minusx1) * pzPlusDz - R(2.0 * t.dz) * (vabs(p.x) - R(t.dx1));
    if (SurfaceT) {
      cout = cout || (cross < -R(t.tolx));
      cin = cin && (cross > R(t.tolx));
    } else {
      cout = cout || (cross < R(0));
    }
    cross = R(t.y2minusy1) * pzPlusDz - R(2.0 * t.dz) * (vabs(p.y) - R(t.dy1));
    if (SurfaceT) {
      cout = cout || (cross < -R(t.toly));
      cin = cin && (cross > R(t.toly));
    } else {
      cout = cout || (cross < R(0));
    }
  }

  template <typename R>
  VGM_DEV static bool contains(const Data &t, V3<R> p)
  {
    bool ci = false, co = false;
    uinside<R, false>(t, p, ci, co);
    return !co;
  }

  template <typename R>
  VGM_DEV static int inside(const Data &t, V3<R> p)
  {
    bool ci = false, co = false;
    uinside<R, true>(t, p, ci, co);
    return co ? kOutsideV : (ci ? kInsideV : kSurfaceV);
  }

  template <typename R>
  VGM_DEV static R distance_to_in(const Data &t, V3<R> point, V3<R> direction, R /*step_max*/)
  {
    R distance = inf_length<R>();
    const bool inz = vabs(point.z) < R(t.dz) - R(kHalfTolerance);
    R distx = R(t.half_x1plusx2) - R(t.fx) * point.z;
    const bool inx = (distx - vabs(point.x)) * R(t.calfx) > R(kHalfTolerance);
    R disty;
    bool iny;
    if (t.vary_y) {
      disty = R(t.half_y1plusy2) - R(t.fy) * point.z;
      iny = (disty - vabs(point.y)) * R(t.calfy) > R(kHalfTolerance);
    } else {
      disty = vabs(point.y) - R(t.dy1);
      iny = disty < R(-kHalfTolerance);
    }
    if (inx && iny && inz) return clamp0(R(-1.));
    bool okz = (point.z * direction.z < R(0)) && !inz;
    if (okz) {
      const R distz = (vabs(point.z) - R(t.dz)) / vabs(direction.z);
      const R hitx = vabs(point.x + distz * direction.x);
      const R hity = vabs(point.y + distz * direction.y);
      const bool okzt = point.z > (R(t.dz) - R(kHalfTolerance)) && hitx <= R(t.dx2) && hity <= R(t.dy2);
      const bool okzb = point.z < (-R(t.dz) + R(kHalfTolerance)) && hitx <= R(t.dx1) && hity <= R(t.dy1);
      if (okzt || okzb) return clamp0(distz);
    }
    if (!inx) {
      bool okx = face<R, false, false, true>(t, point, direction, distx);
      if (okx) distance = distx;
      okx = face<R, false, true, true>(t, point, direction, distx);
      if (okx) return clamp0(distx);
    }
    if (t.vary_y) {
      if (!iny) {
        if (face<R, true, false, true>(t, point, direction, disty)) distance = disty;
        if (face<R, true, true, true>(t, point, direction, disty)) distance = disty;
      }
    } else if (!iny) {
      disty /= vabs(direction.y);
      const R zhit = point.z + disty * direction.z;
      const R xhit = point.x + disty * direction.x;
      const R dx = R(t.half_x1plusx2) - R(t.fx) * zhit;
      if (point.y * direction.y < R(0) && disty > R(-kHalfTolerance) && vabs(xhit) < dx && vabs(zhit) < R(t.dz))
        distance = disty;
    }
    return clamp0(distance);
  }

  template <typename R>
  VGM_DEV static R distance_to_out(const Data &t, V3<R> point, V3<R> dir, R /*step_max*/)
  {
    R distance = R(0);
    const R invdir = R(1.) / vabs(dir.z + R(1e-30));
    const R safz = R(t.dz) - vabs(point.z);
    bool out = safz < R(-kHalfTolerance);
    R distx = R(t.half_x1plusx2) - R(t.fx) * point.z;
    out = out || ((distx - vabs(point.x)) * R(t.calfx) < R(-kHalfTolerance));
    R disty;
    if (t.vary_y) {
      disty = R(t.half_y1plusy2) - R(t.fy) * point.z;
      out = out || ((disty - vabs(point.y)) * R(t.calfy) < R(-kHalfTolerance));
    } else {
      disty = R(t.dy1) - vabs(point.y);
      out = out || (disty < R(-kHalfTolerance));
    }
    if (out) return R(-1.);
    if (dir.z > R(0)) {
      const R distz = (R(t.dz) - point.z) * invdir;
      const R hitx = vabs(point.x + distz * dir.x);
      const R hity = vabs(point.y + distz * dir.y);
      if (hitx <= R(t.dx2) && hity <= R(t.dy2)) return clamp0(distz);
    }
    if (dir.z < R(0)) {
      const R distz = (point.z + R(t.dz)) * invdir;
      const R hitx = vabs(point.x + distz * dir.x);
      const R hity = vabs(point.y + distz * dir.y);
      if (hitx <= R(t.dx1) && hity <= R(t.dy1)) return clamp0(distz);
    }
    if (face<R, false, false, false>(t, point, dir, distx)) return clamp0(distx);
    if (face<R, false, true, false>(t, point, dir, distx)) return clamp0(distx);
    if (t.vary_y) {
      if (face<R, true, false, false>(t, point, dir, disty)) return clamp0(disty);
      if (face<R, true, true, false>(t, point, dir, disty)) distance = disty;
    } else {
      const R plane = dir.y < R(0) ? R(-t.dy1) : R(t.dy1);
      disty = (plane - point.y) / dir.y;
      const R zhit = point.z + disty * dir.z;
      const R xhit = point.x + disty * dir.x;
      const R dx = R(t.half_x1plusx2) - R(t.fx) * zhit;
      if (vabs(xhit) < dx && vabs(zhit) < R(t.dz)) distance = disty;
    }
    return clamp0(distance);
  }

  template <typename R>
  VGM_DEV static R safety_to_in(const Data &t, V3<R> p) { return safety<R, false>(t, p); }
  template <typename R>
  VGM_DEV static R safety_to_out(const Data &t, V3<R> p) { return safety<R, true>(t, p); }

  // UnplacedTrd::Normal (source/UnplacedTrd.cpp:277-352): the faces within
  // tolerance of the nearest one, averaged
  VGM_DEV static bool normal(const Data &t, V3<double> p, V3<double> &norm)
  {
    int noSurfaces = 0;
    V3<double> sumnorm{0., 0., 0.};
    const double distz = vabs(vabs(p.z) - t.dz);
    const double xnorm = 1.0 / __builtin_sqrt(4 * t.dz * t.dz + (t.dx2 - t.dx1) * (t.dx2 - t.dx1));
    const double ynorm = 1.0 / __builtin_sqrt(4 * t.dz * t.dz + (t.dy2 - t.dy1) * (t.dy2 - t.dy1));
    double distmx = -2.0 * t.dz * p.x - (t.dx2 - t.dx1) * p.z - t.dz * (t.dx1 + t.dx2);
    distmx *= xnorm;
    double distpx = 2.0 * t.dz * p.x - (t.dx2 - t.dx1) * p.z - t.dz * (t.dx1 + t.dx2);
    distpx *= xnorm;
    double distmy = -2.0 * t.dz * p.y - (t.dy2 - t.dy1) * p.z - t.dz * (t.dy1 + t.dy2);
    distmy *= ynorm;
    double distpy = 2.0 * t.dz * p.y - (t.dy2 - t.dy1) * p.z - t.dz * (t.dy1 + t.dy2);
    distpy *= ynorm;
    double safmin = vmin(vabs(distmx), vabs(distpx));
    safmin = vmin(safmin, vmin(vabs(distmy), vabs(distpy)));
    safmin = vmin(safmin, vabs(distz));
    const bool valid = safmin <= kHalfTolerance;
    auto acc = [&](double a, double b, double c, double s) {
      sumnorm.x += a * s;
      sumnorm.y += b * s;
      sumnorm.z += c * s;
    };
    if (vabs(distmx) - safmin <= kHalfTolerance) {
      noSurfaces++;
      acc(-2.0 * t.dz, 0.0, -(t.dx2 - t.dx1), xnorm);
    }
    if (vabs(distpx) - safmin <= kHalfTolerance) {
      noSurfaces++;
      acc(2.0 * t.dz, 0.0, -(t.dx2 - t.dx1), xnorm);
    }
    if (vabs(distpy) - safmin <= kHalfTolerance) {
      noSurfaces++;
      acc(0.0, 2.0 * t.dz, -(t.dy2 - t.dy1), ynorm);
    }
    if (vabs(distmy) - safmin <= kHalfTolerance) {
      noSurfaces++;
      acc(0.0, -2.0 * t.dz, -(t.dy2 - t.dy1), ynorm);
    }
    if (vabs(distz) - safmin <= kHalfTolerance) {
      noSurfaces++;
      if (p.z >= 0.) {
        sumnorm.x += 0.;
        sumnorm.y += 0.;
        sumnorm.z += 1.;
      } else {
        sumnorm.x += 0.;
        sumnorm.y += 0.;
        sumnorm.z += -1.;
      }
    }
    if (noSurfaces == 0)
      norm = V3<double>{0., 0., 1.};
    else if (noSurfaces == 1)
      norm = sumnorm;
    else
      norm = v3_unit(sumnorm);
    return valid;
  }
};

}  // namespace vgm
