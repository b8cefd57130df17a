// Tube kernels, UniversalTube runtime-checked variant
// (VecGeom/volumes/kernel/TubeImplementation.h, TUBE_SAFETY_OLD; shape type
// traits VecGeom/volumes/kernel/shapetypes/TubeTypes.h).  One code path per
// thread; the rmin/phi branches are wave-uniform because the parameters are.
#pragma once
#include "../vgm_math.hpp"
#include "tube_data.h"
#include "wedge.hpp"
#include "mixed.hpp"

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
  // mixed precision: Inside/Contains classified in float away from the surface
  static constexpr bool kFloatClassify = true;
  // mixed precision: the float safety less the guard band bounds the FP64 safety
  static constexpr bool kFloatSafety = true;
  // mixed precision: distance_to_in_mixed / distance_to_out_mixed (float selects
  // the surface, FP64 evaluates it)
  static constexpr bool kMixedDistance = true;

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

  // DistanceToInKernel (:507-681), scalar early returns: the tolerance
  // decisions up to and including the z-plane entry (true: `distance` is the
  // result), then the curved and phi surfaces
  template <typename R>
  VGM_DEV static bool din_prologue(const Data &t, V3<R> point, V3<R> dir, R &distance, R &rsq, R &rdotn)
  {
    bool done = false;
    distance = inf_length<R>();
    R distz = vabs(point.z) - R(t.z);
    done = done || (distz > R(kHalfTolerance) && point.z * dir.z >= R(0));
    rsq = point.x * point.x + point.y * point.y;
    rdotn = point.x * dir.x + point.y * dir.y;
    done = done || (rsq > R(t.tolIrmax2) && rdotn >= R(0));
    if (done) return true;
    distance = R(-1.0);
    bool inside = distz < R(-kHalfTolerance);
    inside = inside && (rsq < R(t.tolIrmax2));
    if (t.rmin_treat) inside = inside && (rsq > R(t.tolIrmin2));
    if (t.phi_treat && inside) inside = in_sector<R>(t, point.x, point.y, false);
    if (inside) return true;
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
    return done;
  }

  // the rmax circle (returned at once when hit), rmin circle and phi planes;
  // `which` >= 0 evaluates that candidate alone (0 rmax, 1 rmin, 2/3 phi)
  // and returns it, kInfLength when it is rejected
  template <typename R>
  VGM_DEV static R din_surfaces(const Data &t, V3<R> point, V3<R> dir, R rsq, R rdotn, int which = -1)
  {
    R distance = inf_length<R>();
    const R invnsq = R(1.) / nonzero(R(1.) - dir.z * dir.z);
    const R b = invnsq * rdotn;
    if (which < 0 || which == 0) {
      const R crmax = invnsq * (rsq - R(t.rmax2));
      R dist_rmax;
      bool ok_rmax = circle<R, false, true>(t, b, crmax, point, dir, dist_rmax);
      ok_rmax = ok_rmax && (dist_rmax < distance);
      if (ok_rmax) return dist_rmax;
      if (which == 0) return distance;
    }
    if (t.rmin_treat && (which < 0 || which == 1)) {
      const R crmin = invnsq * (rsq - R(t.rmin2));
      R dist_rmin = -inf_length<R>();
      bool ok_rmin = circle<R, true, true>(t, b, crmin, point, dir, dist_rmin);
      if (ok_rmin && dist_rmin < distance) distance = dist_rmin;
    }
    if (t.phi_treat) {
      R dist_phi;
      if (which < 0 || which == 2) {
        bool ok_phi = phi_plane<R, true>(t, t.along1x, t.along1y, t.normal1x, t.normal1y, point, dir, dist_phi);
        if (ok_phi && dist_phi < distance) distance = dist_phi;
      }
      if (which < 0 || which == 3) {
        bool ok_phi = phi_plane<R, true>(t, t.along2x, t.along2y, t.normal2x, t.normal2y, point, dir, dist_phi);
        if (ok_phi && dist_phi < distance) distance = dist_phi;
      }
    }
    return distance;
  }

  template <typename R>
  VGM_DEV static R distance_to_in_kernel(const Data &t, V3<R> point, V3<R> dir)
  {
    R distance, rsq, rdotn;
    if (din_prologue<R>(t, point, dir, distance, rsq, rdotn)) return distance;
    return din_surfaces<R>(t, point, dir, rsq, rdotn);
  }

  // DistanceToOut (:686-821); UniversalTube takes the last phi branch.
  // The wrong-side tests (true: the point is outside, the result is -1)...
  template <typename R>
  VGM_DEV static bool dout_outside(const Data &t, V3<R> point, R &crmax, R &crmin)
  {
    R distz = R(t.z) - vabs(point.z);
    if (distz < R(-kHalfTolerance)) return true;
    const R rsq = point.x * point.x + point.y * point.y;
    crmax = rsq - R(t.rmax2);
    crmin = rsq;
    if (crmax > R(2.0 * kTolerance) * R(t.rmax)) return true;
    if (t.rmin_treat) {
      crmin -= R(t.rmin2);
      if (crmin < R(-2.0 * kTolerance) * R(t.rmin)) return true;
    }
    if (t.phi_treat) {
      bool ci = false, co = false;
      wedge_generic<R, true>(t.along1x, t.along1y, t.along2x, t.along2y, t.normal1x, t.normal1y, t.normal2x,
                             t.normal2y, t.lt_pi, point, ci, co);
      if (co) return true;
    }
    return false;
  }

  // ...then the smallest of the z plane (0), rmin circle (1), rmax circle
  // (2) and phi planes (3, 4); `which` >= 0 evaluates that candidate alone
  template <typename R>
  VGM_DEV static R dout_surfaces(const Data &t, V3<R> point, V3<R> dir, R crmax, R crmin, int which = -1)
  {
    const R rdotn = dir.x * point.x + dir.y * point.y;
    R distance = inf_length<R>();
    if (which < 0 || which == 0) {
      const R invdirz = R(1.) / nonzero(dir.z);
      R distz = (R(t.z) - point.z) * invdirz;
      if (dir.z < R(0)) distz = (-R(t.z) - point.z) * invdirz;
      if (dir.z != R(0) && distz < distance) distance = distz;
      if (which == 0) return distance;
    }
    const R invnsq = R(1.) / nonzero(R(1.) - dir.z * dir.z);
    const R b = invnsq * rdotn;
    if (t.rmin_treat && (which < 0 || which == 1)) {
      R dist_rmin = inf_length<R>();
      crmin *= invnsq;
      bool ok_rmin = circle<R, false, false>(t, b, crmin, point, dir, dist_rmin);
      if (ok_rmin && dist_rmin < distance) distance = dist_rmin;
    }
    if (which < 0 || which == 2) {
      R dist_rmax = inf_length<R>();
      crmax *= invnsq;
      bool ok_rmax = circle<R, true, false>(t, b, crmax, point, dir, dist_rmax);
      if (ok_rmax && dist_rmax < distance) distance = dist_rmax;
    }
    if (t.phi_treat) {
      R dist_phi;
      if (which < 0 || which == 3) {
        bool ok_phi = phi_plane<R, false>(t, t.along1x, t.along1y, t.normal1x, t.normal1y, point, dir, dist_phi);
        if (ok_phi && dist_phi < distance) distance = dist_phi;
      }
      if (which < 0 || which == 4) {
        bool ok_phi = phi_plane<R, false>(t, t.along2x, t.along2y, t.normal2x, t.normal2y, point, dir, dist_phi);
        if (ok_phi && dist_phi < distance) distance = dist_phi;
      }
    }
    return distance;
  }

  template <typename R>
  VGM_DEV static R distance_to_out(const Data &t, V3<R> point, V3<R> dir, R /*step_max*/)
  {
    R crmax, crmin;
    if (dout_outside<R>(t, point, crmax, crmin)) return R(-1.);
    return dout_surfaces<R>(t, point, dir, crmax, crmin);
  }

  // ---- mixed precision: float selects the surface, FP64 evaluates it -------
  // PointInCyclicalSector at a point known to within e, certified
  VGM_MIX static int in_sector_cert(const Data &t, float x, float y, float e, bool includeSurface)
  {
    const float startCheck = (-x * float(t.along1y) + y * float(t.along1x));
    const float endCheck = (-float(t.along2x) * y + float(t.along2y) * x);
    const float tol = includeSurface ? float(-kHalfTolerance) : float(kHalfTolerance);
    const float m = 2.0f * e + kMixEps * (vabs(x) + vabs(y));
    const int a = cert_ge(startCheck, tol, m), b = cert_ge(endCheck, tol, m);
    return t.le_pi ? cert_and(a, b) : cert_or(a, b);
  }

  // CircleTrajectoryIntersection in float: b, c as in the FP64 kernel,
  // `mag` bounds the magnitudes whose rounding enters the discriminant
  // ((rsq + R^2) / (1 - dz^2)^2), P = |p|_1, R the radius, m the length
  // margin of the query
  template <bool Largest, bool InsectorCheck>
  VGM_MIX static FCand circle_f(const Data &t, float b, float c, float mag, float P, float R, V3<float> pos,
                                V3<float> dir, float m)
  {
    const float d2 = b * b - c;
    const float ed = kMixEps * (mag + b * b + vabs(c));
    if (d2 < -ed) return cand_no();  // the FP64 test delta > 0 (>= 0) fails
    if (d2 <= ed) return cand_unsure();
    const float sq = mix_sqrt(d2);
    const float dist = Largest ? -b + sq : -b - sq;
    const float err = mix_div(ed, sq) + kMixEps * (P + vabs(dist) + R) + m;
    int acc = cert_ge(dist, float(-2.0 * kTolerance), err);
    if (InsectorCheck && acc != kCertNo) {
      const float hitz = pos.z + dist * dir.z;
      acc = cert_and(acc, cert_le(vabs(hitz), float(t.z), err));
      if (t.phi_treat && acc != kCertNo)
        acc = cert_and(acc, in_sector_cert(t, pos.x + dist * dir.x, pos.y + dist * dir.y, err, true));
    }
    return FCand{acc, dist, err};
  }

  // PhiPlaneTrajectoryIntersection in float, certified
  template <bool InsectorCheck>
  VGM_MIX static FCand phi_plane_f(const Data &t, double ax_, double ay_, double nx_, double ny_, V3<float> pos,
                                   V3<float> dir, float m)
  {
    const float ax = float(ax_), ay = float(ay_), nx = float(nx_), ny = float(ny_);
    const float dirDotNorm = dir.x * nx + dir.y * ny;
    int acc = InsectorCheck ? cert_ge(dirDotNorm, 0.0f, kMixEps) : cert_le(dirDotNorm, 0.0f, kMixEps);
    if (acc == kCertNo) return cand_no();
    const float dirDotXY = (dir.y * ax - dir.x * ay);
    if (vabs(dirDotXY) < 1.0e-3f) return cand_unsure();  // grazing: the distance is ill-conditioned
    const float rxy = __builtin_amdgcn_rcpf(dirDotXY);
    const float dist = (ay * pos.x - ax * pos.y) * rxy;
    const float err = (kMixEps * (vabs(pos.x) + vabs(pos.y)) + m) * vabs(rxy) + kMixEps * vabs(dist);
    acc = cert_and(acc, cert_ge(dist * vabs(dirDotNorm), float(-kHalfTolerance), err));
    if (acc == kCertNo) return cand_no();
    const float hitx = pos.x + dist * dir.x;
    const float hity = pos.y + dist * dir.y;
    const float eh = err + kMixEps * vabs(dist);
    if (InsectorCheck) {
      const float hitz = pos.z + dist * dir.z;
      const float r2 = hitx * hitx + hity * hity;
      const float er2 = 2.0f * mix_sqrt(r2) * eh + eh * eh + kMixEps * r2;
      acc = cert_and(acc, cert_le(vabs(hitz), float(t.tolOz), eh));
      acc = cert_and(acc, cert_ge(r2, float(t.tolOrmin2), er2));
      acc = cert_and(acc, cert_le(r2, float(t.tolOrmax2), er2));
      acc = cert_and(acc, cert_ge(hitx * ax + hity * ay, 0.0f, 2.0f * eh));
    } else {
      acc = cert_and(acc, cert_ge(hitx * ax + hity * ay, 0.0f, 2.0f * eh));
    }
    return FCand{acc, dist, err};
  }

  // DistanceToIn, mixed: the FP64 kernel's tolerance decisions (far-point
  // pre-step test, inside, on-surface, z-plane entry) in FP64, the curved and
  // phi surfaces selected in float, the selected one evaluated by the FP64
  // kernel's expressions.  Equal to distance_to_in<double> bit for bit.
  VGM_MIX static double distance_to_in_mixed(const Data &t, V3<double> pt, V3<double> dir, float scale)
  {
    const V3<float> pf{float(pt.x), float(pt.y), float(pt.z)};
    const float lim = 100.0f * float(t.maxval);
    if (!(pf.x * pf.x + pf.y * pf.y + pf.z * pf.z < 0.98f * lim * lim))  // far-point pre-step: FP64
      return distance_to_in<double>(t, pt, dir, kInfLength);
    const double distToMove = 0.0;  // the FP64 kernel moves points beyond 100 maxval only
    double distance, rsq, rdotn;
    if (din_prologue<double>(t, pt, dir, distance, rsq, rdotn)) return distance + distToMove;
    const V3<float> df{float(dir.x), float(dir.y), float(dir.z)};
    const float nsq = float(1.0 - dir.z * dir.z);  // FP64 products: no cancellation error in float
    if (nsq < 1.0e-2f) return din_surfaces<double>(t, pt, dir, rsq, rdotn) + distToMove;  // near-axial
    const float P = vabs(pf.x) + vabs(pf.y) + vabs(pf.z);
    const float m = kMixLen * (P + scale);
    const float invnsq = __builtin_amdgcn_rcpf(nsq);
    const float rsqf = float(rsq), b = invnsq * float(rdotn);
    int which;
    const FCand cmax = circle_f<false, true>(t, b, invnsq * (rsqf - float(t.rmax2)),
                                             (rsqf + float(t.rmax2)) * invnsq * invnsq, P, float(t.rmax), pf, df, m);
    if (cmax.acc == kCertUnsure) return din_surfaces<double>(t, pt, dir, rsq, rdotn) + distToMove;
    if (cmax.acc == kCertYes) {
      which = 0;  // an rmax hit is returned before the other surfaces
    } else {
      FSelect sel;
      if (t.rmin_treat)
        sel.add(1, circle_f<true, true>(t, b, invnsq * (rsqf - float(t.rmin2)),
                                        (rsqf + float(t.rmin2)) * invnsq * invnsq, P, float(t.rmin), pf, df, m));
      if (t.phi_treat) {
        sel.add(2, phi_plane_f<true>(t, t.along1x, t.along1y, t.normal1x, t.normal1y, pf, df, m));
        sel.add(3, phi_plane_f<true>(t, t.along2x, t.along2y, t.normal2x, t.normal2y, pf, df, m));
      }
      if (sel.unsure) return din_surfaces<double>(t, pt, dir, rsq, rdotn) + distToMove;
      if (sel.best < 0) return kInfLength + distToMove;  // certified miss
      which = sel.best;
    }
    const double r = din_surfaces<double>(t, pt, dir, rsq, rdotn, which);
    if (!(r < kInfLength)) return din_surfaces<double>(t, pt, dir, rsq, rdotn) + distToMove;
    return r + distToMove;
  }

  // DistanceToOut, mixed: wrong-side tests in FP64, the smallest surface
  // selected in float and evaluated by the FP64 kernel's expressions.
  VGM_MIX static double distance_to_out_mixed(const Data &t, V3<double> pt, V3<double> dir, float scale)
  {
    double crmax, crmin;
    if (dout_outside<double>(t, pt, crmax, crmin)) return -1.;
    const V3<float> pf{float(pt.x), float(pt.y), float(pt.z)};
    const V3<float> df{float(dir.x), float(dir.y), float(dir.z)};
    const float nsq = float(1.0 - dir.z * dir.z);  // FP64 products: no cancellation error in float
    if (nsq < 1.0e-2f) return dout_surfaces<double>(t, pt, dir, crmax, crmin);  // near-axial
    const float P = vabs(pf.x) + vabs(pf.y) + vabs(pf.z);
    const float m = kMixLen * (P + scale);
    FSelect sel;
    if (dir.z != 0.0) {  // the z plane is accepted exactly when dir.z != 0
      const float zt = dir.z < 0.0 ? -float(t.z) : float(t.z);
      sel.add(0, z_plane_cand(zt, pf.z, df.z, m));
    }
    const float invnsq = __builtin_amdgcn_rcpf(nsq);
    const float rsqf = pf.x * pf.x + pf.y * pf.y;
    const float b = invnsq * (df.x * pf.x + df.y * pf.y);
    if (t.rmin_treat)
      sel.add(1, circle_f<false, false>(t, b, float(crmin) * invnsq, (rsqf + float(t.rmin2)) * invnsq * invnsq, P,
                                        float(t.rmin), pf, df, m));
    sel.add(2, circle_f<true, false>(t, b, float(crmax) * invnsq, (rsqf + float(t.rmax2)) * invnsq * invnsq, P,
                                     float(t.rmax), pf, df, m));
    if (t.phi_treat) {
      sel.add(3, phi_plane_f<false>(t, t.along1x, t.along1y, t.normal1x, t.normal1y, pf, df, m));
      sel.add(4, phi_plane_f<false>(t, t.along2x, t.along2y, t.normal2x, t.normal2y, pf, df, m));
    }
    if (sel.unsure || sel.best < 0) return dout_surfaces<double>(t, pt, dir, crmax, crmin);
    const double r = dout_surfaces<double>(t, pt, dir, crmax, crmin, sel.best);
    if (!(r < kInfLength)) return dout_surfaces<double>(t, pt, dir, crmax, crmin);
    return r;
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

  // IsOnSurfaceGeneric (Wedge_Evolution.h:343-354)
  VGM_DEV static bool on_phi_plane(double ax, double ay, double nx, double ny, V3<double> p)
  {
    if (!(ax * p.x + ay * p.y >= 0.)) return false;
    return vabs(nx * p.x + ny * p.y) < kTolerance;
  }

  // ApproxSurfaceNormalKernel (TubeImplementation.h:969-1015): the nearest
  // surface of a point that is not on one
  VGM_DEV static V3<double> approx_normal(const Data &t, V3<double> p)
  {
    V3<double> norm{0., 0., 0.};
    const double radius = __builtin_sqrt(p.x * p.x + p.y * p.y);
    double distRMax = vabs(radius - t.rmax);
    double distRMin = kInfLength;
    if (distRMax < 0.) distRMax = kInfLength;
    if (t.rmin != 0.) {
      distRMin = vabs(t.rmin - radius);
      if (distRMin < 0.) distRMin = kInfLength;
    }
    double distMin = vmin(distRMin, distRMax);
    double distPhi1 = kInfLength, distPhi2 = kInfLength;
    if (t.dphi != kTwoPi) {
      distPhi1 = p.x * t.normal1x + p.y * t.normal1y;
      distPhi2 = p.x * t.normal2x + p.y * t.normal2y;
      if (distPhi1 < 0.) distPhi1 = kInfLength;
      if (distPhi2 < 0.) distPhi2 = kInfLength;
      distMin = vmin(distMin, vmin(distPhi1, distPhi2));
    }
    double distZ = kInfLength;
    if (p.z < 0.) distZ = vabs(p.z + t.z);
    if (p.z >= 0.) distZ = vabs(p.z - t.z);
    distMin = vmin(distMin, distZ);
    if (t.dphi != 0.) {
      if (distMin == distPhi1) norm = V3<double>{-t.normal1x, -t.normal1y, -0.0};
      if (distMin == distPhi2) norm = V3<double>{-t.normal2x, -t.normal2y, -0.0};
    }
    if ((distMin == distZ) && (p.z < 0.)) norm = V3<double>{0., 0., -1.};
    if ((distMin == distZ) && (p.z >= 0.)) norm = V3<double>{0., 0., 1.};
    if (vabs(p.z) < (t.z + kTolerance)) {
      const V3<double> temp{p.x, p.y, 0.};
      if (distMin == distRMax) norm = v3_unit(temp);
      if (t.rmin != 0.)
        if (distMin == distRMin) {
          const V3<double> u = v3_unit(temp);
          norm = V3<double>{-u.x, -u.y, -u.z};
        }
    }
    return norm;
  }

  // NormalKernel (TubeImplementation.h:1019-1100): the normals of the
  // surfaces the point is on, summed and divided by the root of their number
  VGM_DEV static bool normal(const Data &t, V3<double> p, V3<double> &norm)
  {
    bool ci = false, co = false;
    generic<double, true>(t, p, ci, co);
    if (ci || co) {
      norm = approx_normal(t, p);
      return false;
    }
    norm = V3<double>{0., 0., 0.};
    int nosurface = 0;
    const double x2y2 = __builtin_sqrt(p.x * p.x + p.y * p.y);
    const bool inZ = ((p.z < t.z + kTolerance) && (p.z > -t.z - kTolerance));
    const bool inR = ((x2y2 >= t.rmin - kTolerance) && (x2y2 <= t.rmax + kTolerance));
    if (inR && (vabs(p.z - t.z) <= kTolerance)) {
      norm = V3<double>{0., 0., 1.};
      nosurface++;
    }
    if (inR && (vabs(p.z + t.z) <= kTolerance)) {
      if (nosurface > 0)
        norm.z += -1.;
      else
        norm = V3<double>{0., 0., -1.};
      nosurface++;
    }
    if (t.rmin > 0.) {
      if (inZ && (vabs(x2y2 - t.rmin) <= kTolerance)) {
        const double inv = 1. / x2y2;
        if (nosurface == 0) {
          norm.x = -p.x * inv;
          norm.y = -p.y * inv;
          norm.z = 0.;
        } else {
          norm.x += -p.x * inv;
          norm.y += -p.y * inv;
        }
        nosurface++;
      }
    }
    if (inZ && (vabs(x2y2 - t.rmax) <= kTolerance)) {
      const double inv = 1. / x2y2;
      if (nosurface > 0) {
        norm.x += p.x * inv;
        norm.y += p.y * inv;
      } else {
        norm.x = p.x * inv;
        norm.y = p.y * inv;
        norm.z = 0.;
      }
      nosurface++;
    }
    if (t.dphi < kTwoPi) {
      if (inR && on_phi_plane(t.along1x, t.along1y, t.normal1x, t.normal1y, p)) {
        if (nosurface == 0) {
          norm = V3<double>{-t.normal1x, -t.normal1y, -0.0};
        } else {
          norm.x += -t.normal1x;
          norm.y += -t.normal1y;
          norm.z += -0.0;
        }
        nosurface++;
      }
      if (inR && on_phi_plane(t.along2x, t.along2y, t.normal2x, t.normal2y, p)) {
        if (nosurface == 0) {
          norm = V3<double>{-t.normal2x, -t.normal2y, -0.0};
        } else {
          norm.x += -t.normal2x;
          norm.y += -t.normal2y;
          norm.z += -0.0;
        }
        nosurface++;
      }
    }
    if (nosurface > 1) {
      const double s = __builtin_sqrt(1. * nosurface);
      norm = V3<double>{norm.x / s, norm.y / s, norm.z / s};
    }
    return nosurface != 0;
  }
};

}  // namespace vgm
