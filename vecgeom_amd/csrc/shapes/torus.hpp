// Torus kernels: VecGeom/volumes/kernel/TorusImplementation2.h (the quartic
// of SolveCubic / SolveQuartic, Newton refinement, the bounding-tube entry)
// and UnplacedTorus2::Normal (source/UnplacedTorus2.cpp:201-285).  Scalar-
// backend semantics; the transport runs it in FP64 only (no mixed path).
// The reference's behaviour is kept as written: Contains ignores the phi
// cut, DistanceToIn takes a phi-plane hit only within kTolerance of the
// point, and ToBoundary's small-|dir.z| branch sorts a rejected root along
// with the accepted ones.
//
// SolveCubic calls cbrt, pow, cos and atan2 (the device's own); Newton
// refinement on the quartic follows, so the distances agree with a host
// evaluation to the refinement's tolerance, not to the last bit.  Written on
// a scalar type R (double on the device; the host FLOP counter's counting
// double, which does not count the four transcendental calls).
#pragma once
#include "../vgm_math.hpp"
#include "torus_data.h"
#include "tube.hpp"
#include "wedge.hpp"

namespace vgm {

struct TorusShape {
  using Data = TorusData;
  static constexpr bool kStage = false;  // parameter block staged in LDS
  // waves per SIMD the transport kernel asks the compiler for (1: no request)
  static constexpr int kWavesPerEU = 1;
  // safety is the distance to the surface (1-Lipschitz): carried bounds are valid
  static constexpr bool kSafetyIsDistance = false;
  // the phi planes' safeties can exceed a torus-surface distance that skips
  // a phi-plane entry: the transport takes the safety-first FP64 step
  static constexpr bool kSafetyBelowDistance = false;
  // mixed precision: none (the quartic is solved in FP64)
  static constexpr bool kFloatClassify = false;
  static constexpr bool kFloatSafety = false;
  static constexpr bool kMixedDistance = false;

  template <typename R>
  VGM_DEV static R t_cbrt(R x)
  {
    return R(cbrt(double(x)));
  }
  template <typename R>
  VGM_DEV static R t_pow(R x, R y)
  {
    return R(pow(double(x), double(y)));
  }
  template <typename R>
  VGM_DEV static R t_cos(R x)
  {
    return R(cos(double(x)));
  }
  template <typename R>
  VGM_DEV static R t_atan2(R y, R x)
  {
    return R(atan2(double(y), double(x)));
  }

  // slot i of a four-slot root array, written without a dynamic index
  template <typename R>
  VGM_DEV static void put(R (&x)[4], int i, R v)
  {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k == i) x[k] = v;
  }
  template <typename R>
  VGM_DEV static void cmp_swap(R &a, R &b)
  {
    if (a > b) {
      const R c = b;
      b = a;
      a = c;
    }
  }
  // Sort4 (TorusImplementation2.h:66-91)
  template <typename R>
  VGM_DEV static void sort4(R (&x)[4])
  {
    cmp_swap(x[0], x[2]);
    cmp_swap(x[1], x[3]);
    cmp_swap(x[0], x[1]);
    cmp_swap(x[2], x[3]);
    cmp_swap(x[1], x[2]);
  }

  // SolveCubic (:24-64): one or three real roots of x^3 + a x^2 + b x + c
  template <typename R>
  VGM_DEV static int solve_cubic(R a, R b, R c, R (&x)[4])
  {
    const R ott = R(1.) / R(3.);
    const R sq3 = vsqrt(R(3.));
    const R inv6sq3 = R(1.) / (R(6.) * sq3);
    int ireal = 1;
    const R p = b - a * a * ott;
    const R q = c - a * b * ott + R(2.) * a * a * a * ott * ott * ott;
    R delta = R(4) * p * p * p + R(27.) * q * q;
    R t, u;
    if (delta >= R(0)) {
      delta = vsqrt(delta);
      t = (R(-3) * q * sq3 + delta) * inv6sq3;
      u = (R(3) * q * sq3 + delta) * inv6sq3;
      x[0] = vcopysign(R(1.), t) * t_cbrt(vabs(t)) - vcopysign(R(1.), u) * t_cbrt(vabs(u)) - a * ott;
    } else {
      delta = vsqrt(-delta);
      t = R(-0.5) * q;
      u = delta * inv6sq3;
      x[0] = R(2.) * t_pow(t * t + u * u, R(0.5) * ott) * t_cos(ott * t_atan2(u, t));
      x[0] -= a * ott;
    }
    t = x[0] * x[0] + a * x[0] + b;
    u = a + x[0];
    delta = u * u - R(4.) * t;
    if (delta >= R(0)) {
      ireal = 3;
      delta = vsqrt(delta);
      x[1] = R(0.5) * (-u - delta);
      x[2] = R(0.5) * (-u + delta);
    }
    return ireal;
  }

  // SolveQuartic (:97-174): real roots of x^4 + a x^3 + b x^2 + c x + d,
  // sorted into x (slots past the count keep what the caller put there)
  template <typename R>
  VGM_DEV static int solve_quartic(R a, R b, R c, R d, R (&x)[4])
  {
    const R e = b - R(3.) * a * a / R(8.);
    const R f = c + a * a * a / R(8.) - R(0.5) * a * b;
    const R g = d - R(3.) * a * a * a * a / R(256.) + a * a * b / R(16.) - a * c / R(4.);
    const R inf = inf_length<R>();
    R xx[4] = {inf, inf, inf, inf};
    R delta;
    R h = R(0.);
    int ireal = 0;
    if (vabs(f) < R(1E-6)) {  // biquadratic
      delta = e * e - R(4.) * g;
      if (delta < R(0.)) return 0;
      delta = vsqrt(delta);
      h = R(0.5) * (-e - delta);
      if (h >= R(0)) {
        h = vsqrt(h);
        put(x, ireal++, -h - R(0.25) * a);
        put(x, ireal++, h - R(0.25) * a);
      }
      h = R(0.5) * (-e + delta);
      if (h >= R(0)) {
        h = vsqrt(h);
        put(x, ireal++, -h - R(0.25) * a);
        put(x, ireal++, h - R(0.25) * a);
      }
      sort4(x);
      return ireal;
    }
    if (vabs(g) < R(1E-6)) {  // one root at -a/4, the cubic for the rest
      x[0] = R(-0.25) * a;
      const int ncubicroots = solve_cubic(R(0), e, f, xx);
      x[1] = xx[0] - R(0.25) * a;
      ireal = 2;
      if (ncubicroots == 3) {
        x[2] = xx[1] - R(0.25) * a;
        x[3] = xx[2] - R(0.25) * a;
        ireal = 4;
      }
      sort4(x);
      return ireal;
    }
    ireal = solve_cubic(R(2.) * e, e * e - R(4.) * g, -f * f, xx);
    if (ireal == 1) {
      if (xx[0] <= R(0)) return 0;
      h = vsqrt(xx[0]);
    } else {
      // the first non-negative of the three roots, else the last
      h = xx[0];
      if (!(h >= R(0))) {
        h = xx[1];
        if (!(h >= R(0))) h = xx[2];
      }
      if (h <= R(0)) return 0;
      h = vsqrt(h);
    }
    const R j = R(0.5) * (e + h * h - f / h);
    ireal = 0;
    delta = h * h - R(4.) * j;
    if (delta >= R(0)) {
      delta = vsqrt(delta);
      x[0] = R(0.5) * (-h - delta) - R(0.25) * a;
      x[1] = R(0.5) * (-h + delta) - R(0.25) * a;
      ireal = 2;
    }
    delta = h * h - R(4.) * g / j;
    if (delta >= R(0)) {
      delta = vsqrt(delta);
      put(x, ireal++, R(0.5) * (h - delta) - R(0.25) * a);
      put(x, ireal++, R(0.5) * (h + delta) - R(0.25) * a);
    }
    sort4(x);
    return ireal;
  }

  // ---- phi wedge (Wedge_Evolution.h) -------------------------------------
  template <typename R>
  VGM_DEV static void wedge(const Data &s, V3<R> p, bool forInside, bool &cin, bool &cout)
  {
    if (forInside)
      wedge_generic<R, true>(s.w_a1x, s.w_a1y, s.w_a2x, s.w_a2y, s.w_n1x, s.w_n1y, s.w_n2x, s.w_n2y, s.w_lt_pi, p,
                             cin, cout);
    else
      wedge_generic<R, false>(s.w_a1x, s.w_a1y, s.w_a2x, s.w_a2y, s.w_n1x, s.w_n1y, s.w_n2x, s.w_n2y, s.w_lt_pi, p,
                              cin, cout);
  }
  // Wedge::DistanceToIn / DistanceToOut: the plane hits ahead of the point,
  // entering (in) or leaving (out) the wedge
  template <typename R>
  VGM_DEV static void wedge_distances(const Data &s, V3<R> p, V3<R> d, bool in, R &w1, R &w2)
  {
    const R c1 = d.x * R(s.w_n1x) + d.y * R(s.w_n1y);
    const R c2 = d.x * R(s.w_n2x) + d.y * R(s.w_n2y);
    w1 = inf_length<R>();
    w2 = inf_length<R>();
    if (in ? c1 > R(0.) : c1 < R(0.)) {
      const R t = -(p.x * R(s.w_n1x) + p.y * R(s.w_n1y)) / c1;
      if (t > R(0.)) w1 = t;
    }
    if (in ? c2 > R(0.) : c2 < R(0.)) {
      const R t = -(p.x * R(s.w_n2x) + p.y * R(s.w_n2y)) / c2;
      if (t > R(0.)) w2 = t;
    }
  }

  // GenericKernelForContainsAndInside (:331-384)
  template <typename R>
  VGM_DEV static void generic(const Data &s, V3<R> p, bool forInside, bool notForDisk, bool &cin, bool &cout)
  {
    const R tol = R(100. * kTolerance);
    const R rxy = vsqrt(p.x * p.x + p.y * p.y);
    const R radsq = (rxy - R(s.rtor)) * (rxy - R(s.rtor)) + p.z * p.z;
    cin = false;
    if (forInside) {
      cout = radsq > (tol * R(s.rmax) + R(s.rmax2));
      cin = radsq < (-tol * R(s.rmax) + R(s.rmax2));
    } else {
      cout = radsq > R(s.rmax2);
    }
    if (cout) return;
    if (forInside) {
      cout |= radsq < (-tol * R(s.rmin) + R(s.rmin2));
      cin &= radsq > (tol * R(s.rmin) + R(s.rmin2));
    } else {
      cout |= radsq < R(s.rmin2);
    }
    if (cout) return;
    if (s.has_phi && notForDisk) {
      bool ci = false, co = false;
      wedge<R>(s, p, forInside, ci, co);
      cout |= co;
      if (forInside) cin &= ci;
    }
  }

  // Contains (:520-529): the tolerant test without the phi cut
  template <typename R>
  VGM_DEV static bool contains(const Data &s, V3<R> p)
  {
    bool ci = false, co = false;
    generic<R>(s, p, true, false, ci, co);
    return !co;
  }

  // InsideKernel (:315-326)
  template <typename R>
  VGM_DEV static int inside(const Data &s, V3<R> p)
  {
    bool ci = false, co = false;
    generic<R>(s, p, true, true, ci, co);
    int in = kSurfaceV;
    if (co) in = kOutsideV;
    if (ci) in = kInsideV;
    return in;
  }

  // DistSqrToTorusR (:190-196)
  template <typename R>
  VGM_DEV static R dist_sqr_to_torus_r(const Data &s, V3<R> p, V3<R> d, R dist)
  {
    const R x = p.x + d.x * dist, y = p.y + d.y * dist, z = p.z + d.z * dist;
    const R rxy = vsqrt(x * x + y * y);
    return (rxy - R(s.rtor)) * (rxy - R(s.rtor)) + z * z;
  }

  // ToBoundary (:389-496): distance along dir to the torus of tube radius
  // `radius`, leaving (out) or entering.  Kept out of line: the entry and
  // exit distances call it twice each.
  template <typename R>
  __host__ __device__ __noinline__ static R to_boundary(const Data &s, V3<R> pt, V3<R> dir, R radius, bool out)
  {
    const R inf = inf_length<R>();
    R sv = inf;
    const R tol = R(100. * kTolerance);
    const R rtor2 = R(s.rtor2);
    const R r0sq = pt.x * pt.x + pt.y * pt.y + pt.z * pt.z;
    const R rdotn = pt.x * dir.x + pt.y * dir.y + pt.z * dir.z;
    const R rsumsq = rtor2 + radius * radius;
    const R a = R(4.) * rdotn;
    const R b = R(2.) * (r0sq + R(2.) * rdotn * rdotn - rsumsq + R(2.) * rtor2 * dir.z * dir.z);
    const R c = R(4.) * (r0sq * rdotn - rsumsq * rdotn + R(2.) * rtor2 * pt.z * dir.z);
    const R d = r0sq * r0sq - R(2.) * r0sq * rsumsq + R(4.) * rtor2 * pt.z * pt.z +
                (rtor2 - radius * radius) * (rtor2 - radius * radius);
    R x[4] = {inf, inf, inf, inf};
    int nsol = 0;
    if (vabs(dir.z) < R(1E-3) && vabs(pt.z) < R(0.1) * radius) {
      // nearly horizontal ray near the mid-plane: the two circles of the
      // section at this z, each a quadratic in the ray parameter
      R r0 = R(s.rtor) - vsqrt((radius - pt.z) * (radius + pt.z));
      const R invdirxy2 = R(1.) / (R(1) - dir.z * dir.z);
      const R b0 = (pt.x * dir.x + pt.y * dir.y) * invdirxy2;
      R c0 = (pt.x * pt.x + (pt.y - r0) * (pt.y + r0)) * invdirxy2;
      R delta = b0 * b0 - c0;
      if (delta > R(0)) {
        R v = -b0 - vsqrt(delta);
        put(x, nsol, v);
        if (v > -tol) nsol++;
        v = -b0 + vsqrt(delta);
        put(x, nsol, v);
        if (v > -tol) nsol++;
      }
      r0 = R(s.rtor) + vsqrt((radius - pt.z) * (radius + pt.z));
      c0 = (pt.x * pt.x + (pt.y - r0) * (pt.y + r0)) * invdirxy2;
      delta = b0 * b0 - c0;
      if (delta > R(0)) {
        R v = -b0 - vsqrt(delta);
        put(x, nsol, v);
        if (v > -tol) nsol++;
        v = -b0 + vsqrt(delta);
        put(x, nsol, v);
        if (v > -tol) nsol++;
      }
      if (nsol) sort4(x);
    } else {
      nsol = solve_quartic(a, b, c, d, x);
    }
    if (!nsol) return inf;
    const bool inner = vabs(radius - R(s.rmin)) < R(kTolerance);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (i >= nsol) break;
      if (x[i] < R(-10)) continue;
      // the point's projection onto the tube's centre circle, and the
      // outward direction from it
      V3<R> r0{pt.x + dir.x * x[i], pt.y + dir.y * x[i], pt.z + dir.z * x[i]};
      const V3<R> pn = r0;
      r0.z = R(0.);
      const R inv = R(1.) / vsqrt(r0.x * r0.x + r0.y * r0.y + r0.z * r0.z);  // Vector3D::Normalize
      r0 = V3<R>{r0.x * inv, r0.y * inv, r0.z * inv};
      r0 = V3<R>{r0.x * R(s.rtor), r0.y * R(s.rtor), r0.z * R(s.rtor)};
      const V3<R> norm{pn.x - r0.x, pn.y - r0.y, pn.z - r0.z};
      const R ndotd = norm.x * dir.x + norm.y * dir.y + norm.z * dir.z;
      if (inner ^ out) {
        if (ndotd < R(0)) continue;
      } else {
        if (ndotd > R(0)) continue;
      }
      if (s.has_phi) {
        bool ci = false, co = false;
        wedge<R>(s, r0, true, ci, co);  // ContainsWithBoundary
        if (co) continue;
      }
      sv = x[i];
      R eps = inf;
      R delta = sv * sv * sv * sv + a * sv * sv * sv + b * sv * sv + c * sv + d;
      R eps0 = -delta / (R(4.) * sv * sv * sv + R(3.) * a * sv * sv + R(2.) * b * sv + c);
      int ntry = 0;
      while (vabs(eps) > R(kTolerance)) {
        if (vabs(eps0) > R(100)) break;
        sv += eps0;
        if (vabs(sv + eps0) < R(kTolerance)) break;
        delta = sv * sv * sv * sv + a * sv * sv * sv + b * sv * sv + c * sv + d;
        eps = -delta / (R(4.) * sv * sv * sv + R(3.) * a * sv * sv + R(2.) * b * sv + c);
        if (vabs(eps) >= vabs(eps0)) break;
        ntry++;
        if (ntry > 100) break;
        eps0 = eps;
      }
      if (sv < -tol) continue;
      return vmax(R(0.), sv);
    }
    return inf;
  }

  // DistanceToIn (:540-658): from the bounding tube's entry point, a
  // phi-plane hit within kTolerance, else the nearer torus surface
  template <typename R>
  VGM_DEV static R distance_to_in(const Data &s, V3<R> point, V3<R> direction, R step_max)
  {
    const R inf = inf_length<R>();
    R distance = inf;
    R tubeDistance = inf;
    const bool inBounds = TubeShape::contains<R>(s.tube, point);
    if (!inBounds)
      tubeDistance = TubeShape::distance_to_in<R>(s.tube, point, direction, step_max);
    else
      tubeDistance = R(0.);
    if (inBounds) {
      if (inside<R>(s, point) == kInsideV) return R(-1.);
    } else if (tubeDistance == inf) {
      return distance;
    }
    const V3<R> lp{point.x + direction.x * tubeDistance, point.y + direction.y * tubeDistance,
                   point.z + direction.z * tubeDistance};
    if (s.has_phi) {
      R d1, d2;
      wedge_distances<R>(s, lp, direction, true, d1, d2);
      if (d1 != inf) {
        const R daxis = dist_sqr_to_torus_r<R>(s, lp, direction, d1);
        if (daxis >= R(s.rmin2) && daxis < R(s.rmax2)) {
          distance = d1;
          if (vabs(distance) < R(kTolerance)) return distance + tubeDistance;
        }
      }
      if (d2 != inf) {
        const R daxis = dist_sqr_to_torus_r<R>(s, lp, direction, d2);
        if (daxis >= R(s.rmin2) && daxis < R(s.rmax2)) {
          distance = vmin(d2, distance);
          if (vabs(distance) < R(kTolerance)) return distance + tubeDistance;
        }
      }
      distance = inf;
    }
    R dd = to_boundary<R>(s, lp, direction, R(s.rmax), false);
    if (s.rmin > 0.) dd = vmin(dd, to_boundary<R>(s, lp, direction, R(s.rmin), false));
    distance = vmin(distance, dd);
    distance += tubeDistance;
    if (vabs(distance) > inf) distance = inf;
    return distance;
  }

  // DistanceToOut (:201-292)
  template <typename R>
  VGM_DEV static R distance_to_out(const Data &s, V3<R> point, V3<R> dir, R /*step_max*/)
  {
    const R inf = inf_length<R>();
    const R distz = vabs(point.z) - R(s.rmax);
    bool done = distz > R(kHalfTolerance);
    const R rsq = point.x * point.x + point.y * point.y;
    const R outerExclRadius = R(s.rtor) + R(s.rmax) + R(kHalfTolerance);
    done |= rsq > outerExclRadius * outerExclRadius;
    const R innerExclRadius = R(s.rtor) - R(s.rmax) - R(kHalfTolerance);
    done |= rsq < innerExclRadius * innerExclRadius;
    if (done) return R(-1.);
    if (inside<R>(s, point) == kOutsideV) return R(-1.);
    const R dout = to_boundary<R>(s, point, dir, R(s.rmax), true);
    R din = inf;
    if (s.rmin > 0) din = to_boundary<R>(s, point, dir, R(s.rmin), true);
    R distance = vmin(dout, din);
    if (s.has_phi) {
      R distPhi1, distPhi2;
      wedge_distances<R>(s, point, dir, false, distPhi1, distPhi2);
      if (distPhi1 < distance) {
        const V3<R> ip{point.x + dir.x * distPhi1, point.y + dir.y * distPhi1, point.z + dir.z * distPhi1};
        bool ci = false, co = false;
        generic<R>(s, ip, false, false, ci, co);  // UnplacedContainsDisk
        if (!co) {
          const R diri = ip.x * R(s.w_a1x) + ip.y * R(s.w_a1y);
          if (diri >= R(0)) distance = distPhi1;
        }
      }
      if (distPhi2 < distance) {
        const V3<R> ip{point.x + dir.x * distPhi2, point.y + dir.y * distPhi2, point.z + dir.z * distPhi2};
        bool ci = false, co = false;
        generic<R>(s, ip, false, false, ci, co);
        if (!co) {
          const R diri2 = ip.x * R(s.w_a2x) + ip.y * R(s.w_a2y);
          if (diri2 >= R(0.)) distance = distPhi2;
        }
      }
    }
    if (distance >= inf) distance = R(-1.);
    return distance;
  }

  // SafetyToIn (:662-679)
  template <typename R>
  VGM_DEV static R safety_to_in(const Data &s, V3<R> p)
  {
    const R rxy = vsqrt(p.x * p.x + p.y * p.y);
    const R rad = vsqrt((rxy - R(s.rtor)) * (rxy - R(s.rtor)) + p.z * p.z);
    R safety = rad - R(s.rmax);
    if (s.rmin != 0.) safety = vmax(R(s.rmin) - rad, rad - R(s.rmax));
    if (s.has_phi && rxy != R(0.)) {
      // Wedge::SafetyToIn
      const R d1 = p.x * R(s.w_n1x) + p.y * R(s.w_n1y);
      const R d2 = p.x * R(s.w_n2x) + p.y * R(s.w_n2y);
      const R safetyPhi = s.w_lt_pi ? vmax(R(-1) * d1, R(-1) * d2) : vmin(R(-1) * d1, R(-1) * d2);
      safety = vmax(safetyPhi, safety);
    }
    return safety;
  }

  // SafetyToOut (:500-516)
  template <typename R>
  VGM_DEV static R safety_to_out(const Data &s, V3<R> p)
  {
    const R rxy = vsqrt(p.x * p.x + p.y * p.y);
    const R rad = vsqrt((rxy - R(s.rtor)) * (rxy - R(s.rtor)) + p.z * p.z);
    R safety = R(s.rmax) - rad;
    if (s.rmin != 0.) safety = vmin(rad - R(s.rmin), R(s.rmax) - rad);
    if (s.has_phi) safety = vmin(wedge_safety_to_out<R>(s.w_n1x, s.w_n1y, s.w_n2x, s.w_n2y, s.w_lt_pi, p), safety);
    return safety;
  }

  // UnplacedTorus2::Normal (source/UnplacedTorus2.cpp:201-285): the sum of the
  // normals of the surfaces within delta (radial) or 10 kTolerance (phi) of
  // the point; one surface's normal as is, a sum made a unit vector; no
  // surface leaves the normal at zero and reports invalid.  The phi-plane
  // normals (sin sphi, -cos sphi, 0) and (-sin ephi, cos ephi, 0) are the
  // negated wedge normals.
  VGM_DEV static bool normal(const Data &s, V3<double> p, V3<double> &norm)
  {
    int noSurfaces = 0;
    double distRMin = kInfLength, distSPhi = kInfLength, distEPhi = kInfLength;
    const double delta = s.normal_delta;
    const double dAngle = 10.0 * kTolerance;
    V3<double> nR{0., 0., 0.}, sum{0., 0., 0.};
    const double rho2 = p.x * p.x + p.y * p.y;
    const double rho = vsqrt(rho2);
    double pt2 = rho2 + p.z * p.z + s.rtor * (s.rtor - 2 * rho);
    pt2 = vmax(pt2, 0.);
    const double pt = vsqrt(pt2);
    const double distRMax = vabs(pt - s.rmax);
    if (s.rmin != 0.) distRMin = vabs(pt - s.rmin);
    if (rho > delta && pt != 0.) {
      const double redFactor = (rho - s.rtor) / rho;
      const double ipt = 1. / pt;
      nR = V3<double>{p.x * redFactor * ipt, p.y * redFactor * ipt, p.z * ipt};
    }
    if (s.has_phi && rho != 0.) {
      double pPhi = atan2(p.y, p.x);
      if (pPhi < s.sphi - delta)
        pPhi += kTwoPi;
      else if (pPhi > s.sphi + s.dphi + delta)
        pPhi -= kTwoPi;
      distSPhi = vabs(pPhi - s.sphi);
      distEPhi = vabs(pPhi - s.sphi - s.dphi);
    }
    if (distRMax <= delta) {
      noSurfaces++;
      sum = V3<double>{sum.x + nR.x, sum.y + nR.y, sum.z + nR.z};
    } else if (s.rmin != 0. && distRMin <= delta) {
      noSurfaces++;
      sum = V3<double>{sum.x - nR.x, sum.y - nR.y, sum.z - nR.z};
    }
    if (s.has_phi && (s.rmin - delta <= pt) && (pt <= (s.rmax + delta))) {
      if (distSPhi <= dAngle) {
        noSurfaces++;
        sum = V3<double>{sum.x + -s.w_n1x, sum.y + -s.w_n1y, sum.z + 0.};
      }
      if (distEPhi <= dAngle) {
        noSurfaces++;
        sum = V3<double>{sum.x + -s.w_n2x, sum.y + -s.w_n2y, sum.z + 0.};
      }
    }
    if (noSurfaces == 0) {
      norm = V3<double>{0., 0., 0.};
      return false;
    }
    if (noSurfaces == 1) {
      norm = sum;
    } else {
      const double m = vsqrt(sum.x * sum.x + sum.y * sum.y + sum.z * sum.z + DBL_MIN);  // Vector3D::Unit
      norm = V3<double>{sum.x / m, sum.y / m, sum.z / m};
    }
    return true;
  }
};

}  // namespace vgm
