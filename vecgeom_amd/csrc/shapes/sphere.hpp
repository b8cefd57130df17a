// Sphere kernels: VecGeom/volumes/kernel/SphereImplementation.h with
// SphereUtilities.h, the theta cone of ThetaCone_Evolution.h and the phi
// wedge of Wedge_Evolution.h.  Scalar-backend semantics; the theta cone's
// branches on its angles are wave-uniform.  The reference's quirks are
// kept: the wedge and the theta cone use the constructor's raw angles, and a
// theta range below the equator can give a negative DistanceToOut (the
// second cone root is not rejected when negative).
#pragma once
#include "../vgm_math.hpp"
#include "sphere_data.h"

namespace vgm {

struct SphereShape {
  using Data = SphereData;
  static constexpr bool kStage = false;  // parameter block staged in LDS
  // waves per SIMD the transport kernel asks the compiler for (1: no request)
  static constexpr int kWavesPerEU = 1;
  // safety is the distance to the surface (1-Lipschitz): carried bounds are valid
  static constexpr bool kSafetyIsDistance = false;
  // the theta cone's safeties are not bounded by its distances everywhere:
  // the transport takes the safety-first FP64 step
  static constexpr bool kSafetyBelowDistance = false;
  // mixed precision: Inside/Contains classified in float away from the surface
  static constexpr bool kFloatClassify = false;
  // mixed precision: the float safety less the guard band bounds the FP64 safety
  static constexpr bool kFloatSafety = false;
  // mixed precision: distance_to_in_mixed / distance_to_out_mixed (float selects
  // the surface, FP64 evaluates it)
  static constexpr bool kMixedDistance = false;

  template <typename R>
  VGM_DEV static R plus_tol_sq(R x)  // MakePlusTolerantSquare<true>
  {
    return x * (x + R(2.0 * kTolerance));
  }
  template <typename R>
  VGM_DEV static R minus_tol_sq(R x)  // MakeMinusTolerantSquare<true>
  {
    return vabs(x * (x - R(2.0 * kTolerance)));
  }
  template <typename R>
  VGM_DEV static R mag2(V3<R> p)
  {
    return p.x * p.x + p.y * p.y + p.z * p.z;
  }
  template <typename R>
  VGM_DEV static R dot(V3<R> a, V3<R> b)
  {
    return a.x * b.x + a.y * b.y + a.z * b.z;
  }
  template <typename R>
  VGM_DEV static V3<R> ray(V3<R> p, V3<R> d, R t)  // point + direction * t
  {
    return V3<R>{p.x + d.x * t, p.y + d.y * t, p.z + d.z * t};
  }

  // ---- phi wedge (Wedge_Evolution.h) -------------------------------------
  template <typename R>
  VGM_DEV static bool wedge_on_surface(double ax, double ay, double nx, double ny, V3<R> p)
  {
    if (!(R(ax) * p.x + R(ay) * p.y >= R(0))) return false;
    return vabs(R(nx) * p.x + R(ny) * p.y) < R(kTolerance);
  }
  // GenericKernelForContainsAndInside (:296-338)
  template <typename R>
  VGM_DEV static void wedge_generic(const Data &s, V3<R> p, bool forInside, bool &cin, bool &cout)
  {
    const R startCheck = (-p.x * R(s.w_a1y) + p.y * R(s.w_a1x));
    const R endCheck = (-R(s.w_a2x) * p.y + R(s.w_a2y) * p.x);
    cout = startCheck < R(0);
    if (s.w_dphi < kPi)
      cout |= endCheck < R(0);
    else
      cout &= endCheck < R(0);
    if (forInside) {
      const bool onSurface = wedge_on_surface<R>(s.w_a1x, s.w_a1y, s.w_n1x, s.w_n1y, p) ||
                             wedge_on_surface<R>(s.w_a2x, s.w_a2y, s.w_n2x, s.w_n2y, p);
      cout &= !onSurface;
      cin = !onSurface && !cout;
    }
  }
  template <typename R>
  VGM_DEV static bool wedge_contains(const Data &s, V3<R> p)
  {
    bool ci = false, co = false;
    wedge_generic<R>(s, p, false, ci, co);
    return !co;
  }
  template <typename R>
  VGM_DEV static R wedge_safety_to_in(const Data &s, V3<R> p)
  {
    const R d1 = p.x * R(s.w_n1x) + p.y * R(s.w_n1y);
    const R d2 = p.x * R(s.w_n2x) + p.y * R(s.w_n2y);
    return (s.w_dphi < kPi) ? vmax(R(-1) * d1, R(-1) * d2) : vmin(R(-1) * d1, R(-1) * d2);
  }
  template <typename R>
  VGM_DEV static R wedge_safety_to_out(const Data &s, V3<R> p)
  {
    const R d1 = p.x * R(s.w_n1x) + p.y * R(s.w_n1y);
    const R d2 = p.x * R(s.w_n2x) + p.y * R(s.w_n2y);
    return (s.w_dphi < kPi) ? vmin(d1, d2) : vmax(d1, d2);
  }
  // DistanceToIn / DistanceToOut (:398-455): the plane hits ahead of the
  // point, entering (in) or leaving (out) the wedge
  template <typename R>
  VGM_DEV static void wedge_distances(const Data &s, V3<R> p, V3<R> d, bool in, R &w1, R &w2)
  {
    const R c1 = d.x * R(s.w_n1x) + d.y * R(s.w_n1y);
    const R c2 = d.x * R(s.w_n2x) + d.y * R(s.w_n2y);
    w1 = inf_length<R>();
    w2 = inf_length<R>();
    if (in ? c1 > R(0) : c1 < R(0)) {
      const R t = -(p.x * R(s.w_n1x) + p.y * R(s.w_n1y)) / c1;
      if (t > R(0)) w1 = t;
    }
    if (in ? c2 > R(0) : c2 < R(0)) {
      const R t = -(p.x * R(s.w_n2x) + p.y * R(s.w_n2y)) / c2;
      if (t > R(0)) w2 = t;
    }
  }

  // ---- theta cone (ThetaCone_Evolution.h) --------------------------------
  template <typename R>
  VGM_DEV static V3<R> tc_normal1(const Data &s, V3<R> p)  // GetNormal1
  {
    const V3<R> n{R(2.) * p.x, R(2.) * p.y, R(-2.) * R(s.tan_s2) * p.z};
    return (s.tc_s <= kPi / 2.) ? V3<R>{-n.x, -n.y, -n.z} : n;
  }
  template <typename R>
  VGM_DEV static V3<R> tc_normal2(const Data &s, V3<R> p)  // GetNormal2
  {
    const V3<R> n{R(2) * p.x, R(2) * p.y, R(-2) * R(s.tan_e2) * p.z};
    return (s.tc_e <= kPi / 2.) ? n : V3<R>{-n.x, -n.y, -n.z};
  }
  template <typename R>
  VGM_DEV static bool tc_on_surface(const Data &s, V3<R> p, bool start)  // IsOnSurfaceGeneric
  {
    const R rhs = start ? vabs(R(s.tan_s) * p.z) : vabs(R(s.tan_e) * p.z);
    const R rho2 = p.x * p.x + p.y * p.y;
    return rho2 >= minus_tol_sq<R>(rhs) && rho2 <= plus_tol_sq<R>(rhs);
  }
  template <typename R>
  VGM_DEV static bool tc_on_surface_moving_out(const Data &s, V3<R> p, V3<R> d, bool start, bool movingOut)
  {
    if (!tc_on_surface<R>(s, p, start)) return false;
    V3<R> n;
    if (start) {
      n = V3<R>{p.x, p.y, -R(s.tan_s2) * p.z};
      if (s.tc_s <= 0.5 * kPi) n = V3<R>{-n.x, -n.y, -n.z};
    } else {
      n = V3<R>{p.x, p.y, -R(s.tan_e2) * p.z};
      if (!(s.tc_e <= 0.5 * kPi)) n = V3<R>{-n.x, -n.y, -n.z};
    }
    const R dn = dot<R>(d, n);
    return movingOut ? dn > R(0) : dn < R(0);
  }
  template <typename R>
  VGM_DEV static R tc_line_distance(double slope, R x, R y)  // DistanceToLine
  {
    const R dist = (y - R(slope) * x) / vsqrt(R(1.) + R(slope) * R(slope));
    return vabs(dist);
  }

  template <typename R>
  VGM_DEV static bool tc_completely_inside(const Data &s, V3<R> p)
  {
    const double kAngTol = kTolerance, half = 0.5 * kTolerance;
    const double sT = s.tc_s, eT = s.tc_e;
    const R rho = vsqrt(mag2<R>(p) - (p.z * p.z));
    const R cone1Radius = vabs(p.z * R(s.tan_s));
    const R cone2Radius = vabs(p.z * R(s.tan_e));
    const bool onZ = p.z != R(0) && p.x == R(0) && p.y == R(0);
    const bool onXY = p.z == R(0) && (p.x != R(0) || p.y != R(0));
    bool ci = (onZ && ((sT == 0. && eT == kPi) || (p.z > R(0) && sT == 0.) || (p.z < R(0) && eT == kPi)));
    ci |= (!ci && (onXY && (sT < 0.5 * kPi && eT > 0.5 * kPi && (0.5 * kPi - sT) > kAngTol &&
                            (eT - 0.5 * kPi) > kTolerance)));
    const R tol20 = R(2 * kAngTol * 10.);
    if (sT < 0.5 * kPi + half) {
      if (eT < 0.5 * kPi + half) {
        if (sT < eT) {
          const R tolAngMin = cone1Radius + tol20, tolAngMax = cone2Radius - tol20;
          ci |= (!ci && (((rho <= tolAngMax) && (rho >= tolAngMin) && (p.z > R(0)) && (sT != 0.)) ||
                         ((rho <= tolAngMax) && (sT == 0.) && (p.z > R(0)))));
        }
      }
      if (eT > 0.5 * kPi + half) {
        if (sT < eT) {
          const R tolAngMin = cone1Radius + tol20, tolAngMax = cone2Radius + tol20;
          ci |= (!ci && (((rho >= tolAngMin) && (p.z > R(0))) || ((rho >= tolAngMax) && (p.z < R(0)))));
        }
      }
      if (eT >= 0.5 * kPi - half && eT <= 0.5 * kPi + half) ci &= !(vabs(p.z) < R(half));
    }
    if (eT > 0.5 * kPi + half) {
      if (sT >= 0.5 * kPi - half && sT <= 0.5 * kPi + half) ci &= !(vabs(p.z) < R(half));
      if (sT > 0.5 * kPi + half) {
        if (sT < eT) {
          const R tolAngMin = cone1Radius - tol20, tolAngMax = cone2Radius + tol20;
          ci |= (!ci && (((rho <= tolAngMin) && (rho >= tolAngMax) && (p.z < R(0)) && (eT != kPi)) ||
                         ((rho <= tolAngMin) && (p.z < R(0)) && (eT == kPi))));
        }
      }
    }
    return ci;
  }

  template <typename R>
  VGM_DEV static bool tc_completely_outside(const Data &s, V3<R> p)
  {
    const double kAngTol = kTolerance, half = 0.5 * kTolerance;
    const double sT = s.tc_s, eT = s.tc_e;
    const R rho = vsqrt(p.x * p.x + p.y * p.y);
    const R cone1Radius = vabs(p.z * R(s.tan_s));
    const R cone2Radius = vabs(p.z * R(s.tan_e));
    const bool onZ = p.z != R(0) && p.x == R(0) && p.y == R(0);
    const bool onXY = p.z == R(0) && (p.x != R(0) || p.y != R(0));
    bool co = (onZ && ((sT != 0. && eT != kPi) || (p.z > R(0) && sT != 0.) || (p.z < R(0) && eT != kPi)));
    co |= (!co && (onXY && (((sT < 0.5 * kPi) && (eT < 0.5 * kPi) && ((0.5 * kPi - sT) > kAngTol) &&
                             ((0.5 * kPi - eT) > kTolerance)) ||
                            ((sT > 0.5 * kPi && eT > 0.5 * kPi) && ((sT - 0.5 * kPi) > kAngTol) &&
                             ((eT - 0.5 * kPi) > kTolerance)))));
    const R tol20 = R(2 * kAngTol * 10.);
    if (sT < 0.5 * kPi + half) {
      if (eT < 0.5 * kPi + half) {
        if (sT < eT) {
          const R tolAngMin2 = cone1Radius - tol20, tolAngMax2 = cone2Radius + tol20;
          co |= (!co && ((rho < tolAngMin2) || (rho > tolAngMax2) || (p.z < R(0))));
        }
      }
      if (eT > 0.5 * kPi + half) {
        if (sT < eT) {
          const R tolAngMin2 = cone1Radius - tol20, tolAngMax2 = cone2Radius - tol20;
          co |= (!co && (((rho < tolAngMin2) && (p.z > R(0))) || ((rho < tolAngMax2) && (p.z < R(0)))));
        }
      }
      if (eT >= 0.5 * kPi - half && eT <= 0.5 * kPi + half) co &= !(vabs(p.z) < R(half));
    }
    if (eT > 0.5 * kPi + half) {
      if (sT >= 0.5 * kPi - half && sT <= 0.5 * kPi + half) co &= !(vabs(p.z) < R(half));
      if (sT > 0.5 * kPi + half) {
        if (sT < eT) {
          const R tolAngMin2 = cone1Radius + tol20, tolAngMax2 = cone2Radius - tol20;
          co |= (!co && ((rho < tolAngMax2) || (rho > tolAngMin2) || (p.z > R(0))));
        }
      }
    }
    return co;
  }

  template <typename R>
  VGM_DEV static bool tc_contains(const Data &s, V3<R> p)
  {
    return !tc_completely_outside<R>(s, p);
  }

  template <typename R>
  VGM_DEV static R tc_safety_to_in(const Data &s, V3<R> p)
  {
    const double half = 0.5 * kTolerance, sT = s.tc_s, eT = s.tc_e;
    const R pointRad = vsqrt(p.x * p.x + p.y * p.y);
    const R sfTh1 = tc_line_distance<R>(s.slope1, pointRad, p.z);
    const R sfTh2 = tc_line_distance<R>(s.slope2, pointRad, p.z);
    R safeTheta = vmin(sfTh1, sfTh2);
    if (tc_contains<R>(s, p)) return R(0);
    if (sT < kPi / 2 + half) {
      if (eT < kPi / 2 + half) {
        if (sT < eT && p.z < R(0)) safeTheta = sfTh2;
      }
      if (eT > kPi / 2 + half) {
        if (sT < eT) {
          if (p.z > R(0)) safeTheta = sfTh1;
          if (p.z < R(0)) safeTheta = sfTh2;
        }
      }
    }
    if (eT > kPi / 2 + half) {
      if (sT > kPi / 2 + half) {
        if (sT < eT && p.z > R(0)) safeTheta = sfTh1;
      }
    }
    return safeTheta;
  }

  template <typename R>
  VGM_DEV static R tc_safety_to_out(const Data &s, V3<R> p)
  {
    const double half = 0.5 * kTolerance, sT = s.tc_s, eT = s.tc_e;
    const R pointRad = vsqrt(p.x * p.x + p.y * p.y);
    const R bisectorRad = vabs(p.z * R(s.tan_bis));
    bool condition = false;
    const R sfTh1 = tc_line_distance<R>(s.slope1, pointRad, p.z);
    const R sfTh2 = tc_line_distance<R>(s.slope2, pointRad, p.z);
    if (sT < kPi / 2 + half) {
      if (eT < kPi / 2 + half) {
        if (sT < eT) condition = (pointRad < bisectorRad) && (sT != 0.);
      }
      if (eT > kPi / 2 + half) {
        if (sT < eT) condition = sfTh1 < sfTh2;
      }
    }
    if (eT > kPi / 2 + half) {
      if (sT > kPi / 2 + half) {
        if (sT < eT) condition = !((pointRad < bisectorRad) && (eT != kPi));
      }
    }
    return condition ? sfTh1 : sfTh2;
  }

  // the root cleanup of ThetaCone::DistanceToIn: |root| < 3 kTolerance snaps to 0
  template <typename R>
  VGM_DEV static void tc_snap(R &root, bool &done)
  {
    done |= (vabs(root) < R(3.) * R(kTolerance));
    if (vabs(root) < R(3.) * R(kTolerance)) root = R(0);
  }

  template <typename R>
  VGM_DEV static void tc_distance_to_in(const Data &s, V3<R> p, V3<R> d, R &dist1, R &dist2, bool &int1, bool &int2)
  {
    const double half = 0.5 * kTolerance, sT = s.tc_s, eT = s.tc_e;
    const R inf = inf_length<R>();
    bool done = false;
    R firstRoot = inf, secondRoot = inf;
    const R pDotV2d = p.x * d.x + p.y * d.y;
    const R rho2 = p.x * p.x + p.y * p.y;
    const R dirRho2 = d.x * d.x + d.y * d.y;
    const R b = pDotV2d - p.z * d.z * R(s.tan_s2);
    const R a = dirRho2 - d.z * d.z * R(s.tan_s2);
    const R c = rho2 - p.z * p.z * R(s.tan_s2);
    const R d2 = b * b - a * c;
    const R aInv = R(1.) / nonzero(a);
    if (d2 > R(0)) firstRoot = (-b + vsqrt(vabs(d2))) * aInv;
    done |= (vabs(firstRoot) < R(3.) * R(kTolerance));
    if (vabs(firstRoot) < R(3.) * R(kTolerance)) firstRoot = R(0);
    if (!done && (firstRoot < R(0))) firstRoot = inf;
    const R b2 = pDotV2d - p.z * d.z * R(s.tan_e2);
    const R a2 = dirRho2 - d.z * d.z * R(s.tan_e2);
    const R c2 = rho2 - p.z * p.z * R(s.tan_e2);
    const R d22 = b2 * b2 - a2 * c2;
    const R a2Inv = R(1.) / nonzero(a2);
    if (d22 > R(0)) secondRoot = (-b2 - vsqrt(vabs(d22))) * a2Inv;
    if (!done && (vabs(secondRoot) < R(3.) * R(kTolerance))) secondRoot = R(0);
    done |= (vabs(secondRoot) < R(3.) * R(kTolerance));
    if (!done && (secondRoot < R(0))) secondRoot = inf;

    if (sT < 0.5 * kPi + half) {
      if (eT < 0.5 * kPi + half) {
        if (sT < eT) {
          dist1 = firstRoot;
          dist2 = secondRoot;
          const R z1 = (p.z + dist1 * d.z), z2 = (p.z + dist2 * d.z);
          int1 = ((d2 > R(0)) && (z1 > R(0)));
          int2 = ((d22 > R(0)) && (z2 > R(0)));
        }
      }
      if (eT >= 0.5 * kPi - half && eT <= 0.5 * kPi + half) {
        if (d.z > R(0)) dist2 = -p.z / d.z;
        const R z2 = (p.z + dist2 * d.z);
        int2 = ((dist2 != inf) && (vabs(z2) < R(half)));
      }
      if (eT > 0.5 * kPi + half) {
        if (sT < eT) {
          dist1 = firstRoot;
          if (d22 > R(0)) secondRoot = (-b2 + vsqrt(vabs(d22))) * a2Inv;
          done = false;
          tc_snap<R>(secondRoot, done);
          if (!done && (secondRoot < R(0))) secondRoot = inf;
          dist2 = secondRoot;
          const R z1 = (p.z + dist1 * d.z), z2 = (p.z + dist2 * d.z);
          int1 = ((d2 > R(0)) && (dist1 != inf) && (z1 > R(0)));
          int2 = ((d22 > R(0)) && (dist2 != inf) && (z2 < R(0)));
        }
      }
    }
    if (sT >= 0.5 * kPi - half) {
      if (eT > 0.5 * kPi + half) {
        if (sT < eT) {
          if (d2 > R(0)) firstRoot = (-b - vsqrt(vabs(d2))) * aInv;
          done = false;
          tc_snap<R>(firstRoot, done);
          if (!done && (firstRoot < R(0))) firstRoot = inf;
          dist1 = firstRoot;
          if (d22 > R(0)) secondRoot = (-b2 + vsqrt(vabs(d22))) * a2Inv;
          done = false;
          tc_snap<R>(secondRoot, done);
          if (!done && (secondRoot < R(0))) secondRoot = inf;
          dist2 = secondRoot;
          const R z1 = (p.z + dist1 * d.z), z2 = (p.z + dist2 * d.z);
          int1 = ((d2 > R(0)) && (dist1 != inf) && (z1 < R(0)));
          int2 = ((d22 > R(0)) && (dist2 != inf) && (z2 < R(0)));
        }
      }
    }
    if (sT >= 0.5 * kPi - half && sT <= 0.5 * kPi + half) {
      if (d.z < R(0)) dist1 = -p.z / d.z;
      const R z1 = (p.z + dist1 * d.z);
      int1 = ((dist1 != inf) && (vabs(z1) < R(half)));
    }
  }

  template <typename R>
  VGM_DEV static void tc_distance_to_out(const Data &s, V3<R> p, V3<R> d, R &dist1, R &dist2, bool &int1,
                                         bool &int2)
  {
    const double half = 0.5 * kTolerance, sT = s.tc_s, eT = s.tc_e;
    const R inf = inf_length<R>();
    R firstRoot = inf, secondRoot = inf;
    const R pDotV2d = p.x * d.x + p.y * d.y;
    const R rho2 = p.x * p.x + p.y * p.y;
    const R b = pDotV2d - p.z * d.z * R(s.tan_s2);
    const R a = d.x * d.x + d.y * d.y - d.z * d.z * R(s.tan_s2);
    const R c = rho2 - p.z * p.z * R(s.tan_s2);
    R d2 = b * b - a * c;
    if (d2 < R(0) && vabs(d2) < R(kHalfTolerance)) d2 = R(0);
    if ((d2 >= R(0)) && b >= R(0) && a != R(0)) firstRoot = ((-b - vsqrt(vabs(d2))) / nonzero(a));
    if ((d2 >= R(0)) && b < R(0)) firstRoot = ((c) / nonzero(-b + vsqrt(vabs(d2))));
    if (firstRoot < R(0)) firstRoot = inf;
    const R b2 = p.x * d.x + p.y * d.y - p.z * d.z * R(s.tan_e2);
    const R a2 = d.x * d.x + d.y * d.y - d.z * d.z * R(s.tan_e2);
    const R c2 = p.x * p.x + p.y * p.y - p.z * p.z * R(s.tan_e2);
    R d22 = (b2 * b2) - (a2 * c2);
    if (d22 < R(0) && vabs(d22) < R(kHalfTolerance)) d22 = R(0);
    if ((d22 >= R(0)) && b2 >= R(0)) secondRoot = ((c2) / nonzero(-b2 - vsqrt(vabs(d22))));
    if ((d22 >= R(0)) && b2 < R(0) && a2 != R(0)) secondRoot = ((-b2 + vsqrt(vabs(d22))) / nonzero(a2));
    if (secondRoot < R(0) && vabs(secondRoot) > R(kTolerance)) secondRoot = inf;
    if (vabs(secondRoot) < R(kTolerance)) secondRoot = R(0);

    if (sT < kPi / 2 + half) {
      if (eT < kPi / 2 + half) {
        if (sT < eT) {
          dist1 = firstRoot;
          dist2 = secondRoot;
          const R z1 = (p.z + dist1 * d.z), z2 = (p.z + dist2 * d.z);
          int1 = ((d2 > R(0)) && (dist1 != inf) && ((z1) > R(-kHalfTolerance)));
          int2 = ((d22 > R(0)) && (dist2 != inf) && ((z2) > R(-kHalfTolerance)));
          const R dirRho2 = d.x * d.x + d.y * d.y;
          R zs = inf;
          if (sT != 0.) zs = dirRho2 / R(s.tan_s);
          R ze = inf;
          if (eT != 0.) ze = dirRho2 / R(s.tan_e);
          const bool cond = (p.x == R(0) && p.y == R(0) && p.z == R(0) && d.z < zs && d.z < ze);
          if (cond) {
            dist1 = R(0);
            dist2 = R(0);
          }
          int1 |= cond;
          int2 |= cond;
        }
      }
      if (eT >= kPi / 2 - half && eT <= kPi / 2 + half) {
        dist1 = firstRoot;
        dist2 = inf;
        if (d.z < R(0)) dist2 = R(-1.) * p.z / d.z;
        const R z2 = (p.z + dist2 * d.z), z1 = (p.z + dist1 * d.z);
        int2 = ((dist2 != inf) && (vabs(z2) < R(kHalfTolerance)) && !(d.z == R(0)));
        int1 = ((d2 >= R(0)) && (dist1 != inf) && (vabs(z1) < R(kHalfTolerance)) && !(d.z == R(0)));
      }
      if (eT > kPi / 2 + half) {
        if (sT < eT) {
          dist1 = firstRoot;
          if ((d22 >= R(0)) && b2 > R(0) && a2 != R(0)) secondRoot = ((-b2 - vsqrt(vabs(d22))) / nonzero(a2));
          if ((d22 >= R(0)) && b2 <= R(0)) secondRoot = ((c2) / nonzero(-b2 + vsqrt(vabs(d22))));
          if (secondRoot < R(0)) secondRoot = inf;
          dist2 = secondRoot;
          const R z1 = (p.z + dist1 * d.z), z2 = (p.z + dist2 * d.z);
          int1 = ((d2 >= R(0)) && (dist1 != inf) && ((z1) > R(-kHalfTolerance)));
          int2 = ((d22 >= R(0)) && (dist2 != inf) && ((z2) < R(kHalfTolerance)));
        }
      }
    }
    if (eT > kPi / 2 + half) {
      if (sT < eT) {
        secondRoot = inf;
        if ((d22 >= R(0)) && b2 > R(0) && a2 != R(0)) secondRoot = ((-b2 - vsqrt(vabs(d22))) / nonzero(a2));
        if ((d22 >= R(0)) && b2 <= R(0)) secondRoot = ((c2) / nonzero(-b2 + vsqrt(vabs(d22))));
        dist2 = secondRoot;
        if (sT > kPi / 2 + half) {
          if ((d2 >= R(0)) && b > R(0)) firstRoot = ((c) / nonzero(-b - vsqrt(vabs(d2))));
          if ((d2 >= R(0)) && b <= R(0) && a != R(0)) firstRoot = ((-b + vsqrt(vabs(d2))) / nonzero(a));
          if (firstRoot < R(0)) firstRoot = inf;
          dist1 = firstRoot;
          const R z1 = (p.z + dist1 * d.z);
          int1 = ((d2 > R(0)) && (dist1 != inf) && ((z1) < R(kHalfTolerance)));
          const R z2 = (p.z + dist2 * d.z);
          int2 = ((d22 > R(0)) && (dist2 != inf) && ((z2) < R(kHalfTolerance)));
          const R dirRho2 = d.x * d.x + d.y * d.y;
          R zs = -inf;
          if (s.tan_s != 0.) zs = -dirRho2 / R(s.tan_s);
          R ze = -inf;
          if (s.tan_e != 0.) ze = -dirRho2 / R(s.tan_e);
          const bool cond = (p.x == R(0) && p.y == R(0) && p.z == R(0) && d.z > zs && d.z > ze);
          if (cond) {
            dist1 = R(0);
            dist2 = R(0);
          }
          int1 |= cond;
          int2 |= cond;
        }
      }
    }
    if (sT >= kPi / 2 - half && sT <= kPi / 2 + half) {
      dist2 = secondRoot;
      dist1 = inf;
      if (d.z > R(0)) dist1 = R(-1.) * p.z / nonzero(d.z);
      const R z1 = (p.z + dist1 * d.z), z2 = (p.z + dist2 * d.z);
      int1 = ((dist1 != inf) && (vabs(z1) < R(kHalfTolerance)) && (d.z != R(0)));
      int2 = ((d22 >= R(0)) && (dist2 != inf) && (vabs(z2) < R(kHalfTolerance)) && (d.z != R(0)));
    }
  }

  // ---- sphere ---------------------------------------------------------------
  // GenericKernelForContainsAndInside (SphereImplementation.h:78-120)
  template <typename R>
  VGM_DEV static void generic(const Data &s, V3<R> p, bool forInside, bool &cin, bool &cout)
  {
    const R rad2 = mag2<R>(p);
    const R rmax = R(s.rmax), rmin = R(s.rmin);
    if (s.rmin != 0.)
      cin = rad2 <= minus_tol_sq<R>(rmax) && rad2 >= plus_tol_sq<R>(rmin);
    else
      cin = rad2 <= minus_tol_sq<R>(rmax);
    if (s.rmin != 0.)
      cout = rad2 >= plus_tol_sq<R>(rmax) || rad2 <= minus_tol_sq<R>(rmin);
    else
      cout = rad2 >= plus_tol_sq<R>(rmax);
    if (!s.full_phi) {
      bool ci = false, co = false;
      wedge_generic<R>(s, p, forInside, ci, co);
      cout |= co;
      if (forInside) cin &= ci;
    }
    if (!s.full_theta) {
      const bool ci = forInside ? tc_completely_inside<R>(s, p) : false;
      const bool co = tc_completely_outside<R>(s, p);
      cout |= co;
      if (forInside) cin &= ci;
    }
  }

  template <typename R>
  VGM_DEV static bool contains(const Data &s, V3<R> p)
  {
    bool ci = false, co = false;
    generic<R>(s, p, false, ci, co);
    return !co;
  }

  template <typename R>
  VGM_DEV static int inside(const Data &s, V3<R> p)
  {
    bool ci = false, co = false;
    generic<R>(s, p, true, ci, co);
    int in = kSurfaceV;
    if (co) in = kOutsideV;
    if (ci) in = kInsideV;
    return in;
  }

  // SphereUtilities::IsCompletelyInside / IsCompletelyOutside (:112-140)
  template <typename R>
  VGM_DEV static bool util_completely(const Data &s, V3<R> p, bool inside_test)
  {
    const R rad = vsqrt(mag2<R>(p));
    bool inPhi = false, outPhi = false;
    wedge_generic<R>(s, p, true, inPhi, outPhi);
    if (inside_test) {
      const bool insideRadiusRange = (rad < (R(s.rmax) - R(kTolerance))) && (rad > (R(s.rmin) + R(kTolerance)));
      return insideRadiusRange && inPhi && tc_completely_inside<R>(s, p);
    }
    const bool outsideRadiusRange = (rad > (R(s.rmax) + R(kTolerance))) || (rad < (R(s.rmin) - R(kTolerance)));
    return outsideRadiusRange || outPhi || tc_completely_outside<R>(s, p);
  }

  template <typename R>
  VGM_DEV static bool on_radius(R r, V3<R> p)
  {
    const R m2 = mag2<R>(p);
    return m2 <= plus_tol_sq<R>(r) && m2 >= minus_tol_sq<R>(r);
  }

  // SphereUtilities::IsPointOnSurfaceAndMovingOut (:164-199)
  template <typename R>
  VGM_DEV static bool on_surface_moving_out(const Data &s, V3<R> p, V3<R> d, bool movingOut)
  {
    const R kSqrtTolerance = R(3.1622777e-5);  // VecGeom/base/Math.h:29
    const R dp = dot<R>(d, p), dmp = dot<R>(d, V3<R>{-p.x, -p.y, -p.z});
    const bool tempOuterRad = on_radius<R>(R(s.rmax), p) && (movingOut ? dp > kSqrtTolerance : dp < -kSqrtTolerance);
    bool tempInnerRad = false, tempStartPhi = false, tempEndPhi = false, tempStartTheta = false,
         tempEndTheta = false;
    if (s.rmin != 0.)
      tempInnerRad = on_radius<R>(R(s.rmin), p) && (movingOut ? dmp > kSqrtTolerance : dmp < -kSqrtTolerance);
    if (s.dphi < (kTwoPi - kHalfTolerance)) {
      const R n1 = dot<R>(d, V3<R>{-R(s.w_n1x), -R(s.w_n1y), -R(0)});
      const R n2 = dot<R>(d, V3<R>{-R(s.w_n2x), -R(s.w_n2y), -R(0)});
      const R lim = R(0.005 * kHalfTolerance);
      tempStartPhi = wedge_on_surface<R>(s.w_a1x, s.w_a1y, s.w_n1x, s.w_n1y, p) && (movingOut ? n1 > lim : n1 < lim);
      tempEndPhi = wedge_on_surface<R>(s.w_a2x, s.w_a2y, s.w_n2x, s.w_n2y, p) && (movingOut ? n2 > lim : n2 < lim);
    }
    if (s.dtheta < (kPi - kHalfTolerance)) {
      tempStartTheta = tc_on_surface_moving_out<R>(s, p, d, true, movingOut);
      tempEndTheta = tc_on_surface_moving_out<R>(s, p, d, false, movingOut);
    }
    const R m2 = mag2<R>(p);
    const R rmin2 = R(s.rmin) * R(s.rmin), rmax2 = R(s.rmax) * R(s.rmax);
    return ((tempOuterRad || tempInnerRad) && wedge_contains<R>(s, p) && tc_contains<R>(s, p)) ||
           ((tempStartPhi || tempEndPhi) && (m2 >= rmin2) && (m2 <= rmax2) && tc_contains<R>(s, p)) ||
           ((tempStartTheta || tempEndTheta) && (m2 >= rmin2) && (m2 <= rmax2) && wedge_contains<R>(s, p));
  }

  // GetMinDistFromPhi<DistToIn> (:285-330)
  template <typename R>
  VGM_DEV static void min_dist_from_phi(const Data &s, V3<R> p, V3<R> d, bool toIn, R &distance)
  {
    R distPhi1, distPhi2;
    wedge_distances<R>(s, p, d, toIn, distPhi1, distPhi2);
    auto check = [&](R dist) {
      const V3<R> t = ray<R>(p, d, dist);
      const R rad2 = mag2<R>(t);
      const bool tempCond = ((dist == distPhi1) && wedge_on_surface<R>(s.w_a1x, s.w_a1y, s.w_n1x, s.w_n1y, t)) ||
                            ((dist == distPhi2) && wedge_on_surface<R>(s.w_a2x, s.w_a2y, s.w_n2x, s.w_n2y, t));
      return tempCond && (rad2 > R(s.rmin) * R(s.rmin)) && (rad2 < R(s.rmax) * R(s.rmax)) && tc_contains<R>(s, t);
    };
    R dist = vmin(distPhi1, distPhi2);
    const bool containsCond1 = check(dist);
    if (containsCond1) distance = vmin(dist, distance);
    dist = vmax(distPhi1, distPhi2);
    const bool containsCond2 = check(dist);
    if (!containsCond1 && containsCond2) distance = vmin(dist, distance);
  }

  // DistanceToIn (:128-222)
  template <typename R>
  VGM_DEV static R distance_to_in(const Data &s, V3<R> point, V3<R> direction, R /*step_max*/)
  {
    const R inf = inf_length<R>();
    R distance = inf;
    const R rad2 = mag2<R>(point);
    const R pDotV3d = dot<R>(point, direction);
    R c = rad2 - R(s.rmax) * R(s.rmax);
    if (util_completely<R>(s, point, true)) return R(-1.0);
    if (on_surface_moving_out<R>(s, point, direction, false)) return R(0.0);
    R sd1 = inf, sd2 = inf;
    R d2 = (pDotV3d * pDotV3d - c);
    if (d2 < R(0) || ((c > R(0)) && (pDotV3d > R(0)))) return distance;
    if (d2 >= R(0)) sd1 = (-pDotV3d - vsqrt(vabs(d2)));
    R outerDist = inf, innerDist = inf;
    if (s.full_sphere) {
      if (sd1 >= R(0)) outerDist = sd1;
    } else {
      const V3<R> t = ray<R>(point, direction, sd1);
      if (wedge_contains<R>(s, t) && tc_contains<R>(s, t) && (sd1 >= R(0))) outerDist = sd1;
    }
    if (s.rmin != 0.) {
      c = rad2 - R(s.rmin) * R(s.rmin);
      d2 = pDotV3d * pDotV3d - c;
      if (d2 >= R(0)) sd2 = (-pDotV3d + vsqrt(vabs(d2)));
      if (s.full_sphere) {
        if (sd2 >= R(0)) innerDist = sd2;
      } else {
        const V3<R> t = ray<R>(point, direction, sd2);
        if ((sd2 >= R(0)) && wedge_contains<R>(s, t) && tc_contains<R>(s, t)) innerDist = sd2;
      }
    }
    distance = vmin(outerDist, innerDist);
    if (!s.full_phi) min_dist_from_phi<R>(s, point, direction, true, distance);
    R distThetaMin = inf;
    if (!s.full_theta) {
      bool intsect1 = false, intsect2 = false;
      R distTheta1 = inf, distTheta2 = inf;
      tc_distance_to_in<R>(s, point, direction, distTheta1, distTheta2, intsect1, intsect2);
      const V3<R> cp1 = ray<R>(point, direction, distTheta1);
      const R distCone1 = mag2<R>(cp1);
      const V3<R> cp2 = ray<R>(point, direction, distTheta2);
      const R distCone2 = mag2<R>(cp2);
      const R rmin2 = R(s.rmin) * R(s.rmin), rmax2 = R(s.rmax) * R(s.rmax);
      bool isValidCone1 = (distCone1 >= rmin2 && distCone1 <= rmax2) && intsect1;
      bool isValidCone2 = (distCone2 >= rmin2 && distCone2 <= rmax2) && intsect2;
      if (!s.full_phi) {
        isValidCone1 &= wedge_contains<R>(s, cp1);
        isValidCone2 &= wedge_contains<R>(s, cp2);
      }
      if (isValidCone2 && !isValidCone1) distThetaMin = distTheta2;
      if (isValidCone1 && !isValidCone2) distThetaMin = distTheta1;
      if (isValidCone1 && isValidCone2) distThetaMin = vmin(distTheta1, distTheta2);
    }
    distance = vmin(distThetaMin, distance);
    const V3<R> dd{R(0.) - point.x, R(0.) - point.y, R(0.) - point.z};
    const R dd2 = mag2<R>(dd);
    const R newDist = vsqrt(dd2);
    const R um = vsqrt(dd2 + R(DBL_MIN));  // Vector3D::Unit
    if ((s.stheta > kHalfTolerance || s.etheta < (kPi - kHalfTolerance)) &&
        (vabs(dd.x / um - direction.x) < R(kHalfTolerance)) && (vabs(dd.y / um - direction.y) < R(kHalfTolerance)) &&
        (vabs(dd.z / um - direction.z) < R(kHalfTolerance)))
      distance = vmin(distance, newDist);
    return distance;
  }

  // DistanceToOut (:227-280)
  template <typename R>
  VGM_DEV static R distance_to_out(const Data &s, V3<R> point, V3<R> direction, R /*step_max*/)
  {
    const R inf = inf_length<R>();
    R distance = inf;
    const R pDotV3d = dot<R>(point, direction);
    const R rad2 = mag2<R>(point);
    R c = rad2 - R(s.rmax) * R(s.rmax);
    R sd1 = inf, sd2 = inf;
    if (util_completely<R>(s, point, false)) return R(-1.0);
    if (on_surface_moving_out<R>(s, point, direction, true)) return R(0.0);
    R d2 = (pDotV3d * pDotV3d - c);
    if (d2 >= R(0)) sd1 = (-pDotV3d + vsqrt(vabs(d2)));
    if (s.rmin != 0.) {
      c = rad2 - R(s.rmin) * R(s.rmin);
      d2 = (pDotV3d * pDotV3d - c);
      if ((d2 >= R(0)) && (pDotV3d < R(0))) sd2 = (-pDotV3d - vsqrt(vabs(d2)));
    }
    const R snxt = vmin(sd1, sd2);
    const bool condSemi = ((s.stheta == 0. && s.etheta == kPi / 2.) && direction.z >= R(0)) ||
                          ((s.stheta == kPi / 2. && s.etheta == kPi) && direction.z <= R(0));
    if (condSemi) return snxt;
    R distThetaMin = inf;
    if (!s.full_theta) {
      bool intsect1 = false, intsect2 = false;
      R distTheta1 = inf, distTheta2 = inf;
      tc_distance_to_out<R>(s, point, direction, distTheta1, distTheta2, intsect1, intsect2);
      if (intsect2 && !intsect1) distThetaMin = distTheta2;
      if (!intsect2 && intsect1) distThetaMin = distTheta1;
      if (intsect2 && intsect1) distThetaMin = vmin(distTheta1, distTheta2);
    }
    distance = vmin(distThetaMin, snxt);
    if (!s.full_phi) {
      if (s.dphi <= kPi) {
        R distPhi1, distPhi2;
        wedge_distances<R>(s, point, direction, false, distPhi1, distPhi2);
        const R distPhiMin = vmin(distPhi1, distPhi2);
        distance = vmin(distPhiMin, distance);
      } else {
        min_dist_from_phi<R>(s, point, direction, false, distance);
      }
    }
    return distance;
  }

  // SafetyToIn (:335-372)
  template <typename R>
  VGM_DEV static R safety_to_in(const Data &s, V3<R> point)
  {
    const R rad = vsqrt(mag2<R>(point));
    bool ci = false, co = false;
    generic<R>(s, point, true, ci, co);
    if (ci) return R(-1.0);
    if (!ci && !co) return R(0.0);
    R safety;
    if (s.rmin != 0.) {
      const R safeRMin = R(s.rmin) - rad, safeRMax = rad - R(s.rmax);
      safety = (safeRMin > safeRMax) ? safeRMin : safeRMax;
    } else {
      safety = (rad - R(s.rmax));
    }
    if (!s.full_phi) safety = vmax(wedge_safety_to_in<R>(s, point), safety);
    if (!s.full_theta) safety = vmax(tc_safety_to_in<R>(s, point), safety);
    return safety;
  }

  // SafetyToOut (:377-413)
  template <typename R>
  VGM_DEV static R safety_to_out(const Data &s, V3<R> point)
  {
    const R rad = vsqrt(mag2<R>(point));
    bool ci = false, co = false;
    generic<R>(s, point, true, ci, co);
    if (co) return R(-1.0);
    if (!ci && !co) return R(0.0);
    R safety;
    if (s.rmin != 0.) {
      const R safeRMin = (rad - R(s.rmin)), safeRMax = (R(s.rmax) - rad);
      safety = (safeRMin < safeRMax) ? safeRMin : safeRMax;
    } else {
      safety = (R(s.rmax) - rad);
    }
    if (!s.full_phi) safety = vmin(wedge_safety_to_out<R>(s, point), safety);
    if (!s.full_theta) safety = vmin(tc_safety_to_out<R>(s, point), safety);
    return safety;
  }

  // ApproxSurfaceNormalKernel (:423-465)
  VGM_DEV static V3<double> approx_normal(const Data &s, V3<double> p)
  {
    V3<double> norm{0., 0., 0.};
    const double rr = mag2<double>(p);
    const double radius = __builtin_sqrt(rr);
    const double distRMax = vabs(radius - s.rmax);
    double distRMin = kInfLength, distPhi1 = kInfLength, distPhi2 = kInfLength, distTheta1 = kInfLength,
           distTheta2 = kInfLength;
    double distMin = distRMax;
    if (s.rmin > 0.) {
      distRMin = vabs(s.rmin - radius);
      distMin = vmin(distRMin, distRMax);
    }
    if (!s.full_phi) {
      distPhi1 = vabs(p.x * s.w_n1x + p.y * s.w_n1y);
      distPhi2 = vabs(p.x * s.w_n2x + p.y * s.w_n2y);
      distMin = vmin(vmin(distMin, distPhi1), distPhi2);
    }
    if (!s.full_theta) {
      const double rho = __builtin_sqrt(p.x * p.x + p.y * p.y);
      distTheta1 = tc_line_distance<double>(s.slope1, rho, p.z);
      distTheta2 = tc_line_distance<double>(s.slope2, rho, p.z);
      distMin = vmin(vmin(distMin, distTheta1), distTheta2);
    }
    const V3<double> u = v3_unit(p);
    if (distMin == distRMax) norm = u;
    if (distMin == distRMin) norm = V3<double>{-u.x, -u.y, -u.z};
    if (distMin == distPhi1) norm = V3<double>{-s.w_n1x, -s.w_n1y, -0.0};
    if (distMin == distPhi2) norm = V3<double>{-s.w_n2x, -s.w_n2y, -0.0};
    if (distMin == distTheta1) {
      const V3<double> t = tc_normal1<double>(s, p);
      norm = V3<double>{norm.x + t.x, norm.y + t.y, norm.z + t.z};
    }
    if (distMin == distTheta2) {
      const V3<double> t = tc_normal2<double>(s, p);
      norm = V3<double>{norm.x + t.x, norm.y + t.y, norm.z + t.z};
    }
    return norm;
  }

  // Normal (:470-545): the sum of the normals of the surfaces the point is
  // on (starting from (1e-30, 1e-30, 1e-30)), normalised
  VGM_DEV static bool normal(const Data &s, V3<double> p, V3<double> &out)
  {
    V3<double> n{1e-30, 1e-30, 1e-30};
    bool isIn = false, isOut = false;
    generic<double>(s, p, true, isIn, isOut);
    if (isIn || isOut) n = approx_normal(s, p);
    double noSurfaces = 0.;
    auto add = [&](const V3<double> &v, double sign) {
      n = V3<double>{n.x + sign * v.x, n.y + sign * v.y, n.z + sign * v.z};
    };
    if (on_radius<double>(s.rmax, p)) {
      noSurfaces = noSurfaces + 1;
      if (!isOut) {
        const V3<double> u = v3_unit(p);
        n = V3<double>{n.x + u.x, n.y + u.y, n.z + u.z};
      }
    }
    if (s.rmin != 0.) {
      if (on_radius<double>(s.rmin, p)) {
        noSurfaces = noSurfaces + 1;
        if (!isOut) {
          const V3<double> u = v3_unit(p);
          n = V3<double>{n.x - u.x, n.y - u.y, n.z - u.z};
        }
      }
    }
    if (!s.full_phi) {
      const bool onStart = wedge_on_surface<double>(s.w_a1x, s.w_a1y, s.w_n1x, s.w_n1y, p);
      const bool onEnd = wedge_on_surface<double>(s.w_a2x, s.w_a2y, s.w_n2x, s.w_n2y, p);
      if (onStart) noSurfaces = noSurfaces + 1;
      if (onEnd) noSurfaces = noSurfaces + 1;
      if (!isOut && onStart) n = V3<double>{n.x - s.w_n1x, n.y - s.w_n1y, n.z - 0.0};
      if (!isOut && onEnd) n = V3<double>{n.x - s.w_n2x, n.y - s.w_n2y, n.z - 0.0};
    }
    if (!s.full_theta) {
      const bool onStartTheta = tc_on_surface<double>(s, p, true);
      const bool onEndTheta = tc_on_surface<double>(s, p, false);
      if (onStartTheta) noSurfaces = noSurfaces + 1;
      if (!isOut && onStartTheta) add(tc_normal1<double>(s, p), 1.0);
      if (onEndTheta) noSurfaces = noSurfaces + 1;
      if (!isOut && onEndTheta) add(tc_normal2<double>(s, p), 1.0);
      if (!isOut && onStartTheta && onEndTheta && (s.etheta <= kPi / 2.)) n = V3<double>{0., 0., -1.};
      if (!isOut && onStartTheta && onEndTheta && (s.stheta >= kPi / 2.)) n = V3<double>{0., 0., 1.};
    }
    out = v3_normalized(n);
    return noSurfaces > 0.;
  }
};

}  // namespace vgm
