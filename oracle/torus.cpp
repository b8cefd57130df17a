// Oracle: torus with a phi cut.  TorusStruct2 (VecGeom/volumes/TorusStruct2.h:25-43),
// kernels (VecGeom/volumes/kernel/TorusImplementation2.h), Normal
// (source/UnplacedTorus2.cpp:201-285) and Extent (VecGeom/volumes/UnplacedTorus2.h:210,
// the bounding tube's).  The quartic is solved as the reference solves it:
// SolveCubic with std::cbrt / std::pow / std::cos / std::atan2, SolveQuartic,
// then Newton refinement on the quartic.  TEST ONLY.
//
// Reference behaviour kept as written:
//   * Contains ignores the phi cut (GenericKernelForContainsAndInside<true, false>);
//   * DistanceToIn takes a phi-plane hit only when it lies within kTolerance of
//     the point (TorusImplementation2.h:600-640), other entries come from the
//     torus surface alone;
//   * ToBoundary's small-|dir.z| branch sorts all four slots, including a root
//     it rejected (left in its slot), and then walks the first nsol of them.
#include "core.h"
#include "tube.h"
#include "wedge.h"

namespace orc {

namespace {

// SolveCubic (TorusImplementation2.h:24-64): one or three real roots of
// x^3 + a x^2 + b x + c, unsorted
unsigned SolveCubic(double a, double b, double c, double *x)
{
  const double ott = 1. / 3.;
  const double sq3 = std::sqrt(3.);
  const double inv6sq3 = 1. / (6. * sq3);
  unsigned ireal = 1;
  const double p = b - a * a * ott;
  const double q = c - a * b * ott + 2. * a * a * a * ott * ott * ott;
  double delta = 4 * p * p * p + 27. * q * q;
  double t, u;
  if (delta >= 0) {
    delta = std::sqrt(delta);
    t = (-3 * q * sq3 + delta) * inv6sq3;
    u = (3 * q * sq3 + delta) * inv6sq3;
    x[0] = std::copysign(1., t) * std::cbrt(std::fabs(t)) - std::copysign(1., u) * std::cbrt(std::fabs(u)) - a * ott;
  } else {
    delta = std::sqrt(-delta);
    t = -0.5 * q;
    u = delta * inv6sq3;
    x[0] = 2. * std::pow(t * t + u * u, 0.5 * ott) * std::cos(ott * std::atan2(u, t));
    x[0] -= a * ott;
  }
  t = x[0] * x[0] + a * x[0] + b;
  u = a + x[0];
  delta = u * u - 4. * t;
  if (delta >= 0) {
    ireal = 3;
    delta = std::sqrt(delta);
    x[1] = 0.5 * (-u - delta);
    x[2] = 0.5 * (-u + delta);
  }
  return ireal;
}

// Sort4 (:66-91): the five compare-and-swaps of a 4-element network
void CmpSwap(double *v, int i, int j)
{
  if (v[i] > v[j]) {
    const double c = v[j];
    v[j] = v[i];
    v[i] = c;
  }
}
void Sort4(double *v)
{
  CmpSwap(v, 0, 2);
  CmpSwap(v, 1, 3);
  CmpSwap(v, 0, 1);
  CmpSwap(v, 2, 3);
  CmpSwap(v, 1, 2);
}

// SolveQuartic (:97-174): real roots of x^4 + a x^3 + b x^2 + c x + d,
// sorted into x[0..3] (slots past the count keep what the caller put there)
int SolveQuartic(double a, double b, double c, double d, double *x)
{
  const double e = b - 3. * a * a / 8.;
  const double f = c + a * a * a / 8. - 0.5 * a * b;
  const double g = d - 3. * a * a * a * a / 256. + a * a * b / 16. - a * c / 4.;
  double xx[4] = {kInfLength, kInfLength, kInfLength, kInfLength};
  double delta;
  double h = 0.;
  unsigned ireal = 0;
  if (Abs(f) < 1E-6) {  // biquadratic
    delta = e * e - 4. * g;
    if (delta < 0.) return 0;
    delta = std::sqrt(delta);
    h = 0.5 * (-e - delta);
    if (h >= 0) {
      h = std::sqrt(h);
      x[ireal++] = -h - 0.25 * a;
      x[ireal++] = h - 0.25 * a;
    }
    h = 0.5 * (-e + delta);
    if (h >= 0) {
      h = std::sqrt(h);
      x[ireal++] = -h - 0.25 * a;
      x[ireal++] = h - 0.25 * a;
    }
    Sort4(x);
    return ireal;
  }
  if (Abs(g) < 1E-6) {  // one root at -a/4, the cubic for the rest
    x[ireal++] = -0.25 * a;
    const unsigned ncubicroots = SolveCubic(0, e, f, xx);
    for (unsigned i = 0; i < ncubicroots; i++) x[ireal++] = xx[i] - 0.25 * a;
    Sort4(x);
    return ireal;
  }
  ireal = SolveCubic(2. * e, e * e - 4. * g, -f * f, xx);
  if (ireal == 1) {
    if (xx[0] <= 0) return 0;
    h = std::sqrt(xx[0]);
  } else {
    for (unsigned i = 0; i < 3; i++) {
      h = xx[i];
      if (h >= 0) break;
    }
    if (h <= 0) return 0;
    h = std::sqrt(h);
  }
  const double j = 0.5 * (e + h * h - f / h);
  ireal = 0;
  delta = h * h - 4. * j;
  if (delta >= 0) {
    delta = std::sqrt(delta);
    x[ireal++] = 0.5 * (-h - delta) - 0.25 * a;
    x[ireal++] = 0.5 * (-h + delta) - 0.25 * a;
  }
  delta = h * h - 4. * g / j;
  if (delta >= 0) {
    delta = std::sqrt(delta);
    x[ireal++] = 0.5 * (h - delta) - 0.25 * a;
    x[ireal++] = 0.5 * (h + delta) - 0.25 * a;
  }
  Sort4(x);
  return ireal;
}

struct Torus final : Shape {
  double fRmin, fRmax, fRtor, fSphi, fDphi, fRmin2, fRmax2, fRtor2;
  Wedge fPhiWedge;
  Tube fBoundingTube;

  Torus(double rmin, double rmax, double rtor, double sphi, double dphi)
      : fRmin(rmin), fRmax(rmax), fRtor(rtor), fSphi(sphi), fDphi(dphi), fRmin2(rmin * rmin), fRmax2(rmax * rmax),
        fRtor2(rtor * rtor), fPhiWedge(dphi, sphi),
        fBoundingTube(rtor - rmax - kTolerance, rtor + rmax + kTolerance, rmax, sphi, dphi)
  {
  }

  bool HasPhi() const { return fDphi < kTwoPi; }

  // DistSqrToTorusR (:190-196)
  double DistSqrToTorusR(const Vec3 &point, const Vec3 &dir, double dist) const
  {
    const Vec3 p = point + dir * dist;
    const double rxy = p.Perp();
    return (rxy - fRtor) * (rxy - fRtor) + p.z * p.z;
  }

  // GenericKernelForContainsAndInside (:331-384)
  void Generic(const Vec3 &point, bool forInside, bool notForDisk, bool &cin, bool &cout) const
  {
    const double tol = 100. * kTolerance;
    const double rxy = std::sqrt(point[0] * point[0] + point[1] * point[1]);
    const double radsq = (rxy - fRtor) * (rxy - fRtor) + point[2] * point[2];
    cin = false;
    if (forInside) {
      cout = radsq > (tol * fRmax + fRmax2);
      cin = radsq < (-tol * fRmax + fRmax2);
    } else {
      cout = radsq > fRmax2;
    }
    if (cout) return;
    if (forInside) {
      cout |= radsq < (-tol * fRmin + fRmin2);
      cin &= radsq > (tol * fRmin + fRmin2);
    } else {
      cout |= radsq < fRmin2;
    }
    if (cout) return;
    if (HasPhi() && notForDisk) {
      bool cinPhi = false, coutPhi = false;
      fPhiWedge.GenericKernel(point, forInside, cinPhi, coutPhi);
      cout |= coutPhi;
      if (forInside) cin &= cinPhi;
    }
  }

  // UnplacedContainsDisk (:296-310): ContainsKernel<false> without the wedge
  bool ContainsDisk(const Vec3 &p) const
  {
    bool ci, co;
    Generic(p, false, false, ci, co);
    return !co;
  }

  // Contains (:520-529): the tolerant test, phi cut ignored
  bool Contains(const Vec3 &p) const override
  {
    bool ci, co;
    Generic(p, true, false, ci, co);
    return !co;
  }

  // InsideKernel (:315-326)
  int Inside(const Vec3 &p) const override
  {
    bool ci, co;
    Generic(p, true, true, ci, co);
    int inside = kSurface;
    if (co) inside = kOutside;
    if (ci) inside = kInside;
    return inside;
  }

  // ToBoundary (:389-496): distance along dir to the torus of tube radius
  // `radius`, leaving (out) or entering
  double ToBoundary(const Vec3 &pt, const Vec3 &dir, double radius, bool out) const
  {
    double s = kInfLength;
    const double tol = 100. * kTolerance;
    const double r0sq = pt[0] * pt[0] + pt[1] * pt[1] + pt[2] * pt[2];
    const double rdotn = pt[0] * dir[0] + pt[1] * dir[1] + pt[2] * dir[2];
    const double rsumsq = fRtor2 + radius * radius;
    const double a = 4. * rdotn;
    const double b = 2. * (r0sq + 2. * rdotn * rdotn - rsumsq + 2. * fRtor2 * dir[2] * dir[2]);
    const double c = 4. * (r0sq * rdotn - rsumsq * rdotn + 2. * fRtor2 * pt[2] * dir[2]);
    const double d = r0sq * r0sq - 2. * r0sq * rsumsq + 4. * fRtor2 * pt[2] * pt[2] +
                     (fRtor2 - radius * radius) * (fRtor2 - radius * radius);
    double x[4] = {kInfLength, kInfLength, kInfLength, kInfLength};
    int nsol = 0;
    if (Abs(dir[2]) < 1E-3 && Abs(pt[2]) < 0.1 * radius) {
      // nearly horizontal ray near the mid-plane: the two circles of the
      // section at this z, each a quadratic in the ray parameter
      double r0 = fRtor - std::sqrt((radius - pt[2]) * (radius + pt[2]));
      const double invdirxy2 = 1. / (1 - dir.z * dir.z);
      const double b0 = (pt[0] * dir[0] + pt[1] * dir[1]) * invdirxy2;
      double c0 = (pt[0] * pt[0] + (pt[1] - r0) * (pt[1] + r0)) * invdirxy2;
      double delta = b0 * b0 - c0;
      if (delta > 0) {
        x[nsol] = -b0 - std::sqrt(delta);
        if (x[nsol] > -tol) nsol++;
        x[nsol] = -b0 + std::sqrt(delta);
        if (x[nsol] > -tol) nsol++;
      }
      r0 = fRtor + std::sqrt((radius - pt[2]) * (radius + pt[2]));
      c0 = (pt[0] * pt[0] + (pt[1] - r0) * (pt[1] + r0)) * invdirxy2;
      delta = b0 * b0 - c0;
      if (delta > 0) {
        x[nsol] = -b0 - std::sqrt(delta);
        if (x[nsol] > -tol) nsol++;
        x[nsol] = -b0 + std::sqrt(delta);
        if (x[nsol] > -tol) nsol++;
      }
      if (nsol) Sort4(x);
    } else {
      nsol = SolveQuartic(a, b, c, d, x);
    }
    if (!nsol) return kInfLength;
    const bool inner = Abs(radius - fRmin) < kTolerance;
    for (int i = 0; i < nsol; i++) {
      if (x[i] < -10) continue;
      Vec3 r0 = pt + dir * x[i];
      Vec3 norm = r0;
      r0.z = 0.;
      r0 = r0.Normalized();
      r0 = r0 * fRtor;
      norm = norm - r0;
      const double ndotd = Dot(norm, dir);
      if (inner ^ out) {
        if (ndotd < 0) continue;
      } else {
        if (ndotd > 0) continue;
      }
      if (HasPhi()) {
        bool ci, co;
        fPhiWedge.GenericKernel(r0, true, ci, co);  // ContainsWithBoundary
        if (co) continue;
      }
      s = x[i];
      double eps = kInfLength;
      double delta = s * s * s * s + a * s * s * s + b * s * s + c * s + d;
      double eps0 = -delta / (4. * s * s * s + 3. * a * s * s + 2. * b * s + c);
      int ntry = 0;
      while (Abs(eps) > kTolerance) {
        if (Abs(eps0) > 100) break;
        s += eps0;
        if (Abs(s + eps0) < kTolerance) break;
        delta = s * s * s * s + a * s * s * s + b * s * s + c * s + d;
        eps = -delta / (4. * s * s * s + 3. * a * s * s + 2. * b * s + c);
        if (Abs(eps) >= Abs(eps0)) break;
        ntry++;
        if (ntry > 100) break;
        eps0 = eps;
      }
      if (s < -tol) continue;
      return Max(0., s);
    }
    return kInfLength;
  }

  // DistanceToIn (:540-658)
  double DistanceToIn(const Vec3 &point, const Vec3 &direction, double stepMax) const override
  {
    Vec3 localPoint = point;
    double distance = kInfLength;
    double tubeDistance = kInfLength;
    const bool inBounds = fBoundingTube.Contains(localPoint);
    if (!inBounds)
      tubeDistance = fBoundingTube.DistanceToIn(localPoint, direction, stepMax);
    else
      tubeDistance = 0.;
    if (inBounds) {
      if (Inside(point) == kInside) return -1.;
    } else if (tubeDistance == kInfLength) {
      return distance;
    }
    localPoint += direction * tubeDistance;
    if (HasPhi()) {
      double d1, d2;
      fPhiWedge.DistanceToIn(localPoint, direction, d1, d2);
      if (d1 != kInfLength) {
        const double daxis = DistSqrToTorusR(localPoint, direction, d1);
        if (daxis >= fRmin2 && daxis < fRmax2) {
          distance = d1;
          if (Abs(distance) < kTolerance) return distance + tubeDistance;
        }
      }
      if (d2 != kInfLength) {
        const double daxis = DistSqrToTorusR(localPoint, direction, d2);
        if (daxis >= fRmin2 && daxis < fRmax2) {
          distance = Min(d2, distance);
          if (Abs(distance) < kTolerance) return distance + tubeDistance;
        }
      }
      distance = kInfLength;
    }
    double dd = ToBoundary(localPoint, direction, fRmax, false);
    if (fRmin > 0.) dd = Min(dd, ToBoundary(localPoint, direction, fRmin, false));
    distance = Min(distance, dd);
    distance += tubeDistance;
    if (Abs(distance) > kInfLength) distance = kInfLength;
    return distance;
  }

  // DistanceToOut (:201-292)
  double DistanceToOut(const Vec3 &point, const Vec3 &dir, double) const override
  {
    const double distz = Abs(point.z) - fRmax;
    bool done = distz > kHalfTolerance;
    const double rsq = point.x * point.x + point.y * point.y;
    const double outerExclRadius = fRtor + fRmax + kHalfTolerance;
    done |= rsq > outerExclRadius * outerExclRadius;
    const double innerExclRadius = fRtor - fRmax - kHalfTolerance;
    done |= rsq < innerExclRadius * innerExclRadius;
    if (done) return -1.;
    if (Inside(point) == kOutside) return -1.;
    const double dout = ToBoundary(point, dir, fRmax, true);
    double din = kInfLength;
    if (fRmin > 0) din = ToBoundary(point, dir, fRmin, true);
    double distance = Min(dout, din);
    if (HasPhi()) {
      double distPhi1, distPhi2;
      fPhiWedge.DistanceToOut(point, dir, distPhi1, distPhi2);
      if (distPhi1 < distance) {
        const Vec3 ip = point + dir * distPhi1;
        if (ContainsDisk(ip)) {
          const double diri = ip.x * fPhiWedge.fAlong1.x + ip.y * fPhiWedge.fAlong1.y;
          if (diri >= 0) distance = distPhi1;
        }
      }
      if (distPhi2 < distance) {
        const Vec3 ip = point + dir * distPhi2;
        if (ContainsDisk(ip)) {
          const double diri2 = ip.x * fPhiWedge.fAlong2.x + ip.y * fPhiWedge.fAlong2.y;
          if (diri2 >= 0.) distance = distPhi2;
        }
      }
    }
    if (distance >= kInfLength) distance = -1.;
    return distance;
  }

  // SafetyToIn (:662-679)
  double SafetyToIn(const Vec3 &p) const override
  {
    const double rxy = std::sqrt(p[0] * p[0] + p[1] * p[1]);
    const double rad = std::sqrt((rxy - fRtor) * (rxy - fRtor) + p[2] * p[2]);
    double safety = rad - fRmax;
    if (fRmin != 0.) safety = Max(fRmin - rad, rad - fRmax);
    if (HasPhi() && rxy != 0.) safety = Max(fPhiWedge.SafetyToIn(p), safety);
    return safety;
  }

  // SafetyToOut (:500-516)
  double SafetyToOut(const Vec3 &p) const override
  {
    const double rxy = std::sqrt(p[0] * p[0] + p[1] * p[1]);
    const double rad = std::sqrt((rxy - fRtor) * (rxy - fRtor) + p[2] * p[2]);
    double safety = fRmax - rad;
    if (fRmin != 0.) safety = Min(rad - fRmin, fRmax - rad);
    if (HasPhi()) safety = Min(fPhiWedge.SafetyToOut(p), safety);
    return safety;
  }

  // UnplacedTorus2::Normal (source/UnplacedTorus2.cpp:201-285): the sum of the
  // normals of the surfaces within delta (radial) or dAngle (phi) of the
  // point; a single surface's normal is taken as is, a sum is made a unit
  // vector; no surface leaves the normal at zero and reports invalid
  bool Normal(const Vec3 &point, Vec3 &norm) const override
  {
    int noSurfaces = 0;
    bool valid = true;
    double distRMin = kInfLength, distSPhi = kInfLength, distEPhi = kInfLength;
    const double delta = Max(10.0 * kTolerance, 1.0e-8 * (fRtor + fRmax));
    const double dAngle = 10.0 * kTolerance;
    Vec3 nR, nPs, nPe, sumnorm(0., 0., 0.);
    const double rho2 = point.x * point.x + point.y * point.y;
    const double rho = std::sqrt(rho2);
    double pt2 = rho2 + point.z * point.z + fRtor * (fRtor - 2 * rho);
    pt2 = Max(pt2, 0.);
    const double pt = std::sqrt(pt2);
    const double distRMax = Abs(pt - fRmax);
    if (fRmin != 0.) distRMin = Abs(pt - fRmin);
    if (rho > delta && pt != 0.) {
      const double redFactor = (rho - fRtor) / rho;
      nR = Vec3(point.x * redFactor, point.y * redFactor, point.z);
      nR = nR * (1. / pt);
    }
    if (HasPhi()) {
      if (rho != 0.) {
        double pPhi = std::atan2(point.y, point.x);
        if (pPhi < fSphi - delta)
          pPhi += kTwoPi;
        else if (pPhi > fSphi + fDphi + delta)
          pPhi -= kTwoPi;
        distSPhi = Abs(pPhi - fSphi);
        distEPhi = Abs(pPhi - fSphi - fDphi);
      }
      nPs = Vec3(std::sin(fSphi), -std::cos(fSphi), 0);
      nPe = Vec3(-std::sin(fSphi + fDphi), std::cos(fSphi + fDphi), 0);
    }
    if (distRMax <= delta) {
      noSurfaces++;
      sumnorm += nR;
    } else if (fRmin != 0. && (distRMin <= delta)) {
      noSurfaces++;
      sumnorm = sumnorm - nR;
    }
    if (HasPhi() && (fRmin - delta <= pt) && (pt <= (fRmax + delta))) {
      if (distSPhi <= dAngle) {
        noSurfaces++;
        sumnorm += nPs;
      }
      if (distEPhi <= dAngle) {
        noSurfaces++;
        sumnorm += nPe;
      }
    }
    if (noSurfaces == 0)
      valid = false;
    else if (noSurfaces == 1)
      norm = sumnorm;
    else
      norm = sumnorm.Unit();
    return valid;
  }
  bool HasNormal() const override { return true; }

  // the bounding tube's box (UnplacedTorus2.h:210); the full box of the
  // tube, which contains the phi-trimmed one
  void Extent(Vec3 &mn, Vec3 &mx) const override { fBoundingTube.Extent(mn, mx); }
};

}  // namespace

Shape *make_torus(double rmin, double rmax, double rtor, double sphi, double dphi)
{
  return new Torus(rmin, rmax, rtor, sphi, dphi);
}

}  // namespace orc
