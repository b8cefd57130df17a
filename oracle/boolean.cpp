// Oracle: boolean solids of two placed operands.  Intersection:
// VecGeom/volumes/kernel/BooleanIntersectionImplementation.h; subtraction:
// VecGeom/volumes/kernel/BooleanImplementation.h:39-290; union:
// VecGeom/volumes/kernel/BooleanUnionImplementation.h; Extent:
// source/UnplacedBooleanVolume.cpp:403-482.  An operand's Normal is the
// placed volume's (source/PlacedVolume.cpp:126-137: transform the point,
// the unplaced normal, inverse-transform the direction); the booleans call
// it on points they have already transformed, as the reference does.  Placed-volume conventions of
// the operands: Contains, Inside, DistanceToIn, PlacedDistanceToOut and
// SafetyToIn transform the point into the operand's frame, DistanceToOut and
// SafetyToOut take it as given.  The reference's unbounded search loops stop
// after kMaxIterations passes with a miss.  TEST ONLY.
#include "core.h"

namespace orc {

namespace {
constexpr int kMaxIterations = 1000;

struct Placed {  // a VPlacedVolume: unplaced shape and transformation
  const Shape *s;
  Transform tr;
  bool Contains(const Vec3 &p) const { return s->Contains(tr.Transform_(p)); }
  int Inside(const Vec3 &p) const { return s->Inside(tr.Transform_(p)); }
  double DistanceToIn(const Vec3 &p, const Vec3 &d, double stepMax = kInfLength) const
  {
    return s->DistanceToIn(tr.Transform_(p), tr.TransformDirection(d), stepMax);
  }
  double PlacedDistanceToOut(const Vec3 &p, const Vec3 &d, double stepMax = kInfLength) const
  {
    return s->DistanceToOut(tr.Transform_(p), tr.TransformDirection(d), stepMax);
  }
  double DistanceToOut(const Vec3 &p, const Vec3 &d) const { return s->DistanceToOut(p, d, kInfLength); }
  double SafetyToIn(const Vec3 &p) const { return s->SafetyToIn(tr.Transform_(p)); }
  double SafetyToOut(const Vec3 &p) const { return s->SafetyToOut(p); }
  bool Normal(const Vec3 &p, Vec3 &n) const  // VPlacedVolume::Normal
  {
    Vec3 ln;
    const bool valid = s->Normal(tr.Transform_(p), ln);
    n = tr.InverseTransformDirection(ln);
    return valid;
  }
  // the booleans' pattern: transform, placed Normal, inverse-transform
  bool NormalOfLocal(const Vec3 &p, Vec3 &n, double sign = 1.0) const
  {
    Vec3 localNorm;
    const bool valid = Normal(tr.Transform_(p), localNorm);
    if (sign != 1.0) localNorm = Vec3(localNorm.x * sign, localNorm.y * sign, localNorm.z * sign);
    n = tr.InverseTransformDirection(localNorm);
    return valid;
  }
  void TransformedExtent(Vec3 &aMin, Vec3 &aMax) const  // :403-442
  {
    Vec3 lower, upper;
    s->Extent(lower, upper);
    const Vec3 delta = upper - lower;
    double mn[3] = {kInfLength, kInfLength, kInfLength}, mx[3] = {-kInfLength, -kInfLength, -kInfLength};
    for (int x = 0; x <= 1; ++x)
      for (int y = 0; y <= 1; ++y)
        for (int z = 0; z <= 1; ++z) {
          const Vec3 corner(lower.x + x * delta.x, lower.y + y * delta.y, lower.z + z * delta.z);
          const Vec3 t = tr.InverseTransform(corner);
          for (int k = 0; k < 3; ++k) {
            mn[k] = std::min(mn[k], t[k]);
            mx[k] = std::max(mx[k], t[k]);
          }
        }
    aMin = Vec3(mn[0], mn[1], mn[2]);
    aMax = Vec3(mx[0], mx[1], mx[2]);
  }
};
}  // namespace

struct Boolean final : Shape {
  int op;  // 1 intersection, 2 subtraction, 3 union
  Placed A, B;

  bool Contains(const Vec3 &p) const override
  {
    if (op == 3) return A.Contains(p) || B.Contains(p);
    if (op == 1) {
      const bool insideA = A.Contains(p), insideB = B.Contains(p);
      return insideA && insideB;
    }
    bool inside = A.Contains(p);
    if (!inside) return false;
    return inside && !B.Contains(p);
  }

  // union Inside (BooleanUnionImplementation.h:71-113): on the surfaces of
  // both, opposite normals mean touching solids (inside)
  int UnionInside(const Vec3 &p) const
  {
    const int positionA = A.Inside(p);
    if (positionA == kInside) return kInside;
    const int positionB = B.Inside(p);
    if (positionB == kInside) return kInside;
    if ((positionA == kSurface) && (positionB == kSurface)) {
      Vec3 normalA, normalB;
      A.NormalOfLocal(p, normalA);
      B.NormalOfLocal(p, normalB);
      if (Dot(normalA, normalB) < 0) return kInside;
      return kSurface;
    }
    if ((positionB == kSurface) || (positionA == kSurface)) return kSurface;
    return kOutside;
  }

  int Inside(const Vec3 &p) const override
  {
    if (op == 3) return UnionInside(p);
    const int positionA = A.Inside(p);
    if (positionA == kOutside) return kOutside;
    const int positionB = B.Inside(p);
    if (op == 1) {
      if (positionA == kInside && positionB == kInside) return kInside;
      if ((positionA == kInside && positionB == kSurface) || (positionB == kInside && positionA == kSurface) ||
          (positionA == kSurface && positionB == kSurface))
        return kSurface;
      return kOutside;
    }
    if (positionA == kInside && positionB == kOutside) return kInside;
    if ((positionA == kInside && positionB == kSurface) || (positionB == kOutside && positionA == kSurface))
      return kSurface;
    return kOutside;
  }

  double DistanceToIn(const Vec3 &p, const Vec3 &dir, double stepMax) const override
  {
    if (op == 3) {  // union (:118-126)
      const double d1 = A.DistanceToIn(p, dir, stepMax);
      const double d2 = B.DistanceToIn(p, dir, stepMax);
      return Min(d1, d2);
    }
    if (op == 1) {  // intersection
      Vec3 hitpoint = p;
      bool inleft = A.Contains(hitpoint), inright = B.Contains(hitpoint);
      double d1 = 0., d2 = 0., snext = 0.0;
      if (inleft && inright) {
        d1 = A.PlacedDistanceToOut(hitpoint, dir, stepMax);
        d2 = B.PlacedDistanceToOut(hitpoint, dir, stepMax);
        if (d1 < 2 * kTolerance) inleft = false;
        if (d2 < 2 * kTolerance) inright = false;
        if (inleft && inright) return 0.0;
      }
      for (int it = 0; it < kMaxIterations; ++it) {
        d1 = d2 = 0;
        if (!inleft) {
          d1 = Max(A.DistanceToIn(hitpoint, dir), kTolerance);
          if (d1 > 1E20) return kInfLength;
        }
        if (!inright) {
          d2 = Max(B.DistanceToIn(hitpoint, dir), kTolerance);
          if (d2 > 1E20) return kInfLength;
        }
        if (d1 > d2) {
          snext += d1;
          inleft = true;
          hitpoint = hitpoint + dir * d1;
          inright = B.Contains(hitpoint + dir * kTolerance);
          if (inright) return snext;
        } else {
          snext += d2;
          inright = true;
          hitpoint = hitpoint + dir * d2;
          inleft = A.Contains(hitpoint + dir * kTolerance);
          if (inleft) return snext;
        }
      }
      return kInfLength;
    }
    // subtraction
    double d1, d2, snxt = 0.;
    Vec3 hitpoint = p;
    bool insideRight = B.Contains(p);
    const double epsil = kPushTolerance;
    for (int it = 0; it < kMaxIterations; ++it) {
      if (insideRight) {
        d1 = B.PlacedDistanceToOut(hitpoint, dir, stepMax);
        snxt += (d1 >= 0. && d1 < kInfLength) ? (d1 + epsil) : 0.;
        hitpoint = hitpoint + dir * ((d1 >= 0. && d1 < kInfLength) ? (d1 + epsil) : 0.);
        if (A.Contains(hitpoint)) {
          const double check = A.PlacedDistanceToOut(hitpoint, dir);
          if (check > epsil) return snxt;
        }
      }
      d2 = Max(A.DistanceToIn(hitpoint, dir, stepMax), 0.);
      if (d2 == kInfLength) return kInfLength;
      d1 = B.DistanceToIn(hitpoint, dir, stepMax);
      if (d2 < d1 - kTolerance) {
        snxt += d2 + epsil;
        return snxt;
      }
      snxt += (d1 >= 0. && d1 < kInfLength) ? d1 + epsil : 0.;
      hitpoint = hitpoint + dir * ((d1 >= 0. && d1 < kInfLength) ? (d1 + epsil) : epsil);
      insideRight = true;
    }
    return kInfLength;
  }

  // union DistanceToOut (:128-179): leave the operand the point is in,
  // crossing into the other while it overlaps or connects
  double UnionDistanceToOut(const Vec3 &point, const Vec3 &dir) const
  {
    double dist = 0.;
    const double pushdist = kPushTolerance;
    const int positionA = A.Inside(point);
    Vec3 nextp = point;
    bool connectingstep = false;
    auto kernel = [&](const Placed &P, const Placed &Q) {
      int it = 0;
      do {
        connectingstep = false;
        const double disTmp = P.PlacedDistanceToOut(nextp, dir);
        dist += (disTmp >= 0. && disTmp < kInfLength) ? disTmp : 0;
        dist += pushdist;
        nextp = point + dir * dist;
        if (Q.Inside(nextp) != kOutside) {
          const double disTmp2 = Q.PlacedDistanceToOut(nextp, dir);
          dist += (disTmp2 >= 0. && disTmp2 < kInfLength) ? disTmp2 : 0;
          dist += pushdist;
          nextp = point + dir * dist;
          connectingstep = true;
        }
      } while (connectingstep && (P.Inside(nextp) != kOutside) && ++it < kMaxIterations);
    };
    if (positionA != kOutside)
      kernel(A, B);
    else
      kernel(B, A);
    double distance = dist - pushdist;
    if (distance < kTolerance && positionA == kOutside && B.Inside(point) == kOutside) distance = -kTolerance;
    return distance;
  }

  double DistanceToOut(const Vec3 &p, const Vec3 &dir, double stepMax) const override
  {
    if (op == 3) return UnionDistanceToOut(p, dir);
    if (op == 1) return Min(A.DistanceToOut(p, dir), B.PlacedDistanceToOut(p, dir));
    const double distancel = A.PlacedDistanceToOut(p, dir, stepMax);
    const double dinright = B.DistanceToIn(p, dir, stepMax);
    return Min(distancel, dinright);
  }

  double SafetyToIn(const Vec3 &p) const override
  {
    if (op == 3) return Min(A.SafetyToIn(p), B.SafetyToIn(p));  // :184-194
    if (op == 1) {
      const bool insideA = A.Contains(p), insideB = B.Contains(p);
      if (!insideA && insideB) return A.SafetyToIn(p);
      if (!insideB && insideA) return B.SafetyToIn(p);
      return Min(A.SafetyToIn(p), B.SafetyToIn(p));
    }
    if (A.Contains(p) && B.Contains(p)) return B.SafetyToOut(B.tr.Transform_(p));
    return A.SafetyToIn(p);
  }

  double SafetyToOut(const Vec3 &p) const override
  {
    if (op == 3) {  // :199-223
      const int insideA = A.Inside(p), insideB = B.Inside(p);
      if (insideA == kOutside && insideB == kOutside) return -kTolerance;
      if (insideA != kOutside && insideB != kOutside)
        return Max(A.SafetyToOut(p), B.SafetyToOut(B.tr.Transform_(p)));
      if (insideA == kSurface || insideB == kSurface) return -kTolerance;
      if (insideA == kOutside) return B.SafetyToOut(B.tr.Transform_(p));
      return A.SafetyToOut(p);
    }
    if (op == 1) {
      double safety = Min(A.SafetyToOut(p), B.SafetyToOut(B.tr.Transform_(p)));
      if (safety < 0.) safety = 0.;
      return safety;
    }
    return Min(A.SafetyToOut(p), B.SafetyToIn(p));
  }

  // NormalKernel of each operation: union BooleanUnionImplementation.h:228-279,
  // subtraction BooleanImplementation.h:239-288, intersection
  // BooleanIntersectionImplementation.h:234-274
  bool Normal(const Vec3 &point, Vec3 &normal) const override
  {
    if (op == 3) {
      if (A.Contains(point)) return A.NormalOfLocal(point, normal);
      if (B.Contains(point)) return B.NormalOfLocal(point, normal);
      const double safetyA = A.SafetyToIn(point), safetyB = B.SafetyToIn(point);
      if (safetyA < safetyB) return A.NormalOfLocal(point, normal);
      return B.NormalOfLocal(point, normal);
    }
    if (op == 2) {
      if (B.Contains(point)) return B.NormalOfLocal(point, normal, -1.);
      if (!A.Contains(point)) return A.NormalOfLocal(point, normal);
      const double safetyA = A.SafetyToOut(A.tr.Transform_(point));
      const double safetyB = B.SafetyToIn(point);
      if (safetyA < safetyB) return A.NormalOfLocal(point, normal);
      return B.NormalOfLocal(point, normal, -1.);
    }
    const double safetyA = A.Contains(point) ? A.SafetyToOut(A.tr.Transform_(point)) : A.SafetyToIn(point);
    const double safetyB = B.Contains(point) ? B.SafetyToOut(B.tr.Transform_(point)) : B.SafetyToIn(point);
    if (safetyA < safetyB) return A.NormalOfLocal(point, normal);
    return B.NormalOfLocal(point, normal);
  }
  bool HasNormal() const override { return A.s->HasNormal() && B.s->HasNormal(); }

  void Extent(Vec3 &aMin, Vec3 &aMax) const override  // :444-482
  {
    Vec3 minLeft, maxLeft, minRight, maxRight;
    A.TransformedExtent(minLeft, maxLeft);
    B.TransformedExtent(minRight, maxRight);
    if (op == 2) {
      aMin = minLeft;
      aMax = maxLeft;
      return;
    }
    if (op == 3) {
      aMin = Vec3(Min(minLeft.x, minRight.x), Min(minLeft.y, minRight.y), Min(minLeft.z, minRight.z));
      aMax = Vec3(Max(maxLeft.x, maxRight.x), Max(maxLeft.y, maxRight.y), Max(maxLeft.z, maxRight.z));
      return;
    }
    aMin = Vec3(Max(minLeft.x, minRight.x), Max(minLeft.y, minRight.y), Max(minLeft.z, minRight.z));
    aMax = Vec3(Min(maxLeft.x, maxRight.x), Min(maxLeft.y, maxRight.y), Min(maxLeft.z, maxRight.z));
  }
};

Shape *make_boolean(int op, const Shape *left, const double *ltr, const Shape *right, const double *rtr)
{
  auto *b = new Boolean();
  b->op = op;
  b->A.s = left;
  b->B.s = right;
  for (int k = 0; k < 3; ++k) b->A.tr.t[k] = ltr ? ltr[k] : 0.0, b->B.tr.t[k] = rtr ? rtr[k] : 0.0;
  for (int k = 0; k < 9; ++k) {
    b->A.tr.r[k] = ltr ? ltr[3 + k] : ((k % 4 == 0) ? 1.0 : 0.0);
    b->B.tr.r[k] = rtr ? rtr[3 + k] : ((k % 4 == 0) ? 1.0 : 0.0);
  }
  return b;
}

}  // namespace orc
