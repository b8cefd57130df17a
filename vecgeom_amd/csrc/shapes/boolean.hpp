// Boolean solids of two placed operands: intersection
// (VecGeom/volumes/kernel/BooleanIntersectionImplementation.h), subtraction
// (VecGeom/volumes/kernel/BooleanImplementation.h:39-290) and union
// (VecGeom/volumes/kernel/BooleanUnionImplementation.h).  The operands'
// placed-volume calls follow the reference's conventions: Contains, Inside,
// DistanceToIn, PlacedDistanceToOut and SafetyToIn take the point in the
// boolean's frame and transform it; DistanceToOut and SafetyToOut take it as
// given.  The reference's unbounded search loops are capped at
// kBoolMaxIterations passes, after which the ray counts as a miss.
// Operands are the families of kBoolOperandMask (no booleans, no tessellated
// solids, no tori); the builder refuses the others.
#pragma once
#include "../vgm_math.hpp"
#include "../transform.hpp"
#include "boolean_data.h"

namespace vgm {

constexpr int kBoolMaxIterations = 1000;

enum BoolOperandQuery : int { kOpContains, kOpInside, kOpDistIn, kOpDistOut, kOpSafIn, kOpSafOut };

template <typename R>
struct BoolOperandCall {
  const double *params;
  int query;
  V3<R> p, d;
  R step;
  R result;
  template <class S>
  VGM_DEV void operator()()
  {
    const auto &s = *reinterpret_cast<const typename S::Data *>(params);
    switch (query) {
    case kOpContains: result = S::template contains<R>(s, p) ? R(1) : R(0); break;
    case kOpInside: result = R(S::template inside<R>(s, p)); break;
    case kOpDistIn: result = S::template distance_to_in<R>(s, p, d, step); break;
    case kOpDistOut: result = S::template distance_to_out<R>(s, p, d, step); break;
    case kOpSafIn: result = S::template safety_to_in<R>(s, p); break;
    default: result = S::template safety_to_out<R>(s, p); break;
    }
  }
};

// An operand's unplaced Normal at a point of its own frame.
struct BoolOperandNormal {
  const double *params;
  V3<double> p, n;
  bool valid;
  template <class S>
  VGM_DEV void operator()()
  {
    valid = S::normal(*reinterpret_cast<const typename S::Data *>(params), p, n);
  }
};

template <unsigned Mask, class F>
__host__ __device__ inline bool dispatch_shape_mask(int type, F &&f);

struct BooleanShape {
  using Data = BooleanData;
  static constexpr bool kStage = false;
  static constexpr int kWavesPerEU = 1;
  static constexpr bool kSafetyIsDistance = false;
  static constexpr bool kSafetyBelowDistance = false;
  // mixed precision: Inside/Contains classified in float away from the surface
  static constexpr bool kFloatClassify = false;
  // mixed precision: the float safety less the guard band bounds the FP64 safety
  static constexpr bool kFloatSafety = false;
  // mixed precision: distance_to_in_mixed / distance_to_out_mixed (float selects
  // the surface, FP64 evaluates it)
  static constexpr bool kMixedDistance = false;

  // one operand query; `placed` transforms the point (and direction) into the
  // operand's frame.  Not inlined: the dispatch over every operand family and
  // query is compiled once per kernel rather than at each of its call sites.
  template <typename R>
  __host__ __device__ __attribute__((noinline)) static R op(const Data &b, bool right, int q, V3<R> p, V3<R> d = V3<R>{R(0), R(0), R(0)},
                      R step = inf_length<R>(), bool placed = true)
  {
    const Transform &T = right ? b.rtr : b.ltr;
    if (placed) {
      p = to_local_point<R>(T, p);
      if (q == kOpDistIn || q == kOpDistOut) d = to_local_dir<R>(T, d);
    }
    BoolOperandCall<R> c{reinterpret_cast<const double *>(&b) + (right ? b.roff : b.loff), q, p, d, step, R(0)};
    dispatch_shape_mask<kBoolOperandMask>(right ? b.rtype : b.ltype, c);
    return c.result;
  }
  // VPlacedVolume::Normal of an operand (source/PlacedVolume.cpp:126-137):
  // the point into the operand's frame, the unplaced normal, its direction
  // back.  Not inlined, for the reason op is not.
  __host__ __device__ __attribute__((noinline)) static bool op_normal(const Data &b, bool right, V3<double> p,
                                                                       V3<double> &n)
  {
    const Transform &T = right ? b.rtr : b.ltr;
    BoolOperandNormal c{reinterpret_cast<const double *>(&b) + (right ? b.roff : b.loff), to_local_point<double>(T, p),
                        V3<double>{0., 0., 0.}, false};
    dispatch_shape_mask<kBoolOperandMask>(right ? b.rtype : b.ltype, c);
    n = to_master_dir<double>(T, c.n);
    return c.valid;
  }
  // The booleans' pattern: transform the point into the operand's frame,
  // then call the operand's placed Normal (which transforms it again), and
  // turn the normal (negated for a subtracted operand) back.
  VGM_DEV static bool normal_of_local(const Data &b, bool right, V3<double> p, V3<double> &n, double sign = 1.0)
  {
    const Transform &T = right ? b.rtr : b.ltr;
    V3<double> ln;
    const bool valid = op_normal(b, right, to_local_point<double>(T, p), ln);
    if (sign != 1.0) ln = V3<double>{ln.x * sign, ln.y * sign, ln.z * sign};
    n = to_master_dir<double>(T, ln);
    return valid;
  }

  template <typename R>
  VGM_DEV static bool contains_op(const Data &b, bool right, V3<R> p)
  {
    return op<R>(b, right, kOpContains, p) != R(0);
  }
  template <typename R>
  VGM_DEV static int inside_op(const Data &b, bool right, V3<R> p)
  {
    return int(op<R>(b, right, kOpInside, p));
  }
  template <typename R>
  VGM_DEV static V3<R> advance(V3<R> p, R s, V3<R> d)
  {
    return V3<R>{p.x + s * d.x, p.y + s * d.y, p.z + s * d.z};
  }

  template <typename R>
  VGM_DEV static bool contains(const Data &b, V3<R> p)
  {
    if (b.op == kBoolUnion) {  // BooleanUnionImplementation.h:61-67
      if (contains_op<R>(b, false, p)) return true;
      return contains_op<R>(b, true, p);
    }
    if (b.op == kBoolIntersection) {
      const bool insideA = contains_op<R>(b, false, p);
      const bool insideB = contains_op<R>(b, true, p);
      return insideA && insideB;
    }
    if (!contains_op<R>(b, false, p)) return false;
    return !contains_op<R>(b, true, p);
  }

  template <typename R>
  VGM_DEV static int inside(const Data &b, V3<R> p)
  {
    if (b.op == kBoolUnion) {  // BooleanUnionImplementation.h:71-113
      const int positionA = inside_op<R>(b, false, p);
      if (positionA == kInsideV) return kInsideV;
      const int positionB = inside_op<R>(b, true, p);
      if (positionB == kInsideV) return kInsideV;
      if ((positionA == kSurfaceV) && (positionB == kSurfaceV)) {
        // opposite normals: touching solids, the point is inside
        V3<double> normalA, normalB;
        const V3<double> pd{double(p.x), double(p.y), double(p.z)};
        normal_of_local(b, false, pd, normalA);
        normal_of_local(b, true, pd, normalB);
        if (normalA.x * normalB.x + normalA.y * normalB.y + normalA.z * normalB.z < 0) return kInsideV;
        return kSurfaceV;
      }
      if ((positionB == kSurfaceV) || (positionA == kSurfaceV)) return kSurfaceV;
      return kOutsideV;
    }
    const int positionA = inside_op<R>(b, false, p);
    if (positionA == kOutsideV) return kOutsideV;
    const int positionB = inside_op<R>(b, true, p);
    if (b.op == kBoolIntersection) {
      if (positionA == kInsideV && positionB == kInsideV) return kInsideV;
      if ((positionA == kInsideV && positionB == kSurfaceV) || (positionB == kInsideV && positionA == kSurfaceV) ||
          (positionA == kSurfaceV && positionB == kSurfaceV))
        return kSurfaceV;
      return kOutsideV;
    }
    if (positionA == kInsideV && positionB == kOutsideV) return kInsideV;
    if ((positionA == kInsideV && positionB == kSurfaceV) || (positionB == kOutsideV && positionA == kSurfaceV))
      return kSurfaceV;
    return kOutsideV;
  }

  template <typename R>
  VGM_DEV static R distance_to_in(const Data &b, V3<R> p, V3<R> dir, R stepMax)
  {
    const R big = inf_length<R>();
    if (b.op == kBoolUnion) {  // BooleanUnionImplementation.h:118-126
      const R d1 = op<R>(b, false, kOpDistIn, p, dir, stepMax);
      const R d2 = op<R>(b, true, kOpDistIn, p, dir, stepMax);
      return vmin(d1, d2);
    }
    if (b.op == kBoolIntersection) {  // BooleanIntersectionImplementation.h DistanceToIn
      V3<R> hitpoint = p;
      bool inleft = contains_op<R>(b, false, hitpoint);
      bool inright = contains_op<R>(b, true, hitpoint);
      R d1 = R(0), d2 = R(0), snext = R(0);
      if (inleft && inright) {
        d1 = op<R>(b, false, kOpDistOut, hitpoint, dir, stepMax);
        d2 = op<R>(b, true, kOpDistOut, hitpoint, dir, stepMax);
        if (d1 < R(2 * kTolerance)) inleft = false;
        if (d2 < R(2 * kTolerance)) inright = false;
        if (inleft && inright) return R(0);
      }
      for (int it = 0; it < kBoolMaxIterations; ++it) {
        d1 = d2 = R(0);
        if (!inleft) {
          d1 = vmax(op<R>(b, false, kOpDistIn, hitpoint, dir), R(kTolerance));
          if (d1 > R(1e20)) return big;
        }
        if (!inright) {
          d2 = vmax(op<R>(b, true, kOpDistIn, hitpoint, dir), R(kTolerance));
          if (d2 > R(1e20)) return big;
        }
        if (d1 > d2) {
          snext += d1;
          inleft = true;
          hitpoint = advance<R>(hitpoint, d1, dir);
          inright = contains_op<R>(b, true, advance<R>(hitpoint, R(kTolerance), dir));
          if (inright) return snext;
        } else {
          snext += d2;
          inright = true;
          hitpoint = advance<R>(hitpoint, d2, dir);
          inleft = contains_op<R>(b, false, advance<R>(hitpoint, R(kTolerance), dir));
          if (inleft) return snext;
        }
      }
      return big;
    }
    // subtraction (BooleanImplementation.h:133-194)
    R d1, d2, snxt = R(0);
    V3<R> hitpoint = p;
    bool insideRight = contains_op<R>(b, true, p);
    const R epsil = R(kPushTolerance);
    for (int it = 0; it < kBoolMaxIterations; ++it) {
      if (insideRight) {
        d1 = op<R>(b, true, kOpDistOut, hitpoint, dir, stepMax);
        const bool ok = d1 >= R(0) && d1 < big;
        snxt += ok ? (d1 + epsil) : R(0);
        hitpoint = advance<R>(hitpoint, ok ? (d1 + epsil) : R(0), dir);
        if (contains_op<R>(b, false, hitpoint)) {
          const R check = op<R>(b, false, kOpDistOut, hitpoint, dir);
          if (check > epsil) return snxt;
        }
      }
      d2 = vmax(op<R>(b, false, kOpDistIn, hitpoint, dir, stepMax), R(0));
      if (d2 == big) return big;
      d1 = op<R>(b, true, kOpDistIn, hitpoint, dir, stepMax);
      if (d2 < d1 - R(kTolerance)) {
        snxt += d2 + epsil;
        return snxt;
      }
      const bool ok = d1 >= R(0) && d1 < big;
      snxt += ok ? d1 + epsil : R(0);
      hitpoint = advance<R>(hitpoint, ok ? (d1 + epsil) : epsil, dir);
      insideRight = true;
    }
    return big;
  }

  template <typename R>
  VGM_DEV static R distance_to_out(const Data &b, V3<R> p, V3<R> dir, R stepMax)
  {
    if (b.op == kBoolUnion) {
      // BooleanUnionImplementation.h:128-179: leave the operand the point is
      // in, crossing into the other while it overlaps or connects; the
      // reference's unbounded loop stops after kBoolMaxIterations passes
      R dist = R(0);
      const R pushdist = R(kPushTolerance);
      const int positionA = inside_op<R>(b, false, p);
      V3<R> nextp = p;
      const bool first = positionA != kOutsideV;  // leave A first, else B
      bool connectingstep = false;
      int it = 0;
      do {
        connectingstep = false;
        const R disTmp = op<R>(b, !first, kOpDistOut, nextp, dir);
        dist += (disTmp >= R(0) && disTmp < inf_length<R>()) ? disTmp : R(0);
        dist += pushdist;
        nextp = advance<R>(p, dist, dir);
        if (inside_op<R>(b, first, nextp) != kOutsideV) {
          const R disTmp2 = op<R>(b, first, kOpDistOut, nextp, dir);
          dist += (disTmp2 >= R(0) && disTmp2 < inf_length<R>()) ? disTmp2 : R(0);
          dist += pushdist;
          nextp = advance<R>(p, dist, dir);
          connectingstep = true;
        }
      } while (connectingstep && (inside_op<R>(b, !first, nextp) != kOutsideV) && ++it < kBoolMaxIterations);
      R distance = dist - pushdist;
      if (distance < R(kTolerance) && positionA == kOutsideV && inside_op<R>(b, true, p) == kOutsideV)
        distance = R(-kTolerance);
      return distance;
    }
    if (b.op == kBoolIntersection)  // the left operand's DistanceToOut takes the point as given
      return vmin(op<R>(b, false, kOpDistOut, p, dir, inf_length<R>(), false),
                  op<R>(b, true, kOpDistOut, p, dir));
    const R distancel = op<R>(b, false, kOpDistOut, p, dir, stepMax);
    const R dinright = op<R>(b, true, kOpDistIn, p, dir, stepMax);
    return vmin(distancel, dinright);
  }

  template <typename R>
  VGM_DEV static R safety_to_in(const Data &b, V3<R> p)
  {
    if (b.op == kBoolUnion) {  // BooleanUnionImplementation.h:184-194
      const R distA = op<R>(b, false, kOpSafIn, p);
      const R distB = op<R>(b, true, kOpSafIn, p);
      return vmin(distA, distB);
    }
    if (b.op == kBoolIntersection) {
      const bool insideA = contains_op<R>(b, false, p);
      const bool insideB = contains_op<R>(b, true, p);
      if (!insideA && insideB) return op<R>(b, false, kOpSafIn, p);
      if (!insideB && insideA) return op<R>(b, true, kOpSafIn, p);
      return vmin(op<R>(b, false, kOpSafIn, p), op<R>(b, true, kOpSafIn, p));
    }
    if (contains_op<R>(b, false, p) && contains_op<R>(b, true, p))
      return op<R>(b, true, kOpSafOut, p);  // SafetyToOut of the transformed point
    return op<R>(b, false, kOpSafIn, p);
  }

  template <typename R>
  VGM_DEV static R safety_to_out(const Data &b, V3<R> p)
  {
    const V3<R> z{R(0), R(0), R(0)};
    if (b.op == kBoolUnion) {  // BooleanUnionImplementation.h:199-223
      const int insideA = inside_op<R>(b, false, p);
      const int insideB = inside_op<R>(b, true, p);
      if (insideA == kOutsideV && insideB == kOutsideV) return R(-kTolerance);
      if (insideA != kOutsideV && insideB != kOutsideV)
        return vmax(op<R>(b, false, kOpSafOut, p, z, inf_length<R>(), false), op<R>(b, true, kOpSafOut, p));
      if (insideA == kSurfaceV || insideB == kSurfaceV) return R(-kTolerance);
      if (insideA == kOutsideV) return op<R>(b, true, kOpSafOut, p);
      return op<R>(b, false, kOpSafOut, p, z, inf_length<R>(), false);
    }
    if (b.op == kBoolIntersection) {
      R safety = vmin(op<R>(b, false, kOpSafOut, p, z, inf_length<R>(), false), op<R>(b, true, kOpSafOut, p));
      if (safety < R(0)) safety = R(0);
      return safety;
    }
    const R safetyleft = op<R>(b, false, kOpSafOut, p, z, inf_length<R>(), false);
    const R safetyright = op<R>(b, true, kOpSafIn, p);
    return vmin(safetyleft, safetyright);
  }

  // NormalKernel of each operation: union BooleanUnionImplementation.h:228-279,
  // subtraction BooleanImplementation.h:239-288 (the subtracted operand's
  // normal negated), intersection BooleanIntersectionImplementation.h:234-274
  VGM_DEV static bool normal(const Data &b, V3<double> p, V3<double> &n)
  {
    const V3<double> z{0., 0., 0.};
    if (b.op == kBoolUnion) {
      if (contains_op<double>(b, false, p)) return normal_of_local(b, false, p, n);
      if (contains_op<double>(b, true, p)) return normal_of_local(b, true, p, n);
      const double safetyA = op<double>(b, false, kOpSafIn, p), safetyB = op<double>(b, true, kOpSafIn, p);
      if (safetyA < safetyB) return normal_of_local(b, false, p, n);
      return normal_of_local(b, true, p, n);
    }
    if (b.op == kBoolSubtraction) {
      if (contains_op<double>(b, true, p)) return normal_of_local(b, true, p, n, -1.);
      if (!contains_op<double>(b, false, p)) return normal_of_local(b, false, p, n);
      const double safetyA = op<double>(b, false, kOpSafOut, p);
      const double safetyB = op<double>(b, true, kOpSafIn, p);
      if (safetyA < safetyB) return normal_of_local(b, false, p, n);
      return normal_of_local(b, true, p, n, -1.);
    }
    const double safetyA = contains_op<double>(b, false, p) ? op<double>(b, false, kOpSafOut, p)
                                                             : op<double>(b, false, kOpSafIn, p);
    const double safetyB = contains_op<double>(b, true, p) ? op<double>(b, true, kOpSafOut, p)
                                                            : op<double>(b, true, kOpSafIn, p);
    (void)z;
    if (safetyA < safetyB) return normal_of_local(b, false, p, n);
    return normal_of_local(b, true, p, n);
  }
};

}  // namespace vgm
