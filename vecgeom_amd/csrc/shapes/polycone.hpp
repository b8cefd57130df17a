// Polycone kernels: z-ordered cone sections with the reference's section
// walks (VecGeom/volumes/kernel/PolyconeImplementation.h:66-515).  Section
// lookup is the same linear scan over the z planes (PolyconeStruct::
// GetSectionIndex); the tables are wave-uniform and stay in the scalar cache.
#pragma once
#include "cone.hpp"
#include "polycone_data.h"

namespace vgm {

struct PolyconeShape {
  using Data = PolyconeData;
  static constexpr bool kStage = true;  // parameter block staged in LDS
  // waves per SIMD the transport kernel asks the compiler for: the latency
  // of the section walks hides better at 6 waves (80 VGPRs) than at the 3
  // its unconstrained register allocation allows, although some values then
  // spill to scratch (50 steps of 5M tracks, FP64 step: 41.5 ms at 4 waves,
  // 37.2 at 5, 34.9 at 6)
  static constexpr int kWavesPerEU = 6;
  // safety is the distance to the surface (1-Lipschitz): carried bounds are valid
  static constexpr bool kSafetyIsDistance = false;
  // SafetyToIn can exceed DistanceToIn outside the solid beyond its
  // z-range, where the reference takes the end section's safety alone
  // (PolyconeImplementation.h:388-403); distance_first_ok tells the points
  // where the distance-first step keeps the FP64 decisions
  static constexpr bool kSafetyBelowDistance = false;
  // mixed precision: Inside/Contains classified in float away from the surface
  static constexpr bool kFloatClassify = true;
  // mixed precision: the float safety less the guard band bounds the FP64 safety
  static constexpr bool kFloatSafety = false;
  // mixed precision: distance_to_in_mixed / distance_to_out_mixed (float selects
  // the surface, FP64 evaluates it)
  static constexpr bool kMixedDistance = false;

  // GetSectionIndex: linear search over the z-planes, compared in the
  // arithmetic type of the query
  template <typename R>
  VGM_DEV static int section_index(const Data &d, R zp)
  {
    const double *zs = d.zs();
    if (zp < R(zs[0])) return -1;
    for (int i = 0; i < d.nsec; ++i)
      if (zp >= R(zs[i]) && zp <= R(zs[i + 1])) return i;
    return -2;
  }

  // Inside the solid, and outside it within its z-range, the safety is the
  // smallest section safety of the sections within reach, each a lower
  // bound of the distance to its section: then it cannot exceed the
  // boundary distance and the distance-first step takes the FP64 decisions
  template <typename R>
  VGM_DEV static bool distance_first_ok(const Data &d, V3<R> p, bool in)
  {
    return in || section_index<R>(d, p.z) >= 0;
  }

  template <typename R>
  VGM_DEV static V3<R> shifted(const Data &d, int i, V3<R> p)
  {
    return V3<R>{p.x - R(0), p.y - R(0), p.z - R(d.shift()[i])};
  }

  template <typename R, bool ForInside>
  VGM_DEV static void section_generic(const Data &d, int i, V3<R> p, bool &in, bool &out)
  {
    if (i < 0) {
      in = false;
      out = true;
      return;
    }
    ConeShape::generic<R, ForInside>(d.sec()[i], shifted<R>(d, i, p), in, out);
  }

  template <typename R, bool ForInside>
  VGM_DEV static void generic(const Data &d, V3<R> p, bool &cin, bool &cout)
  {
    const int lo = section_index(d, p.z - R(kTolerance));
    const int hi = section_index(d, p.z + R(kTolerance));
    const int last = d.nsec - 1;
    if (lo < 0 && hi < 0) {
      cout = true;
      return;
    }
    if (lo < 0 && hi == 0) {
      section_generic<R, ForInside>(d, 0, p, cin, cout);
      return;
    }
    if (hi < 0 && lo == last) {
      section_generic<R, ForInside>(d, last, p, cin, cout);
      return;
    }
    if (lo >= 0 && hi >= 0) {
      if (lo == hi) {
        section_generic<R, ForInside>(d, lo, p, cin, cout);
        return;
      }
      bool inLo = false, outLo = false, inHi = false, outHi = false;
      section_generic<R, ForInside>(d, lo, p, inLo, outLo);
      section_generic<R, ForInside>(d, hi, p, inHi, outHi);
      if (!inLo && !outLo && !inHi && !outHi) {
        cin = true;
        return;
      }
      if (outLo && outHi) cout = true;
    }
  }

  template <typename R>
  VGM_DEV static bool contains(const Data &d, V3<R> p)
  {
    bool ci = false, co = false;
    generic<R, false>(d, p, ci, co);
    return !co;
  }

  template <typename R>
  VGM_DEV static int inside(const Data &d, V3<R> p)
  {
    bool ci = false, co = false;
    generic<R, true>(d, p, ci, co);
    return ci ? kInsideV : (co ? kOutsideV : kSurfaceV);
  }

  template <typename R>
  VGM_DEV static R distance_to_in(const Data &d, V3<R> p, V3<R> v, R step_max)
  {
    R distance = inf_length<R>();
    int increment = (v.z > R(0)) ? 1 : -1;
    if (vabs(v.z) < R(kTolerance)) increment = 0;
    int index = section_index(d, p.z);
    if (index == -1) index = 0;
    if (index == -2) index = d.nsec - 1;
    do {
      distance = ConeShape::distance_to_in<R>(d.sec()[index], shifted<R>(d, index, p), v, step_max);
      if (distance < inf_length<R>() || !increment) break;
      index += increment;
    } while (index >= 0 && index < d.nsec);
    return distance;
  }

  template <typename R>
  VGM_DEV static int cone_inside(const Data &d, int i, V3<R> p)
  {
    return ConeShape::inside<R>(d.sec()[i], p);
  }

  template <typename R>
  VGM_DEV static R distance_to_out(const Data &d, V3<R> point, V3<R> dir, R step_max)
  {
    if (d.nsec == 1) return ConeShape::distance_to_out<R>(d.sec()[0], shifted<R>(d, 0, point), dir, step_max);
    const int lo = section_index(d, point.z - R(kTolerance));
    const int hi = section_index(d, point.z + R(kTolerance));
    int index = 0;
    if (lo < 0 && hi < 0) return R(-1);
    if (lo < 0 && hi >= 0) {
      index = hi;
      if (cone_inside<R>(d, index, shifted<R>(d, index, point)) == kOutsideV) return R(-1);
    } else if (lo != hi && lo >= 0) {
      index = (cone_inside<R>(d, lo, shifted<R>(d, lo, point)) == kOutsideV) ? hi : lo;
    } else {
      index = lo;
      if (index < 0) index = section_index(d, point.z);
    }
    if (index < 0) return R(0);
    if (cone_inside<R>(d, index, shifted<R>(d, index, point)) == kOutsideV) return R(-1);
    R totalDistance = R(0);
    int increment = (dir.z > R(0)) ? 1 : -1;
    if (vabs(dir.z) < R(kTolerance)) increment = 0;
    int istep = 0;
    V3<R> pn = point;
    do {
      const double sh = d.shift()[index];
      if ((totalDistance != R(0)) || (istep < 2)) {
        pn.x = point.x + totalDistance * dir.x;
        pn.y = point.y + totalDistance * dir.y;
        pn.z = point.z + totalDistance * dir.z;
        pn.z -= R(sh);
        if (cone_inside<R>(d, index, pn) == kOutsideV) break;
      } else {
        pn.z -= R(sh);
      }
      istep++;
      const R dist = ConeShape::distance_to_out<R>(d.sec()[index], pn, dir, step_max);
      if (dist == R(-1)) return inf_length<R>();
      if (vabs(dist) < R(0.5 * kTolerance)) {
        int index1 = index;
        if ((index > 0) && (index < d.nsec - 1)) {
          index1 += increment;
        } else {
          if ((index == 0) && (increment > 0)) index1 += increment;
          if ((index == d.nsec - 1) && (increment < 0)) index1 += increment;
        }
        const R t = totalDistance + dist;
        V3<R> pte{point.x + t * dir.x, point.y + t * dir.y, point.z + t * dir.z};
        pte.z -= R(d.shift()[index1]);
        if (cone_inside<R>(d, index1, pte) == kOutsideV || increment == 0) break;
      }
      totalDistance += dist;
      index += increment;
    } while (increment != 0 && index >= 0 && index < d.nsec);
    return totalDistance;
  }

  template <typename R>
  VGM_DEV static R neighbour_safety(const Data &d, int index, V3<R> p, R minSafety)
  {
    const double *zs = d.zs();
    double zbase = zs[index + 1];
    for (int i = index + 1; i < d.nsec; ++i) {
      const R dz = R(zs[i] - zbase);
      if (dz >= minSafety) break;
      const R s = ConeShape::safety_to_in<R>(d.sec()[i], shifted<R>(d, i, p));
      if (s < minSafety) minSafety = s;
    }
    if (index > 0) {
      zbase = zs[index - 1];
      for (int i = index - 1; i >= 0; --i) {
        const R dz = R(zbase - zs[i]);
        if (dz >= minSafety) break;
        const R s = ConeShape::safety_to_in<R>(d.sec()[i], shifted<R>(d, i, p));
        if (s < minSafety) minSafety = s;
      }
    }
    return minSafety;
  }

  template <typename R>
  VGM_DEV static R safety_to_in(const Data &d, V3<R> p)
  {
    int index = section_index(d, p.z);
    bool needZ = false;
    if (index < 0) {
      needZ = true;
      index = (index == -1) ? 0 : d.nsec - 1;
    }
    const R safety = ConeShape::safety_to_in<R>(d.sec()[index], shifted<R>(d, index, p));
    if (needZ || safety < R(kTolerance)) return safety;
    return neighbour_safety<R>(d, index, p, safety);
  }

  template <typename R>
  VGM_DEV static R safety_to_out(const Data &d, V3<R> point)
  {
    bool ci = false, co = false;
    generic<R, true>(d, point, ci, co);
    if (co) return R(-1);
    const int index = section_index(d, point.z);
    if (index < 0) return R(-1);
    const R safety = ConeShape::safety_to_out<R>(d.sec()[index], shifted<R>(d, index, point));
    if (safety == inf_length<R>()) return R(0);
    if (safety < R(kTolerance)) return R(0);
    return neighbour_safety<R>(d, index, point, safety);
  }

  // UnplacedPolycone::Normal (source/UnplacedPolycone.cpp:782-832): the
  // section's cone normal; on the plane between two sections the next
  // section's too
  VGM_DEV static bool normal(const Data &d, V3<double> point, V3<double> &norm)
  {
    bool valid = true;
    const int index = section_index<double>(d, point.z - kTolerance);
    if (index < 0) {
      valid = false;
      if (index == -1) norm = V3<double>{0., 0., -1.};
      if (index == -2) norm = V3<double>{0., 0., 1.};
      return valid;
    }
    valid = ConeShape::normal(d.sec()[index], shifted<double>(d, index, point), norm);
    if (index + 1 < d.nsec && vabs(point.z - d.zs()[index + 1]) < kTolerance) {
      V3<double> norm2{0., 0., 0.};
      const bool valid2 = ConeShape::normal(d.sec()[index + 1], shifted<double>(d, index + 1, point), norm2);
      if (!valid && valid2) {
        norm = norm2;
        valid = valid2;
      }
      if (valid && valid2) {
        const double s = kTolerance * 10;
        const V3<double> q{point.x + s * norm2.x, point.y + s * norm2.y, point.z + s * norm2.z};
        const bool c2 = ConeShape::contains<double>(d.sec()[index + 1], shifted<double>(d, index + 1, q));
        if (c2) {
          norm = norm2;
        } else {
          norm = V3<double>{norm.x + norm2.x, norm.y + norm2.y, norm.z + norm2.z};
          if (v3_mag2(norm) < kTolerance) norm = V3<double>{point.x, point.y, 0};
        }
      }
    }
    if (valid) {
      const double m = __builtin_sqrt(v3_mag2(norm));
      norm = V3<double>{norm.x / m, norm.y / m, norm.z / m};
    }
    return valid;
  }
};

}  // namespace vgm
