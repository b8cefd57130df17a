// Every shape family known to the kernels, and the type switch that turns a
// runtime ShapeRec into a call of the matching templated shape.
#pragma once
#include <type_traits>
#include "box.hpp"
#include "tube.hpp"
#include "cone.hpp"
#include "trd.hpp"
#include "trapezoid.hpp"
#include "polycone.hpp"
#include "polyhedron.hpp"
#include "tessellated.hpp"
#include "orb.hpp"
#include "cut_tube.hpp"
#include "boolean.hpp"
#include "sphere.hpp"
#include "torus.hpp"

namespace vgm {

// Call f.template operator()<Shape>() for the shape family of `type`.
// Returns false for an unknown type.
template <class F>
__host__ __device__ inline bool dispatch_shape_type(int type, F &&f)
{
  switch (type) {
  case kBox:
    f.template operator()<BoxShape>();
    return true;
  case kTube:
    f.template operator()<TubeShape>();
    return true;
  case kCone:
    f.template operator()<ConeShape>();
    return true;
  case kTrd:
    f.template operator()<TrdShape>();
    return true;
  case kTrapezoid:
    f.template operator()<TrapezoidShape>();
    return true;
  case kPolycone:
    f.template operator()<PolyconeShape>();
    return true;
  case kPolyhedron:
    f.template operator()<PolyhedronShape>();
    return true;
  case kTessellated:
    f.template operator()<TessellatedShape>();
    return true;
  case kOrb:
    f.template operator()<OrbShape>();
    return true;
  case kCutTube:
    f.template operator()<CutTubeShape>();
    return true;
  case kBoolean:
    f.template operator()<BooleanShape>();
    return true;
  case kSphere:
    f.template operator()<SphereShape>();
    return true;
  case kTorus:
    f.template operator()<TorusShape>();
    return true;
  default:
    return false;
  }
}

// Distance-first eligibility of a track: the distance-first step takes the
// FP64 step's decisions where the safety on the track's side cannot exceed
// its boundary distance.  A family guarantees that everywhere
// (kSafetyBelowDistance) or tells per point (distance_first_ok).
template <class S, class = void>
struct HasDistanceFirstOk : std::false_type {};
template <class S>
struct HasDistanceFirstOk<S, std::void_t<decltype(&S::template distance_first_ok<double>)>> : std::true_type {};

template <class S>
constexpr bool kDistanceFirstAny = S::kSafetyBelowDistance || HasDistanceFirstOk<S>::value;

template <class S, typename R>
VGM_DEV bool distance_first_ok(const typename S::Data &s, const V3<R> &p, bool in)
{
  if constexpr (S::kSafetyBelowDistance)
    return true;
  else if constexpr (HasDistanceFirstOk<S>::value)
    return S::template distance_first_ok<R>(s, p, in);
  else
    return false;
}

// Shape-family sets, one bit per ShapeType.
constexpr unsigned kShapeBit(int type) { return 1u << type; }
constexpr unsigned kAllShapes = (1u << kNumShapeTypes) - 1u;

// dispatch_shape_type restricted to the families of Mask: only their code is
// instantiated; a type outside the set returns false without a call.
template <unsigned Mask, class F>
__host__ __device__ inline bool dispatch_shape_mask(int type, F &&f)
{
  if (type < 0 || type >= kNumShapeTypes || !(Mask & (1u << type))) return false;
  switch (type) {
  case kBox:
    if constexpr ((Mask & (1u << kBox)) != 0) f.template operator()<BoxShape>();
    return true;
  case kTube:
    if constexpr ((Mask & (1u << kTube)) != 0) f.template operator()<TubeShape>();
    return true;
  case kCone:
    if constexpr ((Mask & (1u << kCone)) != 0) f.template operator()<ConeShape>();
    return true;
  case kTrd:
    if constexpr ((Mask & (1u << kTrd)) != 0) f.template operator()<TrdShape>();
    return true;
  case kTrapezoid:
    if constexpr ((Mask & (1u << kTrapezoid)) != 0) f.template operator()<TrapezoidShape>();
    return true;
  case kPolycone:
    if constexpr ((Mask & (1u << kPolycone)) != 0) f.template operator()<PolyconeShape>();
    return true;
  case kPolyhedron:
    if constexpr ((Mask & (1u << kPolyhedron)) != 0) f.template operator()<PolyhedronShape>();
    return true;
  case kTessellated:
    if constexpr ((Mask & (1u << kTessellated)) != 0) f.template operator()<TessellatedShape>();
    return true;
  case kOrb:
    if constexpr ((Mask & (1u << kOrb)) != 0) f.template operator()<OrbShape>();
    return true;
  case kCutTube:
    if constexpr ((Mask & (1u << kCutTube)) != 0) f.template operator()<CutTubeShape>();
    return true;
  case kBoolean:
    if constexpr ((Mask & (1u << kBoolean)) != 0) f.template operator()<BooleanShape>();
    return true;
  case kSphere:
    if constexpr ((Mask & (1u << kSphere)) != 0) f.template operator()<SphereShape>();
    return true;
  case kTorus:
    if constexpr ((Mask & (1u << kTorus)) != 0) f.template operator()<TorusShape>();
    return true;
  default:
    return false;
  }
}

}  // namespace vgm
