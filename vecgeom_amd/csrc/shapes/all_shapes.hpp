// Every shape family known to the kernels, and the type switch that turns a
// runtime ShapeRec into a call of the matching templated shape.
#pragma once
#include "box.hpp"
#include "tube.hpp"
#include "cone.hpp"
#include "trd.hpp"
#include "trapezoid.hpp"
#include "polycone.hpp"
#include "polyhedron.hpp"
#include "tessellated.hpp"

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
  default:
    return false;
  }
}

}  // namespace vgm
