// Phi-wedge tests used by tube, cone and polycone
// (VecGeom/volumes/Wedge_Evolution.h:200-420).
#pragma once
#include "../vgm_math.hpp"

namespace vgm {

template <typename R>
VGM_DEV bool wedge_on_surface(double ax, double ay, double nx, double ny, V3<R> p)
{
  if (!(R(ax) * p.x + R(ay) * p.y >= R(0))) return false;
  return vabs(R(nx) * p.x + R(ny) * p.y) < R(kTolerance);
}

// Wedge::GenericKernelForContainsAndInside (completelyinside, completelyoutside)
template <typename R, bool ForInside>
VGM_DEV void wedge_generic(double a1x, double a1y, double a2x, double a2y, double n1x, double n1y, double n2x,
                           double n2y, bool lt_pi, V3<R> p, bool &cin, bool &cout)
{
  const R startCheck = (-p.x * R(a1y) + p.y * R(a1x));
  const R endCheck = (-R(a2x) * p.y + R(a2y) * p.x);
  cout = startCheck < R(0);
  if (lt_pi)
    cout = cout || (endCheck < R(0));
  else
    cout = cout && (endCheck < R(0));
  if (ForInside) {
    const bool onSurface =
        wedge_on_surface<R>(a1x, a1y, n1x, n1y, p) || wedge_on_surface<R>(a2x, a2y, n2x, n2y, p);
    cout = cout && !onSurface;
    cin = !onSurface && !cout;
  }
}

// Wedge::SafetyToOut
template <typename R>
VGM_DEV R wedge_safety_to_out(double n1x, double n1y, double n2x, double n2y, bool lt_pi, V3<R> p)
{
  const R d1 = p.x * R(n1x) + p.y * R(n1y);
  const R d2 = p.x * R(n2x) + p.y * R(n2y);
  return lt_pi ? vmin(d1, d2) : vmax(d1, d2);
}

}  // namespace vgm
