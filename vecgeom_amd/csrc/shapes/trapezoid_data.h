// Trapezoid parameter block: half-length in z plus the four side planes of
// PlaneShell<4> (VecGeom/base/PlaneShell.h), a*x + b*y + c*z + d <= 0 inside.
#pragma once

namespace vgm {

struct TrapezoidData {
  double dz;
  double a[4], b[4], c[4], d[4];
};

}  // namespace vgm
