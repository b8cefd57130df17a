// Trd parameter block: TrdStruct cached values (VecGeom/volumes/TrdStruct.h).
#pragma once
#include <stdint.h>

namespace vgm {

struct TrdData {
  double dx1, dx2, dy1, dy2, dz;
  double x2minusx1, y2minusy1, half_x1plusx2, half_y1plusy2;
  double calfx, calfy, tolx, toly, fx, fy;
  int32_t vary_y;  // dy1 != dy2
  int32_t pad;
};

}  // namespace vgm
