// Oracle geometry store (TEST INFRASTRUCTURE ONLY): shapes, logical volumes
// with daughter lists, placed volumes with their Transformation3D.
#pragma once
#include <memory>
#include <vector>
#include "core.h"

namespace orc {

struct LogicalVolume {
  int shape;
  std::vector<int> daughters;  // placed-volume ids
};

struct PlacedVolume {
  int lv = -1;
  Transform tr;
};

struct Geometry {
  std::vector<std::unique_ptr<Shape>> shapes;
  std::vector<LogicalVolume> lvs;
  std::vector<PlacedVolume> pvs;
};

}  // namespace orc
