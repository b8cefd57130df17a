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
  void *nav = nullptr;  // navigation tables (navigation.cpp), built by orc_nav_close
  void (*nav_deleter)(void *) = nullptr;
  Geometry() = default;
  Geometry(const Geometry &) = delete;
  Geometry &operator=(const Geometry &) = delete;
  ~Geometry()
  {
    if (nav && nav_deleter) nav_deleter(nav);
  }
};

}  // namespace orc
