// VecGeom-shaped C++ adapter over the C ABI (vgm.h): the classes and method
// signatures a caller of the reference's SOA3D volume interface and
// navigator already uses, forwarding to the gfx950 kernels.  Header-only;
// link against libvgm.so.
//
//   reference                                               here
//   SOA3D<Precision> (VecGeom/base/SOA3D.h:36-41,178-183)   vgm::compat::SOA3D
//   Transformation3D (VecGeom/base/Transformation3D.h)      vgm::compat::Transformation3D
//   UnplacedBox/Tube/Cone/Trd/Trapezoid/Polycone/
//     Polyhedron/Tessellated/Orb/CutTube, LogicalVolume, PlaceDaughter  GeoManager::Box(...), ..., Place
//   VPlacedVolume SOA3D virtuals
//     (VecGeom/volumes/PlacedVolume.h:273,299,332,372,405,421)  PlacedVolume::Contains/Inside/
//                                                            DistanceToIn/DistanceToOut/SafetyToIn/
//                                                            SafetyToOut/PlacedDistanceToOut
//   VNavigator::ComputeStepsAndPropagatedStates
//     (VecGeom/navigation/VNavigator.h:128-132)             Navigator::ComputeStepsAndPropagatedStates
//   VNavigator::ComputeStep (VNavigator.h:494-540)         Navigator::ComputeSteps
//   (ComputeStepAndNextVolume, the same as Propagated)    Navigator::ComputeStepsAndNextVolumes
//   VNavigator::ComputeStepsAndSafetiesAndPropagatedStates
//     (VecGeom/navigation/VNavigator.h:136-148)             Navigator::ComputeStepsAndSafetiesAndPropagatedStates
//   NavIndexTable::ValidateState (NavIndexTable.cpp:86-140),
//     NavStatePath::At                                      Navigator::PathToState / StateToPath
//   GlobalLocator::LocateGlobalPoint (GlobalLocator.h:35)   Navigator::LocateGlobalPoints
//   VSafetyEstimator::ComputeVectorSafety (:45-80)          Navigator::ComputeSafeties
//
// Differences, by design: navigation states are NavStateIndex-style uint32
// indices held in plain arrays (NavStateIndex.h:39-41) instead of
// NavigationState objects; the reference returns sentinels and never reports
// errors, here a failing call throws vgm::compat::Error carrying
// vgm_last_error().  SOA3D arrays may live on the host or on the device
// (set `on_device`); host arrays are staged by the library.
#pragma once
#include <stdexcept>
#include <string>
#include <cmath>
#include <vector>
#include "vgm.h"

namespace vgm {
namespace compat {

using Precision = double;
using Inside_t = int;
constexpr Inside_t kInside = 1, kSurface = 2, kOutside = 3;  // VecGeom/base/Global.h:71-88

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc)
{
  if (rc != VGM_OK) throw Error(rc, vgm_last_error());
}

// Non-owning SOA view: three coordinate arrays of `size` entries.
template <typename T = Precision>
struct SOA3D {
  T *fX = nullptr, *fY = nullptr, *fZ = nullptr;
  size_t fSize = 0;
  bool on_device = false;
  SOA3D() = default;
  SOA3D(T *x, T *y, T *z, size_t size, bool device = false) : fX(x), fY(y), fZ(z), fSize(size), on_device(device) {}
  size_t size() const { return fSize; }
  T *x() const { return fX; }
  T *y() const { return fY; }
  T *z() const { return fZ; }
};

// Translation and rotation matrix in the reference's storage order.
struct Transformation3D {
  double t[3] = {0, 0, 0};
  double r[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  Transformation3D() = default;
  Transformation3D(double tx, double ty, double tz) : t{tx, ty, tz} {}
  Transformation3D(const double trans[3], const double rot[9])
  {
    for (int k = 0; k < 3; ++k) t[k] = trans[k];
    for (int k = 0; k < 9; ++k) r[k] = rot[k];
  }
  void pack(double out[12]) const
  {
    for (int k = 0; k < 3; ++k) out[k] = t[k];
    for (int k = 0; k < 9; ++k) out[3 + k] = r[k];
  }
};

class GeoManager;

class PlacedVolume {
 public:
  PlacedVolume() = default;
  int id() const { return pv_; }

  void Contains(SOA3D<> const &p, bool *out) const
  {
    std::vector<uint8_t> b(p.size());
    if (p.on_device) throw Error(VGM_EINVAL, "Contains into bool* takes host arrays; use ContainsBytes");
    check(vgm_contains(ctx_, pv_, p.x(), p.y(), p.z(), p.size(), b.data(), VGM_HOST_PTRS));
    for (size_t i = 0; i < p.size(); ++i) out[i] = b[i] != 0;
  }
  void ContainsBytes(SOA3D<> const &p, uint8_t *out) const
  {
    check(vgm_contains(ctx_, pv_, p.x(), p.y(), p.z(), p.size(), out, flags(p)));
  }
  void Inside(SOA3D<> const &p, Inside_t *out) const
  {
    check(vgm_inside(ctx_, pv_, p.x(), p.y(), p.z(), p.size(), out, flags(p)));
  }
  void DistanceToIn(SOA3D<> const &p, SOA3D<> const &d, Precision const *step_max, Precision *out) const
  {
    check(vgm_distance_to_in(ctx_, pv_, p.x(), p.y(), p.z(), d.x(), d.y(), d.z(), step_max, p.size(), out,
                             flags(p)));
  }
  // local-frame points, as VPlacedVolume::DistanceToOut (PlacedVolImplHelper.h:76-124)
  void DistanceToOut(SOA3D<> const &p, SOA3D<> const &d, Precision const *step_max, Precision *out) const
  {
    check(vgm_distance_to_out(ctx_, pv_, p.x(), p.y(), p.z(), d.x(), d.y(), d.z(), step_max, p.size(), out,
                              flags(p)));
  }
  void PlacedDistanceToOut(SOA3D<> const &p, SOA3D<> const &d, Precision const *step_max, Precision *out) const
  {
    check(vgm_placed_distance_to_out(ctx_, pv_, p.x(), p.y(), p.z(), d.x(), d.y(), d.z(), step_max, p.size(),
                                     out, flags(p)));
  }
  void SafetyToIn(SOA3D<> const &p, Precision *out) const
  {
    check(vgm_safety_to_in(ctx_, pv_, p.x(), p.y(), p.z(), p.size(), out, flags(p)));
  }
  void SafetyToOut(SOA3D<> const &p, Precision *out) const
  {
    check(vgm_safety_to_out(ctx_, pv_, p.x(), p.y(), p.z(), p.size(), out, flags(p)));
  }
  // VPlacedVolume::Normal(point, normal) -> valid (PlacedVolume.h:437-439),
  // master-frame point and normal; the batched form takes SOA3D arrays
  bool Normal(const double point[3], double normal[3]) const
  {
    uint8_t valid = 0;
    check(vgm_normal(ctx_, pv_, &point[0], &point[1], &point[2], 1, &normal[0], &normal[1], &normal[2], &valid,
                     VGM_HOST_PTRS));
    return valid != 0;
  }
  void Normal(SOA3D<> const &p, SOA3D<> &n, uint8_t *valid) const
  {
    check(vgm_normal(ctx_, pv_, p.x(), p.y(), p.z(), p.size(), n.x(), n.y(), n.z(), valid, flags(p)));
  }

 private:
  friend class GeoManager;
  PlacedVolume(vgm_ctx *c, int pv) : ctx_(c), pv_(pv) {}
  static unsigned flags(SOA3D<> const &p) { return p.on_device ? VGM_DEVICE_PTRS : VGM_HOST_PTRS; }
  vgm_ctx *ctx_ = nullptr;
  int pv_ = -1;
};

struct LogicalVolume {
  int id = -1;
};

// Owns one library context: the geometry (unplaced shapes, logical and
// placed volumes) and the stream everything runs on.
class GeoManager {
 public:
  explicit GeoManager(int device = 0, void *stream = nullptr) { check(vgm_ctx_create(device, stream, &ctx_)); }
  ~GeoManager()
  {
    if (ctx_) vgm_ctx_destroy(ctx_);
  }
  GeoManager(const GeoManager &) = delete;
  GeoManager &operator=(const GeoManager &) = delete;
  vgm_ctx *context() const { return ctx_; }

  // unplaced shapes (constructor arguments of the reference's classes)
  int Box(double dx, double dy, double dz) { return make(vgm_box_create(ctx_, dx, dy, dz, &s_)); }
  int Orb(double r) { return make(vgm_orb_create(ctx_, r, &s_)); }
  int Sphere(double rmin, double rmax, double sphi = 0., double dphi = 2 * M_PI, double stheta = 0.,
             double dtheta = M_PI)
  {
    return make(vgm_sphere_create(ctx_, rmin, rmax, sphi, dphi, stheta, dtheta, &s_));
  }
  int Torus(double rmin, double rmax, double rtor, double sphi = 0., double dphi = 2 * M_PI)
  {
    return make(vgm_torus_create(ctx_, rmin, rmax, rtor, sphi, dphi, &s_));
  }
  // UnplacedBooleanVolume<op>: op 0 union, 1 intersection, 2 subtraction; placements
  // of 12 doubles (translation, rotation) or null for the identity
  int Boolean(int op, int left, const double *left_tr, int right, const double *right_tr)
  {
    return make(vgm_boolean_create(ctx_, op, left, left_tr, right, right_tr, &s_));
  }
  int CutTube(double rmin, double rmax, double z, double sphi, double dphi, double bx, double by, double bz, double tx,
              double ty, double tz)
  {
    return make(vgm_cut_tube_create(ctx_, rmin, rmax, z, sphi, dphi, bx, by, bz, tx, ty, tz, &s_));
  }
  int Tube(double rmin, double rmax, double dz, double sphi, double dphi)
  {
    return make(vgm_tube_create(ctx_, rmin, rmax, dz, sphi, dphi, &s_));
  }
  int Cone(double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi, double dphi)
  {
    return make(vgm_cone_create(ctx_, rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi, &s_));
  }
  int Trd(double x1, double x2, double y1, double y2, double z)
  {
    return make(vgm_trd_create(ctx_, x1, x2, y1, y2, z, &s_));
  }
  int Trapezoid(double dz, double theta, double phi, double dy1, double dx1, double dx2, double alpha1, double dy2,
                double dx3, double dx4, double alpha2)
  {
    return make(vgm_trapezoid_create(ctx_, dz, theta, phi, dy1, dx1, dx2, alpha1, dy2, dx3, dx4, alpha2, &s_));
  }
  int Polycone(double sphi, double dphi, int nz, const double *z, const double *rmin, const double *rmax)
  {
    return make(vgm_polycone_create(ctx_, sphi, dphi, nz, z, rmin, rmax, &s_));
  }
  int Polyhedron(double sphi, double dphi, int nsides, int nz, const double *z, const double *rmin,
                 const double *rmax)
  {
    return make(vgm_polyhedron_create(ctx_, sphi, dphi, nsides, nz, z, rmin, rmax, &s_));
  }
  // triangles: nfacets x 9 doubles, anti-clockwise seen from outside
  int Tessellated(int nfacets, const double *triangles)
  {
    return make(vgm_tessellated_create(ctx_, nfacets, triangles, nullptr, &s_));
  }

  LogicalVolume NewLogicalVolume(int shape)
  {
    LogicalVolume lv;
    check(vgm_lv_create(ctx_, shape, &lv.id));
    return lv;
  }
  PlacedVolume Place(LogicalVolume lv, Transformation3D const &tr = Transformation3D())
  {
    double t[12];
    tr.pack(t);
    int pv = -1;
    check(vgm_pv_create(ctx_, lv.id, t, &pv));
    return PlacedVolume(ctx_, pv);
  }
  // LogicalVolume::PlaceDaughter
  PlacedVolume PlaceDaughter(LogicalVolume mother, LogicalVolume daughter,
                             Transformation3D const &tr = Transformation3D())
  {
    PlacedVolume pv = Place(daughter, tr);
    check(vgm_lv_add_daughter(ctx_, mother.id, pv.id()));
    return pv;
  }

 private:
  int make(int rc)
  {
    check(rc);
    return s_;
  }
  vgm_ctx *ctx_ = nullptr;
  int s_ = -1;
};

// Batched BVH navigation over the closed geometry under `world`.  States are
// NavStateIndex-style uint32 indices (1 = world, 0 = outside); a track also
// carries its last exited state and on-boundary flag.
class Navigator {
 public:
  Navigator(GeoManager &g, PlacedVolume world) : ctx_(g.context())
  {
    check(vgm_nav_close(ctx_, world.id(), &nnodes_));
  }
  int NumberOfStates() const { return nnodes_; }

  void LocateGlobalPoints(SOA3D<> const &p, uint32_t *state) const
  {
    check(vgm_nav_locate(ctx_, p.x(), p.y(), p.z(), p.size(), state, flags(p)));
  }
  // VNavigator::ComputeStepsAndPropagatedStates: on return state/last/boundary
  // hold the propagated states and step the step lengths.
  void ComputeStepsAndPropagatedStates(SOA3D<> const &p, SOA3D<> const &d, Precision const *step_max,
                                       uint32_t *state, uint32_t *last, uint8_t *boundary, Precision *step) const
  {
    check(vgm_nav_step(ctx_, p.x(), p.y(), p.z(), d.x(), d.y(), d.z(), step_max, p.size(), state, last, boundary,
                       step, flags(p)));
  }
  // The same call under its transport-code name: the step and, in state,
  // the next volume.
  void ComputeStepsAndNextVolumes(SOA3D<> const &p, SOA3D<> const &d, Precision const *step_max, uint32_t *state,
                                  uint32_t *last, uint8_t *boundary, Precision *step) const
  {
    check(vgm_nav_compute_step_and_next_volume(ctx_, p.x(), p.y(), p.z(), d.x(), d.y(), d.z(), step_max, p.size(),
                                               state, last, boundary, step, flags(p)));
  }
  // VNavigator::ComputeStep (VNavigator.h:494-540) per track: the step and
  // PrepareOutState's state, with no relocation after the boundary.
  void ComputeSteps(SOA3D<> const &p, SOA3D<> const &d, Precision const *step_max, uint32_t *state, uint32_t *last,
                    uint8_t *boundary, Precision *step) const
  {
    check(vgm_nav_compute_step(ctx_, p.x(), p.y(), p.z(), d.x(), d.y(), d.z(), step_max, p.size(), state, last,
                               boundary, step, flags(p)));
  }
  // ...AndSafeties...: safety[i] of the start point where calcsafety[i]
  // (calcsafety may be null: every track), else 0, then the step.
  void ComputeStepsAndSafetiesAndPropagatedStates(SOA3D<> const &p, SOA3D<> const &d, Precision const *step_max,
                                                  uint32_t *state, uint32_t *last, uint8_t *boundary,
                                                  Precision *step, bool const *calcsafety, Precision *safety) const
  {
    std::vector<uint8_t> calc;
    if (calcsafety) {
      if (p.on_device) throw Error(VGM_EINVAL, "calcsafety as bool* takes host arrays");
      calc.resize(p.size());
      for (size_t i = 0; i < p.size(); ++i) calc[i] = calcsafety[i] ? 1 : 0;
    }
    check(vgm_nav_step_safety(ctx_, p.x(), p.y(), p.z(), d.x(), d.y(), d.z(), step_max, p.size(), state, last,
                              boundary, step, calcsafety ? calc.data() : nullptr, safety, flags(p)));
  }
  // NavStatePath (placed volumes from the world down) <-> state index
  uint32_t PathToState(std::vector<int> const &path) const
  {
    uint32_t s = 0;
    check(vgm_nav_path_to_state(ctx_, path.data(), int(path.size()), &s));
    return s;
  }
  std::vector<int> StateToPath(uint32_t state) const
  {
    int pv = 0, level = 0, mother = 0;
    check(vgm_nav_node(ctx_, state, &pv, &level, &mother));
    std::vector<int> path(size_t(level) + 1);
    int depth = 0;
    check(vgm_nav_state_to_path(ctx_, state, path.data(), int(path.size()), &depth));
    return path;
  }
  // BVH::ApproachNextDaughter (source/BVH.cpp:395-449) per track: step (in and
  // out) shortened to the nearest daughter bounding box of the track's volume
  void ApproachNextDaughters(SOA3D<> const &p, SOA3D<> const &d, uint32_t const *state, uint32_t const *last,
                             Precision *step) const
  {
    check(vgm_nav_approach(ctx_, p.x(), p.y(), p.z(), d.x(), d.y(), d.z(), p.size(), state, last, step, flags(p)));
  }
  // VLevelLocator::LevelLocateExclVol (BVHLevelLocator.h:73-95) per point,
  // in the frame of the volume of state[i]: the daughter's state (0: none)
  // and the point in its frame; with a direction (dir non-null) a surface
  // point is taken only when the direction enters the daughter; excl may be
  // null
  void LevelLocate(SOA3D<> const &p, SOA3D<> const *dir, uint32_t const *state, uint32_t const *excl,
                   uint32_t *child, SOA3D<> &daughter_point) const
  {
    check(vgm_nav_level_locate(ctx_, p.x(), p.y(), p.z(), dir ? dir->x() : nullptr, dir ? dir->y() : nullptr,
                               dir ? dir->z() : nullptr, p.size(), state, excl, child, daughter_point.x(),
                               daughter_point.y(), daughter_point.z(), flags(p)));
  }
  void ComputeSafeties(SOA3D<> const &p, uint32_t const *state, Precision *safety) const
  {
    check(vgm_nav_safety(ctx_, p.x(), p.y(), p.z(), p.size(), state, safety, flags(p)));
  }
  // pv, level and mother state of a state (NavStateIndex::Top/GetLevel/Pop)
  void Decode(uint32_t state, int *pv, int *level, uint32_t *mother) const
  {
    int m = 0;
    check(vgm_nav_node(ctx_, state, pv, level, &m));
    *mother = uint32_t(m);
  }

 private:
  static unsigned flags(SOA3D<> const &p) { return p.on_device ? VGM_DEVICE_PTRS : VGM_HOST_PTRS; }
  vgm_ctx *ctx_;
  int nnodes_ = 0;
};

}  // namespace compat
}  // namespace vgm
