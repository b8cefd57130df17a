// Host-side geometry model: the flat tables the device consumes, plus the
// shape builders that derive each family's cached parameters from its
// constructor arguments.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "vgm_types.h"
#include "nav_data.h"

namespace vgm {

struct HostGeometry {
  std::vector<ShapeRec> shapes;
  std::vector<double> pool;  // parameter blocks, 8-byte granular
  std::vector<LogicalRec> lvs;
  std::vector<std::vector<int32_t>> lv_daughters;  // placed-volume ids per logical volume
  std::vector<PlacedRec> pvs;
  // navigation tables, built by close_navigation
  int world_pv = -1;
  std::vector<NavNode> nav_nodes;      // [0] unused: node 0 is "outside"
  std::vector<int32_t> nav_children;
  std::vector<BvhRec> bvhs;
  std::vector<int32_t> nav_ipool;
  std::vector<double> nav_bpool;
  uint32_t nav_shape_mask = 0;  // shape families of the navigation tree (bit per ShapeType)
  uint64_t version = 1;  // bumped on every change; the device copy follows it

  // Appends a parameter block of `ndoubles` and registers a shape; returns its id.
  int add_shape(int type, const void *block, size_t ndoubles, int n0, int n1, const double bmin[3],
                const double bmax[3]);
};

struct ConeData;
struct TubeData;
// TubeStruct cached values (tube builder, and the polyhedron's bounding tube).
void make_tube_data(TubeData &t, double rmin, double rmax, double z, double sphi, double dphi);
// Cone parameter derivation shared by the cone and polycone builders.
bool make_cone_data(ConeData &c, double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi,
                    double dphi, std::string &err);

// Navigation tables for the tree under `world_pv`; returns the node count.
int close_navigation(HostGeometry &g, int world_pv, std::string &err);

// Shape builders.  Return the new shape id, or -1 with `err` set.
int build_box(HostGeometry &g, double dx, double dy, double dz, std::string &err);
int build_orb(HostGeometry &g, double r, std::string &err);
int build_torus(HostGeometry &g, double rmin, double rmax, double rtor, double sphi, double dphi, std::string &err);
int build_sphere(HostGeometry &g, double rmin, double rmax, double sphi, double dphi, double stheta, double dtheta,
                 std::string &err);
int build_boolean(HostGeometry &g, int op, int left, const double *ltr, int right, const double *rtr,
                  std::string &err);
int build_cut_tube(HostGeometry &g, double rmin, double rmax, double z, double sphi, double dphi, const double bot[3],
                   const double top[3], std::string &err);
int build_cone(HostGeometry &g, double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi,
               double dphi, std::string &err);
int build_trd(HostGeometry &g, double x1, double x2, double y1, double y2, double z, std::string &err);
int build_trapezoid(HostGeometry &g, const double p[11], std::string &err);
int build_polycone(HostGeometry &g, double sphi, double dphi, int nz, const double *z, const double *rmin,
                   const double *rmax, std::string &err);
int build_tube(HostGeometry &g, double rmin, double rmax, double z, double sphi, double dphi, std::string &err);
// triangles: nfacets x 9 doubles; test_dir may be NULL (library default).
int build_tessellated(HostGeometry &g, int nfacets, const double *triangles, const double *test_dir,
                      std::string &err);
int build_polyhedron(HostGeometry &g, double sphi, double dphi, int nsides, int nz, const double *z,
                     const double *rmin, const double *rmax, std::string &err);

}  // namespace vgm
