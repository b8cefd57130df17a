/* FP64 CPU oracle C API — TEST INFRASTRUCTURE ONLY (tests/, smoke, bench CPU
 * baseline).  Mirrors the shape constructors and SOA3D queries of the
 * reference so that the HIP library can be checked against it. */
#ifndef ORC_ORACLE_H
#define ORC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_geom orc_geom;

enum {
  ORC_CONTAINS = 0,
  ORC_INSIDE = 1,
  ORC_DISTANCE_TO_IN = 2,
  ORC_DISTANCE_TO_OUT = 3,
  ORC_SAFETY_TO_IN = 4,
  ORC_SAFETY_TO_OUT = 5,
  ORC_PLACED_DISTANCE_TO_OUT = 6
};

orc_geom *orc_geom_create(void);
void orc_geom_destroy(orc_geom *g);

int orc_box_create(orc_geom *g, double dx, double dy, double dz);
int orc_orb_create(orc_geom *g, double r);
/* UnplacedSphere(rmin, rmax, sphi, dphi, stheta, dtheta) */
int orc_sphere_create(orc_geom *g, double rmin, double rmax, double sphi, double dphi, double stheta, double dtheta);
/* UnplacedTorus2(rmin, rmax, rtor, sphi, dphi) */
int orc_torus_create(orc_geom *g, double rmin, double rmax, double rtor, double sphi, double dphi);
/* op: 1 intersection, 2 subtraction, 3 union */
int orc_boolean_create(orc_geom *g, int op, int left, const double *ltr, int right, const double *rtr);
int orc_cut_tube_create(orc_geom *g, double rmin, double rmax, double z, double sphi, double dphi, double bx, double by,
                        double bz, double tx, double ty, double tz);
int orc_tube_create(orc_geom *g, double rmin, double rmax, double z, double sphi, double dphi);
int orc_cone_create(orc_geom *g, double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi,
                    double dphi);
int orc_trd_create(orc_geom *g, double x1, double x2, double y1, double y2, double z);
int orc_trapezoid_create(orc_geom *g, double dz, double theta, double phi, double dy1, double dx1, double dx2,
                         double alpha1, double dy2, double dx3, double dx4, double alpha2);
int orc_polycone_create(orc_geom *g, double sphi, double dphi, int nz, const double *z, const double *rmin,
                        const double *rmax);
int orc_polyhedron_create(orc_geom *g, double sphi, double dphi, int nsides, int nz, const double *z,
                          const double *rmin, const double *rmax);
/* Triangular facets (anti-clockwise from outside), fTestDir for Inside. */
int orc_tessellated_create(orc_geom *g, int nfacets, const double *triangles, const double *test_dir,
                           int brute_force);
int orc_shape_extent(orc_geom *g, int shape, double *mn, double *mx);
int orc_lv_create(orc_geom *g, int shape);
int orc_pv_create(orc_geom *g, int lv, const double *tr);
int orc_lv_add_daughter(orc_geom *g, int mother, int pv);

int orc_query(orc_geom *g, int pv, int method, const double *px, const double *py, const double *pz,
              const double *dx, const double *dy, const double *dz, const double *step, size_t n, void *out,
              int nthreads);

/* VPlacedVolume::Normal over a batch of master-frame points: the unit
 * normal of the nearest surface (master frame) and whether the point is on
 * the surface.  Returns -1 for a family without Normal (tessellated). */
int orc_normal(orc_geom *g, int pv, const double *px, const double *py, const double *pz, size_t n, double *nx,
               double *ny, double *nz, uint8_t *valid, int nthreads);

/* Navigation (navigation.cpp): states are node numbers of a depth-first
 * walk of the placement tree from the world (1 = world, 0 = outside).
 * orc_nav_close returns the number of nodes. */
int orc_nav_close(orc_geom *g, int world_pv, int use_bvh);
int orc_nav_node(orc_geom *g, unsigned node, int *pv, int *level, int *mother);
int orc_nav_locate(orc_geom *g, const double *x, const double *y, const double *z, size_t n, uint32_t *state,
                   int nthreads);
int orc_nav_step(orc_geom *g, const double *x, const double *y, const double *z, const double *dx, const double *dy,
                 const double *dz, const double *step_limit, size_t n, uint32_t *state, uint32_t *last,
                 uint8_t *boundary, double *step, int nthreads);
/* relocate = 0: VNavigator::ComputeStep, no relocation after the boundary */
int orc_nav_step2(orc_geom *g, const double *x, const double *y, const double *z, const double *dx,
                  const double *dy, const double *dz, const double *step_limit, size_t n, uint32_t *state,
                  uint32_t *last, uint8_t *boundary, double *step, int relocate, int nthreads);
int orc_nav_transport_limited(orc_geom *g, double *x, double *y, double *z, const double *dx, const double *dy,
                              const double *dz, const double *step_limit, size_t n, uint32_t *state, uint32_t *last,
                              uint8_t *boundary, int nsteps, int reuse_safety, int32_t *steps, int32_t *geo_steps,
                              int nthreads);
int orc_nav_approach(orc_geom *g, const double *x, const double *y, const double *z, const double *dx,
                     const double *dy, const double *dz, size_t n, const uint32_t *state, const uint32_t *last,
                     double *step, int nthreads);
int orc_nav_transport(orc_geom *g, double *x, double *y, double *z, const double *dx, const double *dy,
                      const double *dz, size_t n, uint32_t *state, uint32_t *last, uint8_t *boundary, int nsteps,
                      int32_t *steps, int nthreads);
/* LevelLocate in the volume of each state, directional when dx is given */
int orc_nav_level_locate(orc_geom *g, const double *x, const double *y, const double *z, const double *dx,
                         const double *dy, const double *dz, size_t n, const uint32_t *state, const uint32_t *excl,
                         uint32_t *child, double *qx, double *qy, double *qz, int nthreads);
int orc_nav_safety(orc_geom *g, const double *x, const double *y, const double *z, size_t n, const uint32_t *state,
                   double *out, int nthreads);

/* Batched track-step transport (project-defined loop, see transport.cpp). */
double orc_physics_step(double mean, uint64_t seed, uint64_t track, uint32_t step);
int orc_batch_transport(orc_geom *g, int nsolids, const int *shapes, const int64_t *offsets, const double *mean_step,
                        double *x, double *y, double *z, const double *dx, const double *dy, const double *dz,
                        int8_t *side, int32_t *crossings, int nsteps, int step0, uint64_t seed, int classify,
                        int nthreads);

#ifdef __cplusplus
}
#endif
#endif
