// Oracle C API (TEST INFRASTRUCTURE ONLY): geometry store + SOA3D query loops
// mirroring the placed-volume SOA3D interface
// (VecGeom/volumes/SpecializedPlacedVolImplHelper.h:184-257,
//  VecGeom/volumes/PlacedVolImplHelper.h:76-124).
#include <omp.h>
#include <memory>
#include "oracle.h"
#include "geom.h"

using namespace orc;

extern "C" {

orc_geom *orc_geom_create(void) { return reinterpret_cast<orc_geom *>(new Geometry()); }
void orc_geom_destroy(orc_geom *g) { delete reinterpret_cast<Geometry *>(g); }

static int add_shape(orc_geom *g, Shape *s)
{
  auto *G = reinterpret_cast<Geometry *>(g);
  if (!s) return -1;
  G->shapes.emplace_back(s);
  return int(G->shapes.size()) - 1;
}

int orc_box_create(orc_geom *g, double dx, double dy, double dz) { return add_shape(g, make_box(dx, dy, dz)); }
int orc_boolean_create(orc_geom *g, int op, int left, const double *ltr, int right, const double *rtr)
{
  auto *G = reinterpret_cast<Geometry *>(g);
  const int n = int(G->shapes.size());
  if ((op < 1 || op > 3) || left < 0 || left >= n || right < 0 || right >= n) return -1;
  return add_shape(g, make_boolean(op, G->shapes[left].get(), ltr, G->shapes[right].get(), rtr));
}
int orc_orb_create(orc_geom *g, double r) { return r > 0 ? add_shape(g, make_orb(r)) : -1; }
int orc_sphere_create(orc_geom *g, double rmin, double rmax, double sphi, double dphi, double stheta, double dtheta)
{
  if (!(rmin >= 0 && rmin < rmax && dphi > 0 && stheta >= 0 && stheta <= kPi && dtheta > 0)) return -1;
  return add_shape(g, make_sphere(rmin, rmax, sphi, dphi, stheta, dtheta));
}
int orc_torus_create(orc_geom *g, double rmin, double rmax, double rtor, double sphi, double dphi)
{
  if (!(rmin >= 0 && rmin < rmax && rtor > 0 && dphi > 0)) return -1;
  return add_shape(g, make_torus(rmin, rmax, rtor, sphi, dphi));
}
int orc_cut_tube_create(orc_geom *g, double rmin, double rmax, double z, double sphi, double dphi, double bx, double by,
                        double bz, double tx, double ty, double tz)
{
  if (!(rmin >= 0 && rmin < rmax && z > 0 && dphi > 0) || bz >= 0 || tz <= 0) return -1;
  const double bot[3] = {bx, by, bz}, top[3] = {tx, ty, tz};
  return add_shape(g, make_cut_tube(rmin, rmax, z, sphi, dphi, bot, top));
}

int orc_tube_create(orc_geom *g, double rmin, double rmax, double z, double sphi, double dphi)
{
  return add_shape(g, make_tube(rmin, rmax, z, sphi, dphi));
}

int orc_cone_create(orc_geom *g, double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi,
                    double dphi)
{
  return add_shape(g, make_cone(rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi));
}

int orc_trd_create(orc_geom *g, double x1, double x2, double y1, double y2, double z)
{
  return add_shape(g, make_trd(x1, x2, y1, y2, z));
}

int orc_trapezoid_create(orc_geom *g, double dz, double theta, double phi, double dy1, double dx1, double dx2,
                         double alpha1, double dy2, double dx3, double dx4, double alpha2)
{
  return add_shape(g, make_trapezoid(dz, theta, phi, dy1, dx1, dx2, alpha1, dy2, dx3, dx4, alpha2));
}

int orc_polycone_create(orc_geom *g, double sphi, double dphi, int nz, const double *z, const double *rmin,
                        const double *rmax)
{
  return add_shape(g, make_polycone(sphi, dphi, nz, z, rmin, rmax));
}

int orc_polyhedron_create(orc_geom *g, double sphi, double dphi, int nsides, int nz, const double *z,
                          const double *rmin, const double *rmax)
{
  return add_shape(g, make_polyhedron(sphi, dphi, nsides, nz, z, rmin, rmax));
}

int orc_tessellated_create(orc_geom *g, int nfacets, const double *triangles, const double *test_dir,
                           int brute_force)
{
  return add_shape(g, make_tessellated(nfacets, triangles, test_dir, brute_force != 0));
}

int orc_shape_extent(orc_geom *g, int shape, double *mn, double *mx)
{
  auto *G = reinterpret_cast<Geometry *>(g);
  Vec3 a, b;
  G->shapes[shape]->Extent(a, b);
  for (int k = 0; k < 3; ++k) {
    mn[k] = a[k];
    mx[k] = b[k];
  }
  return 0;
}

int orc_lv_create(orc_geom *g, int shape)
{
  auto *G = reinterpret_cast<Geometry *>(g);
  G->lvs.push_back(LogicalVolume{shape, {}});
  return int(G->lvs.size()) - 1;
}

int orc_pv_create(orc_geom *g, int lv, const double *tr)
{
  auto *G = reinterpret_cast<Geometry *>(g);
  PlacedVolume p;
  p.lv = lv;
  if (tr) {
    for (int k = 0; k < 3; ++k) p.tr.t[k] = tr[k];
    for (int k = 0; k < 9; ++k) p.tr.r[k] = tr[3 + k];
  }
  G->pvs.push_back(p);
  return int(G->pvs.size()) - 1;
}

int orc_lv_add_daughter(orc_geom *g, int mother, int pv)
{
  auto *G = reinterpret_cast<Geometry *>(g);
  G->lvs[mother].daughters.push_back(pv);
  return 0;
}

int orc_query(orc_geom *g, int pv, int method, const double *px, const double *py, const double *pz,
              const double *dx, const double *dy, const double *dz, const double *step, size_t n, void *out,
              int nthreads)
{
  auto *G = reinterpret_cast<Geometry *>(g);
  const PlacedVolume &P = G->pvs[pv];
  const Shape &S = *G->shapes[G->lvs[P.lv].shape];
  const Transform &T = P.tr;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (long long i = 0; i < (long long)n; ++i) {
    const Vec3 m(px[i], py[i], pz[i]);
    switch (method) {
    case ORC_CONTAINS: static_cast<uint8_t *>(out)[i] = S.Contains(T.Transform_(m)) ? 1 : 0; break;
    case ORC_INSIDE: static_cast<int32_t *>(out)[i] = S.Inside(T.Transform_(m)); break;
    case ORC_SAFETY_TO_IN: static_cast<double *>(out)[i] = S.SafetyToIn(T.Transform_(m)); break;
    case ORC_SAFETY_TO_OUT: static_cast<double *>(out)[i] = S.SafetyToOut(m); break;
    case ORC_DISTANCE_TO_IN: {
      const Vec3 d(dx[i], dy[i], dz[i]);
      static_cast<double *>(out)[i] =
          S.DistanceToIn(T.Transform_(m), T.TransformDirection(d), step ? step[i] : kInfLength);
      break;
    }
    case ORC_DISTANCE_TO_OUT: {
      const Vec3 d(dx[i], dy[i], dz[i]);
      static_cast<double *>(out)[i] = S.DistanceToOut(m, d, step ? step[i] : kInfLength);
      break;
    }
    case ORC_PLACED_DISTANCE_TO_OUT: {
      const Vec3 d(dx[i], dy[i], dz[i]);
      static_cast<double *>(out)[i] =
          S.DistanceToOut(T.Transform_(m), T.TransformDirection(d), step ? step[i] : kInfLength);
      break;
    }
    default: break;
    }
  }
  return 0;
}

// VPlacedVolume::Normal (source/PlacedVolume.cpp:126-137): master-frame
// point, unplaced normal in the local frame, direction back to the master.
int orc_normal(orc_geom *g, int pv, const double *px, const double *py, const double *pz, size_t n, double *nx,
               double *ny, double *nz, uint8_t *valid, int nthreads)
{
  auto *G = reinterpret_cast<Geometry *>(g);
  const PlacedVolume &P = G->pvs[pv];
  const Shape &S = *G->shapes[G->lvs[P.lv].shape];
  if (!S.HasNormal()) return -1;
  const Transform &T = P.tr;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (long long i = 0; i < (long long)n; ++i) {
    Vec3 ln;
    const bool ok = S.Normal(T.Transform_(Vec3(px[i], py[i], pz[i])), ln);
    const Vec3 m = T.InverseTransformDirection(ln);
    nx[i] = m.x;
    ny[i] = m.y;
    nz[i] = m.z;
    valid[i] = ok ? 1 : 0;
  }
  return 0;
}

}  // extern "C"
