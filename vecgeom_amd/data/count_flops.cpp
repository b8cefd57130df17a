// Algorithmic FLOP count of the FP64 track-step, per solid.
//
// The shape kernels are templates on their scalar type and compile for the
// host as well as for gfx950.  Instantiated here with `Counted`, a double
// that counts every +, -, *, / and sqrt it performs (comparisons, abs,
// min/max and copysign are not counted), they run the exact FP64 track-step
// of the batched transport (safety, physics step, boundary distance, carried
// safety bound) on a sample of tracks; the count divided by the track-steps
// is the algorithmic work per track-step of that solid.  Random-number and
// direction sampling are bookkeeping and are not counted.  The same sample is
// also stepped with plain double and must end on bit-identical positions, so
// the counted run is the FP64 path itself.
//
// Built as a host-only shared library (Makefile in this folder) and driven
// from make_flops.py, which writes track_step_flops.json for the benchmark's
// roofline.
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include "../csrc/geometry.h"
#include "../csrc/vgm_math.hpp"

namespace vgm {

inline uint64_t g_flops = 0;

struct Counted {
  double v;
  Counted() = default;
  constexpr Counted(double x) : v(x) {}
  explicit operator double() const { return v; }
  explicit operator float() const { return float(v); }
  explicit operator int() const { return int(v); }
  Counted &operator+=(Counted o) { ++g_flops; v = v + o.v; return *this; }
  Counted &operator-=(Counted o) { ++g_flops; v = v - o.v; return *this; }
  Counted &operator*=(Counted o) { ++g_flops; v = v * o.v; return *this; }
  Counted &operator/=(Counted o) { ++g_flops; v = v / o.v; return *this; }
};
inline Counted operator+(Counted a, Counted b) { ++g_flops; return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { ++g_flops; return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { ++g_flops; return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { ++g_flops; return Counted(a.v / b.v); }
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }
inline Counted vabs(Counted a) { return Counted(__builtin_fabs(a.v)); }
inline Counted vsqrt(Counted a) { ++g_flops; return Counted(__builtin_sqrt(a.v)); }
inline Counted vsign(Counted a) { return Counted(__builtin_copysign(1.0, a.v)); }
inline Counted vcopysign(Counted a, Counted b) { return Counted(__builtin_copysign(a.v, b.v)); }
inline Counted nonzero(Counted x) { ++g_flops; return Counted(x.v + __builtin_copysign(1.0e-30, x.v)); }
inline Counted nonzero_abs(Counted x) { ++g_flops; return Counted(__builtin_fabs(x.v) + 1.0e-30); }
template <>
VGM_DEV Counted inf_length<Counted>() { return Counted(DBL_MAX); }

}  // namespace vgm

#include "../csrc/shapes/all_shapes.hpp"
#include "../csrc/stepping.hpp"
#include "../csrc/navigation.hpp"

namespace vgm {
namespace {

template <typename R>
struct Track {
  V3<R> p, d;
  bool in;
  int32_t crossings;
  R carry;
};

// The distance-first track-step (same decisions as step(), for families with
// kSafetyBelowDistance), on scalar type R.
template <class S, typename R>
bool step_distance_first(const typename S::Data &s, Track<R> &t, R l)
{
  R db = t.in ? S::template distance_to_out<R>(s, t.p, t.d, inf_length<R>())
              : S::template distance_to_in<R>(s, t.p, t.d, inf_length<R>());
  if (db < R(0.0)) db = R(0.0);
  if (l < db) {
    t.carry = R(0.0);
    t.p.x = t.p.x + l * t.d.x;
    t.p.y = t.p.y + l * t.d.y;
    t.p.z = t.p.z + l * t.d.z;
    return false;
  }
  if (l <= db + R(kDistanceFirstWindow)) {
    if constexpr (S::kSafetyIsDistance) {
      const R saf = S::template surface_distance<R>(s, t.p, t.in);
      if (l <= saf && S::template contains<R>(s, t.p) == t.in) {
        t.carry = saf - l - carry_margin(t.p);
        t.p.x = t.p.x + l * t.d.x;
        t.p.y = t.p.y + l * t.d.y;
        t.p.z = t.p.z + l * t.d.z;
        return false;
      }
    } else {
      const R saf = t.in ? S::template safety_to_out<R>(s, t.p) : S::template safety_to_in<R>(s, t.p);
      if (l <= saf) {
        t.p.x = t.p.x + l * t.d.x;
        t.p.y = t.p.y + l * t.d.y;
        t.p.z = t.p.z + l * t.d.z;
        return false;
      }
    }
  }
  t.carry = R(0.0);
  const R m = db + R(kBoundaryPush);
  t.p.x = t.p.x + m * t.d.x;
  t.p.y = t.p.y + m * t.d.y;
  t.p.z = t.p.z + m * t.d.z;
  t.in = !t.in;
  t.crossings++;
  return true;
}

// The FP64 track-step of the batched transport kernel, on scalar type R.
template <class S, typename R>
bool step(const typename S::Data &s, Track<R> &t, R l)
{
  if constexpr (S::kSafetyIsDistance) {
    if (l <= t.carry) {
      const R m = carry_margin(t.p);
      t.p.x = t.p.x + l * t.d.x;
      t.p.y = t.p.y + l * t.d.y;
      t.p.z = t.p.z + l * t.d.z;
      t.carry = t.carry - l - m;
      return false;
    }
  }
  R saf;
  if constexpr (S::kSafetyIsDistance) {
    saf = S::template surface_distance<R>(s, t.p, t.in);
    if (l <= saf && S::template contains<R>(s, t.p) != t.in) saf = R(-1.0);
  } else {
    saf = t.in ? S::template safety_to_out<R>(s, t.p) : S::template safety_to_in<R>(s, t.p);
  }
  if (l <= saf) {
    if constexpr (S::kSafetyIsDistance) t.carry = saf - l - carry_margin(t.p);
    t.p.x = t.p.x + l * t.d.x;
    t.p.y = t.p.y + l * t.d.y;
    t.p.z = t.p.z + l * t.d.z;
    return false;
  }
  t.carry = R(0.0);
  R db = t.in ? S::template distance_to_out<R>(s, t.p, t.d, inf_length<R>())
              : S::template distance_to_in<R>(s, t.p, t.d, inf_length<R>());
  if (db < R(0.0)) db = R(0.0);
  if (l < db) {
    t.p.x = t.p.x + l * t.d.x;
    t.p.y = t.p.y + l * t.d.y;
    t.p.z = t.p.z + l * t.d.z;
    return false;
  }
  const R m = db + R(kBoundaryPush);
  t.p.x = t.p.x + m * t.d.x;
  t.p.y = t.p.y + m * t.d.y;
  t.p.z = t.p.z + m * t.d.z;
  t.in = !t.in;
  t.crossings++;
  return true;
}

template <typename R>
V3<double> dbl(const V3<R> &a)
{
  return V3<double>{double(a.x), double(a.y), double(a.z)};
}

// Steps track i of the sample through nsteps track-steps on scalar type R;
// returns the final position.  `mode` bit 0: the distance-first step where
// the family allows it (else the FP64 step); bit 1: outside moves clear of
// the bounding box skip the queries (the culled mode of the kernels).
template <class S, typename R>
V3<double> run_track(const typename S::Data &s, const ShapeRec &rec, double lambda, uint64_t seed, uint64_t i,
                     const double *p0, const double *d0, int nsteps, int mode, bool &in_out)
{
  double wc[3], wh[3], bc[3], bh[3];
  for (int k = 0; k < 3; ++k) {
    wc[k] = 0.5 * (rec.bmin[k] + rec.bmax[k]);
    wh[k] = 1.3 * 0.5 * (rec.bmax[k] - rec.bmin[k]);
    bc[k] = 0.5 * (rec.bmin[k] + rec.bmax[k]);
    bh[k] = 0.5 * (rec.bmax[k] - rec.bmin[k]);
  }
  const bool distance_first = (mode & 1) != 0;
  const bool cull = (mode & 2) != 0;
  Track<R> t;
  t.p = V3<R>{R(p0[0]), R(p0[1]), R(p0[2])};
  t.d = V3<R>{R(d0[0]), R(d0[1]), R(d0[2])};
  t.in = S::template contains<double>(s, V3<double>{p0[0], p0[1], p0[2]});
  t.crossings = 0;
  t.carry = R(0.0);
  for (int k = 0; k < nsteps; ++k) {
    const uint32_t st = uint32_t(k);
    const double l = physics_step(lambda, seed, i, st);
    // a step under the carried bound is the same in both orders (step()
    // takes it first)
    const bool carried = S::kSafetyIsDistance && R(l) <= t.carry;
    bool crossed;
    if (cull && !carried && !t.in && move_clears_box<R>(bc, bh, t.p, t.d, R(l))) {
      const R m = carry_margin(t.p);
      t.p.x = t.p.x + R(l) * t.d.x;
      t.p.y = t.p.y + R(l) * t.d.y;
      t.p.z = t.p.z + R(l) * t.d.z;
      if constexpr (S::kSafetyIsDistance) t.carry = t.carry - R(l) - m;
      crossed = false;
    } else {
      crossed = (distance_first && !carried && distance_first_ok<S, R>(s, t.p, t.in))
                    ? step_distance_first<S, R>(s, t, R(l))
                    : step<S, R>(s, t, R(l));
    }
    if (!crossed) {
      const V3<double> d = new_direction(seed, i, st, dbl(t.d));
      t.d = V3<R>{R(d.x), R(d.y), R(d.z)};
    }
    if (!t.in) {
      const V3<double> q = dbl(t.p);
      const V3<double> w{wrap(q.x, wc[0], wh[0]), wrap(q.y, wc[1], wh[1]), wrap(q.z, wc[2], wh[2])};
      if (w.x != q.x || w.y != q.y || w.z != q.z) {
        t.p = V3<R>{R(w.x), R(w.y), R(w.z)};
        t.carry = R(0.0);
      }
    }
  }
  in_out = t.in;
  return dbl(t.p);
}

struct Counter {
  const HostGeometry &g;
  const ShapeRec &rec;
  double lambda;
  uint64_t seed;
  uint64_t first;
  size_t n;
  const double *const *soa;  // x, y, z, dx, dy, dz
  int nsteps;
  int mode;
  uint64_t flops = 0;
  size_t mismatches = 0;

  template <class S>
  void operator()()
  {
    const auto &s = *reinterpret_cast<const typename S::Data *>(g.pool.data() + rec.off);
    for (size_t i = 0; i < n; ++i) {
      const double p0[3] = {soa[0][i], soa[1][i], soa[2][i]};
      const double d0[3] = {soa[3][i], soa[4][i], soa[5][i]};
      bool in_d = false, in_c = false;
      const V3<double> a = run_track<S, double>(s, rec, lambda, seed, first + i, p0, d0, nsteps, 0, in_d);
      const uint64_t f0 = g_flops;
      const V3<double> b = run_track<S, Counted>(s, rec, lambda, seed, first + i, p0, d0, nsteps, mode, in_c);
      flops += g_flops - f0;
      if (std::memcmp(&a, &b, sizeof a) != 0 || in_d != in_c) ++mismatches;
    }
  }
};

}  // namespace
}  // namespace vgm

using namespace vgm;

struct cf_geom {
  HostGeometry g;
  std::string err;
};

extern "C" {

cf_geom *cf_create() { return new cf_geom(); }
void cf_destroy(cf_geom *h) { delete h; }
const char *cf_error(cf_geom *h) { return h->err.c_str(); }

int cf_box(cf_geom *h, double dx, double dy, double dz) { return build_box(h->g, dx, dy, dz, h->err); }
int cf_tube(cf_geom *h, double rmin, double rmax, double dz, double sphi, double dphi)
{
  return build_tube(h->g, rmin, rmax, dz, sphi, dphi, h->err);
}
int cf_cone(cf_geom *h, double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi, double dphi)
{
  return build_cone(h->g, rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi, h->err);
}
int cf_trd(cf_geom *h, double x1, double x2, double y1, double y2, double z)
{
  return build_trd(h->g, x1, x2, y1, y2, z, h->err);
}
int cf_trapezoid(cf_geom *h, const double *p11) { return build_trapezoid(h->g, p11, h->err); }
int cf_polycone(cf_geom *h, double sphi, double dphi, int nz, const double *z, const double *rmin, const double *rmax)
{
  return build_polycone(h->g, sphi, dphi, nz, z, rmin, rmax, h->err);
}
int cf_polyhedron(cf_geom *h, double sphi, double dphi, int nsides, int nz, const double *z, const double *rmin,
                  const double *rmax)
{
  return build_polyhedron(h->g, sphi, dphi, nsides, nz, z, rmin, rmax, h->err);
}
int cf_tessellated(cf_geom *h, int nfacets, const double *triangles)
{
  return build_tessellated(h->g, nfacets, triangles, nullptr, h->err);
}
int cf_extent(cf_geom *h, int shape, double *mn, double *mx)
{
  if (shape < 0 || size_t(shape) >= h->g.shapes.size()) return -1;
  const ShapeRec &r = h->g.shapes[size_t(shape)];
  for (int k = 0; k < 3; ++k) {
    mn[k] = r.bmin[k];
    mx[k] = r.bmax[k];
  }
  return 0;
}

// ---- navigation ---------------------------------------------------------------
// The placement tree, as the library's vgm_lv_create / vgm_pv_create /
// vgm_lv_add_daughter build it, and the navigation tables of close_navigation.
int cf_lv(cf_geom *h, int shape)
{
  if (shape < 0 || size_t(shape) >= h->g.shapes.size()) return -1;
  h->g.lvs.push_back(LogicalRec{shape, 0, 0, -1});
  h->g.lv_daughters.emplace_back();
  return int(h->g.lvs.size()) - 1;
}
int cf_place(cf_geom *h, int lv, const double *tr)
{
  if (lv < 0 || size_t(lv) >= h->g.lvs.size()) return -1;
  PlacedRec p{};
  p.lv = lv;
  p.id = int32_t(h->g.pvs.size());
  for (int k = 0; k < 3; ++k) p.tr.t[k] = tr ? tr[k] : 0.0;
  for (int k = 0; k < 9; ++k) p.tr.r[k] = tr ? tr[3 + k] : ((k % 4 == 0) ? 1.0 : 0.0);
  h->g.pvs.push_back(p);
  return p.id;
}
int cf_add_daughter(cf_geom *h, int mother, int pv)
{
  if (mother < 0 || size_t(mother) >= h->g.lvs.size() || pv < 0 || size_t(pv) >= h->g.pvs.size()) return -1;
  h->g.lv_daughters[size_t(mother)].push_back(pv);
  return 0;
}
int cf_nav_close(cf_geom *h, int world_pv) { return close_navigation(h->g, world_pv, h->err); }

// FP64 FLOPs of the straight-line navigation transport of n tracks (located
// first; the location is not counted), up to nsteps ComputeStepAndPropagatedState
// steps each with the move of the global point, counted on the counting type
// and checked to follow the plain-double run bit for bit.  steps_out: the
// steps taken.  Returns 0, or -2 when a counted track left the FP64 path.
int cf_nav_count(cf_geom *h, size_t n, const double *x, const double *y, const double *z, const double *dx,
                 const double *dy, const double *dz, int nsteps, uint64_t *flops, uint64_t *steps_out)
{
  HostGeometry &g = h->g;
  if (g.nav_nodes.size() < 2) return -1;
  std::vector<LogicalRec> lvs = g.lvs;
  std::vector<int32_t> dl;
  for (size_t i = 0; i < lvs.size(); ++i) {
    lvs[i].dbegin = int32_t(dl.size());
    lvs[i].dcount = int32_t(g.lv_daughters[i].size());
    dl.insert(dl.end(), g.lv_daughters[i].begin(), g.lv_daughters[i].end());
  }
  const GeomView view{g.shapes.data(), g.pool.data(), lvs.data(), dl.data(), g.pvs.data(), g.nav_nodes.data(),
                      g.nav_children.data(), g.bvhs.data(), g.nav_ipool.data(), g.nav_bpool.data(), g.world_pv};
  Navigator<kAllShapes, double> nd{view};
  Navigator<kAllShapes, Counted> nc{view};
  uint64_t f = 0, st = 0;
  size_t bad = 0;
  for (size_t i = 0; i < n; ++i) {
    const V3<double> p0{x[i], y[i], z[i]}, d0{dx[i], dy[i], dz[i]};
    NavState sd{nd.locate_global(p0), 0u, false}, sc = sd;
    V3<double> pd = p0;
    V3<Counted> pc{Counted(p0.x), Counted(p0.y), Counted(p0.z)};
    const V3<Counted> dc{Counted(d0.x), Counted(d0.y), Counted(d0.z)};
    for (int k = 0; k < nsteps && sd.ind; ++k) {
      NavState od, oc;
      const double a = nd.step(pd, d0, kInfLength, sd, od);
      pd.x = pd.x + a * d0.x;
      pd.y = pd.y + a * d0.y;
      pd.z = pd.z + a * d0.z;
      sd = od;
      const uint64_t f0 = g_flops;
      const Counted b = nc.step(pc, dc, Counted(kInfLength), sc, oc);
      pc.x = pc.x + b * dc.x;
      pc.y = pc.y + b * dc.y;
      pc.z = pc.z + b * dc.z;
      sc = oc;
      f += g_flops - f0;
      ++st;
      if (double(b) != a || sc.ind != sd.ind) ++bad;
    }
    if (pd.x != double(pc.x) || pd.y != double(pc.y) || pd.z != double(pc.z)) ++bad;
  }
  *flops = f;
  *steps_out = st;
  return bad ? -2 : 0;
}

// FP64 FLOPs of `nsteps` track-steps of n sample tracks on `shape` (`mode`
// as run_track's); sample track i
// draws its random numbers as track first + i of the transport.  Returns 0,
// or -1 for a bad shape, or -2 when the counted run left the FP64 trajectory
// of some track.
int cf_count(cf_geom *h, int shape, double lambda, uint64_t seed, uint64_t first, size_t n, const double *x,
             const double *y, const double *z, const double *dx, const double *dy, const double *dz, int nsteps,
             int mode, uint64_t *flops)
{
  if (shape < 0 || size_t(shape) >= h->g.shapes.size()) return -1;
  const double *soa[6] = {x, y, z, dx, dy, dz};
  const ShapeRec &rec = h->g.shapes[size_t(shape)];
  Counter c{h->g, rec, lambda, seed, first, n, soa, nsteps, mode};
  if (!dispatch_shape_type(rec.type, c)) return -1;
  *flops = c.flops;
  return c.mismatches ? -2 : 0;
}

}  // extern "C"
