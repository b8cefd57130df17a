// Navigation on the device: one lane per track, FP64.
//   ComputeStepAndPropagatedState of VNavigatorHelper<BVHNavigator<false>>
//     (VecGeom/navigation/VNavigator.h:436-492): global -> local with the
//     state's cached matrix, DistanceToOut of the mother clamped at 0
//     (TreatDistanceToMother :241-249), daughters through the BVH
//     (BVH::CheckDaughterIntersections, VecGeom/base/BVH.h:96-146),
//     PrepareOutState (:192-235), and relocation of the point pushed 1e-6 mm
//     past the boundary (MovePointAfterBoundary :168-173, Relocate :1070-1091)
//     with GlobalLocator::LocateGlobalPoint / LocateGlobalPointExclVolume /
//     RelocatePointFromPathForceDifferent (VecGeom/navigation/GlobalLocator.h:35-200)
//     and BVH::LevelLocate (VecGeom/base/BVH.h:219-280).
//   Safety: BVHSafetyEstimator::ComputeSafetyForLocalPoint
//     (VecGeom/navigation/BVHSafetyEstimator.h:39-47) with BVH::ComputeSafety
//     (VecGeom/base/BVH.h:172-209).
// Shapes are reached through a dispatcher over the families the geometry
// uses (a template mask): inlined for small sets such as boxes and tubes,
// one out-of-line function for the set of all families.
#pragma once
#include <type_traits>
#include "vgm_math.hpp"
#include "transform.hpp"
#include "nav_data.h"
#include "shapes/all_shapes.hpp"

namespace vgm {

struct GeomView {
  const ShapeRec *shapes;
  const double *pool;
  const LogicalRec *lvs;
  const int32_t *daughters;
  const PlacedRec *pvs;
  const NavNode *nodes;
  const int32_t *children;
  const BvhRec *bvhs;
  const int32_t *ipool;
  const double *bpool;
  int32_t world_pv;
};

// kQInside returns the Inside code; kQNormalDot the dot product of the
// unplaced Normal at p with d (FP64 only), the entering test of the
// directional LevelLocate.
enum ShapeQuery : int32_t {
  kQContains = 0,
  kQDistIn = 1,
  kQDistOut = 2,
  kQSafIn = 3,
  kQSafOut = 4,
  kQInside = 5,
  kQNormalDot = 6
};

template <typename R = double>
struct ShapeCall {
  const double *params;
  int query;
  V3<R> p, d;
  R step;
  R result;
  template <class S>
  __host__ __device__ void operator()()
  {
    const auto &s = *reinterpret_cast<const typename S::Data *>(params);
    switch (query) {
    case kQContains: result = S::template contains<R>(s, p) ? R(1.0) : R(0.0); break;
    case kQDistIn: result = S::template distance_to_in<R>(s, p, d, step); break;
    case kQDistOut: result = S::template distance_to_out<R>(s, p, d, step); break;
    case kQSafIn: result = S::template safety_to_in<R>(s, p); break;
    case kQInside: result = R(double(S::template inside<R>(s, p))); break;
    case kQNormalDot:
      if constexpr (std::is_same_v<R, double>) {
        V3<double> n{0.0, 0.0, 0.0};
        S::normal(s, p, n);
        result = n.x * d.x + n.y * d.y + n.z * d.z;  // Vector3D::Dot
      } else {
        result = R(0.0);
      }
      break;
    default: result = S::template safety_to_out<R>(s, p); break;
    }
  }
};

template <unsigned Mask, typename R = double>
__host__ __device__ __forceinline__ R shape_query_inline(const ShapeRec *shapes, const double *pool, int shape,
                                                         int query, V3<R> p, V3<R> d, R step)
{
  const ShapeRec &r = shapes[shape];
  ShapeCall<R> c{pool + r.off, query, p, d, step, R(0.0)};
  dispatch_shape_mask<Mask>(r.type, c);
  return c.result;
}

static __host__ __device__ __noinline__ double shape_query_all(const ShapeRec *shapes, const double *pool, int shape,
                                                               int query, V3<double> p, V3<double> d, double step)
{
  return shape_query_inline<kAllShapes>(shapes, pool, shape, query, p, d, step);
}

// Shapes of the families in Mask; FP64 calls over every family go through
// one out-of-line function, other scalar types (the host FLOP counter) inline.
template <unsigned Mask, typename R = double>
__host__ __device__ __forceinline__ R shape_query(const ShapeRec *shapes, const double *pool, int shape, int query,
                                                  V3<R> p, V3<R> d, R step)
{
  if constexpr (Mask == kAllShapes && std::is_same_v<R, double>)
    return shape_query_all(shapes, pool, shape, query, p, d, step);
  else
    return shape_query_inline<Mask, R>(shapes, pool, shape, query, p, d, step);
}

// BVH walk stack.  On the device it lives in LDS, one column of 16-bit node
// indices per thread of the workgroup (a private array is kept in registers
// with indexed moves, 32 of them per thread); on the host it is a local
// array.  BVHs have at most 2^15 - 1 nodes (checked when they are built) and
// a walk holds at most depth + 1 < kNavWalkDepth pending nodes.  The walks
// (daughter intersections, level locate, daughter safety) never nest.
constexpr int kNavWalkDepth = 32;
constexpr int kNavWalkThreads = 128;  // workgroup of the navigation kernels
#if defined(__HIP_DEVICE_COMPILE__)
static __shared__ uint16_t nav_walk_lds[kNavWalkDepth * kNavWalkThreads];
struct NavWalkStack {
  uint16_t *col;
  __device__ NavWalkStack() : col(nav_walk_lds + threadIdx.x) {}
  __device__ void set(int i, unsigned v) { col[i * kNavWalkThreads] = uint16_t(v); }
  __device__ unsigned get(int i) const { return col[i * kNavWalkThreads]; }
};
#else
struct NavWalkStack {
  unsigned a[kNavWalkDepth];
  void set(int i, unsigned v) { a[i] = v; }
  unsigned get(int i) const { return a[i]; }
};
#endif

// R: the scalar type of the arithmetic (double on the device; the host FLOP
// counter instantiates a counting type)
template <unsigned Mask = kAllShapes, typename R = double>
struct Navigator {
  const GeomView &g;

  __host__ __device__ int shape_of_pv(int pv) const { return g.lvs[g.pvs[pv].lv].shape; }

  // ---- AABB (VecGeom/base/AABB.h) ---------------------------------------------
  __host__ __device__ static bool box_contains(const double *b, V3<R> p)
  {
    return p.x >= b[0] && p.x <= b[3] && p.y >= b[1] && p.y <= b[4] && p.z >= b[2] && p.z <= b[5];
  }
  __host__ __device__ static void box_intersect(const double *b, V3<R> p, V3<R> inv, R &tmin, R &tmax)
  {
    R a0 = (b[0] - p.x) * inv.x, a1 = (b[1] - p.y) * inv.y, a2 = (b[2] - p.z) * inv.z;
    R c0 = (b[3] - p.x) * inv.x, c1 = (b[4] - p.y) * inv.y, c2 = (b[5] - p.z) * inv.z;
    if (a0 > c0) {
      const R t = a0;
      a0 = c0;
      c0 = t;
    }
    if (a1 > c1) {
      const R t = a1;
      a1 = c1;
      c1 = t;
    }
    if (a2 > c2) {
      const R t = a2;
      a2 = c2;
      c2 = t;
    }
    tmin = vmax(vmax(a0, a1), a2);
    tmax = vmin(vmin(c0, c1), c2) * (1. + 2 * 2.220446049250313e-16);
  }
  __host__ __device__ static R box_safety(const double *b, V3<R> p)
  {
    const R cx = 0.5 * (b[3] + b[0]), cy = 0.5 * (b[4] + b[1]), cz = 0.5 * (b[5] + b[2]);
    const R sx = b[3] - b[0], sy = b[4] - b[1], sz = b[5] - b[2];
    return vmax(vmax(vabs(p.x - cx) - 0.5 * sx, vabs(p.y - cy) - 0.5 * sy), vabs(p.z - cz) - 0.5 * sz);
  }

  // ---- daughters through the BVH --------------------------------------------------
  // BVH::CheckDaughterIntersections: step and hit (daughter index) updated
  __host__ __device__ void check_daughters(int lv, V3<R> p, V3<R> d, R &step, int last_pv, int &hit) const
  {
    const LogicalRec &L = g.lvs[lv];
    if (L.bvh < 0) return;
    const BvhRec &B = g.bvhs[L.bvh];
    const int32_t *prim = g.ipool + B.prim, *offset = g.ipool + B.offset, *nchild = g.ipool + B.nchild;
    const double *nbox = g.bpool + B.node_box, *pbox = g.bpool + B.prim_box;
    const V3<R> inv{1.0 / nonzero(d.x), 1.0 / nonzero(d.y), 1.0 / nonzero(d.z)};
    NavWalkStack stack;
    int sp = 0;
    stack.set(sp++, 0);
    do {
      const unsigned id = stack.get(--sp);
      const int nc = nchild[id];
      if (nc >= 0) {
        for (int i = 0; i < nc; ++i) {
          const int k = prim[offset[id] + i];
          R tmin, tmax;
          box_intersect(pbox + 6 * k, p, inv, tmin, tmax);
          if (tmin <= tmax && tmax >= 0.0 && tmin < step) {
            const int v = g.daughters[L.dbegin + k];
            const Transform &T = g.pvs[v].tr;
            const R dist = shape_query<Mask, R>(g.shapes, g.pool, shape_of_pv(v), kQDistIn, to_local_point(T, p),
                                            to_local_dir(T, d), step);
            if (dist < step && !(dist <= 0.0 && v == last_pv)) {
              step = dist;
              hit = k;
            }
          }
        }
      } else {
        const unsigned cl = 2 * id + 1, cr = 2 * id + 2;
        R tminL, tmaxL, tminR, tmaxR;
        box_intersect(nbox + 6 * cl, p, inv, tminL, tmaxL);
        box_intersect(nbox + 6 * cr, p, inv, tminR, tmaxR);
        const bool tl = tminL <= tmaxL && tmaxL >= 0.0 && tminL < step;
        const bool tr = tminR <= tmaxR && tmaxR >= 0.0 && tminR < step;
        if (tminR < tminL) {
          if (tr) stack.set(sp++, cr);
          if (tl) stack.set(sp++, cl);
        } else {
          if (tl) stack.set(sp++, cl);
          if (tr) stack.set(sp++, cr);
        }
      }
    } while (sp > 0);
  }

  // UnplacedVolume::ApproachSolid (VecGeom/volumes/UnplacedVolume.h:272-281):
  // BoxImplementation::IntersectCachedKernel2 (BoxImplementation.h:326-368) on
  // the extent with t0 = 0, t1 = kInfLength; negative inside the box
  __host__ __device__ static R approach_solid(const double *mn, const double *mx, V3<R> p, V3<R> inv)
  {
    const bool sx = inv.x < 0, sy = inv.y < 0, sz = inv.z < 0;
    R tmin = ((sx ? mx[0] : mn[0]) - p.x) * inv.x;
    const R tymax = ((sy ? mn[1] : mx[1]) - p.y) * inv.y;
    if (tmin > tymax) return kInfLength;
    R tmax = ((sx ? mn[0] : mx[0]) - p.x) * inv.x;
    const R tymin = ((sy ? mx[1] : mn[1]) - p.y) * inv.y;
    if (tymin > tmax) return kInfLength;
    tmin = vmax(tmin, tymin);
    tmax = vmin(tmax, tymax);
    const R tzmin = ((sz ? mx[2] : mn[2]) - p.z) * inv.z;
    const R tzmax = ((sz ? mn[2] : mx[2]) - p.z) * inv.z;
    if (tmin > tzmax || tzmin > tmax) return kInfLength;
    tmin = vmax(tmin, tzmin);
    tmax = vmin(tmax, tzmax);
    if (!((tmin <= kInfLength + kTolerance) && (tmax > 0.0 - kTolerance))) return kInfLength;
    return tmin;
  }

  // BVH::ApproachNextDaughter (source/BVH.cpp:395-449): the daughter traversal
  // of check_daughters with the daughter's bounding box (ApproachSolid) in
  // place of its shape
  __host__ __device__ void approach_next_daughter(int lv, V3<R> p, V3<R> d, R &step, int last_pv) const
  {
    const LogicalRec &L = g.lvs[lv];
    if (L.bvh < 0) return;
    const BvhRec &B = g.bvhs[L.bvh];
    const int32_t *prim = g.ipool + B.prim, *offset = g.ipool + B.offset, *nchild = g.ipool + B.nchild;
    const double *nbox = g.bpool + B.node_box, *pbox = g.bpool + B.prim_box;
    const V3<R> inv{1.0 / nonzero(d.x), 1.0 / nonzero(d.y), 1.0 / nonzero(d.z)};
    NavWalkStack stack;
    int sp = 0;
    stack.set(sp++, 0);
    do {
      const unsigned id = stack.get(--sp);
      const int nc = nchild[id];
      if (nc >= 0) {
        for (int i = 0; i < nc; ++i) {
          const int k = prim[offset[id] + i];
          R tmin, tmax;
          box_intersect(pbox + 6 * k, p, inv, tmin, tmax);
          if (tmin <= tmax && tmax >= 0.0 && tmin < step) {
            const int v = g.daughters[L.dbegin + k];
            const Transform &T = g.pvs[v].tr;
            const V3<R> lp = to_local_point(T, p), ld = to_local_dir(T, d);
            const V3<R> li{1.0 / nonzero(ld.x), 1.0 / nonzero(ld.y), 1.0 / nonzero(ld.z)};
            const ShapeRec &S = g.shapes[shape_of_pv(v)];
            const R dist = approach_solid(S.bmin, S.bmax, lp, li);
            if (dist < step && !(dist <= 0.0 && v == last_pv)) step = dist;
          }
        }
      } else {
        const unsigned cl = 2 * id + 1, cr = 2 * id + 2;
        R tminL, tmaxL, tminR, tmaxR;
        box_intersect(nbox + 6 * cl, p, inv, tminL, tmaxL);
        box_intersect(nbox + 6 * cr, p, inv, tminR, tmaxR);
        const bool tl = tminL <= tmaxL && tmaxL >= 0.0 && tminL < step;
        const bool tr = tminR <= tmaxR && tmaxR >= 0.0 && tminR < step;
        if (tminR < tminL) {
          if (tr) stack.set(sp++, cr);
          if (tl) stack.set(sp++, cl);
        } else {
          if (tl) stack.set(sp++, cl);
          if (tr) stack.set(sp++, cr);
        }
      }
    } while (sp > 0);
  }

  // ApproachNextDaughter from a global point in the volume of `in`; the
  // volume last exited (in.last) is not approached again from its surface
  __host__ __device__ R approach(V3<R> gp, V3<R> gd, R step, const NavState &in) const
  {
    const NavNode &N = g.nodes[in.ind];
    const V3<R> lp = to_local_point(N.global, gp), ld = to_local_dir(N.global, gd);
    const int last_pv = in.last ? g.nodes[in.last].pv : -1;
    approach_next_daughter(g.pvs[N.pv].lv, lp, ld, step, last_pv);
    return step;
  }

  // BVH::LevelLocate(exclvol, localpoint, pvol, daughterlocalpoint): first
  // daughter in tree order (skipping excl_pv) that contains the point
  __host__ __device__ int level_locate(int lv, int excl_pv, V3<R> p, V3<R> &dp) const
  {
    const LogicalRec &L = g.lvs[lv];
    if (L.bvh < 0) return -1;
    const BvhRec &B = g.bvhs[L.bvh];
    const int32_t *prim = g.ipool + B.prim, *offset = g.ipool + B.offset, *nchild = g.ipool + B.nchild;
    const double *nbox = g.bpool + B.node_box, *pbox = g.bpool + B.prim_box;
    NavWalkStack stack;
    int sp = 0;
    stack.set(sp++, 0);
    do {
      const unsigned id = stack.get(--sp);
      const int nc = nchild[id];
      if (nc >= 0) {
        for (int i = 0; i < nc; ++i) {
          const int k = prim[offset[id] + i];
          if (!box_contains(pbox + 6 * k, p)) continue;
          const int v = g.daughters[L.dbegin + k];
          if (v == excl_pv) continue;
          const V3<R> q = to_local_point(g.pvs[v].tr, p);
          if (shape_query<Mask, R>(g.shapes, g.pool, shape_of_pv(v), kQContains, q, q, 0.0) != 0.0) {
            dp = q;
            return k;
          }
        }
      } else {
        const unsigned cl = 2 * id + 1, cr = 2 * id + 2;
        if (box_contains(nbox + 6 * cl, p)) stack.set(sp++, cl);
        if (box_contains(nbox + 6 * cr, p)) stack.set(sp++, cr);
      }
    } while (sp > 0);
    return -1;
  }

  // BVH::LevelLocate(exclvol, localpoint, localdirection, pvol,
  // daughterlocalpoint) (VecGeom/base/BVH.h:297-344): first daughter in tree
  // order (skipping excl_pv) whose unplaced Inside at the daughter-local
  // point is kInside, or kSurface with the direction entering it (the
  // unplaced Normal there against the daughter-local direction, < 0)
  __host__ __device__ int level_locate_dir(int lv, int excl_pv, V3<R> p, V3<R> d, V3<R> &dp) const
  {
    const LogicalRec &L = g.lvs[lv];
    if (L.bvh < 0) return -1;
    const BvhRec &B = g.bvhs[L.bvh];
    const int32_t *prim = g.ipool + B.prim, *offset = g.ipool + B.offset, *nchild = g.ipool + B.nchild;
    const double *nbox = g.bpool + B.node_box, *pbox = g.bpool + B.prim_box;
    NavWalkStack stack;
    int sp = 0;
    stack.set(sp++, 0);
    do {
      const unsigned id = stack.get(--sp);
      const int nc = nchild[id];
      if (nc >= 0) {
        for (int i = 0; i < nc; ++i) {
          const int k = prim[offset[id] + i];
          if (!box_contains(pbox + 6 * k, p)) continue;
          const int v = g.daughters[L.dbegin + k];
          if (v == excl_pv) continue;
          const V3<R> q = to_local_point(g.pvs[v].tr, p);
          const int inside = int(shape_query<Mask, R>(g.shapes, g.pool, shape_of_pv(v), kQInside, q, q, R(0.0)));
          bool take = inside == kInsideV;
          if (inside == kSurfaceV) {
            const V3<R> qd = to_local_dir(g.pvs[v].tr, d);
            take = shape_query<Mask, R>(g.shapes, g.pool, shape_of_pv(v), kQNormalDot, q, qd, R(0.0)) < R(0.0);
          }
          if (take) {
            dp = q;
            return k;
          }
        }
      } else {
        const unsigned cl = 2 * id + 1, cr = 2 * id + 2;
        if (box_contains(nbox + 6 * cl, p)) stack.set(sp++, cl);
        if (box_contains(nbox + 6 * cr, p)) stack.set(sp++, cr);
      }
    } while (sp > 0);
    return -1;
  }

  // LocateGlobalPoint(vol, point, path, top = false) from the node of vol
  __host__ __device__ unsigned descend(unsigned node, V3<R> p, int excl_pv) const
  {
    for (;;) {
      const int lv = g.pvs[g.nodes[node].pv].lv;
      V3<R> dp;
      const int k = level_locate(lv, excl_pv, p, dp);
      if (k < 0) return node;
      node = unsigned(g.children[g.nodes[node].child0 + k]);
      p = dp;
    }
  }

  __host__ __device__ unsigned locate_global(V3<R> gp) const
  {
    if (shape_query<Mask, R>(g.shapes, g.pool, shape_of_pv(g.world_pv), kQContains, gp, gp, 0.0) == 0.0) return 0;
    return descend(1, gp, -1);
  }

  // ---- one step -------------------------------------------------------------------
  // relocate = false is ComputeStep (VNavigator.h:494-540): the out state is
  // PrepareOutState's (the hit daughter pushed, or the current volume with
  // its boundary flag), with no relocation after the boundary.
  __host__ __device__ R step(V3<R> gp, V3<R> gd, R step_limit, const NavState &in, NavState &out,
                                  bool relocate = true) const
  {
    const NavNode &N = g.nodes[in.ind];
    const V3<R> lp = to_local_point(N.global, gp), ld = to_local_dir(N.global, gd);
    const int pv = N.pv;
    const int lv = g.pvs[pv].lv;
    R st = shape_query<Mask, R>(g.shapes, g.pool, g.lvs[lv].shape, kQDistOut, lp, ld, step_limit);
    if (st < 0.) st = 0.;
    int hit = -1;
    const int last_pv = in.last ? g.nodes[in.last].pv : -1;
    check_daughters(lv, lp, ld, st, last_pv, hit);
    out = in;
    if (st == kInfLength && step_limit > 0.) {
      out.boundary = true;
      out.ind = unsigned(N.mother);
      return kTolerance;
    }
    if (st > step_limit) {
      out.boundary = false;
      return step_limit;
    }
    out.boundary = true;
    out.last = out.ind;
    if (hit >= 0) out.ind = unsigned(g.children[N.child0 + hit]);
    if (st < 0.) st = 0.;
    if (!relocate) return st;
    const R e = st + kPushTolerance;
    V3<R> q{lp.x + e * ld.x, lp.y + e * ld.y, lp.z + e * ld.z};
    if (out.ind == in.ind) {
      // RelocatePointFromPathForceDifferent: climb while the point is outside
      // (the entry volume is always left), then descend excluding it
      unsigned cur = out.ind;
      const unsigned entry = cur;
      const int entry_pv = pv;
      while (cur) {
        const int cpv = g.nodes[cur].pv;
        if (cur == entry ||
            shape_query<Mask, R>(g.shapes, g.pool, shape_of_pv(cpv), kQContains, q, q, 0.0) == 0.0) {
          q = to_master_point(g.pvs[cpv].tr, q);
          cur = unsigned(g.nodes[cur].mother);
        } else {
          break;
        }
      }
      out.ind = cur ? descend(cur, q, entry_pv) : 0u;
    } else {
      const unsigned next = out.ind;
      out.ind = descend(next, to_local_point(g.pvs[g.nodes[next].pv].tr, q), -1);
    }
    return st;
  }

  // ---- safety -----------------------------------------------------------------------
  __host__ __device__ R daughter_safety(int lv, V3<R> p, R safety) const
  {
    const LogicalRec &L = g.lvs[lv];
    const BvhRec &B = g.bvhs[L.bvh];
    const int32_t *prim = g.ipool + B.prim, *offset = g.ipool + B.offset, *nchild = g.ipool + B.nchild;
    const double *nbox = g.bpool + B.node_box, *pbox = g.bpool + B.prim_box;
    NavWalkStack stack;
    int sp = 0;
    stack.set(sp++, 0);
    do {
      const unsigned id = stack.get(--sp);
      const int nc = nchild[id];
      if (nc >= 0) {
        for (int i = 0; i < nc; ++i) {
          const int k = prim[offset[id] + i];
          if (box_safety(pbox + 6 * k, p) < safety) {
            const int v = g.daughters[L.dbegin + k];
            const R dist =
                shape_query<Mask, R>(g.shapes, g.pool, shape_of_pv(v), kQSafIn, to_local_point(g.pvs[v].tr, p), p, 0.0);
            if (dist < safety) safety = dist;
          }
        }
      } else {
        const unsigned cl = 2 * id + 1, cr = 2 * id + 2;
        const R sL = box_safety(nbox + 6 * cl, p), sR = box_safety(nbox + 6 * cr, p);
        const bool tl = sL < safety, tr = sR < safety;
        if (sR < sL) {
          if (tr) stack.set(sp++, cr);
          if (tl) stack.set(sp++, cl);
        } else {
          if (tl) stack.set(sp++, cl);
          if (tr) stack.set(sp++, cr);
        }
      }
    } while (sp > 0);
    return safety;
  }

  __host__ __device__ R safety(V3<R> gp, unsigned node) const
  {
    if (!node) return 0.0;
    const NavNode &N = g.nodes[node];
    const V3<R> lp = to_local_point(N.global, gp);
    const int lv = g.pvs[N.pv].lv;
    R s = shape_query<Mask, R>(g.shapes, g.pool, g.lvs[lv].shape, kQSafOut, lp, lp, 0.0);
    if (s > 0.0 && g.lvs[lv].bvh >= 0) s = daughter_safety(lv, lp, s);
    return s;
  }
};

}  // namespace vgm
