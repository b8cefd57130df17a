// Oracle: navigation (TEST INFRASTRUCTURE ONLY), FP64 scalar.
//   navigation states: NavStateIndex semantics (VecGeom/navigation/NavStateIndex.h:
//     Push :289, Pop :297, SetLastExited :281, Top :305), states numbered by a
//     depth-first walk of the placement tree from the world (node 1; 0 = none)
//   cached global matrices: NavStatePath::TopMatrix (VecGeom/navigation/NavStatePath.h:616-628)
//     through Transformation3D::MultiplyFromRight (VecGeom/base/Transformation3D.h:878-947)
//     with the flags of SetProperties / GenerateRotationCode (source/Transformation3D.cpp:118-185)
//   step: VNavigatorHelper::ComputeStepAndPropagatedState for MotherIsConvex = false
//     (VecGeom/navigation/VNavigator.h:436-492), TreatDistanceToMother (:241-249),
//     PrepareOutState (:192-235), MovePointAfterBoundary (:168-173), Relocate (:1070-1091)
//   locators: GlobalLocator::LocateGlobalPoint / LocateGlobalPointExclVolume /
//     RelocatePointFromPathForceDifferent (VecGeom/navigation/GlobalLocator.h:35-200)
//   daughters: either a linear scan in daughter order (NewSimpleNavigator /
//     SimpleLevelLocator semantics, VecGeom/navigation/NewSimpleNavigator.h:30-70)
//     or the reference's BVH (source/BVH.cpp construction, VecGeom/base/BVH.h
//     traversals, AABBs from ABBoxManager::ComputeABBox source/ABBoxManager.cpp:109-143).
//     Both give the same steps and locations on geometries without ties or
//     overlaps; ComputeSafety prunes in the tree's order, so it always uses
//     the tree.
#include <omp.h>
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>
#include "geom.h"
#include "oracle.h"

namespace orc {

namespace {

constexpr double kEps = std::numeric_limits<double>::epsilon();

struct Box3 {  // AABB (VecGeom/base/AABB.h)
  Vec3 mn, mx;
  Vec3 Center() const { return Vec3(0.5 * (mx.x + mn.x), 0.5 * (mx.y + mn.y), 0.5 * (mx.z + mn.z)); }
  Vec3 Size() const { return mx - mn; }
  double SurfaceArea() const
  {
    const Vec3 e = Size();
    return 2. * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0]);
  }
  bool Contains(const Vec3 &p) const
  {
    return p[0] >= mn[0] && p[0] <= mx[0] && p[1] >= mn[1] && p[1] <= mx[1] && p[2] >= mn[2] && p[2] <= mx[2];
  }
  double Safety(const Vec3 &p) const
  {
    const Vec3 c = Center(), s = Size();
    const Vec3 a = (p - c).Abs();
    const Vec3 v(a.x - 0.5 * s.x, a.y - 0.5 * s.y, a.z - 0.5 * s.z);
    return v.MaxC();
  }
  void IntersectInvDir(const Vec3 &p, const Vec3 &inv, double &tmin, double &tmax) const
  {
    Vec3 t0((mn.x - p.x) * inv.x, (mn.y - p.y) * inv.y, (mn.z - p.z) * inv.z);
    Vec3 t1((mx.x - p.x) * inv.x, (mx.y - p.y) * inv.y, (mx.z - p.z) * inv.z);
    for (int k = 0; k < 3; ++k)
      if (t0[k] > t1[k]) std::swap(t0[k], t1[k]);
    tmin = t0.MaxC();
    tmax = t1.MinC() * (1. + 2 * kEps);
  }
  bool HitWithin(const Vec3 &p, const Vec3 &inv, double step) const
  {
    double a, b;
    IntersectInvDir(p, inv, a, b);
    return a <= b && b >= 0.0 && a < step;
  }
  static Box3 Union(const Box3 &A, const Box3 &B)
  {
    Box3 r;
    r.mn = Vec3(Min(A.mn[0], B.mn[0]), Min(A.mn[1], B.mn[1]), Min(A.mn[2], B.mn[2]));
    r.mx = Vec3(Max(A.mx[0], B.mx[0]), Max(A.mx[1], B.mx[1]), Max(A.mx[2], B.mx[2]));
    return r;
  }
};

// BVH of one logical volume (source/BVH.cpp:24-87, 275-329)
struct Bvh {
  std::vector<Box3> aabbs, nodes;
  std::vector<int> prim, offset, nchild;
  int depth = 0;

  int *SahSplit(int *begin, int *end)
  {
    int bestSplitAxis = -1;
    double bestMetric = double(end - begin);
    int bestSplitObject = -1;
    const long nObj = end - begin;
    int axis_ = 0;
    auto key = [&](int a) {
      const Vec3 c = aabbs[a].Center(), m = aabbs[a].mn;
      const double shift = 0.01;
      return Vec3(c.x + shift * (c.x - m.x), c.y + shift * (c.y - m.y), c.z + shift * (c.z - m.z));
    };
    auto less3D = [&](const Vec3 &l, const Vec3 &r) {
      const int s = axis_;
      return l[s] < r[s] || (l[s] == r[s] && (l[(s + 1) % 3] < r[(s + 1) % 3] ||
                                               (l[(s + 1) % 3] == r[(s + 1) % 3] && l[(s + 2) % 3] < r[(s + 2) % 3])));
    };
    auto sorter = [&](int a, int b) { return less3D(key(a), key(b)); };
    for (int axis = 0; axis <= 2; ++axis) {
      axis_ = axis;
      std::sort(begin, end, sorter);
      // sweepSurfaceArea
      std::vector<std::pair<double, double>> areas(end - begin, {0., 0.});
      Box3 box = aabbs[*begin];
      for (int *it = begin + 1; it < end; ++it) {
        areas[it - begin].first = box.SurfaceArea();
        box = Box3::Union(box, aabbs[*it]);
      }
      Box3 box2 = aabbs[*(end - 1)];
      for (int *it = end - 1; it >= begin; --it) {
        box2 = Box3::Union(box2, aabbs[*it]);
        areas[it - begin].second = box2.SurfaceArea();
      }
      const double tot = areas.front().second;
      for (int *so = begin; so < end; ++so) {
        const double left = areas[so - begin].first / tot;
        const double right = areas[so - begin].second / tot;
        const long k = so - begin;
        const double metric = left * k + right * (end - so) + 0.1 * std::abs(nObj / 2 - k / nObj);
        if (metric < bestMetric) {
          bestMetric = metric;
          bestSplitAxis = axis;
          bestSplitObject = *so;
        }
      }
    }
    if (bestSplitAxis == -1) return end;
    axis_ = bestSplitAxis;
    return std::partition(begin, end, [&](size_t i) { return sorter(int(i), bestSplitObject); });
  }

  void ComputeNodes(unsigned id, int *first, int *last, unsigned nn)
  {
    if (id >= nn) return;
    nchild[id] = int(last - first);
    offset[id] = int(first - prim.data());
    if (first == last) return;
    nodes[id] = aabbs[*first];
    for (int *it = first + 1; it != last; ++it) nodes[id] = Box3::Union(nodes[id], aabbs[*it]);
    if (first + 1 == last) return;
    int *pivot = SahSplit(first, last);
    ComputeNodes(2 * id + 1, first, pivot, nn);
    ComputeNodes(2 * id + 2, pivot, last, nn);
  }

  void Build(std::vector<Box3> boxes)
  {
    const int n = int(boxes.size());
    aabbs = std::move(boxes);
    prim.resize(n);
    std::iota(prim.begin(), prim.end(), 0);
    const int half = n / 2;
    depth = std::min(std::max(0, half > 0 ? int(std::log2(double(half))) : 0), 32);
    const unsigned nn = (2u << depth) - 1;
    nchild.assign(nn, 0);
    offset.assign(nn, -1);
    nodes.assign(nn, Box3());
    ComputeNodes(0, prim.data(), prim.data() + n, nn);
    for (unsigned id = 0; id < nn / 2; ++id)
      if (nchild[id] > 8 && (nchild[id] == nchild[2 * id + 1] + nchild[2 * id + 2])) nchild[id] = -1;
  }
};

struct XForm {  // Transformation3D with its property flags
  Transform T;
  bool hasT = false, hasR = false, ident = true;
  void SetProperties()
  {
    hasT = std::fabs(T.t[0]) > kTolerance || std::fabs(T.t[1]) > kTolerance || std::fabs(T.t[2]) > kTolerance;
    int code = 0;
    for (int i = 0; i < 9; ++i) code |= (1 << i) * (std::fabs(T.r[i]) > kTolerance);
    const bool identity_rot = (code == 0x111) && T.r[0] == 1. && T.r[4] == 1. && T.r[8] == 1.;
    hasR = !identity_rot;
    ident = !hasT && !hasR;
  }
  void MultiplyFromRight(const XForm &rhs)
  {
    if (rhs.ident) return;
    ident = false;
    double *t = T.t, *r = T.r;
    const double *rt = rhs.T.t, *rr = rhs.T.r;
    if (rhs.hasT) {
      hasT = true;
      t[0] += r[0] * rt[0];
      t[0] += r[1] * rt[1];
      t[0] += r[2] * rt[2];
      t[1] += r[3] * rt[0];
      t[1] += r[4] * rt[1];
      t[1] += r[5] * rt[2];
      t[2] += r[6] * rt[0];
      t[2] += r[7] * rt[1];
      t[2] += r[8] * rt[2];
    }
    if (rhs.hasR) {
      hasR = true;
      for (int row = 0; row < 3; ++row) {
        const double tx = r[3 * row], ty = r[3 * row + 1], tz = r[3 * row + 2];
        r[3 * row] = tx * rr[0];
        r[3 * row + 1] = tx * rr[1];
        r[3 * row + 2] = tx * rr[2];
        r[3 * row] += ty * rr[3];
        r[3 * row + 1] += ty * rr[4];
        r[3 * row + 2] += ty * rr[5];
        r[3 * row] += tz * rr[6];
        r[3 * row + 1] += tz * rr[7];
        r[3 * row + 2] += tz * rr[8];
      }
    }
  }
};

struct Node {
  int pv = -1, mother = 0, level = 0;
  std::vector<int> children;  // node of daughter k of the pv's logical volume
  Transform global;
};

struct State {
  unsigned ind = 0, last = 0;
  bool boundary = false;
};

}  // namespace

struct Navigator {
  Geometry &G;
  int world = -1;
  bool use_bvh = false;
  std::vector<Node> nodes;  // index 0 unused
  std::vector<Bvh> bvh;     // per logical volume (empty if no daughters)
  std::vector<XForm> pvx;   // per placed volume

  explicit Navigator(Geometry &g) : G(g) {}

  const Shape &ShapeOf(int pv) const { return *G.shapes[G.lvs[G.pvs[pv].lv].shape]; }
  const LogicalVolume &LvOf(int pv) const { return G.lvs[G.pvs[pv].lv]; }

  void Close(int world_pv)
  {
    world = world_pv;
    pvx.resize(G.pvs.size());
    for (size_t i = 0; i < G.pvs.size(); ++i) {
      pvx[i].T = G.pvs[i].tr;
      pvx[i].SetProperties();
    }
    // ABBoxManager::ComputeABBox per daughter, then one BVH per logical volume
    bvh.assign(G.lvs.size(), Bvh());
    for (size_t l = 0; l < G.lvs.size(); ++l) {
      const LogicalVolume &lv = G.lvs[l];
      if (lv.daughters.empty()) continue;
      std::vector<Box3> boxes;
      for (int d : lv.daughters) {
        Vec3 lo, hi;
        ShapeOf(d).Extent(lo, hi);
        const Vec3 delta = hi - lo;
        double mnx = kInfLength, mny = kInfLength, mnz = kInfLength;
        double mxx = -kInfLength, mxy = -kInfLength, mxz = -kInfLength;
        for (int x = 0; x <= 1; ++x)
          for (int y = 0; y <= 1; ++y)
            for (int z = 0; z <= 1; ++z) {
              const Vec3 c(lo.x + x * delta.x, lo.y + y * delta.y, lo.z + z * delta.z);
              const Vec3 tc = G.pvs[d].tr.InverseTransform(c);
              mnx = std::min(mnx, tc.x);
              mny = std::min(mny, tc.y);
              mnz = std::min(mnz, tc.z);
              mxx = std::max(mxx, tc.x);
              mxy = std::max(mxy, tc.y);
              mxz = std::max(mxz, tc.z);
            }
        Box3 b;
        b.mn = Vec3(mnx - 1E-3, mny - 1E-3, mnz - 1E-3);
        b.mx = Vec3(mxx + 1E-3, mxy + 1E-3, mxz + 1E-3);
        boxes.push_back(b);
      }
      bvh[l].Build(boxes);
    }
    // state tree, depth first from the world
    nodes.assign(1, Node());
    XForm ident;
    AddNode(world_pv, 0, 0, ident);
  }

  int AddNode(int pv, int mother, int level, const XForm &parent)
  {
    const int id = int(nodes.size());
    nodes.push_back(Node());
    nodes[id].pv = pv;
    nodes[id].mother = mother;
    nodes[id].level = level;
    XForm g = parent;
    if (level > 0) g.MultiplyFromRight(pvx[pv]);  // TopMatrix skips the world's own placement
    nodes[id].global = g.T;
    const LogicalVolume &lv = LvOf(pv);
    std::vector<int> kids;
    for (int d : lv.daughters) kids.push_back(AddNode(d, id, level + 1, g));
    nodes[id].children = kids;
    return id;
  }

  // ---- level locators ------------------------------------------------------
  // First daughter (skipping `excl`) whose Contains holds; returns its index.
  int LevelLocate(const LogicalVolume &lv, int lvid, int excl, const Vec3 &p, Vec3 &dp) const
  {
    if (!use_bvh) {
      for (size_t k = 0; k < lv.daughters.size(); ++k) {
        const int d = lv.daughters[k];
        if (d == excl) continue;
        const Vec3 q = G.pvs[d].tr.Transform_(p);
        if (ShapeOf(d).Contains(q)) {
          dp = q;
          return int(k);
        }
      }
      return -1;
    }
    const Bvh &B = bvh[lvid];
    unsigned stack[40], *ptr = &stack[1];
    stack[0] = 0;
    do {
      const unsigned id = *--ptr;
      if (B.nchild[id] >= 0) {
        for (int i = 0; i < B.nchild[id]; ++i) {
          const int k = B.prim[B.offset[id] + i];
          if (B.aabbs[k].Contains(p)) {
            const int d = lv.daughters[k];
            if (d == excl) continue;
            const Vec3 q = G.pvs[d].tr.Transform_(p);
            if (ShapeOf(d).Contains(q)) {
              dp = q;
              return k;
            }
          }
        }
      } else {
        const unsigned L = 2 * id + 1, R = 2 * id + 2;
        if (B.nodes[L].Contains(p)) *ptr++ = L;
        if (B.nodes[R].Contains(p)) *ptr++ = R;
      }
    } while (ptr > stack);
    return -1;
  }

  // BVH::LevelLocate(exclvol, localpoint, localdirection, pvol,
  // daughterlocalpoint) (VecGeom/base/BVH.h:297-344): the first daughter
  // (skipping `excl`) whose unplaced Inside is kInside, or kSurface with
  // the direction entering it (the unplaced Normal against the daughter-local
  // direction < 0); a linear scan in daughter order without the BVH.
  int LevelLocateDir(const LogicalVolume &lv, int lvid, int excl, const Vec3 &p, const Vec3 &d, Vec3 &dp) const
  {
    auto holds = [&](int v, Vec3 &q) {
      const Transform &T = G.pvs[v].tr;
      q = T.Transform_(p);
      const Shape &u = ShapeOf(v);
      const int inside = u.Inside(q);
      if (inside == kInside) return true;
      if (inside != kSurface) return false;
      const Vec3 dir = T.TransformDirection(d);
      Vec3 normal;
      u.Normal(q, normal);
      return Dot(normal, dir) < 0.0;
    };
    if (!use_bvh) {
      for (size_t k = 0; k < lv.daughters.size(); ++k) {
        const int v = lv.daughters[k];
        Vec3 q;
        if (v != excl && holds(v, q)) {
          dp = q;
          return int(k);
        }
      }
      return -1;
    }
    const Bvh &B = bvh[lvid];
    unsigned stack[40], *ptr = &stack[1];
    stack[0] = 0;
    do {
      const unsigned id = *--ptr;
      if (B.nchild[id] >= 0) {
        for (int i = 0; i < B.nchild[id]; ++i) {
          const int k = B.prim[B.offset[id] + i];
          if (!B.aabbs[k].Contains(p)) continue;
          const int v = lv.daughters[k];
          if (v == excl) continue;
          Vec3 q;
          if (holds(v, q)) {
            dp = q;
            return k;
          }
        }
      } else {
        const unsigned L = 2 * id + 1, R = 2 * id + 2;
        if (B.nodes[L].Contains(p)) *ptr++ = L;
        if (B.nodes[R].Contains(p)) *ptr++ = R;
      }
    } while (ptr > stack);
    return -1;
  }

  // LevelLocate in the volume of state `node` (directional when d is given),
  // skipping the daughter of state `excl`: the daughter's state or 0
  unsigned LevelLocateState(unsigned node, unsigned excl, const Vec3 &p, const Vec3 *d, Vec3 &q) const
  {
    q = p;
    if (!node) return 0u;
    const int pv = nodes[node].pv;
    const LogicalVolume &lv = LvOf(pv);
    if (lv.daughters.empty()) return 0u;
    const int ex = excl ? nodes[excl].pv : -1;
    Vec3 dp;
    const int k = d ? LevelLocateDir(lv, G.pvs[pv].lv, ex, p, *d, dp) : LevelLocate(lv, G.pvs[pv].lv, ex, p, dp);
    if (k < 0) return 0u;
    q = dp;
    return unsigned(nodes[node].children[size_t(k)]);
  }

  // LocateGlobalPoint(vol, point, path, top = false) with `node` the node of vol
  unsigned Descend(unsigned node, Vec3 p, int excl) const
  {
    for (;;) {
      const int pv = nodes[node].pv;
      const LogicalVolume &lv = LvOf(pv);
      if (lv.daughters.empty()) return node;
      Vec3 dp;
      const int k = LevelLocate(lv, G.pvs[pv].lv, excl, p, dp);
      if (k < 0) return node;
      node = unsigned(nodes[node].children[k]);
      p = dp;
    }
  }

  unsigned LocateGlobal(const Vec3 &gp) const
  {
    if (!ShapeOf(world).Contains(gp)) return 0;
    return Descend(1, gp, -1);
  }

  // ---- daughters -------------------------------------------------------------
  void CheckDaughters(const LogicalVolume &lv, int lvid, const Vec3 &p, const Vec3 &d, double &step, int last,
                      int &hit) const
  {
    if (lv.daughters.empty()) return;
    if (!use_bvh) {
      for (size_t k = 0; k < lv.daughters.size(); ++k) {
        const int v = lv.daughters[k];
        const double dist = ShapeOf(v).DistanceToIn(G.pvs[v].tr.Transform_(p), G.pvs[v].tr.TransformDirection(d), step);
        if (dist < step && !(dist <= 0.0 && v == last)) {
          step = dist;
          hit = int(k);
        }
      }
      return;
    }
    const Bvh &B = bvh[lvid];
    const Vec3 inv(1.0 / NonZero(d[0]), 1.0 / NonZero(d[1]), 1.0 / NonZero(d[2]));
    unsigned stack[40], *ptr = &stack[1];
    stack[0] = 0;
    do {
      const unsigned id = *--ptr;
      if (B.nchild[id] >= 0) {
        for (int i = 0; i < B.nchild[id]; ++i) {
          const int k = B.prim[B.offset[id] + i];
          if (B.aabbs[k].HitWithin(p, inv, step)) {
            const int v = lv.daughters[k];
            const double dist =
                ShapeOf(v).DistanceToIn(G.pvs[v].tr.Transform_(p), G.pvs[v].tr.TransformDirection(d), step);
            if (dist < step && !(dist <= 0.0 && v == last)) {
              step = dist;
              hit = k;
            }
          }
        }
      } else {
        const unsigned L = 2 * id + 1, R = 2 * id + 2;
        double tminL = kInfLength, tmaxL = -kInfLength, tminR = kInfLength, tmaxR = -kInfLength;
        B.nodes[L].IntersectInvDir(p, inv, tminL, tmaxL);
        B.nodes[R].IntersectInvDir(p, inv, tminR, tmaxR);
        const bool tl = tminL <= tmaxL && tmaxL >= 0.0 && tminL < step;
        const bool tr = tminR <= tmaxR && tmaxR >= 0.0 && tminR < step;
        if (tminR < tminL) {
          if (tr) *ptr++ = R;
          if (tl) *ptr++ = L;
        } else {
          if (tl) *ptr++ = L;
          if (tr) *ptr++ = R;
        }
      }
    } while (ptr > stack);
  }

  // UnplacedVolume::ApproachSolid (VecGeom/volumes/UnplacedVolume.h:272-281)
  // = BoxImplementation::IntersectCachedKernel2 (BoxImplementation.h:326-368)
  // on the extent, t0 = 0, t1 = kInfLength
  static double ApproachSolid(const Vec3 corners[2], const Vec3 &point, const Vec3 &inverseray)
  {
    const int signx = inverseray.x < 0, signy = inverseray.y < 0, signz = inverseray.z < 0;
    double tmin = (corners[signx].x - point.x) * inverseray.x;
    double tymax = (corners[1 - signy].y - point.y) * inverseray.y;
    if (tmin > tymax) return kInfLength;
    double tmax = (corners[1 - signx].x - point.x) * inverseray.x;
    double tymin = (corners[signy].y - point.y) * inverseray.y;
    if (tymin > tmax) return kInfLength;
    tmin = Max(tmin, tymin);
    tmax = Min(tmax, tymax);
    double tzmin = (corners[signz].z - point.z) * inverseray.z;
    double tzmax = (corners[1 - signz].z - point.z) * inverseray.z;
    if ((tmin > tzmax) || (tzmin > tmax)) return kInfLength;
    tmin = Max(tmin, tzmin);
    tmax = Min(tmax, tzmax);
    if (!((tmin <= kInfLength + kTolerance) && (tmax > 0.0 - kTolerance))) return kInfLength;
    return tmin;
  }

  // BVH::ApproachNextDaughter (source/BVH.cpp:395-449)
  void ApproachNextDaughter(const LogicalVolume &lv, int lvid, const Vec3 &point, const Vec3 &dir, double &step,
                            int last) const
  {
    if (lv.daughters.empty()) return;
    if (!use_bvh) {  // linear scan over the daughters, no bounding-box pruning
      for (int vol : lv.daughters) {
        const Vec3 localdir = G.pvs[vol].tr.TransformDirection(dir);
        const Vec3 invlocaldir(1.0 / NonZero(localdir[0]), 1.0 / NonZero(localdir[1]), 1.0 / NonZero(localdir[2]));
        Vec3 bbox[2];
        ShapeOf(vol).Extent(bbox[0], bbox[1]);
        const double dist = ApproachSolid(bbox, G.pvs[vol].tr.Transform_(point), invlocaldir);
        if (dist < step && !(dist <= 0.0 && vol == last)) step = dist;
      }
      return;
    }
    const Bvh &B = bvh[lvid];
    const Vec3 invdir(1.0 / NonZero(dir[0]), 1.0 / NonZero(dir[1]), 1.0 / NonZero(dir[2]));
    unsigned stack[40] = {0}, *ptr = &stack[1];
    do {
      const unsigned id = *--ptr;
      if (B.nchild[id] >= 0) {
        for (int i = 0; i < B.nchild[id]; ++i) {
          const int prim = B.prim[B.offset[id] + i];
          if (B.aabbs[prim].HitWithin(point, invdir, step)) {
            const int vol = lv.daughters[prim];
            const Vec3 localpoint = G.pvs[vol].tr.Transform_(point);
            const Vec3 localdir = G.pvs[vol].tr.TransformDirection(dir);
            const Vec3 invlocaldir(1.0 / NonZero(localdir[0]), 1.0 / NonZero(localdir[1]),
                                   1.0 / NonZero(localdir[2]));
            Vec3 bbox[2];
            ShapeOf(vol).Extent(bbox[0], bbox[1]);
            const double dist = ApproachSolid(bbox, localpoint, invlocaldir);
            if (dist < step && !(dist <= 0.0 && vol == last)) step = dist;
          }
        }
      } else {
        const unsigned childL = 2 * id + 1, childR = 2 * id + 2;
        double tminL = kInfLength, tmaxL = -kInfLength, tminR = kInfLength, tmaxR = -kInfLength;
        B.nodes[childL].IntersectInvDir(point, invdir, tminL, tmaxL);
        B.nodes[childR].IntersectInvDir(point, invdir, tminR, tmaxR);
        const bool traverseL = tminL <= tmaxL && tmaxL >= 0.0 && tminL < step;
        const bool traverseR = tminR <= tmaxR && tmaxR >= 0.0 && tminR < step;
        if (tminR < tminL) {
          if (traverseR) *ptr++ = childR;
          if (traverseL) *ptr++ = childL;
        } else {
          if (traverseL) *ptr++ = childL;
          if (traverseR) *ptr++ = childR;
        }
      }
    } while (ptr > stack);
  }

  double Approach(const Vec3 &gp, const Vec3 &gd, double step, const State &in) const
  {
    const Transform &T = nodes[in.ind].global;
    const int pv = nodes[in.ind].pv;
    const int last = in.last ? nodes[in.last].pv : -1;
    ApproachNextDaughter(LvOf(pv), G.pvs[pv].lv, T.Transform_(gp), T.TransformDirection(gd), step, last);
    return step;
  }

  // BVH::ComputeSafety (VecGeom/base/BVH.h:172-209)
  double DaughterSafety(const LogicalVolume &lv, int lvid, const Vec3 &p, double safety) const
  {
    const Bvh &B = bvh[lvid];
    unsigned stack[40], *ptr = &stack[1];
    stack[0] = 0;
    do {
      const unsigned id = *--ptr;
      if (B.nchild[id] >= 0) {
        for (int i = 0; i < B.nchild[id]; ++i) {
          const int k = B.prim[B.offset[id] + i];
          if (B.aabbs[k].Safety(p) < safety) {
            const int v = lv.daughters[k];
            const double dist = ShapeOf(v).SafetyToIn(G.pvs[v].tr.Transform_(p));
            if (dist < safety) safety = dist;
          }
        }
      } else {
        const unsigned L = 2 * id + 1, R = 2 * id + 2;
        const double sL = B.nodes[L].Safety(p), sR = B.nodes[R].Safety(p);
        const bool tl = sL < safety, tr = sR < safety;
        if (sR < sL) {
          if (tr) *ptr++ = R;
          if (tl) *ptr++ = L;
        } else {
          if (tl) *ptr++ = L;
          if (tr) *ptr++ = R;
        }
      }
    } while (ptr > stack);
    return safety;
  }

  // VSafetyEstimator::ComputeSafety with BVHSafetyEstimator::ComputeSafetyForLocalPoint
  double Safety(const Vec3 &gp, unsigned node) const
  {
    if (!node) return 0;
    const Transform &T = nodes[node].global;
    const Vec3 lp = T.Transform_(gp);
    const int pv = nodes[node].pv;
    double s = ShapeOf(pv).SafetyToOut(lp);
    const LogicalVolume &lv = LvOf(pv);
    if (s > 0.0 && !lv.daughters.empty()) s = DaughterSafety(lv, G.pvs[pv].lv, lp, s);
    return s;
  }

  // ---- one step ----------------------------------------------------------------
  // relocate = false: VNavigator::ComputeStep (VNavigator.h:494-540), the
  // PrepareOutState result without the relocation
  double Step(const Vec3 &gp, const Vec3 &gd, double step_limit, const State &in, State &out,
              bool relocate = true) const
  {
    const Transform &T = nodes[in.ind].global;
    const Vec3 lp = T.Transform_(gp), ld = T.TransformDirection(gd);
    const int pv = nodes[in.ind].pv;
    const LogicalVolume &lv = LvOf(pv);
    // TreatDistanceToMother
    double step = ShapeOf(pv).DistanceToOut(lp, ld, step_limit);
    if (step < 0.) step = 0.;
    int hit = -1;
    const int last = in.last ? nodes[in.last].pv : -1;
    CheckDaughters(lv, G.pvs[pv].lv, lp, ld, step, last, hit);
    // PrepareOutState
    out = in;
    if (step == kInfLength && step_limit > 0.) {
      out.boundary = true;
      out.ind = unsigned(nodes[out.ind].mother);
      return kTolerance;
    }
    if (step > step_limit) {
      out.boundary = false;
      return step_limit;
    }
    out.boundary = true;
    out.last = out.ind;
    if (hit >= 0) out.ind = unsigned(nodes[out.ind].children[hit]);
    if (step < 0.) step = 0.;
    if (!relocate) return step;
    // Relocate(MovePointAfterBoundary(localpoint, localdir, step))
    const double e = step + 1E-6;
    Vec3 q(lp.x + e * ld.x, lp.y + e * ld.y, lp.z + e * ld.z);
    if (out.ind == in.ind) {
      // RelocatePointFromPathForceDifferent
      unsigned cur = out.ind;
      const int entry_pv = nodes[cur].pv;
      const unsigned entry = cur;
      while (cur) {
        const int cpv = nodes[cur].pv;
        if (cur == entry || !ShapeOf(cpv).Contains(q)) {
          q = G.pvs[cpv].tr.InverseTransform(q);
          cur = unsigned(nodes[cur].mother);
        } else {
          break;
        }
      }
      out.ind = cur ? Descend(cur, q, entry_pv) : 0;
    } else {
      // enter the hit daughter and descend from it
      const unsigned next = out.ind;
      out.ind = Descend(next, G.pvs[nodes[next].pv].tr.Transform_(q), -1);
    }
    return step;
  }
};

}  // namespace orc

using namespace orc;

static Navigator *nav_of(orc_geom *g)
{
  auto *G = reinterpret_cast<Geometry *>(g);
  return reinterpret_cast<Navigator *>(G->nav);
}

extern "C" {

int orc_nav_close(orc_geom *g, int world_pv, int use_bvh)
{
  auto *G = reinterpret_cast<Geometry *>(g);
  if (world_pv < 0 || size_t(world_pv) >= G->pvs.size()) return -1;
  delete reinterpret_cast<Navigator *>(G->nav);
  auto *N = new Navigator(*G);
  N->use_bvh = use_bvh != 0;
  N->Close(world_pv);
  G->nav = N;
  G->nav_deleter = [](void *p) { delete reinterpret_cast<Navigator *>(p); };
  return int(N->nodes.size()) - 1;
}

int orc_nav_node(orc_geom *g, unsigned node, int *pv, int *level, int *mother)
{
  Navigator *N = nav_of(g);
  if (!N || node == 0 || node >= N->nodes.size()) return -1;
  *pv = N->nodes[node].pv;
  *level = N->nodes[node].level;
  *mother = N->nodes[node].mother;
  return 0;
}

int orc_nav_locate(orc_geom *g, const double *x, const double *y, const double *z, size_t n, uint32_t *state,
                   int nthreads)
{
  Navigator *N = nav_of(g);
  if (!N) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024)
  for (long i = 0; i < long(n); ++i) state[i] = N->LocateGlobal(Vec3(x[i], y[i], z[i]));
  return 0;
}

int orc_nav_step2(orc_geom *g, const double *x, const double *y, const double *z, const double *dx, const double *dy,
                  const double *dz, const double *step_limit, size_t n, uint32_t *state, uint32_t *last,
                  uint8_t *boundary, double *step, int relocate, int nthreads)
{
  Navigator *N = nav_of(g);
  if (!N) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024)
  for (long i = 0; i < long(n); ++i) {
    if (!state[i]) {
      step[i] = 0;
      continue;
    }
    State in{state[i], last[i], boundary[i] != 0}, out;
    step[i] = N->Step(Vec3(x[i], y[i], z[i]), Vec3(dx[i], dy[i], dz[i]), step_limit ? step_limit[i] : kInfLength, in,
                      out, relocate != 0);
    state[i] = out.ind;
    last[i] = out.last;
    boundary[i] = out.boundary;
  }
  return 0;
}

int orc_nav_step(orc_geom *g, const double *x, const double *y, const double *z, const double *dx, const double *dy,
                 const double *dz, const double *step_limit, size_t n, uint32_t *state, uint32_t *last,
                 uint8_t *boundary, double *step, int nthreads)
{
  return orc_nav_step2(g, x, y, z, dx, dy, dz, step_limit, n, state, last, boundary, step, 1, nthreads);
}

// Straight-line transport: nsteps of ComputeStepAndPropagatedState with an
// unlimited physics step, moving the global point by each step; a track that
// leaves the world stops.  steps[i] counts the steps taken.
int orc_nav_transport(orc_geom *g, double *x, double *y, double *z, const double *dx, const double *dy,
                      const double *dz, size_t n, uint32_t *state, uint32_t *last, uint8_t *boundary, int nsteps,
                      int32_t *steps, int nthreads)
{
  Navigator *N = nav_of(g);
  if (!N) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
  for (long i = 0; i < long(n); ++i) {
    State s{state[i], last[i], boundary[i] != 0};
    Vec3 p(x[i], y[i], z[i]);
    const Vec3 d(dx[i], dy[i], dz[i]);
    int k = 0;
    for (; k < nsteps && s.ind; ++k) {
      State out;
      const double st = N->Step(p, d, kInfLength, s, out);
      p = Vec3(p.x + st * d.x, p.y + st * d.y, p.z + st * d.z);
      s = out;
    }
    x[i] = p.x;
    y[i] = p.y;
    z[i] = p.z;
    state[i] = s.ind;
    last[i] = s.last;
    boundary[i] = s.boundary;
    steps[i] += k;
  }
  return 0;
}

// Transport under a physics step limit: nsteps of
// ComputeStepAndSafetyAndPropagatedState with calcsafety
// (VecGeom/navigation/VNavigator.h:543-592): when the safety covers the
// limit the track moves by it with no boundary distance (safetydone);
// otherwise the step limited to it.  reuse_safety carries safety - limit
// after such a move; a geometry step forces a fresh safety.
int orc_nav_transport_limited(orc_geom *g, double *x, double *y, double *z, const double *dx, const double *dy,
                              const double *dz, const double *step_limit, size_t n, uint32_t *state, uint32_t *last,
                              uint8_t *boundary, int nsteps, int reuse_safety, int32_t *steps, int32_t *geo_steps,
                              int nthreads)
{
  Navigator *N = nav_of(g);
  if (!N) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
  for (long i = 0; i < long(n); ++i) {
    State s{state[i], last[i], boundary[i] != 0};
    Vec3 p(x[i], y[i], z[i]);
    const Vec3 d(dx[i], dy[i], dz[i]);
    const double L = step_limit[i];
    double safety = -1.0;
    int k = 0, geo = 0;
    for (; k < nsteps && s.ind; ++k) {
      const bool carried = reuse_safety && safety >= L;
      if (!carried) safety = N->Safety(p, s.ind);
      double st;
      if (safety >= L) {  // safetydone
        st = L;
        s.boundary = false;
        safety = safety - L;
      } else {
        State out;
        st = N->Step(p, d, L, s, out);
        s = out;
        safety = -1.0;
        ++geo;
      }
      p = Vec3(p.x + st * d.x, p.y + st * d.y, p.z + st * d.z);
    }
    x[i] = p.x;
    y[i] = p.y;
    z[i] = p.z;
    state[i] = s.ind;
    last[i] = s.last;
    boundary[i] = s.boundary;
    steps[i] += k;
    geo_steps[i] += geo;
  }
  return 0;
}

int orc_nav_approach(orc_geom *g, const double *x, const double *y, const double *z, const double *dx,
                     const double *dy, const double *dz, size_t n, const uint32_t *state, const uint32_t *last,
                     double *step, int nthreads)
{
  Navigator *N = nav_of(g);
  if (!N) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024)
  for (long i = 0; i < long(n); ++i) {
    if (!state[i]) continue;
    step[i] = N->Approach(Vec3(x[i], y[i], z[i]), Vec3(dx[i], dy[i], dz[i]), step[i], State{state[i], last[i], false});
  }
  return 0;
}

int orc_nav_level_locate(orc_geom *g, const double *x, const double *y, const double *z, const double *dx,
                         const double *dy, const double *dz, size_t n, const uint32_t *state, const uint32_t *excl,
                         uint32_t *child, double *qx, double *qy, double *qz, int nthreads)
{
  Navigator *N = nav_of(g);
  if (!N) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024)
  for (long i = 0; i < long(n); ++i) {
    const Vec3 d = dx ? Vec3(dx[i], dy[i], dz[i]) : Vec3();
    Vec3 q;
    child[i] = N->LevelLocateState(state[i], excl ? excl[i] : 0u, Vec3(x[i], y[i], z[i]), dx ? &d : nullptr, q);
    qx[i] = q.x;
    qy[i] = q.y;
    qz[i] = q.z;
  }
  return 0;
}

int orc_nav_safety(orc_geom *g, const double *x, const double *y, const double *z, size_t n, const uint32_t *state,
                   double *out, int nthreads)
{
  Navigator *N = nav_of(g);
  if (!N) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024)
  for (long i = 0; i < long(n); ++i) out[i] = N->Safety(Vec3(x[i], y[i], z[i]), state[i]);
  return 0;
}

}  // extern "C"
