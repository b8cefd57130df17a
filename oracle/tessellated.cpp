// Oracle: tessellated solid of triangular facets (TEST INFRASTRUCTURE ONLY),
// scalar backend.
//   facets: Tile<3,double>::AddVertex (VecGeom/volumes/Tile.h:75-150)
//   per-facet queries as computed lane-wise by TessellatedCluster
//     (VecGeom/volumes/TessellatedCluster.h: DistanceToCluster :247-290,
//     InsideCluster :228-236, SafetySq :503-545) with
//     DistanceToLineSegmentSquared2 (VecGeom/volumes/kernel/GenericKernels.h:224-249)
//   solid queries: VecGeom/volumes/kernel/TessellatedImplementation.h:62-318
//     (Contains/Inside by a ray along fTestDir, DistanceToSolid with its
//     special cases, SafetySq), bounding-box pre-check
//     BoxImplementation::IntersectCachedKernel2 (BoxImplementation.h:326-368).
//
// The reference visits facet clusters in the order of a float box tree and
// keeps the minimum with strict '<'.  Up to exact ties (and hits within 1e-9
// of a cluster box face) that equals the minimum over all facets; here ties
// go to the lowest facet index.  Results beyond a finite step_max are not
// pinned (the reference stops its cluster walk there).  The reference draws
// fTestDir at random in Close(); here it is a constructor argument.
//
// Queries run either over all facets (brute force, the definition) or over a
// bounding-volume tree with conservative pruning; the tests check that both
// agree.
#include <algorithm>
#include <memory>
#include <vector>
#include "core.h"

namespace orc {

namespace {

struct Facet {
  Vec3 v[3];
  Vec3 sv[3];  // side vectors n x (v[i+1] - v[i]), normalised
  Vec3 n;
  double d = 0;
};

Vec3 Normalized(const Vec3 &v) { return v * (1. / v.Mag()); }

// Tile<3>::AddVertex for the last vertex: normal from the first two
// non-degenerate edges, distance from the first vertex, side vectors.
bool make_facet(Facet &f, const Vec3 &a, const Vec3 &b, const Vec3 &c)
{
  f.v[0] = a;
  f.v[1] = b;
  f.v[2] = c;
  int nvert = 3;
  for (int i = 0; i < 3; ++i)
    if ((f.v[(i + 1) % 3] - f.v[i]).Mag2() < kTolerance) nvert--;
  if (nvert < 3) return false;
  bool degenerated = true;
  for (int i = 0; i < 2; ++i) {
    Vec3 e1 = f.v[i + 1] - f.v[i];
    if (e1.Mag2() < kTolerance) continue;
    for (int j = i + 1; j < 3; ++j) {
      Vec3 e2 = f.v[(j + 1) % 3] - f.v[j];
      if (e2.Mag2() < kTolerance) continue;
      f.n = Cross(e1, e2);
      if (f.n.Mag2() < kTolerance) continue;
      f.n = f.n * (1. / f.n.Mag());
      degenerated = false;
      break;
    }
    if (!degenerated) break;
  }
  if (degenerated) return false;
  f.d = -Dot(f.n, f.v[0]);
  // TessellatedCluster::AddFacet recomputes the side vectors per edge
  for (int i = 0; i < 3; ++i) f.sv[i] = Normalized(Cross(f.n, f.v[(i + 1) % 3] - f.v[i]));
  return true;
}

// InsideCluster for one facet
bool facet_contains(const Facet &f, const Vec3 &p)
{
  bool inside = true;
  for (int i = 0; i < 3; ++i) inside &= Dot(p - f.v[i], f.sv[i]) > -kTolerance;
  return inside;
}

// DistanceToLineSegmentSquared2 for one lane with mask = true
double segment_dist2(const Vec3 &c0, const Vec3 &c1, const Vec3 &p)
{
  const Vec3 line = c1 - c0;
  const Vec3 projection = p - c0;
  const double dot0 = Dot(projection, line);
  if (dot0 <= 0) return (p - c0).Mag2();
  const double dot1 = line.Mag2();
  double result = kInfLength;
  if (dot1 <= dot0) result = (p - c1).Mag2();
  if (result < kInfLength) return result;
  return ((c0 + line * (dot0 / dot1)) - p).Mag2();
}

// TessellatedCluster::SafetySq for one lane
template <bool ToIn>
double facet_safety2(const Facet &f, const Vec3 &p)
{
  double s = Dot(p, f.n) + f.d;
  const Vec3 x = p - f.n * s;
  bool within = facet_contains(f, x);
  if (ToIn)
    within &= s > -kTolerance;
  else
    within &= s < kTolerance;
  s *= s;
  if (within) return s;
  const double e0 = segment_dist2(f.v[0], f.v[1], p);
  const double e1 = segment_dist2(f.v[1], f.v[2], p);
  const double e2 = segment_dist2(f.v[2], f.v[0], p);
  return Min(Min(e0, e1), e2);
}

struct Node {
  Vec3 lo, hi;
  int left = -1, right = -1;  // children, or -1 for a leaf
  int first = 0, count = 0;   // facet range in `order` for leaves
};

}  // namespace

struct Tessellated final : Shape {
  std::vector<Facet> f;
  Vec3 mn{kInfLength, kInfLength, kInfLength}, mx{-kInfLength, -kInfLength, -kInfLength};
  Vec3 testDir;
  bool brute = false;
  std::vector<Node> nodes;
  std::vector<int> order;
  double margin = 0;

  bool Build(int n, const double *tri, const double *test_dir)
  {
    // AddTriangularFacet drops a degenerate facet (Tile::AddVertex fails)
    f.clear();
    f.reserve(n);
    for (int i = 0; i < n; ++i) {
      const double *t = tri + 9 * i;
      Facet ft;
      if (!make_facet(ft, Vec3(t[0], t[1], t[2]), Vec3(t[3], t[4], t[5]), Vec3(t[6], t[7], t[8]))) continue;
      f.push_back(ft);
      // TessellatedStruct::AddFacet extent update
      for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) {
          mn[c] = Min(mn[c], ft.v[k][c]);
          mx[c] = Max(mx[c], ft.v[k][c]);
        }
    }
    n = int(f.size());
    if (n < 4) return false;
    testDir = Vec3(test_dir[0], test_dir[1], test_dir[2]);
    double big = 0;
    for (int c = 0; c < 3; ++c) big = Max(big, Max(Abs(mn[c]), Abs(mx[c])));
    margin = 1e-7 + 1e-12 * big;
    order.resize(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    nodes.clear();
    nodes.reserve(2 * n);
    build_node(0, n);
    return true;
  }

  // Median split on the longest centroid axis; leaves of up to 4 facets.
  int build_node(int b, int e)
  {
    const int id = int(nodes.size());
    nodes.push_back(Node());
    Vec3 lo(kInfLength, kInfLength, kInfLength), hi(-kInfLength, -kInfLength, -kInfLength);
    Vec3 clo = lo, chi = hi;
    for (int i = b; i < e; ++i) {
      const Facet &t = f[order[i]];
      for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) {
          lo[c] = Min(lo[c], t.v[k][c]);
          hi[c] = Max(hi[c], t.v[k][c]);
        }
      for (int c = 0; c < 3; ++c) {
        const double m = (t.v[0][c] + t.v[1][c] + t.v[2][c]) / 3.;
        clo[c] = Min(clo[c], m);
        chi[c] = Max(chi[c], m);
      }
    }
    for (int c = 0; c < 3; ++c) {
      lo[c] -= margin;
      hi[c] += margin;
    }
    nodes[id].lo = lo;
    nodes[id].hi = hi;
    if (e - b <= 4) {
      nodes[id].first = b;
      nodes[id].count = e - b;
      return id;
    }
    int axis = 0;
    for (int c = 1; c < 3; ++c)
      if (chi[c] - clo[c] > chi[axis] - clo[axis]) axis = c;
    const int mid = (b + e) / 2;
    std::nth_element(order.begin() + b, order.begin() + mid, order.begin() + e, [&](int x, int y) {
      const double mx_ = f[x].v[0][axis] + f[x].v[1][axis] + f[x].v[2][axis];
      const double my_ = f[y].v[0][axis] + f[y].v[1][axis] + f[y].v[2][axis];
      return mx_ < my_ || (mx_ == my_ && x < y);
    });
    const int l = build_node(b, mid);
    const int r = build_node(mid, e);
    nodes[id].left = l;
    nodes[id].right = r;
    return id;
  }

  // Entry parameter of the ray into a node box, or +inf when it misses.
  static double box_entry(const Node &nd, const Vec3 &p, const Vec3 &inv)
  {
    double t0 = -kInfLength, t1 = kInfLength;
    for (int c = 0; c < 3; ++c) {
      double a = (nd.lo[c] - p[c]) * inv[c], b = (nd.hi[c] - p[c]) * inv[c];
      if (a > b) std::swap(a, b);
      t0 = Max(t0, a);
      t1 = Min(t1, b);
    }
    return (t0 <= t1) ? t0 : kInfLength;
  }

  static double box_dist2(const Node &nd, const Vec3 &p)
  {
    double s = 0;
    for (int c = 0; c < 3; ++c) {
      double e = 0;
      if (p[c] < nd.lo[c]) e = nd.lo[c] - p[c];
      if (p[c] > nd.hi[c]) e = p[c] - nd.hi[c];
      s += e * e;
    }
    return s;
  }

  // Minimum entering (ToIn) and exiting (ToOut) facet hits along the ray.
  // `to_in` selects which of the two bounds the tree walk: hits of the other
  // kind beyond the first kind's minimum are not needed by DistanceToSolid.
  void ray_hits(const Vec3 &p, const Vec3 &d, bool to_in, double &dIn, int &iIn, double &dOut, int &iOut) const
  {
    dIn = dOut = kInfLength;
    iIn = iOut = -1;
    auto visit = [&](int i) {
      const Facet &t = f[i];
      const double ndd = NonZero(Dot(d, t.n));
      const double saf = Dot(p, t.n) + t.d;
      const bool vIn = ndd < 0. && saf > -kTolerance;
      const bool vOut = ndd > 0. && saf < kTolerance;
      if (!vIn && !vOut) return;
      const double dist = -saf / ndd;
      if (!facet_contains(t, p + d * dist)) return;
      if (vIn) {
        if (dist < dIn || (dist == dIn && i < iIn)) {
          dIn = dist;
          iIn = i;
        }
      } else if (dist < dOut || (dist == dOut && i < iOut)) {
        dOut = dist;
        iOut = i;
      }
    };
    if (brute) {
      for (int i = 0; i < int(f.size()); ++i) visit(i);
      return;
    }
    const Vec3 inv(1. / NonZero(d.x), 1. / NonZero(d.y), 1. / NonZero(d.z));
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
      const Node &nd = nodes[stack[--sp]];
      const double bound = to_in ? (iIn >= 0 ? dIn : kInfLength) : (iOut >= 0 ? dOut : kInfLength);
      const double te = box_entry(nd, p, inv);
      if (te == kInfLength || te > bound) continue;
      if (nd.left < 0) {
        for (int k = nd.first; k < nd.first + nd.count; ++k) visit(order[k]);
      } else if (box_entry(nodes[nd.left], p, inv) <= box_entry(nodes[nd.right], p, inv)) {
        stack[sp++] = nd.right;  // nearer child last: visited first
        stack[sp++] = nd.left;
      } else {
        stack[sp++] = nd.left;
        stack[sp++] = nd.right;
      }
    }
  }

  template <bool ToIn>
  double safety2(const Vec3 &p) const
  {
    double best = kInfLength;
    if (brute) {
      for (const Facet &t : f) {
        const double s = facet_safety2<ToIn>(t, p);
        if (s < best) best = s;
      }
      return best;
    }
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
      const Node &nd = nodes[stack[--sp]];
      if (box_dist2(nd, p) > best) continue;
      if (nd.left < 0) {
        for (int k = nd.first; k < nd.first + nd.count; ++k) {
          const double s = facet_safety2<ToIn>(f[order[k]], p);
          if (s < best) best = s;
        }
      } else if (box_dist2(nodes[nd.left], p) <= box_dist2(nodes[nd.right], p)) {
        stack[sp++] = nd.right;  // nearer child last: visited first
        stack[sp++] = nd.left;
      } else {
        stack[sp++] = nd.left;
        stack[sp++] = nd.right;
      }
    }
    return best;
  }

  // BoxImplementation::IntersectCachedKernel2 with t0 = -kTolerance, t1 = kInfLength
  double bbox_intersect(const Vec3 &p, const Vec3 &d) const
  {
    const Vec3 inv(1.0 / NonZero(d.x), 1.0 / NonZero(d.y), 1.0 / NonZero(d.z));
    const int sx = inv.x < 0, sy = inv.y < 0, sz = inv.z < 0;
    const Vec3 *corners[2] = {&mn, &mx};
    double tmin = ((*corners[sx]).x - p.x) * inv.x;
    double tymax = ((*corners[1 - sy]).y - p.y) * inv.y;
    if (tmin > tymax) return kInfLength;
    double tmax = ((*corners[1 - sx]).x - p.x) * inv.x;
    double tymin = ((*corners[sy]).y - p.y) * inv.y;
    if (tymin > tmax) return kInfLength;
    tmin = Max(tmin, tymin);
    tmax = Min(tmax, tymax);
    const double tzmin = ((*corners[sz]).z - p.z) * inv.z;
    const double tzmax = ((*corners[1 - sz]).z - p.z) * inv.z;
    if ((tmin > tzmax) || (tzmin > tmax)) return kInfLength;
    tmin = Max(tmin, tzmin);
    tmax = Min(tmax, tzmax);
    if (!((tmin <= kInfLength + kTolerance) && (tmax > -kTolerance - kTolerance))) return kInfLength;
    return tmin;
  }

  // TessellatedImplementation::DistanceToSolid<ToIn>
  template <bool ToIn>
  void distance_to_solid(const Vec3 &p, const Vec3 &d, double stepMax, double &distance, int &isurf,
                         double &distother, int &isurfother) const
  {
    isurf = -1;
    isurfother = -1;
    if (ToIn) {
      distance = bbox_intersect(p, d);
      if (distance >= stepMax) return;
    }
    double dIn, dOut;
    int iIn, iOut;
    ray_hits(p, d, ToIn, dIn, iIn, dOut, iOut);
    distance = kInfLength;
    distother = kInfLength;
    if (ToIn) {
      if (iIn >= 0) {
        isurf = iIn;
        distance = dIn;
      }
      if (iOut >= 0) {
        isurfother = iOut;
        distother = dOut;
      }
      if (iIn < 0) {
        if (iOut >= 0 && dOut * Dot(d, f[iOut].n) > kTolerance) distance = -1.;
      } else {
        if (iOut >= 0 && dOut > kTolerance && dOut < dIn) distance = -1.;
      }
    } else {
      if (iOut >= 0) {
        isurf = iOut;
        distance = dOut;
      }
      if (iIn >= 0) {
        isurfother = iIn;
        distother = dIn;
      }
      if (iOut < 0) {
        distance = -1.;
      } else if (iIn >= 0 && dIn < dOut && dIn * Dot(d, f[iIn].n) < -kTolerance) {
        distance = -1.;
        isurf = -1;
      }
    }
  }

  bool Contains(const Vec3 &p) const override
  {
    double dOut, dIn;
    int iOut, iIn;
    distance_to_solid<false>(p, testDir, kInfLength, dOut, iOut, dIn, iIn);
    return iOut >= 0;
  }

  int Inside(const Vec3 &p) const override
  {
    double dOut, dIn;
    int iOut, iIn;
    distance_to_solid<false>(p, testDir, kInfLength, dOut, iOut, dIn, iIn);
    if (iOut < 0) return kOutside;
    if (dOut < 0 || dOut * Dot(testDir, f[iOut].n) < kTolerance) return kSurface;
    if (iIn < 0 || dOut < dIn) return kInside;
    if (dIn < 0 || dIn * Dot(testDir, f[iIn].n) > -kTolerance) return kSurface;
    return kOutside;
  }

  double DistanceToIn(const Vec3 &p, const Vec3 &d, double stepMax) const override
  {
    double dist, other;
    int is, io;
    distance_to_solid<true>(p, d, stepMax, dist, is, other, io);
    return dist;
  }

  double DistanceToOut(const Vec3 &p, const Vec3 &d, double stepMax) const override
  {
    double dist, other;
    int is, io;
    distance_to_solid<false>(p, d, stepMax, dist, is, other, io);
    return dist;
  }

  double SafetyToIn(const Vec3 &p) const override
  {
    if (Contains(p)) return -1.;
    return Sqrt(safety2<true>(p));
  }

  double SafetyToOut(const Vec3 &p) const override
  {
    if (!Contains(p)) return -1.;
    return Sqrt(safety2<false>(p));
  }

  // UnplacedTessellated::Normal (source/UnplacedTessellated.cpp:83-89,
  // TessellatedImplementation.h:163-172): the normal of the facet nearest to
  // p by SafetySq<ToIn = false>, valid everywhere; over all facets in
  // definition order with strict '<', so ties go to the facet defined first.
  bool Normal(const Vec3 &p, Vec3 &n) const override
  {
    double best = kInfLength;
    int isurf = -1;
    for (size_t k = 0; k < f.size(); ++k) {
      const double s = facet_safety2<false>(f[k], p);
      if (s < best) {
        best = s;
        isurf = int(k);
      }
    }
    n = f[size_t(isurf)].n;
    return true;
  }
  bool HasNormal() const override { return true; }

  void Extent(Vec3 &a, Vec3 &b) const override
  {
    a = mn;
    b = mx;
  }
};

Shape *make_tessellated(int nfacets, const double *triangles, const double *test_dir, bool brute_force)
{
  if (nfacets < 4 || !triangles || !test_dir) return nullptr;
  auto *t = new Tessellated();
  t->brute = brute_force;
  if (!t->Build(nfacets, triangles, test_dir)) {
    delete t;
    return nullptr;
  }
  return t;
}

}  // namespace orc
