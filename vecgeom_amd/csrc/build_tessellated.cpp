// Tessellated builder: facets as Tile<3>::AddVertex makes them
// (VecGeom/volumes/Tile.h:75-150; degenerate facets are dropped, as
// TessellatedStruct::AddTriangularFacet does, VecGeom/volumes/TessellatedStruct.h:596-610),
// side vectors as TessellatedCluster::AddFacet recomputes them
// (VecGeom/volumes/TessellatedCluster.h:180-200), extent as
// TessellatedStruct::AddFacet accumulates it.  The facets are then ordered by
// a bounding-volume tree built here (the reference's cluster grid and float
// box tree are replaced; they only prune).
#include <algorithm>
#include <cmath>
#include <vector>
#include "geometry.h"
#include "shapes/tessellated_data.h"
#include "shapes/tessellated.hpp"

namespace vgm {

namespace {

constexpr double kTol = 1e-9;

struct T3 {
  double x, y, z;
};
T3 sub(const T3 &a, const T3 &b) { return T3{a.x - b.x, a.y - b.y, a.z - b.z}; }
double dot(const T3 &a, const T3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
T3 cross(const T3 &a, const T3 &b) { return T3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
T3 scale(const T3 &a, double s) { return T3{a.x * s, a.y * s, a.z * s}; }
double mag2(const T3 &a) { return dot(a, a); }

bool make_facet(TessFacet &f, const T3 v[3], int orig)
{
  int nvert = 3;
  for (int i = 0; i < 3; ++i)
    if (mag2(sub(v[(i + 1) % 3], v[i])) < kTol) nvert--;
  if (nvert < 3) return false;
  T3 n{0, 0, 0};
  bool degenerated = true;
  for (int i = 0; i < 2 && degenerated; ++i) {
    const T3 e1 = sub(v[i + 1], v[i]);
    if (mag2(e1) < kTol) continue;
    for (int j = i + 1; j < 3; ++j) {
      const T3 e2 = sub(v[(j + 1) % 3], v[j]);
      if (mag2(e2) < kTol) continue;
      n = cross(e1, e2);
      if (mag2(n) < kTol) continue;
      n = scale(n, 1. / std::sqrt(mag2(n)));  // Vector3D::Normalize
      degenerated = false;
      break;
    }
  }
  if (degenerated) return false;
  f.n[0] = n.x;
  f.n[1] = n.y;
  f.n[2] = n.z;
  f.d = -dot(n, v[0]);
  for (int i = 0; i < 3; ++i) {
    f.v[i][0] = v[i].x;
    f.v[i][1] = v[i].y;
    f.v[i][2] = v[i].z;
    const T3 c = cross(n, sub(v[(i + 1) % 3], v[i]));
    const T3 s = scale(c, 1. / std::sqrt(mag2(c)));  // Vector3D::Normalized
    f.sv[i][0] = s.x;
    f.sv[i][1] = s.y;
    f.sv[i][2] = s.z;
  }
  f.orig = orig;
  f.pad0 = 0;
  f.pad1 = 0;
  return true;
}

// Facet tree.  Splits follow a binned surface-area heuristic over the facet
// centroids (fewer boxes per ray than median splits); a tree deeper than the
// device walk allows is rebuilt with median splits on the longest axis.
// Node boxes are the facets' vertex boxes enlarged by `margin`.
struct Builder {
  // binary tree of the build: node boxes, children (-1: leaf), facet range
  struct Node {
    double lo[3], hi[3];
    int left, right, first, count;
  };
  std::vector<TessFacet> &facets;  // reordered in place into leaf order
  std::vector<Node> nodes;
  double margin;
  int leaf = 1;  // facets per leaf
  bool sah = true;
  int max_depth = 0;

  static constexpr int kBins = 32;

  static double centroid(const TessFacet &f, int c) { return f.v[0][c] + f.v[1][c] + f.v[2][c]; }

  struct Box {
    double lo[3] = {1e308, 1e308, 1e308}, hi[3] = {-1e308, -1e308, -1e308};
    void add(const TessFacet &f)
    {
      for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) {
          lo[c] = std::min(lo[c], f.v[k][c]);
          hi[c] = std::max(hi[c], f.v[k][c]);
        }
    }
    void add(const Box &o)
    {
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::min(lo[c], o.lo[c]);
        hi[c] = std::max(hi[c], o.hi[c]);
      }
    }
    double half_area() const
    {
      const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
      return dx < 0 ? 0.0 : dx * dy + dy * dz + dz * dx;
    }
  };

  static int bin(double x, double lo, double scale)
  {
    const int k = int((x - lo) * scale);
    return k < 0 ? 0 : (k >= kBins ? kBins - 1 : k);
  }

  // SAH split of [b, e): returns the split position, or -1 when no bin
  // boundary separates the centroids.
  int sah_split(int b, int e, const double *clo, const double *chi)
  {
    double best = 1e308;
    int best_axis = -1, best_k = -1;
    for (int c = 0; c < 3; ++c) {
      if (!(chi[c] > clo[c])) continue;
      const double scale = kBins / (chi[c] - clo[c]);
      Box bins[kBins];
      int cnt[kBins] = {0};
      for (int i = b; i < e; ++i) {
        const int k = bin(centroid(facets[i], c), clo[c], scale);
        ++cnt[k];
        bins[k].add(facets[i]);
      }
      Box right[kBins];
      int rcnt[kBins];
      Box acc;
      int n = 0;
      for (int k = kBins - 1; k >= 0; --k) {
        acc.add(bins[k]);
        n += cnt[k];
        right[k] = acc;
        rcnt[k] = n;
      }
      Box left;
      int lc = 0;
      for (int k = 0; k + 1 < kBins; ++k) {
        left.add(bins[k]);
        lc += cnt[k];
        if (lc == 0 || rcnt[k + 1] == 0) continue;
        const double cost = left.half_area() * lc + right[k + 1].half_area() * rcnt[k + 1];
        if (cost < best) {
          best = cost;
          best_axis = c;
          best_k = k;
        }
      }
    }
    if (best_axis < 0) return -1;
    const double scale = kBins / (chi[best_axis] - clo[best_axis]);
    const auto it = std::stable_partition(facets.begin() + b, facets.begin() + e, [&](const TessFacet &f) {
      return bin(centroid(f, best_axis), clo[best_axis], scale) <= best_k;
    });
    const int mid = int(it - facets.begin());
    return (mid == b || mid == e) ? -1 : mid;
  }

  int median_split(int b, int e, const double *clo, const double *chi)
  {
    int axis = 0;
    for (int c = 1; c < 3; ++c)
      if (chi[c] - clo[c] > chi[axis] - clo[axis]) axis = c;
    const int mid = b + (e - b) / 2;
    std::nth_element(facets.begin() + b, facets.begin() + mid, facets.begin() + e,
                     [axis](const TessFacet &x, const TessFacet &y) {
                       const double cx = centroid(x, axis), cy = centroid(y, axis);
                       return cx < cy || (cx == cy && x.orig < y.orig);
                     });
    return mid;
  }

  int build(int b, int e, int depth = 0)
  {
    max_depth = std::max(max_depth, depth);
    const int id = int(nodes.size());
    nodes.push_back(Node{});
    Box box;
    double clo[3] = {1e308, 1e308, 1e308}, chi[3] = {-1e308, -1e308, -1e308};
    for (int i = b; i < e; ++i) {
      box.add(facets[i]);
      for (int c = 0; c < 3; ++c) {
        clo[c] = std::min(clo[c], centroid(facets[i], c));
        chi[c] = std::max(chi[c], centroid(facets[i], c));
      }
    }
    for (int c = 0; c < 3; ++c) {
      nodes[id].lo[c] = box.lo[c] - margin;
      nodes[id].hi[c] = box.hi[c] + margin;
    }
    if (e - b <= leaf) {
      nodes[id].left = nodes[id].right = -1;
      nodes[id].first = b;
      nodes[id].count = e - b;
      return id;
    }
    int mid = sah ? sah_split(b, e, clo, chi) : -1;
    if (mid < 0) mid = median_split(b, e, clo, chi);
    const int l = build(b, mid, depth + 1);
    const int r = build(mid, e, depth + 1);
    nodes[id].left = l;
    nodes[id].right = r;
    nodes[id].first = nodes[id].count = 0;
    return id;
  }

  // Children of the node record of binary node o: each child of o, or, for
  // an internal child, its two children (grandchildren pulled up).
  void record_slots(int o, int *slot, int &ns) const
  {
    ns = 0;
    const Node &nd = nodes[size_t(o)];
    for (const int c : {nd.left, nd.right}) {
      const Node &cn = nodes[size_t(c)];
      if (cn.left >= 0) {
        slot[ns++] = cn.left;
        slot[ns++] = cn.right;
      } else {
        slot[ns++] = c;
      }
    }
  }

  // The device layout: one record per node of the four-wide tree, preorder,
  // holding the boxes of its children; leaves are folded into their parents.
  // A tree that is a single leaf gets a root with one leaf child.  `depth`
  // is the number of records on the longest root-to-leaf path.
  std::vector<TessNode> wide(int &depth) const
  {
    std::vector<TessNode> out;
    depth = 1;
    if (nodes[0].left < 0) {
      TessNode w{};
      for (int k = 0; k < 3; ++k) {
        w.lo[0][k] = nodes[0].lo[k];
        w.hi[0][k] = nodes[0].hi[k];
      }
      for (int i = 0; i < kTessWidth; ++i) w.child[i] = -1;
      w.first[0] = nodes[0].first;
      w.count[0] = nodes[0].count;
      out.push_back(w);
      return out;
    }
    std::vector<int> map(nodes.size(), -1), level(nodes.size(), 0);
    std::vector<int> todo{0}, order;
    level[0] = 1;
    while (!todo.empty()) {  // preorder numbering of the records
      const int o = todo.back();
      todo.pop_back();
      map[size_t(o)] = int(order.size());
      order.push_back(o);
      depth = std::max(depth, level[size_t(o)]);
      int slot[kTessWidth], ns;
      record_slots(o, slot, ns);
      for (int i = ns - 1; i >= 0; --i)
        if (nodes[size_t(slot[i])].left >= 0) {
          level[size_t(slot[i])] = level[size_t(o)] + 1;
          todo.push_back(slot[i]);
        }
    }
    for (const int o : order) {
      TessNode w{};
      int slot[kTessWidth], ns;
      record_slots(o, slot, ns);
      for (int i = 0; i < kTessWidth; ++i) {
        w.child[i] = -1;
        if (i >= ns) continue;  // unused: an empty leaf with a point box
        const Node &c = nodes[size_t(slot[i])];
        for (int k = 0; k < 3; ++k) {
          w.lo[i][k] = c.lo[k];
          w.hi[i][k] = c.hi[k];
        }
        if (c.left >= 0) {
          w.child[i] = map[size_t(slot[i])];
        } else {
          w.first[i] = c.first;
          w.count[i] = c.count;
        }
      }
      out.push_back(w);
    }
    return out;
  }
};

}  // namespace

int build_tessellated(HostGeometry &g, int nfacets, const double *tri, const double *test_dir, std::string &err)
{
  if (nfacets < 4 || !tri) {
    err = "tessellated solid needs at least 4 facets";
    return -1;
  }
  std::vector<TessFacet> facets;
  facets.reserve(nfacets);
  double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
  double mx[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
  for (int i = 0; i < nfacets; ++i) {
    const double *t = tri + 9 * i;
    const T3 v[3] = {{t[0], t[1], t[2]}, {t[3], t[4], t[5]}, {t[6], t[7], t[8]}};
    TessFacet f;
    if (!make_facet(f, v, int(facets.size()))) continue;
    facets.push_back(f);
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c) {
        mn[c] = mn[c] < f.v[k][c] ? mn[c] : f.v[k][c];
        mx[c] = mx[c] > f.v[k][c] ? mx[c] : f.v[k][c];
      }
  }
  const int n = int(facets.size());
  if (n < 4) {
    err = "tessellated solid has fewer than 4 valid facets";
    return -1;
  }
  double big = 0;
  for (int c = 0; c < 3; ++c) big = std::max(big, std::max(std::fabs(mn[c]), std::fabs(mx[c])));
  // One facet per leaf walks fastest (box tests are cheaper than facet
  // tests); larger solids use the smallest leaf size whose tree keeps node
  // indices within 16 bits.  The input facet order is kept so that every
  // rebuild starts from the same sequence.
  // The device walks keep 16-bit node indices on a stack of kTessWalkDepth
  // entries; a nearest-first walk defers at most kTessWidth - 1 siblings per
  // level, so the four-wide depth D needs (kTessWidth - 1) D + 1 entries.
  const std::vector<TessFacet> input = facets;
  Builder B{facets, {}, 1e-7 + 1e-12 * big};
  std::vector<TessNode> wide;
  int wdepth = 0;
  auto fits = [&]() { return (kTessWidth - 1) * wdepth + 1 <= kTessWalkDepth; };
  for (;;) {
    facets = input;
    B.nodes.clear();
    B.max_depth = 0;
    B.nodes.reserve(2 * size_t(n) / size_t(B.leaf) + 1);
    B.build(0, n);
    wide = B.wide(wdepth);
    if (!fits() && B.sah) {
      B.sah = false;  // median splits: depth log2(n / leaf)
      continue;
    }
    if (wide.size() <= 65535 || B.leaf >= n) break;
    B.leaf *= 2;
  }
  const int nn = int(wide.size());
  if (nn > 65535 || !fits()) {
    err = "tessellated solid too large for the facet tree walk (" + std::to_string(n) + " facets)";
    return -1;
  }
  std::vector<double> block(TessData::bytes(n, nn) / sizeof(double), 0.0);
  TessData *h = reinterpret_cast<TessData *>(block.data());
  h->nfacets = n;
  h->nnodes = nn;
  for (int c = 0; c < 3; ++c) {
    h->mn[c] = mn[c];
    h->mx[c] = mx[c];
    h->test_dir[c] = test_dir ? test_dir[c] : kTessTestDir[c];
  }
  std::copy(facets.begin(), facets.end(), const_cast<TessFacet *>(h->facets()));
  std::copy(wide.begin(), wide.end(), const_cast<TessNode *>(h->nodes()));
  return g.add_shape(kTessellated, block.data(), block.size(), n, nn, mn, mx);
}

}  // namespace vgm
