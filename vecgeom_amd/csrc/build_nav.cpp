// Navigation tables: placement-tree nodes with cached global matrices and one
// BVH per logical volume with daughters.
//
// The BVH must be the reference's own tree, not just any valid tree: the
// daughter-safety walk (BVH::ComputeSafety) prunes with the running minimum,
// so its result depends on the visiting order.  Construction therefore
// follows source/BVH.cpp: daughter boxes from ABBoxManager::ComputeABBox
// (source/ABBoxManager.cpp:109-143: the 8 corners of the shape's extent taken
// to the mother frame, +-1e-3 mm), a complete binary tree of depth
// min(log2(n/2), 32), surface-area splits over centroid-sorted primitives with
// the reference's comparator, sweep and cost (source/BVH.cpp:190-259), the
// split applied with std::partition, and internal nodes marked by a negative
// count (source/BVH.cpp:24-87).  std::sort and std::partition are the
// standard library's, called on the same sequences with the same
// predicates, so the resulting order is the same.
//
// Global matrices: NavStatePath::TopMatrix multiplies the placements of
// levels 1..N (the world's own placement is skipped) from the right,
// Transformation3D::MultiplyFromRight with the identity/translation/rotation
// flags of Transformation3D::SetProperties (source/Transformation3D.cpp:118-185).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>
#include "geometry.h"
#include "nav_data.h"

namespace vgm {

namespace {


struct AB {
  double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
  void center(double c[3]) const
  {
    for (int k = 0; k < 3; ++k) c[k] = 0.5 * (mx[k] + mn[k]);
  }
  double area() const
  {
    const double e0 = mx[0] - mn[0], e1 = mx[1] - mn[1], e2 = mx[2] - mn[2];
    return 2. * (e0 * e1 + e1 * e2 + e2 * e0);
  }
  static AB unite(const AB &a, const AB &b)
  {
    AB r;
    for (int k = 0; k < 3; ++k) {
      r.mn[k] = a.mn[k] < b.mn[k] ? a.mn[k] : b.mn[k];
      r.mx[k] = a.mx[k] > b.mx[k] ? a.mx[k] : b.mx[k];
    }
    return r;
  }
};

struct TreeBuild {
  const std::vector<AB> &box;
  std::vector<int> prim, offset, nchild;
  std::vector<AB> node;

  // sort key: centroid pushed 1% away from the box minimum
  void key(int a, double k[3]) const
  {
    double c[3];
    box[a].center(c);
    for (int i = 0; i < 3; ++i) k[i] = c[i] + 0.01 * (c[i] - box[a].mn[i]);
  }
  bool less(int a, int b, int axis) const
  {
    double l[3], r[3];
    key(a, l);
    key(b, r);
    const int s1 = (axis + 1) % 3, s2 = (axis + 2) % 3;
    return l[axis] < r[axis] || (l[axis] == r[axis] && (l[s1] < r[s1] || (l[s1] == r[s1] && l[s2] < r[s2])));
  }

  int *split(int *begin, int *end)
  {
    int best_axis = -1, best_obj = -1;
    double best = double(end - begin);
    const long n = end - begin;
    std::vector<std::pair<double, double>> sweep;
    sweep.resize(static_cast<size_t>(n));
    for (int axis = 0; axis <= 2; ++axis) {
      std::sort(begin, end, [&](int a, int b) { return less(a, b, axis); });
      for (auto &s : sweep) s = {0., 0.};
      AB acc = box[*begin];
      for (int *it = begin + 1; it < end; ++it) {
        sweep[it - begin].first = acc.area();
        acc = AB::unite(acc, box[*it]);
      }
      AB acc2 = box[*(end - 1)];
      for (int *it = end - 1; it >= begin; --it) {
        acc2 = AB::unite(acc2, box[*it]);
        sweep[it - begin].second = acc2.area();
      }
      const double total = sweep.front().second;
      for (int *so = begin; so < end; ++so) {
        const long k = so - begin;
        const double left = sweep[k].first / total, right = sweep[k].second / total;
        // the balance term is integer arithmetic in the reference: n/2 - k/n
        const double cost = left * k + right * (end - so) + 0.1 * std::labs(n / 2 - k / n);
        if (cost < best) {
          best = cost;
          best_axis = axis;
          best_obj = *so;
        }
      }
    }
    if (best_axis == -1) return end;
    return std::partition(begin, end, [&](size_t i) { return less(int(i), best_obj, best_axis); });
  }

  void nodes_from(unsigned id, int *first, int *last, unsigned nn)
  {
    if (id >= nn) return;
    nchild[id] = int(last - first);
    offset[id] = int(first - prim.data());
    if (first == last) return;
    node[id] = box[*first];
    for (int *it = first + 1; it != last; ++it) node[id] = AB::unite(node[id], box[*it]);
    if (first + 1 == last) return;
    int *pivot = split(first, last);
    nodes_from(2 * id + 1, first, pivot, nn);
    nodes_from(2 * id + 2, pivot, last, nn);
  }
};

// Transformation3D with its property flags
struct XF {
  Transform T{{0, 0, 0}, {1, 0, 0, 0, 1, 0, 0, 0, 1}};
  bool has_t = false, has_r = false, identity = true;
  void set_properties()
  {
    const double tol = 1e-9;
    has_t = std::fabs(T.t[0]) > tol || std::fabs(T.t[1]) > tol || std::fabs(T.t[2]) > tol;
    int code = 0;
    for (int i = 0; i < 9; ++i) code |= (1 << i) * (std::fabs(T.r[i]) > tol);
    const bool ident_rot = code == 0x111 && T.r[0] == 1. && T.r[4] == 1. && T.r[8] == 1.;
    has_r = !ident_rot;
    identity = !has_t && !has_r;
  }
  void multiply_from_right(const XF &b)
  {
    if (b.identity) return;
    identity = false;
    double *t = T.t, *r = T.r;
    const double *bt = b.T.t, *br = b.T.r;
    if (b.has_t) {
      has_t = true;
      for (int row = 0; row < 3; ++row) {
        t[row] += r[3 * row] * bt[0];
        t[row] += r[3 * row + 1] * bt[1];
        t[row] += r[3 * row + 2] * bt[2];
      }
    }
    if (b.has_r) {
      has_r = true;
      for (int row = 0; row < 3; ++row) {
        const double x = r[3 * row], y = r[3 * row + 1], z = r[3 * row + 2];
        for (int col = 0; col < 3; ++col) r[3 * row + col] = x * br[col];
        for (int col = 0; col < 3; ++col) r[3 * row + col] += y * br[3 + col];
        for (int col = 0; col < 3; ++col) r[3 * row + col] += z * br[6 + col];
      }
    }
  }
};

// local -> master (Transformation3D::InverseTransform)
void to_master(const Transform &T, const double l[3], double m[3])
{
  for (int i = 0; i < 3; ++i) {
    m[i] = T.t[i];
    m[i] += l[0] * T.r[3 * i];
    m[i] += l[1] * T.r[3 * i + 1];
    m[i] += l[2] * T.r[3 * i + 2];
  }
}

}  // namespace

int close_navigation(HostGeometry &g, int world_pv, std::string &err)
{
  if (world_pv < 0 || size_t(world_pv) >= g.pvs.size()) {
    err = "invalid world placed volume";
    return -1;
  }
  g.nav_nodes.assign(1, NavNode{});
  g.nav_children.clear();
  g.bvhs.clear();
  g.nav_ipool.clear();
  g.nav_bpool.clear();
  std::vector<XF> pvx(g.pvs.size());
  for (size_t i = 0; i < g.pvs.size(); ++i) {
    pvx[i].T = g.pvs[i].tr;
    pvx[i].set_properties();
  }
  // trees
  for (size_t l = 0; l < g.lvs.size(); ++l) {
    const std::vector<int32_t> &dl = g.lv_daughters[l];
    g.lvs[l].bvh = -1;
    if (dl.empty()) continue;
    std::vector<AB> boxes(dl.size());
    for (size_t k = 0; k < dl.size(); ++k) {
      const PlacedRec &P = g.pvs[dl[k]];
      const ShapeRec &S = g.shapes[g.lvs[P.lv].shape];
      double lo[3], delta[3];
      for (int c = 0; c < 3; ++c) {
        lo[c] = S.bmin[c];
        delta[c] = S.bmax[c] - S.bmin[c];
      }
      double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
      double mx[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
      for (int x = 0; x <= 1; ++x)
        for (int y = 0; y <= 1; ++y)
          for (int z = 0; z <= 1; ++z) {
            const double c[3] = {lo[0] + x * delta[0], lo[1] + y * delta[1], lo[2] + z * delta[2]};
            double m[3];
            to_master(P.tr, c, m);
            for (int a = 0; a < 3; ++a) {
              mn[a] = std::min(mn[a], m[a]);
              mx[a] = std::max(mx[a], m[a]);
            }
          }
      for (int a = 0; a < 3; ++a) {
        boxes[k].mn[a] = mn[a] - 1E-3;
        boxes[k].mx[a] = mx[a] + 1E-3;
      }
    }
    const int n = int(dl.size());
    TreeBuild T{boxes, {}, {}, {}, {}};
    T.prim.resize(n);
    std::iota(T.prim.begin(), T.prim.end(), 0);
    const int half = n / 2;
    const int depth = std::min(std::max(0, half > 0 ? int(std::log2(double(half))) : 0), 32);
    if (depth > 14) {  // the device walk keeps 16-bit node indices (at most 2^15 - 1 nodes)
      err = "too many daughters in one volume for the BVH walk";
      return -1;
    }
    const unsigned nn = (2u << depth) - 1;
    T.nchild.assign(nn, 0);
    T.offset.assign(nn, -1);
    T.node.assign(nn, AB());
    T.nodes_from(0, T.prim.data(), T.prim.data() + n, nn);
    for (unsigned id = 0; id < nn / 2; ++id)
      if (T.nchild[id] > 8 && T.nchild[id] == T.nchild[2 * id + 1] + T.nchild[2 * id + 2]) T.nchild[id] = -1;
    BvhRec R{};
    R.nnodes = int(nn);
    R.nprim = n;
    R.prim = int(g.nav_ipool.size());
    g.nav_ipool.insert(g.nav_ipool.end(), T.prim.begin(), T.prim.end());
    R.offset = int(g.nav_ipool.size());
    g.nav_ipool.insert(g.nav_ipool.end(), T.offset.begin(), T.offset.end());
    R.nchild = int(g.nav_ipool.size());
    g.nav_ipool.insert(g.nav_ipool.end(), T.nchild.begin(), T.nchild.end());
    R.node_box = int64_t(g.nav_bpool.size());
    for (const AB &b : T.node) {
      g.nav_bpool.insert(g.nav_bpool.end(), b.mn, b.mn + 3);
      g.nav_bpool.insert(g.nav_bpool.end(), b.mx, b.mx + 3);
    }
    R.prim_box = int64_t(g.nav_bpool.size());
    for (const AB &b : boxes) {
      g.nav_bpool.insert(g.nav_bpool.end(), b.mn, b.mn + 3);
      g.nav_bpool.insert(g.nav_bpool.end(), b.mx, b.mx + 3);
    }
    g.lvs[l].bvh = int32_t(g.bvhs.size());
    g.bvhs.push_back(R);
  }
  // placement tree, depth first; a node's children list is written after its subtrees
  auto add = [&](auto &self, int pv, int mother, int level, const XF &parent) -> int {
    const int id = int(g.nav_nodes.size());
    g.nav_nodes.push_back(NavNode{});
    XF m = parent;
    if (level > 0) m.multiply_from_right(pvx[pv]);
    NavNode nd{};
    nd.pv = pv;
    nd.mother = mother;
    nd.level = level;
    nd.child0 = 0;
    nd.global = m.T;
    const std::vector<int32_t> &dl = g.lv_daughters[g.pvs[pv].lv];
    std::vector<int32_t> kids;
    kids.reserve(dl.size());
    for (int32_t d : dl) kids.push_back(self(self, d, id, level + 1, m));
    nd.child0 = int32_t(g.nav_children.size());
    g.nav_children.insert(g.nav_children.end(), kids.begin(), kids.end());
    g.nav_nodes[id] = nd;
    return id;
  };
  XF ident;
  add(add, world_pv, 0, 0, ident);
  if (g.nav_nodes.size() > 0x7fffffffu) {
    err = "placement tree too large";
    return -1;
  }
  g.world_pv = world_pv;
  g.nav_shape_mask = 0;
  for (size_t k = 1; k < g.nav_nodes.size(); ++k)
    g.nav_shape_mask |= 1u << g.shapes[g.lvs[g.pvs[g.nav_nodes[k].pv].lv].shape].type;
  ++g.version;
  return int(g.nav_nodes.size()) - 1;
}

}  // namespace vgm
