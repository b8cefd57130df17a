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

struct Builder {
  std::vector<TessFacet> &facets;  // reordered in place into leaf order
  std::vector<TessNode> nodes;
  double margin;
  int leaf = 1;  // facets per leaf

  static double centroid(const TessFacet &f, int c) { return f.v[0][c] + f.v[1][c] + f.v[2][c]; }

  int max_depth = 0;

  int build(int b, int e, int depth = 0)
  {
    max_depth = std::max(max_depth, depth);
    const int id = int(nodes.size());
    nodes.push_back(TessNode{});
    double lo[3] = {1e308, 1e308, 1e308}, hi[3] = {-1e308, -1e308, -1e308};
    double clo[3] = {1e308, 1e308, 1e308}, chi[3] = {-1e308, -1e308, -1e308};
    for (int i = b; i < e; ++i)
      for (int c = 0; c < 3; ++c) {
        for (int k = 0; k < 3; ++k) {
          lo[c] = std::min(lo[c], facets[i].v[k][c]);
          hi[c] = std::max(hi[c], facets[i].v[k][c]);
        }
        clo[c] = std::min(clo[c], centroid(facets[i], c));
        chi[c] = std::max(chi[c], centroid(facets[i], c));
      }
    for (int c = 0; c < 3; ++c) {
      nodes[id].lo[c] = lo[c] - margin;
      nodes[id].hi[c] = hi[c] + margin;
    }
    if (e - b <= leaf) {
      nodes[id].left = nodes[id].right = -1;
      nodes[id].first = b;
      nodes[id].count = e - b;
      return id;
    }
    int axis = 0;
    for (int c = 1; c < 3; ++c)
      if (chi[c] - clo[c] > chi[axis] - clo[axis]) axis = c;
    const int mid = b + (e - b) / 2;
    std::nth_element(facets.begin() + b, facets.begin() + mid, facets.begin() + e,
                     [axis](const TessFacet &x, const TessFacet &y) {
                       const double cx = centroid(x, axis), cy = centroid(y, axis);
                       return cx < cy || (cx == cy && x.orig < y.orig);
                     });
    const int l = build(b, mid, depth + 1);
    const int r = build(mid, e, depth + 1);
    nodes[id].left = l;
    nodes[id].right = r;
    nodes[id].first = nodes[id].count = 0;
    return id;
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
  // indices within 16 bits.
  Builder B{facets, {}, 1e-7 + 1e-12 * big};
  for (;;) {
    B.nodes.clear();
    B.max_depth = 0;
    B.nodes.reserve(2 * size_t(n) / size_t(B.leaf) + 1);
    B.build(0, n);
    if (B.nodes.size() <= 65535 || B.leaf >= n) break;
    B.leaf *= 2;
  }
  const int nn = int(B.nodes.size());
  // the device walks keep 16-bit node indices, at most one deferred sibling
  // per level, on a stack of kTessWalkDepth entries
  if (nn > 65535 || B.max_depth > kTessWalkDepth - 2) {
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
  std::copy(B.nodes.begin(), B.nodes.end(), const_cast<TessNode *>(h->nodes()));
  return g.add_shape(kTessellated, block.data(), block.size(), n, nn, mn, mx);
}

}  // namespace vgm
