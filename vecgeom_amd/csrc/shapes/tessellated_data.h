// Tessellated parameter block (TessellatedStruct<3,double>,
// VecGeom/volumes/TessellatedStruct.h:160-200): a header, then the facets in
// the leaf order of a bounding-volume tree, then the tree nodes.  The tree
// replaces the reference's float cluster boxes (HybridManager2); it only
// prunes, so results are the minimum over all facets.
#pragma once
#include <stdint.h>

namespace vgm {

struct TessFacet {   // 24 doubles
  double n[3], d;    // outward unit normal, plane offset (n.p + d)
  double v[3][3];    // vertices, anti-clockwise seen from outside
  double sv[3][3];   // side vectors n x (v[i+1] - v[i]), normalised
  int32_t orig;      // index in definition order (ties go to the lowest)
  int32_t pad0;
  double pad1;
};
static_assert(sizeof(TessFacet) == 24 * sizeof(double), "facet is 24 doubles");

// Node of the facet tree: up to four children, whose boxes it holds so that
// a walk reads one node per level (a binary tree with each node's
// grandchildren pulled up: half the levels, so half the dependent loads on a
// walk's path).  A child is another node (child >= 0) or a leaf: the facets
// [first, first + count) of the leaf order (child < 0; count 0 for an
// unused slot).
constexpr int kTessWidth = 4;
struct TessNode {      // 32 doubles
  double lo[kTessWidth][3], hi[kTessWidth][3];  // child boxes, enlarged by a margin (conservative pruning)
  int32_t child[kTessWidth];
  int32_t first[kTessWidth], count[kTessWidth];
  int32_t pad[kTessWidth];
};
static_assert(sizeof(TessNode) == 32 * sizeof(double), "node is 32 doubles");

struct TessData {
  int32_t nfacets, nnodes;
  double mn[3], mx[3];  // fMinExtent, fMaxExtent (exact vertex extent)
  double test_dir[3];   // fTestDir: direction of the Contains/Inside ray

  __host__ __device__ const TessFacet *facets() const { return reinterpret_cast<const TessFacet *>(this + 1); }
  __host__ __device__ const TessNode *nodes() const
  {
    return reinterpret_cast<const TessNode *>(facets() + nfacets);
  }
  __host__ __device__ static size_t bytes(int nfacets, int nnodes)
  {
    return sizeof(TessData) + sizeof(TessFacet) * nfacets + sizeof(TessNode) * nnodes;
  }
};
static_assert(sizeof(TessData) % sizeof(double) == 0, "header is whole doubles");

// Default Contains/Inside test direction: (sin 1.1 cos 0.7, sin 1.1 sin 0.7, cos 1.1).
constexpr double kTessTestDir[3] = {0.681632986593423, 0.5741315443479861, 0.4535961214255773};

}  // namespace vgm
