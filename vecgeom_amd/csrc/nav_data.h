// Navigation tables shared by the host builder (build_nav.cpp) and the
// navigation kernels (nav_kernels.hip).
//
// States.  A navigation state is a node of the placement tree: the path from
// the world down to one placed volume.  Nodes are numbered by a depth-first
// walk from the world (1 = world, 0 = outside / none), which plays the role
// of NavStateIndex's uint32 index (VecGeom/navigation/NavStateIndex.h:39-41):
// Push is "child k of this node", Pop is "mother", and every node carries its
// cached global matrix (NavIndexTable, source/NavIndexTable.cpp:9-84).
//
// BVH.  One tree per logical volume with daughters, with the reference's
// layout (complete binary tree in an array, VecGeom/base/BVH.h:364-370): per
// tree node a child count (negative = internal), an offset into the
// primitive-id list and a box; per daughter its box (AABB, 6 doubles).
#pragma once
#include <stdint.h>
#include "vgm_types.h"

namespace vgm {

struct NavNode {
  int32_t pv;        // placed volume at the top of this path
  int32_t mother;    // node of the path one level up (0 for the world)
  int32_t level;     // 0 = world
  int32_t child0;    // children[child0 + k] = node of daughter k
  Transform global;  // master -> local of this path (NavStatePath::TopMatrix)
};

struct BvhRec {
  int32_t nnodes;    // tree nodes, (2 << depth) - 1
  int32_t nprim;     // daughters
  int32_t prim;      // ipool offset: primitive (daughter) ids in tree order
  int32_t offset;    // ipool offset: per node, first primitive
  int32_t nchild;    // ipool offset: per node, primitive count (< 0: internal)
  int32_t pad;
  int64_t node_box;  // bpool offset: per node, 6 doubles (min, max)
  int64_t prim_box;  // bpool offset: per daughter (original order), 6 doubles
};

// A track's navigation state: NavStateIndex's index, last exited index and
// boundary flag.
struct NavState {
  uint32_t ind, last;
  bool boundary;
};

}  // namespace vgm
