// Library context: one device, one stream, one geometry and its device copy.
// Shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "../../include/vgm/vgm.h"
#include "geometry.h"
#include "mlp.h"
#include "vgm_internal.h"

namespace vgm {

// Device mirror of HostGeometry: one allocation, 256-byte aligned sections.
struct DeviceGeometry {
  uint64_t version = 0;
  char *buf = nullptr;
  size_t cap = 0;
  ShapeRec *shapes = nullptr;
  double *pool = nullptr;
  LogicalRec *lvs = nullptr;
  int32_t *daughters = nullptr;
  PlacedRec *pvs = nullptr;
  NavNode *nav_nodes = nullptr;
  int32_t *nav_children = nullptr;
  BvhRec *bvhs = nullptr;
  int32_t *nav_ipool = nullptr;
  double *nav_bpool = nullptr;
};

// The graph-captured navigation step loop of the last VGM_GRAPH transport
// and the arguments it was captured with: replayed while they stay the same.
struct NavGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void *ptrs[10] = {};
  size_t n = 0;
  int nsteps = 0;
  uint64_t version = 0;
  unsigned mask = 0;
};

// Growable device scratch for host-pointer queries.
struct Scratch {
  char *buf = nullptr;
  size_t cap = 0;
};

int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
int sync_geometry(vgm_ctx *c);
int ensure_scratch(vgm_ctx *c, size_t bytes);
size_t align256(size_t x);

}  // namespace vgm

struct vgm_ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  vgm::HostGeometry host;
  vgm::DeviceGeometry dev;
  vgm::Scratch scratch;
  unsigned long long *nav_queue = nullptr;  // work-queue counter of the navigation transports
  vgm::NavGraph nav_graph;
  std::vector<vgm::MlpState> mlps;  // learned safety, indexed by shape id
};

#define VGM_HIP(call)                                           \
  do {                                                          \
    hipError_t e_ = (call);                                     \
    if (e_ != hipSuccess) return ::vgm::hip_fail(e_, #call);    \
  } while (0)
