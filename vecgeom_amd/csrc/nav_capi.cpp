// Navigation entry points of the C ABI (include/vgm/vgm.h).
#include <cstring>
#include <initializer_list>
#include "context.h"
#include "nav_kernels.h"

using namespace vgm;

namespace {

// One array argument of a navigation call.  Host arrays are staged through
// the context's device scratch; device arrays are used in place.  `slot` is
// the NavArgs field that receives the device address.
struct Arr {
  const void *user;
  size_t bytes;
  bool in, out;
  void **slot;
  bool optional = false;  // may be null
};

// The straight-line transport as nsteps launches of the one-step kernel,
// captured into a HIP graph on the context stream and replayed; captured
// again when the arrays, sizes, step count or geometry change.
int nav_graph_transport(vgm_ctx *c, const NavArgs &a)
{
  NavGraph &G = c->nav_graph;
  const void *ptrs[10] = {a.x, a.y, a.z, a.dx, a.dy, a.dz, a.state, a.last, a.boundary, a.steps};
  bool same = G.exec && G.n == a.n && G.nsteps == a.nsteps && G.version == c->dev.version && G.mask == a.shape_mask;
  for (int k = 0; k < 10 && same; ++k) same = G.ptrs[k] == ptrs[k];
  if (!same) {
    if (G.exec) VGM_HIP(hipGraphExecDestroy(G.exec));
    if (G.graph) VGM_HIP(hipGraphDestroy(G.graph));
    G = NavGraph{};
    VGM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int k = 0; k < a.nsteps && e == hipSuccess; ++k) e = launch_nav(c->stream, kNavStepMove, a);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return hip_fail(e != hipSuccess ? e : e2, "capturing the navigation step loop");
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return hip_fail(e, "hipGraphInstantiate");
    }
    G.graph = graph;
    G.exec = exec;
    for (int k = 0; k < 10; ++k) G.ptrs[k] = ptrs[k];
    G.n = a.n;
    G.nsteps = a.nsteps;
    G.version = c->dev.version;
    G.mask = a.shape_mask;
  }
  VGM_HIP(hipGraphLaunch(G.exec, c->stream));
  return VGM_OK;
}

int run_nav(vgm_ctx *c, int kind, NavArgs &a, std::initializer_list<Arr> arrs, unsigned flags)
{
  if (!c) return fail(VGM_EINVAL, "null context");
  if (c->host.nav_nodes.size() < 2) return fail(VGM_ESTATE, "navigation tables not built: call vgm_nav_close");
  for (const Arr &r : arrs)
    if (!r.user && !r.optional && a.n) return fail(VGM_EINVAL, "null array argument");
  VGM_HIP(hipSetDevice(c->device));
  int rc = sync_geometry(c);
  if (rc) return rc;
  const DeviceGeometry &d = c->dev;
  a.g = GeomView{d.shapes, d.pool, d.lvs, d.daughters, d.pvs, d.nav_nodes, d.nav_children, d.bvhs,
                 d.nav_ipool, d.nav_bpool, c->host.world_pv};
  a.shape_mask = c->host.nav_shape_mask;
  a.num_cus = c->num_cus;
  if (a.n == 0) return VGM_OK;
  if ((kind == kNavTransport || kind == kNavTransportLimited) && !c->nav_queue)
    VGM_HIP(hipMalloc(&c->nav_queue, 256));
  a.queue = c->nav_queue;
  if (flags & VGM_DEVICE_PTRS) {
    for (const Arr &r : arrs) *r.slot = const_cast<void *>(r.user);
    if ((flags & VGM_GRAPH) && kind == kNavTransport) {
      if (a.nsteps > 0 && (rc = nav_graph_transport(c, a))) return rc;
      if (!(flags & VGM_ASYNC)) VGM_HIP(hipStreamSynchronize(c->stream));
      return VGM_OK;
    }
    if ((flags & VGM_STEP_KERNELS) && kind == kNavTransport) {
      for (int k = 0; k < a.nsteps; ++k) VGM_HIP(launch_nav(c->stream, kNavStepMove, a));
      if (!(flags & VGM_ASYNC)) VGM_HIP(hipStreamSynchronize(c->stream));
      return VGM_OK;
    }
    VGM_HIP(launch_nav(c->stream, kind, a));
    if (!(flags & VGM_ASYNC)) VGM_HIP(hipStreamSynchronize(c->stream));
    return VGM_OK;
  }
  if (flags & (VGM_GRAPH | VGM_STEP_KERNELS))
    return fail(VGM_EINVAL, "VGM_GRAPH and VGM_STEP_KERNELS take device arrays (VGM_DEVICE_PTRS)");
  size_t total = 0;
  for (const Arr &r : arrs)
    if (r.user) total += align256(r.bytes);
  if ((rc = ensure_scratch(c, total))) return rc;
  char *b = c->scratch.buf;
  for (const Arr &r : arrs) {
    if (!r.user) {
      *r.slot = nullptr;
      continue;
    }
    *r.slot = b;
    if (r.in) VGM_HIP(hipMemcpyAsync(b, r.user, r.bytes, hipMemcpyHostToDevice, c->stream));
    b += align256(r.bytes);
  }
  VGM_HIP(launch_nav(c->stream, kind, a));
  for (const Arr &r : arrs)
    if (r.user && r.out)
      VGM_HIP(hipMemcpyAsync(const_cast<void *>(r.user), *r.slot, r.bytes, hipMemcpyDeviceToHost, c->stream));
  VGM_HIP(hipStreamSynchronize(c->stream));
  return VGM_OK;
}

// address of a NavArgs pointer field, whatever its pointee type
template <typename T>
void **slot(T *&p)
{
  return (void **)&p;
}

}  // namespace

namespace {
int nav_step(vgm_ctx *c, const double *x, const double *y, const double *z, const double *dx, const double *dy,
             const double *dz, const double *step_limit, size_t n, uint32_t *state, uint32_t *last,
             uint8_t *boundary, double *step, unsigned flags, bool relocate)
{
  if (n && (!state || !last || !boundary || !step)) return fail(VGM_EINVAL, "null state or output array");
  NavArgs a{};
  a.n = n;
  a.no_relocate = relocate ? 0 : 1;
  const size_t b8 = n * sizeof(double), b4 = n * sizeof(uint32_t);
  return run_nav(c, kNavStep, a,
                 {{x, b8, true, false, slot(a.x)}, {y, b8, true, false, slot(a.y)}, {z, b8, true, false, slot(a.z)},
                  {dx, b8, true, false, slot(a.dx)}, {dy, b8, true, false, slot(a.dy)},
                  {dz, b8, true, false, slot(a.dz)}, {step_limit, b8, true, false, slot(a.step_limit), true},
                  {state, b4, true, true, slot(a.state)}, {last, b4, true, true, slot(a.last)},
                  {boundary, n, true, true, slot(a.boundary)}, {step, b8, false, true, slot(a.step)}},
                 flags);
}

}  // namespace

extern "C" {

int vgm_nav_close(vgm_ctx *c, int world_pv, int *nnodes)
{
  if (!c) return fail(VGM_EINVAL, "null context");
  std::string err;
  const int n = close_navigation(c->host, world_pv, err);
  if (n < 0) return fail(VGM_EINVAL, err);
  if (nnodes) *nnodes = n;
  return VGM_OK;
}

int vgm_nav_node(vgm_ctx *c, uint32_t node, int *pv, int *level, int *mother)
{
  if (!c || !pv || !level || !mother) return fail(VGM_EINVAL, "null argument");
  if (node == 0 || node >= c->host.nav_nodes.size()) return fail(VGM_EINVAL, "invalid state index");
  const NavNode &N = c->host.nav_nodes[node];
  *pv = N.pv;
  *level = N.level;
  *mother = N.mother;
  return VGM_OK;
}

int vgm_nav_path_to_state(vgm_ctx *c, const int *pvs, int depth, uint32_t *state)
{
  if (!c || !pvs || !state) return fail(VGM_EINVAL, "null argument");
  const HostGeometry &g = c->host;
  if (g.nav_nodes.size() < 2) return fail(VGM_ESTATE, "navigation tables not built: call vgm_nav_close");
  if (depth < 1) return fail(VGM_EINVAL, "a path holds at least the world");
  if (pvs[0] != g.world_pv) return fail(VGM_EINVAL, "path does not start at the world volume");
  uint32_t node = 1;
  for (int lev = 1; lev < depth; ++lev) {
    const NavNode &N = g.nav_nodes[node];
    const std::vector<int32_t> &dl = g.lv_daughters[size_t(g.pvs[size_t(N.pv)].lv)];
    int k = 0;
    while (k < int(dl.size()) && dl[size_t(k)] != pvs[lev]) ++k;  // GetChildId
    if (k == int(dl.size()))
      return fail(VGM_EINVAL, "placed volume " + std::to_string(pvs[lev]) + " at level " + std::to_string(lev) +
                                  " is not a daughter of placed volume " + std::to_string(N.pv));
    node = uint32_t(g.nav_children[size_t(N.child0 + k)]);
  }
  *state = node;
  return VGM_OK;
}

int vgm_nav_state_to_path(vgm_ctx *c, uint32_t state, int *pvs, int cap, int *depth)
{
  if (!c || !depth || (cap > 0 && !pvs)) return fail(VGM_EINVAL, "null argument");
  const HostGeometry &g = c->host;
  if (state == 0 || state >= g.nav_nodes.size()) return fail(VGM_EINVAL, "invalid state index");
  const int d = g.nav_nodes[state].level + 1;
  if (d > cap) return fail(VGM_EINVAL, "path needs " + std::to_string(d) + " entries");
  for (uint32_t s = state; s != 0; s = uint32_t(g.nav_nodes[s].mother)) pvs[g.nav_nodes[s].level] = g.nav_nodes[s].pv;
  *depth = d;
  return VGM_OK;
}

int vgm_nav_locate(vgm_ctx *c, const double *x, const double *y, const double *z, size_t n, uint32_t *state,
                   unsigned flags)
{
  NavArgs a{};
  a.n = n;
  const size_t b8 = n * sizeof(double), b4 = n * sizeof(uint32_t);
  return run_nav(c, kNavLocate, a,
                 {{x, b8, true, false, slot(a.x)}, {y, b8, true, false, slot(a.y)}, {z, b8, true, false, slot(a.z)},
                  {state, b4, false, true, slot(a.state)}},
                 flags);
}

int vgm_nav_step(vgm_ctx *c, const double *x, const double *y, const double *z, const double *dx, const double *dy,
                 const double *dz, const double *step_limit, size_t n, uint32_t *state, uint32_t *last,
                 uint8_t *boundary, double *step, unsigned flags)
{
  return nav_step(c, x, y, z, dx, dy, dz, step_limit, n, state, last, boundary, step, flags, true);
}

int vgm_nav_compute_step_and_next_volume(vgm_ctx *c, const double *x, const double *y, const double *z,
                                         const double *dx, const double *dy, const double *dz,
                                         const double *step_limit, size_t n, uint32_t *state, uint32_t *last,
                                         uint8_t *boundary, double *step, unsigned flags)
{
  return nav_step(c, x, y, z, dx, dy, dz, step_limit, n, state, last, boundary, step, flags, true);
}

int vgm_nav_compute_step(vgm_ctx *c, const double *x, const double *y, const double *z, const double *dx,
                         const double *dy, const double *dz, const double *step_limit, size_t n, uint32_t *state,
                         uint32_t *last, uint8_t *boundary, double *step, unsigned flags)
{
  return nav_step(c, x, y, z, dx, dy, dz, step_limit, n, state, last, boundary, step, flags, false);
}

int vgm_nav_step_safety(vgm_ctx *c, const double *x, const double *y, const double *z, const double *dx,
                        const double *dy, const double *dz, const double *step_limit, size_t n, uint32_t *state,
                        uint32_t *last, uint8_t *boundary, double *step, const uint8_t *calcsafety, double *safety,
                        unsigned flags)
{
  if (n && (!state || !last || !boundary || !step || !safety)) return fail(VGM_EINVAL, "null state or output array");
  NavArgs a{};
  a.n = n;
  const size_t b8 = n * sizeof(double), b4 = n * sizeof(uint32_t);
  return run_nav(c, kNavStep, a,
                 {{x, b8, true, false, slot(a.x)}, {y, b8, true, false, slot(a.y)}, {z, b8, true, false, slot(a.z)},
                  {dx, b8, true, false, slot(a.dx)}, {dy, b8, true, false, slot(a.dy)},
                  {dz, b8, true, false, slot(a.dz)}, {step_limit, b8, true, false, slot(a.step_limit), true},
                  {state, b4, true, true, slot(a.state)}, {last, b4, true, true, slot(a.last)},
                  {boundary, n, true, true, slot(a.boundary)}, {step, b8, false, true, slot(a.step)},
                  {calcsafety, n, true, false, slot(a.calcsafety), true}, {safety, b8, false, true, slot(a.safety)}},
                 flags);
}

int vgm_nav_transport(vgm_ctx *c, double *x, double *y, double *z, const double *dx, const double *dy,
                      const double *dz, size_t n, uint32_t *state, uint32_t *last, uint8_t *boundary, int nsteps,
                      int32_t *steps, unsigned flags)
{
  if (n && (!state || !last || !boundary || !steps)) return fail(VGM_EINVAL, "null state or counter array");
  if (nsteps < 0) return fail(VGM_EINVAL, "negative step count");
  NavArgs a{};
  a.n = n;
  a.nsteps = nsteps;
  const size_t b8 = n * sizeof(double), b4 = n * sizeof(uint32_t);
  return run_nav(c, kNavTransport, a,
                 {{x, b8, true, true, slot(a.x)}, {y, b8, true, true, slot(a.y)}, {z, b8, true, true, slot(a.z)},
                  {dx, b8, true, false, slot(a.dx)}, {dy, b8, true, false, slot(a.dy)},
                  {dz, b8, true, false, slot(a.dz)}, {state, b4, true, true, slot(a.state)},
                  {last, b4, true, true, slot(a.last)}, {boundary, n, true, true, slot(a.boundary)},
                  {steps, b4, true, true, slot(a.steps)}},
                 flags);
}

int vgm_nav_transport_limited(vgm_ctx *c, double *x, double *y, double *z, const double *dx, const double *dy,
                              const double *dz, const double *step_limit, size_t n, uint32_t *state, uint32_t *last,
                              uint8_t *boundary, int nsteps, int reuse_safety, int32_t *steps, int32_t *geo_steps,
                              unsigned flags)
{
  if (n && (!state || !last || !boundary || !steps || !geo_steps)) return fail(VGM_EINVAL, "null state or counter array");
  if (n && !step_limit) return fail(VGM_EINVAL, "null physics step array");
  if (nsteps < 0) return fail(VGM_EINVAL, "negative step count");
  NavArgs a{};
  a.n = n;
  a.nsteps = nsteps;
  a.reuse_safety = reuse_safety ? 1 : 0;
  const size_t b8 = n * sizeof(double), b4 = n * sizeof(uint32_t);
  return run_nav(c, kNavTransportLimited, a,
                 {{x, b8, true, true, slot(a.x)}, {y, b8, true, true, slot(a.y)}, {z, b8, true, true, slot(a.z)},
                  {dx, b8, true, false, slot(a.dx)}, {dy, b8, true, false, slot(a.dy)},
                  {dz, b8, true, false, slot(a.dz)}, {step_limit, b8, true, false, slot(a.step_limit)},
                  {state, b4, true, true, slot(a.state)}, {last, b4, true, true, slot(a.last)},
                  {boundary, n, true, true, slot(a.boundary)}, {steps, b4, true, true, slot(a.steps)},
                  {geo_steps, b4, true, true, slot(a.geo_steps)}},
                 flags);
}

int vgm_nav_approach(vgm_ctx *c, const double *x, const double *y, const double *z, const double *dx,
                     const double *dy, const double *dz, size_t n, const uint32_t *state, const uint32_t *last,
                     double *step, unsigned flags)
{
  if (n && (!state || !last || !step)) return fail(VGM_EINVAL, "null state or step array");
  NavArgs a{};
  a.n = n;
  const size_t b8 = n * sizeof(double), b4 = n * sizeof(uint32_t);
  return run_nav(c, kNavApproach, a,
                 {{x, b8, true, false, slot(a.x)}, {y, b8, true, false, slot(a.y)}, {z, b8, true, false, slot(a.z)},
                  {dx, b8, true, false, slot(a.dx)}, {dy, b8, true, false, slot(a.dy)},
                  {dz, b8, true, false, slot(a.dz)}, {state, b4, true, false, slot(a.state)},
                  {last, b4, true, false, slot(a.last)}, {step, b8, true, true, slot(a.step)}},
                 flags);
}

int vgm_nav_level_locate(vgm_ctx *c, const double *x, const double *y, const double *z, const double *dx,
                         const double *dy, const double *dz, size_t n, const uint32_t *state, const uint32_t *excl,
                         uint32_t *child, double *qx, double *qy, double *qz, unsigned flags)
{
  if (n && (!state || !child || !qx || !qy || !qz)) return fail(VGM_EINVAL, "null state or output array");
  if ((dx || dy || dz) && !(dx && dy && dz)) return fail(VGM_EINVAL, "direction needs all three components");
  NavArgs a{};
  a.n = n;
  const size_t b8 = n * sizeof(double), b4 = n * sizeof(uint32_t);
  return run_nav(c, kNavLevelLocate, a,
                 {{x, b8, true, false, slot(a.x)}, {y, b8, true, false, slot(a.y)}, {z, b8, true, false, slot(a.z)},
                  {dx, b8, true, false, slot(a.dx), true}, {dy, b8, true, false, slot(a.dy), true},
                  {dz, b8, true, false, slot(a.dz), true}, {state, b4, true, false, slot(a.state)},
                  {excl, b4, true, false, slot(a.last), true}, {child, b4, false, true, slot(a.child)},
                  {qx, b8, false, true, slot(a.ox)}, {qy, b8, false, true, slot(a.oy)},
                  {qz, b8, false, true, slot(a.oz)}},
                 flags);
}

int vgm_nav_safety(vgm_ctx *c, const double *x, const double *y, const double *z, size_t n, const uint32_t *state,
                   double *out, unsigned flags)
{
  if (n && !out) return fail(VGM_EINVAL, "null output array");
  NavArgs a{};
  a.n = n;
  const size_t b8 = n * sizeof(double), b4 = n * sizeof(uint32_t);
  return run_nav(c, kNavSafety, a,
                 {{x, b8, true, false, slot(a.x)}, {y, b8, true, false, slot(a.y)}, {z, b8, true, false, slot(a.z)},
                  {state, b4, true, false, slot(a.state)}, {out, b8, false, true, slot(a.step)}},
                 flags);
}

}  // extern "C"
