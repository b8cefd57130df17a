// C-ABI layer: contexts, geometry registration/upload and SOA query dispatch.
#include <hip/hip_runtime.h>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include "../../include/vgm/vgm.h"
#include "context.h"
#include "mlp_train.h"

namespace vgm {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

}  // namespace vgm

using namespace vgm;

// Staging array of vgm_query_noalloc: each device holds its own copy from the
// library's load until the process ends.  Host code never uses this name as
// an address (on the host it is only a stand-in); it uses the per-device
// address that hipGetSymbolAddress returns, kept in g_noalloc_ptr.
static __device__ __align__(256) char scratch[VGM_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
char *g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

namespace vgm {

// Looks up the device address of `scratch` on `device` (the current device)
// once; called by vgm_ctx_create.
int init_noalloc_scratch(int device)
{
  if (device < 0 || device >= kMaxDevices) return fail(VGM_EINVAL, "device index above the library's limit");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return VGM_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  if (!p) return fail(VGM_EHIP, "hipGetSymbolAddress(scratch) returned a null address");
  g_noalloc_ptr[device] = static_cast<char *>(p);
  return VGM_OK;
}

int fail(int code, const std::string &msg)
{
  set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char *what)
{
  return fail(VGM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

int ensure_scratch(vgm_ctx *c, size_t bytes)
{
  if (c->scratch.cap >= bytes) return VGM_OK;
  if (c->scratch.buf) {
    VGM_HIP(hipStreamSynchronize(c->stream));
    VGM_HIP(hipFree(c->scratch.buf));
    c->scratch.buf = nullptr;
    c->scratch.cap = 0;
  }
  size_t cap = align256(bytes + bytes / 4);
  VGM_HIP(hipMalloc(&c->scratch.buf, cap));
  c->scratch.cap = cap;
  return VGM_OK;
}

// Pack the host tables into one staging buffer and upload when stale.
int sync_geometry(vgm_ctx *c)
{
  HostGeometry &h = c->host;
  DeviceGeometry &d = c->dev;
  if (d.version == h.version && d.buf) return VGM_OK;
  std::vector<int32_t> dlist;
  std::vector<LogicalRec> lvs = h.lvs;
  for (size_t i = 0; i < lvs.size(); ++i) {
    lvs[i].dbegin = int32_t(dlist.size());
    lvs[i].dcount = int32_t(h.lv_daughters[i].size());
    dlist.insert(dlist.end(), h.lv_daughters[i].begin(), h.lv_daughters[i].end());
  }
  const size_t s_shapes = align256(h.shapes.size() * sizeof(ShapeRec) + 1);
  const size_t s_pool = align256(h.pool.size() * sizeof(double) + 1);
  const size_t s_lvs = align256(lvs.size() * sizeof(LogicalRec) + 1);
  const size_t s_dl = align256(dlist.size() * sizeof(int32_t) + 1);
  const size_t s_pvs = align256(h.pvs.size() * sizeof(PlacedRec) + 1);
  const size_t s_nn = align256(h.nav_nodes.size() * sizeof(NavNode) + 1);
  const size_t s_nc = align256(h.nav_children.size() * sizeof(int32_t) + 1);
  const size_t s_bv = align256(h.bvhs.size() * sizeof(BvhRec) + 1);
  const size_t s_ip = align256(h.nav_ipool.size() * sizeof(int32_t) + 1);
  const size_t s_bp = align256(h.nav_bpool.size() * sizeof(double) + 1);
  const size_t total = s_shapes + s_pool + s_lvs + s_dl + s_pvs + s_nn + s_nc + s_bv + s_ip + s_bp;
  std::vector<char> stage(total, 0);
  size_t o = 0;
  auto put = [&](const void *src, size_t n, size_t span) {
    if (n) std::memcpy(stage.data() + o, src, n);
    size_t at = o;
    o += span;
    return at;
  };
  const size_t o_shapes = put(h.shapes.data(), h.shapes.size() * sizeof(ShapeRec), s_shapes);
  const size_t o_pool = put(h.pool.data(), h.pool.size() * sizeof(double), s_pool);
  const size_t o_lvs = put(lvs.data(), lvs.size() * sizeof(LogicalRec), s_lvs);
  const size_t o_dl = put(dlist.data(), dlist.size() * sizeof(int32_t), s_dl);
  const size_t o_pvs = put(h.pvs.data(), h.pvs.size() * sizeof(PlacedRec), s_pvs);
  const size_t o_nn = put(h.nav_nodes.data(), h.nav_nodes.size() * sizeof(NavNode), s_nn);
  const size_t o_nc = put(h.nav_children.data(), h.nav_children.size() * sizeof(int32_t), s_nc);
  const size_t o_bv = put(h.bvhs.data(), h.bvhs.size() * sizeof(BvhRec), s_bv);
  const size_t o_ip = put(h.nav_ipool.data(), h.nav_ipool.size() * sizeof(int32_t), s_ip);
  const size_t o_bp = put(h.nav_bpool.data(), h.nav_bpool.size() * sizeof(double), s_bp);
  if (d.cap < total) {
    if (d.buf) {
      VGM_HIP(hipStreamSynchronize(c->stream));
      VGM_HIP(hipFree(d.buf));
      d.buf = nullptr;
      d.cap = 0;
    }
    VGM_HIP(hipMalloc(&d.buf, total));
    d.cap = total;
  }
  VGM_HIP(hipMemcpyAsync(d.buf, stage.data(), total, hipMemcpyHostToDevice, c->stream));
  VGM_HIP(hipStreamSynchronize(c->stream));
  d.shapes = reinterpret_cast<ShapeRec *>(d.buf + o_shapes);
  d.pool = reinterpret_cast<double *>(d.buf + o_pool);
  d.lvs = reinterpret_cast<LogicalRec *>(d.buf + o_lvs);
  d.daughters = reinterpret_cast<int32_t *>(d.buf + o_dl);
  d.pvs = reinterpret_cast<PlacedRec *>(d.buf + o_pvs);
  d.nav_nodes = reinterpret_cast<NavNode *>(d.buf + o_nn);
  d.nav_children = reinterpret_cast<int32_t *>(d.buf + o_nc);
  d.bvhs = reinterpret_cast<BvhRec *>(d.buf + o_bv);
  d.nav_ipool = reinterpret_cast<int32_t *>(d.buf + o_ip);
  d.nav_bpool = reinterpret_cast<double *>(d.buf + o_bp);
  d.version = h.version;
  return VGM_OK;
}

}  // namespace vgm

namespace {

int check_pv(vgm_ctx *c, int pv)
{
  if (!c) return fail(VGM_EINVAL, "null context");
  if (pv < 0 || size_t(pv) >= c->host.pvs.size()) return fail(VGM_EINVAL, "invalid placed-volume id");
  return VGM_OK;
}

// Bytes of device staging for nin input arrays and one output of n points.
size_t staged_bytes(size_t nin, size_t n, size_t out_elem)
{
  return nin * align256(n * sizeof(double)) + align256(n * out_elem);
}

// Host-pointer query through the staging buffer `b` (staged_bytes long):
// copy in, launch, copy out, wait.
int run_staged(vgm_ctx *c, char *b, int method, int precision, const ShapeRec &S, const void *params,
               const Transform &tr, bool use_tr, const double *px, const double *py, const double *pz,
               const double *dx, const double *dy, const double *dz, const double *step, size_t n, void *out,
               size_t out_elem)
{
  const bool with_dir = (method == kDistanceToIn || method == kDistanceToOut);
  const size_t nin = with_dir ? (step ? 7 : 6) : 3;
  const size_t bytes_in = align256(n * sizeof(double));
  const double *src[7] = {px, py, pz, dx, dy, dz, step};
  double *dst[7] = {nullptr};
  for (size_t k = 0; k < nin; ++k) {
    dst[k] = reinterpret_cast<double *>(b + k * bytes_in);
    VGM_HIP(hipMemcpyAsync(dst[k], src[k], n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  void *dout = b + nin * bytes_in;
  SoaArgs a{};
  a.n = n;
  a.px = dst[0], a.py = dst[1], a.pz = dst[2];
  if (with_dir) {
    a.dx = dst[3], a.dy = dst[4], a.dz = dst[5];
    a.step = step ? dst[6] : nullptr;
  }
  a.out = dout;
  VGM_HIP(launch_soa(c->stream, method, precision, S, params, tr, use_tr, a));
  VGM_HIP(hipMemcpyAsync(out, dout, n * out_elem, hipMemcpyDeviceToHost, c->stream));
  VGM_HIP(hipStreamSynchronize(c->stream));
  return VGM_OK;
}

// Common driver for the seven SOA queries.
int run_soa(vgm_ctx *c, int pv, int method, bool use_tr, const double *px, const double *py, const double *pz,
            const double *dx, const double *dy, const double *dz, const double *step, size_t n, void *out,
            size_t out_elem, unsigned flags)
{
  int rc = check_pv(c, pv);
  if (rc) return rc;
  const bool with_dir = (method == kDistanceToIn || method == kDistanceToOut);
  if (n && (!px || !py || !pz || !out || (with_dir && (!dx || !dy || !dz))))
    return fail(VGM_EINVAL, "null input or output pointer");
  VGM_HIP(hipSetDevice(c->device));
  rc = sync_geometry(c);
  if (rc) return rc;
  const PlacedRec &P = c->host.pvs[pv];
  const ShapeRec &S = c->host.shapes[c->host.lvs[P.lv].shape];
  const void *params = c->dev.pool + S.off;
  const int precision = (flags & VGM_MIXED) ? kMixed : kFP64;
  SoaArgs a{};
  a.n = n;
  if (n == 0) return VGM_OK;
  if (flags & VGM_DEVICE_PTRS) {
    a.px = px, a.py = py, a.pz = pz;
    a.dx = dx, a.dy = dy, a.dz = dz;
    a.step = step;
    a.out = out;
    VGM_HIP(launch_soa(c->stream, method, precision, S, params, P.tr, use_tr, a));
    if (!(flags & VGM_ASYNC)) VGM_HIP(hipStreamSynchronize(c->stream));
    return VGM_OK;
  }
  const size_t nin = with_dir ? (step ? 7 : 6) : 3;
  rc = ensure_scratch(c, staged_bytes(nin, n, out_elem));
  if (rc) return rc;
  return run_staged(c, c->scratch.buf, method, precision, S, params, P.tr, use_tr, px, py, pz, dx, dy, dz, step, n,
                    out, out_elem);
}

}  // namespace

extern "C" {

const char *vgm_last_error(void) { return g_last_error.c_str(); }

int vgm_ctx_create(int device, void *stream, vgm_ctx **out)
{
  if (!out) return fail(VGM_EINVAL, "null output pointer");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return fail(VGM_EHIP, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(VGM_EINVAL, "device index out of range");
  VGM_HIP(hipSetDevice(device));
  vgm_ctx *c = new vgm_ctx();
  c->device = device;
  if (hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    c->num_cus = 256;
  if (stream) {
    c->stream = static_cast<hipStream_t>(stream);
  } else {
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      return hip_fail(e, "hipStreamCreate");
    }
    c->own_stream = true;
  }
  const int rc = init_noalloc_scratch(device);
  if (rc) {
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return VGM_OK;
}

int vgm_ctx_destroy(vgm_ctx *c)
{
  if (!c) return VGM_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->dev.buf) (void)hipFree(c->dev.buf);
  if (c->scratch.buf) (void)hipFree(c->scratch.buf);
  if (c->nav_queue) (void)hipFree(c->nav_queue);
  if (c->nav_graph.exec) (void)hipGraphExecDestroy(c->nav_graph.exec);
  if (c->nav_graph.graph) (void)hipGraphDestroy(c->nav_graph.graph);
  mlp_release(c);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return VGM_OK;
}

int vgm_ctx_synchronize(vgm_ctx *c)
{
  if (!c) return fail(VGM_EINVAL, "null context");
  VGM_HIP(hipStreamSynchronize(c->stream));
  return VGM_OK;
}

void *vgm_ctx_stream(vgm_ctx *c) { return c ? static_cast<void *>(c->stream) : nullptr; }

int vgm_box_create(vgm_ctx *c, double dx, double dy, double dz, int *shape)
{
  if (!c || !shape) return fail(VGM_EINVAL, "null argument");
  std::string err;
  int id = build_box(c->host, dx, dy, dz, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_boolean_create(vgm_ctx *c, int op, int left, const double *left_tr, int right, const double *right_tr,
                       int *shape)
{
  if (!c || !shape) return fail(VGM_EINVAL, "null argument");
  std::string err;
  int id = build_boolean(c->host, op, left, left_tr, right, right_tr, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_cut_tube_create(vgm_ctx *c, double rmin, double rmax, double z, double sphi, double dphi, double bx,
                        double by, double bz, double tx, double ty, double tz, int *shape)
{
  if (!c || !shape) return fail(VGM_EINVAL, "null argument");
  std::string err;
  const double bot[3] = {bx, by, bz}, top[3] = {tx, ty, tz};
  int id = build_cut_tube(c->host, rmin, rmax, z, sphi, dphi, bot, top, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_orb_create(vgm_ctx *c, double r, int *shape)
{
  if (!c || !shape) return fail(VGM_EINVAL, "null argument");
  std::string err;
  int id = build_orb(c->host, r, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_tube_create(vgm_ctx *c, double rmin, double rmax, double dz, double sphi, double dphi, int *shape)
{
  if (!c || !shape) return fail(VGM_EINVAL, "null argument");
  std::string err;
  int id = build_tube(c->host, rmin, rmax, dz, sphi, dphi, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_cone_create(vgm_ctx *c, double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi,
                    double dphi, int *shape)
{
  if (!c || !shape) return fail(VGM_EINVAL, "null argument");
  std::string err;
  int id = build_cone(c->host, rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_trd_create(vgm_ctx *c, double x1, double x2, double y1, double y2, double z, int *shape)
{
  if (!c || !shape) return fail(VGM_EINVAL, "null argument");
  std::string err;
  int id = build_trd(c->host, x1, x2, y1, y2, z, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_trapezoid_create(vgm_ctx *c, double dz, double theta, double phi, double dy1, double dx1, double dx2,
                         double alpha1, double dy2, double dx3, double dx4, double alpha2, int *shape)
{
  if (!c || !shape) return fail(VGM_EINVAL, "null argument");
  std::string err;
  const double p[11] = {dz, theta, phi, dy1, dx1, dx2, alpha1, dy2, dx3, dx4, alpha2};
  int id = build_trapezoid(c->host, p, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_polycone_create(vgm_ctx *c, double sphi, double dphi, int nz, const double *z, const double *rmin,
                        const double *rmax, int *shape)
{
  if (!c || !shape) return fail(VGM_EINVAL, "null argument");
  std::string err;
  int id = build_polycone(c->host, sphi, dphi, nz, z, rmin, rmax, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_polyhedron_create(vgm_ctx *c, double sphi, double dphi, int nsides, int nz, const double *z,
                          const double *rmin, const double *rmax, int *shape)
{
  if (!c || !shape) return fail(VGM_EINVAL, "null argument");
  std::string err;
  int id = build_polyhedron(c->host, sphi, dphi, nsides, nz, z, rmin, rmax, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_tessellated_create(vgm_ctx *c, int nfacets, const double *triangles, const double *test_dir, int *shape)
{
  if (!c || !shape || !triangles) return fail(VGM_EINVAL, "null argument");
  std::string err;
  int id = build_tessellated(c->host, nfacets, triangles, test_dir, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_shape_extent(vgm_ctx *c, int shape, double *mn, double *mx)
{
  if (!c || !mn || !mx) return fail(VGM_EINVAL, "null argument");
  if (shape < 0 || size_t(shape) >= c->host.shapes.size()) return fail(VGM_EINVAL, "invalid shape id");
  const ShapeRec &r = c->host.shapes[shape];
  for (int k = 0; k < 3; ++k) {
    mn[k] = r.bmin[k];
    mx[k] = r.bmax[k];
  }
  return VGM_OK;
}

int vgm_lv_create(vgm_ctx *c, int shape, int *lv)
{
  if (!c || !lv) return fail(VGM_EINVAL, "null argument");
  if (shape < 0 || size_t(shape) >= c->host.shapes.size()) return fail(VGM_EINVAL, "invalid shape id");
  LogicalRec r{shape, 0, 0, -1};
  c->host.lvs.push_back(r);
  c->host.lv_daughters.emplace_back();
  ++c->host.version;
  *lv = int(c->host.lvs.size()) - 1;
  return VGM_OK;
}

int vgm_pv_create(vgm_ctx *c, int lv, const double *tr, int *pv)
{
  if (!c || !pv) return fail(VGM_EINVAL, "null argument");
  if (lv < 0 || size_t(lv) >= c->host.lvs.size()) return fail(VGM_EINVAL, "invalid logical-volume id");
  PlacedRec p{};
  p.lv = lv;
  p.id = int32_t(c->host.pvs.size());
  if (tr) {
    for (int k = 0; k < 3; ++k) p.tr.t[k] = tr[k];
    for (int k = 0; k < 9; ++k) p.tr.r[k] = tr[3 + k];
  } else {
    for (int k = 0; k < 3; ++k) p.tr.t[k] = 0;
    for (int k = 0; k < 9; ++k) p.tr.r[k] = (k % 4 == 0) ? 1.0 : 0.0;
  }
  c->host.pvs.push_back(p);
  ++c->host.version;
  *pv = p.id;
  return VGM_OK;
}

int vgm_lv_add_daughter(vgm_ctx *c, int mother, int pv)
{
  if (!c) return fail(VGM_EINVAL, "null context");
  if (mother < 0 || size_t(mother) >= c->host.lvs.size()) return fail(VGM_EINVAL, "invalid logical-volume id");
  if (pv < 0 || size_t(pv) >= c->host.pvs.size()) return fail(VGM_EINVAL, "invalid placed-volume id");
  c->host.lv_daughters[mother].push_back(pv);
  ++c->host.version;
  return VGM_OK;
}

int vgm_contains(vgm_ctx *c, int pv, const double *x, const double *y, const double *z, size_t n, uint8_t *out,
                 unsigned flags)
{
  return run_soa(c, pv, kContains, true, x, y, z, nullptr, nullptr, nullptr, nullptr, n, out, 1, flags);
}

int vgm_inside(vgm_ctx *c, int pv, const double *x, const double *y, const double *z, size_t n, int32_t *out,
               unsigned flags)
{
  return run_soa(c, pv, kInside, true, x, y, z, nullptr, nullptr, nullptr, nullptr, n, out, 4, flags);
}

int vgm_distance_to_in(vgm_ctx *c, int pv, const double *px, const double *py, const double *pz, const double *dx,
                       const double *dy, const double *dz, const double *step, size_t n, double *out,
                       unsigned flags)
{
  return run_soa(c, pv, kDistanceToIn, true, px, py, pz, dx, dy, dz, step, n, out, 8, flags);
}

int vgm_distance_to_out(vgm_ctx *c, int pv, const double *px, const double *py, const double *pz, const double *dx,
                        const double *dy, const double *dz, const double *step, size_t n, double *out,
                        unsigned flags)
{
  return run_soa(c, pv, kDistanceToOut, false, px, py, pz, dx, dy, dz, step, n, out, 8, flags);
}

int vgm_placed_distance_to_out(vgm_ctx *c, int pv, const double *px, const double *py, const double *pz,
                               const double *dx, const double *dy, const double *dz, const double *step, size_t n,
                               double *out, unsigned flags)
{
  return run_soa(c, pv, kDistanceToOut, true, px, py, pz, dx, dy, dz, step, n, out, 8, flags);
}

int vgm_safety_to_in(vgm_ctx *c, int pv, const double *x, const double *y, const double *z, size_t n, double *out,
                     unsigned flags)
{
  return run_soa(c, pv, kSafetyToIn, true, x, y, z, nullptr, nullptr, nullptr, nullptr, n, out, 8, flags);
}

int vgm_safety_to_out(vgm_ctx *c, int pv, const double *x, const double *y, const double *z, size_t n, double *out,
                      unsigned flags)
{
  return run_soa(c, pv, kSafetyToOut, false, x, y, z, nullptr, nullptr, nullptr, nullptr, n, out, 8, flags);
}

int vgm_sphere_create(vgm_ctx *c, double rmin, double rmax, double sphi, double dphi, double stheta, double dtheta,
                      int *shape)
{
  if (!c || !shape) return fail(VGM_EINVAL, "null argument");
  std::string err;
  int id = build_sphere(c->host, rmin, rmax, sphi, dphi, stheta, dtheta, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_torus_create(vgm_ctx *c, double rmin, double rmax, double rtor, double sphi, double dphi, int *shape)
{
  if (!c || !shape) return fail(VGM_EINVAL, "null argument");
  std::string err;
  int id = build_torus(c->host, rmin, rmax, rtor, sphi, dphi, err);
  if (id < 0) return fail(VGM_EINVAL, err);
  *shape = id;
  return VGM_OK;
}

int vgm_normal(vgm_ctx *c, int pv, const double *x, const double *y, const double *z, size_t n, double *nx,
               double *ny, double *nz, uint8_t *valid, unsigned flags)
{
  int rc = check_pv(c, pv);
  if (rc) return rc;
  if (n && (!x || !y || !z || !nx || !ny || !nz || !valid)) return fail(VGM_EINVAL, "null input or output pointer");
  const PlacedRec &P = c->host.pvs[pv];
  const ShapeRec &S = c->host.shapes[c->host.lvs[P.lv].shape];
  if (!family_has_normal(S.type)) return fail(VGM_EINVAL, "this shape family has no Normal");
  if (flags & VGM_MIXED) return fail(VGM_EINVAL, "Normal is FP64 only");
  if (n == 0) return VGM_OK;
  VGM_HIP(hipSetDevice(c->device));
  rc = sync_geometry(c);
  if (rc) return rc;
  const void *params = c->dev.pool + S.off;
  if (flags & VGM_DEVICE_PTRS) {
    NormalArgs a{x, y, z, n, nx, ny, nz, valid};
    VGM_HIP(launch_normal(c->stream, S, params, P.tr, a));
    if (!(flags & VGM_ASYNC)) VGM_HIP(hipStreamSynchronize(c->stream));
    return VGM_OK;
  }
  const size_t col = align256(n * sizeof(double));
  rc = ensure_scratch(c, 6 * col + align256(n));
  if (rc) return rc;
  char *b = c->scratch.buf;
  double *d[6];
  for (int k = 0; k < 6; ++k) d[k] = reinterpret_cast<double *>(b + k * col);
  uint8_t *dv = reinterpret_cast<uint8_t *>(b + 6 * col);
  const double *src[3] = {x, y, z};
  for (int k = 0; k < 3; ++k)
    VGM_HIP(hipMemcpyAsync(d[k], src[k], n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  NormalArgs a{d[0], d[1], d[2], n, d[3], d[4], d[5], dv};
  VGM_HIP(launch_normal(c->stream, S, params, P.tr, a));
  double *dst[3] = {nx, ny, nz};
  for (int k = 0; k < 3; ++k)
    VGM_HIP(hipMemcpyAsync(dst[k], d[3 + k], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  VGM_HIP(hipMemcpyAsync(valid, dv, n, hipMemcpyDeviceToHost, c->stream));
  VGM_HIP(hipStreamSynchronize(c->stream));
  return VGM_OK;
}

int vgm_query_noalloc(vgm_ctx *c, int pv, int method, const double *px, const double *py, const double *pz,
                      const double *dx, const double *dy, const double *dz, const double *step, size_t n, void *out,
                      unsigned flags)
{
  // VGM_Q_* -> (kernel method, master-frame input, output element size)
  static const int kMethod[7] = {kContains, kInside, kDistanceToIn, kDistanceToOut, kDistanceToOut, kSafetyToIn,
                                 kSafetyToOut};
  static const bool kMaster[7] = {true, true, true, false, true, true, false};
  static const size_t kElem[7] = {1, 4, 8, 8, 8, 8, 8};
  int rc = check_pv(c, pv);
  if (rc) return rc;
  if (method < VGM_Q_CONTAINS || method > VGM_Q_SAFETY_TO_OUT) return fail(VGM_EINVAL, "unknown query method");
  if (flags & VGM_DEVICE_PTRS) return fail(VGM_EINVAL, "vgm_query_noalloc takes host arrays only");
  const int m = kMethod[method];
  const bool with_dir = (m == kDistanceToIn || m == kDistanceToOut);
  if (n && (!px || !py || !pz || !out || (with_dir && (!dx || !dy || !dz))))
    return fail(VGM_EINVAL, "null input or output pointer");
  const size_t nin = with_dir ? (step ? 7 : 6) : 3;
  const size_t need = staged_bytes(nin, n, kElem[method]);
  if (need > size_t(VGM_NOALLOC_SCRATCH_BYTES))
    return fail(VGM_EINVAL, "vgm_query_noalloc: " + std::to_string(n) + " points need " + std::to_string(need) +
                                " bytes of staging, above VGM_NOALLOC_SCRATCH_BYTES (" +
                                std::to_string(size_t(VGM_NOALLOC_SCRATCH_BYTES)) + ")");
  if (c->dev.version != c->host.version || !c->dev.buf)
    return fail(VGM_ESTATE, "vgm_query_noalloc: the geometry is not on the device; run an allocating query first");
  char *buf = (c->device >= 0 && c->device < kMaxDevices) ? g_noalloc_ptr[c->device] : nullptr;
  if (!buf) return fail(VGM_ESTATE, "vgm_query_noalloc: no staging array on this device");
  if (n == 0) return VGM_OK;
  VGM_HIP(hipSetDevice(c->device));
  const PlacedRec &P = c->host.pvs[pv];
  const ShapeRec &S = c->host.shapes[c->host.lvs[P.lv].shape];
  return run_staged(c, buf, m, (flags & VGM_MIXED) ? kMixed : kFP64, S, c->dev.pool + S.off, P.tr, kMaster[method],
                    px, py, pz, dx, dy, dz, step, n, out, kElem[method]);
}

}  // extern "C"
