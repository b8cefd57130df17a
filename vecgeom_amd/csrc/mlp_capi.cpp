// Learned-safety entry points of the C ABI (include/vgm/vgm.h).
#include <cstring>
#include "context.h"
#include "mlp_train.h"

using namespace vgm;

namespace {

int query(vgm_ctx *c, int shape, const double *x, const double *y, const double *z, size_t n, double *out,
          unsigned flags, bool bound)
{
  if (!c) return fail(VGM_EINVAL, "null context");
  if (shape < 0 || size_t(shape) >= c->mlps.size() || !c->mlps[size_t(shape)].ready)
    return fail(VGM_ESTATE, "no trained network for this shape");
  if (n && (!x || !y || !z || !out)) return fail(VGM_EINVAL, "null array argument");
  VGM_HIP(hipSetDevice(c->device));
  const MlpState &st = c->mlps[size_t(shape)];
  if (n == 0) return VGM_OK;
  int rc;
  if (flags & VGM_DEVICE_PTRS) {
    if ((rc = mlp_infer(c, st, x, y, z, n, out, bound))) return rc;
    if (!(flags & VGM_ASYNC)) VGM_HIP(hipStreamSynchronize(c->stream));
    return VGM_OK;
  }
  const size_t b = align256(n * sizeof(double));
  if ((rc = ensure_scratch(c, 4 * b))) return rc;
  double *d[4];
  for (int k = 0; k < 4; ++k) d[k] = reinterpret_cast<double *>(c->scratch.buf + size_t(k) * b);
  const double *src[3] = {x, y, z};
  for (int k = 0; k < 3; ++k)
    VGM_HIP(hipMemcpyAsync(d[k], src[k], n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if ((rc = mlp_infer(c, st, d[0], d[1], d[2], n, d[3], bound))) return rc;
  VGM_HIP(hipMemcpyAsync(out, d[3], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  VGM_HIP(hipStreamSynchronize(c->stream));
  return VGM_OK;
}

}  // namespace

extern "C" {

void vgm_mlp_config_default(vgm_mlp_config *cfg)
{
  if (!cfg) return;
  std::memset(cfg, 0, sizeof *cfg);
  cfg->samples = 6250000;
  cfg->validation = 1000000;
  cfg->batch = 8192;
  cfg->epochs = 4;
  cfg->lr = 2e-3f;
  cfg->over_weight = 4.0f;
  cfg->margin_factor = 1.25;
  cfg->margin_abs = 1e-4;
  cfg->domain = 1.3;
  cfg->seed = 42;
}

int vgm_mlp_train(vgm_ctx *c, int shape, const vgm_mlp_config *cfg, vgm_mlp_report *report)
{
  if (!c) return fail(VGM_EINVAL, "null context");
  if (shape < 0 || size_t(shape) >= c->host.shapes.size()) return fail(VGM_EINVAL, "invalid shape id");
  vgm_mlp_config d;
  vgm_mlp_config_default(&d);
  vgm_mlp_report r{};
  VGM_HIP(hipSetDevice(c->device));
  const int rc = mlp_train(c, shape, cfg ? *cfg : d, r);
  if (rc == VGM_OK && report) *report = r;
  return rc;
}

int vgm_mlp_predict(vgm_ctx *c, int shape, const double *x, const double *y, const double *z, size_t n,
                    double *out, unsigned flags)
{
  return query(c, shape, x, y, z, n, out, flags, false);
}

int vgm_mlp_bound(vgm_ctx *c, int shape, const double *x, const double *y, const double *z, size_t n,
                  double *out, unsigned flags)
{
  return query(c, shape, x, y, z, n, out, flags, true);
}

int vgm_mlp_get(vgm_ctx *c, int shape, float *params, double *geom8)
{
  if (!c || !params || !geom8) return fail(VGM_EINVAL, "null argument");
  VGM_HIP(hipSetDevice(c->device));
  return mlp_get(c, shape, params, geom8);
}

int vgm_mlp_set(vgm_ctx *c, int shape, const float *params, const double *geom8)
{
  if (!c || !params || !geom8) return fail(VGM_EINVAL, "null argument");
  for (int k = 3; k < 7; ++k)
    if (!(geom8[k] > 0)) return fail(VGM_EINVAL, "domain half-widths and scale must be positive");
  if (!(geom8[7] >= 0)) return fail(VGM_EINVAL, "margin must be non-negative");
  VGM_HIP(hipSetDevice(c->device));
  int rc = sync_geometry(c);
  if (rc) return rc;
  return mlp_set(c, shape, params, geom8);
}

int vgm_mlp_gradient(vgm_ctx *c, const float *params, const float *u, const float *target, size_t n,
                     float over_weight, float *grad, float *loss)
{
  if (!c || !params || !u || !target || !grad || !loss) return fail(VGM_EINVAL, "null argument");
  VGM_HIP(hipSetDevice(c->device));
  return mlp_gradient(c, params, u, target, n, over_weight, grad, loss);
}

}  // extern "C"
