// Host interfaces of the learned-safety module (mlp.hip) used by the C-ABI
// layer and by the batched transport.
#pragma once
#include "context.h"
#include "mlp.h"

namespace vgm {

int mlp_train(vgm_ctx *c, int shape, const vgm_mlp_config &cfg, vgm_mlp_report &rep);
// One forward and backward pass of a batch with the given public-order
// parameters; n is a multiple of 64, u holds n x 3 normalised coordinates.
int mlp_gradient(vgm_ctx *c, const float *params_pub, const float *u, const float *target, size_t n,
                 float over_weight, float *grad_pub, float *loss);
int mlp_set(vgm_ctx *c, int shape, const float *params_pub, const double *geom8);
int mlp_get(vgm_ctx *c, int shape, float *params_pub, double *geom8);
// Prediction (or served bound) for device points in the shape's frame.
int mlp_infer(vgm_ctx *c, const MlpState &st, const double *x, const double *y, const double *z, size_t n,
              double *out, bool bound);
int mlp_pack(vgm_ctx *c, MlpState &st);
MlpDesc mlp_desc(const MlpState &st);
void mlp_release(vgm_ctx *c);
void mlp_public_to_internal(const float *pub, float *internal);
void mlp_internal_to_public(const float *internal, float *pub);

}  // namespace vgm
