// lpc.hpp -- learned CG preconditioner (3-layer conv net, BF16 on MFMA).
#pragma once
#include "context.hpp"

namespace rsg {
int lpc_param_count();
int lpc_set_params(rsg_ctx *ctx, const float *params);
int lpc_get_params(rsg_ctx *ctx, float *params);
int lpc_set_potential(rsg_ctx *ctx, const double *vloc_host);
// P g for nb bands g[b*ld + p] (device), precision 64 or 32; pg may equal
// neither g nor overlap it
int lpc_apply(rsg_ctx *ctx, const void *g, void *pg, int64_t ld, int nb, int precision);
int lpc_train_step(rsg_ctx *ctx, const float *g, const float *t, int nsamp, float kappa, float lr, double *loss,
                   float *grads);
int lpc_reset_optimizer(rsg_ctx *ctx);
}  // namespace rsg
