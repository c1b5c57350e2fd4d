// solver.hpp -- Gram-Schmidt, Rayleigh-Ritz, CG and esp on the band store.
#pragma once
#include "context.hpp"

namespace rsg {
int gram_schmidt(rsg_ctx *ctx, int n0, int n1);
int subspace_diag(rsg_ctx *ctx, double *esp);
int conjugate_gradient(rsg_ctx *ctx, int ncg, int pc_kind, int mloop, double *esp, double *res);
int esp_calc(rsg_ctx *ctx, double *esp);
void solver_release(rsg_ctx *ctx);

// Per-band dot products alpha * sum_i x_q(i,b) y_q(i,b) for up to 6 pairs
// of column-major N x nb blocks (leading dimension N); out[q*nb + b] on the
// device.  Deterministic two-pass reduction.
int multidot(rsg_ctx *ctx, int npair, const double *const *x, const double *const *y, int nb,
             double alpha, double *out);
}  // namespace rsg
