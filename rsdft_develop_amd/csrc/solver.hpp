// solver.hpp -- Gram-Schmidt, Rayleigh-Ritz, CG and esp on the band store.
#pragma once
#include "context.hpp"

namespace rsg {
int gram_schmidt(rsg_ctx *ctx, int n0, int n1);
int subspace_diag(rsg_ctx *ctx, double *esp);
int conjugate_gradient(rsg_ctx *ctx, int ncg, int pc_kind, int mloop, double *esp, double *res);
int conjugate_gradient_f32(rsg_ctx *ctx, int ncg, int pc_kind, int mloop, double *esp, double *res);
int precond_hprsdft_dev(rsg_ctx *ctx, int mloop, const double *g, double *pg, int64_t ld, int nb);
int autotune_gram(rsg_ctx *ctx, int nb, int reps, float *best_ms, int64_t *best_kc);
int precond_hprsdft_dev(rsg_ctx *ctx, int mloop, const float *g, float *pg, int64_t ld, int nb);
int axpby_f32(rsg_ctx *ctx, float *y, const float *x, float a, float b, int64_t n);
int lpc_target(rsg_ctx *ctx, const float *g, float *t, int nb, int steps, double shift, double vweight);
int band_residuals_f32(rsg_ctx *ctx, float *g);
int esp_calc(rsg_ctx *ctx, double *esp);
int random_bands(rsg_ctx *ctx, int n0, int nb, uint64_t seed);
// 32: the bands move to the FP32 store (subspace_diag, CG (FP32 only),
// esp_calc and the density then work on it; Gram-Schmidt goes through the
// FP64 store); 64: back to the FP64 store
int set_band_precision(rsg_ctx *ctx, int bits);
// RSG_ERR_STATE while the bands are in the FP32 store
int check_store64(rsg_ctx *ctx, const char *who);
void solver_release(rsg_ctx *ctx);

// Per-band dot products alpha * sum_i x_q(i,b) y_q(i,b) for up to 6 pairs
// of column-major N x nb blocks (leading dimension N); out[q*nb + b] on the
// device.  Deterministic two-pass reduction.
int multidot(rsg_ctx *ctx, int npair, const double *const *x, const double *const *y, int nb,
             double alpha, double *out);
int multidot(rsg_ctx *ctx, int npair, const float *const *x, const float *const *y, int nb, double alpha,
             double *out);
// scf.hip
int scf_set_local_gth(rsg_ctx *ctx, int natom, const double *frac, const int *kind, int nkind, const double *zps,
                      const double *rloc, const double *cloc);
int scf_constants(rsg_ctx *ctx, double *eewald, double *nel);
int scf_atom_count(const rsg_ctx *ctx);
int scf_field(rsg_ctx *ctx, int field, double *host, bool get);
int scf_init_density(rsg_ctx *ctx, double width);
int scf_normalize(rsg_ctx *ctx);
int scf_calc_density(rsg_ctx *ctx, const double *occ);
int scf_hartree_xc(rsg_ctx *ctx, int field, double *eh, double *exc);
int scf_update_vloc(rsg_ctx *ctx);
int scf_density_integrals(rsg_ctx *ctx, int field, double *eloc, double *eion);
int scf_mix_init(rsg_ctx *ctx, int mmix, double beta, int imix);
int scf_mix_pulay(rsg_ctx *ctx);
void scf_release(rsg_ctx *ctx);
void scf_reset(rsg_ctx *ctx);
}  // namespace rsg
