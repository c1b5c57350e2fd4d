/*
 * oracle.h -- CPU FP64 restatement of RSDFT's real8 Gamma-point inner loop.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or as the timed CPU baseline).  The shipped GPU path never calls it.
 *
 * Each function names the reference routine (file:line, paths relative to the
 * reference tree's src/) whose arithmetic it restates.  Conventions:
 *   - grid point (i1,i2,i3), 0-based, linear index j = i1 + i2*N1 + i3*N1*N2
 *     (x fastest), as d_kinetic_sol_module.f90:129;
 *   - wavefunction blocks are column-major psi[j + ib*ld] (band ib contiguous),
 *     as tpsi(ML0,nb) in hamiltonian_module.f90:118-122;
 *   - lattice aa[3*i + r] = component r of lattice vector i (aa(r,i) in the
 *     reference), fractional atom coordinates frac[3*a + i].
 *
 * Pinning status of each function is recorded in tests/test_oracle_*.py.
 */
#ifndef RSDFT_ORACLE_H
#define RSDFT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXORB 4

typedef struct orc_ham orc_ham;

/* fd_module.f90:18-54 (get_coef_lapla_fd): lap[m+md] for m=-md..md */
void orc_fd_lap(int md, double *lap);

/* kinetic_module.f90:92-101 + get_ggg_kinetic :196-211.
 * coef_lap[axis*md + (m-1)]; returns nonzero for a non-orthogonal cell. */
int orc_kinetic_coef(const double *aa, const int *ngrid, int md,
                     double *coef_lap0, double *coef_lap, double *ggg6);

/* Hamiltonian context: grid, FD coefficients, Vloc, projector tables. */
orc_ham *orc_ham_create(int n1, int n2, int n3, const double *aa, int md);
void orc_ham_destroy(orc_ham *h);
double orc_ham_dv(const orc_ham *h);
void orc_ham_hgrid(const orc_ham *h, double *hgrid3);
void orc_ham_set_vloc(orc_ham *h, const double *vloc);
/* Replace the projector tables with caller-supplied ones (JJP 0-based). */
int orc_ham_set_proj(orc_ham *h, int nzlma, int maxmjj, const int *mjj,
                     const int *jjp, const double *uvk, const int *iuv);

/* GTH projector tables, restating prep_ps_nloc2 for ippform==4:
 *   init_ps_nloc_gth       ps_nloc_gth_module.f90:137-180  (cutoff radius)
 *   make_minimal_box       minimal_box_module.f90:14-165
 *   prep_ps_nloc_gth       ps_nloc_gth_module.f90:19-134   (radial values)
 *   Ylm                    ylm_module.f90:11-73,133-175
 *   uV = uV_tmp*Ylm        ps_nloc2_module.f90:760-790
 *   MJJ count / fold       ps_nloc2_module.f90:800-818, 1070-1167
 *   sort_index_sub         hsort_module.f90 (ascending JJP per lma)
 * Per-kind arrays are [kind*ORC_MAXORB + iorb]; hnl < 0 flips iuV
 * (ps_gth_module.f90:129-138). */
int orc_ham_build_gth(orc_ham *h, int natom, const double *frac,
                      const int *kind, int nkind, const int *norb,
                      const int *lo, const int *no, const double *rps,
                      const double *hnl);
void orc_ham_proj_sizes(const orc_ham *h, int *nzlma, int *maxmjj);
void orc_ham_proj_get(const orc_ham *h, int *mjj, int *jjp, double *uvk,
                      int *iuv);
/* GTH cutoff radius, ps_nloc_gth_module.f90:137-180 */
double orc_gth_cutoff(int L, int n, double rl);
/* Real spherical harmonic, ylm_module.f90:11-73 */
double orc_ylm(double x, double y, double z, int L, int m);

/* d_op_kinetic_sol with vloc: d_kinetic_sol_module.f90:244-322
 * (htpsi overwritten by (T + Vloc) psi). */
void orc_op_kinetic(const orc_ham *h, const double *psi, double *hpsi,
                    int64_t ld, int nb);
/* d_op_ps_nloc2_hp: d_ps_nloc2_op_module.f90:146-288 (accumulates). */
void orc_op_nonlocal(const orc_ham *h, const double *psi, double *hpsi,
                     int64_t ld, int nb);
/* d_hamiltonian1: hamiltonian_module.f90:118-144 */
void orc_hpsi(const orc_ham *h, const double *psi, double *hpsi, int64_t ld,
              int nb);

/* dot_product: ext0/dotp.f90:1-27 (m==1 form) */
double orc_dot(const double *a, const double *b, int64_t n, double alpha);

/* gram_schmidt_t / gram_schmidt_sub: gram_schmidt_t_module.f90:85-308,
 * NBLK = MB (one band rank), NBLK1 = 4. Acts on all mb columns. */
void orc_gram_schmidt_t(double *u, int64_t ld, int64_t n, int mb, double dv);
/* d_gram_schmidt_lusl: gram_schmidt_lusl_module.f90:241-328 (Cholesky-QR).
 * Returns nonzero if the overlap is not positive definite. */
int orc_gram_schmidt_lusl(double *u, int64_t ld, int64_t n, int mb,
                          double dv);

/* Dense helpers standing in for LAPACK (third-party, absent here). */
int orc_syev_jacobi(int n, double *a, double *w); /* DSYEV('V','U') role */
int orc_potrf_upper(int n, double *a);            /* DPOTRF('U') role */
int orc_trtri_upper(int n, double *a);            /* DTRTRI('U','N') role */
/* DSYGV(itype=1,'V','U') on a 2x2 pencil, cg_module.f90:305-316 role.
 * a,b column-major 2x2; returns w[2] ascending and B-normalised vectors
 * in a (column k = eigenvector k). */
int orc_sygv2(double *a, double *b, double *w);

/* subspace_diag_la: subspace_diag_la_module.f90:18-223 */
int orc_subspace_diag(const orc_ham *h, double *u, int64_t ld, int mb,
                      double *esp);

/* precond_hprsdft: cgpc_hprsdft_module.f90:133-361 (E0 = 0, mloop steps) */
void orc_precond_hprsdft(const orc_ham *h, int mloop, const double *gk,
                         double *pgk, int64_t ld, int nb);

/* d_conjugate_gradient_1: cg_module.f90:112-431 with MB_d = band block mbd,
 * IPC = 4 (pc_kind 4) or 0 (no preconditioner). */
int orc_conjugate_gradient(const orc_ham *h, double *u, int64_t ld, int mb,
                           int mbd, int ncg, int pc_kind, int mloop,
                           double *esp, double *res);
/* The same with the SWGS keyword (cg_module.f90:42-50): 1 or 2 applies
 * cggs (cggs_module.f90:14-127) to each preconditioned gradient. */
int orc_conjugate_gradient_gs(const orc_ham *h, double *u, int64_t ld, int mb,
                              int mbd, int ncg, int pc_kind, int mloop, int swgs,
                              double *esp, double *res);

/* d_esp_calc: esp_calc_module.f90:17-68 (dv = 1/|psi_1|^2 on first use) */
void orc_esp_calc(const orc_ham *h, const double *u, int64_t ld, int mb,
                  double *esp);

/* Dense H (N x N, column-major) for tiny grids: kinetic+local built column
 * by column from orc_hpsi on unit vectors. */
void orc_dense_h(const orc_ham *h, double *H);

void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
