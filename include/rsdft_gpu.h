/*
 * rsdft_gpu.h -- C ABI of the MI355X (gfx950) Kohn-Sham inner loop.
 *
 * One context owns one GPU's copy of the operator state (grid, FD
 * coefficients, local potential, projector tables) and the device-resident
 * wavefunction store, the way RSDFT's module globals own them on the CPU.
 * Every entry point returns 0 on success and a nonzero RSG_ERR_* code on
 * failure; rsg_last_error() then names the cause.  Calls on one context are
 * not re-entrant; work is queued on the context's own HIP stream and entry
 * points that return host-visible results synchronise that stream.
 *
 * Layout conventions (identical to the reference):
 *   grid point (i1,i2,i3), 0-based, j = i1 + i2*N1 + i3*N1*N2 (x fastest),
 *     kinetic/d_kinetic_sol_module.f90:129;
 *   band blocks are column-major psi[j + ib*ld], ld >= N, as tpsi(ML0,nb)
 *     in hamiltonian_module.f90:118-122;
 *   lattice aa[3*i + r] = component r of lattice vector i (aa(r,i));
 *   projector tables JJP(maxMJJ,nzlma) / uVk(maxMJJ,nzlma) column-major with
 *     0-based JJP, MJJ(nzlma), iuV(nzlma), ps_nloc2_variables.f90.
 */
#ifndef RSDFT_GPU_H
#define RSDFT_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSG_OK 0
#define RSG_ERR_ARG 1      /* bad argument or size */
#define RSG_ERR_HIP 2      /* HIP runtime / library failure */
#define RSG_ERR_STATE 3    /* call out of order (e.g. no grid yet) */
#define RSG_ERR_NOTPD 4    /* overlap matrix not positive definite */
#define RSG_ERR_SOLVER 5   /* dense eigensolver failed */
#define RSG_ERR_UNSUPPORTED 6

typedef struct rsg_ctx rsg_ctx;

/* Context lifetime.  No reference counterpart (module initialisation in
 * main.f90:266-388 plays this role). */
int rsg_create(rsg_ctx **ctx, int device);
int rsg_destroy(rsg_ctx *ctx);
const char *rsg_last_error(const rsg_ctx *ctx);
int rsg_synchronize(rsg_ctx *ctx);
/* The context's hipStream_t, for callers that record events on it. */
void *rsg_stream(rsg_ctx *ctx);

/* Grid and FD kinetic coefficients.  Replaces Init_Rgrid
 * (rgrid_sol_module.f90:26-39: Hgrid, dV) and init_kinetic
 * (kinetic/kinetic_module.f90:28-180; get_coef_lapla_fd, fd_module.f90:18-54).
 * Orthorhombic cells only (the Gamma-path cross terms of
 * d_kinetic_sol_module.f90:351-529 vanish). */
int rsg_set_grid(rsg_ctx *ctx, int n1, int n2, int n3, const double *aa,
                 int md);
/* dv, hgrid[3], coef_lap0 and coef_lap[axis*md + m-1] as the context uses
 * them. */
int rsg_get_grid(rsg_ctx *ctx, double *dv, double *hgrid, double *coef_lap0,
                 double *coef_lap);

/* Local potential Vloc(:,s) (host array of N doubles).  Replaces the module
 * array of localpot_module.f90:14 set by update_localpot (:209-215). */
int rsg_set_vloc(rsg_ctx *ctx, const double *vloc);

/* Projector tables in the reference's layout (ps_nloc2_variables.f90:
 * MJJ, JJP, d_uVk, iuV) plus an optional lma -> atom map (amap, may be
 * NULL) that lets the context gather the projectors of one atom together. */
int rsg_set_projectors(rsg_ctx *ctx, int nzlma, int maxmjj, const int *mjj,
                       const int *jjp, const double *uvk, const int *iuv,
                       const int *amap);
/* Build the GTH tables for a set of atoms.  Replaces prep_ps_nloc2 for
 * ippform 4 (ps_nloc2_module.f90:42-1167 with prep_ps_nloc_gth,
 * ps_nloc_gth_module.f90:19-180, and Ylm, ylm_module.f90:11-73).
 * frac[3*a+i] fractional positions, kind[a] species; per-species arrays are
 * [kind*4 + iorb] (norb[kind] <= 4), h < 0 flips iuV.  Where the analytic
 * form (below) applies, the reference-layout tables are built on first
 * use: by rsg_get_projector_sizes / rsg_get_projectors (host tables) or
 * rsg_set_nonlocal_mode(ctx, 1) (tables and their device layouts). */
int rsg_build_gth_projectors(rsg_ctx *ctx, int natom, const double *frac,
                             const int *kind, int nkind, const int *norb,
                             const int *lo, const int *no, const double *rps,
                             const double *hnl);
/* How hpsi applies V_NL: 0 = automatic (analytic when available), 1 =
 * the tables, 2 = analytic (RSG_ERR_STATE at hpsi time if unavailable).
 * The analytic form evaluates GTH (l <= 1, n = 1) projectors on the fly
 * from per-atom separable Gaussian factors instead of streaming uVk; it is
 * built by rsg_build_gth_projectors for orthorhombic cells longer than
 * twice the largest cutoff, and dropped by rsg_set_projectors.  No
 * reference counterpart (an implementation choice of this library). */
int rsg_set_nonlocal_mode(rsg_ctx *ctx, int mode);
int rsg_get_nonlocal_mode(rsg_ctx *ctx, int *mode, int *analytic_available);
int rsg_get_projector_sizes(rsg_ctx *ctx, int *nzlma, int *maxmjj,
                            int64_t *nentries);
int rsg_get_projectors(rsg_ctx *ctx, int *mjj, int *jjp, double *uvk,
                       int *iuv);

/* hpsi on device buffers: hpsi = (T + Vloc) psi + V_NL psi, overwriting
 * hpsi.  Replaces d_hamiltonian1 (hamiltonian_module.f90:118-144): the fused
 * kinetic+local stencil (d_kinetic_sol_module.f90:244-292) and
 * d_op_ps_nloc2_hp (d_ps_nloc2_op_module.f90:146-288).  psi and hpsi must
 * not overlap.  Asynchronous on the context stream. */
int rsg_hpsi_f64(rsg_ctx *ctx, const double *d_psi, double *d_hpsi,
                 int64_t ld, int nb);
/* rsg_hpsi_f64 that allocates nothing: no hipMalloc, hipFree or allocating
 * library call, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  No counterpart
 * in the reference.  The nonlocal projections go to a fixed 16 MiB device
 * scratch array (one per GPU, held until the process ends) instead of
 * context buffers; the bands run in chunks small enough to fit, and a call
 * whose single band's projections exceed the scratch fails with
 * RSG_ERR_ARG.  Results equal rsg_hpsi_f64's.  The scratch holds the last
 * call's projections (every call writes what it reads); calls that overlap
 * on one GPU, from any context, race on it.  The context must be fully set
 * up (grid, Vloc, projectors and nonlocal mode) beforehand. */
int rsg_hpsi_f64_noalloc(rsg_ctx *ctx, const double *d_psi, double *d_hpsi,
                         int64_t ld, int nb);
/* The same operator in FP32 storage with FP32 accumulation (tables rounded
 * to FP32).  No reference counterpart (mixed-precision addition). */
int rsg_hpsi_f32(rsg_ctx *ctx, const float *d_psi, float *d_hpsi, int64_t ld,
                 int nb);
/* Local part of hpsi alone: hpsi (+)= (T + Vloc) psi, accumulating when
 * accumulate != 0.  Replaces d_op_kinetic with its Vloc argument
 * (kinetic_module.f90:214-243), which hamiltonian0
 * (hamiltonian_module.f90:62-116) calls before op_nonlocal.  Same kernel
 * and launch shape as rsg_hpsi_*; asynchronous on the context stream. */
int rsg_op_kinetic_f64(rsg_ctx *ctx, const double *d_psi, double *d_hpsi,
                       int64_t ld, int nb, int accumulate);
int rsg_op_kinetic_f32(rsg_ctx *ctx, const float *d_psi, float *d_hpsi,
                       int64_t ld, int nb, int accumulate);
/* hpsi on host arrays (copied through device scratch); the entry point the
 * Fortran `hamiltonian` shim binds.  Synchronous. */
int rsg_hpsi_f64_host(rsg_ctx *ctx, const double *psi, double *hpsi,
                      int64_t ld, int nb);

/* Launch shape of the hpsi kernels for the current grid, per precision
 * (64 or 32): rows per thread ry (1 or 2), tile rows ty, z-planes per
 * workgroup zlen, bands per Infinity-Cache chunk, band tile bt of the
 * nonlocal apply (4/8/16/20/32: 16, 32, 64, 80 or 128 bands per matrix-core
 * wave of the analytic form; the table form's scatter takes 4/8/16, and 8
 * for the others).  Zero fields are derived from the grid.  No reference
 * counterpart (launch tuning of this implementation). */
int rsg_set_hpsi_shape(rsg_ctx *ctx, int precision, int ry, int ty, int zlen,
                       int chunk, int bt);
/* Order of the two passes over H psi with analytic projectors: 0 = the
 * nonlocal kernel writes H psi and the stencil adds to it, 1 = the stencil
 * writes and the nonlocal kernel adds.  No reference counterpart. */
int rsg_set_hpsi_order(rsg_ctx *ctx, int precision, int order);
int rsg_get_hpsi_order(rsg_ctx *ctx, int precision, int *order);
/* Form of the analytic projections (d_ps_nloc2_op_module.f90:182-191):
 * 0 = per sphere segment (one atom's points on one plane; 32 bands per
 * wave), 2 = the same with 64 bands per wave, 1 = by row tiles (grid rows
 * staged once per plane, every atom meeting them).  The getter reports
 * what an nb-band call would use (1 falls back to 0 where the row tiles do
 * not fit).  No reference counterpart. */
int rsg_set_hpsi_proj(rsg_ctx *ctx, int precision, int proj);
/* LDS (KiB, 0..64) reserved by each wave of the segment projections (forms
 * 0 and 2), unused by the kernel: fewer waves in flight per CU, so the
 * sphere chords they share stay in L2 (the autotuner tries 4-32).  No
 * reference counterpart. */
int rsg_set_hpsi_proj_lds(rsg_ctx *ctx, int precision, int kib);
int rsg_get_hpsi_proj_lds(rsg_ctx *ctx, int precision, int *kib);
/* Segment projections (forms 0 and 2) in 1 (default), 2, 4 or 8 pieces
 * per sphere segment, each piece a wave of its own with its own partial
 * sums: shorter waves read the chords that neighbouring spheres share in a
 * plane closer together in time, so the second read finds them in L2 (the
 * autotuner tries 2-8).  The partial sums are added in a fixed order; the
 * result differs from one piece only by rounding.  No reference
 * counterpart. */
int rsg_set_hpsi_proj_split(rsg_ctx *ctx, int precision, int pieces);
int rsg_get_hpsi_proj_split(rsg_ctx *ctx, int precision, int *pieces);
/* Where rsg_build_gth_projectors builds the analytic projector
 * description (the set-up half of ps_nloc2_module.f90:42-1167 for
 * l <= 1, n = 1 GTH sets): 0 (default) = with device kernels, 1 = on the
 * host.  Both give the same arrays (factor tables up to the rounding of
 * exp); takes effect at the next build. */
int rsg_set_projector_setup(rsg_ctx *ctx, int on_host);
/* Copy one array of the analytic description (FP64 form) to host memory
 * and report its length: which = 0 sphere points, 1 segment begins,
 * 2 segment ends, 3 segment atoms, 4 segment launch order, 5 segments per
 * atom (offsets), 6 per-point hit words, 7 tile offsets, 8 tile atoms,
 * 9 s factors, 10 p factors, 11 axis displacements, 12 amplitudes,
 * 13 scales.  out may be NULL to query the length only.  For tests and
 * inspection; no reference counterpart. */
int rsg_get_analytic_array(rsg_ctx *ctx, int which, void *out, int64_t *count);
int rsg_get_hpsi_proj(rsg_ctx *ctx, int precision, int *proj);
/* Grid points along x per stencil thread: 1, or 2 for the pair kernel
 * (FP32 with md = 6, n1 even, tile rows dividing n2; elsewhere the call
 * falls back to 1).  The getter reports what an nb-band call would use.
 * No reference counterpart. */
int rsg_set_hpsi_xpoints(rsg_ctx *ctx, int precision, int rx);
int rsg_get_hpsi_xpoints(rsg_ctx *ctx, int precision, int nb, int *rx);
/* The shape an nb-band call would use. */
int rsg_get_hpsi_shape(rsg_ctx *ctx, int precision, int nb, int *ry, int *ty,
                       int *zlen, int *chunk, int *bt);
/* Time candidate shapes for nb bands (reps calls each, on internal scratch
 * bands) and keep the fastest for this context.  best_ms may be NULL. */
int rsg_autotune_hpsi(rsg_ctx *ctx, int precision, int nb, int reps,
                      float *best_ms);

/* The Gram products U^T X (N x nb by N x nb, K = N) of Gram-Schmidt and
 * subspace_diag: time one rocBLAS GEMM over all of K against K-splits into
 * 8..512 chunks (one strided-batched GEMM plus a chunk sum), reps calls
 * each on scratch bands, and keep the fastest for (N, nb) in this context.
 * chunk receives the points per chunk (0 = one GEMM); best_ms and chunk
 * may be NULL.  No reference counterpart (launch tuning). */
int rsg_autotune_gram(rsg_ctx *ctx, int nb, int reps, float *best_ms,
                      int64_t *chunk);

/* Start bands on the device: columns n0 .. n0+nb-1 (0-based) of the band
 * store get uniform [0, 1) values, one counter-based draw per point and band
 * (SplitMix64 of seed + (b N + j + 1) * 0x9E3779B97F4A7C15, top 53 bits), so
 * the result depends only on seed.  Replaces random_initial_wf
 * (wf_module.f90:166-191), which draws uniform numbers per point and band
 * with the Fortran generator.  Asynchronous on the context stream. */
int rsg_random_bands(rsg_ctx *ctx, int n0, int nb, uint64_t seed);

/* Device-resident wavefunction store, the counterpart of
 * wf_module::unk(ML_0:ML_1, MB, k, s) (wf_module.f90:131).  Upload and
 * download move nb bands starting at store band n0 (0-based). */
int rsg_alloc_bands(rsg_ctx *ctx, int mb);
int rsg_upload_bands(rsg_ctx *ctx, const double *u, int64_t ld, int n0,
                     int nb);
int rsg_download_bands(rsg_ctx *ctx, double *u, int64_t ld, int n0, int nb);
int rsg_bands_device_ptr(rsg_ctx *ctx, double **d_u, int64_t *ld);
/* Storage of the band store for the mixed-precision phase (no reference
 * counterpart; unk is real(8) there): 32 moves the bands to an FP32 store,
 * on which rsg_subspace_diag (FP32 hpsi and products summed in FP64, the
 * FP64 eigensolve, an FP32 rotation) and rsg_conjugate_gradient_f32 (in
 * place) then work.  rsg_gram_schmidt orthonormalises in FP64 through the
 * FP64 store, which then holds the same bands in FP64: rsg_esp_calc and
 * rsg_calc_density read it until the next change to the FP32 store (before
 * that, FP32 hpsi with FP64 sums).  64 moves the bands back (no copy when
 * the FP64 store is current).  Upload, download, the device pointer, random bands, the FP64
 * CG and SWGS 1/2 return RSG_ERR_STATE / RSG_ERR_UNSUPPORTED meanwhile.
 * rsg_alloc_bands starts a store in FP64. */
int rsg_set_band_precision(rsg_ctx *ctx, int bits);
int rsg_get_band_precision(rsg_ctx *ctx, int *bits);

/* Orthonormalise bands n0..n1 (1-based, inclusive) of the store under
 * <u,v> = dV u^T v.  Replaces gram_schmidt(n0,n1,k,s)
 * (gram_schmidt/gram_schmidt_module.f90:35-79) with the Cholesky-QR form of
 * d_gram_schmidt_lusl (gram_schmidt_lusl_module.f90:241-328): S = dV U^T U,
 * S = R^T R, U <- U R^-1.  RSG_ERR_NOTPD if S is not positive definite. */
int rsg_gram_schmidt(rsg_ctx *ctx, int n0, int n1);

/* Rayleigh-Ritz on the whole store.  Replaces subspace_diag(k,s,unk,esp)
 * (subspace_diag/subspace_diag_module.f90:25-65 -> subspace_diag_la,
 * subspace_diag_la_module.f90:18-223).  esp (host, mb) ascending. */
int rsg_subspace_diag(rsg_ctx *ctx, double *esp);

/* Band-block CG update of the whole store.  Replaces
 * conjugate_gradient(n1,n2,k,s,unk,esp,res) (conjugate_gradient/
 * cg_module.f90:53-431).  pc_kind 4 = hprsdft 7-point preconditioner with
 * mloop steps (cgpc_hprsdft_module.f90:133-361), 0 = none.  esp, res (host,
 * mb). */
int rsg_conjugate_gradient(rsg_ctx *ctx, int ncg, int pc_kind, int mloop,
                           double *esp, double *res);
/* enable != 0: rsg_conjugate_gradient(_f32) with pc_kind 0 or 4 and SWGS 0
 * replays each band block's kernel sequence from a captured HIP graph
 * (captured on the second call with the same block, buffers and launch
 * parameters, replayed from the third on; any change of those captures
 * anew).  Results equal the direct launches bit for bit.  0 (default)
 * launches directly and frees the captured graphs.  No reference
 * counterpart (launch-overhead reduction for many small cells). */
int rsg_set_cg_graphs(rsg_ctx *ctx, int enable);
/* CG band blocks run from a captured graph since the context was created. */
int rsg_cg_graph_launches(rsg_ctx *ctx, int64_t *launches);
/* Captures the runtime refused since the context was created (those band
 * blocks ran, and their keys keep running, with direct launches). */
int rsg_cg_graph_failures(rsg_ctx *ctx, int64_t *failures);
/* The SWGS keyword of the CG (read_cg, cg_module.f90:42-50): 0 (default)
 * leaves the preconditioned gradients as they are; 1 and 2 project each
 * one against the other bands of the store before the direction update,
 * cggs (cggs_module.f90:14-127, called at cg_module.f90:249-251): 1 against
 * the bands before it, 2 against all bands but itself. */
int rsg_set_cg_swgs(rsg_ctx *ctx, int swgs);
/* pc_kind 5 = the learned preconditioner declared below; pc_kind 6 = the
 * hybrid of both, P g = scale * P_hprsdft g + net(g)
 * (rsg_lpc_set_hybrid_scale), with the same descent check. */

/* The hprsdft preconditioner on its own: pg = P g for nb device bands of
 * leading dimension ld (g, pg), mloop CG steps on -1/2 lap_7 x = g from
 * x0 = g.  Replaces precond_hprsdft (cgpc_hprsdft_module.f90:133-270) with
 * Pgk entering as gk (its caller, cg_module.f90, sets Pgk = gk first); E
 * is not an argument, since the reference's E0 is zero.  Asynchronous on
 * rsg_stream(ctx): the reference's all(rr1 < 1e-24) early exit is decided
 * on the device (every step then takes a zero step, leaving pg = g). */
int rsg_precond_hprsdft_f64(rsg_ctx *ctx, int mloop, const double *g, double *pg,
                            int64_t ld, int nb);
int rsg_precond_hprsdft_f32(rsg_ctx *ctx, int mloop, const float *g, float *pg,
                            int64_t ld, int nb);
/* Storage of the hprsdft preconditioner's internal vectors (its residual,
 * direction and next direction) in FP32 calls: 32 (default) or 16
 * (bfloat16, half their bytes; g, P g, the stencil sums and every update
 * stay FP32, the dot products FP64).  A precision choice of this build with
 * no reference counterpart: precond_hprsdft (cgpc_hprsdft_module.f90:
 * 133-270) is FP64.  Applies to rsg_precond_hprsdft_f32 and to
 * rsg_conjugate_gradient_f32 when every band column allows 2- or 4-wide
 * vectors (else the FP32 kernels run). */
int rsg_set_precond_storage(rsg_ctx *ctx, int bits);
/* The same band update with every CG and preconditioner vector, hpsi
 * included, in FP32 storage with FP32 arithmetic; dot products accumulate
 * in FP64, the 2x2 Ritz step is FP64 and the accepted update lands in the
 * FP64 store.  No reference counterpart (mixed-precision addition for the
 * early SCF iterations). */
int rsg_conjugate_gradient_f32(rsg_ctx *ctx, int ncg, int pc_kind, int mloop,
                               double *esp, double *res);

/* e_n = dV <psi_n|H psi_n>.  Replaces d_esp_calc (esp_calc_module.f90:17-68). */
int rsg_esp_calc(rsg_ctx *ctx, double *esp);

/* ---- density/potential half of the SCF step (FP64, device-resident) ----
 * Fields the context holds on the grid (N doubles each). */
#define RSG_FIELD_RHO 0     /* input density of the step (mixing state) */
#define RSG_FIELD_RHO_OUT 1 /* density of the current bands */
#define RSG_FIELD_VION 2    /* local pseudopotential */
#define RSG_FIELD_VH 3      /* Hartree potential */
#define RSG_FIELD_VXC 4     /* exchange-correlation potential */
#define RSG_FIELD_VLOC 5    /* Vion + VH + Vxc, the potential hpsi applies */

/* GTH local potential in G-space on the density grid's cutoff sphere
 * (ps_local_gth_module.f90:13-58; construct_ps_local_fft0,
 * ps_local_fft0_module.f90:14-86), the Ewald energy of the ions
 * (ewald_module.f90:50-62) and the electron count.  Per species k:
 * zps[k], rloc[k], cloc[4k..4k+3] = C1..C4. */
int rsg_set_local_gth(rsg_ctx *ctx, int natom, const double *frac,
                      const int *kind, int nkind, const double *zps,
                      const double *rloc, const double *cloc);
int rsg_get_scf_constants(rsg_ctx *ctx, double *eewald, double *nelectron);
int rsg_get_field(rsg_ctx *ctx, int field, double *out);
/* Only RSG_FIELD_RHO, RSG_FIELD_RHO_OUT and RSG_FIELD_VLOC can be set. */
int rsg_set_field(rsg_ctx *ctx, int field, const double *in);
/* Initial density: atom-centred Gaussians of the given width (bohr)
 * normalised to the electron count.  No reference counterpart: the
 * reference starts GTH runs from a compiler-RNG random density
 * (construct_RandomInitrho, ps_initrho_module.f90:58-134). */
int rsg_init_density(rsg_ctx *ctx, double width);
/* rho <- rho * Nel / integral(rho) (normalize_density). */
int rsg_normalize_density(rsg_ctx *ctx);
/* RHO_OUT = sum_n occ[n] psi_n^2 over the band store
 * (calc_density, density_module.f90:120-151). */
int rsg_calc_density(rsg_ctx *ctx, const double *occ);
/* VH and Vxc (LDA-PZ81) of field RHO or RHO_OUT; returns the Hartree
 * energy 1/2 int rho VH and Exc (hartree_sol_fftw_module.f90:29-130,
 * xc/xc_ldapz81_module.f90:17-216). */
int rsg_hartree_xc(rsg_ctx *ctx, int field, double *e_hartree, double *e_xc);
/* Vloc = Vion + VH + Vxc (scf_module.f90:446-456), FP64 and FP32 copies. */
int rsg_update_vloc(rsg_ctx *ctx);
/* int rho Vloc and int rho Vion for field RHO or RHO_OUT
 * (total_energy_module.f90:320-337). */
int rsg_density_integrals(rsg_ctx *ctx, int field, double *e_loc,
                          double *e_ion);
/* Pulay density mixing (mixing_pulay_module.f90:29-121; simple mixing
 * until mmix steps are held): mix_init stores RHO as the first input; each
 * mix_pulay call takes RHO_OUT as the output of the latest input and writes
 * the next input density into RHO. */
int rsg_mix_init(rsg_ctx *ctx, int mmix, double beta);
/* The same with the reference's IMIX choice: 10 = Pulay on the density in
 * real space (pulay_mixing, the default), 30 = Pulay on its Fourier
 * coefficients with Kerker weights (pulay_g_mixing,
 * mixing_pulay_module.f90:124-337; mmix <= 4), which damps the
 * long-wavelength charge sloshing of large cells. */
int rsg_mix_init_imix(rsg_ctx *ctx, int mmix, double beta, int imix);
int rsg_mix_pulay(rsg_ctx *ctx);

/* ---- a whole SCF run in native code ----
 * The SCF loop of calc_scf (scf_module.f90:183-481, Ndiag = 1) on the
 * context's cell, with nothing crossing to the caller until it ends: start
 * density (atomic Gaussians of init_width), uniform [0, 1) start bands from
 * seed (random_initial_wf), Gram-Schmidt (main.f90:392), then per
 * iteration subspace_diag, conjugate_gradient, gram_schmidt, esp_calc,
 * Fermi-Dirac occupations at ekbt, the Harris and Kohn-Sham energies, and
 * Pulay mixing (IMIX 10 or 30) with new Hartree/XC/Vloc, until
 * |Etot_k - Etot_(k-1)| / natom < etol_per_atom between two FP64
 * iterations or max_iter.  mixed = 1 runs FP32 CG/hpsi while that
 * difference is at least fp64_below (0: twice etol_per_atom) and FP32
 * keeps improving within fp32_stall iterations, then FP64.  The grid, the
 * projectors (rsg_build_gth_projectors) and the local part
 * (rsg_set_local_gth) must be set.  No reference counterpart as one call:
 * it lets many cells run concurrently from host threads without returning
 * to the caller per step (the Python and Fortran drivers make the same
 * calls one by one). */
typedef struct {
  int nband;
  int ncg, pc_kind, mloop; /* NCG, IPC (0/4/5/6), MLOOP */
  int mmix, imix;          /* MMIX, IMIX (10 or 30) */
  double beta, ekbt;       /* BETA, EKBT (Ha) */
  double etol_per_atom;    /* stop test on |dEtot| / natom */
  int max_iter;
  double init_width;       /* bohr */
  uint64_t seed;           /* start bands */
  int mixed;
  double fp64_below;
  int fp32_stall;
  int fp32_store;          /* mixed: FP32 iterations on the FP32 band store */
  int precond_bits;        /* FP32 hprsdft storage: 32, or 16 (bfloat16) */
} rsg_scf_params;
typedef struct {
  int converged, iterations, fp32_iterations;
  double etot, fene, efermi, eewald, wall_s;
} rsg_scf_summary;
/* etot_history (max_iter doubles) and esp (nband doubles, final band
 * energies) may be NULL. */
int rsg_scf_run(rsg_ctx *ctx, const rsg_scf_params *params, rsg_scf_summary *out, double *etot_history,
                double *esp);

/* ---- learned CG preconditioner (pc_kind = 5 in rsg_conjugate_gradient*).
 * No reference counterpart: it replaces the kinetic smoother
 * precond_hprsdft (cgpc_hprsdft_module.f90:133-361) with a 3-layer 3x3x3
 * convolution net on the periodic grid, P g = s * net(g/s, v), s =
 * rms(g), v the cell's normalised Vloc; bfloat16 weights and activations,
 * FP32 accumulation on MFMA.  A block whose <Pg, g> is not positive falls
 * back to precond_hprsdft, for that step and the rest of the CG call.
 * Parameters (FP32, this order):
 * W1[16][2][27] b1[16] W2[16][16][27] b2[16] W3[1][16][27] b3[1], W[co][ci]
 * [tap], tap = d3*9 + d2*3 + d1 (offsets d - 1 along z, y, x). */
int rsg_lpc_param_count(void);
int rsg_lpc_set_params(rsg_ctx *ctx, const float *params, int n);
int rsg_lpc_get_params(rsg_ctx *ctx, float *params, int n);
/* normalised potential input; vloc NULL takes the context's current Vloc */
int rsg_lpc_set_potential(rsg_ctx *ctx, const double *vloc);
int rsg_lpc_apply_f64(rsg_ctx *ctx, const double *d_g, double *d_pg,
                      int64_t ld, int nb);
int rsg_lpc_apply_f32(rsg_ctx *ctx, const float *d_g, float *d_pg,
                      int64_t ld, int nb);
/* Training targets t = A^-1 g, A = -1/2 lap_7 + vweight (Vloc - min Vloc)
 * + shift, by `steps` CG steps from t = g (FP32, device, leading
 * dimension N): with vweight = 0 the converged form of the kinetic solve
 * that precond_hprsdft truncates after mloop steps. */
int rsg_lpc_target_f32(rsg_ctx *ctx, const float *d_g, float *d_t, int nb,
                       int steps, double shift, double vweight);
/* One Adam step (lr > 0) or a gradient evaluation only (lr = 0) on nsamp
 * residual/target pairs (device FP32, leading dimension N); loss = mean
 * ((P g - kappa t)/s)^2, kappa > 0 a fixed scale of the targets; grads
 * (host, may be NULL) in the parameter order. */
int rsg_lpc_train_step(rsg_ctx *ctx, const float *d_g, const float *d_t,
                       int nsamp, float kappa, float lr, double *loss,
                       float *grads);
int rsg_lpc_reset_optimizer(rsg_ctx *ctx);
/* CG gradients g_n = -2 (H u_n - e_n u_n) of every band in the store,
 * written FP32 to d_g (N x mb, device): the first gradient of a CG call
 * (cg_module.f90:207), exported as training residuals. */
int rsg_band_residuals_f32(rsg_ctx *ctx, float *d_g);
/* CG band blocks that fell back to precond_hprsdft since creation */
int rsg_lpc_fallbacks(rsg_ctx *ctx, int64_t *count);
/* Rejection limit of the learned preconditioner (no reference
 * counterpart; the reference has no learned preconditioner): once `limit`
 * CG calls in a row had a band block whose network output failed the
 * descent check, the following pc_kind 5/6 calls run precond_hprsdft
 * without trying the network, until the limit is set again (which also
 * clears the streak).  0 (the default) tries the network in every call. */
int rsg_lpc_set_reject_limit(rsg_ctx *ctx, int limit);
/* the current streak of rejected calls and the calls run without the
 * network because of the limit, since creation */
int rsg_lpc_reject_state(rsg_ctx *ctx, int *streak, int64_t *skipped_calls);
/* pc_kind 6: the factor on the hprsdft output the network corrects
 * (default 1). */
int rsg_lpc_set_hybrid_scale(rsg_ctx *ctx, double scale);
/* y = a y + b x on n device floats, on rsg_stream(ctx): training-target
 * plumbing for the hybrid (t - mu P_hprsdft g).  No reference
 * counterpart. */
int rsg_axpby_f32(rsg_ctx *ctx, float *d_y, const float *d_x, float a,
                  float b, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
