/*
 * solver.c -- oracle: Rayleigh-Ritz (subspace_diag), band-block conjugate
 * gradient with the hprsdft preconditioner, per-band eigenvalues.
 * TEST INFRASTRUCTURE (see oracle.h).
 */
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* subspace_diag_la: subspace_diag_la_module.f90:18-223.
 * psit = H U (:79), Hsub = 0.5 dV (U^T psit + psit^T U) (DSYR2K 'U','T' :86),
 * DSYEV (:155), U <- U Hsub (:190). */
int orc_subspace_diag(const orc_ham *h, double *u, int64_t ld, int mb,
                      double *esp) {
  const int64_t n = h->N;
  double *psit = (double *)malloc(sizeof(double) * n * mb);
  for (int m = 0; m < mb; ++m) orc_hpsi(h, u + (int64_t)m * ld, psit + (int64_t)m * n, n, 1);
  double *hs = (double *)calloc((size_t)mb * mb, sizeof(double));
  const double zz = 0.5 * h->dv;
#pragma omp parallel for schedule(dynamic)
  for (int j = 0; j < mb; ++j)
    for (int i = 0; i <= j; ++i) {
      const double *ui = u + (int64_t)i * ld, *uj = u + (int64_t)j * ld;
      const double *pi = psit + (int64_t)i * n, *pj = psit + (int64_t)j * n;
      double s = 0.0;
      for (int64_t k = 0; k < n; ++k) s += ui[k] * pj[k] + pi[k] * uj[k];
      hs[i + (size_t)j * mb] = zz * s;
    }
  int info = orc_syev_jacobi(mb, hs, esp);
  /* rotation U <- U * V */
#pragma omp parallel
  {
    double *row = (double *)malloc(sizeof(double) * mb);
#pragma omp for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
      for (int j = 0; j < mb; ++j) row[j] = u[k + (int64_t)j * ld];
      for (int j = 0; j < mb; ++j) {
        double s = 0.0;
        for (int i = 0; i < mb; ++i) s += row[i] * hs[i + (size_t)j * mb];
        u[k + (int64_t)j * ld] = s;
      }
    }
    free(row);
  }
  free(hs);
  free(psit);
  return info;
}

/* precond_cg_mat2_omp: cgpc_hprsdft_module.f90:288-361 -- the 7-point
 * -1/2 Laplacian f = (c0 - E) g + c1 (g(x+1)+g(x-1)) + c2 (...) + c3 (...)
 * after the 1-deep periodic halo fill bcset_1; c0..c3 from :125-128. */
static void pc_mat(const orc_ham *h, const double *g, double *f) {
  const int n1 = h->n[0], n2 = h->n[1], n3 = h->n[2];
  double c0 = h->ggg[0] / (h->hgrid[0] * h->hgrid[0]) +
              h->ggg[1] / (h->hgrid[1] * h->hgrid[1]) +
              h->ggg[2] / (h->hgrid[2] * h->hgrid[2]);
  double c1 = -0.5 / (h->hgrid[0] * h->hgrid[0]) * h->ggg[0];
  double c2 = -0.5 / (h->hgrid[1] * h->hgrid[1]) * h->ggg[1];
  double c3 = -0.5 / (h->hgrid[2] * h->hgrid[2]) * h->ggg[2];
  const double d = c0 - 0.0; /* E0 = 0 (:149) */
  const int64_t n12 = (int64_t)n1 * n2;
#pragma omp parallel for schedule(static)
  for (int i3 = 0; i3 < n3; ++i3)
    for (int i2 = 0; i2 < n2; ++i2)
      for (int i1 = 0; i1 < n1; ++i1) {
        int64_t j = i1 + (int64_t)i2 * n1 + i3 * n12;
        int xp = i1 + 1 == n1 ? 0 : i1 + 1, xm = i1 == 0 ? n1 - 1 : i1 - 1;
        int yp = i2 + 1 == n2 ? 0 : i2 + 1, ym = i2 == 0 ? n2 - 1 : i2 - 1;
        int zp = i3 + 1 == n3 ? 0 : i3 + 1, zm = i3 == 0 ? n3 - 1 : i3 - 1;
        f[j] = d * g[j] + c1 * (g[xp + i2 * n1 + i3 * n12] + g[xm + i2 * n1 + i3 * n12]) +
               c2 * (g[i1 + yp * n1 + i3 * n12] + g[i1 + ym * n1 + i3 * n12]) +
               c3 * (g[i1 + i2 * n1 + zp * n12] + g[i1 + i2 * n1 + zm * n12]);
      }
}

static double plain_dot(const double *a, const double *b, int64_t n) {
  return orc_dot(a, b, n, 1.0); /* fixed summation order */
}

/* precond_hprsdft: cgpc_hprsdft_module.f90:133-270 -- mloop CG steps on
 * A x = gk started from x0 = Pgk (= gk on entry from the CG driver). */
void orc_precond_hprsdft(const orc_ham *h, int mloop, const double *gk,
                         double *pgk, int64_t ld, int nb) {
  if (mloop <= 0) return;
  const int64_t n = h->N;
  double *f = (double *)malloc(sizeof(double) * n);
  double *r = (double *)malloc(sizeof(double) * n * nb);
  double *p = (double *)malloc(sizeof(double) * n * nb);
  double *rr1 = (double *)malloc(sizeof(double) * nb);
  double *rr0 = (double *)malloc(sizeof(double) * nb);
  int all_small = 1;
  for (int b = 0; b < nb; ++b) {
    const double *g = gk + (int64_t)b * ld;
    pc_mat(h, g, f);
    double *rb = r + (int64_t)b * n, *pb = p + (int64_t)b * n;
    for (int64_t i = 0; i < n; ++i) {
      rb[i] = g[i] - f[i];
      pb[i] = rb[i];
    }
    rr1[b] = plain_dot(rb, rb, n);
    if (!(rr1[b] < 1e-24)) all_small = 0;
  }
  if (!all_small) {
    for (int it = 1; it <= mloop; ++it) {
      for (int b = 0; b < nb; ++b) {
        double *rb = r + (int64_t)b * n, *pb = p + (int64_t)b * n;
        double *x = pgk + (int64_t)b * ld;
        pc_mat(h, pb, f);
        double pap = plain_dot(pb, f, n);
        rr0[b] = rr1[b];
        /* a band at rest (zero residual, so pAp = 0) takes a zero step
         * rather than 0/0; the reference has no such band in practice */
        double a = pap != 0.0 ? rr0[b] / pap : 0.0;
        for (int64_t i = 0; i < n; ++i) x[i] += a * pb[i];
        if (it == mloop) continue;
        for (int64_t i = 0; i < n; ++i) rb[i] -= a * f[i];
        rr1[b] = plain_dot(rb, rb, n);
        double bb = rr0[b] != 0.0 ? rr1[b] / rr0[b] : 0.0;
        for (int64_t i = 0; i < n; ++i) pb[i] = rb[i] + bb * pb[i];
      }
    }
  }
  free(f);
  free(r);
  free(p);
  free(rr1);
  free(rr0);
}

/* cggs: cggs_module.f90:14-127 -- classical Gram-Schmidt of one
 * preconditioned gradient g against the store bands u(:, 0..mb-1): swgs 1
 * projects out bands 0..nt-1 (cggs_1 with ns = nt - 1, 1-based nt = this
 * band's store index + 1), swgs 2 every band but nt (cggs_2); any other
 * value does nothing.  All coefficients dV <u_m, g> come from the incoming
 * g, then g -= sum_m u_m c_m, as the reference's two loops do. */
static void cggs(const orc_ham *h, int swgs, const double *u, int64_t ld, int mb, int nt, double *g) {
  if (swgs != 1 && swgs != 2) return;
  const int64_t n = h->N;
  const int hi = swgs == 1 ? nt : mb;
  if (hi <= 0) return;
  double *c = (double *)calloc((size_t)hi, sizeof(double));
  for (int m = 0; m < hi; ++m) {
    if (m == nt) continue;
    const double *um = u + (int64_t)m * ld;
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += um[i] * g[i];
    c[m] = s * h->dv;
  }
  for (int m = 0; m < hi; ++m) {
    if (m == nt) continue;
    const double *um = u + (int64_t)m * ld;
    for (int64_t i = 0; i < n; ++i) g[i] -= um[i] * c[m];
  }
  free(c);
}

int orc_conjugate_gradient(const orc_ham *h, double *u, int64_t ld, int mb,
                           int mbd, int ncg, int pc_kind, int mloop,
                           double *esp, double *res) {
  return orc_conjugate_gradient_gs(h, u, ld, mb, mbd, ncg, pc_kind, mloop, 0, esp, res);
}

/* d_conjugate_gradient_1: cg_module.f90:112-431 */
int orc_conjugate_gradient_gs(const orc_ham *h, double *u, int64_t ld, int mb,
                              int mbd, int ncg, int pc_kind, int mloop, int swgs,
                              double *esp, double *res) {
  const int64_t n = h->N;
  const double dv = h->dv;
  if (ld != n) return 1; /* work arrays share the leading dimension */
  if (mbd < 1) mbd = 1;
  if (mbd > mb) mbd = mb;
  size_t blk = (size_t)n * mbd;
  double *hxk = (double *)malloc(sizeof(double) * blk);
  double *hpk = (double *)malloc(sizeof(double) * blk);
  double *gk = (double *)malloc(sizeof(double) * blk);
  double *pgk = (double *)malloc(sizeof(double) * blk);
  double *pk = (double *)malloc(sizeof(double) * blk);
  double *E = (double *)malloc(sizeof(double) * mbd);
  double *E1 = (double *)malloc(sizeof(double) * mbd);
  double *rb = (double *)malloc(sizeof(double) * mbd);
  double *gkgk = (double *)malloc(sizeof(double) * mbd);
  double *ut3 = (double *)malloc(sizeof(double) * 2 * mbd);
  for (int i = 0; i < mb; ++i) res[i] = esp[i] = 0.0;
  for (int ns = 0; ns < mb; ns += mbd) {
    int ne = ns + mbd < mb ? ns + mbd : mb;
    int nn = ne - ns;
    for (int b = 0; b < nn; ++b) E1[b] = 1.0e10;
    for (size_t i = 0; i < (size_t)n * nn; ++i) hxk[i] = 0.0;
    orc_hpsi(h, u + (int64_t)ns * ld, hxk, n, nn);
    for (int b = 0; b < nn; ++b) {
      const double *x = u + (int64_t)(ns + b) * ld;
      double *hx = hxk + (int64_t)b * n, *g = gk + (int64_t)b * n;
      E[b] = orc_dot(x, hx, n, dv);
      for (int64_t i = 0; i < n; ++i) g[i] = -2.0 * (hx[i] - E[b] * x[i]);
      rb[b] = orc_dot(g, g, n, dv);
    }
    for (int icg = 1; icg <= ncg + 1; ++icg) {
      memcpy(pgk, gk, sizeof(double) * n * nn);
      for (int b = 0; b < nn; ++b) res[ns + b] = rb[b] / 4.0;
      int conv = 1;
      for (int b = 0; b < nn; ++b)
        if (!(fabs(E[b] - E1[b]) < 1e-15)) conv = 0;
      if (conv) break;
      if (icg == ncg + 1) break;
      if (pc_kind == 4) orc_precond_hprsdft(h, mloop, gk, pgk, n, nn);
      for (int b = 0; b < nn; ++b) cggs(h, swgs, u, ld, mb, ns + b, pgk + (int64_t)b * n); /* :249-251 */
      for (int b = 0; b < nn; ++b)
        rb[b] = orc_dot(pgk + (int64_t)b * n, gk + (int64_t)b * n, n, dv);
      if (icg == 1) {
        memcpy(pk, pgk, sizeof(double) * n * nn);
      } else {
        for (int b = 0; b < nn; ++b) {
          /* a zero previous <Pg, g> (vanished gradient) restarts the
           * direction instead of forming 0/0 */
          double bk = gkgk[b] != 0.0 ? rb[b] / gkgk[b] : 0.0;
          double *p = pk + (int64_t)b * n, *q = pgk + (int64_t)b * n;
          for (int64_t i = 0; i < n; ++i) p[i] = q[i] + bk * p[i];
        }
      }
      for (int b = 0; b < nn; ++b) gkgk[b] = rb[b];
      for (size_t i = 0; i < (size_t)n * nn; ++i) hpk[i] = 0.0;
      orc_hpsi(h, pk, hpk, n, nn);
      for (int b = 0; b < nn; ++b) {
        const double *x = u + (int64_t)(ns + b) * ld;
        double *p = pk + (int64_t)b * n, *hx = hxk + (int64_t)b * n,
               *hp = hpk + (int64_t)b * n, *g = gk + (int64_t)b * n;
        double w1 = orc_dot(x, x, n, dv), w2 = orc_dot(p, x, n, dv),
               w3 = orc_dot(p, p, n, dv), w4 = orc_dot(x, hx, n, dv),
               w5 = orc_dot(p, hx, n, dv), w6 = orc_dot(p, hp, n, dv);
        double bm[4] = {w1, w2, w2, w3}, um[4] = {w4, w5, w5, w6}, W[2];
        /* info is ignored at :316; here B not positive definite (p zero or
         * parallel to x) keeps the band: u = (1, 0), E unchanged */
        if (orc_sygv2(um, bm, W) != 0 || !isfinite(um[0]) || !isfinite(um[2]) || !isfinite(W[0]) ||
            !isfinite(W[1])) {
          um[0] = um[2] = 1.0;
          um[1] = um[3] = 0.0;
          W[0] = W[1] = E[b];
        }
        if (fabs(W[0] - E[b]) > 1e-1 && fabs(W[1] - E[b]) <= 1e-1) {
          um[0] = um[2];
          um[1] = um[3];
          W[0] = W[1];
        }
        if (um[0] < 0.0) {
          um[0] = -um[0];
          um[1] = -um[1];
        }
        ut3[2 * b] = um[0];
        ut3[2 * b + 1] = um[1];
        E1[b] = E[b];
        E[b] = W[0];
        for (int64_t i = 0; i < n; ++i) hx[i] = um[0] * hx[i] + um[1] * hp[i];
        for (int64_t i = 0; i < n; ++i)
          g[i] = -2.0 * (hx[i] - W[0] * (um[0] * x[i] + um[1] * p[i]));
        rb[b] = orc_dot(g, g, n, dv);
      }
      for (int b = 0; b < nn; ++b) {
        if (!isfinite(rb[b]) || rb[b] / res[ns + b] > 1.0e8) {
          E[b] = E1[b];
          continue;
        }
        double *x = u + (int64_t)(ns + b) * ld, *p = pk + (int64_t)b * n;
        for (int64_t i = 0; i < n; ++i) x[i] = ut3[2 * b] * x[i] + ut3[2 * b + 1] * p[i];
      }
    }
    for (int b = 0; b < nn; ++b) esp[ns + b] = E[b];
  }
  free(hxk);
  free(hpk);
  free(gk);
  free(pgk);
  free(pk);
  free(E);
  free(E1);
  free(rb);
  free(gkgk);
  free(ut3);
  return 0;
}

/* d_esp_calc: esp_calc_module.f90:17-68.  The reference's dV there is
 * 1/sum|psi_1|^2 taken at first use (:124-131), which equals the grid dV
 * for a dV-normalised psi_1; the grid dV is used here. */
void orc_esp_calc(const orc_ham *h, const double *u, int64_t ld, int mb,
                  double *esp) {
  const int64_t n = h->N;
  double *hw = (double *)malloc(sizeof(double) * n);
  for (int m = 0; m < mb; ++m) {
    const double *w = u + (int64_t)m * ld;
    orc_hpsi(h, w, hw, n, 1);
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += w[i] * hw[i];
    esp[m] = s * h->dv;
  }
  free(hw);
}
