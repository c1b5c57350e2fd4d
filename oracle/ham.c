/*
 * ham.c -- oracle: FD coefficients, kinetic+local stencil, nonlocal operator,
 * hpsi.  TEST INFRASTRUCTURE (see oracle.h).
 */
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* fd_module.f90:18-54 -- central-difference Laplacian weights by the
 * reference's Lagrange-product loop (integer products, then doubles). */
void orc_fd_lap(int md, double *lap) {
  double s = 0.0;
  for (int m = 0; m < 2 * md + 1; ++m) lap[m] = 0.0;
  for (int i = -md; i <= md; ++i) {
    if (i == 0) continue;
    for (int j = -md; j <= md; ++j) {
      if (j == i || j == 0) continue;
      s += 1.0 / (double)(i * j);
    }
  }
  lap[md] = s;
  for (int j = 1; j <= md; ++j) {
    double t = 1.0;
    for (int i = -md; i <= md; ++i) {
      if (i == j) continue;
      t *= (double)(j - i);
    }
    s = 0.0;
    for (int k = -md; k <= md; ++k) {
      if (k == j || k == 0) continue;
      double s0 = 1.0;
      for (int i = -md; i <= md; ++i) {
        if (i == k || i == j || i == 0) continue;
        s0 *= (double)(-i);
      }
      s += s0;
    }
    lap[md + j] = 2.0 * s / t;
    lap[md - j] = lap[md + j];
  }
}

/* Reciprocal vectors with a_i . b_j = 2 pi delta_ij (bb[3*j + r]). */
static void recip(const double *aa, double *bb) {
  const double *a1 = aa, *a2 = aa + 3, *a3 = aa + 6;
  double c[3][3];
  c[0][0] = a2[1] * a3[2] - a2[2] * a3[1];
  c[0][1] = a2[2] * a3[0] - a2[0] * a3[2];
  c[0][2] = a2[0] * a3[1] - a2[1] * a3[0];
  c[1][0] = a3[1] * a1[2] - a3[2] * a1[1];
  c[1][1] = a3[2] * a1[0] - a3[0] * a1[2];
  c[1][2] = a3[0] * a1[1] - a3[1] * a1[0];
  c[2][0] = a1[1] * a2[2] - a1[2] * a2[1];
  c[2][1] = a1[2] * a2[0] - a1[0] * a2[2];
  c[2][2] = a1[0] * a2[1] - a1[1] * a2[0];
  double vol = a1[0] * c[0][0] + a1[1] * c[0][1] + a1[2] * c[0][2];
  double f = 2.0 * acos(-1.0) / vol;
  for (int j = 0; j < 3; ++j)
    for (int r = 0; r < 3; ++r) bb[3 * j + r] = f * c[j][r];
}

double orc_det3(const double *aa) {
  const double *a1 = aa, *a2 = aa + 3, *a3 = aa + 6;
  return a1[0] * (a2[1] * a3[2] - a2[2] * a3[1]) -
         a1[1] * (a2[0] * a3[2] - a2[2] * a3[0]) +
         a1[2] * (a2[0] * a3[1] - a2[1] * a3[0]);
}

/* kinetic_module.f90:92-101 with get_ggg_kinetic :196-211 */
int orc_kinetic_coef(const double *aa, const int *ngrid, int md,
                     double *coef_lap0, double *coef_lap, double *ggg) {
  double bb[9], lap[2 * 16 + 1], a[3], hg[3], c[3];
  if (md < 1 || md > 16) return 2;
  recip(aa, bb);
  const double cst = 1.0 / (4.0 * acos(-1.0) * acos(-1.0));
  for (int i = 0; i < 3; ++i)
    a[i] = sqrt(aa[3 * i] * aa[3 * i] + aa[3 * i + 1] * aa[3 * i + 1] +
                aa[3 * i + 2] * aa[3 * i + 2]);
  double g[6];
  for (int i = 0; i < 3; ++i) {
    double s = 0;
    for (int r = 0; r < 3; ++r) s += bb[3 * i + r] * bb[3 * i + r];
    g[i] = a[i] * a[i] * s * cst;
  }
  int pr[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  for (int q = 0; q < 3; ++q) {
    int i = pr[q][0], j = pr[q][1];
    double s = 0;
    for (int r = 0; r < 3; ++r) s += bb[3 * i + r] * bb[3 * j + r];
    g[3 + q] = a[i] * a[j] * s * cst;
  }
  if (ggg)
    for (int i = 0; i < 6; ++i) ggg[i] = g[i];
  orc_fd_lap(md, lap);
  for (int i = 0; i < 3; ++i) {
    hg[i] = a[i] / ngrid[i];
    c[i] = -0.5 * g[i] / (hg[i] * hg[i]);
  }
  *coef_lap0 = lap[md] * (c[0] + c[1] + c[2]);
  for (int n = 1; n <= md; ++n)
    for (int ax = 0; ax < 3; ++ax) coef_lap[ax * md + n - 1] = lap[md + n] * c[ax];
  /* cross terms (non-orthogonal cells) are outside this restatement */
  if (fabs(g[3]) > 1e-12 || fabs(g[4]) > 1e-12 || fabs(g[5]) > 1e-12) return 1;
  return 0;
}

orc_ham *orc_ham_create(int n1, int n2, int n3, const double *aa, int md) {
  orc_ham *h = (orc_ham *)calloc(1, sizeof(orc_ham));
  h->n[0] = n1;
  h->n[1] = n2;
  h->n[2] = n3;
  h->N = (int64_t)n1 * n2 * n3;
  h->md = md;
  memcpy(h->aa, aa, 9 * sizeof(double));
  h->coef_lap = (double *)calloc(3 * md, sizeof(double));
  if (orc_kinetic_coef(aa, h->n, md, &h->coef_lap0, h->coef_lap, h->ggg)) {
    free(h->coef_lap);
    free(h);
    return NULL;
  }
  for (int i = 0; i < 3; ++i)
    h->hgrid[i] = sqrt(aa[3 * i] * aa[3 * i] + aa[3 * i + 1] * aa[3 * i + 1] +
                       aa[3 * i + 2] * aa[3 * i + 2]) /
                  h->n[i];
  h->dv = fabs(orc_det3(aa)) / (double)h->N;
  h->vloc = (double *)calloc(h->N, sizeof(double));
  return h;
}

void orc_ham_free_proj(orc_ham *h) {
  free(h->mjj);
  free(h->jjp);
  free(h->uvk);
  free(h->iuv);
  h->mjj = h->jjp = h->iuv = NULL;
  h->uvk = NULL;
  h->nzlma = h->maxmjj = 0;
}

void orc_ham_destroy(orc_ham *h) {
  if (!h) return;
  orc_ham_free_proj(h);
  free(h->vloc);
  free(h->coef_lap);
  free(h);
}

double orc_ham_dv(const orc_ham *h) { return h->dv; }
void orc_ham_hgrid(const orc_ham *h, double *hg) {
  for (int i = 0; i < 3; ++i) hg[i] = h->hgrid[i];
}

void orc_ham_set_vloc(orc_ham *h, const double *vloc) {
  memcpy(h->vloc, vloc, h->N * sizeof(double));
}

int orc_ham_set_proj(orc_ham *h, int nzlma, int maxmjj, const int *mjj,
                     const int *jjp, const double *uvk, const int *iuv) {
  orc_ham_free_proj(h);
  h->nzlma = nzlma;
  h->maxmjj = maxmjj;
  size_t tot = (size_t)nzlma * (size_t)(maxmjj > 0 ? maxmjj : 1);
  h->mjj = (int *)calloc(nzlma > 0 ? nzlma : 1, sizeof(int));
  h->iuv = (int *)calloc(nzlma > 0 ? nzlma : 1, sizeof(int));
  h->jjp = (int *)calloc(tot, sizeof(int));
  h->uvk = (double *)calloc(tot, sizeof(double));
  if (nzlma > 0) {
    memcpy(h->mjj, mjj, nzlma * sizeof(int));
    memcpy(h->iuv, iuv, nzlma * sizeof(int));
    memcpy(h->jjp, jjp, tot * sizeof(int));
    memcpy(h->uvk, uvk, tot * sizeof(double));
  }
  return 0;
}

void orc_ham_proj_sizes(const orc_ham *h, int *nzlma, int *maxmjj) {
  *nzlma = h->nzlma;
  *maxmjj = h->maxmjj;
}

void orc_ham_proj_get(const orc_ham *h, int *mjj, int *jjp, double *uvk,
                      int *iuv) {
  size_t tot = (size_t)h->nzlma * (size_t)h->maxmjj;
  if (h->nzlma <= 0) return;
  memcpy(mjj, h->mjj, h->nzlma * sizeof(int));
  memcpy(iuv, h->iuv, h->nzlma * sizeof(int));
  memcpy(jjp, h->jjp, tot * sizeof(int));
  memcpy(uvk, h->uvk, tot * sizeof(double));
}

static inline int wrap(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

/* d_kinetic_sol_module.f90:125-139 copy into the halo array www and the
 * periodic halo fill bcset_3 (bc/bcset_3_module.f90:15-151, self-exchange
 * at one rank), then the Md=6 fused loop :244-292 (or the general-Md loop
 * :294-322).  Summation order follows the reference expression. */
void orc_op_kinetic(const orc_ham *h, const double *psi, double *hpsi,
                    int64_t ld, int nb) {
  const int n1 = h->n[0], n2 = h->n[1], n3 = h->n[2], md = h->md;
  const double c = h->coef_lap0; /* + const_k2(k) = 0 at Gamma */
  const double *cl = h->coef_lap;
  const int64_t n12 = (int64_t)n1 * n2;
  for (int ib = 0; ib < nb; ++ib) {
    const double *w = psi + (int64_t)ib * ld;
    double *o = hpsi + (int64_t)ib * ld;
#pragma omp parallel for schedule(static)
    for (int i3 = 0; i3 < n3; ++i3) {
      for (int i2 = 0; i2 < n2; ++i2) {
        for (int i1 = 0; i1 < n1; ++i1) {
          int64_t j = i1 + (int64_t)i2 * n1 + i3 * n12;
#define WX(d) w[wrap(i1 + (d), n1) + (int64_t)i2 * n1 + i3 * n12]
#define WY(d) w[i1 + (int64_t)wrap(i2 + (d), n2) * n1 + i3 * n12]
#define WZ(d) w[i1 + (int64_t)i2 * n1 + (int64_t)wrap(i3 + (d), n3) * n12]
          if (md == 6) {
            const double *c1 = cl, *c2 = cl + 6, *c3 = cl + 12;
            double t = +c1[5] * WX(-6) + c1[4] * WX(-5) + c1[3] * WX(-4) +
                       c1[2] * WX(-3) + c1[1] * WX(-2) + c1[0] * WX(-1) +
                       (c + h->vloc[j]) * w[j] + c1[0] * WX(1) + c1[1] * WX(2) +
                       c1[2] * WX(3) + c1[3] * WX(4) + c1[4] * WX(5) +
                       c1[5] * WX(6) + c2[5] * WY(-6) + c2[4] * WY(-5) +
                       c2[3] * WY(-4) + c2[2] * WY(-3) + c2[1] * WY(-2) +
                       c2[0] * WY(-1) + c2[0] * WY(1) + c2[1] * WY(2) +
                       c2[2] * WY(3) + c2[3] * WY(4) + c2[4] * WY(5) +
                       c2[5] * WY(6) + c3[5] * WZ(-6) + c3[4] * WZ(-5) +
                       c3[3] * WZ(-4) + c3[2] * WZ(-3) + c3[1] * WZ(-2) +
                       c3[0] * WZ(-1) + c3[0] * WZ(1) + c3[1] * WZ(2) +
                       c3[2] * WZ(3) + c3[3] * WZ(4) + c3[4] * WZ(5) +
                       c3[5] * WZ(6);
            o[j] = t;
          } else {
            double t = (c + h->vloc[j]) * w[j];
            for (int m = 1; m <= md; ++m) {
              t = t + cl[m - 1] * (WX(-m) + WX(m)) +
                  cl[md + m - 1] * (WY(-m) + WY(m)) +
                  cl[2 * md + m - 1] * (WZ(-m) + WZ(m));
            }
            o[j] = t;
          }
#undef WX
#undef WY
#undef WZ
        }
      }
    }
  }
}

/* d_ps_nloc2_op_module.f90:182-191 (gather-dot, uVunk = iuV*dV*sum) and
 * :274-281 (scatter-add); the neighbour exchange :195-269 is empty at one
 * rank. */
void orc_op_nonlocal(const orc_ham *h, const double *psi, double *hpsi,
                     int64_t ld, int nb) {
  if (h->nzlma <= 0) return;
  const int nz = h->nzlma, mx = h->maxmjj;
  double *uvunk = (double *)malloc(sizeof(double) * (size_t)nz * nb);
#pragma omp parallel for collapse(2) schedule(static)
  for (int ib = 0; ib < nb; ++ib) {
    for (int lma = 0; lma < nz; ++lma) {
      const double *t = psi + (int64_t)ib * ld;
      double s = 0.0;
      for (int j = 0; j < h->mjj[lma]; ++j)
        s = s + h->uvk[(size_t)lma * mx + j] * t[h->jjp[(size_t)lma * mx + j]];
      uvunk[lma + (size_t)ib * nz] = h->iuv[lma] * h->dv * s;
    }
  }
  /* scatter: per band, lma in order (the reference's thread split by grid
   * range keeps this order per point) */
#pragma omp parallel for schedule(static)
  for (int ib = 0; ib < nb; ++ib) {
    double *o = hpsi + (int64_t)ib * ld;
    for (int lma = 0; lma < nz; ++lma) {
      double u = uvunk[lma + (size_t)ib * nz];
      for (int j = 0; j < h->mjj[lma]; ++j)
        o[h->jjp[(size_t)lma * mx + j]] += h->uvk[(size_t)lma * mx + j] * u;
    }
  }
  free(uvunk);
}

/* hamiltonian_module.f90:118-144 (kin_select = 0; op_fock is a no-op for
 * LDA, xc/fock_module.f90:64) */
void orc_hpsi(const orc_ham *h, const double *psi, double *hpsi, int64_t ld,
              int nb) {
  orc_op_kinetic(h, psi, hpsi, ld, nb);
  orc_op_nonlocal(h, psi, hpsi, ld, nb);
}

/* ext0/dotp.f90:1-27.  Fixed-order sum: partials over fixed 4096-point
 * chunks, added in chunk order, so the result does not depend on the thread
 * count or on the order in which threads finish (an OpenMP reduction's
 * combining order is unspecified, and the CG's rounding-level branches then
 * made repeated runs differ). */
double orc_dot(const double *a, const double *b, int64_t n, double alpha) {
  enum { kChunk = 4096 };
  const int64_t nc = (n + kChunk - 1) / kChunk;
  if (nc <= 0) return 0.0;
  double *part = (double *)malloc(sizeof(double) * (size_t)nc);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nc; ++c) {
    const int64_t i1 = c * kChunk, i2 = i1 + kChunk < n ? i1 + kChunk : n;
    double s = 0.0;
    for (int64_t i = i1; i < i2; ++i) s += a[i] * b[i];
    part[c] = s;
  }
  double tmp = 0.0;
  for (int64_t c = 0; c < nc; ++c) tmp += part[c];
  free(part);
  return tmp * alpha;
}

void orc_dense_h(const orc_ham *h, double *H) {
  const int64_t N = h->N;
  double *e = (double *)calloc(N, sizeof(double));
  for (int64_t k = 0; k < N; ++k) {
    e[k] = 1.0;
    orc_hpsi(h, e, H + k * N, N, 1);
    e[k] = 0.0;
  }
  free(e);
}
