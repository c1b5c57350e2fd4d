/*
 * gs.c -- oracle: Gram-Schmidt orthonormalisation under <u,v> = dV u^T v.
 * TEST INFRASTRUCTURE (see oracle.h).
 */
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double *u;
  int64_t ld, n;
  double dv;
  int nblk1;
} gs_state;

/* 1-based band index -> column pointer */
#define COL(st, m) ((st)->u + ((int64_t)(m)-1) * (st)->ld)

static void normalize_col(gs_state *st, int m) {
  double *c = COL(st, m);
  double s = 0.0;
  for (int64_t i = 0; i < st->n; ++i) s += c[i] * c[i];
  s = 1.0 / sqrt(s * st->dv);
  for (int64_t i = 0; i < st->n; ++i) c[i] *= s;
}

/* gram_schmidt_sub: gram_schmidt_t_module.f90:155-306 (1-based bounds) */
static void gs_sub(gs_state *st, int mm1, int mm2, int nn1, int nn2, int mblk) {
  const int64_t n = st->n;
  for (int ms = mm1; ms <= mm2; ms += mblk) {
    int me = ms + mblk - 1 < mm2 ? ms + mblk - 1 : mm2;
    int mm = me - ms + 1;
    for (int ns = nn1; ns <= nn2; ns += mblk) {
      int ne = ns + mblk - 1 < nn2 ? ns + mblk - 1 : nn2;
      if (ne > me - 1) ne = me - 1;
      int nn = ne - ns + 1;
      if (nn <= 0) continue;
      if (ms >= ne + 1) {
        /* utmp2 = -dV U_ns^T U_ms (DGEMM :181-183), U_ms += U_ns utmp2 */
        double *ut = (double *)malloc(sizeof(double) * nn * mm);
#pragma omp parallel for collapse(2) schedule(static)
        for (int b = 0; b < mm; ++b)
          for (int a = 0; a < nn; ++a) {
            const double *x = COL(st, ns + a), *y = COL(st, ms + b);
            double s = 0.0;
            for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
            ut[a + b * nn] = -st->dv * s;
          }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i)
          for (int b = 0; b < mm; ++b) {
            double s = COL(st, ms + b)[i];
            for (int a = 0; a < nn; ++a) s += COL(st, ns + a)[i] * ut[a + b * nn];
            COL(st, ms + b)[i] = s;
          }
        free(ut);
        if (ms == ne + 1) normalize_col(st, ms);
      } else if (mm <= st->nblk1) {
        double ut[64];
        for (int m = ms; m <= me; ++m) {
          int nend = m - 1 < ne ? m - 1 : ne;
          int cnt = nend - ns + 1;
          if (cnt > 0) {
            /* DGEMV :241-243 then :249-251 */
            for (int a = 0; a < cnt; ++a) {
              const double *x = COL(st, ns + a), *y = COL(st, m);
              double s = 0.0;
              for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
              ut[a] = -st->dv * s;
            }
            double *y = COL(st, m);
            for (int64_t i = 0; i < n; ++i) {
              double s = y[i];
              for (int a = 0; a < cnt; ++a) s += COL(st, ns + a)[i] * ut[a];
              y[i] = s;
            }
          }
          if (m == 1 || (nend == m - 1 && m != ns)) normalize_col(st, m);
        }
      } else {
        int mblkh = mblk / 2 > st->nblk1 ? mblk / 2 : st->nblk1;
        gs_sub(st, ms, me, ns, ne, mblkh);
      }
    }
  }
}

/* gram_schmidt_t: gram_schmidt_t_module.f90:85-150 with np_band = 1,
 * NBLK = MB, NBLK1 = 4 -> one cycle, gram_schmidt_sub(1,MB,1,MB,MB). */
void orc_gram_schmidt_t(double *u, int64_t ld, int64_t n, int mb, double dv) {
  gs_state st = {u, ld, n, dv, 4};
  int nblk = mb;
  int ncycle = (mb - 1) / nblk + 1;
  for (int k1 = 1; k1 <= ncycle; ++k1) {
    int ns = nblk * (k1 - 1) + 1;
    int ne = ns + nblk - 1 < mb ? ns + nblk - 1 : mb;
    gs_sub(&st, ns, ne, ns, ne, nblk);
    if (ns <= mb - nblk) {
      for (int nbss = k1 + 1; nbss <= ncycle; ++nbss) {
        int ms = nblk * (nbss - 1) + 1;
        int me = ms + nblk - 1 < mb ? ms + nblk - 1 : mb;
        if (ms <= me) gs_sub(&st, ms, me, ns, ne, nblk);
      }
    }
  }
}

/* d_gram_schmidt_lusl: gram_schmidt_lusl_module.f90:241-328.
 * S = dV U^T U (upper, calc_dsyrk3), S = R^T R (PDPOTRF 'U'),
 * R^-1 (PDTRTRI), U <- U R^-1 (calc_dtrmm3). */
int orc_gram_schmidt_lusl(double *u, int64_t ld, int64_t n, int mb,
                          double dv) {
  double *S = (double *)calloc((size_t)mb * mb, sizeof(double));
#pragma omp parallel for schedule(dynamic)
  for (int j = 0; j < mb; ++j)
    for (int i = 0; i <= j; ++i) {
      const double *x = u + (int64_t)i * ld, *y = u + (int64_t)j * ld;
      double s = 0.0;
      for (int64_t k = 0; k < n; ++k) s += x[k] * y[k];
      S[i + (size_t)j * mb] = dv * s;
    }
  int info = orc_potrf_upper(mb, S);
  if (info) {
    free(S);
    return info;
  }
  orc_trtri_upper(mb, S);
  double *row = NULL;
#pragma omp parallel private(row)
  {
    row = (double *)malloc(sizeof(double) * mb);
#pragma omp for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
      for (int j = 0; j < mb; ++j) row[j] = u[k + (int64_t)j * ld];
      for (int j = 0; j < mb; ++j) {
        double s = 0.0;
        for (int i = 0; i <= j; ++i) s += row[i] * S[i + (size_t)j * mb];
        u[k + (int64_t)j * ld] = s;
      }
    }
    free(row);
  }
  free(S);
  return 0;
}
