/*
 * dense.c -- oracle: small dense symmetric kernels standing in for the
 * LAPACK routines the reference links (DSYEV, DPOTRF, DTRTRI, DSYGV).
 * LAPACK is third-party and absent here; these are textbook methods
 * (cyclic Jacobi, Cholesky-Banachiewicz, back substitution) whose results
 * are pinned against numpy.linalg in tests/test_oracle_dense.py.
 * TEST INFRASTRUCTURE (see oracle.h).
 */
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (column-
 * major, upper triangle read, as DSYEV('V','U')).  On return a holds the
 * eigenvectors (column k <-> w[k]), w ascending. */
int orc_syev_jacobi(int n, double *a, double *w) {
  double *A = (double *)malloc(sizeof(double) * n * n);
  double *V = (double *)calloc((size_t)n * n, sizeof(double));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double v = i <= j ? a[i + (size_t)j * n] : a[j + (size_t)i * n];
      A[i + (size_t)j * n] = v;
    }
  for (int i = 0; i < n; ++i) V[i + (size_t)i * n] = 1.0;
  int info = 1;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double v = A[i + (size_t)j * n] * A[i + (size_t)j * n];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-30 * tot || off == 0.0) {
      info = 0;
      break;
    }
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p + (size_t)q * n];
        if (apq == 0.0) continue;
        double app = A[p + (size_t)p * n], aqq = A[q + (size_t)q * n];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) { /* columns p,q */
          double akp = A[k + (size_t)p * n], akq = A[k + (size_t)q * n];
          A[k + (size_t)p * n] = c * akp - s * akq;
          A[k + (size_t)q * n] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) { /* rows p,q */
          double apk = A[p + (size_t)k * n], aqk = A[q + (size_t)k * n];
          A[p + (size_t)k * n] = c * apk - s * aqk;
          A[q + (size_t)k * n] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k + (size_t)p * n], vkq = V[k + (size_t)q * n];
          V[k + (size_t)p * n] = c * vkp - s * vkq;
          V[k + (size_t)q * n] = s * vkp + c * vkq;
        }
      }
  }
  /* sort ascending (selection) */
  int *idx = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) {
    idx[i] = i;
    w[i] = A[i + (size_t)i * n];
  }
  for (int i = 0; i < n; ++i) {
    int m = i;
    for (int k = i + 1; k < n; ++k)
      if (w[idx[k]] < w[idx[m]]) m = k;
    int t = idx[i];
    idx[i] = idx[m];
    idx[m] = t;
  }
  double *ws = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) ws[i] = w[idx[i]];
  for (int k = 0; k < n; ++k) {
    w[k] = ws[k];
    memcpy(a + (size_t)k * n, V + (size_t)idx[k] * n, sizeof(double) * n);
  }
  free(ws);
  free(idx);
  free(A);
  free(V);
  return info;
}

/* S = R^T R with R upper triangular (column-major, upper read/written,
 * strict lower zeroed).  Returns k>0 if the leading k x k minor is not
 * positive definite. */
int orc_potrf_upper(int n, double *a) {
  for (int j = 0; j < n; ++j) {
    double s = a[j + (size_t)j * n];
    for (int k = 0; k < j; ++k) s -= a[k + (size_t)j * n] * a[k + (size_t)j * n];
    if (!(s > 0.0)) return j + 1;
    double rjj = sqrt(s);
    a[j + (size_t)j * n] = rjj;
    for (int i = j + 1; i < n; ++i) {
      double t = a[j + (size_t)i * n];
      for (int k = 0; k < j; ++k) t -= a[k + (size_t)j * n] * a[k + (size_t)i * n];
      a[j + (size_t)i * n] = t / rjj;
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) a[i + (size_t)j * n] = 0.0;
  return 0;
}

/* In-place inverse of an upper triangular matrix. */
int orc_trtri_upper(int n, double *a) {
  for (int j = 0; j < n; ++j) {
    if (a[j + (size_t)j * n] == 0.0) return j + 1;
    a[j + (size_t)j * n] = 1.0 / a[j + (size_t)j * n];
    double ajj = -a[j + (size_t)j * n];
    /* column j above the diagonal: x = inv(R11) * R12 (in place, rows in
     * increasing order so R12[k] for k >= i is still unmodified), then
     * scaled by -1/rjj */
    for (int i = 0; i < j; ++i) {
      double s = 0.0;
      for (int k = i; k < j; ++k) s += a[i + (size_t)k * n] * a[k + (size_t)j * n];
      a[i + (size_t)j * n] = s;
    }
    for (int i = 0; i < j; ++i) a[i + (size_t)j * n] *= ajj;
  }
  return 0;
}

/* 2x2 symmetric-definite pencil (A, B), DSYGV(itype=1) semantics:
 * B = U^T U, C = U^-T A U^-1, C = Y diag(w) Y^T, X = U^-1 Y (X^T B X = I).
 * a <- X, w ascending. */
int orc_sygv2(double *a, double *b, double *w) {
  double b11 = b[0], b12 = b[2], b22 = b[3];
  if (!(b11 > 0.0)) return 1;
  double u11 = sqrt(b11), u12 = b12 / u11, d = b22 - u12 * u12;
  if (!(d > 0.0)) return 2;
  double u22 = sqrt(d);
  /* Ui = U^-1 (upper): [1/u11, -u12/(u11 u22); 0, 1/u22] */
  double i11 = 1.0 / u11, i12 = -u12 / (u11 * u22), i22 = 1.0 / u22;
  double a11 = a[0], a12 = a[2], a22 = a[3];
  /* C = Ui^T A Ui */
  double t11 = a11 * i11, t12 = a11 * i12 + a12 * i22; /* row1 of A*Ui */
  double t21 = a12 * i11, t22 = a12 * i12 + a22 * i22; /* row2 of A*Ui */
  double c11 = i11 * t11;
  double c12 = i11 * t12;
  double c22 = i12 * t12 + i22 * t22;
  (void)t21;
  /* symmetric 2x2 eigen by a Jacobi rotation */
  double y11, y12, y21, y22, l1, l2;
  if (c12 == 0.0) {
    y11 = 1; y21 = 0; y12 = 0; y22 = 1; l1 = c11; l2 = c22;
  } else {
    double theta = (c22 - c11) / (2.0 * c12);
    double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
    l1 = c11 - t * c12;
    l2 = c22 + t * c12;
    y11 = c; y21 = -s; y12 = s; y22 = c;
  }
  if (l2 < l1) {
    double tl = l1; l1 = l2; l2 = tl;
    double t1 = y11, t2 = y21;
    y11 = y12; y21 = y22; y12 = t1; y22 = t2;
  }
  w[0] = l1;
  w[1] = l2;
  a[0] = i11 * y11 + i12 * y21;
  a[1] = i22 * y21;
  a[2] = i11 * y12 + i12 * y22;
  a[3] = i22 * y22;
  return 0;
}
