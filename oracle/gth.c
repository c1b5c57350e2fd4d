/*
 * gth.c -- oracle: GTH separable projector tables (JJP, uVk, MJJ, iuV) for
 * the periodic real-space grid.  TEST INFRASTRUCTURE (see oracle.h).
 */
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double ORC_PI = 3.141592653589793238;

/* Fortran nint: round half away from zero */
static int f_nint(double x) { return (int)(x >= 0 ? floor(x + 0.5) : -floor(-x + 0.5)); }

static double ipow(double r, int e) {
  double v = 1.0;
  for (int i = 0; i < e; ++i) v *= r;
  return v;
}

/* sqrt(pi) * prod_{i=1}^{L+2n-1} (i - 1/2) = Gamma(L + 2n - 1/2),
 * ps_nloc_gth_module.f90:63-66 */
static double gth_gamma(int L, int n) {
  double g = sqrt(acos(-1.0));
  for (int i = 1; i <= L + 2 * n - 1; ++i) g *= (i - 0.5);
  return g;
}

/* init_ps_nloc_gth: ps_nloc_gth_module.f90:137-180 -- first r on the
 * dr = 0.01 mesh (r <= 30) where the unscaled radial value drops below
 * 1e-10. */
double orc_gth_cutoff(int L, int n, double rl) {
  const double rmax = 30.0, ep = 1e-10;
  const int mmr = 3000;
  const double dr = rmax / mmr;
  double const1 = sqrt(2.0) / (pow(rl, L + 2 * n - 0.5) * sqrt(gth_gamma(L, n)));
  double const2 = 0.5 / (rl * rl);
  for (int i = 1; i <= mmr; ++i) {
    double r = i * dr;
    double v = const1 * ipow(r, L + 2 * n - 2) * exp(-r * r * const2);
    if (fabs(v) < ep) return r;
  }
  return 0.0;
}

/* Plm: ylm_module.f90:133-175 */
static double plm(int L, int M, double xi) {
  double x = xi;
  if (fabs(fabs(x) - 1.0) < 1e-13) x = (x >= 0 ? 1.0 : -1.0);
  double pmm = 1.0;
  if (M > 0) {
    double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double fact = 1.0;
    for (int i = 1; i <= M; ++i) {
      pmm = -pmm * fact * somx2;
      fact += 2.0;
    }
  }
  if (L == M) return pmm;
  double pmmp1 = x * (2.0 * M + 1.0) * pmm;
  if (L == M + 1) return pmmp1;
  double pll = 0.0;
  for (int ll = M + 2; ll <= L; ++ll) {
    pll = (x * (2.0 * ll - 1.0) * pmmp1 - (ll + M - 1.0) * pmm) / (ll - M);
    pmm = pmmp1;
    pmmp1 = pll;
  }
  return pll;
}

/* Ylm: ylm_module.f90:11-73 (real harmonics) */
double orc_ylm(double x, double y, double z, int L, int m_in) {
  int M = m_in < 0 ? -m_in : m_in;
  const double pi4 = 4.0 * ORC_PI;
  if (L == 0) return 1.0 / sqrt(pi4);
  if (x == 0.0 && y == 0.0 && z == 0.0) return 0.0;
  double c = 1.0;
  for (int k = L - M + 1; k <= L + M; ++k) c *= (double)k;
  double clm = sqrt((2.0 * L + 1.0) / (c * pi4));
  double r = sqrt(x * x + y * y + z * z);
  double ct = z / r, phi;
  if (fabs(x) < 1e-10) {
    phi = 0.5 * ORC_PI * (y >= 0 ? 1.0 : -1.0);
  } else {
    phi = atan(y / x);
    if (x < 0) phi += ORC_PI;
  }
  if (m_in > 0) return sqrt(2.0) * clm * plm(L, M, ct) * cos(phi * M);
  if (m_in == 0) return clm * plm(L, M, ct);
  if (M % 2 == 0) return -sqrt(2.0) * clm * plm(L, M, ct) * sin(phi * M);
  return sqrt(2.0) * clm * plm(L, M, ct) * sin(phi * M);
}

typedef struct {
  int i0[3]; /* wrapped grid index */
  int k[3];  /* image shift */
  double v;  /* radial value (uV_tmp) */
} gpt;

static int cmp_int_pair(const void *a, const void *b) {
  const int *x = (const int *)a, *y = (const int *)b;
  return (x[0] > y[0]) - (x[0] < y[0]);
}

int orc_ham_build_gth(orc_ham *h, int natom, const double *frac,
                      const int *kind, int nkind, const int *norb,
                      const int *lo, const int *no, const double *rps0,
                      const double *hnl) {
  const int *ng = h->n;
  const double *aa = h->aa;
  double rcut[64 * ORC_MAXORB];
  if (nkind > 64) return 3;
  /* cutoff radii (init_ps_nloc_gth) */
  double rmaxc = 0.0;
  for (int ik = 0; ik < nkind; ++ik)
    for (int io = 0; io < norb[ik]; ++io) {
      int q = ik * ORC_MAXORB + io;
      rcut[q] = orc_gth_cutoff(lo[q], no[q], rps0[q]);
      if (rcut[q] > rmaxc) rmaxc = rcut[q];
    }
  /* minimal box: ps_nloc2_module.f90:158-159, minimal_box_module.f90:14-165
   * -- every offset with |aa.(i/N)|^2 <= Rc^2, i3 outer, i1 inner. */
  double hmax = h->hgrid[0];
  if (h->hgrid[1] > hmax) hmax = h->hgrid[1];
  if (h->hgrid[2] > hmax) hmax = h->hgrid[2];
  const double Rc = rmaxc + hmax + 1e-8, R0 = Rc * Rc;
  double bb[9];
  {
    /* |b_k| N_k / 2pi bounds the fractional half-extent of the sphere */
    const double *a1 = aa, *a2 = aa + 3, *a3 = aa + 6;
    double vol = orc_det3(aa);
    double c[3][3] = {{a2[1] * a3[2] - a2[2] * a3[1], a2[2] * a3[0] - a2[0] * a3[2],
                       a2[0] * a3[1] - a2[1] * a3[0]},
                      {a3[1] * a1[2] - a3[2] * a1[1], a3[2] * a1[0] - a3[0] * a1[2],
                       a3[0] * a1[1] - a3[1] * a1[0]},
                      {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2],
                       a1[0] * a2[1] - a1[1] * a2[0]}};
    for (int j = 0; j < 3; ++j)
      for (int r = 0; r < 3; ++r) bb[3 * j + r] = c[j][r] / vol;
  }
  int mm[3];
  for (int k = 0; k < 3; ++k) {
    double bn = sqrt(bb[3 * k] * bb[3 * k] + bb[3 * k + 1] * bb[3 * k + 1] +
                     bb[3 * k + 2] * bb[3 * k + 2]);
    mm[k] = (int)ceil(Rc * bn * ng[k]) + 1;
  }
  int nbox = 0, cap = (2 * mm[0] + 1) * (2 * mm[1] + 1) * (2 * mm[2] + 1);
  int *box = (int *)malloc(sizeof(int) * 3 * cap);
  for (int i3 = -mm[2]; i3 <= mm[2]; ++i3)
    for (int i2 = -mm[1]; i2 <= mm[1]; ++i2)
      for (int i1 = -mm[0]; i1 <= mm[0]; ++i1) {
        double d1 = i1 * (1.0 / ng[0]), d2 = i2 * (1.0 / ng[1]), d3 = i3 * (1.0 / ng[2]);
        double x = d1 * aa[0] + d2 * aa[3] + d3 * aa[6];
        double y = d1 * aa[1] + d2 * aa[4] + d3 * aa[7];
        double z = d1 * aa[2] + d2 * aa[5] + d3 * aa[8];
        if (x * x + y * y + z * z <= R0) {
          box[3 * nbox] = i1;
          box[3 * nbox + 1] = i2;
          box[3 * nbox + 2] = i3;
          ++nbox;
        }
      }

  /* count lma and collect sphere points per (atom, orbital):
   * prep_ps_nloc_gth, ps_nloc_gth_module.f90:19-134 */
  int nsph = 0;
  for (int a = 0; a < natom; ++a) nsph += norb[kind[a]];
  gpt **pts = (gpt **)calloc(nsph, sizeof(gpt *));
  int *npts = (int *)calloc(nsph, sizeof(int));
  int q = 0;
  const double c1 = 1.0 / ng[0], c2 = 1.0 / ng[1], c3 = 1.0 / ng[2];
  for (int a = 0; a < natom; ++a) {
    const double *fa = frac + 3 * a;
    double Rx = aa[0] * fa[0] + aa[3] * fa[1] + aa[6] * fa[2];
    double Ry = aa[1] * fa[0] + aa[4] * fa[1] + aa[7] * fa[2];
    double Rz = aa[2] * fa[0] + aa[5] * fa[1] + aa[8] * fa[2];
    int ic1 = f_nint(fa[0] * ng[0]), ic2 = f_nint(fa[1] * ng[1]), ic3 = f_nint(fa[2] * ng[2]);
    int ik = kind[a];
    for (int io = 0; io < norb[ik]; ++io, ++q) {
      int qq = ik * ORC_MAXORB + io;
      int L = lo[qq], n = no[qq];
      double rl = rps0[qq];
      double rps2 = rcut[qq] * rcut[qq];
      double const1 = sqrt(fabs(hnl[qq])) * sqrt(2.0) /
                      (pow(rl, L + 2 * n - 0.5) * sqrt(gth_gamma(L, n)));
      double const2 = 0.5 / (rl * rl);
      pts[q] = (gpt *)malloc(sizeof(gpt) * (nbox > 0 ? nbox : 1));
      int j = 0;
      for (int i = 0; i < nbox; ++i) {
        int id[3] = {ic1 + box[3 * i], ic2 + box[3 * i + 1], ic3 + box[3 * i + 2]};
        int kk[3], i0[3];
        for (int t = 0; t < 3; ++t) {
          /* floor division as in :84-86 */
          kk[t] = id[t] / ng[t];
          if (id[t] < 0) kk[t] = (id[t] + 1) / ng[t] - 1;
          i0[t] = id[t] - kk[t] * ng[t];
        }
        double d1 = id[0] * c1, d2 = id[1] * c2, d3 = id[2] * c3;
        double x = aa[0] * d1 + aa[3] * d2 + aa[6] * d3 - Rx;
        double y = aa[1] * d1 + aa[4] * d2 + aa[7] * d3 - Ry;
        double z = aa[2] * d1 + aa[5] * d2 + aa[8] * d3 - Rz;
        double r2 = x * x + y * y + z * z;
        if (r2 > rps2 + 1e-10) continue;
        double r = sqrt(r2);
        gpt *p = &pts[q][j++];
        for (int t = 0; t < 3; ++t) {
          p->i0[t] = i0[t];
          p->k[t] = kk[t];
        }
        p->v = const1 * ipow(r, L + 2 * n - 2) * exp(-r2 * const2);
      }
      npts[q] = j;
    }
  }
  free(box);

  /* lma enumeration: atom, orbital, m = -L..L for spheres with points
   * (ps_nloc2_module.f90:482-497, 560) */
  int nz = 0;
  q = 0;
  for (int a = 0; a < natom; ++a)
    for (int io = 0; io < norb[kind[a]]; ++io, ++q)
      if (npts[q] > 0) nz += 2 * lo[kind[a] * ORC_MAXORB + io] + 1;

  /* capacity per lma while filling: the largest sphere */
  int maxmjj = 0;
  for (q = 0; q < nsph; ++q)
    if (npts[q] > maxmjj) maxmjj = npts[q];

  orc_ham_free_proj(h);
  h->nzlma = nz;
  h->maxmjj = maxmjj;
  size_t tot = (size_t)(nz > 0 ? nz : 1) * (size_t)(maxmjj > 0 ? maxmjj : 1);
  h->mjj = (int *)calloc(nz > 0 ? nz : 1, sizeof(int));
  h->iuv = (int *)calloc(nz > 0 ? nz : 1, sizeof(int));
  h->jjp = (int *)calloc(tot, sizeof(int));
  h->uvk = (double *)calloc(tot, sizeof(double));
  int *mark = (int *)calloc(h->N, sizeof(int)); /* icheck_tmp4 */
  int *pairs = (int *)malloc(sizeof(int) * 2 * (maxmjj > 0 ? maxmjj : 1));
  double *tmpv = (double *)malloc(sizeof(double) * (maxmjj > 0 ? maxmjj : 1));

  int lma = 0;
  q = 0;
  for (int a = 0; a < natom; ++a) {
    const double *fa = frac + 3 * a;
    double Rx = aa[0] * fa[0] + aa[3] * fa[1] + aa[6] * fa[2];
    double Ry = aa[1] * fa[0] + aa[4] * fa[1] + aa[7] * fa[2];
    double Rz = aa[2] * fa[0] + aa[5] * fa[1] + aa[8] * fa[2];
    int ik = kind[a];
    for (int io = 0; io < norb[ik]; ++io, ++q) {
      if (npts[q] == 0) continue;
      int qq = ik * ORC_MAXORB + io;
      int L = lo[qq];
      for (int m = -L; m <= L; ++m, ++lma) {
        h->iuv[lma] = hnl[qq] < 0.0 ? -1 : 1;
        /* uV(j,lma) = uV_tmp * Ylm at the unwrapped position
         * (ps_nloc2_module.f90:770-789), then fold duplicates keeping the
         * first |uV| > 1e-14 occurrence (:800-818, :1113-1127) */
        int j = 0;
        for (int i = 0; i < npts[q]; ++i) {
          const gpt *p = &pts[q][i];
          double d1 = c1 * p->i0[0] + p->k[0], d2 = c2 * p->i0[1] + p->k[1],
                 d3 = c3 * p->i0[2] + p->k[2];
          double x = aa[0] * d1 + aa[3] * d2 + aa[6] * d3 - Rx;
          double y = aa[1] * d1 + aa[4] * d2 + aa[7] * d3 - Ry;
          double z = aa[2] * d1 + aa[5] * d2 + aa[8] * d3 - Rz;
          double u3 = p->v * orc_ylm(x, y, z, L, m);
          int64_t lin = p->i0[0] + (int64_t)p->i0[1] * ng[0] +
                        (int64_t)p->i0[2] * ng[0] * ng[1];
          int j3 = mark[lin];
          if (j3 == 0 && fabs(u3) > 1e-14) {
            mark[lin] = ++j;
            pairs[2 * (j - 1)] = (int)lin;
            tmpv[j - 1] = u3;
          } else if (j3 != 0) {
            tmpv[j3 - 1] += u3;
          }
        }
        for (int t = 0; t < j; ++t) {
          mark[pairs[2 * t]] = 0;
          pairs[2 * t + 1] = t;
        }
        /* sort_index_sub: ascending JJP within the lma */
        qsort(pairs, j, 2 * sizeof(int), cmp_int_pair);
        h->mjj[lma] = j;
        for (int t = 0; t < j; ++t) {
          h->jjp[(size_t)lma * maxmjj + t] = pairs[2 * t];
          h->uvk[(size_t)lma * maxmjj + t] = tmpv[pairs[2 * t + 1]];
        }
      }
    }
  }
  /* JJP/uVk leading dimension is maxval(MJJ) after the |uV| <= 1e-14 drop
   * (d_allocate_ps_nloc2, ps_nloc2_variables.f90:153-161) */
  int mx = 0;
  for (int l = 0; l < nz; ++l)
    if (h->mjj[l] > mx) mx = h->mjj[l];
  if (mx < maxmjj && nz > 0) {
    for (int l = 1; l < nz; ++l) {
      memmove(h->jjp + (size_t)l * mx, h->jjp + (size_t)l * maxmjj, sizeof(int) * mx);
      memmove(h->uvk + (size_t)l * mx, h->uvk + (size_t)l * maxmjj, sizeof(double) * mx);
    }
  }
  h->maxmjj = mx;
  free(mark);
  free(pairs);
  free(tmpv);
  for (q = 0; q < nsph; ++q) free(pts[q]);
  free(pts);
  free(npts);
  return 0;
}
