/* oracle_internal.h -- private state of the oracle (test infrastructure). */
#ifndef RSDFT_ORACLE_INTERNAL_H
#define RSDFT_ORACLE_INTERNAL_H
#include "oracle.h"

struct orc_ham {
  int n[3];
  int64_t N;
  int md;
  double aa[9];
  double hgrid[3];
  double dv;
  double ggg[6];
  double coef_lap0;
  double *coef_lap; /* [axis*md + m-1] */
  double *vloc;     /* N */
  /* ps_nloc2_variables: MJJ(nzlma), JJP(maxMJJ,nzlma), d_uVk, iuV */
  int nzlma, maxmjj;
  int *mjj, *jjp, *iuv;
  double *uvk;
};

double orc_det3(const double *aa);
void orc_ham_free_proj(orc_ham *h);

#endif
