// scf.hip -- the density/potential half of an SCF step on the device:
// GTH local potential (G-space form factor + hipFFT), density from the band
// store, FFT Hartree, LDA-PZ81, Pulay mixing of the density, and the
// energy integrals and the Ewald energy.  FP64 throughout.
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "solver.hpp"

namespace rsg {

struct ScfState {
  bool have_local = false;
  double eewald = 0.0, nelectron = 0.0;
  DevBuf rho, rho_out, vion, vh, vxc, zbuf, work, occ;
  // reciprocal-space cutoff of the density grid (ggrid_module.f90:55-91)
  double ecut = 0.0;
  double bb[9];
  hipfftHandle plan = 0;
  bool have_plan = false;
  // Pulay history (mixing_pulay_module.f90): pointer rings of mmix slots
  int mmix = 0, count = 0;
  double beta = 0.3;
  std::vector<DevBuf> xin, xou;
  DevBuf res;  // mmix residuals Xou_i - Xin_i, N doubles each
  // IMIX 30: Pulay on Fourier coefficients (complex histories), Kerker
  // weights wq (residual metric) and wq2 (update damping)
  int imix = 10;
  std::vector<DevBuf> gxin, gxou;
  DevBuf wq, wq2, gpart;
  // atom positions, kinds and species parameters of the local potential
  // (kept: a free would wait for the whole device)
  DevBuf dfrac, dkind, dsp;
  DevBuf ewpart;  // Ewald terms before their ordered sums
  // GTH local parameters per species
  std::vector<double> zps, rloc, cloc;
  std::vector<double> frac;
  std::vector<int> kind;
};

namespace {

constexpr int kT = 256;
const double kPi = 3.14159265358979323846;

ScfState *state(rsg_ctx *ctx) {
  if (!ctx->scf) ctx->scf = new ScfState();
  return ctx->scf;
}

struct GridG {
  int n1, n2, n3;
  double b[9];  // reciprocal vectors b[3*k + r]
  double ecut;
};

// Reduced frequency of FFT index i, or a sentinel outside -(n-1)/2..(n-1)/2
__device__ inline bool freq(int i, int n, int &m) {
  m = i <= (n - 1) / 2 ? i : i - n;
  return m >= -(n - 1) / 2 && m <= (n - 1) / 2;
}

__device__ inline bool gvec(const GridG &g, int64_t j, double &G2, int m[3]) {
  const int i1 = int(j % g.n1), i2 = int((j / g.n1) % g.n2), i3 = int(j / (int64_t(g.n1) * g.n2));
  if (!freq(i1, g.n1, m[0]) || !freq(i2, g.n2, m[1]) || !freq(i3, g.n3, m[2])) return false;
  double gx = g.b[0] * m[0] + g.b[3] * m[1] + g.b[6] * m[2];
  double gy = g.b[1] * m[0] + g.b[4] * m[1] + g.b[7] * m[2];
  double gz = g.b[2] * m[0] + g.b[5] * m[1] + g.b[8] * m[2];
  G2 = gx * gx + gy * gy + gz * gz;
  return G2 < g.ecut - 1e-14;
}

// V_loc(G) of one GTH species (ps_local_gth_module.f90:13-58), divided by
// the cell volume.
__device__ inline double gth_local_g(double G2, double zps, double rl, const double *c, double inv_vol) {
  const double twopi3 = sqrt(8.0 * kPi * kPi * kPi);
  if (G2 == 0.0)
    return inv_vol * (2.0 * kPi * zps * rl * rl + twopi3 * rl * rl * rl * (c[0] + 3.0 * c[1] + 15.0 * c[2] + 105.0 * c[3]));
  const double x2 = G2 * rl * rl, e = exp(-0.5 * x2);
  double v = -4.0 * kPi * zps * e / G2;
  v += twopi3 * rl * rl * rl * e *
       (c[0] + c[1] * (3.0 - x2) + c[2] * (15.0 - 10.0 * x2 + x2 * x2) +
        c[3] * (105.0 - 105.0 * x2 + 21.0 * x2 * x2 - x2 * x2 * x2));
  return inv_vol * v;
}

// vg(G) = sum_a V_s(a)(|G|) exp(-i G.R_a) inside the cutoff sphere
// (construct_ps_local_fft0, ps_local_fft0_module.f90:14-86)
__global__ void k_vion_g(GridG g, int natom, const double *__restrict__ frac, const int *__restrict__ kind,
                         const double *__restrict__ sp /* per species: zps, rloc, c1..c4 */, double inv_vol,
                         double2 *__restrict__ z) {
  const int64_t N = int64_t(g.n1) * g.n2 * g.n3;
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= N) return;
  double G2;
  int m[3];
  double2 acc = make_double2(0.0, 0.0);
  if (gvec(g, j, G2, m)) {
    for (int a = 0; a < natom; ++a) {
      const double *p = sp + 6 * kind[a];
      const double v = gth_local_g(G2, p[0], p[1], p + 2, inv_vol);
      const double ph = -2.0 * kPi * (m[0] * frac[3 * a] + m[1] * frac[3 * a + 1] + m[2] * frac[3 * a + 2]);
      double s, c;
      sincos(ph, &s, &c);
      acc.x += v * c;
      acc.y += v * s;
    }
  }
  z[j] = acc;
}

__global__ void k_real_part(const double2 *__restrict__ z, double *__restrict__ out, int64_t N) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < N) out[j] = z[j].x;
}

// dst = s * src (complex)
__global__ void k_zscale(const double2 *__restrict__ src, double2 *__restrict__ dst, double sc, int64_t N) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < N) dst[j] = make_double2(sc * src[j].x, sc * src[j].y);
}

__global__ void k_to_complex(const double *__restrict__ r, double2 *__restrict__ z, int64_t N) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < N) z[j] = make_double2(r[j], 0.0);
}

// 4 pi / (N G^2) inside the cutoff, G = 0 dropped (hartree_sol_fftw_module.f90:29-130)
__global__ void k_hartree_g(GridG g, double2 *__restrict__ z) {
  const int64_t N = int64_t(g.n1) * g.n2 * g.n3;
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= N) return;
  double G2;
  int m[3];
  double f = 0.0;
  if (gvec(g, j, G2, m) && !(m[0] == 0 && m[1] == 0 && m[2] == 0)) f = 4.0 * kPi / (double(N) * G2);
  z[j] = make_double2(z[j].x * f, z[j].y * f);
}

// rho = sum_n occ_n psi_n^2 (density_module.f90:120-151), accumulated in
// FP64 also from the FP32 band store
template <typename U>
__global__ void k_density(const U *__restrict__ u, int64_t N, int mb, const double *__restrict__ occ,
                          double *__restrict__ rho) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= N) return;
  double s = 0.0;
  for (int n = 0; n < mb; ++n) {
    const double o = occ[n];
    if (o != 0.0) {
      const double v = double(u[j + int64_t(n) * N]);
      s += o * v * v;
    }
  }
  rho[j] = s;
}

// LDA Perdew-Zunger 81, spin-unpolarised (xc/xc_ldapz81_module.f90:78-216).
// out: vxc, and exc_dens = rho*e_x + |rho|*e_c for the energy sum.
__global__ void k_ldapz81(const double *__restrict__ rho, double *__restrict__ vxc, double *__restrict__ ed,
                          int64_t N) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const double r = rho[j];
  double vx = 0.0, ex = 0.0;
  if (r > 0.0) {
    const double exd = -0.75 * cbrt(3.0 / kPi * r);
    ex = r * exd;
    vx = (4.0 / 3.0) * exd;
  }
  double vc = 0.0, ec = 0.0;
  const double t = fabs(r);  // rhoa + rhob with rhoa = rhob = |rho|/2
  if (t > 0.0) {
    const double gam = -0.1423, b1 = 1.0529, b2 = 0.3334;
    const double A = 0.0311, B = -0.048, C = 0.002, D = -0.0116;
    const double rs = cbrt(3.0 / (4.0 * kPi) / t);
    double ecd, mu;
    if (rs >= 1.0) {
      const double sq = sqrt(rs);
      ecd = gam / (1.0 + b1 * sq + b2 * rs);
      mu = ecd + ecd * (1.0 / 3.0) * (0.5 * b1 / sq + b2) / (1.0 / rs + b1 / sq + b2);
    } else {
      const double ln = log(rs);
      ecd = A * ln + B + C * rs * ln + D * rs;
      mu = ecd - (1.0 / 3.0) * (A + C * rs * (1.0 + ln) + rs * D);
    }
    ec = t * ecd;
    vc = mu;
  }
  vxc[j] = vx + vc;
  ed[j] = ex + ec;
}

__global__ void k_add3(const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
                       double *__restrict__ out, float *__restrict__ out32, int64_t N) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const double v = a[j] + b[j] + c[j];
  out[j] = v;
  out32[j] = float(v);
}

__global__ void k_axpby(double *__restrict__ y, const double *__restrict__ x, double a, double b, int64_t N) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < N) y[j] = a * x[j] + b * y[j];
}

__global__ void k_gauss_rho(GridG g, const double *__restrict__ aa, int natom, const double *__restrict__ frac,
                            const double *__restrict__ zat, double width, double *__restrict__ rho) {
  const int64_t N = int64_t(g.n1) * g.n2 * g.n3;
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const int i1 = int(j % g.n1), i2 = int((j / g.n1) % g.n2), i3 = int(j / (int64_t(g.n1) * g.n2));
  const double f[3] = {double(i1) / g.n1, double(i2) / g.n2, double(i3) / g.n3};
  double s = 0.0;
  const double w2 = width * width;
  for (int a = 0; a < natom; ++a) {
    double d[3];
    for (int k = 0; k < 3; ++k) {
      d[k] = f[k] - frac[3 * a + k];
      d[k] -= rint(d[k]);
    }
    const double x = aa[0] * d[0] + aa[3] * d[1] + aa[6] * d[2];
    const double y = aa[1] * d[0] + aa[4] * d[1] + aa[7] * d[2];
    const double z = aa[2] * d[0] + aa[5] * d[1] + aa[8] * d[2];
    s += zat[a] * exp(-(x * x + y * y + z * z) / w2);
  }
  rho[j] = s;
}

__global__ void k_sub(double *__restrict__ y, const double *__restrict__ a, const double *__restrict__ b, int64_t N) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < N) y[j] = a[j] - b[j];
}

// x += beta * w2 * (y - x) on complex arrays
__global__ void k_gmix_step(double2 *__restrict__ x, const double2 *__restrict__ y, const double *__restrict__ w2,
                            double beta, int64_t N) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const double f = beta * w2[j];
  x[j].x += f * (y[j].x - x[j].x);
  x[j].y += f * (y[j].y - x[j].y);
}

struct GSlots {
  const double2 *xin[4];
  const double2 *xou[4];
};

// per-block partial sums of Re(conj(R_i) R_j) wq over the m0 residuals
// R_i = xou_i - xin_i, pairs (i, j <= i) in row order
template <int M0>
__global__ __launch_bounds__(256) void k_gram_g(GSlots sl, const double *__restrict__ wq, int64_t N,
                                                double *__restrict__ part) {
  constexpr int NP = M0 * (M0 + 1) / 2;
  __shared__ double red[4][NP];
  double acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = 0.0;
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < N; j += int64_t(gridDim.x) * blockDim.x) {
    double rr[M0], ri[M0];
#pragma unroll
    for (int i = 0; i < M0; ++i) {
      rr[i] = sl.xou[i][j].x - sl.xin[i][j].x;
      ri[i] = sl.xou[i][j].y - sl.xin[i][j].y;
    }
    const double w = wq[j];
    int q = 0;
#pragma unroll
    for (int i = 0; i < M0; ++i)
#pragma unroll
      for (int k = 0; k <= i; ++k) acc[q++] += w * (rr[i] * rr[k] + ri[i] * ri[k]);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    double v = acc[q];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wid][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < NP) part[int64_t(blockIdx.x) * NP + threadIdx.x] =
      red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

struct GCoef {
  double c[4];
};

// out = X + beta w2 (Y - X), X = sum c_i xin_i, Y = sum c_i xou_i
__global__ void k_gmix_combine(GSlots sl, GCoef cf, int m0, const double *__restrict__ w2, double beta, int64_t N,
                               double2 *__restrict__ out) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= N) return;
  double xr = 0, xi = 0, yr = 0, yi = 0;
  for (int i = 0; i < m0; ++i) {
    xr += cf.c[i] * sl.xin[i][j].x;
    xi += cf.c[i] * sl.xin[i][j].y;
    yr += cf.c[i] * sl.xou[i][j].x;
    yi += cf.c[i] * sl.xou[i][j].y;
  }
  const double f = beta * w2[j];
  out[j] = make_double2(xr + f * (yr - xr), xi + f * (yi - xi));
}

__global__ void k_fill_value(double *__restrict__ y, double v, int64_t N) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < N) y[j] = v;
}

inline unsigned nblk(int64_t N) { return unsigned((N + kT - 1) / kT); }

GridG make_grid(const rsg_ctx *ctx, const ScfState *st) {
  GridG g;
  g.n1 = ctx->n[0];
  g.n2 = ctx->n[1];
  g.n3 = ctx->n[2];
  for (int i = 0; i < 9; ++i) g.b[i] = st->bb[i];
  g.ecut = st->ecut;
  return g;
}

double *field_ptr(rsg_ctx *ctx, int field) {
  ScfState *st = ctx->scf;
  switch (field) {
    case RSG_FIELD_RHO: return st->rho.as<double>();
    case RSG_FIELD_RHO_OUT: return st->rho_out.as<double>();
    case RSG_FIELD_VION: return st->vion.as<double>();
    case RSG_FIELD_VH: return st->vh.as<double>();
    case RSG_FIELD_VXC: return st->vxc.as<double>();
    case RSG_FIELD_VLOC: return ctx->vloc64.as<double>();
    default: return nullptr;
  }
}

int ensure_state(rsg_ctx *ctx) {
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "scf: grid not set");
  ScfState *st = state(ctx);
  const size_t n8 = size_t(ctx->N) * 8;
  if (st->rho.ensure(n8) || st->rho_out.ensure(n8) || st->vion.ensure(n8) || st->vh.ensure(n8) ||
      st->vxc.ensure(n8) || st->work.ensure(n8) || st->zbuf.ensure(2 * n8))
    return ctx->fail(RSG_ERR_HIP, "scf: field allocation failed");
  if (!st->have_plan) {
    DeviceOpGuard guard;
    if (hipfftPlan3d(&st->plan, ctx->n[2], ctx->n[1], ctx->n[0], HIPFFT_Z2Z) != HIPFFT_SUCCESS)
      return ctx->fail(RSG_ERR_HIP, "scf: hipfftPlan3d failed");
    if (hipfftSetStream(st->plan, ctx->stream) != HIPFFT_SUCCESS)
      return ctx->fail(RSG_ERR_HIP, "scf: hipfftSetStream failed");
    st->have_plan = true;
  }
  // reciprocal vectors and the density-grid cutoff of CalcCutoff_Ggrid
  // (orthorhombic cell: Gcut = min_k |b_k| (2*floor((N_k-1)/2)+1) / 2)
  const double *aa = ctx->aa;
  const double vol = aa[0] * (aa[4] * aa[8] - aa[5] * aa[7]) - aa[1] * (aa[3] * aa[8] - aa[5] * aa[6]) +
                     aa[2] * (aa[3] * aa[7] - aa[4] * aa[6]);
  const double *a1 = aa, *a2 = aa + 3, *a3 = aa + 6;
  const double f = 2.0 * kPi / vol;
  double c[3][3] = {{a2[1] * a3[2] - a2[2] * a3[1], a2[2] * a3[0] - a2[0] * a3[2], a2[0] * a3[1] - a2[1] * a3[0]},
                    {a3[1] * a1[2] - a3[2] * a1[1], a3[2] * a1[0] - a3[0] * a1[2], a3[0] * a1[1] - a3[1] * a1[0]},
                    {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2], a1[0] * a2[1] - a1[1] * a2[0]}};
  double gc = 1e300;
  for (int k = 0; k < 3; ++k) {
    double bn = 0.0;
    for (int r = 0; r < 3; ++r) {
      st->bb[3 * k + r] = f * c[k][r];
      bn += st->bb[3 * k + r] * st->bb[3 * k + r];
    }
    gc = std::min(gc, std::sqrt(bn) * (2 * ((ctx->n[k] - 1) / 2) + 1));
  }
  gc *= 0.5;
  st->ecut = gc * gc;
  return RSG_OK;
}

int fft(rsg_ctx *ctx, int dir) {
  ScfState *st = ctx->scf;
  hipfftDoubleComplex *z = reinterpret_cast<hipfftDoubleComplex *>(st->zbuf.p);
  DeviceOpGuard guard;
  if (hipfftExecZ2Z(st->plan, z, z, dir) != HIPFFT_SUCCESS) return ctx->fail(RSG_ERR_HIP, "scf: hipfftExecZ2Z failed");
  return RSG_OK;
}

int dot1(rsg_ctx *ctx, const double *x, const double *y, double alpha, double *out_host) {
  ScfState *st = ctx->scf;
  if (st->occ.ensure(64 * 8)) return ctx->fail(RSG_ERR_HIP, "scf: scratch allocation failed");
  const double *xs[1] = {x}, *ys[1] = {y};
  double *d = st->occ.as<double>() + 48;  // scalar slot past the occupations block
  int rc = multidot(ctx, 1, xs, ys, 1, alpha, d);
  if (rc) return rc;
  RSG_HIP(hipMemcpyAsync(out_host, d, 8, hipMemcpyDeviceToHost, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

// Ewald energy of point ions in a neutralising background (the lattice sum
// cal_ewald computes, ewald_module.f90:50-62), by the standard split into
// real-space erfc and reciprocal Gaussian sums, evaluated on the device: one
// atom pair per thread over its images, one reciprocal vector per thread over
// the atoms, each set of terms summed in a fixed order (k_sum_ordered), so
// the result does not vary from run to run.
struct EwaldArgs {
  double aa[9], bb[9];
  double se, rcut, gcut2;
  int natom, nr[3], ng[3];
};

__global__ void k_ewald_real(EwaldArgs e, const double *__restrict__ frac, const int *__restrict__ kind,
                             const double *__restrict__ sp, double *__restrict__ part) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(e.natom) * e.natom) return;
  const int a = int(t / e.natom), b = int(t - int64_t(a) * e.natom);
  const double f0 = frac[3 * b] - frac[3 * a], f1 = frac[3 * b + 1] - frac[3 * a + 1],
               f2 = frac[3 * b + 2] - frac[3 * a + 2];
  double s = 0.0;
  for (int l1 = -e.nr[0]; l1 <= e.nr[0]; ++l1)
    for (int l2 = -e.nr[1]; l2 <= e.nr[1]; ++l2)
      for (int l3 = -e.nr[2]; l3 <= e.nr[2]; ++l3) {
        if (a == b && l1 == 0 && l2 == 0 && l3 == 0) continue;
        const double d0 = f0 + l1, d1 = f1 + l2, d2 = f2 + l3;
        const double x = e.aa[0] * d0 + e.aa[3] * d1 + e.aa[6] * d2;
        const double y = e.aa[1] * d0 + e.aa[4] * d1 + e.aa[7] * d2;
        const double z = e.aa[2] * d0 + e.aa[5] * d1 + e.aa[8] * d2;
        const double r = sqrt(x * x + y * y + z * z);
        if (r < e.rcut) s += erfc(e.se * r) / r;
      }
  part[t] = sp[6 * kind[a]] * sp[6 * kind[b]] * s;
}

// exp(-G^2 / 4 eta) / G^2 |sum_a Z_a exp(i G.R_a)|^2 for one reciprocal
// vector G = m1 b1 + m2 b2 + m3 b3 of the box |m_k| <= ng_k
__global__ void k_ewald_recip(EwaldArgs e, const double *__restrict__ frac, const int *__restrict__ kind,
                              const double *__restrict__ sp, double inv4eta, double *__restrict__ part) {
  const int w1 = 2 * e.ng[0] + 1, w2 = 2 * e.ng[1] + 1, w3 = 2 * e.ng[2] + 1;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(w1) * w2 * w3) return;
  const int m1 = int(t % w1) - e.ng[0], m2 = int((t / w1) % w2) - e.ng[1], m3 = int(t / (int64_t(w1) * w2)) - e.ng[2];
  const double gx = e.bb[0] * m1 + e.bb[3] * m2 + e.bb[6] * m3;
  const double gy = e.bb[1] * m1 + e.bb[4] * m2 + e.bb[7] * m3;
  const double gz = e.bb[2] * m1 + e.bb[5] * m2 + e.bb[8] * m3;
  const double g2 = gx * gx + gy * gy + gz * gz;
  double v = 0.0;
  if ((m1 | m2 | m3) != 0 && g2 <= e.gcut2) {
    double sr = 0.0, si = 0.0;
    for (int a = 0; a < e.natom; ++a) {
      const double ph = 2.0 * kPi * (m1 * frac[3 * a] + m2 * frac[3 * a + 1] + m3 * frac[3 * a + 2]);
      double sn, cs;
      sincos(ph, &sn, &cs);
      const double za = sp[6 * kind[a]];
      sr += za * cs;
      si += za * sn;
    }
    v = exp(-g2 * inv4eta) / g2 * (sr * sr + si * si);
  }
  part[t] = v;
}

// out[0] = sum of n values in a fixed order: one block, strided partial
// sums, then a shared-memory tree
__global__ __launch_bounds__(1024) void k_sum_ordered(const double *__restrict__ v, int64_t n,
                                                      double *__restrict__ out) {
  __shared__ double sh[1024];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) s += v[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

int ewald_energy(rsg_ctx *ctx, ScfState *st, double vol, int natom, const double *zat, double *energy) {
  double q = 0.0, q2 = 0.0;
  for (int a = 0; a < natom; ++a) {
    q += zat[a];
    q2 += zat[a] * zat[a];
  }
  // Gaussian split parameter.  The energy does not depend on it; eta = pi
  // natom^(1/3) / vol^(2/3) balances the real-space pair-image sum against
  // the reciprocal sum (pi / vol^(2/3) alone makes the real-space cutoff grow
  // with the cell: 13^3 image cells per pair at Si216, and its large
  // background term cancels to 1e-7 Ha there in double precision).
  const double eta = kPi * std::cbrt(double(std::max(natom, 1))) / std::pow(vol, 2.0 / 3.0);
  EwaldArgs e{};
  for (int k = 0; k < 9; ++k) {
    e.aa[k] = ctx->aa[k];
    e.bb[k] = st->bb[k];
  }
  e.se = std::sqrt(eta);
  e.rcut = 8.0 / e.se;
  const double gcut = 16.0 * e.se;
  e.gcut2 = gcut * gcut;
  e.natom = natom;
  auto len = [](const double *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
  for (int k = 0; k < 3; ++k) {
    e.nr[k] = int(std::ceil(e.rcut * len(st->bb + 3 * k) / (2 * kPi))) + 1;
    e.ng[k] = int(std::ceil(gcut / len(st->bb + 3 * k))) + 1;
  }
  const int64_t npair = int64_t(natom) * natom;
  const int64_t ngv = int64_t(2 * e.ng[0] + 1) * (2 * e.ng[1] + 1) * (2 * e.ng[2] + 1);
  if (st->ewpart.ensure(size_t(std::max(npair, ngv) + 2) * 8) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "ewald: allocation failed");
  double *part = st->ewpart.as<double>(), *sums = part + std::max(npair, ngv);
  double h[2] = {0.0, 0.0};
  if (natom > 0) {
    hipLaunchKernelGGL(k_ewald_real, dim3(nblk(npair)), dim3(kT), 0, ctx->stream, e, st->dfrac.as<double>(),
                       st->dkind.as<int>(), st->dsp.as<double>(), part);
    hipLaunchKernelGGL(k_sum_ordered, dim3(1), dim3(1024), 0, ctx->stream, part, npair, sums);
    hipLaunchKernelGGL(k_ewald_recip, dim3(nblk(ngv)), dim3(kT), 0, ctx->stream, e, st->dfrac.as<double>(),
                       st->dkind.as<int>(), st->dsp.as<double>(), 1.0 / (4.0 * eta), part);
    hipLaunchKernelGGL(k_sum_ordered, dim3(1), dim3(1024), 0, ctx->stream, part, ngv, sums + 1);
    RSG_HIP(hipGetLastError());
    RSG_HIP(hipMemcpyAsync(h, sums, 16, hipMemcpyDeviceToHost, ctx->stream));
    RSG_HIP(hipStreamSynchronize(ctx->stream));
  }
  *energy = 0.5 * h[0] + 2.0 * kPi / vol * h[1] - std::sqrt(eta / kPi) * q2 - kPi * q * q / (2.0 * eta * vol);
  return RSG_OK;
}

// small dense solve A x = b (partial pivoting), in place
bool solve_small(int n, std::vector<double> &A, std::vector<double> &b) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i + n * k]) > std::fabs(A[p + n * k])) p = i;
    if (A[p + n * k] == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k + n * j], A[p + n * j]);
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i + n * k] / A[k + n * k];
      for (int j = k; j < n; ++j) A[i + n * j] -= f * A[k + n * j];
      b[i] -= f * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < n; ++j) s -= A[k + n * j] * b[j];
    b[k] = s / A[k + n * k];
  }
  return true;
}

}  // namespace

int scf_set_local_gth(rsg_ctx *ctx, int natom, const double *frac, const int *kind, int nkind, const double *zps,
                      const double *rloc, const double *cloc) {
  if (natom < 0 || nkind < 1 || (natom > 0 && (!frac || !kind)) || !zps || !rloc || !cloc)
    return ctx->fail(RSG_ERR_ARG, "set_local_gth: bad arguments");
  int rc = ensure_state(ctx);
  if (rc) return rc;
  ScfState *st = ctx->scf;
  const int64_t N = ctx->N;
  std::vector<double> sp(static_cast<size_t>(6 * nkind));
  for (int k = 0; k < nkind; ++k) {
    if (!(rloc[k] > 0)) return ctx->fail(RSG_ERR_ARG, "set_local_gth: rloc must be > 0");
    sp[6 * k] = zps[k];
    sp[6 * k + 1] = rloc[k];
    for (int i = 0; i < 4; ++i) sp[6 * k + 2 + i] = cloc[4 * k + i];
  }
  for (int a = 0; a < natom; ++a)
    if (kind[a] < 0 || kind[a] >= nkind) return ctx->fail(RSG_ERR_ARG, "set_local_gth: kind out of range");
  DevBuf &dfrac = st->dfrac, &dkind = st->dkind, &dsp = st->dsp;
  if (dfrac.ensure(size_t(std::max(natom, 1)) * 24) || dkind.ensure(size_t(std::max(natom, 1)) * 4) ||
      dsp.ensure(sp.size() * 8))
    return ctx->fail(RSG_ERR_HIP, "set_local_gth: allocation failed");
  if (natom > 0) {
    RSG_HIP(hipMemcpyAsync(dfrac.p, frac, size_t(natom) * 24, hipMemcpyHostToDevice, ctx->stream));
    RSG_HIP(hipMemcpyAsync(dkind.p, kind, size_t(natom) * 4, hipMemcpyHostToDevice, ctx->stream));
  }
  RSG_HIP(hipMemcpyAsync(dsp.p, sp.data(), sp.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  const double vol = ctx->dv * double(N);
  GridG g = make_grid(ctx, st);
  hipLaunchKernelGGL(k_vion_g, dim3(nblk(N)), dim3(kT), 0, ctx->stream, g, natom, dfrac.as<double>(),
                     dkind.as<int>(), dsp.as<double>(), 1.0 / vol, st->zbuf.as<double2>());
  if ((rc = fft(ctx, HIPFFT_BACKWARD))) return rc;
  hipLaunchKernelGGL(k_real_part, dim3(nblk(N)), dim3(kT), 0, ctx->stream, st->zbuf.as<double2>(),
                     st->vion.as<double>(), N);
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<double> zat(static_cast<size_t>(natom));
  st->nelectron = 0.0;
  for (int a = 0; a < natom; ++a) {
    zat[size_t(a)] = zps[kind[a]];
    st->nelectron += zat[size_t(a)];
  }
  if ((rc = ewald_energy(ctx, st, vol, natom, zat.data(), &st->eewald))) return rc;
  st->frac.assign(frac, frac + 3 * natom);
  st->kind.assign(kind, kind + natom);
  st->zps.assign(zps, zps + nkind);
  st->have_local = true;
  return RSG_OK;
}

int scf_atom_count(const rsg_ctx *ctx) { return ctx->scf ? int(ctx->scf->kind.size()) : 0; }

int scf_constants(rsg_ctx *ctx, double *eewald, double *nel) {
  if (!ctx->scf || !ctx->scf->have_local) return ctx->fail(RSG_ERR_STATE, "scf: local potential not set");
  if (eewald) *eewald = ctx->scf->eewald;
  if (nel) *nel = ctx->scf->nelectron;
  return RSG_OK;
}

int scf_field(rsg_ctx *ctx, int field, double *host, bool get) {
  int rc = ensure_state(ctx);
  if (rc) return rc;
  double *d = field_ptr(ctx, field);
  if (!d) return ctx->fail(RSG_ERR_ARG, "scf: unknown field");
  if (get) {
    RSG_HIP(hipMemcpyAsync(host, d, size_t(ctx->N) * 8, hipMemcpyDeviceToHost, ctx->stream));
  } else {
    RSG_HIP(hipMemcpyAsync(d, host, size_t(ctx->N) * 8, hipMemcpyHostToDevice, ctx->stream));
    if (field == RSG_FIELD_VLOC) {
      std::vector<float> v(host, host + ctx->N);
      RSG_HIP(hipMemcpyAsync(ctx->vloc32.p, v.data(), size_t(ctx->N) * 4, hipMemcpyHostToDevice, ctx->stream));
      RSG_HIP(hipStreamSynchronize(ctx->stream));
      ctx->have_vloc = true;
    }
  }
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

int scf_init_density(rsg_ctx *ctx, double width) {
  int rc = ensure_state(ctx);
  if (rc) return rc;
  ScfState *st = ctx->scf;
  if (!st->have_local) return ctx->fail(RSG_ERR_STATE, "init_density: local potential not set");
  if (!(width > 0)) return ctx->fail(RSG_ERR_ARG, "init_density: width must be > 0");
  const int natom = int(st->kind.size());
  const int64_t N = ctx->N;
  std::vector<double> zat(static_cast<size_t>(natom));
  for (int a = 0; a < natom; ++a) zat[size_t(a)] = st->zps[size_t(st->kind[size_t(a)])];
  DevBuf dfrac, dz, daa;
  if (dfrac.ensure(size_t(std::max(natom, 1)) * 24) || dz.ensure(size_t(std::max(natom, 1)) * 8) || daa.ensure(72))
    return ctx->fail(RSG_ERR_HIP, "init_density: allocation failed");
  if (natom > 0) {
    RSG_HIP(hipMemcpyAsync(dfrac.p, st->frac.data(), size_t(natom) * 24, hipMemcpyHostToDevice, ctx->stream));
    RSG_HIP(hipMemcpyAsync(dz.p, zat.data(), size_t(natom) * 8, hipMemcpyHostToDevice, ctx->stream));
  }
  RSG_HIP(hipMemcpyAsync(daa.p, ctx->aa, 72, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_gauss_rho, dim3(nblk(N)), dim3(kT), 0, ctx->stream, make_grid(ctx, st), daa.as<double>(),
                     natom, dfrac.as<double>(), dz.as<double>(), width, st->rho.as<double>());
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return scf_normalize(ctx);
}

// rho <- rho * Nel / (dV sum rho)  (normalize_initrho, ps_initrho_module.f90:550-560)
int scf_normalize(rsg_ctx *ctx) {
  int rc = ensure_state(ctx);
  if (rc) return rc;
  ScfState *st = ctx->scf;
  const int64_t N = ctx->N;
  double *ones = st->work.as<double>();
  hipLaunchKernelGGL(k_fill_value, dim3(nblk(N)), dim3(kT), 0, ctx->stream, ones, 1.0, N);
  double q = 0.0;
  if ((rc = dot1(ctx, st->rho.as<double>(), ones, ctx->dv, &q))) return rc;
  if (!(q > 0.0)) return ctx->fail(RSG_ERR_ARG, "normalize: density integrates to a non-positive charge");
  hipLaunchKernelGGL(k_axpby, dim3(nblk(N)), dim3(kT), 0, ctx->stream, st->rho.as<double>(), st->rho.as<double>(),
                     0.0, st->nelectron / q, N);
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

int scf_calc_density(rsg_ctx *ctx, const double *occ) {
  int rc = ensure_state(ctx);
  if (rc) return rc;
  if (ctx->mb <= 0) return ctx->fail(RSG_ERR_STATE, "calc_density: no band store");
  ScfState *st = ctx->scf;
  if (st->occ.ensure(size_t(std::max(ctx->mb, 64)) * 8 + 64 * 8))
    return ctx->fail(RSG_ERR_HIP, "calc_density: allocation failed");
  double *docc = st->occ.as<double>() + 64;  // slots 0..63 are dot scratch
  RSG_HIP(hipMemcpyAsync(docc, occ, size_t(ctx->mb) * 8, hipMemcpyHostToDevice, ctx->stream));
  if (ctx->band32 && !ctx->unk_fresh)
    hipLaunchKernelGGL(k_density<float>, dim3(nblk(ctx->N)), dim3(kT), 0, ctx->stream, ctx->unk32.as<float>(),
                       ctx->N, ctx->mb, docc, st->rho_out.as<double>());
  else
    hipLaunchKernelGGL(k_density<double>, dim3(nblk(ctx->N)), dim3(kT), 0, ctx->stream, ctx->unk.as<double>(),
                       ctx->N, ctx->mb, docc, st->rho_out.as<double>());
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

int scf_hartree_xc(rsg_ctx *ctx, int field, double *eh, double *exc) {
  int rc = ensure_state(ctx);
  if (rc) return rc;
  if (field != RSG_FIELD_RHO && field != RSG_FIELD_RHO_OUT)
    return ctx->fail(RSG_ERR_ARG, "hartree_xc: field must be RHO or RHO_OUT");
  ScfState *st = ctx->scf;
  const int64_t N = ctx->N;
  const double *rho = field_ptr(ctx, field);
  hipLaunchKernelGGL(k_to_complex, dim3(nblk(N)), dim3(kT), 0, ctx->stream, rho, st->zbuf.as<double2>(), N);
  if ((rc = fft(ctx, HIPFFT_FORWARD))) return rc;
  hipLaunchKernelGGL(k_hartree_g, dim3(nblk(N)), dim3(kT), 0, ctx->stream, make_grid(ctx, st), st->zbuf.as<double2>());
  if ((rc = fft(ctx, HIPFFT_BACKWARD))) return rc;
  hipLaunchKernelGGL(k_real_part, dim3(nblk(N)), dim3(kT), 0, ctx->stream, st->zbuf.as<double2>(),
                     st->vh.as<double>(), N);
  hipLaunchKernelGGL(k_ldapz81, dim3(nblk(N)), dim3(kT), 0, ctx->stream, rho, st->vxc.as<double>(),
                     st->work.as<double>(), N);
  double e = 0.0;
  if ((rc = dot1(ctx, rho, st->vh.as<double>(), 0.5 * ctx->dv, &e))) return rc;
  if (eh) *eh = e;
  // exc: dV * sum of the energy density (dotted with ones held in zbuf)
  double *ones = reinterpret_cast<double *>(st->zbuf.p);
  hipLaunchKernelGGL(k_fill_value, dim3(nblk(N)), dim3(kT), 0, ctx->stream, ones, 1.0, N);
  if ((rc = dot1(ctx, st->work.as<double>(), ones, ctx->dv, &e))) return rc;
  if (exc) *exc = e;
  return RSG_OK;
}

int scf_update_vloc(rsg_ctx *ctx) {
  int rc = ensure_state(ctx);
  if (rc) return rc;
  ScfState *st = ctx->scf;
  if (!st->have_local) return ctx->fail(RSG_ERR_STATE, "update_vloc: local potential not set");
  hipLaunchKernelGGL(k_add3, dim3(nblk(ctx->N)), dim3(kT), 0, ctx->stream, st->vion.as<double>(),
                     st->vh.as<double>(), st->vxc.as<double>(), ctx->vloc64.as<double>(), ctx->vloc32.as<float>(),
                     ctx->N);
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_vloc = true;
  return RSG_OK;
}

int scf_density_integrals(rsg_ctx *ctx, int field, double *eloc, double *eion) {
  int rc = ensure_state(ctx);
  if (rc) return rc;
  if (field != RSG_FIELD_RHO && field != RSG_FIELD_RHO_OUT)
    return ctx->fail(RSG_ERR_ARG, "density_integrals: field must be RHO or RHO_OUT");
  const double *rho = field_ptr(ctx, field);
  double e = 0.0;
  if ((rc = dot1(ctx, rho, ctx->vloc64.as<double>(), ctx->dv, &e))) return rc;
  if (eloc) *eloc = e;
  if ((rc = dot1(ctx, rho, ctx->scf->vion.as<double>(), ctx->dv, &e))) return rc;
  if (eion) *eion = e;
  return RSG_OK;
}

namespace {

// Fourier coefficients FFT(r) / N into a complex history slot
int to_g(rsg_ctx *ctx, const double *r, double2 *slot) {
  ScfState *st = ctx->scf;
  const int64_t N = ctx->N;
  hipLaunchKernelGGL(k_to_complex, dim3(nblk(N)), dim3(kT), 0, ctx->stream, r, st->zbuf.as<double2>(), N);
  int rc = fft(ctx, HIPFFT_FORWARD);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zscale, dim3(nblk(N)), dim3(kT), 0, ctx->stream, st->zbuf.as<double2>(), slot,
                     1.0 / double(N), N);
  return RSG_OK;
}

// Kerker weights of get_Kerker_weights (mixing_pulay_module.f90:275-337)
int kerker_weights(rsg_ctx *ctx) {
  ScfState *st = ctx->scf;
  const int64_t N = ctx->N;
  const int n1 = ctx->n[0], n2 = ctx->n[1], n3 = ctx->n[2];
  std::vector<double> gg(static_cast<size_t>(N));
  double gmin = 1e300, gmax = -1e300;
  auto fr = [](int i, int n) {
    int j = i - n;
    return j < -(n - 1) / 2 ? i : j;
  };
  for (int i3 = 0; i3 < n3; ++i3)
    for (int i2 = 0; i2 < n2; ++i2)
      for (int i1 = 0; i1 < n1; ++i1) {
        const int j1 = fr(i1, n1), j2 = fr(i2, n2), j3 = fr(i3, n3);
        const double gx = j1 * st->bb[0] + j2 * st->bb[3] + j3 * st->bb[6];
        const double gy = j1 * st->bb[1] + j2 * st->bb[4] + j3 * st->bb[7];
        const double gz = j1 * st->bb[2] + j2 * st->bb[5] + j3 * st->bb[8];
        const double g2 = gx * gx + gy * gy + gz * gz;
        gg[size_t(i1 + int64_t(n1) * (i2 + int64_t(n2) * i3))] = g2;
        if (j1 == 0 && j2 == 0 && j3 == 0) continue;
        gmin = std::min(gmin, g2);
        gmax = std::max(gmax, g2);
      }
  const double b1 = 19.0 * gmin * gmax / (gmax - 20.0 * gmin);
  const double b2 = std::pow((1.5 * 0.529177) * (1.5 * 0.529177), 2.0);
  std::vector<double> w(static_cast<size_t>(N)), w2(static_cast<size_t>(N));
  for (int64_t i = 0; i < N; ++i) {
    const double g2 = gg[size_t(i)];
    w[size_t(i)] = g2 == 0.0 ? 1.0 : (g2 + b1) / g2;
    w2[size_t(i)] = g2 == 0.0 ? 1.0 : g2 / (g2 + b2);
  }
  if (st->wq.ensure(size_t(N) * 8) || st->wq2.ensure(size_t(N) * 8))
    return ctx->fail(RSG_ERR_HIP, "mix_init: weight allocation failed");
  RSG_HIP(hipMemcpyAsync(st->wq.p, w.data(), size_t(N) * 8, hipMemcpyHostToDevice, ctx->stream));
  RSG_HIP(hipMemcpyAsync(st->wq2.p, w2.data(), size_t(N) * 8, hipMemcpyHostToDevice, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

}  // namespace

// init_mixing (mixing_module.f90:120-126): history slot mmix holds rho_in
// (IMIX 10, real space) or its Fourier coefficients (IMIX 30).
int scf_mix_init(rsg_ctx *ctx, int mmix, double beta, int imix) {
  int rc = ensure_state(ctx);
  if (rc) return rc;
  if (mmix < 1 || mmix > 32 || !(beta > 0.0 && beta <= 1.0))
    return ctx->fail(RSG_ERR_ARG, "mix_init: mmix in 1..32 and 0 < beta <= 1");
  if (imix != 10 && imix != 30) return ctx->fail(RSG_ERR_ARG, "mix_init: imix must be 10 or 30");
  if (imix == 30 && mmix > 4) return ctx->fail(RSG_ERR_ARG, "mix_init: imix 30 takes mmix <= 4");
  ScfState *st = ctx->scf;
  // history slots are kept across SCF runs and only grown: a free waits for
  // the whole device, which stalls every other context's stream
  auto resize = [](std::vector<DevBuf> &v, size_t n) {
    for (size_t i = n; i < v.size(); ++i) v[i].release();
    v.resize(n);
  };
  st->imix = imix;
  const size_t n8 = size_t(ctx->N) * 8;
  if (imix == 10) {
    resize(st->gxin, 0);
    resize(st->gxou, 0);
    resize(st->xin, size_t(mmix));
    resize(st->xou, size_t(mmix));
    for (int k = 0; k < mmix; ++k)
      if (st->xin[size_t(k)].ensure(n8) || st->xou[size_t(k)].ensure(n8))
        return ctx->fail(RSG_ERR_HIP, "mix_init: history allocation failed");
    if (st->res.ensure(n8 * size_t(mmix)) || st->occ.ensure(size_t(mmix) * mmix * 8 + 64 * 8))
      return ctx->fail(RSG_ERR_HIP, "mix_init: residual allocation failed");
  } else {
    resize(st->xin, 0);
    resize(st->xou, 0);
    resize(st->gxin, size_t(mmix));
    resize(st->gxou, size_t(mmix));
    for (int k = 0; k < mmix; ++k)
      if (st->gxin[size_t(k)].ensure(2 * n8) || st->gxou[size_t(k)].ensure(2 * n8))
        return ctx->fail(RSG_ERR_HIP, "mix_init: history allocation failed");
    if (st->gpart.ensure(size_t(1024) * 10 * 8)) return ctx->fail(RSG_ERR_HIP, "mix_init: allocation failed");
    if ((rc = kerker_weights(ctx))) return rc;
  }
  st->mmix = mmix;
  st->beta = beta;
  st->count = 0;
  if (imix == 10) {
    RSG_HIP(hipMemcpyAsync(st->xin[size_t(mmix - 1)].p, st->rho.p, n8, hipMemcpyDeviceToDevice, ctx->stream));
  } else if ((rc = to_g(ctx, st->rho.as<double>(), st->gxin[size_t(mmix - 1)].as<double2>()))) {
    return rc;
  }
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

namespace {

// pulay_g_mixing (mixing_pulay_module.f90:124-227) on f = rho_out
int mix_pulay_g(rsg_ctx *ctx) {
  ScfState *st = ctx->scf;
  const int mm = st->mmix;
  const int64_t N = ctx->N;
  const size_t n16 = size_t(N) * 16;
  int rc = to_g(ctx, st->rho_out.as<double>(), st->gxou[size_t(mm - 1)].as<double2>());
  if (rc) return rc;
  st->count += 1;
  const int m0 = std::min(st->count, mm);
  auto rotate = [&](std::vector<DevBuf> &h) -> int {
    DevBuf first = h[0];
    for (int i = 1; i < mm; ++i) h[size_t(i - 1)] = h[size_t(i)];
    h[size_t(mm - 1)] = first;
    if (mm > 1 && hipMemcpyAsync(h[size_t(mm - 1)].p, h[size_t(mm - 2)].p, n16, hipMemcpyDeviceToDevice, ctx->stream))
      return RSG_ERR_HIP;
    return RSG_OK;
  };
  if (mm == 1 || m0 < mm) {
    if (rotate(st->gxin) || rotate(st->gxou)) return ctx->fail(RSG_ERR_HIP, "mix_pulay: history copy failed");
    hipLaunchKernelGGL(k_gmix_step, dim3(nblk(N)), dim3(kT), 0, ctx->stream, st->gxin[size_t(mm - 1)].as<double2>(),
                       st->gxou[size_t(mm - 1)].as<double2>(), st->wq2.as<double>(), st->beta, N);
  } else {
    const int base = mm - m0;
    GSlots sl{};
    for (int i = 0; i < m0; ++i) {
      sl.xin[i] = st->gxin[size_t(base + i)].as<double2>();
      sl.xou[i] = st->gxou[size_t(base + i)].as<double2>();
    }
    const int nblock = 1024, np = m0 * (m0 + 1) / 2;
    double *part = st->gpart.as<double>();
    switch (m0) {
      case 1: hipLaunchKernelGGL(k_gram_g<1>, dim3(nblock), dim3(256), 0, ctx->stream, sl, st->wq.as<double>(), N, part); break;
      case 2: hipLaunchKernelGGL(k_gram_g<2>, dim3(nblock), dim3(256), 0, ctx->stream, sl, st->wq.as<double>(), N, part); break;
      case 3: hipLaunchKernelGGL(k_gram_g<3>, dim3(nblock), dim3(256), 0, ctx->stream, sl, st->wq.as<double>(), N, part); break;
      default: hipLaunchKernelGGL(k_gram_g<4>, dim3(nblock), dim3(256), 0, ctx->stream, sl, st->wq.as<double>(), N, part); break;
    }
    std::vector<double> ph(size_t(nblock) * np);
    RSG_HIP(hipMemcpyAsync(ph.data(), part, ph.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    RSG_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<double> A(size_t(m0) * m0), b(size_t(m0), 1.0);
    int q = 0;
    for (int i = 0; i < m0; ++i)
      for (int k = 0; k <= i; ++k, ++q) {
        double v = 0.0;
        for (int blk = 0; blk < nblock; ++blk) v += ph[size_t(blk) * np + q];
        A[size_t(i) + size_t(m0) * k] = A[size_t(k) + size_t(m0) * i] = v;
      }
    if (!solve_small(m0, A, b)) return ctx->fail(RSG_ERR_SOLVER, "mix_pulay: singular residual matrix");
    double sb = 0.0;
    for (double v : b) sb += v;
    GCoef cf{};
    for (int i = 0; i < m0; ++i) cf.c[i] = b[size_t(i)] / sb;
    // the combination goes to zbuf, then becomes the newest input slot
    hipLaunchKernelGGL(k_gmix_combine, dim3(nblk(N)), dim3(kT), 0, ctx->stream, sl, cf, m0, st->wq2.as<double>(),
                       st->beta, N, st->zbuf.as<double2>());
    if (rotate(st->gxin) || rotate(st->gxou)) return ctx->fail(RSG_ERR_HIP, "mix_pulay: history copy failed");
    RSG_HIP(hipMemcpyAsync(st->gxin[size_t(mm - 1)].p, st->zbuf.p, n16, hipMemcpyDeviceToDevice, ctx->stream));
  }
  // rho = Re IFFT(Xin(mmix)) (coefficients carry the 1/N)
  RSG_HIP(hipMemcpyAsync(st->zbuf.p, st->gxin[size_t(mm - 1)].p, n16, hipMemcpyDeviceToDevice, ctx->stream));
  if ((rc = fft(ctx, HIPFFT_BACKWARD))) return rc;
  hipLaunchKernelGGL(k_real_part, dim3(nblk(N)), dim3(kT), 0, ctx->stream, st->zbuf.as<double2>(),
                     st->rho.as<double>(), N);
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

}  // namespace

// pulay_mixing (mixing_pulay_module.f90:29-121) on f = rho_out; the mixed
// input density lands in history slot mmix and in rho.
int scf_mix_pulay(rsg_ctx *ctx) {
  int rc = ensure_state(ctx);
  if (rc) return rc;
  ScfState *st = ctx->scf;
  if (st->mmix >= 1 && st->imix == 30) return mix_pulay_g(ctx);
  const int mm = st->mmix;
  if (mm < 1) return ctx->fail(RSG_ERR_STATE, "mix_pulay: mixing not initialised");
  const int64_t N = ctx->N;
  const size_t n8 = size_t(N) * 8;
  auto rotate = [&](std::vector<DevBuf> &h) -> int {
    // Xin(:,:,i-1) = Xin(:,:,i), i = 2..mmix; slot mmix keeps its value
    DevBuf first = h[0];
    for (int i = 1; i < mm; ++i) h[size_t(i - 1)] = h[size_t(i)];
    h[size_t(mm - 1)] = first;
    if (mm > 1) {
      if (hipMemcpyAsync(h[size_t(mm - 1)].p, h[size_t(mm - 2)].p, n8, hipMemcpyDeviceToDevice, ctx->stream))
        return RSG_ERR_HIP;
    }
    return RSG_OK;
  };
  RSG_HIP(hipMemcpyAsync(st->xou[size_t(mm - 1)].p, st->rho_out.p, n8, hipMemcpyDeviceToDevice, ctx->stream));
  st->count += 1;
  const int m0 = std::min(st->count, mm);
  double *X = st->work.as<double>();
  if (mm == 1 || m0 < mm) {
    if (rotate(st->xin) || rotate(st->xou)) return ctx->fail(RSG_ERR_HIP, "mix_pulay: history copy failed");
    double *xin = st->xin[size_t(mm - 1)].as<double>();
    const double *xou = st->xou[size_t(mm - 1)].as<double>();
    // Xin = Xin + beta (Xou - Xin)
    hipLaunchKernelGGL(k_axpby, dim3(nblk(N)), dim3(kT), 0, ctx->stream, xin, xou, st->beta, 1.0 - st->beta, N);
  } else {
    // A(i,j) = <R_i, R_j> over explicit residuals R_i = Xou_i - Xin_i, so the
    // small residuals near convergence do not cancel against |rho|^2
    std::vector<double> A(size_t(m0) * m0), b(size_t(m0), 1.0);
    const int base = mm - m0;
    double *R = st->res.as<double>();
    for (int i = 0; i < m0; ++i)
      hipLaunchKernelGGL(k_sub, dim3(nblk(N)), dim3(kT), 0, ctx->stream, R + size_t(i) * N,
                         st->xou[size_t(base + i)].as<double>(), st->xin[size_t(base + i)].as<double>(), N);
    std::vector<std::pair<int, int>> pairs;
    for (int i = 0; i < m0; ++i)
      for (int j = 0; j <= i; ++j) pairs.emplace_back(i, j);
    double *d = st->occ.as<double>();
    for (size_t q0 = 0; q0 < pairs.size(); q0 += 6) {
      const int np = int(std::min<size_t>(6, pairs.size() - q0));
      const double *xs[6], *ys[6];
      for (int q = 0; q < np; ++q) {
        xs[q] = R + size_t(pairs[q0 + q].first) * N;
        ys[q] = R + size_t(pairs[q0 + q].second) * N;
      }
      if ((rc = multidot(ctx, np, xs, ys, 1, 1.0, d + q0))) return rc;
    }
    std::vector<double> h(pairs.size());
    RSG_HIP(hipMemcpyAsync(h.data(), d, h.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    RSG_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t q = 0; q < pairs.size(); ++q) {
      const int i = pairs[q].first, j = pairs[q].second;
      A[size_t(i) + size_t(m0) * j] = A[size_t(j) + size_t(m0) * i] = h[q];
    }
    if (!solve_small(m0, A, b)) return ctx->fail(RSG_ERR_SOLVER, "mix_pulay: singular residual matrix");
    double sb = 0.0;
    for (double v : b) sb += v;
    for (double &v : b) v /= sb;
    // X = sum b_i Xin_i, Y = sum b_i Xou_i (Y into rho_out scratch)
    double *Y = st->rho_out.as<double>();
    hipLaunchKernelGGL(k_fill_value, dim3(nblk(N)), dim3(kT), 0, ctx->stream, X, 0.0, N);
    hipLaunchKernelGGL(k_fill_value, dim3(nblk(N)), dim3(kT), 0, ctx->stream, Y, 0.0, N);
    for (int i = 0; i < m0; ++i) {
      hipLaunchKernelGGL(k_axpby, dim3(nblk(N)), dim3(kT), 0, ctx->stream, X, st->xin[size_t(base + i)].as<double>(),
                         b[size_t(i)], 1.0, N);
      hipLaunchKernelGGL(k_axpby, dim3(nblk(N)), dim3(kT), 0, ctx->stream, Y, st->xou[size_t(base + i)].as<double>(),
                         b[size_t(i)], 1.0, N);
    }
    if (rotate(st->xin) || rotate(st->xou)) return ctx->fail(RSG_ERR_HIP, "mix_pulay: history copy failed");
    // Xin(mmix) = X + beta (Y - X)
    double *xin = st->xin[size_t(mm - 1)].as<double>();
    RSG_HIP(hipMemcpyAsync(xin, X, n8, hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(k_axpby, dim3(nblk(N)), dim3(kT), 0, ctx->stream, xin, Y, st->beta, 1.0 - st->beta, N);
    // restore rho_out (the unmixed output) from the history
    RSG_HIP(hipMemcpyAsync(Y, st->xou[size_t(mm - 1)].p, n8, hipMemcpyDeviceToDevice, ctx->stream));
  }
  RSG_HIP(hipMemcpyAsync(st->rho.p, st->xin[size_t(mm - 1)].p, n8, hipMemcpyDeviceToDevice, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

// Forget the cell (local potential, constants, mixing history) but keep the
// fields and the FFT plan, which depend on the grid dimensions only.
void scf_reset(rsg_ctx *ctx) {
  ScfState *st = ctx->scf;
  if (!st) return;
  st->have_local = false;
  st->eewald = st->nelectron = 0.0;
  st->mmix = st->count = 0;
  st->zps.clear();
  st->rloc.clear();
  st->cloc.clear();
  st->frac.clear();
  st->kind.clear();
}

void scf_release(rsg_ctx *ctx) {
  ScfState *st = ctx->scf;
  if (!st) return;
  if (st->have_plan) {
    DeviceOpGuard guard;
    hipfftDestroy(st->plan);
  }
  for (DevBuf *b : {&st->rho, &st->rho_out, &st->vion, &st->vh, &st->vxc, &st->zbuf, &st->work, &st->occ}) b->release();
  for (auto &b : st->xin) b.release();
  for (auto &b : st->xou) b.release();
  for (auto &b : st->gxin) b.release();
  for (auto &b : st->gxou) b.release();
  for (DevBuf *b : {&st->res, &st->wq, &st->wq2, &st->gpart, &st->dfrac, &st->dkind, &st->dsp, &st->ewpart})
    b->release();
  delete st;
  ctx->scf = nullptr;
}

}  // namespace rsg
