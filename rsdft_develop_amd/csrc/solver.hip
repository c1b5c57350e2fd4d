// solver.hip -- Gram-Schmidt (Cholesky-QR), Rayleigh-Ritz, band CG with the
// hprsdft preconditioner, and per-band eigenvalues on the device band store.
//
// Dense contractions with K = N (overlap, Hsub, rotation, R^-1 apply) go to
// rocBLAS (MFMA FP64 on gfx950); the small dense factorisations to rocSOLVER.
// Band-wise vector work, dot products and the per-band 2x2 Ritz step are the
// kernels below.
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "lpc.hpp"
#include "solver.hpp"

namespace rsg {
namespace {

constexpr int kThreads = 256;

// ---------------------------------------------------------------- reductions
// The six products of the CG Ritz step, (x.x, p.x, p.p, x.Hx, p.Hx, p.Hp),
// from one load of each of the four vectors per point; partial sums in the
// k_multidot layout, partial[(q nb + b) nseg + s].
template <typename W>
__global__ __launch_bounds__(kThreads) void k_ritz_dots(const W *x, const W *p, const W *hx, const W *hp, int64_t N,
                                                        int nb, int nseg, int64_t seglen, double *partial) {
  constexpr int NP = 6;
  __shared__ double red[kThreads / 64][NP];
  const int s = blockIdx.x, b = blockIdx.y;
  const int64_t i0 = int64_t(s) * seglen, i1 = min(N, i0 + seglen);
  double acc[NP] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int64_t off = int64_t(b) * N;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
    const double xv = double(x[off + i]), pv = double(p[off + i]);
    const double hxv = double(hx[off + i]), hpv = double(hp[off + i]);
    acc[0] += xv * xv;
    acc[1] += pv * xv;
    acc[2] += pv * pv;
    acc[3] += xv * hxv;
    acc[4] += pv * hxv;
    acc[5] += pv * hpv;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    double v = acc[q];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wid][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < NP) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < kThreads / 64; ++k) v += red[k][threadIdx.x];
    partial[(int64_t(threadIdx.x) * nb + b) * nseg + s] = v;
  }
}

template <typename W> struct DotArgs {
  const W *x[6];
  const W *y[6];
};

// products and sums in FP64 for either storage type
template <int NP, typename W>
__global__ __launch_bounds__(kThreads) void k_multidot(DotArgs<W> a, int64_t N, int nb, int nseg,
                                                       int64_t seglen, double *partial) {
  __shared__ double red[kThreads / 64][NP];
  const int s = blockIdx.x, b = blockIdx.y;
  const int64_t i0 = int64_t(s) * seglen, i1 = min(N, i0 + seglen);
  double acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = 0.0;
  const int64_t off = int64_t(b) * N;
  // four independent points per trip keep more loads in flight
  int64_t i = i0 + threadIdx.x;
  for (; i + 3 * kThreads < i1; i += 4 * kThreads) {
    double part[4][NP];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int q = 0; q < NP; ++q)
        part[u][q] = double(a.x[q][off + i + u * kThreads]) * double(a.y[q][off + i + u * kThreads]);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int q = 0; q < NP; ++q) acc[q] += part[u][q];
  }
  for (; i < i1; i += kThreads) {
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[q] += double(a.x[q][off + i]) * double(a.y[q][off + i]);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    double v = acc[q];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wid][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < NP) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < kThreads / 64; ++k) v += red[k][threadIdx.x];
    partial[(int64_t(threadIdx.x) * nb + b) * nseg + s] = v;
  }
}

__global__ void k_reduce_partials(const double *__restrict__ partial, int nseg, int count,
                                  double alpha, double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  double v = 0.0;
  for (int s = 0; s < nseg; ++s) v += partial[int64_t(t) * nseg + s];
  out[t] = alpha * v;
}

// --------------------------------------------------------- band-wise updates
// gk = -2 (hx - E x)          (cg_module.f90:207)
template <typename W>
__global__ void k_cg_grad(const W *__restrict__ hx, const W *__restrict__ x, const double *__restrict__ E,
                          W *__restrict__ g, int64_t N) {
  const int b = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t k = int64_t(b) * N + i;
  const W e = W(E[b]);
  g[k] = W(-2) * (hx[k] - e * x[k]);
}

// y = x (first step) or y = x + (rb/gkgk) y   (cg_module.f90:266-279);
// a zero previous <Pg, g> (a band whose gradient vanished) restarts the
// direction instead of forming 0/0
template <typename W>
__global__ void k_cg_dir(const W *__restrict__ pg, W *__restrict__ pk, const double *__restrict__ rb,
                         const double *__restrict__ gkgk, int first, int64_t N) {
  const int b = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t k = int64_t(b) * N + i;
  const double bk = gkgk[b] != 0.0 ? rb[b] / gkgk[b] : 0.0;
  pk[k] = first ? pg[k] : pg[k] + W(bk) * pk[k];
}

// 2x2 generalised Ritz step per band (cg_module.f90:305-334): DSYGV on
// (<x,p|H|x,p>, <x,p|x,p>), root switch, sign fix.  d6 holds
// (xx, px, pp, xHx, pHx, pHp) * dV as [q*nb + b].
__global__ void k_cg_ritz(const double *__restrict__ d6, int nb, double *__restrict__ E,
                          double *__restrict__ E1, double *__restrict__ uc, const int *__restrict__ done) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  if (*done) {
    // the block has converged (the reference leaves the CG loop here): the
    // remaining steps keep every band, u = (1, 0), E unchanged
    uc[2 * b] = 1.0;
    uc[2 * b + 1] = 0.0;
    E1[b] = E[b];
    return;
  }
  const double b11 = d6[0 * nb + b], b12 = d6[1 * nb + b], b22 = d6[2 * nb + b];
  const double a11 = d6[3 * nb + b], a12 = d6[4 * nb + b], a22 = d6[5 * nb + b];
  // B = U^T U, C = U^-T A U^-1, C = Y diag(w) Y^T, X = U^-1 Y
  const double u11 = sqrt(b11), u12 = b12 / u11, u22 = sqrt(b22 - u12 * u12);
  const double i11 = 1.0 / u11, i12 = -u12 / (u11 * u22), i22 = 1.0 / u22;
  const double t11 = a11 * i11, t12 = a11 * i12 + a12 * i22, t22 = a12 * i12 + a22 * i22;
  const double c11 = i11 * t11, c12 = i11 * t12, c22 = i12 * t12 + i22 * t22;
  double y11, y12, y21, y22, l1, l2;
  if (c12 == 0.0) {
    y11 = 1; y21 = 0; y12 = 0; y22 = 1; l1 = c11; l2 = c22;
  } else {
    const double th = (c22 - c11) / (2.0 * c12);
    const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
    l1 = c11 - t * c12;
    l2 = c22 + t * c12;
    y11 = c; y21 = -s; y12 = s; y22 = c;
  }
  if (l2 < l1) {
    double tl = l1; l1 = l2; l2 = tl;
    double p = y11, q = y21;
    y11 = y12; y21 = y22; y12 = p; y22 = q;
  }
  double v1 = i11 * y11 + i12 * y21, v2 = i22 * y21;
  const double w1 = i11 * y12 + i12 * y22, w2 = i22 * y22;
  double W = l1;
  const double e = E[b];
  if (fabs(l1 - e) > 1e-1 && fabs(l2 - e) <= 1e-1) {
    v1 = w1;
    v2 = w2;
    W = l2;
  }
  if (v1 < 0.0) {
    v1 = -v1;
    v2 = -v2;
  }
  // B not positive definite (p parallel to x, or zero) leaves non-finite
  // coefficients: keep the band as it is (u = (1, 0), E unchanged); the
  // guard below then rejects the step
  if (!(std::isfinite(v1) && std::isfinite(v2) && std::isfinite(W))) {
    v1 = 1.0;
    v2 = 0.0;
    W = e;
  }
  uc[2 * b] = v1;
  uc[2 * b + 1] = v2;
  E1[b] = e;
  E[b] = W;
}

// hx <- u1 hx + u2 hp ; g <- -2 (hx - W (u1 x + u2 p))   (cg_module.f90:336-352)
template <typename W>
__global__ void k_cg_update(W *__restrict__ hx, const W *__restrict__ hp, const W *__restrict__ x,
                            const W *__restrict__ p, W *__restrict__ g, const double *__restrict__ uc,
                            const double *__restrict__ E, int64_t N) {
  const int b = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t k = int64_t(b) * N + i;
  const W u1 = W(uc[2 * b]), u2 = W(uc[2 * b + 1]), ev = W(E[b]);
  const W h = u1 * hx[k] + u2 * hp[k];
  hx[k] = h;
  g[k] = W(-2) * (h - ev * (u1 * x[k] + u2 * p[k]));
}

// divergence guard (cg_module.f90:358-364): reject when rb/res > 1e8, and
// also when the new gradient norm is not finite (the reference would carry
// a NaN into the band); an exactly zero gradient before and after (0/0)
// keeps the band, which the Ritz step left unchanged
__global__ void k_cg_guard(const double *__restrict__ rb, const double *__restrict__ res,
                           double *__restrict__ E, const double *__restrict__ E1,
                           int *__restrict__ accept, int nb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int ok = std::isfinite(rb[b]) && !(rb[b] / res[b] > 1.0e8);
  accept[b] = ok;
  if (!ok) E[b] = E1[b];
}

// x <- u1 x + u2 p for accepted bands (cg_module.f90:366)
// (mixed precision: x is the FP64 store, x32 its FP32 copy for the block)
template <typename W>
__global__ void k_cg_accept(double *__restrict__ x, W *__restrict__ x32, const W *__restrict__ p,
                            const double *__restrict__ uc, const int *__restrict__ accept, int64_t N) {
  const int b = blockIdx.y;
  if (!accept[b]) return;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t k = int64_t(b) * N + i;
  const double v = uc[2 * b] * x[k] + uc[2 * b + 1] * double(p[k]);
  x[k] = v;
  if (x32) x32[k] = W(v);
}

// The same update on bands held in the FP32 store only
template <typename W>
__global__ void k_cg_accept32(W *__restrict__ x, const W *__restrict__ p, const double *__restrict__ uc,
                              const int *__restrict__ accept, int64_t N) {
  const int b = blockIdx.y;
  if (!accept[b]) return;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t k = int64_t(b) * N + i;
  x[k] = W(uc[2 * b] * double(x[k]) + uc[2 * b + 1] * double(p[k]));
}

// y = a y + b x (n values)
template <typename W>
__global__ void k_axpby(W *__restrict__ y, const W *__restrict__ x, W a, W b, int64_t n) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) y[k] = a * y[k] + b * x[k];
}

// y = x (n values): a kernel rather than a copy node, so that a captured
// CG block holds kernels only
template <typename W>
__global__ void k_copy(W *__restrict__ y, const W *__restrict__ x, int64_t n) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) y[k] = x[k];
}

// x32 = (W) x for an nn-band block
template <typename W>
__global__ void k_to_work(const double *__restrict__ x, W *__restrict__ x32, int64_t total) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < total) x32[k] = W(x[k]);
}

// x64 = (double) x for an nn-band block
template <typename W>
__global__ void k_to_double(const W *__restrict__ x, double *__restrict__ x64, int64_t total) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < total) x64[k] = double(x[k]);
}

// cggs coefficient mask (cggs_module.f90:35-127): column j of C (mb x nn)
// belongs to store band nt = n0 + j; swgs 1 keeps rows m < nt, swgs 2
// every row but nt
__global__ void k_cggs_mask(double *__restrict__ C, int mb, int nn, int n0, int swgs) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (m >= mb || j >= nn) return;
  const int nt = n0 + j;
  const bool keep = swgs == 1 ? m < nt : m != nt;
  if (!keep) C[m + int64_t(mb) * j] = 0.0;
}

__global__ void k_scale_copy(const double *__restrict__ a, double s, double *__restrict__ b, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = s * a[t];
}

__global__ void k_fill(double *__restrict__ a, double v, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

// ------------------------------------------------ hprsdft preconditioner
// f = A g with A the 7-point -1/2 Laplacian (E0 = 0): precond_cg_mat2_omp,
// cgpc_hprsdft_module.f90:288-361, periodic 1-deep halo (bcset_1).
template <typename W>
__global__ __launch_bounds__(kThreads) void k_pc_mat(const W *__restrict__ g, W *__restrict__ f, int n1,
                                                     int n2, int n3, W c0, W c1, W c2, W c3,
                                                     const float *__restrict__ diag) {
  const int64_t n12 = int64_t(n1) * n2, N = n12 * n3;
  const int b = blockIdx.y;
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const int i1 = int(j % n1), i2 = int((j / n1) % n2), i3 = int(j / n12);
  const W *w = g + int64_t(b) * N;
  const int xp = i1 + 1 == n1 ? 0 : i1 + 1, xm = i1 == 0 ? n1 - 1 : i1 - 1;
  const int yp = i2 + 1 == n2 ? 0 : i2 + 1, ym = i2 == 0 ? n2 - 1 : i2 - 1;
  const int zp = i3 + 1 == n3 ? 0 : i3 + 1, zm = i3 == 0 ? n3 - 1 : i3 - 1;
  const int64_t row = int64_t(i2) * n1 + i3 * n12;
  const W d = diag ? W(diag[j]) : W(0);
  f[int64_t(b) * N + j] = (c0 + d) * w[j] + c1 * (w[xp + row] + w[xm + row]) +
                          c2 * (w[i1 + int64_t(yp) * n1 + i3 * n12] + w[i1 + int64_t(ym) * n1 + i3 * n12]) +
                          c3 * (w[i1 + int64_t(i2) * n1 + zp * n12] + w[i1 + int64_t(i2) * n1 + zm * n12]);
}

// r = g - f ; p = r
template <typename W>
__global__ void k_pc_init(const W *__restrict__ g, const W *__restrict__ f, W *__restrict__ r,
                          W *__restrict__ p, int64_t total) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const W v = g[k] - f[k];
  r[k] = v;
  p[k] = v;
}

// a = rr1/pAp, rr0 = rr1 (per band).  A band whose residual is exactly
// zero (its direction too, so pAp = 0) takes a zero step instead of 0/0.
__global__ void k_pc_alpha(const double *__restrict__ rr1, const double *__restrict__ pap,
                           double *__restrict__ rr0, double *__restrict__ a, int nb,
                           const int *__restrict__ skip) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  rr0[b] = rr1[b];
  // skip: every first residual below the early-exit bound, so the result
  // stays x = g (the reference returns before its first step)
  a[b] = (skip && *skip) ? 0.0 : (pap[b] != 0.0 ? rr1[b] / pap[b] : 0.0);
}

// flag = all(v[b] < bound) over nb values (one block); the hprsdft early
// exit all(rr1 < 1e-24) (cgpc_hprsdft_module.f90:160) decided on the device
__global__ void k_all_below(const double *__restrict__ v, int nb, double bound, int *__restrict__ flag) {
  __shared__ int any_above;
  if (threadIdx.x == 0) any_above = 0;
  __syncthreads();
  int above = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) above |= !(v[b] < bound);
  if (above) atomicOr(&any_above, 1);
  __syncthreads();
  if (threadIdx.x == 0) *flag = any_above ? 0 : 1;
}

// flag |= all(|E - E1| < 1e-15): the CG block's convergence test
// (cg_module.f90:236), sticky for the rest of the call
__global__ void k_cg_conv(const double *__restrict__ E, const double *__restrict__ E1, int nb,
                          int *__restrict__ flag) {
  __shared__ int any_moving;
  if (threadIdx.x == 0) any_moving = 0;
  __syncthreads();
  int moving = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) moving |= !(fabs(E[b] - E1[b]) < 1e-15);
  if (moving) atomicOr(&any_moving, 1);
  __syncthreads();
  if (threadIdx.x == 0 && !any_moving) *flag = 1;
}

__global__ void k_fill_int(int *__restrict__ a, int v, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

// rr1/rr0 with a zero previous residual (a band already at rest) taken as 0
__device__ inline double pc_beta(double rr1, double rr0) { return rr0 != 0.0 ? rr1 / rr0 : 0.0; }

// x += a p ; optionally r -= a f
template <typename W>
__global__ void k_pc_step(W *__restrict__ x, const W *__restrict__ p, W *__restrict__ r,
                          const W *__restrict__ f, const double *__restrict__ a, int update_r, int64_t N) {
  const int b = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t k = int64_t(b) * N + i;
  const W s = W(a[b]);
  x[k] += s * p[k];
  if (update_r) r[k] -= s * f[k];
}

// p = r + (rr1/rr0) p
template <typename W>
__global__ void k_pc_dir(const W *__restrict__ r, W *__restrict__ p, const double *__restrict__ rr1,
                         const double *__restrict__ rr0, int64_t N) {
  const int b = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t k = int64_t(b) * N + i;
  p[k] = r[k] + W(pc_beta(rr1[b], rr0[b])) * p[k];
}

// ---- hprsdft preconditioner, z-streaming passes ----
// The same mloop-step CG on the 7-point operator A = -1/2 lap_7 as
// precond_hprsdft, one pass per step.  A block owns kThreads consecutive
// points of an xy-plane and a run of z-planes of one band; each thread
// walks its z-column holding the column's values at z-1, z and z+1 in
// registers, so every stencil input is read from HBM once, and takes the
// four in-plane neighbours from its neighbours' loads of the same plane
// (L1/L2).  A p is never stored: the step pass recomputes it from p, and
// the direction pass evaluates A p' for p' = r + beta p from r and p
// (cgpc_hprsdft_module.f90:193-197 applies precond_cg_mat to the new
// direction the same way).  The new direction is written to the second
// buffer, since other blocks still read the old one as neighbours; the
// host swaps the two.  Per-band dot products: one partial per block,
// summed by k_reduce_partials.

enum PzOp { kPzFirst, kPzPap, kPzDir, kPzStepF, kPzStepM, kPzStepL, kPzStepFL };

struct PzGeom {
  int n1, n2, n3;
  int64_t N, n12;
  int64_t ld;  // leading dimension of g and x (the work vectors r, p use N)
  int nq, nzr, zlen, nseg;
};

// block sum of v into partial[b * nseg + s]
__device__ inline void block_partial(double v, double *__restrict__ partial, int b, int s, int nseg) {
  __shared__ double red[kThreads / 64];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kThreads / 64; ++k) t += red[k];
    partial[int64_t(b) * nseg + s] = t;
  }
}

// Pointers of one pass: the stencil input is g (first), p (pap, steps) or
// r + beta p (dir); x accumulates P g (it is g before the first step).
// S is the storage type of the internal vectors r, p and the new direction
// (W, or bfloat16 bits for W = float); g and x are always W.
template <typename W, typename S = W> struct PzArgs {
  const W *g;
  W *x;
  S *r;
  const S *p;
  S *pw;  // first: p = r; dir: the new direction
  const double *a, *rr1, *rr0;
  double *partial;
};

template <typename T> struct TypeTag {
  typedef T type;
};

// Storage conversions of the internal preconditioner vectors: identity for
// S = W; for bfloat16 bits (uint16_t) with W = float, the upper half of the
// float on load and round-to-nearest-even on store.
template <typename W, typename S> struct PzStore {
  __device__ static W ld(S v) { return W(v); }
  __device__ static S st(W v) { return S(v); }
};
template <> struct PzStore<float, uint16_t> {
  __device__ static float ld(uint16_t v) { return __uint_as_float(uint32_t(v) << 16); }
  __device__ static uint16_t st(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t((u >> 16) | ((u & 0x7fffffu) ? 0x40u : 0u));
    return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
  }
};

// Raw values of one plane the thread needs: the four in-plane neighbours
// of the stencil input (r and p separately for the direction pass, which
// combines them where they are used) and the pointwise streams (g or x,
// and r).
template <typename W> struct PzPlane {
  W n[4], m[4];  // neighbours x+, x-, y+, y- of the input (m: p for the direction pass)
  W u, v;        // pointwise: first stream (g or x), r
};

template <typename W, int OP>
__global__ __launch_bounds__(kThreads) void k_pcz(PzArgs<W> A, PzGeom q, W c0, W c1, W c2, W c3) {
  constexpr bool kStencil = OP != kPzStepL && OP != kPzStepFL;
  constexpr bool kStep = OP == kPzStepF || OP == kPzStepM;
  // logical block: plane chunk fastest, then z-run, then band; the 8 XCDs
  // each take one contiguous range, so neighbouring chunks share an L2
  const unsigned nblk = gridDim.x, xc = blockIdx.x % 8;
  const unsigned lb = xc * (nblk / 8) + min(xc, nblk % 8) + blockIdx.x / 8;
  const int qb = int(lb % unsigned(q.nq));
  const int zr = int((lb / unsigned(q.nq)) % unsigned(q.nzr));
  const int b = int(lb / unsigned(q.nq * q.nzr));
  const int64_t qi = int64_t(qb) * kThreads + threadIdx.x;
  const bool act = qi < q.n12;
  const int64_t qc = act ? qi : 0;
  const int i1 = int(qc % q.n1), i2 = int(qc / q.n1);
  const int64_t nb4[4] = {i1 == q.n1 - 1 ? qc - (q.n1 - 1) : qc + 1, i1 == 0 ? qc + q.n1 - 1 : qc - 1,
                          i2 == q.n2 - 1 ? qc - int64_t(q.n2 - 1) * q.n1 : qc + q.n1,
                          i2 == 0 ? qc + int64_t(q.n2 - 1) * q.n1 : qc - q.n1};
  const int z0 = zr * q.zlen, z1 = min(z0 + q.zlen, q.n3);
  const int64_t off = int64_t(b) * q.N;
  const W *gb = A.g + int64_t(b) * q.ld, *pb = A.p + off;
  W *xb = A.x + int64_t(b) * q.ld, *rb = A.r + off;
  const W al = (kStep || OP == kPzStepL || OP == kPzStepFL) ? W(A.a[b]) : W(0);
  const W beta = OP == kPzDir ? W(pc_beta(A.rr1[b], A.rr0[b])) : W(0);
  const W *src = OP == kPzFirst ? gb : (OP == kPzDir ? rb : pb);  // stencil input (r for the direction pass)
  auto wrapz = [&](int z) { return ((z % q.n3) + q.n3) % q.n3; };  // any n3 >= 1
  double acc = 0.0;
  if constexpr (kStencil) {
    // loads of a plane are issued one iteration before its use (the centre
    // column two), ahead of the stores of the current plane
    auto load = [&](int z, PzPlane<W> &P) {
      const int64_t o = int64_t(z) * q.n12;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        P.n[t] = src[o + nb4[t]];
        if (OP == kPzDir) P.m[t] = pb[o + nb4[t]];
      }
      if (kStep) {
        P.u = OP == kPzStepF ? gb[o + qc] : xb[o + qc];
        P.v = rb[o + qc];
      }
    };
    auto val = [&](W a, W m) -> W { return OP == kPzDir ? std::fma(beta, m, a) : a; };
    W cm = src[int64_t(wrapz(z0 - 1)) * q.n12 + qc], c0v = src[int64_t(z0) * q.n12 + qc];
    W cp = src[int64_t(wrapz(z0 + 1)) * q.n12 + qc];
    W dm = W(0), d0 = W(0), dp = W(0);  // p of the centre column (direction pass)
    if (OP == kPzDir) {
      dm = pb[int64_t(wrapz(z0 - 1)) * q.n12 + qc];
      d0 = pb[int64_t(z0) * q.n12 + qc];
      dp = pb[int64_t(wrapz(z0 + 1)) * q.n12 + qc];
    }
    PzPlane<W> cur, nxt;
    load(z0, nxt);
    for (int z = z0; z < z1; ++z) {
      cur = nxt;
      load(wrapz(z + 1), nxt);
      const int64_t o2 = int64_t(wrapz(z + 2)) * q.n12 + qc;
      const W cn = src[o2];
      const W dn = OP == kPzDir ? pb[o2] : W(0);
      const W v0 = val(c0v, d0);
      const W av = c0 * v0 + c1 * (val(cur.n[0], cur.m[0]) + val(cur.n[1], cur.m[1])) +
                   c2 * (val(cur.n[2], cur.m[2]) + val(cur.n[3], cur.m[3])) + c3 * (val(cp, dp) + val(cm, dm));
      const int64_t k = int64_t(z) * q.n12 + qc;
      if (act) {
        if (OP == kPzFirst) {  // r = g - A g ; p = r
          const W rv = v0 - av;
          rb[k] = rv;
          A.pw[off + k] = rv;
          acc += double(rv) * double(rv);
        } else if (OP == kPzPap) {
          acc += double(v0) * double(av);
        } else if (OP == kPzDir) {  // p' = r + beta p (written apart)
          A.pw[off + k] = v0;
          acc += double(v0) * double(av);
        } else {  // steps before the last: x += a p ; r -= a A p
          const W xv = cur.u + al * v0;
          xb[k] = xv;
          const W rv = cur.v - al * av;
          rb[k] = rv;
          acc += double(rv) * double(rv);
        }
      }
      cm = c0v;
      c0v = cp;
      cp = cn;
      dm = d0;
      d0 = dp;
      dp = dn;
    }
  } else {  // last step: x += a p ; partial <x, g>
    W gn = gb[int64_t(z0) * q.n12 + qc], pn = pb[int64_t(z0) * q.n12 + qc];
    W xn = OP == kPzStepFL ? W(0) : xb[int64_t(z0) * q.n12 + qc];
    for (int z = z0; z < z1; ++z) {
      const W gv = gn, pv = pn, xo = xn;
      const int64_t o1 = int64_t(wrapz(z + 1)) * q.n12 + qc;
      gn = gb[o1];
      pn = pb[o1];
      if (OP != kPzStepFL) xn = xb[o1];
      const int64_t k = int64_t(z) * q.n12 + qc;
      if (act) {
        const W xv = (OP == kPzStepFL ? gv : xo) + al * pv;
        xb[k] = xv;
        acc += double(xv) * double(gv);
      }
    }
  }
  block_partial(acc, A.partial, b, zr * q.nq + qb, q.nseg);
}

// VP x-points per thread (n1 and the leading dimension multiples of VP):
// the same passes with VP-wide vector loads and stores, 1/VP of the load
// instructions per point.  The in-plane neighbours of the group are its two
// x-outer points (scalar loads) and the y-rows above and below (vector
// loads); each point's stencil sum is formed in the same order as in k_pcz.
template <typename W, int VP> struct alignas(VP * sizeof(W)) PzVec {
  W v[VP];
};

template <typename W, int OP, int VP, typename S = W>
__global__ __launch_bounds__(kThreads) void k_pczv(PzArgs<W, S> A, PzGeom q, W c0, W c1, W c2, W c3) {
  typedef PzVec<W, VP> V;
  typedef PzVec<S, VP> VS;
  typedef PzStore<W, S> St;
  constexpr bool kStencil = OP != kPzStepL && OP != kPzStepFL;
  constexpr bool kStep = OP == kPzStepF || OP == kPzStepM;
  const unsigned nblk = gridDim.x, xc = blockIdx.x % 8;
  const unsigned lb = xc * (nblk / 8) + min(xc, nblk % 8) + blockIdx.x / 8;
  const int qb = int(lb % unsigned(q.nq));
  const int zr = int((lb / unsigned(q.nq)) % unsigned(q.nzr));
  const int b = int(lb / unsigned(q.nq * q.nzr));
  const int hv = q.n1 / VP;  // groups per row
  const int64_t qi = int64_t(qb) * kThreads + threadIdx.x;
  const bool act = qi < q.n12 / VP;
  const int64_t qc = act ? qi : 0;
  const int x = VP * int(qc % hv), y = int(qc / hv);
  const int64_t e0 = x + int64_t(y) * q.n1;  // first point of the group in its plane
  const int64_t exl = (x == 0 ? q.n1 - 1 : x - 1) + int64_t(y) * q.n1;
  const int64_t exr = (x + VP == q.n1 ? 0 : x + VP) + int64_t(y) * q.n1;
  const int64_t eyu = x + int64_t(y == q.n2 - 1 ? 0 : y + 1) * q.n1;
  const int64_t eyd = x + int64_t(y == 0 ? q.n2 - 1 : y - 1) * q.n1;
  // the groups left and right of this one (periodic in x) are the groups
  // of lanes sl_l and sr_l when those lie in the same wave: their combined
  // centre values are taken by lane shuffle, and only the other lanes load
  const int lane = int(threadIdx.x & 63);
  const int sl_l = lane + (x == 0 ? hv - 1 : -1), sr_l = lane + (x + VP == q.n1 ? 1 - hv : 1);
  const bool sl_in = sl_l >= 0 && sl_l < 64, sr_in = sr_l >= 0 && sr_l < 64;
  const int sl_src = sl_in ? sl_l : lane, sr_src = sr_in ? sr_l : lane;
  // the same for the rows above and below (periodic in y) when a row holds
  // at most half a wave of groups, so that most lanes find them in-wave
  const bool ysh = hv <= 32;
  const int yu_l = lane + (y == q.n2 - 1 ? (1 - q.n2) * hv : hv), yd_l = lane + (y == 0 ? (q.n2 - 1) * hv : -hv);
  const bool yu_in = yu_l >= 0 && yu_l < 64, yd_in = yd_l >= 0 && yd_l < 64;
  const int yu_src = yu_in ? yu_l : lane, yd_src = yd_in ? yd_l : lane;
  const int z0 = zr * q.zlen, z1 = min(z0 + q.zlen, q.n3);
  const int64_t off = int64_t(b) * q.N;
  const W *gb = A.g + int64_t(b) * q.ld;
  const S *pb = A.p + off;
  W *xb = A.x + int64_t(b) * q.ld;
  S *rb = A.r + off;
  const W al = (kStep || OP == kPzStepL || OP == kPzStepFL) ? W(A.a[b]) : W(0);
  const W beta = OP == kPzDir ? W(pc_beta(A.rr1[b], A.rr0[b])) : W(0);
  // W-typed vector and scalar loads of a W or S stream
  auto ldw = [](const W *base, int64_t e) { return *reinterpret_cast<const V *>(base + e); };
  auto lds = [](const S *base, int64_t e) {
    const VS t = *reinterpret_cast<const VS *>(base + e);
    V v;
#pragma unroll
    for (int i = 0; i < VP; ++i) v.v[i] = St::ld(t.v[i]);
    return v;
  };
  // stencil input: g (first), r (direction, with p beside it) or p
  auto ld_src = [&](int64_t e) {
    if constexpr (OP == kPzFirst)
      return ldw(gb, e);
    else
      return lds(OP == kPzDir ? rb : pb, e);
  };
  auto src_at = [&](int64_t e) -> W {
    if constexpr (OP == kPzFirst)
      return gb[e];
    else
      return St::ld((OP == kPzDir ? rb : pb)[e]);
  };
  auto stw = [](W *base, int64_t e, const W *w) {
    V v;
#pragma unroll
    for (int i = 0; i < VP; ++i) v.v[i] = w[i];
    *reinterpret_cast<V *>(base + e) = v;
  };
  // stores w rounded to S, and leaves the rounded values in w
  auto sts = [](S *base, int64_t e, W *w) {
    VS v;
#pragma unroll
    for (int i = 0; i < VP; ++i) {
      v.v[i] = St::st(w[i]);
      w[i] = St::ld(v.v[i]);
    }
    *reinterpret_cast<VS *>(base + e) = v;
  };
  auto comb = [&](W r, W m) -> W { return OP == kPzDir ? std::fma(beta, m, r) : r; };
  auto wrapz = [&](int z) { return ((z % q.n3) + q.n3) % q.n3; };
  double acc = 0.0;
  if constexpr (kStencil) {
    // ring of the group at z-1, z, z+1 (raw input; p alongside for the direction pass)
    V cm = ld_src(int64_t(wrapz(z0 - 1)) * q.n12 + e0), cz = ld_src(int64_t(z0) * q.n12 + e0);
    V cp = ld_src(int64_t(wrapz(z0 + 1)) * q.n12 + e0);
    V dm{}, dz{}, dp{};
    if (OP == kPzDir) {
      dm = lds(pb, int64_t(wrapz(z0 - 1)) * q.n12 + e0);
      dz = lds(pb, int64_t(z0) * q.n12 + e0);
      dp = lds(pb, int64_t(wrapz(z0 + 1)) * q.n12 + e0);
    }
    int zn = wrapz(z0 + 1);
    for (int z = z0; z < z1; ++z) {
      const int64_t o = int64_t(z) * q.n12;
      zn = zn + 1 == q.n3 ? 0 : zn + 1;  // plane z + 2
      const V cn = ld_src(int64_t(zn) * q.n12 + e0);
      V dn{};
      if (OP == kPzDir) dn = lds(pb, int64_t(zn) * q.n12 + e0);
      W v[VP], wu[VP], wd[VP], av[VP];
#pragma unroll
      for (int i = 0; i < VP; ++i) v[i] = comb(cz.v[i], dz.v[i]);
      auto ldc = [&](int64_t e, W *w) {  // combined input of the group at e
        const V a = ld_src(e);
        V m{};
        if (OP == kPzDir) m = lds(pb, e);
#pragma unroll
        for (int i = 0; i < VP; ++i) w[i] = comb(a.v[i], m.v[i]);
      };
      // every lane of the wave is at plane z here (z0, z1 are block-uniform)
      W sl = __shfl(v[VP - 1], sl_src, 64), sr = __shfl(v[0], sr_src, 64);
      if (!sl_in) sl = comb(src_at(o + exl), OP == kPzDir ? St::ld(pb[o + exl]) : W(0));
      if (!sr_in) sr = comb(src_at(o + exr), OP == kPzDir ? St::ld(pb[o + exr]) : W(0));
      if (ysh) {  // uniform over the launch
#pragma unroll
        for (int i = 0; i < VP; ++i) {
          wu[i] = __shfl(v[i], yu_src, 64);
          wd[i] = __shfl(v[i], yd_src, 64);
        }
        if (!yu_in) ldc(o + eyu, wu);
        if (!yd_in) ldc(o + eyd, wd);
      } else {
        ldc(o + eyu, wu);
        ldc(o + eyd, wd);
      }
#pragma unroll
      for (int i = 0; i < VP; ++i) {
        const W xp = i + 1 < VP ? v[i + 1 < VP ? i + 1 : i] : sr;
        const W xm = i > 0 ? v[i > 0 ? i - 1 : i] : sl;
        av[i] = c0 * v[i] + c1 * (xp + xm) + c2 * (wu[i] + wd[i]) + c3 * (comb(cp.v[i], dp.v[i]) + comb(cm.v[i], dm.v[i]));
      }
      const int64_t k = o + e0;
      if (act) {
        if (OP == kPzFirst) {  // r = g - A g ; p = r
          W rv[VP];
#pragma unroll
          for (int i = 0; i < VP; ++i) rv[i] = v[i] - av[i];
          sts(rb, k, rv);
          sts(A.pw + off, k, rv);
#pragma unroll
          for (int i = 0; i < VP; ++i) acc += double(rv[i]) * double(rv[i]);
        } else if (OP == kPzPap) {
#pragma unroll
          for (int i = 0; i < VP; ++i) acc += double(v[i]) * double(av[i]);
        } else if (OP == kPzDir) {  // p' = r + beta p (written apart)
          sts(A.pw + off, k, v);
#pragma unroll
          for (int i = 0; i < VP; ++i) acc += double(v[i]) * double(av[i]);
        } else {  // steps before the last: x += a p ; r -= a A p
          const V u = ldw(OP == kPzStepF ? gb : xb, k), rr = lds(rb, k);
          W xv[VP], rv[VP];
#pragma unroll
          for (int i = 0; i < VP; ++i) {
            xv[i] = u.v[i] + al * v[i];
            rv[i] = rr.v[i] - al * av[i];
          }
          stw(xb, k, xv);
          sts(rb, k, rv);
#pragma unroll
          for (int i = 0; i < VP; ++i) acc += double(rv[i]) * double(rv[i]);
        }
      }
      cm = cz;
      cz = cp;
      cp = cn;
      dm = dz;
      dz = dp;
      dp = dn;
    }
  } else {  // last step: x += a p ; partial <x, g>
    for (int z = z0; z < z1; ++z) {
      const int64_t k = int64_t(z) * q.n12 + e0;
      if (act) {
        const V gv = ldw(gb, k), pv = lds(pb, k);
        const V xo = OP == kPzStepFL ? gv : ldw(xb, k);
        W xv[VP];
#pragma unroll
        for (int i = 0; i < VP; ++i) {
          xv[i] = xo.v[i] + al * pv.v[i];
          acc += double(xv[i]) * double(gv.v[i]);
        }
        stw(xb, k, xv);
      }
    }
  }
  block_partial(acc, A.partial, b, zr * q.nq + qb, q.nseg);
}

// A <- s (A + A^T) in place (both triangles), for an n x n matrix
__global__ void k_symmetrize(double *__restrict__ A, int n, double s) {
  const int i = blockIdx.x * 16 + threadIdx.x, j = blockIdx.y * 16 + threadIdx.y;
  if (i >= n || j >= n || i > j) return;
  const double v = s * (A[i + int64_t(n) * j] + A[j + int64_t(n) * i]);
  A[i + int64_t(n) * j] = v;
  A[j + int64_t(n) * i] = v;
}

inline dim3 grid_nb(int64_t N, int nb) { return dim3(unsigned((N + kThreads - 1) / kThreads), nb); }

struct Scratch {
  // per-band device scalars, each nb doubles
  double *E, *E1, *rb, *gkgk, *res, *d6, *uc, *rr0, *rr1, *a, *pap;
  int *accept;
  // device-side decisions: flags[0] the CG block has converged, flags[1]
  // the preconditioner's first residuals are all below its exit bound
  int *flags;
};

int ensure_scalars(rsg_ctx *ctx, int nb, int nseg, Scratch &s) {
  const size_t per = size_t(nb);
  (void)nseg;
  const size_t nd = per * (9 + 6 + 2);
  if (ctx->scal.ensure(nd * sizeof(double) + per * sizeof(int) + 256) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "solver: scalar scratch allocation failed");
  double *p = ctx->scal.as<double>();
  s.E = p; p += per;
  s.E1 = p; p += per;
  s.rb = p; p += per;
  s.gkgk = p; p += per;
  s.res = p; p += per;
  s.rr0 = p; p += per;
  s.rr1 = p; p += per;
  s.a = p; p += per;
  s.pap = p; p += per;
  s.d6 = p; p += 6 * per;
  s.uc = p; p += 2 * per;
  s.accept = reinterpret_cast<int *>(p);
  s.flags = s.accept + per;  // 4 ints inside the 256-byte tail
  return RSG_OK;
}

int seg_count(int64_t N, int nb) {
  int nseg = 1;
  // enough blocks to fill the chip (8 four-wave blocks per CU), each
  // segment at least 4096 points
  while (nseg < 64 && int64_t(nseg) * nb < 4096 && N / (nseg * 2) >= 4096) nseg *= 2;
  return nseg;
}

}  // namespace

namespace {

template <typename W>
int multidot_t(rsg_ctx *ctx, int npair, const W *const *x, const W *const *y, int nb, double alpha, double *out) {
  if (npair < 1 || npair > 6) return ctx->fail(RSG_ERR_ARG, "multidot: 1..6 pairs");
  if (nb == 0) return RSG_OK;
  const int64_t N = ctx->N;
  const int nseg = seg_count(N, nb);
  const int64_t seglen = (N + nseg - 1) / nseg;
  if (ctx->dotpart.ensure(size_t(npair) * nb * nseg * sizeof(double)) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "multidot: scratch allocation failed");
  double *partial = ctx->dotpart.as<double>();
  DotArgs<W> a{};
  for (int q = 0; q < npair; ++q) {
    a.x[q] = x[q];
    a.y[q] = y[q];
  }
  dim3 grid(nseg, nb);
  switch (npair) {
#define RSG_DOT_CASE(P)                                                                                  \
  case P:                                                                                                \
    hipLaunchKernelGGL((k_multidot<P, W>), grid, dim3(kThreads), 0, ctx->stream, a, N, nb, nseg, seglen, \
                       partial);                                                                         \
    break;
    RSG_DOT_CASE(1)
    RSG_DOT_CASE(2)
    RSG_DOT_CASE(3)
    RSG_DOT_CASE(4)
    RSG_DOT_CASE(5)
    RSG_DOT_CASE(6)
#undef RSG_DOT_CASE
  }
  const int count = npair * nb;
  hipLaunchKernelGGL(k_reduce_partials, dim3((count + 255) / 256), dim3(256), 0, ctx->stream, partial,
                     nseg, count, alpha, out);
  return RSG_OK;
}

}  // namespace

int multidot(rsg_ctx *ctx, int npair, const double *const *x, const double *const *y, int nb,
             double alpha, double *out) {
  return multidot_t<double>(ctx, npair, x, y, nb, alpha, out);
}

namespace {

// out[q*nb + b] = alpha * the six Ritz-step products of band b
template <typename W>
int ritz_dots(rsg_ctx *ctx, const W *x, const W *p, const W *hx, const W *hp, int nb, double alpha, double *out) {
  if (nb == 0) return RSG_OK;
  const int64_t N = ctx->N;
  const int nseg = seg_count(N, nb);
  const int64_t seglen = (N + nseg - 1) / nseg;
  if (ctx->dotpart.ensure(size_t(6) * nb * nseg * sizeof(double)) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "ritz_dots: scratch allocation failed");
  double *partial = ctx->dotpart.as<double>();
  hipLaunchKernelGGL((k_ritz_dots<W>), dim3(nseg, nb), dim3(kThreads), 0, ctx->stream, x, p, hx, hp, N, nb, nseg,
                     seglen, partial);
  const int count = 6 * nb;
  hipLaunchKernelGGL(k_reduce_partials, dim3((count + 255) / 256), dim3(256), 0, ctx->stream, partial, nseg, count,
                     alpha, out);
  return RSG_OK;
}

}  // namespace

int multidot(rsg_ctx *ctx, int npair, const float *const *x, const float *const *y, int nb, double alpha,
             double *out) {
  return multidot_t<float>(ctx, npair, x, y, nb, alpha, out);
}

namespace {

int check_store(rsg_ctx *ctx, const char *who) {
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, std::string(who) + ": grid not set");
  if (ctx->mb <= 0) return ctx->fail(RSG_ERR_STATE, std::string(who) + ": no band store");
  return RSG_OK;
}

int read_info(rsg_ctx *ctx, int *host) {
  RSG_HIP(hipMemcpyAsync(host, ctx->info.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

// hprsdft preconditioner on an nn-band block (cgpc_hprsdft_module.f90:133-270):
// mloop CG steps on A x = g from x0 = pg (= g on entry).
// With diag non-null the operator gains a diagonal term, A = -1/2 lap_7 +
// diag (the shifted, potential-aware solve that trains the learned
// preconditioner).
template <typename W>
int precond_hprsdft(rsg_ctx *ctx, int mloop, const W *g, W *pg, int nn, W *r, W *p, W *f, Scratch &s,
                    const float *diag = nullptr) {
  if (mloop <= 0) return RSG_OK;
  const int64_t N = ctx->N;
  const double c0 = ctx->ggg[0] / (ctx->hgrid[0] * ctx->hgrid[0]) +
                    ctx->ggg[1] / (ctx->hgrid[1] * ctx->hgrid[1]) +
                    ctx->ggg[2] / (ctx->hgrid[2] * ctx->hgrid[2]);
  const double c1 = -0.5 / (ctx->hgrid[0] * ctx->hgrid[0]) * ctx->ggg[0];
  const double c2 = -0.5 / (ctx->hgrid[1] * ctx->hgrid[1]) * ctx->ggg[1];
  const double c3 = -0.5 / (ctx->hgrid[2] * ctx->hgrid[2]) * ctx->ggg[2];
  const dim3 gN = grid_nb(N, nn), blk(kThreads);
  const int64_t tot = N * nn;
  hipLaunchKernelGGL((k_pc_mat<W>), gN, blk, 0, ctx->stream, pg, f, ctx->n[0], ctx->n[1], ctx->n[2], W(c0), W(c1),
                     W(c2), W(c3), diag);
  hipLaunchKernelGGL((k_pc_init<W>), dim3(unsigned((tot + kThreads - 1) / kThreads)), blk, 0, ctx->stream, g, f, r,
                     p, tot);
  {
    const W *xs[1] = {r}, *ys[1] = {r};
    int rc = multidot(ctx, 1, xs, ys, nn, 1.0, s.rr1);
    if (rc) return rc;
  }
  std::vector<double> h(static_cast<size_t>(nn));
  RSG_HIP(hipMemcpyAsync(h.data(), s.rr1, nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  bool small = true;
  for (double v : h) small = small && (v < 1e-24);
  if (small) return RSG_OK;
  const dim3 gb((nn + 63) / 64), b64(64);
  for (int it = 1; it <= mloop; ++it) {
    hipLaunchKernelGGL((k_pc_mat<W>), gN, blk, 0, ctx->stream, p, f, ctx->n[0], ctx->n[1], ctx->n[2], W(c0),
                       W(c1), W(c2), W(c3), diag);
    const W *xs[1] = {p}, *ys[1] = {f};
    int rc = multidot(ctx, 1, xs, ys, nn, 1.0, s.pap);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pc_alpha, gb, b64, 0, ctx->stream, s.rr1, s.pap, s.rr0, s.a, nn, nullptr);
    const int last = it == mloop;
    hipLaunchKernelGGL((k_pc_step<W>), gN, blk, 0, ctx->stream, pg, p, r, f, s.a, last ? 0 : 1, N);
    if (last) break;
    const W *xr[1] = {r}, *yr[1] = {r};
    rc = multidot(ctx, 1, xr, yr, nn, 1.0, s.rr1);
    if (rc) return rc;
    hipLaunchKernelGGL((k_pc_dir<W>), gN, blk, 0, ctx->stream, r, p, s.rr1, s.rr0, N);
  }
  RSG_HIP(hipGetLastError());
  return RSG_OK;
}


// precond_hprsdft in z-streaming passes for the CG step (no diagonal
// shift): pg = P g and rb = dv <pg, g>.  The same iteration as
// precond_hprsdft; f is the second direction buffer.
template <typename W>
int precond_hprsdft_fused(rsg_ctx *ctx, int mloop, const W *g, W *pg, int nn, W *r, W *p, W *f, Scratch &s,
                          double dv, double *rb, int64_t ld = 0) {
  const int64_t N = ctx->N;
  if (ld <= 0) ld = N;
  auto plain = [&]() -> int {  // pg = g, rb = dv <g, g> (rb only for ld = N)
    RSG_HIP(hipMemcpy2DAsync(pg, size_t(ld) * sizeof(W), g, size_t(ld) * sizeof(W), size_t(N) * sizeof(W), nn,
                             hipMemcpyDeviceToDevice, ctx->stream));
    if (!rb || ld != N) return RSG_OK;
    const W *xs[1] = {pg}, *ys[1] = {g};
    return multidot(ctx, 1, xs, ys, nn, dv, rb);
  };
  if (!rb) rb = s.rb;
  if (mloop <= 0) return plain();
  const double c0 = ctx->ggg[0] / (ctx->hgrid[0] * ctx->hgrid[0]) +
                    ctx->ggg[1] / (ctx->hgrid[1] * ctx->hgrid[1]) +
                    ctx->ggg[2] / (ctx->hgrid[2] * ctx->hgrid[2]);
  const W C0 = W(c0), C1 = W(-0.5 / (ctx->hgrid[0] * ctx->hgrid[0]) * ctx->ggg[0]);
  const W C2 = W(-0.5 / (ctx->hgrid[1] * ctx->hgrid[1]) * ctx->ggg[1]);
  const W C3 = W(-0.5 / (ctx->hgrid[2] * ctx->hgrid[2]) * ctx->ggg[2]);
  PzGeom q;
  q.n1 = ctx->n[0];
  q.n2 = ctx->n[1];
  q.n3 = ctx->n[2];
  q.N = N;
  q.n12 = int64_t(q.n1) * q.n2;
  q.ld = ld;
  // VP points per thread where every band column starts on a group: n1 and
  // the leading dimension multiples of VP, caller buffers aligned to the
  // VP-wide vector (4 floats or 2 doubles = 16 bytes, else 2 floats)
  auto fits = [&](int v) {
    const uintptr_t va = uintptr_t(v) * sizeof(W);
    return q.n1 % v == 0 && ld % v == 0 && reinterpret_cast<uintptr_t>(g) % va == 0 &&
           reinterpret_cast<uintptr_t>(pg) % va == 0;
  };
  const int vp = (sizeof(W) == 4 && fits(4)) ? 4 : (fits(2) ? 2 : 1);
  q.nq = int((q.n12 / vp + kThreads - 1) / kThreads);
  // enough blocks to fill the chip; z-runs of at least 4 planes
  const int64_t per_run = int64_t(q.nq) * nn;
  q.nzr = int(std::min<int64_t>(std::max(1, q.n3 / 4), std::max<int64_t>(1, (4096 + per_run - 1) / per_run)));
  q.zlen = (q.n3 + q.nzr - 1) / q.nzr;
  q.nzr = (q.n3 + q.zlen - 1) / q.zlen;
  q.nseg = q.nq * q.nzr;
  if (int64_t(q.nseg) * nn > 0x7fffffff) return ctx->fail(RSG_ERR_UNSUPPORTED, "preconditioner: too many blocks");
  if (ctx->dotpart.ensure(size_t(nn) * q.nseg * sizeof(double)) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "preconditioner: scratch allocation failed");
  // S: storage of the internal vectors r, p and the new direction; with
  // ctx->pc_bf16 an FP32 call keeps them in bfloat16 (half their bytes; g,
  // P g and every product stay FP32, the dot products FP64)
  auto run = [&](auto stag) {
    using S = typename decltype(stag)::type;
    PzArgs<W, S> A;
    A.g = g;
    A.x = pg;
    A.r = reinterpret_cast<S *>(r);
    A.p = reinterpret_cast<S *>(p);
    A.pw = reinterpret_cast<S *>(p);
    A.a = s.a;
    A.rr1 = s.rr1;
    A.rr0 = s.rr0;
    A.partial = ctx->dotpart.as<double>();
    const dim3 grid(unsigned(int64_t(q.nseg) * nn)), blk(kThreads), gb((nn + 63) / 64), b64(64);
    const dim3 gr((nn + 255) / 256);
    auto reduce = [&](double alpha, double *out) {
      hipLaunchKernelGGL(k_reduce_partials, gr, dim3(256), 0, ctx->stream, A.partial, q.nseg, nn, alpha, out);
    };
    auto pass = [&](auto opc) {
      constexpr int OP = decltype(opc)::value;
      if (vp == 4)
        hipLaunchKernelGGL((k_pczv<W, OP, 4, S>), grid, blk, 0, ctx->stream, A, q, C0, C1, C2, C3);
      else if (vp == 2)
        hipLaunchKernelGGL((k_pczv<W, OP, 2, S>), grid, blk, 0, ctx->stream, A, q, C0, C1, C2, C3);
      else if constexpr (std::is_same<S, W>::value)
        hipLaunchKernelGGL((k_pcz<W, OP>), grid, blk, 0, ctx->stream, A, q, C0, C1, C2, C3);
    };
    pass(std::integral_constant<int, kPzFirst>{});
    reduce(1.0, s.rr1);
    // the early exit all(rr1 < 1e-24) decided on the device: every step then
    // takes a zero step, leaving P g = g, with no read-back to the host
    hipLaunchKernelGGL(k_all_below, dim3(1), dim3(256), 0, ctx->stream, s.rr1, nn, 1e-24, s.flags + 1);
    S *pcur = reinterpret_cast<S *>(p), *pnext = reinterpret_cast<S *>(f);
    for (int it = 1; it <= mloop; ++it) {
      A.p = pcur;
      if (it == 1) {
        pass(std::integral_constant<int, kPzPap>{});
      } else {
        A.pw = pnext;
        pass(std::integral_constant<int, kPzDir>{});
        std::swap(pcur, pnext);
        A.p = pcur;
      }
      reduce(1.0, s.pap);
      hipLaunchKernelGGL(k_pc_alpha, gb, b64, 0, ctx->stream, s.rr1, s.pap, s.rr0, s.a, nn, s.flags + 1);
      const bool first = it == 1, last = it == mloop;
      if (first && last)
        pass(std::integral_constant<int, kPzStepFL>{});
      else if (first)
        pass(std::integral_constant<int, kPzStepF>{});
      else if (last)
        pass(std::integral_constant<int, kPzStepL>{});
      else
        pass(std::integral_constant<int, kPzStepM>{});
      reduce(last ? dv : 1.0, last ? rb : s.rr1);
    }
  };
  if constexpr (std::is_same<W, float>::value) {
    if (ctx->pc_bf16 && vp >= 2)
      run(TypeTag<uint16_t>{});
    else
      run(TypeTag<W>{});
  } else {
    run(TypeTag<W>{});
  }
  RSG_HIP(hipGetLastError());
  return RSG_OK;
}

}  // namespace

namespace {
// out[t] = alpha * sum_c part[c * count + t]
__global__ void k_sum_batches(const double *__restrict__ part, int nbat, int count, double alpha,
                              double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  double v = 0.0;
  for (int c = 0; c < nbat; ++c) v += part[int64_t(c) * count + t];
  out[t] = alpha * v;
}

// C (nb x nb) = alpha A^T B with A, B column-major N x nb (leading dim N).
// rocBLAS runs this K = N >> nb shape on a handful of output tiles (one
// workgroup walks all of K when nb is small), so for nb <= 192 K is split
// into chunks: one strided-batched GEMM into per-chunk partial matrices,
// then a deterministic sum over the chunks.
// Points per K chunk for an (N, nb) product: the autotuned choice if there
// is one, else at most 256 chunks of >= 1024 points for nb <= 192 and one
// GEMM above; 0 = one GEMM.
int64_t gram_chunk(const rsg_ctx *ctx, int nb, int64_t N) {
  for (const auto &e : ctx->gram_split)
    if (e[0] == N && e[1] == nb) return e[2];
  if (nb > 192 || N < 4096) return 0;
  return std::max<int64_t>(1024, (N + 255) / 256);
}

int gemm_tn_gram(rsg_ctx *ctx, int nb, int64_t N, double alpha, const double *A, const double *B, double *C,
                 int64_t kc_force = -1) {
  const double zero = 0.0, one = 1.0;
  const int64_t kc = kc_force >= 0 ? kc_force : gram_chunk(ctx, nb, N);
  if (kc <= 0 || kc >= N) {
    RSG_BLAS_GUARDED(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, nb, nb, int(N), &alpha, A,
                           int(N), B, int(N), &zero, C, nb));
    return RSG_OK;
  }
  const int nfull = int(N / kc);
  const int64_t krem = N - int64_t(nfull) * kc;
  const int nbat = nfull + (krem > 0 ? 1 : 0);
  const size_t mat = size_t(nb) * nb;
  if (ctx->grampart.ensure(mat * nbat * sizeof(double)) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "gram product: scratch allocation failed");
  double *part = ctx->grampart.as<double>();
  RSG_BLAS_GUARDED(rocblas_dgemm_strided_batched(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, nb, nb,
                                         int(kc), &one, A, int(N), kc, B, int(N), kc, &zero, part, nb,
                                         rocblas_stride(mat), nfull));
  if (krem > 0)
    RSG_BLAS_GUARDED(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, nb, nb, int(krem), &one,
                           A + int64_t(nfull) * kc, int(N), B + int64_t(nfull) * kc, int(N), &zero,
                           part + mat * nfull, nb));
  const int count = int(mat);
  hipLaunchKernelGGL(k_sum_batches, dim3((count + 255) / 256), dim3(256), 0, ctx->stream, part, nbat, count, alpha, C);
  RSG_HIP(hipGetLastError());
  return RSG_OK;
}
// out[t] = alpha * sum_c part[c * count + t], FP32 partials summed in FP64
__global__ void k_sum_batches_f32(const float *__restrict__ part, int nbat, int count, double alpha,
                                  double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  double v = 0.0;
  for (int c = 0; c < nbat; ++c) v += double(part[int64_t(c) * count + t]);
  out[t] = alpha * v;
}

// C (nb x nb, FP64) = alpha A^T B for FP32 bands: SGEMMs over K chunks of at
// most 16384 points (FP32 sums stay short) into FP32 partial matrices, then
// the chunk sum in FP64.
int gemm_tn_gram_f32(rsg_ctx *ctx, int nb, int64_t N, double alpha, const float *A, const float *B, double *C) {
  const float zero = 0.f, one = 1.f;
  int64_t kc = gram_chunk(ctx, nb, N);
  if (kc <= 0 || kc > 16384) kc = std::min<int64_t>(N, 16384);
  const int nfull = int(N / kc);
  const int64_t krem = N - int64_t(nfull) * kc;
  const int nbat = nfull + (krem > 0 ? 1 : 0);
  const size_t mat = size_t(nb) * nb;
  if (ctx->grampart32.ensure(mat * nbat * sizeof(float)) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "gram product: scratch allocation failed");
  float *part = ctx->grampart32.as<float>();
  if (nfull > 0)
    RSG_BLAS_GUARDED(rocblas_sgemm_strided_batched(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, nb,
                                                   nb, int(kc), &one, A, int(N), kc, B, int(N), kc, &zero, part, nb,
                                                   rocblas_stride(mat), nfull));
  if (krem > 0)
    RSG_BLAS_GUARDED(rocblas_sgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, nb, nb, int(krem),
                                   &one, A + int64_t(nfull) * kc, int(N), B + int64_t(nfull) * kc, int(N), &zero,
                                   part + mat * nfull, nb));
  const int count = int(mat);
  hipLaunchKernelGGL(k_sum_batches_f32, dim3((count + 255) / 256), dim3(256), 0, ctx->stream, part, nbat, count, alpha,
                     C);
  RSG_HIP(hipGetLastError());
  return RSG_OK;
}
}  // namespace

// Time one GEMM over all of K against K-splits into 8..512 chunks of at
// least 512 points, for nb bands of the current grid (scratch bands, reps
// calls each), and keep the fastest for (N, nb).  The products are the same
// up to summation order.
int autotune_gram(rsg_ctx *ctx, int nb, int reps, float *best_ms, int64_t *best_kc) {
  DeviceOpGuard guard;
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "autotune_gram: grid not set");
  if (nb < 1 || reps < 1) return ctx->fail(RSG_ERR_ARG, "autotune_gram: nb and reps must be >= 1");
  const int64_t N = ctx->N;
  double *A = nullptr, *C = nullptr;
  if (hipMalloc(&A, size_t(N) * nb * 8) != hipSuccess) return ctx->fail(RSG_ERR_HIP, "autotune_gram: allocation failed");
  if (hipMalloc(&C, size_t(nb) * nb * 8) != hipSuccess) {
    (void)hipFree(A);
    return ctx->fail(RSG_ERR_HIP, "autotune_gram: allocation failed");
  }
  (void)hipMemsetAsync(A, 0x3f, size_t(N) * nb * 8, ctx->stream);
  std::vector<int64_t> cand{0};
  for (int c = 8; c <= 512; c *= 2)
    if (N / c >= 512) cand.push_back((N + c - 1) / c);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = RSG_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = RSG_ERR_HIP;
  float tbest = 0.f;
  int64_t kbest = 0;
  for (size_t i = 0; i < cand.size() && !rc; ++i) {
    rc = gemm_tn_gram(ctx, nb, N, 1.0, A, A, C, cand[i]);  // warm-up (and scratch sizing)
    if (!rc && hipEventRecord(e0, ctx->stream) != hipSuccess) rc = RSG_ERR_HIP;
    for (int r = 0; r < reps && !rc; ++r) rc = gemm_tn_gram(ctx, nb, N, 1.0, A, A, C, cand[i]);
    float ms = 0.f;
    if (!rc && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = RSG_ERR_HIP;
    if (!rc && (i == 0 || ms < tbest)) {
      tbest = ms;
      kbest = cand[i];
    }
  }
  (void)hipStreamSynchronize(ctx->stream);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(A);
  (void)hipFree(C);
  if (rc) return rc == RSG_ERR_HIP ? ctx->fail(rc, "autotune_gram: timing failed") : rc;
  bool found = false;
  for (auto &e : ctx->gram_split)
    if (e[0] == N && e[1] == nb) {
      e[2] = kbest;
      found = true;
    }
  if (!found) ctx->gram_split.push_back({N, int64_t(nb), kbest});
  if (best_ms) *best_ms = tbest / reps;
  if (best_kc) *best_kc = kbest;
  return RSG_OK;
}

namespace {
int gram_schmidt64(rsg_ctx *ctx, int n0, int n1) {
  int rc = RSG_OK;
  const int nb = n1 - n0 + 1;
  const int64_t N = ctx->N;
  double *U = ctx->unk.as<double>() + size_t(n0 - 1) * N;
  if (ctx->hsub.ensure(size_t(nb) * nb * 8) != hipSuccess || ctx->info.ensure(64) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "gram_schmidt: allocation failed");
  double *S = ctx->hsub.as<double>();
  const double dv = ctx->dv, zero = 0.0, one = 1.0;
  // S = dV U^T U, gram_schmidt_lusl_module.f90:268 (calc_dsyrk3), as a full
  // GEMM: rocBLAS's SYRK runs this K = N shape far below the MFMA GEMM
  rc = gemm_tn_gram(ctx, nb, N, dv, U, U, S);
  if (rc) return rc;
  // S = R^T R (:286, PDPOTRF)
  RSG_BLAS_GUARDED(rocsolver_dpotrf(ctx->blas, rocblas_fill_upper, nb, S, nb, ctx->info.as<rocblas_int>()));
  int info = 0;
  rc = read_info(ctx, &info);
  if (rc) return rc;
  if (info != 0)
    return ctx->fail(RSG_ERR_NOTPD, "gram_schmidt: overlap not positive definite (leading minor " +
                                        std::to_string(info) + ")");
  // U <- U R^-1 (:290 PDTRTRI + :313 calc_dtrmm3, done as one triangular solve)
  RSG_BLAS_GUARDED(rocblas_dtrsm(ctx->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                         rocblas_diagonal_non_unit, int(N), nb, &one, S, nb, U, int(N)));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return rc;
}
}  // namespace

// With the bands in the FP32 store, the orthonormalisation still runs in
// FP64 (the overlap and its Cholesky factor): the range goes to the FP64
// store, through gram_schmidt64, and back.
int gram_schmidt(rsg_ctx *ctx, int n0, int n1) {
  int rc = check_store(ctx, "gram_schmidt");
  if (rc) return rc;
  if (n0 < 1 || n1 > ctx->mb || n0 > n1) return ctx->fail(RSG_ERR_ARG, "gram_schmidt: band range outside 1..mb");
  if (!ctx->band32) return gram_schmidt64(ctx, n0, n1);
  const int64_t tot = ctx->N * int64_t(n1 - n0 + 1), o = ctx->N * int64_t(n0 - 1);
  const dim3 g(unsigned((tot + kThreads - 1) / kThreads)), b(kThreads);
  hipLaunchKernelGGL((k_to_double<float>), g, b, 0, ctx->stream, ctx->unk32.as<float>() + o, ctx->unk.as<double>() + o,
                     tot);
  if ((rc = gram_schmidt64(ctx, n0, n1))) return rc;
  hipLaunchKernelGGL((k_to_work<float>), g, b, 0, ctx->stream, ctx->unk.as<double>() + o, ctx->unk32.as<float>() + o,
                     tot);
  RSG_HIP(hipGetLastError());
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  ctx->unk_fresh = n0 == 1 && n1 == ctx->mb ? true : ctx->unk_fresh;
  return RSG_OK;
}

namespace {
// subspace_diag on the FP32 store: FP32 hpsi, the subspace matrix from FP32
// products summed in FP64, the FP64 eigensolve, and the rotation as an
// FP32 GEMM with the eigenvectors rounded to FP32.
int subspace_diag32(rsg_ctx *ctx, double *esp, double *Hs, double *D, double *Ework) {
  const int mb = ctx->mb;
  const int64_t N = ctx->N;
  if (ctx->hsub32.ensure(size_t(mb) * mb * 4) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "subspace_diag: allocation failed");
  float *U = ctx->unk32.as<float>(), *HU = ctx->work32.as<float>(), *V = ctx->hsub32.as<float>();
  int rc = launch_hpsi<float>(ctx, U, HU, N, mb);
  if (rc) return rc;
  if ((rc = gemm_tn_gram_f32(ctx, mb, N, 1.0, U, HU, Hs))) return rc;
  hipLaunchKernelGGL(k_symmetrize, dim3((mb + 15) / 16, (mb + 15) / 16), dim3(16, 16), 0, ctx->stream, Hs, mb,
                     0.5 * ctx->dv);
  RSG_BLAS_GUARDED(rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_upper, mb, Hs, mb, D, Ework,
                                    ctx->info.as<rocblas_int>()));
  int info = 0;
  if ((rc = read_info(ctx, &info))) return rc;
  if (info != 0) return ctx->fail(RSG_ERR_SOLVER, "subspace_diag: dsyevd did not converge");
  const int64_t mm = int64_t(mb) * mb;
  hipLaunchKernelGGL((k_to_work<float>), dim3(unsigned((mm + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                     ctx->stream, Hs, V, mm);
  const float zero = 0.f, one = 1.f;
  RSG_BLAS_GUARDED(rocblas_sgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, int(N), mb, mb, &one, U,
                                 int(N), V, mb, &zero, HU, int(N)));
  std::swap(ctx->unk32, ctx->work32);
  ctx->unk_fresh = false;
  RSG_HIP(hipMemcpyAsync(esp, D, mb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}
}  // namespace

int subspace_diag(rsg_ctx *ctx, double *esp) {
  int rc = check_store(ctx, "subspace_diag");
  if (rc) return rc;
  const int mb = ctx->mb;
  const int64_t N = ctx->N;
  if (ctx->hsub.ensure(size_t(mb) * mb * 8) != hipSuccess || ctx->esp_dev.ensure(size_t(mb) * 16) != hipSuccess ||
      ctx->info.ensure(64) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "subspace_diag: allocation failed");
  double *U = ctx->unk.as<double>(), *HU = ctx->work.as<double>(), *Hs = ctx->hsub.as<double>();
  double *D = ctx->esp_dev.as<double>(), *Ework = D + mb;
  if (ctx->band32) return subspace_diag32(ctx, esp, Hs, D, Ework);
  // psit = H U for all bands in one band-blocked call (subspace_diag_la_module.f90:79)
  rc = launch_hpsi<double>(ctx, U, HU, N, mb);
  if (rc) return rc;
  // Hsub = 1/2 dV (U^T HU + HU^T U) (:86, DSYR2K 'U','T') as A = U^T HU by
  // one GEMM, then Hsub = dV (A + A^T) / 2 (the SYR2K path of rocBLAS runs
  // this K = N shape far below the MFMA GEMM)
  const double zero = 0.0, one = 1.0;
  rc = gemm_tn_gram(ctx, mb, N, 1.0, U, HU, Hs);
  if (rc) return rc;
  hipLaunchKernelGGL(k_symmetrize, dim3((mb + 15) / 16, (mb + 15) / 16), dim3(16, 16), 0, ctx->stream, Hs, mb,
                     0.5 * ctx->dv);
  // eigen-decomposition, ascending (:155, DSYEV 'V','U')
  RSG_BLAS_GUARDED(rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_upper, mb, Hs, mb, D, Ework,
                            ctx->info.as<rocblas_int>()));
  int info = 0;
  rc = read_info(ctx, &info);
  if (rc) return rc;
  if (info != 0) return ctx->fail(RSG_ERR_SOLVER, "subspace_diag: dsyevd did not converge");
  // rotation U <- U V (:190) into the work buffer, which becomes the store
  RSG_BLAS_GUARDED(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, int(N), mb, mb, &one, U,
                         int(N), Hs, mb, &zero, HU, int(N)));
  std::swap(ctx->unk, ctx->work);
  RSG_HIP(hipMemcpyAsync(esp, D, mb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

int esp_calc(rsg_ctx *ctx, double *esp) {
  int rc = check_store(ctx, "esp_calc");
  if (rc) return rc;
  const int mb = ctx->mb;
  const int64_t N = ctx->N;
  if (ctx->esp_dev.ensure(size_t(mb) * 16) != hipSuccess) return ctx->fail(RSG_ERR_HIP, "esp_calc: allocation failed");
  if (ctx->band32 && !ctx->unk_fresh) {  // FP32 store: FP32 hpsi, FP64 dot products
    float *U = ctx->unk32.as<float>(), *HU = ctx->work32.as<float>();
    if ((rc = launch_hpsi<float>(ctx, U, HU, N, mb))) return rc;
    const float *xs[1] = {U}, *ys[1] = {HU};
    if ((rc = multidot(ctx, 1, xs, ys, mb, ctx->dv, ctx->esp_dev.as<double>()))) return rc;
  } else {
    double *U = ctx->unk.as<double>(), *HU = ctx->work.as<double>();
    rc = launch_hpsi<double>(ctx, U, HU, N, mb);
    if (rc) return rc;
    const double *xs[1] = {U}, *ys[1] = {HU};
    rc = multidot(ctx, 1, xs, ys, mb, ctx->dv, ctx->esp_dev.as<double>());
    if (rc) return rc;
  }
  RSG_HIP(hipMemcpyAsync(esp, ctx->esp_dev.p, mb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

namespace {

// cggs (cggs_module.f90:14-127) on the preconditioned gradients of store
// bands n0 .. n0+nn-1: C = dV U^T Pg over the whole store, rows masked per
// column by the SWGS rule, Pg <- Pg - U C (classical Gram-Schmidt: every
// coefficient from the incoming Pg, as the reference's two loops).  FP32
// blocks go through an FP64 copy in tmp (N x nn doubles).
template <typename W>
int cggs_block(rsg_ctx *ctx, int swgs, int n0, int nn, W *pg, double *tmp) {
  const int mb = ctx->mb;
  const int64_t N = ctx->N;
  if (ctx->cgcoef.ensure(size_t(mb) * nn * sizeof(double)) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "conjugate_gradient: cggs allocation failed");
  double *C = ctx->cgcoef.as<double>();
  const double *U = ctx->unk.as<double>();
  const int64_t tot = N * nn;
  const dim3 g1(unsigned((tot + kThreads - 1) / kThreads)), b1(kThreads);
  double *P = nullptr;
  if constexpr (sizeof(W) == 8) {
    P = reinterpret_cast<double *>(pg);
    (void)tmp;
  } else {
    P = tmp;
    hipLaunchKernelGGL((k_to_double<W>), g1, b1, 0, ctx->stream, pg, P, tot);
  }
  const double zero = 0.0, one = 1.0, mone = -1.0, dv = ctx->dv;
  RSG_BLAS_GUARDED(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, mb, nn, int(N), &dv, U,
                         int(N), P, int(N), &zero, C, mb));
  hipLaunchKernelGGL(k_cggs_mask, dim3((mb + 255) / 256, nn), dim3(256), 0, ctx->stream, C, mb, nn, n0, swgs);
  RSG_BLAS_GUARDED(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, int(N), nn, mb, &mone, U,
                         int(N), C, mb, &one, P, int(N)));
  if constexpr (sizeof(W) != 8)
    hipLaunchKernelGGL((k_to_work<W>), g1, b1, 0, ctx->stream, P, pg, tot);
  RSG_HIP(hipGetLastError());
  return RSG_OK;
}

// Key of a CG band block's kernel sequence: the block, the solver
// parameters, every device buffer address and the host-side launch
// parameters the kernels are built from.  A block replays a captured graph
// only while all of these are unchanged (a buffer that grows moves, and
// so changes the key); the buffers' contents (Vloc, bands) may change.
template <typename V> void key_put(std::vector<unsigned char> &k, const V &v) {
  const unsigned char *p = reinterpret_cast<const unsigned char *>(&v);
  k.insert(k.end(), p, p + sizeof(V));
}

template <typename T> void key_analytic(std::vector<unsigned char> &k, const AnalyticDev<T> &a) {
  for (const DevBuf *b : {&a.es, &a.ep, &a.dd, &a.amp, &a.scale, &a.pts, &a.seg_beg, &a.seg_end, &a.seg_grp,
                          &a.seg_order, &a.grp_seg_off, &a.hits, &a.tile_off, &a.tile_g, &a.rt_piece,
                          &a.rt_piece_off, &a.rt_slot_off, &a.rt_slot_gid, &a.grp_slot_off})
    key_put(k, b->p);
  for (int v : {a.ngroup, a.nseg, a.S, a.ntile, a.rt_ty, a.rt_zr, a.rt_nslot, a.rt_nslot_max}) key_put(k, v);
}

template <typename T> void key_tables(std::vector<unsigned char> &k, const NonlocalDev<T> &t) {
  for (const DevBuf *b : {&t.grp_lma0, &t.grp_nl, &t.grp_seg_off, &t.seg_beg, &t.seg_end, &t.seg_grp, &t.grp_pts,
                          &t.grp_coef, &t.lma_grp, &t.scale, &t.sl_off, &t.sl_len, &t.sl_lma, &t.sl_uv})
    key_put(k, b->p);
  for (int v : {t.ngroup, t.nseg}) key_put(k, v);
}

std::vector<unsigned char> cg_graph_key(const rsg_ctx *ctx, int wsize, int ns, int nn, int mbd, int ncg, int pc_kind,
                                        int mloop) {
  std::vector<unsigned char> k;
  for (int v : {wsize, ns, nn, mbd, ncg, pc_kind, mloop, ctx->n[0], ctx->n[1], ctx->n[2], ctx->md, ctx->nl_mode,
                int(ctx->have_ana), int(ctx->have_nl), ctx->tables.nzlma, int(ctx->pc_bf16), int(ctx->band32)})
    key_put(k, v);
  key_put(k, ctx->N);
  for (const DevBuf *b : {&ctx->unk, &ctx->unk32, &ctx->cgbuf, &ctx->scal, &ctx->dotpart, &ctx->cproj, &ctx->uvpart,
                          &ctx->uvunk, &ctx->vloc64, &ctx->vloc32, &ctx->dead})
    key_put(k, b->p);
  key_analytic(k, ctx->ana64);
  key_analytic(k, ctx->ana32);
  key_tables(k, ctx->nl64);
  key_tables(k, ctx->nl32);
  for (const HpsiShape *h : {&ctx->shape64, &ctx->shape32})
    for (int v : {h->ry, h->ty, h->zlen, h->chunk, h->bt, h->order, h->rx, h->proj, h->plds, h->psplit}) key_put(k, v);
  for (double v : ctx->coef_lap) key_put(k, v);
  for (double v : {ctx->coef_lap0, ctx->dv, ctx->hgrid[0], ctx->hgrid[1], ctx->hgrid[2], ctx->ggg[0], ctx->ggg[1],
                   ctx->ggg[2]})
    key_put(k, v);
  return k;
}

// Run the work `enqueue` queues on the context stream through a captured
// HIP graph: the first call of a key runs eagerly (its buffers reach their
// final size), the second captures and launches the graph, later ones
// replay it.  A capture the runtime refuses leaves the key eager.
template <typename F> int run_graphed(rsg_ctx *ctx, const std::vector<unsigned char> &key, F &&enqueue) {
  CgGraph *e = nullptr;
  for (auto &g : ctx->cg_graph_cache)
    if (g.key == key) e = &g;
  if (!e) {
    constexpr size_t kMaxGraphs = 16;
    if (ctx->cg_graph_cache.size() >= kMaxGraphs) {
      if (ctx->cg_graph_cache.front().exec) {
        // the oldest graph may still run on the stream
        RSG_HIP(hipStreamSynchronize(ctx->stream));
        DeviceOpGuard guard;
        (void)hipGraphExecDestroy(ctx->cg_graph_cache.front().exec);
      }
      ctx->cg_graph_cache.erase(ctx->cg_graph_cache.begin());
    }
    ctx->cg_graph_cache.push_back(CgGraph());
    e = &ctx->cg_graph_cache.back();
    e->key = key;
  }
  if (e->exec) {
    RSG_HIP(hipGraphLaunch(e->exec, ctx->stream));
    ++ctx->cg_graph_launches;
    return RSG_OK;
  }
  if (e->failed || e->seen++ == 0) return enqueue();
  int rc = RSG_OK;
  hipError_t ce = hipErrorUnknown;
  hipGraph_t graph = nullptr;
  {
    // no allocation, free or library call of another thread may run while
    // this stream records
    CaptureGate gate;
    if (!gate.usable()) {
      e->failed = true;
      return enqueue();
    }
    RSG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue();
    ce = hipStreamEndCapture(ctx->stream, &graph);
    // a capture the runtime invalidated can leave the stream recording:
    // end it before anything else is queued there
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ctx->stream, &st) == hipSuccess && st != hipStreamCaptureStatusNone) {
      hipGraph_t g2 = nullptr;
      (void)hipStreamEndCapture(ctx->stream, &g2);
      if (g2) (void)hipGraphDestroy(g2);
      ce = hipErrorStreamCaptureInvalidated;
    }
    if (rc == RSG_OK && ce == hipSuccess && graph &&
        hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0) == hipSuccess) {
      (void)hipGraphDestroy(graph);
      graph = nullptr;
    } else {
      e->exec = nullptr;
    }
  }
  if (e->exec) {
    RSG_HIP(hipGraphLaunch(e->exec, ctx->stream));
    ++ctx->cg_graph_launches;
    return RSG_OK;
  }
  // the capture failed (or the work reported an error while recording):
  // nothing ran, so run the block eagerly and keep this key eager
  if (graph) (void)hipGraphDestroy(graph);
  e->failed = true;
  ++ctx->cg_graph_failures;
  (void)hipGetLastError();
  ctx->err.clear();
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(ctx->stream, &st) != hipSuccess || st != hipStreamCaptureStatusNone)
    return ctx->fail(RSG_ERR_HIP, "conjugate_gradient: the stream is still recording after a failed graph capture");
  return enqueue();
}

// d_conjugate_gradient_1 (cg_module.f90:112-431) with the band block MB_d
// as large as the CG work arrays allow.  W = double is the plain FP64
// method on the store itself.  W = float keeps every CG and preconditioner
// vector, and an FP32 copy of the block's bands, in FP32 storage with FP32
// arithmetic (hpsi included); dot products accumulate in FP64, the 2x2
// Ritz step runs in FP64 and the accepted update is applied to the FP64
// store.
template <typename W>
int cg_impl(rsg_ctx *ctx, int ncg, int pc_kind, int mloop, double *esp, double *res) {
  int rc = check_store(ctx, "conjugate_gradient");
  if (rc) return rc;
  if (ncg < 0 || (pc_kind != 0 && pc_kind != 4 && pc_kind != 5 && pc_kind != 6))
    return ctx->fail(RSG_ERR_ARG, "conjugate_gradient: ncg < 0 or pc_kind not 0/4/5/6");
  const bool learned = pc_kind == 5 || pc_kind == 6;
  if (learned && ctx->lpc.reject_limit > 0 && ctx->lpc.reject_streak >= ctx->lpc.reject_limit) {
    // the network's output was rejected in the last reject_limit calls:
    // this call runs the hprsdft smoother without trying it
    ++ctx->lpc.skipped_calls;
    pc_kind = 4;
  }
  const int64_t fallbacks0 = ctx->lpc.fallbacks;
  constexpr bool kMixed = sizeof(W) == 4;
  if (ctx->band32 && !kMixed)
    return ctx->fail(RSG_ERR_STATE, "conjugate_gradient: the bands are in the FP32 store (FP32 CG only)");
  if (ctx->band32 && ctx->cg_swgs != 0)
    return ctx->fail(RSG_ERR_UNSUPPORTED, "conjugate_gradient: SWGS 1/2 needs the FP64 band store");
  // FP32 store: the CG updates its bands in place
  const bool in_place = kMixed && ctx->band32;
  if (in_place) ctx->unk_fresh = false;
  const int mb = ctx->mb;
  const int64_t N = ctx->N;
  const double dv = ctx->dv;
  // work vectors per band: hxk hpk gk Pgk pk + preconditioner r p f (+ x32)
  constexpr int kVec = kMixed ? 9 : 8;
  const size_t budget = size_t(64) << 30;
  int mbd = int(std::min<size_t>(size_t(mb), std::max<size_t>(1, budget / (size_t(N) * sizeof(W) * kVec))));
  if (ctx->cgbuf.ensure(size_t(N) * mbd * sizeof(W) * kVec) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "conjugate_gradient: work allocation failed");
  Scratch s;
  rc = ensure_scalars(ctx, mbd, 1, s);
  if (rc) return rc;
  W *base = ctx->cgbuf.as<W>();
  const size_t blk = size_t(N) * mbd;
  W *hxk = base, *hpk = base + blk, *gk = base + 2 * blk, *pgk = base + 3 * blk, *pk = base + 4 * blk;
  W *pr = base + 5 * blk, *pp = base + 6 * blk, *pf = base + 7 * blk;
  W *x32 = kMixed ? base + 8 * blk : nullptr;
  // everything one band block queues on the stream; no host read-back for
  // pc_kind 0 and 4 (the learned ones read back their descent check)
  auto enqueue_block = [&](int ns, int nn) -> int {
    int rc = RSG_OK;
    bool net_on = true;  // learned preconditioner still in use for this block
    double *xs64 = ctx->unk.as<double>() + size_t(ns) * N;
    const dim3 gN = grid_nb(N, nn), blkd(kThreads), gb((nn + 63) / 64), b64(64);
    W *x;
    if constexpr (kMixed) {
      if (in_place) {
        x = ctx->unk32.as<W>() + size_t(ns) * N;
      } else {
        const int64_t tot = N * nn;
        hipLaunchKernelGGL((k_to_work<W>), dim3(unsigned((tot + kThreads - 1) / kThreads)), blkd, 0, ctx->stream,
                           xs64, x32, tot);
        x = x32;
      }
    } else {
      x = reinterpret_cast<W *>(xs64);
    }
    hipLaunchKernelGGL(k_fill, gb, b64, 0, ctx->stream, s.E1, 1.0e10, nn);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, s.flags, 0, 2);
    rc = launch_hpsi<W>(ctx, x, hxk, N, nn);
    if (rc) return rc;
    {
      const W *xs[1] = {x}, *ys[1] = {hxk};
      if ((rc = multidot(ctx, 1, xs, ys, nn, dv, s.E))) return rc;
    }
    hipLaunchKernelGGL((k_cg_grad<W>), gN, blkd, 0, ctx->stream, hxk, x, s.E, gk, N);
    {
      const W *xs[1] = {gk}, *ys[1] = {gk};
      if ((rc = multidot(ctx, 1, xs, ys, nn, dv, s.rb))) return rc;
    }
    for (int icg = 1; icg <= ncg + 1; ++icg) {
      hipLaunchKernelGGL(k_scale_copy, gb, b64, 0, ctx->stream, s.rb, 0.25, s.res, nn);
      if (icg == ncg + 1) break;
      // convergence test on the device: once the whole block has
      // |E - E1| < 1e-15 the remaining steps leave it as it is (the
      // reference exits the loop), without a read-back per step
      hipLaunchKernelGGL(k_cg_conv, dim3(1), dim3(256), 0, ctx->stream, s.E, s.E1, nn, s.flags);
      // pgk = P gk and rb = dV <pgk, gk>; once the network's output has
      // failed the descent check in this call, the block keeps the hprsdft
      // smoother for the rest of the call
      const bool use_net = (pc_kind == 5 || pc_kind == 6) && net_on;
      const int pc_step = use_net ? pc_kind : (pc_kind == 0 ? 0 : 4);
      if (pc_step == 4 || pc_step == 6) {
        if ((rc = precond_hprsdft_fused<W>(ctx, mloop, gk, pgk, nn, pr, pp, pf, s, dv, s.rb))) return rc;
        if (pc_step == 6) {
          // hybrid: pgk = scale * P_hprsdft g + network(g); the
          // preconditioner's work vectors are free again, pf takes the net
          if ((rc = lpc_apply(ctx, gk, pf, N, nn, kMixed ? 32 : 64))) return rc;
          const int64_t tot = N * nn;
          hipLaunchKernelGGL((k_axpby<W>), dim3(unsigned((tot + kThreads - 1) / kThreads)), blkd, 0, ctx->stream,
                             pgk, pf, W(ctx->lpc.hybrid_scale), W(1), tot);
          const W *xs[1] = {pgk}, *ys[1] = {gk};
          if ((rc = multidot(ctx, 1, xs, ys, nn, dv, s.rb))) return rc;
        }
      } else {
        if (pc_step == 5) {
          if ((rc = lpc_apply(ctx, gk, pgk, N, nn, kMixed ? 32 : 64))) return rc;
        } else {
          const int64_t tot = N * nn;
          hipLaunchKernelGGL((k_copy<W>), dim3(unsigned((tot + kThreads - 1) / kThreads)), blkd, 0, ctx->stream, pgk,
                             gk, tot);
        }
        const W *xs[1] = {pgk}, *ys[1] = {gk};
        if ((rc = multidot(ctx, 1, xs, ys, nn, dv, s.rb))) return rc;
      }
      if (use_net) {
        // descent check: <Pg, g> must be positive and finite for every band,
        // else the block falls back to the hprsdft smoother
        std::vector<double> rbh(static_cast<size_t>(nn));
        RSG_HIP(hipMemcpyAsync(rbh.data(), s.rb, nn * 8, hipMemcpyDeviceToHost, ctx->stream));
        RSG_HIP(hipStreamSynchronize(ctx->stream));
        bool good = true;
        for (double v : rbh) good = good && std::isfinite(v) && v > 0.0;
        if (!good) {
          ++ctx->lpc.fallbacks;
          net_on = false;
          if ((rc = precond_hprsdft_fused<W>(ctx, mloop, gk, pgk, nn, pr, pp, pf, s, dv, s.rb))) return rc;
        }
      }
      if (ctx->cg_swgs == 1 || ctx->cg_swgs == 2) {
        // orthogonalisation of Pg against the other bands (:249-251), then
        // rb = dV <Pg, g> of the projected Pg (:255-257); the
        // preconditioner's r and p vectors are free here
        // (the FP64 copy of an FP32 block needs 2 N nn floats, 8-byte
        // aligned: r, p and f of the preconditioner hold 3 N mbd floats)
        double *tmp = reinterpret_cast<double *>((reinterpret_cast<uintptr_t>(pr) + 7) & ~uintptr_t(7));
        if ((rc = cggs_block<W>(ctx, ctx->cg_swgs, ns, nn, pgk, tmp))) return rc;
        const W *xs[1] = {pgk}, *ys[1] = {gk};
        if ((rc = multidot(ctx, 1, xs, ys, nn, dv, s.rb))) return rc;
      }
      hipLaunchKernelGGL((k_cg_dir<W>), gN, blkd, 0, ctx->stream, pgk, pk, s.rb, s.gkgk, icg == 1 ? 1 : 0, N);
      hipLaunchKernelGGL(k_scale_copy, gb, b64, 0, ctx->stream, s.rb, 1.0, s.gkgk, nn);
      if ((rc = launch_hpsi<W>(ctx, pk, hpk, N, nn))) return rc;
      if ((rc = ritz_dots<W>(ctx, x, pk, hxk, hpk, nn, dv, s.d6))) return rc;
      hipLaunchKernelGGL(k_cg_ritz, gb, b64, 0, ctx->stream, s.d6, nn, s.E, s.E1, s.uc, s.flags);
      hipLaunchKernelGGL((k_cg_update<W>), gN, blkd, 0, ctx->stream, hxk, hpk, x, pk, gk, s.uc, s.E, N);
      {
        const W *xs[1] = {gk}, *ys[1] = {gk};
        if ((rc = multidot(ctx, 1, xs, ys, nn, dv, s.rb))) return rc;
      }
      hipLaunchKernelGGL(k_cg_guard, gb, b64, 0, ctx->stream, s.rb, s.res, s.E, s.E1, s.accept, nn);
      if (in_place)
        hipLaunchKernelGGL((k_cg_accept32<W>), gN, blkd, 0, ctx->stream, x, pk, s.uc, s.accept, N);
      else
        hipLaunchKernelGGL((k_cg_accept<W>), gN, blkd, 0, ctx->stream, xs64, x32, pk, s.uc, s.accept, N);
      RSG_HIP(hipGetLastError());
    }
    return RSG_OK;
  };
  const bool graphable = ctx->cg_graphs && (pc_kind == 0 || pc_kind == 4) && ctx->cg_swgs == 0;
  for (int ns = 0; ns < mb; ns += mbd) {
    const int nn = std::min(mbd, mb - ns);
    if (graphable) {
      const std::vector<unsigned char> key = cg_graph_key(ctx, int(sizeof(W)), ns, nn, mbd, ncg, pc_kind, mloop);
      rc = run_graphed(ctx, key, [&]() { return enqueue_block(ns, nn); });
    } else {
      rc = enqueue_block(ns, nn);
    }
    if (rc) return rc;
    RSG_HIP(hipMemcpyAsync(esp + ns, s.E, nn * 8, hipMemcpyDeviceToHost, ctx->stream));
    RSG_HIP(hipMemcpyAsync(res + ns, s.res, nn * 8, hipMemcpyDeviceToHost, ctx->stream));
    RSG_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (pc_kind == 5 || pc_kind == 6) ctx->lpc.reject_streak = ctx->lpc.fallbacks > fallbacks0 ? ctx->lpc.reject_streak + 1 : 0;
  return RSG_OK;
}

}  // namespace

int conjugate_gradient(rsg_ctx *ctx, int ncg, int pc_kind, int mloop, double *esp, double *res) {
  return cg_impl<double>(ctx, ncg, pc_kind, mloop, esp, res);
}

int conjugate_gradient_f32(rsg_ctx *ctx, int ncg, int pc_kind, int mloop, double *esp, double *res) {
  return cg_impl<float>(ctx, ncg, pc_kind, mloop, esp, res);
}

int axpby_f32(rsg_ctx *ctx, float *y, const float *x, float a, float b, int64_t n) {
  if (n < 0 || (n > 0 && (!x || !y))) return ctx->fail(RSG_ERR_ARG, "axpby: bad arguments");
  if (n == 0) return RSG_OK;
  hipLaunchKernelGGL((k_axpby<float>), dim3(unsigned((n + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                     ctx->stream, y, x, a, b, n);
  RSG_HIP(hipGetLastError());
  return RSG_OK;
}

namespace {

// pg = P g for nb device bands of leading dimension ld: the hprsdft
// preconditioner on its own, with its own work vectors.
template <typename W> int precond_entry(rsg_ctx *ctx, int mloop, const W *g, W *pg, int64_t ld, int nb) {
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "precond_hprsdft: grid not set");
  if (nb < 0 || ld < ctx->N || mloop < 0) return ctx->fail(RSG_ERR_ARG, "precond_hprsdft: bad nb, ld or mloop");
  if (nb == 0) return RSG_OK;
  if (!g || !pg) return ctx->fail(RSG_ERR_ARG, "precond_hprsdft: null buffer");
  const size_t blk = size_t(ctx->N) * nb;
  if (ctx->pcbuf.ensure(3 * blk * sizeof(W)) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "precond_hprsdft: work allocation failed");
  Scratch s;
  int rc = ensure_scalars(ctx, nb, 1, s);
  if (rc) return rc;
  W *r = ctx->pcbuf.as<W>(), *p = r + blk, *f = p + blk;
  return precond_hprsdft_fused<W>(ctx, mloop, g, pg, nb, r, p, f, s, ctx->dv, nullptr, ld);
}

}  // namespace

int precond_hprsdft_dev(rsg_ctx *ctx, int mloop, const double *g, double *pg, int64_t ld, int nb) {
  return precond_entry<double>(ctx, mloop, g, pg, ld, nb);
}

int precond_hprsdft_dev(rsg_ctx *ctx, int mloop, const float *g, float *pg, int64_t ld, int nb) {
  return precond_entry<float>(ctx, mloop, g, pg, ld, nb);
}

namespace {
__global__ void k_resid32(const double *__restrict__ hx, const double *__restrict__ x, const double *__restrict__ E,
                          float *__restrict__ g, int64_t N) {
  const int b = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t k = int64_t(b) * N + i;
  g[k] = float(-2.0 * (hx[k] - E[b] * x[k]));
}
}  // namespace

// Uniform [0, 1) start bands, one counter-based value per (band, point):
// u(j, b) = splitmix64(seed + (b N + j + 1) * golden) / 2^64 with the 53 top
// bits kept, so the values do not depend on the launch shape.  The same
// formula runs on the host (systems.counter_bands).
__global__ __launch_bounds__(256) void k_random_bands(double *__restrict__ u, int64_t N, int n0, uint64_t seed) {
  const int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (j >= N) return;
  const int64_t b = int64_t(n0) + blockIdx.y;
  uint64_t z = seed + uint64_t(b * N + j + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  u[b * N + j] = double(z >> 11) * 0x1.0p-53;
}

// Fills columns n0 .. n0+nb-1 of the band store (random_initial_wf,
// wf_module.f90:166-191: uniform random numbers per point and band).
int random_bands(rsg_ctx *ctx, int n0, int nb, uint64_t seed) {
  if (ctx->mb == 0) return ctx->fail(RSG_ERR_STATE, "random_bands: no band store");
  if (ctx->band32) return ctx->fail(RSG_ERR_STATE, "random_bands: the bands are in the FP32 store");
  if (n0 < 0 || nb < 0 || n0 + nb > ctx->mb) return ctx->fail(RSG_ERR_ARG, "random_bands: band range out of bounds");
  if (nb == 0) return RSG_OK;
  if (nb > 65535) return ctx->fail(RSG_ERR_ARG, "random_bands: nb above 65535");
  hipLaunchKernelGGL(k_random_bands, dim3(unsigned((ctx->N + 255) / 256), unsigned(nb)), dim3(256), 0, ctx->stream,
                     ctx->unk.as<double>(), ctx->N, n0, seed);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ctx->fail(RSG_ERR_HIP, std::string("random_bands: ") + hipGetErrorString(e));
  return RSG_OK;
}

// CG gradients g_n = -2 (H u_n - e_n u_n), e_n = dV <u_n|H u_n>, of the
// whole store in FP32 (cg_module.f90:207, the first gradient of a CG call):
// the residuals the learned preconditioner is trained on.
int band_residuals_f32(rsg_ctx *ctx, float *g) {
  int rc = check_store64(ctx, "band_residuals");
  if (rc) return rc;
  const int mb = ctx->mb;
  const int64_t N = ctx->N;
  if (ctx->esp_dev.ensure(size_t(mb) * 16) != hipSuccess) return ctx->fail(RSG_ERR_HIP, "band_residuals: allocation failed");
  double *U = ctx->unk.as<double>(), *HU = ctx->work.as<double>();
  if ((rc = launch_hpsi<double>(ctx, U, HU, N, mb))) return rc;
  const double *xs[1] = {U}, *ys[1] = {HU};
  if ((rc = multidot(ctx, 1, xs, ys, mb, ctx->dv, ctx->esp_dev.as<double>()))) return rc;
  hipLaunchKernelGGL(k_resid32, grid_nb(N, mb), dim3(kThreads), 0, ctx->stream, HU, U, ctx->esp_dev.as<double>(), g, N);
  RSG_HIP(hipGetLastError());
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

// Training targets of the learned preconditioner: t = A^-1 g with
// A = -1/2 lap_7 + vweight (Vloc - min Vloc) + shift, by `steps` CG steps
// from t0 = g, for nb FP32 bands (device, leading dimension N).
int lpc_target(rsg_ctx *ctx, const float *g, float *t, int nb, int steps, double shift, double vweight) {
  if (!ctx->have_grid || !ctx->have_vloc) return ctx->fail(RSG_ERR_STATE, "lpc_target: grid or Vloc not set");
  if (nb < 1 || steps < 1 || !(shift > 0.0) || vweight < 0.0)
    return ctx->fail(RSG_ERR_ARG, "lpc_target: nb, steps >= 1, shift > 0, vweight >= 0");
  const int64_t N = ctx->N;
  LpcState &L = ctx->lpc;
  std::vector<double> v(static_cast<size_t>(N));
  RSG_HIP(hipMemcpyAsync(v.data(), ctx->vloc64.p, size_t(N) * 8, hipMemcpyDeviceToHost, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  const double vmin = *std::min_element(v.begin(), v.end());
  std::vector<float> d(static_cast<size_t>(N));
  for (int64_t i = 0; i < N; ++i) d[size_t(i)] = float(vweight * (v[size_t(i)] - vmin) + shift);
  if (L.diag.ensure(size_t(N) * 4) != hipSuccess || L.tbuf.ensure(size_t(N) * nb * 4 * 3) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "lpc_target: allocation failed");
  RSG_HIP(hipMemcpyAsync(L.diag.p, d.data(), size_t(N) * 4, hipMemcpyHostToDevice, ctx->stream));
  Scratch s;
  int rc = ensure_scalars(ctx, nb, 1, s);
  if (rc) return rc;
  float *r = L.tbuf.as<float>(), *p = r + size_t(N) * nb, *f = p + size_t(N) * nb;
  RSG_HIP(hipMemcpyAsync(t, g, size_t(N) * nb * 4, hipMemcpyDeviceToDevice, ctx->stream));
  rc = precond_hprsdft<float>(ctx, steps, g, t, nb, r, p, f, s, L.diag.as<float>());
  if (rc) return rc;
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

int check_store64(rsg_ctx *ctx, const char *who) {
  int rc = check_store(ctx, who);
  if (rc) return rc;
  if (ctx->band32)
    return ctx->fail(RSG_ERR_STATE, std::string(who) + ": the bands are in the FP32 store (set the band precision to 64)");
  return RSG_OK;
}

int set_band_precision(rsg_ctx *ctx, int bits) {
  int rc = check_store(ctx, "set_band_precision");
  if (rc) return rc;
  if (bits != 32 && bits != 64) return ctx->fail(RSG_ERR_ARG, "set_band_precision: bits must be 32 or 64");
  if ((bits == 32) == ctx->band32) return RSG_OK;
  const int64_t tot = ctx->N * int64_t(ctx->mb);
  const dim3 g(unsigned((tot + kThreads - 1) / kThreads)), b(kThreads);
  if (bits == 32) {
    const size_t bytes = size_t(tot) * sizeof(float);
    if (ctx->unk32.ensure(bytes) != hipSuccess || ctx->work32.ensure(bytes) != hipSuccess)
      return ctx->fail(RSG_ERR_HIP, "set_band_precision: out of device memory");
    hipLaunchKernelGGL((k_to_work<float>), g, b, 0, ctx->stream, ctx->unk.as<double>(), ctx->unk32.as<float>(), tot);
  } else if (!ctx->unk_fresh) {  // else unk already holds the bands, in FP64
    hipLaunchKernelGGL((k_to_double<float>), g, b, 0, ctx->stream, ctx->unk32.as<float>(), ctx->unk.as<double>(), tot);
  }
  RSG_HIP(hipGetLastError());
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  ctx->band32 = bits == 32;
  ctx->unk_fresh = ctx->band32;  // the FP32 store starts as a copy of unk
  return RSG_OK;
}

void solver_release(rsg_ctx *ctx) {
  ctx->dotpart.release();
  ctx->grampart.release();
  ctx->grampart32.release();
  ctx->pcbuf.release();
  ctx->unk32.release();
  ctx->work32.release();
  ctx->hsub32.release();
  ctx->band32 = false;
  DeviceOpGuard guard;
  for (auto &g : ctx->cg_graph_cache)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  ctx->cg_graph_cache.clear();
}

}  // namespace rsg
