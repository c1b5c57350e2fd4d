// hpsi.hip -- Hamiltonian application on gfx950:
//   hpsi = (T + Vloc) psi + V_NL psi  over a block of bands.
//
// Two launches per call:
//  1. k_nl_gather: for every atom group (<= 4 projectors sharing one sphere
//     of grid points) and band tile, uVunk(lma,b) = iuV dV sum_j uV(j) psi(j,b)
//     (d_ps_nloc2_op_module.f90:182-191).  psi at each sphere point is read
//     once for all projectors of the atom; partial sums are reduced with
//     wavefront shuffles and one LDS pass.
//  2. k_kinetic_nl: the 13-point-per-axis stencil fused with (c+V)psi
//     (d_kinetic_sol_module.f90:244-292) and, at the same output point, the
//     nonlocal scatter hpsi(j) += sum_lma uV(j,lma) uVunk(lma)
//     (d_ps_nloc2_op_module.f90:274-281) taken from a per-point CSR, so no
//     atomics and one write of hpsi.  Threads own one (x,y) point of a plane
//     (consecutive threads -> consecutive x, coalesced) and stream a run of
//     z-planes with the 2*MD+1 z-neighbours held in registers; the periodic
//     halo of bcset_3 (bc/bcset_3_module.f90:15-151) is modular indexing.
#include <algorithm>

#include "context.hpp"

namespace rsg {
namespace {

constexpr int kMaxMd = 8;
constexpr int kThreads = 256;

template <typename T> struct StencilCoef {
  T c[3][kMaxMd];  // coef_lap(axis, m), m = 1..md
  T c0;            // coef_lap0 (+ const_k2 = 0 at Gamma)
};

template <typename T, int MD>
__global__ __launch_bounds__(kThreads) void k_kinetic_nl(
    const T *__restrict__ psi, T *__restrict__ hpsi, int64_t ld,
    const T *__restrict__ vloc, StencilCoef<T> cf, int n1, int n2, int n3,
    int zlen, const int *__restrict__ pt_off, const int *__restrict__ pt_lma,
    const T *__restrict__ pt_uv, const T *__restrict__ uvunk, int nzlma) {
  const int n12 = n1 * n2;
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n12) return;
  const int ib = blockIdx.z;
  const int z0 = blockIdx.y * zlen;
  const int z1 = min(z0 + zlen, n3);
  const int i1 = p % n1, i2 = p / n1;
  const T *__restrict__ w = psi + int64_t(ib) * ld;
  T *__restrict__ o = hpsi + int64_t(ib) * ld;

  int xp[MD], xm[MD], yp[MD], ym[MD];
#pragma unroll
  for (int m = 1; m <= MD; ++m) {
    xp[m - 1] = (i1 + m) % n1 + i2 * n1;
    xm[m - 1] = ((i1 - m) % n1 + n1) % n1 + i2 * n1;
    yp[m - 1] = i1 + ((i2 + m) % n2) * n1;
    ym[m - 1] = i1 + (((i2 - m) % n2 + n2) % n2) * n1;
  }
  T q[2 * MD + 1];
#pragma unroll
  for (int k = 0; k < 2 * MD + 1; ++k) {
    int z = ((z0 - MD + k) % n3 + n3) % n3;
    q[k] = w[p + int64_t(z) * n12];
  }
  const T *__restrict__ uv = uvunk + int64_t(ib) * nzlma;
  for (int z = z0; z < z1; ++z) {
    const int64_t zb = int64_t(z) * n12;
    const int64_t j = zb + p;
    T acc = (cf.c0 + vloc[j]) * q[MD];
#pragma unroll
    for (int m = 0; m < MD; ++m) acc += cf.c[0][m] * (w[zb + xm[m]] + w[zb + xp[m]]);
#pragma unroll
    for (int m = 0; m < MD; ++m) acc += cf.c[1][m] * (w[zb + ym[m]] + w[zb + yp[m]]);
#pragma unroll
    for (int m = 0; m < MD; ++m) acc += cf.c[2][m] * (q[MD - 1 - m] + q[MD + 1 + m]);
    if (pt_off) {
      const int e1 = pt_off[j + 1];
      for (int e = pt_off[j]; e < e1; ++e) acc += pt_uv[e] * uv[pt_lma[e]];
    }
    o[j] = acc;
#pragma unroll
    for (int k = 0; k < 2 * MD; ++k) q[k] = q[k + 1];
    if (z + 1 < z1) {
      int zn = (z + 1 + MD) % n3;
      q[2 * MD] = w[p + int64_t(zn) * n12];
    }
  }
}

template <typename T> __device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T, int BB>
__global__ __launch_bounds__(kThreads) void k_nl_gather(
    const T *__restrict__ psi, int64_t ld, int nb, const int *__restrict__ grp_off,
    const int *__restrict__ grp_lma0, const int *__restrict__ grp_nl,
    const int *__restrict__ grp_pts, const T *__restrict__ grp_coef,
    const T *__restrict__ scale, T *__restrict__ uvunk, int nzlma) {
  __shared__ T red[kThreads / 64][4 * BB];
  const int g = blockIdx.x;
  const int b0 = blockIdx.y * BB;
  const int nbb = min(BB, nb - b0);
  const int beg = grp_off[g], end = grp_off[g + 1];
  T acc[4][BB];
#pragma unroll
  for (int l = 0; l < 4; ++l)
#pragma unroll
    for (int b = 0; b < BB; ++b) acc[l][b] = T(0);
  const T *__restrict__ w = psi + int64_t(b0) * ld;
  for (int i = beg + threadIdx.x; i < end; i += kThreads) {
    const int pt = grp_pts[i];
    const T c0 = grp_coef[4 * i + 0], c1 = grp_coef[4 * i + 1];
    const T c2 = grp_coef[4 * i + 2], c3 = grp_coef[4 * i + 3];
#pragma unroll
    for (int b = 0; b < BB; ++b) {
      if (b < nbb) {
        const T v = w[pt + int64_t(b) * ld];
        acc[0][b] += c0 * v;
        acc[1][b] += c1 * v;
        acc[2][b] += c2 * v;
        acc[3][b] += c3 * v;
      }
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int l = 0; l < 4; ++l)
#pragma unroll
    for (int b = 0; b < BB; ++b) {
      T s = wave_sum(acc[l][b]);
      if (lane == 0) red[wid][l * BB + b] = s;
    }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < 4 * BB) {
    const int l = t / BB, b = t % BB;
    const int nl = grp_nl[g];
    if (l < nl && b < nbb) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < kThreads / 64; ++k) s += red[k][t];
      const int lma = grp_lma0[g] + l;
      uvunk[lma + int64_t(b0 + b) * nzlma] = scale[lma] * s;
    }
  }
}

template <typename T> StencilCoef<T> make_coef(const rsg_ctx *ctx) {
  StencilCoef<T> cf{};
  for (int a = 0; a < 3; ++a)
    for (int m = 0; m < ctx->md; ++m) cf.c[a][m] = T(ctx->coef_lap[size_t(a) * ctx->md + m]);
  cf.c0 = T(ctx->coef_lap0);
  return cf;
}

template <typename T> NonlocalDev<T> &nl_of(rsg_ctx *ctx);
template <> NonlocalDev<double> &nl_of<double>(rsg_ctx *ctx) { return ctx->nl64; }
template <> NonlocalDev<float> &nl_of<float>(rsg_ctx *ctx) { return ctx->nl32; }
template <typename T> const T *vloc_of(rsg_ctx *ctx);
template <> const double *vloc_of<double>(rsg_ctx *ctx) { return ctx->vloc64.as<double>(); }
template <> const float *vloc_of<float>(rsg_ctx *ctx) { return ctx->vloc32.as<float>(); }

template <typename T, int MD>
void launch_stencil(rsg_ctx *ctx, const T *psi, T *hpsi, int64_t ld, int nb, const int *pt_off,
                    const int *pt_lma, const T *pt_uv, const T *uvunk, int nzlma) {
  const int n12 = ctx->n[0] * ctx->n[1], n3 = ctx->n[2];
  const int bx = (n12 + kThreads - 1) / kThreads;
  // z-run length: long runs amortise the 2*MD-plane register warm-up;
  // shorten them only until the grid holds enough workgroups.
  int zlen = n3;
  while (zlen > 16 && int64_t(bx) * ((n3 + zlen - 1) / zlen) * nb < 4096) zlen = (zlen + 1) / 2;
  dim3 grid(bx, (n3 + zlen - 1) / zlen, nb);
  hipLaunchKernelGGL((k_kinetic_nl<T, MD>), grid, dim3(kThreads), 0, ctx->stream, psi, hpsi, ld,
                     vloc_of<T>(ctx), make_coef<T>(ctx), ctx->n[0], ctx->n[1], n3, zlen, pt_off,
                     pt_lma, pt_uv, uvunk, nzlma);
}

}  // namespace

template <typename T>
int launch_hpsi(rsg_ctx *ctx, const T *psi, T *hpsi, int64_t ld, int nb) {
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "hpsi: grid not set");
  if (nb < 0 || ld < ctx->N) return ctx->fail(RSG_ERR_ARG, "hpsi: bad nb or ld");
  if (nb == 0) return RSG_OK;
  if (nb > 65535) return ctx->fail(RSG_ERR_ARG, "hpsi: nb above 65535");
  if (!psi || !hpsi) return ctx->fail(RSG_ERR_ARG, "hpsi: null buffer");
  NonlocalDev<T> &nl = nl_of<T>(ctx);
  const int nz = ctx->tables.nzlma;
  const int *pt_off = nullptr, *pt_lma = nullptr;
  const T *pt_uv = nullptr;
  T *uv = nullptr;
  if (ctx->have_nl) {
    constexpr int BB = 8;
    if (ctx->uvunk.ensure(size_t(nz) * nb * sizeof(T)) != hipSuccess)
      return ctx->fail(RSG_ERR_HIP, "hpsi: uVunk allocation failed");
    uv = ctx->uvunk.as<T>();
    dim3 grid(nl.ngroup, (nb + BB - 1) / BB);
    hipLaunchKernelGGL((k_nl_gather<T, BB>), grid, dim3(kThreads), 0, ctx->stream, psi, ld, nb,
                       nl.grp_off.template as<int>(), nl.grp_lma0.template as<int>(),
                       nl.grp_nl.template as<int>(), nl.grp_pts.template as<int>(),
                       nl.grp_coef.template as<T>(), nl.scale.template as<T>(), uv, nz);
    pt_off = nl.pt_off.template as<int>();
    pt_lma = nl.pt_lma.template as<int>();
    pt_uv = nl.pt_uv.template as<T>();
  }
  switch (ctx->md) {
#define RSG_MD_CASE(M) \
  case M:              \
    launch_stencil<T, M>(ctx, psi, hpsi, ld, nb, pt_off, pt_lma, pt_uv, uv, nz); break;
    RSG_MD_CASE(1)
    RSG_MD_CASE(2)
    RSG_MD_CASE(3)
    RSG_MD_CASE(4)
    RSG_MD_CASE(5)
    RSG_MD_CASE(6)
    RSG_MD_CASE(7)
    RSG_MD_CASE(8)
#undef RSG_MD_CASE
    default:
      return ctx->fail(RSG_ERR_UNSUPPORTED, "hpsi: md outside 1..8");
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ctx->fail(RSG_ERR_HIP, std::string("hpsi launch: ") + hipGetErrorString(e));
  return RSG_OK;
}

template int launch_hpsi<double>(rsg_ctx *, const double *, double *, int64_t, int);
template int launch_hpsi<float>(rsg_ctx *, const float *, float *, int64_t, int);

}  // namespace rsg
