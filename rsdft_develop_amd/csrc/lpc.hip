// lpc.hip -- learned CG preconditioner: a 3-layer 3x3x3 convolution net on
// the periodic grid that maps a band residual g and the cell's normalised
// local potential v to a correction P g (replacing precond_hprsdft,
// cgpc_hprsdft_module.f90:133-361, which has no learned counterpart in the
// reference):
//
//   s = rms(g); x = [g/s, v]
//   a1 = relu(W1 * x + b1)   (2 -> 16)
//   a2 = relu(W2 * a1 + b2)  (16 -> 16)
//   y  = W3 * a2 + b3        (16 -> 1)
//   P g = s y                (scale-equivariant: P(c g) = c P g)
//
// Storage: weights and activations in bfloat16, channel-last
// (act[sample][point][channel]); every product on MFMA
// (v_mfma_f32_16x16x16bf16_1k) with FP32 accumulation.  A convolution is
// D[co][p] = sum_k W[co][k] X[k][p] over k = tap * CIN + ci, tap = d3*9 +
// d2*3 + d1, with X[k][p] = in[point p shifted by (d - 1), periodic][ci];
// the weight matrix is the MFMA A operand and 16 grid points form the B
// operand, so a lane's output is 4 consecutive channels of one point.
// The input gradient of a layer is the same convolution with the adjoint
// weights W'[ci][tap' * COUT + co] = W[co][ci][26 - tap'].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "lpc.hpp"
#include "solver.hpp"

namespace rsg {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;  // waves per workgroup in the convolution kernels
constexpr int kTiles = 4;  // 16-point tiles per wave

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return uint16_t((u >> 16) | 0x40);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  uint32_t u = uint32_t(h) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

struct Grid3 {
  int n1, n2, n3;
};

// neighbour of point (i1, i2, i3) at tap t (offsets d - 1 in {-1, 0, 1})
__device__ inline int64_t nbr(const Grid3 &g, int i1, int i2, int i3, int t) {
  const int d3 = t / 9, d2 = (t / 3) % 3, d1 = t % 3;
  int x = i1 + d1 - 1, y = i2 + d2 - 1, z = i3 + d3 - 1;
  x = x < 0 ? x + g.n1 : (x >= g.n1 ? x - g.n1 : x);
  y = y < 0 ? y + g.n2 : (y >= g.n2 ? y - g.n2 : y);
  z = z < 0 ? z + g.n3 : (z >= g.n3 ? z - g.n3 : z);
  return x + int64_t(g.n1) * (y + int64_t(g.n2) * z);
}

// B fragment of chunk c for the lane's point: X[k][p], k = 16c + 4(l/16) + i
template <int CIN>
__device__ inline bf16x4 load_b(const uint16_t *__restrict__ in, const Grid3 &g, int i1, int i2, int i3, int c,
                                int lq) {
  bf16x4 v;
  if constexpr (CIN == 16) {
    const int64_t q = nbr(g, i1, i2, i3, c);
    const uint2 w = *reinterpret_cast<const uint2 *>(in + q * 16 + 4 * lq);
    v[0] = short(w.x & 0xffff);
    v[1] = short(w.x >> 16);
    v[2] = short(w.y & 0xffff);
    v[3] = short(w.y >> 16);
  } else if constexpr (CIN == 2) {
    const int t0 = (16 * c + 4 * lq) / 2;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int t = t0 + h;
      uint32_t w = 0;
      if (t < 27) w = *reinterpret_cast<const uint32_t *>(in + nbr(g, i1, i2, i3, t) * 2);
      v[2 * h] = short(w & 0xffff);
      v[2 * h + 1] = short(w >> 16);
    }
  } else {  // CIN == 1
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int t = 16 * c + 4 * lq + i;
      v[i] = t < 27 ? short(in[nbr(g, i1, i2, i3, t)]) : short(0);
    }
  }
  return v;
}

enum ConvMode { kReluOut = 0, kFinalOut = 1, kMaskedGrad = 2 };

// One convolution layer for nsamp samples.  Modes:
//  kReluOut:   out_bf[p][co] = bf16(relu(D + bias))            (16 outputs)
//  kFinalOut:  y = D[0] + bias[0]; yout[p] = y (FP32 if non-null);
//              pg[p] = s y  (the preconditioned residual, type T)
//  kMaskedGrad: out_bf[p][co] = bf16(D * (mask[p][co] > 0))    (16 outputs)
template <int CIN, int KP, int MODE, typename T>
__global__ __launch_bounds__(64 * kWaves) void k_conv(Grid3 g, int64_t N, const uint16_t *__restrict__ in,
                                                      int64_t in_stride, const uint16_t *__restrict__ wbf,
                                                      const float *__restrict__ bias, uint16_t *__restrict__ out_bf,
                                                      const uint16_t *__restrict__ mask, float *__restrict__ yout,
                                                      const T *__restrict__ gsrc, T *__restrict__ pg, int64_t ldg,
                                                      const float *__restrict__ sval) {
  constexpr int NC = KP / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lq = lane >> 4, lr = lane & 15;
  const int smp = blockIdx.y;
  const uint16_t *__restrict__ src = in + int64_t(smp) * in_stride;
  // A fragments: W[co = lr][k = 16c + 4 lq + i]
  bf16x4 a[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const uint2 w = *reinterpret_cast<const uint2 *>(wbf + lr * KP + 16 * c + 4 * lq);
    a[c][0] = short(w.x & 0xffff);
    a[c][1] = short(w.x >> 16);
    a[c][2] = short(w.y & 0xffff);
    a[c][3] = short(w.y >> 16);
  }
  const int64_t base = (int64_t(blockIdx.x) * kWaves + wave) * (16 * kTiles);
  for (int t = 0; t < kTiles; ++t) {
    const int64_t p0 = base + 16 * t;
    if (p0 >= N) break;  // wave-uniform
    const int64_t p = min(p0 + lr, N - 1);
    const int i1 = int(p % g.n1), i2 = int((p / g.n1) % g.n2), i3 = int(p / (int64_t(g.n1) * g.n2));
    // three independent accumulator chains (taps c mod 3): the matrix-core
    // latency of one chain of NC dependent products, not its issue rate,
    // would otherwise bound the tile
    f32x4 acc3[3] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const bf16x4 b = load_b<CIN>(src, g, i1, i2, i3, c, lq);
      acc3[c % 3] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[c], b, acc3[c % 3], 0, 0, 0);
    }
    const f32x4 acc = (acc3[0] + acc3[1]) + acc3[2];
    const bool ok = p0 + lr < N;
    if constexpr (MODE == kReluOut || MODE == kMaskedGrad) {
      uint16_t h[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int co = 4 * lq + i;
        float z = acc[i];
        if constexpr (MODE == kReluOut) {
          z += bias[co];
          z = z > 0.f ? z : 0.f;
        } else {
          const uint16_t m = mask[int64_t(smp) * N * 16 + p * 16 + co];
          z = bf2f(m) > 0.f ? z : 0.f;
        }
        h[i] = f2bf(z);
      }
      if (ok) {
        uint2 w;
        w.x = uint32_t(h[0]) | (uint32_t(h[1]) << 16);
        w.y = uint32_t(h[2]) | (uint32_t(h[3]) << 16);
        *reinterpret_cast<uint2 *>(out_bf + int64_t(smp) * N * 16 + p * 16 + 4 * lq) = w;
      }
    } else {
      // row co = 0 lives in acc[0] of lanes 0..15
      if (lq == 0 && ok) {
        const float y = acc[0] + bias[0];
        if (yout) yout[int64_t(smp) * N + p] = y;
        if (pg) pg[int64_t(smp) * ldg + p] = T(double(sval[smp]) * double(y));
      }
    }
  }
}

// x[smp][p] = (bf16(g/s), bf16(v))
template <typename T>
__global__ void k_prep(const T *__restrict__ gsrc, int64_t ldg, const float *__restrict__ v,
                       const float *__restrict__ sval, int64_t N, uint16_t *__restrict__ x) {
  const int smp = blockIdx.y;
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const float s = sval[smp];
  const float g = float(double(gsrc[int64_t(smp) * ldg + p]) / double(s));
  const uint32_t w = uint32_t(f2bf(g)) | (uint32_t(f2bf(v[p])) << 16);
  *reinterpret_cast<uint32_t *>(x + (int64_t(smp) * N + p) * 2) = w;
}

// s = sqrt(<g,g>/N) per sample, with a floor that keeps 0/0 out
__global__ void k_rms(const double *__restrict__ gg, int nsamp, int64_t N, float *__restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsamp) s[i] = float(fmax(sqrt(gg[i] / double(N)), 1e-30));
}

// Loss of a batch: r = y - kappa t / s, L += sum r^2 / (N nsamp);
// dy = 2 r / (N nsamp) stored as bf16 for the backward convolutions.
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ tgt, const float *__restrict__ y,
                                              const float *__restrict__ sval, float kappa, int64_t N,
                                              double inv_count, uint16_t *__restrict__ dy,
                                              double *__restrict__ lsum) {
  __shared__ double red[4];
  const int smp = blockIdx.y;
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double r2 = 0.0;
  if (p < N) {
    const int64_t o = int64_t(smp) * N + p;
    const float r = y[o] - kappa * tgt[o] / sval[smp];
    r2 = double(r) * r;
    dy[o] = f2bf(float(2.0 * inv_count * r));
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) r2 += __shfl_xor(r2, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r2;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(lsum, (red[0] + red[1] + red[2] + red[3]) * inv_count);
}

// dW[co][k] += sum_p dz[p][co] X[k][p]: MFMA with co x k-chunk outputs and
// 16 points per product.  Each workgroup folds its partial into part[blk].
template <int CIN, int KP, int COUT>
__global__ __launch_bounds__(256) void k_conv_dw(Grid3 g, int64_t N, int nsamp, const uint16_t *__restrict__ in,
                                                 const uint16_t *__restrict__ dz, int64_t pts_per_block,
                                                 float *__restrict__ part) {
  constexpr int NC = KP / 16;
  __shared__ float red[16 * KP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lq = lane >> 4, lr = lane & 15;
  for (int i = threadIdx.x; i < 16 * KP; i += blockDim.x) red[i] = 0.f;
  __syncthreads();
  f32x4 acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int64_t total = N * nsamp;
  const int64_t b0 = int64_t(blockIdx.x) * pts_per_block;
  const int64_t b1 = min(total, b0 + pts_per_block);
  for (int64_t q0 = b0 + 16 * wave; q0 < b1; q0 += 16 * 4) {
    // A: dz^T [co = lr][point 4 lq + i]; B: X [k = 16c + lr][point 4 lq + i]
    bf16x4 av;
    int smp[4], c1[4], c2[4], c3[4];
    bool ok[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t q = q0 + 4 * lq + i;
      ok[i] = q < b1;
      const int64_t qq = ok[i] ? q : b0;
      smp[i] = int(qq / N);
      const int64_t p = qq - int64_t(smp[i]) * N;
      c1[i] = int(p % g.n1);
      c2[i] = int((p / g.n1) % g.n2);
      c3[i] = int(p / (int64_t(g.n1) * g.n2));
      av[i] = (ok[i] && lr < COUT) ? short(dz[(int64_t(smp[i]) * N + p) * COUT + lr]) : short(0);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int k = 16 * c + lr;
      const int tap = k / CIN, ci = k % CIN;
      bf16x4 bv;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        bv[i] = (ok[i] && tap < 27)
                    ? short(in[(int64_t(smp[i]) * N + nbr(g, c1[i], c2[i], c3[i], tap)) * CIN + ci])
                    : short(0);
      }
      acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[c], 0, 0, 0);
    }
  }
  // D[co = 4 lq + i][k = 16c + lr]
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int i = 0; i < 4; ++i) atomicAdd(&red[(4 * lq + i) * KP + 16 * c + lr], acc[c][i]);
  __syncthreads();
  for (int i = threadIdx.x; i < 16 * KP; i += blockDim.x) part[int64_t(blockIdx.x) * 16 * KP + i] = red[i];
}

__global__ void k_sum_parts(const float *__restrict__ part, int nblk, int len, float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[int64_t(b) * len + i];
  out[i] = float(s);
}

// bias gradient: db[co] = sum_p dz[p][co] (COUT channels, bf16)
template <int COUT>
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t *__restrict__ dz, int64_t total,
                                                   float *__restrict__ db) {
  float acc[COUT];
#pragma unroll
  for (int c = 0; c < COUT; ++c) acc[c] = 0.f;
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < total; q += int64_t(gridDim.x) * blockDim.x)
#pragma unroll
    for (int c = 0; c < COUT; ++c) acc[c] += bf2f(dz[q * COUT + c]);
#pragma unroll
  for (int c = 0; c < COUT; ++c) {
    float v = acc[c];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&db[c], v);
  }
}

// Adam on the FP32 master parameters
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ gr, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

// ------------------------------------------------------------ host side

// Parameter vector layout (FP32 master, reference order):
//   W1[16][2][27] | b1[16] | W2[16][16][27] | b2[16] | W3[1][16][27] | b3[1]
// with W[co][ci][tap], tap = d3*9 + d2*3 + d1.
int lpc_param_count() { return 16 * 2 * 27 + 16 + 16 * 16 * 27 + 16 + 16 * 27 + 1; }

namespace {

constexpr int KP1 = 64, KP2 = 432, KP3 = 432, KP3T = 32;
constexpr int OFF_W1 = 0, OFF_B1 = 864, OFF_W2 = 880, OFF_B2 = 7792, OFF_W3 = 7808, OFF_B3 = 8240;

// bf16 weight matrices for the kernels, from the FP32 master vector
void pack_weights(const std::vector<float> &p, std::vector<uint16_t> &w) {
  w.assign(size_t(16) * (KP1 + KP2 + KP3 + KP2 + KP3T), 0);
  uint16_t *w1 = w.data(), *w2 = w1 + 16 * KP1, *w3 = w2 + 16 * KP2, *w2t = w3 + 16 * KP3, *w3t = w2t + 16 * KP2;
  for (int co = 0; co < 16; ++co)
    for (int ci = 0; ci < 2; ++ci)
      for (int t = 0; t < 27; ++t) w1[co * KP1 + t * 2 + ci] = f2bf(p[OFF_W1 + (co * 2 + ci) * 27 + t]);
  for (int co = 0; co < 16; ++co)
    for (int ci = 0; ci < 16; ++ci)
      for (int t = 0; t < 27; ++t) {
        const uint16_t h = f2bf(p[OFF_W2 + (co * 16 + ci) * 27 + t]);
        w2[co * KP2 + t * 16 + ci] = h;
        w2t[ci * KP2 + (26 - t) * 16 + co] = h;  // adjoint: rows ci, flipped taps
      }
  for (int ci = 0; ci < 16; ++ci)
    for (int t = 0; t < 27; ++t) {
      const uint16_t h = f2bf(p[OFF_W3 + ci * 27 + t]);
      w3[0 * KP3 + t * 16 + ci] = h;  // row 0 only
      w3t[ci * KP3T + (26 - t)] = h;  // adjoint: 1 input channel
    }
}

}  // namespace

LpcState::~LpcState() { release(); }

void LpcState::release() {
  for (DevBuf *b : {&wbf, &bias, &vnorm, &x0, &a1, &a2, &y, &dyb, &dz1, &dz2, &sval, &gg, &part, &grad, &master, &mom1,
                    &mom2, &lsum, &diag, &tbuf})
    b->release();
}

int lpc_set_params(rsg_ctx *ctx, const float *params) {
  LpcState &L = ctx->lpc;
  const int n = lpc_param_count();
  std::vector<float> p(params, params + n);
  std::vector<uint16_t> w;
  pack_weights(p, w);
  float b[16 + 16 + 1];
  std::copy(p.begin() + OFF_B1, p.begin() + OFF_B1 + 16, b);
  std::copy(p.begin() + OFF_B2, p.begin() + OFF_B2 + 16, b + 16);
  b[32] = p[OFF_B3];
  if (L.wbf.ensure(w.size() * 2) != hipSuccess || L.bias.ensure(sizeof(b)) != hipSuccess ||
      L.master.ensure(size_t(n) * 4) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "lpc: allocation failed");
  RSG_HIP(hipMemcpyAsync(L.wbf.p, w.data(), w.size() * 2, hipMemcpyHostToDevice, ctx->stream));
  RSG_HIP(hipMemcpyAsync(L.bias.p, b, sizeof(b), hipMemcpyHostToDevice, ctx->stream));
  RSG_HIP(hipMemcpyAsync(L.master.p, p.data(), size_t(n) * 4, hipMemcpyHostToDevice, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  L.have_params = true;
  return RSG_OK;
}

int lpc_get_params(rsg_ctx *ctx, float *params) {
  LpcState &L = ctx->lpc;
  if (!L.have_params) return ctx->fail(RSG_ERR_STATE, "lpc: no parameters");
  RSG_HIP(hipMemcpyAsync(params, L.master.p, size_t(lpc_param_count()) * 4, hipMemcpyDeviceToHost, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

// v = (Vloc - mean) / std of the context's current Vloc, FP32
int lpc_set_potential(rsg_ctx *ctx, const double *vloc_host) {
  LpcState &L = ctx->lpc;
  const int64_t N = ctx->N;
  std::vector<double> v(vloc_host, vloc_host + N);
  double m = 0.0, q = 0.0;
  for (double x : v) m += x;
  m /= double(N);
  for (double x : v) q += (x - m) * (x - m);
  const double sd = std::sqrt(q / double(N)) + 1e-30;
  std::vector<float> f(static_cast<size_t>(N));
  for (int64_t i = 0; i < N; ++i) f[size_t(i)] = float((v[size_t(i)] - m) / sd);
  if (L.vnorm.ensure(size_t(N) * 4) != hipSuccess) return ctx->fail(RSG_ERR_HIP, "lpc: allocation failed");
  RSG_HIP(hipMemcpyAsync(L.vnorm.p, f.data(), size_t(N) * 4, hipMemcpyHostToDevice, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  L.have_v = true;
  L.vN = N;
  return RSG_OK;
}

namespace {

int ensure_acts(rsg_ctx *ctx, int nsamp, bool training) {
  LpcState &L = ctx->lpc;
  const size_t N = size_t(ctx->N), ns = size_t(nsamp);
  bool ok = L.x0.ensure(ns * N * 4) == hipSuccess && L.a1.ensure(ns * N * 32) == hipSuccess &&
            L.a2.ensure(ns * N * 32) == hipSuccess && L.sval.ensure(ns * 4 + 256) == hipSuccess &&
            L.gg.ensure(ns * 8 + 256) == hipSuccess && L.y.ensure(ns * N * 4) == hipSuccess;
  if (training)
    ok = ok && L.dyb.ensure(ns * N * 2) == hipSuccess && L.dz1.ensure(ns * N * 32) == hipSuccess &&
         L.dz2.ensure(ns * N * 32) == hipSuccess;
  return ok ? RSG_OK : ctx->fail(RSG_ERR_HIP, "lpc: activation allocation failed");
}

template <int CIN, int KP, int MODE, typename T>
void launch_conv(rsg_ctx *ctx, int nsamp, const uint16_t *in, int64_t in_stride, const uint16_t *w,
                 const float *bias, uint16_t *out, const uint16_t *mask, float *yout, const T *gsrc, T *pg,
                 int64_t ldg, const float *sval) {
  const Grid3 g{ctx->n[0], ctx->n[1], ctx->n[2]};
  const int64_t per = 16 * kTiles * kWaves;
  dim3 grid(unsigned((ctx->N + per - 1) / per), nsamp);
  hipLaunchKernelGGL((k_conv<CIN, KP, MODE, T>), grid, dim3(64 * kWaves), 0, ctx->stream, g, ctx->N, in, in_stride,
                     w, bias, out, mask, yout, gsrc, pg, ldg, sval);
}

// forward through all layers for nsamp samples g[smp*ldg + p]; y -> L.y;
// pg (if non-null) <- g + s y
template <typename T>
int forward_t(rsg_ctx *ctx, const T *g, int64_t ldg, int nsamp, T *pg, bool training) {
  LpcState &L = ctx->lpc;
  if (!L.have_params || !L.have_v || L.vN != ctx->N)
    return ctx->fail(RSG_ERR_STATE, "lpc: parameters or potential not set for this grid");
  int rc = ensure_acts(ctx, nsamp, training);
  if (rc) return rc;
  const int64_t N = ctx->N;
  // s = rms(g); multidot takes blocks with leading dimension N, so a
  // padded ldg goes one sample per call
  {
    if (ldg == N) {
      const T *xp[1] = {g}, *yp[1] = {g};
      if ((rc = multidot(ctx, 1, xp, yp, nsamp, 1.0, L.gg.as<double>()))) return rc;
    } else {
      for (int s = 0; s < nsamp; ++s) {
        const T *xp[1] = {g + int64_t(s) * ldg}, *yp[1] = {g + int64_t(s) * ldg};
        if ((rc = multidot(ctx, 1, xp, yp, 1, 1.0, L.gg.as<double>() + s))) return rc;
      }
    }
  }
  hipLaunchKernelGGL(k_rms, dim3((nsamp + 63) / 64), dim3(64), 0, ctx->stream, L.gg.as<double>(), nsamp, N,
                     L.sval.as<float>());
  hipLaunchKernelGGL((k_prep<T>), dim3(unsigned((N + 255) / 256), nsamp), dim3(256), 0, ctx->stream, g, ldg,
                     L.vnorm.as<float>(), L.sval.as<float>(), N, L.x0.as<uint16_t>());
  const uint16_t *w1 = L.wbf.as<uint16_t>(), *w2 = w1 + 16 * KP1, *w3 = w2 + 16 * KP2;
  const float *b = L.bias.as<float>();
  launch_conv<2, KP1, kReluOut, T>(ctx, nsamp, L.x0.as<uint16_t>(), N * 2, w1, b, L.a1.as<uint16_t>(), nullptr,
                                   nullptr, nullptr, nullptr, 0, nullptr);
  launch_conv<16, KP2, kReluOut, T>(ctx, nsamp, L.a1.as<uint16_t>(), N * 16, w2, b + 16, L.a2.as<uint16_t>(),
                                    nullptr, nullptr, nullptr, nullptr, 0, nullptr);
  launch_conv<16, KP3, kFinalOut, T>(ctx, nsamp, L.a2.as<uint16_t>(), N * 16, w3, b + 32, nullptr, nullptr,
                                     L.y.as<float>(), g, pg, ldg, L.sval.as<float>());
  RSG_HIP(hipGetLastError());
  return RSG_OK;
}

template <int CIN, int KP, int COUT>
int weight_grad(rsg_ctx *ctx, int nsamp, const uint16_t *in, const uint16_t *dz, float *out) {
  LpcState &L = ctx->lpc;
  const Grid3 g{ctx->n[0], ctx->n[1], ctx->n[2]};
  const int64_t total = ctx->N * nsamp;
  const int nblk = int(std::min<int64_t>(512, (total + 4095) / 4096));
  const int64_t per = (total + nblk - 1) / nblk;
  if (L.part.ensure(size_t(nblk) * 16 * KP * 4) != hipSuccess) return ctx->fail(RSG_ERR_HIP, "lpc: allocation failed");
  hipLaunchKernelGGL((k_conv_dw<CIN, KP, COUT>), dim3(nblk), dim3(256), 0, ctx->stream, g, ctx->N, nsamp, in, dz,
                     per, L.part.as<float>());
  const int len = 16 * KP;
  hipLaunchKernelGGL(k_sum_parts, dim3((len + 255) / 256), dim3(256), 0, ctx->stream, L.part.as<float>(), nblk, len,
                     out);
  return RSG_OK;
}

}  // namespace

// bands per forward pass: activations take 68 bytes per point and band
static int apply_chunk(const rsg_ctx *ctx, int nb) {
  const size_t per = size_t(ctx->N) * 72;
  return int(std::max<size_t>(1, std::min<size_t>(size_t(nb), (size_t(4) << 30) / per)));
}

int lpc_apply(rsg_ctx *ctx, const void *g, void *pg, int64_t ld, int nb, int precision) {
  if (precision != 64 && precision != 32) return ctx->fail(RSG_ERR_ARG, "lpc: precision must be 64 or 32");
  if (ld < ctx->N || nb < 0) return ctx->fail(RSG_ERR_ARG, "lpc: bad ld or nb");
  const int ch = apply_chunk(ctx, std::max(nb, 1));
  for (int b0 = 0; b0 < nb; b0 += ch) {
    const int n = std::min(ch, nb - b0);
    int rc;
    if (precision == 64)
      rc = forward_t<double>(ctx, static_cast<const double *>(g) + int64_t(b0) * ld, ld, n,
                             static_cast<double *>(pg) + int64_t(b0) * ld, false);
    else
      rc = forward_t<float>(ctx, static_cast<const float *>(g) + int64_t(b0) * ld, ld, n,
                            static_cast<float *>(pg) + int64_t(b0) * ld, false);
    if (rc) return rc;
  }
  return RSG_OK;
}

// One training step on nsamp samples (FP32 residuals g and targets t, each
// N x nsamp column-major): forward, loss against kappa t, backward,
// optional Adam update.  grads (host, may be NULL) receives the gradient in
// the parameter layout.
int lpc_train_step(rsg_ctx *ctx, const float *g, const float *t, int nsamp, float kappa, float lr, double *loss,
                   float *grads) {
  LpcState &L = ctx->lpc;
  int rc = forward_t<float>(ctx, g, ctx->N, nsamp, nullptr, true);
  if (rc) return rc;
  const int64_t N = ctx->N;
  const int np = lpc_param_count();
  if (L.lsum.ensure(64) != hipSuccess || L.grad.ensure(size_t(16) * (KP1 + KP2 + KP3) * 4 + 256) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "lpc: allocation failed");
  RSG_HIP(hipMemsetAsync(L.lsum.p, 0, 64, ctx->stream));
  const double inv = 1.0 / (double(N) * nsamp);
  hipLaunchKernelGGL(k_loss, dim3(unsigned((N + 255) / 256), nsamp), dim3(256), 0, ctx->stream, t, L.y.as<float>(),
                     L.sval.as<float>(), kappa, N, inv, L.dyb.as<uint16_t>(), L.lsum.as<double>());
  const uint16_t *w1 = L.wbf.as<uint16_t>(), *w2 = w1 + 16 * KP1, *w3 = w2 + 16 * KP2, *w2t = w3 + 16 * KP3,
                 *w3t = w2t + 16 * KP2;
  // dz2 = (W3' * dy) masked by a2 > 0 ; dz1 = (W2' * dz2) masked by a1 > 0
  launch_conv<1, KP3T, kMaskedGrad, float>(ctx, nsamp, L.dyb.as<uint16_t>(), N, w3t, nullptr, L.dz2.as<uint16_t>(),
                                           L.a2.as<uint16_t>(), nullptr, nullptr, nullptr, 0, nullptr);
  launch_conv<16, KP2, kMaskedGrad, float>(ctx, nsamp, L.dz2.as<uint16_t>(), N * 16, w2t, nullptr,
                                           L.dz1.as<uint16_t>(), L.a1.as<uint16_t>(), nullptr, nullptr, nullptr, 0,
                                           nullptr);
  // weight gradients into grad = [G1 16 x KP1 | G2 16 x KP2 | G3 16 x KP3 (row 0)]
  float *G = L.grad.as<float>();
  if ((rc = weight_grad<2, KP1, 16>(ctx, nsamp, L.x0.as<uint16_t>(), L.dz1.as<uint16_t>(), G))) return rc;
  if ((rc = weight_grad<16, KP2, 16>(ctx, nsamp, L.a1.as<uint16_t>(), L.dz2.as<uint16_t>(), G + 16 * KP1))) return rc;
  if ((rc = weight_grad<16, KP3, 1>(ctx, nsamp, L.a2.as<uint16_t>(), L.dyb.as<uint16_t>(), G + 16 * (KP1 + KP2))))
    return rc;
  float *db = G + 16 * (KP1 + KP2 + KP3);
  RSG_HIP(hipMemsetAsync(db, 0, 33 * 4, ctx->stream));
  const int64_t tot = N * nsamp;
  const unsigned nbb = unsigned(std::min<int64_t>(1024, (tot + 255) / 256));
  hipLaunchKernelGGL(k_bias_grad<16>, dim3(nbb), dim3(256), 0, ctx->stream, L.dz1.as<uint16_t>(), tot, db);
  hipLaunchKernelGGL(k_bias_grad<16>, dim3(nbb), dim3(256), 0, ctx->stream, L.dz2.as<uint16_t>(), tot, db + 16);
  hipLaunchKernelGGL(k_bias_grad<1>, dim3(nbb), dim3(256), 0, ctx->stream, L.dyb.as<uint16_t>(), tot, db + 32);
  RSG_HIP(hipGetLastError());
  // gather into the parameter layout on the host (8241 floats)
  std::vector<float> Gh(size_t(16) * (KP1 + KP2 + KP3) + 33);
  RSG_HIP(hipMemcpyAsync(Gh.data(), G, Gh.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
  double lh = 0.0;
  RSG_HIP(hipMemcpyAsync(&lh, L.lsum.p, 8, hipMemcpyDeviceToHost, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  if (loss) *loss = lh;
  std::vector<float> gp(size_t(np), 0.f);
  const float *g1 = Gh.data(), *g2 = g1 + 16 * KP1, *g3 = g2 + 16 * KP2, *gb = g3 + 16 * KP3;
  for (int co = 0; co < 16; ++co)
    for (int ci = 0; ci < 2; ++ci)
      for (int tp = 0; tp < 27; ++tp) gp[size_t(OFF_W1 + (co * 2 + ci) * 27 + tp)] = g1[co * KP1 + tp * 2 + ci];
  for (int co = 0; co < 16; ++co)
    for (int ci = 0; ci < 16; ++ci)
      for (int tp = 0; tp < 27; ++tp) gp[size_t(OFF_W2 + (co * 16 + ci) * 27 + tp)] = g2[co * KP2 + tp * 16 + ci];
  for (int ci = 0; ci < 16; ++ci)
    for (int tp = 0; tp < 27; ++tp) gp[size_t(OFF_W3 + ci * 27 + tp)] = g3[0 * KP3 + tp * 16 + ci];
  for (int c = 0; c < 16; ++c) gp[size_t(OFF_B1 + c)] = gb[c];
  for (int c = 0; c < 16; ++c) gp[size_t(OFF_B2 + c)] = gb[16 + c];
  gp[size_t(OFF_B3)] = gb[32];
  if (grads) std::copy(gp.begin(), gp.end(), grads);
  if (lr > 0.f) {
    if (L.mom1.ensure(size_t(np) * 4) != hipSuccess || L.mom2.ensure(size_t(np) * 4) != hipSuccess)
      return ctx->fail(RSG_ERR_HIP, "lpc: allocation failed");
    if (L.adam_t == 0) {
      RSG_HIP(hipMemsetAsync(L.mom1.p, 0, size_t(np) * 4, ctx->stream));
      RSG_HIP(hipMemsetAsync(L.mom2.p, 0, size_t(np) * 4, ctx->stream));
    }
    ++L.adam_t;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, float(L.adam_t)), c2 = 1.f - std::pow(b2, float(L.adam_t));
    // the gradient in parameter layout goes up once; the update runs on the device
    RSG_HIP(hipMemcpyAsync(G, gp.data(), size_t(np) * 4, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_adam, dim3((np + 255) / 256), dim3(256), 0, ctx->stream, L.master.as<float>(), G,
                       L.mom1.as<float>(), L.mom2.as<float>(), np, lr, b1, b2, 1e-8f, c1, c2);
    std::vector<float> p(static_cast<size_t>(np));
    RSG_HIP(hipMemcpyAsync(p.data(), L.master.p, size_t(np) * 4, hipMemcpyDeviceToHost, ctx->stream));
    RSG_HIP(hipStreamSynchronize(ctx->stream));
    return lpc_set_params(ctx, p.data());
  }
  return RSG_OK;
}

int lpc_reset_optimizer(rsg_ctx *ctx) {
  ctx->lpc.adam_t = 0;
  return RSG_OK;
}

}  // namespace rsg
