// setup.hip -- the analytic GTH projector description built on the device
// (the set-up half of ps_nloc2_module.f90:42-1167 for l <= 1, n = 1 sets):
// per atom the 1-D Gaussian factor tables, the sphere points in grid order
// with their plane segments, per grid point the (at most four) spheres that
// hold it, and per 16-point tile the atoms that meet it.  The host computes
// only the per-atom parameters (analytic_groups) and the plane bookkeeping
// from per-plane point counts; the result equals the host builder's
// (build_gth_analytic) array for array, with the factor tables equal up to
// the rounding of exp.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "context.hpp"

namespace rsg {
namespace {

constexpr int kPlaneSegMaxDev = 1024;  // points per segment, as the host builder's kPlaneSegMax

// Per group as the kernels read it.
struct GroupDev {
  double R[3];
  double thr_s, thr_p;  // rs^2 + 1e-10, rp^2 + 1e-10 (-1: no such projector)
  double c2s, c2p;
  int ic[3];
  int box[3];     // candidate offsets per axis: 2 mm + 1
  int offbase;    // first entry of this group's sorted offset lists
};

// 1-D factor tables per group and axis at the image nearest the atom
// (the host builder's loop, one thread per entry).
__global__ void k_axis_tables(const GroupDev *__restrict__ grp, int ngrp, int n1, int n2, int n3, double a1, double a2,
                              double a3, double c1, double c2, double c3, double *__restrict__ es64,
                              double *__restrict__ ep64, double *__restrict__ dd64, float *__restrict__ es32,
                              float *__restrict__ ep32, float *__restrict__ dd32) {
#pragma clang fp contract(off)
  const int S = n1 + n2 + n3;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(ngrp) * S) return;
  const int g = int(e / S), k = int(e - int64_t(g) * S);
  const int t = k < n1 ? 0 : (k < n1 + n2 ? 1 : 2);
  const int i = k - (t == 0 ? 0 : (t == 1 ? n1 : n1 + n2));
  const int ng = t == 0 ? n1 : (t == 1 ? n2 : n3);
  const double at = t == 0 ? a1 : (t == 1 ? a2 : a3), ct = t == 0 ? c1 : (t == 1 ? c2 : c3);
  const GroupDev &G = grp[g];
  int id = i;
  while (id - G.ic[t] > ng / 2) id -= ng;
  while (G.ic[t] - id > ng / 2) id += ng;
  const double d = at * (double(id) * ct) - G.R[t];
  const double es = exp(-d * d * G.c2s), ep = exp(-d * d * G.c2p);
  es64[e] = es;
  ep64[e] = ep;
  dd64[e] = d;
  es32[e] = float(es);
  ep32[e] = float(ep);
  dd32[e] = float(d);
}

// Sphere membership of candidate (k1, k2, k3) of group G, with the host
// builder's arithmetic (no fused operations).  off: per axis the offsets in
// ascending order of the wrapped grid index; wrp: those wrapped indices.
struct Cand {
  bool s, p;
  int w[3];
};
__device__ inline Cand candidate(const GroupDev &G, const int *__restrict__ off, const int *__restrict__ wrp, int k1,
                                 int k2, int k3, double a1, double a2, double a3, double c1, double c2, double c3) {
#pragma clang fp contract(off)
  const int b1 = G.offbase, b2 = b1 + G.box[0], b3 = b2 + G.box[1];
  const int id1 = G.ic[0] + off[b1 + k1], id2 = G.ic[1] + off[b2 + k2], id3 = G.ic[2] + off[b3 + k3];
  const double d1 = double(id1) * c1, d2 = double(id2) * c2, d3 = double(id3) * c3;
  // x = aa0 d1 + aa3 d2 + aa6 d3 - R0 with aa3 = aa6 = 0 (orthorhombic);
  // the zero products keep the host's rounding sequence
  const double x = ((a1 * d1 + 0.0 * d2) + 0.0 * d3) - G.R[0];
  const double y = ((0.0 * d1 + a2 * d2) + 0.0 * d3) - G.R[1];
  const double z = ((0.0 * d1 + 0.0 * d2) + a3 * d3) - G.R[2];
  const double r2 = (x * x + y * y) + z * z;
  Cand c;
  c.s = r2 <= G.thr_s;
  c.p = r2 <= G.thr_p;
  c.w[0] = wrp[b1 + k1];
  c.w[1] = wrp[b2 + k2];
  c.w[2] = wrp[b3 + k3];
  return c;
}

constexpr int kSetupThreads = 256;

// Exclusive prefix of a 0/1 flag over the block, in thread order; *total
// gets the block's sum.
__device__ inline int block_rank(bool flag, int *total) {
  __shared__ int wsum[kSetupThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long m = __ballot(flag);
  const int in_wave = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[wv] = __popcll(m);
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kSetupThreads / 64; ++w) {
    before += w < wv ? wsum[w] : 0;
    all += wsum[w];
  }
  __syncthreads();
  *total = all;
  return before + in_wave;
}

// Points per (group, plane of the box): one block per (group, k3).
__global__ __launch_bounds__(kSetupThreads) void k_sphere_count(const GroupDev *__restrict__ grp, int b3max,
                                                                const int *__restrict__ off,
                                                                const int *__restrict__ wrp, double a1, double a2,
                                                                double a3, double c1, double c2, double c3,
                                                                int *__restrict__ cnt) {
  const int g = blockIdx.x / b3max, k3 = blockIdx.x - g * b3max;
  const GroupDev G = grp[g];
  if (k3 >= G.box[2]) {
    if (threadIdx.x == 0) cnt[blockIdx.x] = 0;
    return;
  }
  const int n12 = G.box[0] * G.box[1];
  int sum = 0;
  for (int c0 = 0; c0 < n12; c0 += kSetupThreads) {
    const int q = c0 + threadIdx.x;
    bool f = false;
    if (q < n12) {
      const int k2 = q / G.box[0], k1 = q - k2 * G.box[0];
      const Cand c = candidate(G, off, wrp, k1, k2, k3, a1, a2, a3, c1, c2, c3);
      f = c.s || c.p;
    }
    int tot;
    (void)block_rank(f, &tot);
    sum += tot;
  }
  if (threadIdx.x == 0) cnt[blockIdx.x] = sum;
}

// The points themselves, in grid order from each (group, plane) offset, and
// their sphere entries in the per-point hit words (slot order fixed later).
__global__ __launch_bounds__(kSetupThreads) void k_sphere_fill(
    const GroupDev *__restrict__ grp, int b3max, const int *__restrict__ off, const int *__restrict__ wrp, double a1,
    double a2, double a3, double c1, double c2, double c3, int n1, int n2, const int *__restrict__ poff,
    uint32_t *__restrict__ pts, unsigned *__restrict__ hcnt, unsigned long long *__restrict__ hits,
    int *__restrict__ overflow) {
  const int g = blockIdx.x / b3max, k3 = blockIdx.x - g * b3max;
  const GroupDev G = grp[g];
  if (k3 >= G.box[2]) return;
  const int n12 = G.box[0] * G.box[1];
  int base = poff[blockIdx.x];
  for (int c0 = 0; c0 < n12; c0 += kSetupThreads) {
    const int q = c0 + threadIdx.x;
    Cand c{};
    bool f = false;
    if (q < n12) {
      const int k2 = q / G.box[0], k1 = q - k2 * G.box[0];
      c = candidate(G, off, wrp, k1, k2, k3, a1, a2, a3, c1, c2, c3);
      f = c.s || c.p;
    }
    int tot;
    const int r = block_rank(f, &tot);
    if (f) {
      pts[base + r] = uint32_t(c.w[0]) | (uint32_t(c.w[1]) << 10) | (uint32_t(c.w[2]) << 20) |
                      (uint32_t(c.s) << 30) | (uint32_t(c.p) << 31);
      const int64_t lin = c.w[0] + int64_t(c.w[1]) * n1 + int64_t(c.w[2]) * n1 * n2;
      const unsigned slot = atomicAdd(&hcnt[lin], 1u);
      if (slot < 4) {
        const unsigned long long e = (unsigned long long)(g + 1) | ((unsigned long long)c.s << 14) |
                                     ((unsigned long long)c.p << 15);
        atomicOr(&hits[lin], e << (16 * slot));
      } else {
        atomicOr(overflow, 1);
      }
    }
    base += tot;
  }
}

// Per point: the hit entries in ascending group order (the host builder
// adds groups in order).
__global__ void k_hits_sort(unsigned long long *__restrict__ hits, int64_t N) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= N) return;
  unsigned long long h = hits[j];
  if (!h) return;
  unsigned e[4];
  int n = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const unsigned v = unsigned(h >> (16 * q)) & 0xffffu;
    if (v) e[n++] = v;
  }
  for (int a = 1; a < n; ++a)
    for (int b = a; b > 0 && (e[b - 1] & 0x3fffu) > (e[b] & 0x3fffu); --b) {
      const unsigned t = e[b - 1];
      e[b - 1] = e[b];
      e[b] = t;
    }
  h = 0;
  for (int q = 0; q < n; ++q) h |= (unsigned long long)e[q] << (16 * q);
  hits[j] = h;
}

// One wave per 16-point tile: its 64 hit slots sorted by group (bitonic
// sort across the lanes), duplicates dropped.  Count pass (tile_g null) or
// fill pass.
__global__ __launch_bounds__(256) void k_tile_groups(const unsigned long long *__restrict__ hits, int64_t N,
                                                     int64_t ntile, int *__restrict__ tcnt,
                                                     const int *__restrict__ tile_off, int *__restrict__ tile_g) {
  const int64_t tile = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (tile >= ntile) return;  // whole wave
  const int lane = threadIdx.x & 63;
  const int64_t j = tile * 16 + (lane >> 2);
  const unsigned long long h = j < N ? hits[j] : 0ull;
  const unsigned v = unsigned(h >> (16 * (lane & 3))) & 0x3fffu;
  int key = v ? int(v) - 1 : 0x7fffffff;  // group, or past every group
  for (int k = 2; k <= 64; k <<= 1)
    for (int s = k >> 1; s > 0; s >>= 1) {
      const int other = __shfl_xor(key, s, 64);
      const bool up = (lane & k) == 0, lower = (lane & s) == 0;
      key = (lower == up) ? min(key, other) : max(key, other);
    }
  const int prev = __shfl_up(key, 1, 64);
  const bool keep = key != 0x7fffffff && (lane == 0 || prev != key);
  const unsigned long long m = __ballot(keep);
  if (!tile_g) {
    if (lane == 0) tcnt[tile] = __popcll(m);
    return;
  }
  if (keep) tile_g[tile_off[tile] + __popcll(m & ((1ull << lane) - 1ull))] = key;
}

}  // namespace

int build_gth_analytic_device(rsg_ctx *ctx, int natom, const double *frac, const int *kind, int nkind,
                              const int *norb, const int *lo, const int *no, const double *rps, const double *hnl) {
  AnalyticGroups P;
  int rc = analytic_groups(ctx, natom, frac, kind, nkind, norb, lo, no, rps, hnl, P);
  if (rc) return rc;
  const int ngrp = int(P.g.size());
  if (ngrp == 0) return RSG_ERR_UNSUPPORTED;  // no atom carries projectors: the table form covers it
  const int *ng = ctx->n;
  const int S = P.S;
  const int64_t N = ctx->N;
  const double a1 = ctx->aa[0], a2 = ctx->aa[4], a3 = ctx->aa[8];
  const double c1 = 1.0 / ng[0], c2 = 1.0 / ng[1], c3 = 1.0 / ng[2];
  hipStream_t st = ctx->stream;
  // per group and axis: candidate offsets in ascending wrapped index
  std::vector<GroupDev> gd(static_cast<size_t>(ngrp));
  std::vector<int> offs, wraps;
  int b3max = 0;
  for (int g = 0; g < ngrp; ++g) {
    const AnalyticGroup &A = P.g[size_t(g)];
    GroupDev &G = gd[size_t(g)];
    for (int t = 0; t < 3; ++t) {
      G.R[t] = A.R[t];
      G.ic[t] = A.ic[t];
      G.box[t] = 2 * A.mm[t] + 1;
    }
    G.thr_s = A.has_s ? A.rs * A.rs + 1e-10 : -1.0;
    G.thr_p = A.has_p ? A.rp * A.rp + 1e-10 : -1.0;
    G.c2s = A.c2s;
    G.c2p = A.c2p;
    G.offbase = int(offs.size());
    for (int t = 0; t < 3; ++t) {
      std::vector<std::pair<int, int>> ow;
      for (int o = -A.mm[t]; o <= A.mm[t]; ++o) ow.emplace_back(((A.ic[t] + o) % ng[t] + ng[t]) % ng[t], o);
      std::stable_sort(ow.begin(), ow.end(), [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
        return x.first < y.first;
      });
      for (const auto &p : ow) {
        wraps.push_back(p.first);
        offs.push_back(p.second);
      }
    }
    b3max = std::max(b3max, G.box[2]);
  }
  // device inputs and outputs
  DevBuf &tmp = ctx->setup_tmp;
  const size_t nblk = size_t(ngrp) * b3max;
  const size_t gbytes = gd.size() * sizeof(GroupDev), obytes = offs.size() * sizeof(int);
  const size_t o_g = 0, o_off = (gbytes + 255) / 256 * 256, o_wrp = o_off + (obytes + 255) / 256 * 256,
               o_cnt = o_wrp + (obytes + 255) / 256 * 256, o_hcnt = o_cnt + (nblk * 4 + 255) / 256 * 256,
               o_ovf = o_hcnt + (size_t(N) * 4 + 255) / 256 * 256, o_tcnt = o_ovf + 256,
               o_end = o_tcnt + (size_t((N + 15) / 16) * 4 + 255) / 256 * 256;
  if (tmp.ensure(o_end) != hipSuccess) return ctx->fail(RSG_ERR_HIP, "build_gth: set-up scratch allocation failed");
  char *tb = static_cast<char *>(tmp.p);
  GroupDev *dgrp = reinterpret_cast<GroupDev *>(tb + o_g);
  int *doff = reinterpret_cast<int *>(tb + o_off), *dwrp = reinterpret_cast<int *>(tb + o_wrp);
  int *dcnt = reinterpret_cast<int *>(tb + o_cnt);
  unsigned *hcnt = reinterpret_cast<unsigned *>(tb + o_hcnt);
  int *ovf = reinterpret_cast<int *>(tb + o_ovf);
  int *tcnt = reinterpret_cast<int *>(tb + o_tcnt);
  RSG_HIP(hipMemcpyAsync(dgrp, gd.data(), gbytes, hipMemcpyHostToDevice, st));
  RSG_HIP(hipMemcpyAsync(doff, offs.data(), obytes, hipMemcpyHostToDevice, st));
  RSG_HIP(hipMemcpyAsync(dwrp, wraps.data(), obytes, hipMemcpyHostToDevice, st));
  AnalyticDev<double> &d64 = ctx->ana64;
  AnalyticDev<float> &d32 = ctx->ana32;
  const size_t nS = size_t(ngrp) * S;
  for (DevBuf *b : {&d64.es, &d64.ep, &d64.dd})
    if (b->ensure(nS * 8) != hipSuccess) return ctx->fail(RSG_ERR_HIP, "build_gth: allocation failed");
  for (DevBuf *b : {&d32.es, &d32.ep, &d32.dd})
    if (b->ensure(nS * 4) != hipSuccess) return ctx->fail(RSG_ERR_HIP, "build_gth: allocation failed");
  hipLaunchKernelGGL(k_axis_tables, dim3(unsigned((nS + 255) / 256)), dim3(256), 0, st, dgrp, ngrp, ng[0],
                     ng[1], ng[2], a1, a2, a3, c1, c2, c3, d64.es.as<double>(), d64.ep.as<double>(),
                     d64.dd.as<double>(), d32.es.as<float>(), d32.ep.as<float>(), d32.dd.as<float>());
  hipLaunchKernelGGL(k_sphere_count, dim3(unsigned(nblk)), dim3(kSetupThreads), 0, st, dgrp, b3max, doff, dwrp, a1, a2,
                     a3, c1, c2, c3, dcnt);
  RSG_HIP(hipGetLastError());
  std::vector<int> cnt(nblk);
  RSG_HIP(hipMemcpyAsync(cnt.data(), dcnt, nblk * 4, hipMemcpyDeviceToHost, st));
  RSG_HIP(hipStreamSynchronize(st));
  // point offsets per (group, plane); segments: runs of one plane, at most
  // kPlaneSegMaxDev points each; launch order by plane, then atom
  std::vector<int> poff(nblk, 0), seg_beg, seg_end, seg_grp, grp_seg_off{0}, seg_z;
  int64_t E = 0;
  for (int g = 0; g < ngrp; ++g) {
    const GroupDev &G = gd[size_t(g)];
    for (int k3 = 0; k3 < G.box[2]; ++k3) {
      const size_t b = size_t(g) * b3max + size_t(k3);
      poff[b] = int(E);
      for (int s = 0; s < cnt[b]; s += kPlaneSegMaxDev) {
        seg_beg.push_back(int(E) + s);
        seg_end.push_back(int(E) + std::min(cnt[b], s + kPlaneSegMaxDev));
        seg_grp.push_back(g);
        seg_z.push_back(wraps[size_t(G.offbase + G.box[0] + G.box[1] + k3)]);
      }
      E += cnt[b];
    }
    grp_seg_off.push_back(int(seg_beg.size()));
  }
  if (E > (int64_t(1) << 31) - 1) return RSG_ERR_UNSUPPORTED;
  const int nseg = int(seg_beg.size());
  std::vector<int> seg_order(static_cast<size_t>(nseg));
  for (int i = 0; i < nseg; ++i) seg_order[size_t(i)] = i;
  std::stable_sort(seg_order.begin(), seg_order.end(), [&](int x, int y) { return seg_z[size_t(x)] < seg_z[size_t(y)]; });
  int *dpoff = dcnt;  // the counts are no longer needed on the device
  RSG_HIP(hipMemcpyAsync(dpoff, poff.data(), nblk * 4, hipMemcpyHostToDevice, st));
  if (d64.pts.ensure(size_t(std::max<int64_t>(E, 1)) * 4) != hipSuccess || d64.hits.ensure(size_t(N) * 8) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "build_gth: allocation failed");
  RSG_HIP(hipMemsetAsync(hcnt, 0, size_t(N) * 4, st));
  RSG_HIP(hipMemsetAsync(ovf, 0, 4, st));
  RSG_HIP(hipMemsetAsync(d64.hits.p, 0, size_t(N) * 8, st));
  hipLaunchKernelGGL(k_sphere_fill, dim3(unsigned(nblk)), dim3(kSetupThreads), 0, st, dgrp, b3max, doff, dwrp, a1, a2,
                     a3, c1, c2, c3, ng[0], ng[1], dpoff, d64.pts.as<uint32_t>(), hcnt,
                     d64.hits.as<unsigned long long>(), ovf);
  hipLaunchKernelGGL(k_hits_sort, dim3(unsigned((N + 255) / 256)), dim3(256), 0, st,
                     d64.hits.as<unsigned long long>(), N);
  RSG_HIP(hipGetLastError());
  int overflow = 0;
  RSG_HIP(hipMemcpyAsync(&overflow, ovf, 4, hipMemcpyDeviceToHost, st));
  RSG_HIP(hipStreamSynchronize(st));
  if (overflow) return RSG_ERR_UNSUPPORTED;  // more than four spheres at one point
  // atoms per 16-point tile
  const int64_t ntile = (N + 15) / 16;
  if (d64.tile_off.ensure(size_t(ntile + 1) * 4) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "build_gth: allocation failed");
  int *toff = d64.tile_off.as<int>();
  RSG_HIP(hipMemsetAsync(toff, 0, 4, st));
  const dim3 tgrid(unsigned((ntile + 3) / 4));
  hipLaunchKernelGGL(k_tile_groups, tgrid, dim3(256), 0, st, d64.hits.as<unsigned long long>(), N, ntile, tcnt,
                     nullptr, nullptr);
  size_t scan_bytes = 0;
  RSG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, tcnt, toff + 1, int(ntile), st));
  if (ctx->setup_scan.ensure(std::max<size_t>(scan_bytes, 256)) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "build_gth: allocation failed");
  RSG_HIP(hipcub::DeviceScan::InclusiveSum(ctx->setup_scan.p, scan_bytes, tcnt, toff + 1, int(ntile), st));
  int ntg = 0;
  RSG_HIP(hipMemcpyAsync(&ntg, toff + ntile, 4, hipMemcpyDeviceToHost, st));
  RSG_HIP(hipStreamSynchronize(st));
  if (d64.tile_g.ensure(size_t(std::max(ntg, 1)) * 4) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "build_gth: allocation failed");
  hipLaunchKernelGGL(k_tile_groups, tgrid, dim3(256), 0, st, d64.hits.as<unsigned long long>(), N, ntile, nullptr,
                     toff, d64.tile_g.as<int>());
  RSG_HIP(hipGetLastError());
  // the small host-side arrays: amplitudes, scales, segments
  std::vector<double> amp(size_t(ngrp) * 4), scale(size_t(ngrp) * 4);
  for (int g = 0; g < ngrp; ++g) {
    const AnalyticGroup &A = P.g[size_t(g)];
    amp[4 * size_t(g)] = A.amp_s;
    amp[4 * size_t(g) + 1] = amp[4 * size_t(g) + 2] = amp[4 * size_t(g) + 3] = A.amp_p;
    scale[4 * size_t(g)] = A.has_s ? A.sg_s * ctx->dv : 0.0;
    scale[4 * size_t(g) + 1] = scale[4 * size_t(g) + 2] = scale[4 * size_t(g) + 3] = A.has_p ? A.sg_p * ctx->dv : 0.0;
  }
  std::vector<float> amp32(amp.begin(), amp.end()), scale32(scale.begin(), scale.end());
  auto up = [&](DevBuf &b, const void *src, size_t bytes) -> int {
    if (b.ensure(std::max<size_t>(bytes, 4)) != hipSuccess) return RSG_ERR_HIP;
    if (bytes && hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st) != hipSuccess) return RSG_ERR_HIP;
    return RSG_OK;
  };
  int e = 0;
  e |= up(d64.amp, amp.data(), amp.size() * 8);
  e |= up(d64.scale, scale.data(), scale.size() * 8);
  e |= up(d32.amp, amp32.data(), amp32.size() * 4);
  e |= up(d32.scale, scale32.data(), scale32.size() * 4);
  e |= up(d64.seg_beg, seg_beg.data(), seg_beg.size() * 4);
  e |= up(d64.seg_end, seg_end.data(), seg_end.size() * 4);
  e |= up(d64.seg_grp, seg_grp.data(), seg_grp.size() * 4);
  e |= up(d64.grp_seg_off, grp_seg_off.data(), grp_seg_off.size() * 4);
  e |= up(d64.seg_order, seg_order.data(), seg_order.size() * 4);
  // the FP32 description shares the integer arrays' contents
  auto dup = [&](DevBuf &dst, const DevBuf &src, size_t bytes) -> int {
    if (dst.ensure(std::max<size_t>(bytes, 4)) != hipSuccess) return RSG_ERR_HIP;
    if (bytes && hipMemcpyAsync(dst.p, src.p, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) return RSG_ERR_HIP;
    return RSG_OK;
  };
  e |= dup(d32.pts, d64.pts, size_t(E) * 4);
  e |= dup(d32.hits, d64.hits, size_t(N) * 8);
  e |= dup(d32.tile_off, d64.tile_off, size_t(ntile + 1) * 4);
  e |= dup(d32.tile_g, d64.tile_g, size_t(ntg) * 4);
  e |= dup(d32.seg_beg, d64.seg_beg, seg_beg.size() * 4);
  e |= dup(d32.seg_end, d64.seg_end, seg_end.size() * 4);
  e |= dup(d32.seg_grp, d64.seg_grp, seg_grp.size() * 4);
  e |= dup(d32.grp_seg_off, d64.grp_seg_off, grp_seg_off.size() * 4);
  e |= dup(d32.seg_order, d64.seg_order, seg_order.size() * 4);
  if (e) return ctx->fail(RSG_ERR_HIP, "build_gth: upload of analytic projector data failed");
  d64.ngroup = d32.ngroup = ngrp;
  d64.nseg = d32.nseg = nseg;
  d64.S = d32.S = S;
  d64.ntile = d32.ntile = int(ntile);
  // row tiles of the projection are built from these arrays on first use
  for (int *v : {&d64.rt_ty, &d64.rt_zr, &d64.rt_ny, &d64.rt_nz, &d64.rt_nslot_max, &d64.rt_nslot, &d32.rt_ty,
                 &d32.rt_zr, &d32.rt_ny, &d32.rt_nz, &d32.rt_nslot_max, &d32.rt_nslot})
    *v = 0;
  ctx->ana_rows_pending = true;
  RSG_HIP(hipStreamSynchronize(st));  // the host vectors above go out of scope
  ctx->have_ana = ngrp > 0;
  return ngrp > 0 ? RSG_OK : RSG_ERR_UNSUPPORTED;
}

}  // namespace rsg
