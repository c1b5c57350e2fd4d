// tables.cpp -- host-side operator setup: FD weights, GTH projector tables
// and the two device layouts of the nonlocal operator.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

#include "context.hpp"

namespace rsg {

hipError_t DevBuf::ensure(size_t need) {
  if (need <= bytes && p) return hipSuccess;
  DeviceOpGuard guard;
  release();
  if (need == 0) need = 256;
  hipError_t e = hipMalloc(&p, need);
  if (e != hipSuccess) {
    p = nullptr;
    bytes = 0;
    return e;
  }
  bytes = need;
  return hipSuccess;
}

void DevBuf::release() {
  if (!p) return;
  DeviceOpGuard guard;
  (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

// Central-difference second-derivative weights of half-width md in closed
// form: w(+-m) = 2 (-1)^(m+1) (md!)^2 / (m^2 (md-m)! (md+m)!),
// w(0) = -2 sum 1/m^2.  Same weights as get_coef_lapla_fd
// (fd_module.f90:18-54), which builds them from Lagrange products.
void fd_lap_closed_form(int md, double *lap) {
  auto lfact = [](int k) { return std::lgamma(double(k) + 1.0); };
  double w0 = 0.0;
  for (int m = 1; m <= md; ++m) {
    double mag = std::exp(2.0 * lfact(md) - lfact(md - m) - lfact(md + m));
    double w = 2.0 * ((m % 2) ? 1.0 : -1.0) * mag / double(m * m);
    lap[md + m] = lap[md - m] = w;
    w0 -= 2.0 / double(m * m);
  }
  lap[md] = w0;
}

namespace {

const double kPi = 3.141592653589793238;

int nint_away(double x) { return int(x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5)); }

double powi(double r, int e) {
  double v = 1.0;
  for (int i = 0; i < e; ++i) v *= r;
  return v;
}

// Gamma(L + 2n - 1/2) as sqrt(pi) * prod_{i=1}^{L+2n-1} (i - 1/2)
double gamma_half(int L, int n) {
  double g = std::sqrt(std::acos(-1.0));
  for (int i = 1; i <= L + 2 * n - 1; ++i) g *= (i - 0.5);
  return g;
}

// Projector cutoff: first r = i*0.01 (i <= 3000) where the unscaled GTH
// radial value is below 1e-10 (ps_nloc_gth_module.f90:137-180).
double gth_cutoff(int L, int n, double rl) {
  const double dr = 0.01, ep = 1e-10;
  double c1 = std::sqrt(2.0) / (std::pow(rl, L + 2 * n - 0.5) * std::sqrt(gamma_half(L, n)));
  double c2 = 0.5 / (rl * rl);
  for (int i = 1; i <= 3000; ++i) {
    double r = i * dr;
    if (std::fabs(c1 * powi(r, L + 2 * n - 2) * std::exp(-r * r * c2)) < ep) return r;
  }
  return 0.0;
}

// Associated Legendre P_L^M(x) by upward recursion in L.
double assoc_legendre(int L, int M, double x) {
  if (std::fabs(std::fabs(x) - 1.0) < 1e-13) x = x >= 0 ? 1.0 : -1.0;
  double pmm = 1.0;
  if (M > 0) {
    double s = std::sqrt((1.0 - x) * (1.0 + x)), f = 1.0;
    for (int i = 1; i <= M; ++i) {
      pmm = -pmm * f * s;
      f += 2.0;
    }
  }
  if (L == M) return pmm;
  double p1 = x * (2.0 * M + 1.0) * pmm;
  if (L == M + 1) return p1;
  double pl = 0.0;
  for (int l = M + 2; l <= L; ++l) {
    pl = (x * (2.0 * l - 1.0) * p1 - (l + M - 1.0) * pmm) / (l - M);
    pmm = p1;
    p1 = pl;
  }
  return pl;
}

// Real spherical harmonic with RSDFT's sign/ordering convention
// (ylm_module.f90:11-73).
double real_ylm(double x, double y, double z, int L, int m) {
  const int M = std::abs(m);
  const double pi4 = 4.0 * kPi;
  if (L == 0) return 1.0 / std::sqrt(pi4);
  if (x == 0.0 && y == 0.0 && z == 0.0) return 0.0;
  double c = 1.0;
  for (int k = L - M + 1; k <= L + M; ++k) c *= double(k);
  const double clm = std::sqrt((2.0 * L + 1.0) / (c * pi4));
  const double r = std::sqrt(x * x + y * y + z * z), ct = z / r;
  double phi;
  if (std::fabs(x) < 1e-10) {
    phi = 0.5 * kPi * (y >= 0 ? 1.0 : -1.0);
  } else {
    phi = std::atan(y / x);
    if (x < 0) phi += kPi;
  }
  const double p = assoc_legendre(L, M, ct);
  if (m > 0) return std::sqrt(2.0) * clm * p * std::cos(phi * M);
  if (m == 0) return clm * p;
  if (M % 2 == 0) return -std::sqrt(2.0) * clm * p * std::sin(phi * M);
  return std::sqrt(2.0) * clm * p * std::sin(phi * M);
}

struct SpherePoint {
  int i0[3];
  int k[3];
  double v;
};

}  // namespace

int build_gth_tables(const rsg_ctx *ctx, int natom, const double *frac,
                     const int *kind, int nkind, const int *norb,
                     const int *lo, const int *no, const double *rps,
                     const double *hnl, HostTables &out) {
  const int *ng = ctx->n;
  const double *aa = ctx->aa;
  std::vector<double> rcut(size_t(nkind) * 4, 0.0);
  for (int ik = 0; ik < nkind; ++ik) {
    if (norb[ik] < 0 || norb[ik] > 4) return RSG_ERR_ARG;
    for (int io = 0; io < norb[ik]; ++io) {
      int q = ik * 4 + io;
      if (lo[q] < 0 || lo[q] > 3 || no[q] < 1 || !(rps[q] > 0)) return RSG_ERR_ARG;
      rcut[q] = gth_cutoff(lo[q], no[q], rps[q]);
    }
  }
  for (int a = 0; a < natom; ++a)
    if (kind[a] < 0 || kind[a] >= nkind) return RSG_ERR_ARG;
  const double c[3] = {1.0 / ng[0], 1.0 / ng[1], 1.0 / ng[2]};
  const int64_t n12 = int64_t(ng[0]) * ng[1];

  out = HostTables();
  std::vector<std::vector<int>> lma_pts;
  std::vector<std::vector<double>> lma_val;
  std::vector<int> marker(size_t(ctx->N), 0);

  for (int a = 0; a < natom; ++a) {
    const double *fa = frac + 3 * a;
    const double R[3] = {aa[0] * fa[0] + aa[3] * fa[1] + aa[6] * fa[2],
                         aa[1] * fa[0] + aa[4] * fa[1] + aa[7] * fa[2],
                         aa[2] * fa[0] + aa[5] * fa[1] + aa[8] * fa[2]};
    const int ic[3] = {nint_away(fa[0] * ng[0]), nint_away(fa[1] * ng[1]), nint_away(fa[2] * ng[2])};
    const int ik = kind[a];
    for (int io = 0; io < norb[ik]; ++io) {
      const int q = ik * 4 + io;
      const int L = lo[q], nn = no[q];
      const double rl = rps[q], rc2 = rcut[q] * rcut[q];
      const double c1 = std::sqrt(std::fabs(hnl[q])) * std::sqrt(2.0) /
                        (std::pow(rl, L + 2 * nn - 0.5) * std::sqrt(gamma_half(L, nn)));
      const double c2 = 0.5 / (rl * rl);
      // candidate offsets around the nearest grid point, i3 outer, i1 inner
      int mm[3];
      for (int t = 0; t < 3; ++t) mm[t] = int(std::ceil(rcut[q] / ctx->hgrid[t])) + 2;
      std::vector<SpherePoint> pts;
      for (int o3 = -mm[2]; o3 <= mm[2]; ++o3)
        for (int o2 = -mm[1]; o2 <= mm[1]; ++o2)
          for (int o1 = -mm[0]; o1 <= mm[0]; ++o1) {
            const int id[3] = {ic[0] + o1, ic[1] + o2, ic[2] + o3};
            SpherePoint p;
            for (int t = 0; t < 3; ++t) {
              int kq = id[t] / ng[t];
              if (id[t] < 0) kq = (id[t] + 1) / ng[t] - 1;
              p.k[t] = kq;
              p.i0[t] = id[t] - kq * ng[t];
            }
            const double d1 = id[0] * c[0], d2 = id[1] * c[1], d3 = id[2] * c[2];
            const double x = aa[0] * d1 + aa[3] * d2 + aa[6] * d3 - R[0];
            const double y = aa[1] * d1 + aa[4] * d2 + aa[7] * d3 - R[1];
            const double z = aa[2] * d1 + aa[5] * d2 + aa[8] * d3 - R[2];
            const double r2 = x * x + y * y + z * z;
            if (r2 > rc2 + 1e-10) continue;
            p.v = c1 * powi(std::sqrt(r2), L + 2 * nn - 2) * std::exp(-r2 * c2);
            pts.push_back(p);
          }
      if (pts.empty()) continue;
      for (int m = -L; m <= L; ++m) {
        std::vector<int> idx;
        std::vector<double> val;
        for (const SpherePoint &p : pts) {
          const double d1 = c[0] * p.i0[0] + p.k[0], d2 = c[1] * p.i0[1] + p.k[1],
                       d3 = c[2] * p.i0[2] + p.k[2];
          const double x = aa[0] * d1 + aa[3] * d2 + aa[6] * d3 - R[0];
          const double y = aa[1] * d1 + aa[4] * d2 + aa[7] * d3 - R[1];
          const double z = aa[2] * d1 + aa[5] * d2 + aa[8] * d3 - R[2];
          const double u = p.v * real_ylm(x, y, z, L, m);
          const int64_t lin = p.i0[0] + int64_t(p.i0[1]) * ng[0] + int64_t(p.i0[2]) * n12;
          const int seen = marker[size_t(lin)];
          if (seen == 0 && std::fabs(u) > 1e-14) {
            idx.push_back(int(lin));
            val.push_back(u);
            marker[size_t(lin)] = int(idx.size());
          } else if (seen != 0) {
            val[size_t(seen - 1)] += u;  // periodic image of a point already held
          }
        }
        for (int v : idx) marker[size_t(v)] = 0;
        std::vector<int> order(idx.size());
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int x, int y) { return idx[x] < idx[y]; });
        std::vector<int> si(idx.size());
        std::vector<double> sv(idx.size());
        for (size_t t = 0; t < order.size(); ++t) {
          si[t] = idx[order[t]];
          sv[t] = val[order[t]];
        }
        lma_pts.push_back(std::move(si));
        lma_val.push_back(std::move(sv));
        out.iuv.push_back(hnl[q] < 0.0 ? -1 : 1);
        out.amap.push_back(a);
      }
    }
  }
  out.nzlma = int(lma_pts.size());
  out.maxmjj = 0;
  for (auto &v : lma_pts) out.maxmjj = std::max(out.maxmjj, int(v.size()));
  out.mjj.resize(out.nzlma);
  out.jjp.assign(size_t(out.nzlma) * out.maxmjj, 0);
  out.uvk.assign(size_t(out.nzlma) * out.maxmjj, 0.0);
  for (int l = 0; l < out.nzlma; ++l) {
    out.mjj[l] = int(lma_pts[l].size());
    std::copy(lma_pts[l].begin(), lma_pts[l].end(), out.jjp.begin() + size_t(l) * out.maxmjj);
    std::copy(lma_val[l].begin(), lma_val[l].end(), out.uvk.begin() + size_t(l) * out.maxmjj);
  }
  return RSG_OK;
}

namespace {

constexpr int kSegLen = 4096;       // points per gather segment
constexpr int kPlaneSegMax = 1024;  // points per analytic segment (one plane)

struct HostLayouts {
  std::vector<int> grp_lma0, grp_nl, grp_seg_off{0}, seg_beg, seg_end, seg_grp, grp_pts, lma_grp;
  std::vector<int> sl_off, sl_len, sl_lma;
  std::vector<double> grp_coef, scale, sl_uv;
};

template <typename T>
int upload_layouts(rsg_ctx *ctx, NonlocalDev<T> &d, const HostLayouts &h) {
  // stream-ordered copies (a plain hipMemcpy waits for every stream of the
  // device), synchronised before the host vectors go out of scope
  auto up = [&](DevBuf &b, const void *src, size_t bytes) -> int {
    if (b.ensure(bytes) != hipSuccess) return RSG_ERR_HIP;
    if (bytes && hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
      return RSG_ERR_HIP;
    return RSG_OK;
  };
  std::vector<T> c(h.grp_coef.begin(), h.grp_coef.end()), s(h.scale.begin(), h.scale.end()),
      u(h.sl_uv.begin(), h.sl_uv.end());
  d.ngroup = int(h.grp_lma0.size());
  d.nseg = int(h.seg_beg.size());
  int rc = 0;
  rc |= up(d.grp_lma0, h.grp_lma0.data(), h.grp_lma0.size() * sizeof(int));
  rc |= up(d.grp_nl, h.grp_nl.data(), h.grp_nl.size() * sizeof(int));
  rc |= up(d.grp_seg_off, h.grp_seg_off.data(), h.grp_seg_off.size() * sizeof(int));
  rc |= up(d.seg_beg, h.seg_beg.data(), h.seg_beg.size() * sizeof(int));
  rc |= up(d.seg_end, h.seg_end.data(), h.seg_end.size() * sizeof(int));
  rc |= up(d.seg_grp, h.seg_grp.data(), h.seg_grp.size() * sizeof(int));
  rc |= up(d.grp_pts, h.grp_pts.data(), h.grp_pts.size() * sizeof(int));
  rc |= up(d.grp_coef, c.data(), c.size() * sizeof(T));
  rc |= up(d.lma_grp, h.lma_grp.data(), h.lma_grp.size() * sizeof(int));
  rc |= up(d.scale, s.data(), s.size() * sizeof(T));
  rc |= up(d.sl_off, h.sl_off.data(), h.sl_off.size() * sizeof(int));
  rc |= up(d.sl_len, h.sl_len.data(), h.sl_len.size() * sizeof(int));
  rc |= up(d.sl_lma, h.sl_lma.data(), h.sl_lma.size() * sizeof(int));
  rc |= up(d.sl_uv, u.data(), u.size() * sizeof(T));
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = RSG_ERR_HIP;
  return rc ? ctx->fail(RSG_ERR_HIP, "upload of projector tables failed") : RSG_OK;
}

}  // namespace

// Build the atom-grouped gather layout and the per-point scatter CSR from
// the reference-layout tables in ctx->tables, then upload FP64 and FP32
// copies.
int upload_nonlocal(rsg_ctx *ctx) {
  const HostTables &t = ctx->tables;
  const int nz = t.nzlma, mx = t.maxmjj;
  HostLayouts h;
  h.lma_grp.assign(size_t(nz), 0);
  std::vector<int> mark(size_t(ctx->N), -1);
  for (int l0 = 0; l0 < nz;) {
    int nl = 1;
    if (!t.amap.empty())
      while (l0 + nl < nz && nl < 4 && t.amap[l0 + nl] == t.amap[l0]) ++nl;
    std::vector<int> pts;
    for (int l = l0; l < l0 + nl; ++l)
      for (int j = 0; j < t.mjj[l]; ++j) {
        int p = t.jjp[size_t(l) * mx + j];
        if (mark[size_t(p)] < 0) {
          mark[size_t(p)] = 0;
          pts.push_back(p);
        }
      }
    std::sort(pts.begin(), pts.end());
    for (size_t i = 0; i < pts.size(); ++i) mark[size_t(pts[i])] = int(i);
    const size_t base = h.grp_coef.size();
    const int pbase = int(h.grp_pts.size());
    h.grp_coef.resize(base + pts.size() * 4, 0.0);
    for (int l = l0; l < l0 + nl; ++l) {
      h.lma_grp[size_t(l)] = int(h.grp_lma0.size());
      for (int j = 0; j < t.mjj[l]; ++j) {
        int p = t.jjp[size_t(l) * mx + j];
        h.grp_coef[base + size_t(mark[size_t(p)]) * 4 + (l - l0)] = t.uvk[size_t(l) * mx + j];
      }
    }
    for (int p : pts) mark[size_t(p)] = -1;
    h.grp_pts.insert(h.grp_pts.end(), pts.begin(), pts.end());
    const int g = int(h.grp_lma0.size());
    for (int b = 0; b < int(pts.size()); b += kSegLen) {
      h.seg_beg.push_back(pbase + b);
      h.seg_end.push_back(pbase + std::min(int(pts.size()), b + kSegLen));
      h.seg_grp.push_back(g);
    }
    h.grp_seg_off.push_back(int(h.seg_beg.size()));
    h.grp_lma0.push_back(l0);
    h.grp_nl.push_back(nl);
    l0 += nl;
  }
  h.scale.resize(size_t(nz));
  for (int l = 0; l < nz; ++l) h.scale[size_t(l)] = t.iuv[size_t(l)] * ctx->dv;

  // per-point entries in increasing lma order (CSR), then sliced ELL
  std::vector<int> pt_off(size_t(ctx->N) + 1, 0);
  for (int l = 0; l < nz; ++l)
    for (int j = 0; j < t.mjj[l]; ++j) ++pt_off[size_t(t.jjp[size_t(l) * mx + j]) + 1];
  for (int64_t i = 0; i < ctx->N; ++i) pt_off[size_t(i) + 1] += pt_off[size_t(i)];
  const int64_t E = pt_off[size_t(ctx->N)];
  std::vector<int> pt_lma(static_cast<size_t>(E));
  std::vector<double> pt_uv(static_cast<size_t>(E));
  std::vector<int> fill(pt_off.begin(), pt_off.end() - 1);
  for (int l = 0; l < nz; ++l)
    for (int j = 0; j < t.mjj[l]; ++j) {
      int p = t.jjp[size_t(l) * mx + j];
      int e = fill[size_t(p)]++;
      pt_lma[size_t(e)] = l;
      pt_uv[size_t(e)] = t.uvk[size_t(l) * mx + j];
    }
  const int64_t nslice = (ctx->N + 63) / 64;
  h.sl_off.resize(size_t(nslice));
  h.sl_len.resize(size_t(nslice));
  int64_t slots = 0;
  for (int64_t c = 0; c < nslice; ++c) {
    int K = 0;
    for (int64_t i = c * 64; i < std::min(ctx->N, c * 64 + 64); ++i)
      K = std::max(K, pt_off[size_t(i) + 1] - pt_off[size_t(i)]);
    h.sl_off[size_t(c)] = int(slots);
    h.sl_len[size_t(c)] = K;
    slots += int64_t(K) * 64;
  }
  if (slots > (int64_t(1) << 31) - 1) return ctx->fail(RSG_ERR_UNSUPPORTED, "projector table above 2^31 slots");
  h.sl_lma.assign(size_t(slots), 0);
  h.sl_uv.assign(size_t(slots), 0.0);
  for (int64_t i = 0; i < ctx->N; ++i) {
    const int64_t c = i / 64, lane = i % 64;
    for (int k = 0; k < pt_off[size_t(i) + 1] - pt_off[size_t(i)]; ++k) {
      const size_t e = size_t(h.sl_off[size_t(c)] + int64_t(k) * 64 + lane);
      h.sl_lma[e] = pt_lma[size_t(pt_off[size_t(i)] + k)];
      h.sl_uv[e] = pt_uv[size_t(pt_off[size_t(i)] + k)];
    }
  }
  int rc = upload_layouts<double>(ctx, ctx->nl64, h);
  if (rc) return rc;
  rc = upload_layouts<float>(ctx, ctx->nl32, h);
  if (rc) return rc;
  ctx->have_nl = nz > 0;
  return RSG_OK;
}

namespace {

// Row-tile blocks of the projection: rt_ty grid rows by rt_zr planes (all
// of x).  A piece is the run of one atom's sphere points on one plane that
// falls in the block's rows; an atom's points are in (z, y, x) order, so the
// run is contiguous.  Each block numbers the atoms it meets (local slots,
// in atom order); the global slots of one atom are contiguous, in block
// order, so that the reduction sums one range per atom.
// f(begin, end) over [0, n) split among host threads (at most 16, the
// host share of one GPU); each index belongs to exactly one call.
template <typename F> void parallel_range(int n, F &&f) {
  const int nt = int(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  if (n < 2 * nt || nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) th.emplace_back([&, t]() { f(int(int64_t(n) * t / nt), int(int64_t(n) * (t + 1) / nt)); });
  for (auto &x : th) x.join();
}

constexpr int kRowTilePoints = 288;  // rows per block ~ this many points / n1
constexpr int kRowTileSlots = 32;    // at most this many atoms per block (LDS sums)

}  // namespace

void build_row_tiles(const int *ng, HostAnalytic &out) {
  const int ty = std::max(1, std::min(8, kRowTilePoints / ng[0]));
  const int ny = (ng[1] + ty - 1) / ty;
  struct Piece {
    int blk, z, g, beg, end;
  };
  auto collect = [&](int zr, std::vector<Piece> &pieces) {
    pieces.clear();
    for (int g = 0; g < out.ngroup; ++g) {
      const int s0 = out.grp_seg_off[size_t(g)], s1 = out.grp_seg_off[size_t(g) + 1];
      const int b = out.seg_beg[size_t(s0)], e = out.seg_end[size_t(s1) - 1];
      for (int i = b; i < e;) {
        const uint32_t w = out.pts[size_t(i)];
        const int z = int(w >> 20 & 1023u), t = int(w >> 10 & 1023u) / ty;
        int j = i + 1;
        while (j < e && int(out.pts[size_t(j)] >> 20 & 1023u) == z && int(out.pts[size_t(j)] >> 10 & 1023u) / ty == t)
          ++j;
        pieces.push_back({t + ny * (z / zr), z, g, i, j});
        i = j;
      }
    }
    std::stable_sort(pieces.begin(), pieces.end(), [](const Piece &a, const Piece &b) {
      return a.blk != b.blk ? a.blk < b.blk : (a.z != b.z ? a.z < b.z : a.g < b.g);
    });
  };
  std::vector<Piece> pieces;
  int zr = 8, nz = 0, maxslot = 0;
  for (;;) {
    nz = (ng[2] + zr - 1) / zr;
    collect(zr, pieces);
    // atoms per block
    maxslot = 0;
    for (size_t i = 0; i < pieces.size();) {
      size_t j = i;
      std::vector<int> gs;
      while (j < pieces.size() && pieces[j].blk == pieces[i].blk) gs.push_back(pieces[j++].g);
      std::sort(gs.begin(), gs.end());
      maxslot = std::max(maxslot, int(std::unique(gs.begin(), gs.end()) - gs.begin()));
      i = j;
    }
    if (maxslot <= kRowTileSlots || zr == 1) break;
    zr /= 2;
  }
  const int nrt = ny * nz;
  out.rt_ty = ty;
  out.rt_zr = zr;
  out.rt_ny = ny;
  out.rt_nz = nz;
  out.rt_nslot_max = maxslot;
  out.rt_piece.clear();
  out.rt_piece_off.assign(size_t(nrt) + 1, 0);
  out.rt_slot_off.assign(size_t(nrt) + 1, 0);
  std::vector<std::vector<int>> blk_groups(static_cast<size_t>(nrt));
  std::vector<std::vector<int>> grp_blocks(static_cast<size_t>(out.ngroup));
  size_t i = 0;
  for (int blk = 0; blk < nrt; ++blk) {
    size_t j = i;
    while (j < pieces.size() && pieces[j].blk == blk) ++j;
    std::vector<int> &gs = blk_groups[size_t(blk)];
    for (size_t k = i; k < j; ++k) gs.push_back(pieces[k].g);
    std::sort(gs.begin(), gs.end());
    gs.erase(std::unique(gs.begin(), gs.end()), gs.end());
    for (size_t k = i; k < j; ++k) {
      const int slot = int(std::lower_bound(gs.begin(), gs.end(), pieces[k].g) - gs.begin());
      out.rt_piece.insert(out.rt_piece.end(), {pieces[k].beg, pieces[k].end, pieces[k].z * 256 + slot, pieces[k].g});
    }
    for (int g : gs) grp_blocks[size_t(g)].push_back(blk);
    out.rt_piece_off[size_t(blk) + 1] = int(out.rt_piece.size() / 4);
    out.rt_slot_off[size_t(blk) + 1] = out.rt_slot_off[size_t(blk)] + int(gs.size());
    i = j;
  }
  // global slot ids: one contiguous range per atom, its blocks in order
  out.grp_slot_off.assign(size_t(out.ngroup) + 1, 0);
  for (int g = 0; g < out.ngroup; ++g)
    out.grp_slot_off[size_t(g) + 1] = out.grp_slot_off[size_t(g)] + int(grp_blocks[size_t(g)].size());
  out.rt_slot_gid.assign(size_t(out.rt_slot_off[size_t(nrt)]), 0);
  std::vector<int> next(out.grp_slot_off.begin(), out.grp_slot_off.end() - 1);
  for (int blk = 0; blk < nrt; ++blk) {
    const std::vector<int> &gs = blk_groups[size_t(blk)];
    for (size_t k = 0; k < gs.size(); ++k) out.rt_slot_gid[size_t(out.rt_slot_off[size_t(blk)]) + k] = next[size_t(gs[k])]++;
  }
}

// Per-atom parameters of the analytic description of GTH sets made only of
// (l = 0, n = 1) and (l = 1, n = 1) projectors, at most one of each per
// species (ps_nloc_gth_module.f90:19-180 for these forms).  Returns
// RSG_ERR_UNSUPPORTED (not an error for the caller) when the set, the cell
// or the grid does not fit that form; the table path then stays in use.
int analytic_groups(const rsg_ctx *ctx, int natom, const double *frac, const int *kind, int nkind, const int *norb,
                    const int *lo, const int *no, const double *rps, const double *hnl, AnalyticGroups &out) {
  out = AnalyticGroups();
  const double *aa = ctx->aa;
  if (aa[1] != 0.0 || aa[2] != 0.0 || aa[3] != 0.0 || aa[5] != 0.0 || aa[6] != 0.0 || aa[7] != 0.0)
    return RSG_ERR_UNSUPPORTED;
  const int *ng = ctx->n;
  if (ng[0] > 1023 || ng[1] > 1023 || ng[2] > 1023 || natom > 16382) return RSG_ERR_UNSUPPORTED;
  // per species: orbital index of s and p (or -1), cutoffs
  std::vector<int> qs(size_t(nkind), -1), qp(size_t(nkind), -1);
  std::vector<double> rcs(size_t(nkind), 0.0), rcp(size_t(nkind), 0.0);
  double rmax = 0.0;
  for (int ik = 0; ik < nkind; ++ik) {
    for (int io = 0; io < norb[ik]; ++io) {
      const int q = ik * 4 + io;
      if (no[q] != 1 || lo[q] > 1) return RSG_ERR_UNSUPPORTED;
      int &slot = lo[q] == 0 ? qs[size_t(ik)] : qp[size_t(ik)];
      if (slot >= 0) return RSG_ERR_UNSUPPORTED;
      slot = q;
      const double rc = gth_cutoff(lo[q], 1, rps[q]);
      (lo[q] == 0 ? rcs : rcp)[size_t(ik)] = rc;
      rmax = std::max(rmax, rc);
    }
  }
  for (int t = 0; t < 3; ++t)
    if (!(aa[4 * t] > 2.0 * rmax + 2.0 * ctx->hgrid[t])) return RSG_ERR_UNSUPPORTED;
  for (int a = 0; a < natom; ++a)
    if (kind[a] < 0 || kind[a] >= nkind) return RSG_ERR_ARG;
  out.S = ng[0] + ng[1] + ng[2];
  const double pi4 = 4.0 * kPi;
  // groups (atoms with projectors) in atom order
  for (int a = 0; a < natom; ++a) {
    const int ik = kind[a];
    const int q0 = qs[size_t(ik)], q1 = qp[size_t(ik)];
    if (q0 < 0 && q1 < 0) continue;
    AnalyticGroup G;
    const double *fa = frac + 3 * a;
    // same arithmetic as build_gth_tables, so sphere membership agrees
    G.R[0] = aa[0] * fa[0] + aa[3] * fa[1] + aa[6] * fa[2];
    G.R[1] = aa[1] * fa[0] + aa[4] * fa[1] + aa[7] * fa[2];
    G.R[2] = aa[2] * fa[0] + aa[5] * fa[1] + aa[8] * fa[2];
    for (int t = 0; t < 3; ++t) G.ic[t] = nint_away(fa[t] * ng[t]);
    G.has_s = q0 >= 0;
    G.has_p = q1 >= 0;
    G.rs = q0 >= 0 ? rcs[size_t(ik)] : 0.0;
    G.rp = q1 >= 0 ? rcp[size_t(ik)] : 0.0;
    const double rc = std::max(G.rs, G.rp);
    for (int t = 0; t < 3; ++t) G.mm[t] = int(std::ceil(rc / ctx->hgrid[t])) + 2;
    G.c2s = q0 >= 0 ? 0.5 / (rps[q0] * rps[q0]) : 0.0;
    G.c2p = q1 >= 0 ? 0.5 / (rps[q1] * rps[q1]) : 0.0;
    G.amp_s = q0 >= 0 ? std::sqrt(std::fabs(hnl[q0])) * std::sqrt(2.0) /
                            (std::pow(rps[q0], 1.5) * std::sqrt(gamma_half(0, 1))) / std::sqrt(pi4)
                      : 0.0;
    G.amp_p = q1 >= 0 ? std::sqrt(std::fabs(hnl[q1])) * std::sqrt(2.0) /
                            (std::pow(rps[q1], 2.5) * std::sqrt(gamma_half(1, 1))) * std::sqrt(3.0 / pi4)
                      : 0.0;
    G.sg_s = q0 >= 0 && hnl[q0] < 0.0 ? -1.0 : 1.0;
    G.sg_p = q1 >= 0 && hnl[q1] < 0.0 ? -1.0 : 1.0;
    out.g.push_back(G);
  }
  return RSG_OK;
}

// The analytic description built on the host from analytic_groups: factor
// tables, sphere points, hit words, segments, tile lists and row tiles.
int build_gth_analytic(const rsg_ctx *ctx, int natom, const double *frac, const int *kind, int nkind,
                       const int *norb, const int *lo, const int *no, const double *rps, const double *hnl,
                       HostAnalytic &out) {
  out = HostAnalytic();
  AnalyticGroups P;
  const int prc = analytic_groups(ctx, natom, frac, kind, nkind, norb, lo, no, rps, hnl, P);
  if (prc) return prc;
  const double *aa = ctx->aa;
  const int *ng = ctx->n;
  const int S = P.S;
  const int axoff[3] = {0, ng[0], ng[0] + ng[1]};
  const double c[3] = {1.0 / ng[0], 1.0 / ng[1], 1.0 / ng[2]};
  const int64_t n12 = int64_t(ng[0]) * ng[1];
  const int ngrp = int(P.g.size());
  out.S = S;
  out.amp.assign(size_t(ngrp) * 4, 0.0);
  out.scale.assign(size_t(ngrp) * 4, 0.0);
  out.es.assign(size_t(ngrp) * S, 0.0);
  out.ep.assign(size_t(ngrp) * S, 0.0);
  out.dd.assign(size_t(ngrp) * S, 0.0);
  // Per group, independent of the others: amplitudes, the 1-D factor tables
  // and the sphere points in grid order.  Host threads take groups in turn;
  // everything shared (hit words, point list, segments) is merged afterwards
  // in group order, so the result does not depend on the thread count.
  std::vector<std::vector<std::pair<int64_t, uint32_t>>> gpts(static_cast<size_t>(ngrp));
  auto build_group = [&](int g) {
    const AnalyticGroup &G = P.g[size_t(g)];
    const double *R = G.R;
    const int *ic = G.ic;
    const double rs = G.rs, rp = G.rp, c2s = G.c2s, c2p = G.c2p;
    const size_t g4 = size_t(g) * 4;
    out.amp[g4] = G.amp_s;
    out.amp[g4 + 1] = out.amp[g4 + 2] = out.amp[g4 + 3] = G.amp_p;
    out.scale[g4] = G.has_s ? G.sg_s * ctx->dv : 0.0;
    out.scale[g4 + 1] = out.scale[g4 + 2] = out.scale[g4 + 3] = G.has_p ? G.sg_p * ctx->dv : 0.0;
    // 1-D tables over each axis at the image nearest the atom
    const size_t base = size_t(g) * S;
    for (int t = 0; t < 3; ++t)
      for (int i = 0; i < ng[t]; ++i) {
        int id = i;
        while (id - ic[t] > ng[t] / 2) id -= ng[t];
        while (ic[t] - id > ng[t] / 2) id += ng[t];
        const double d = aa[4 * t] * (id * c[t]) - R[t];
        out.dd[base + size_t(axoff[t] + i)] = d;
        out.es[base + size_t(axoff[t] + i)] = std::exp(-d * d * c2s);
        out.ep[base + size_t(axoff[t] + i)] = std::exp(-d * d * c2p);
      }
    // sphere points, in grid order
    const int *mm = G.mm;
    std::vector<std::pair<int64_t, uint32_t>> &pts = gpts[size_t(g)];
    for (int o3 = -mm[2]; o3 <= mm[2]; ++o3)
      for (int o2 = -mm[1]; o2 <= mm[1]; ++o2)
        for (int o1 = -mm[0]; o1 <= mm[0]; ++o1) {
          const int id[3] = {ic[0] + o1, ic[1] + o2, ic[2] + o3};
          const double d1 = id[0] * c[0], d2 = id[1] * c[1], d3 = id[2] * c[2];
          const double x = aa[0] * d1 + aa[3] * d2 + aa[6] * d3 - R[0];
          const double y = aa[1] * d1 + aa[4] * d2 + aa[7] * d3 - R[1];
          const double z = aa[2] * d1 + aa[5] * d2 + aa[8] * d3 - R[2];
          const double r2 = x * x + y * y + z * z;
          const bool in_s = G.has_s && r2 <= rs * rs + 1e-10;
          const bool in_p = G.has_p && r2 <= rp * rp + 1e-10;
          if (!in_s && !in_p) continue;
          int w[3];
          for (int t = 0; t < 3; ++t) w[t] = ((id[t] % ng[t]) + ng[t]) % ng[t];
          const int64_t lin = w[0] + int64_t(w[1]) * ng[0] + int64_t(w[2]) * n12;
          const uint32_t word = uint32_t(w[0]) | (uint32_t(w[1]) << 10) | (uint32_t(w[2]) << 20) |
                                (uint32_t(in_s) << 30) | (uint32_t(in_p) << 31);
          pts.emplace_back(lin, word);
        }
    std::sort(pts.begin(), pts.end());
  };
  parallel_range(ngrp, [&](int g0, int g1) {
    for (int g = g0; g < g1; ++g) build_group(g);
  });
  // merge in group order: hit words (at most four spheres per point), the
  // point list and its plane segments
  std::vector<unsigned char> nhit(size_t(ctx->N), 0);
  out.hits.assign(size_t(ctx->N), 0);
  int g = 0;
  for (; g < ngrp; ++g) {
    const std::vector<std::pair<int64_t, uint32_t>> &pts = gpts[size_t(g)];
    for (const auto &pw : pts) {
      unsigned char &k = nhit[size_t(pw.first)];
      if (k >= 4) return RSG_ERR_UNSUPPORTED;  // more than four spheres at one point
      const uint64_t e = uint64_t(g + 1) | (uint64_t((pw.second >> 30) & 1u) << 14) | (uint64_t(pw.second >> 31) << 15);
      out.hits[size_t(pw.first)] |= e << (16 * k);
      ++k;
    }
    const int pbase = int(out.pts.size());
    for (auto &pw : pts) out.pts.push_back(pw.second);
    // one segment per plane run (grid order is plane-major), capped in length
    for (int b = 0; b < int(pts.size());) {
      const int64_t plane = pts[size_t(b)].first / n12;
      int e = b + 1;
      while (e < int(pts.size()) && e - b < kPlaneSegMax && pts[size_t(e)].first / n12 == plane) ++e;
      out.seg_beg.push_back(pbase + b);
      out.seg_end.push_back(pbase + e);
      out.seg_grp.push_back(g);
      b = e;
    }
    out.grp_seg_off.push_back(int(out.seg_beg.size()));
    std::vector<std::pair<int64_t, uint32_t>>().swap(gpts[size_t(g)]);
  }
  out.ngroup = g;
  if (g == 0) return RSG_ERR_UNSUPPORTED;
  // launch order: by plane, then atom
  out.seg_order.resize(out.seg_beg.size());
  for (size_t i = 0; i < out.seg_order.size(); ++i) out.seg_order[i] = int(i);
  std::stable_sort(out.seg_order.begin(), out.seg_order.end(), [&](int a, int b) {
    return (out.pts[size_t(out.seg_beg[size_t(a)])] >> 20 & 1023u) < (out.pts[size_t(out.seg_beg[size_t(b)])] >> 20 & 1023u);
  });
  // atoms meeting each 16-point tile: tile ranges in host threads, joined in
  // tile order
  const int64_t ntile = (ctx->N + 15) / 16;
  out.tile_off.assign(size_t(ntile) + 1, 0);
  out.tile_g.clear();
  const int nchunk = int(std::min<int64_t>(ntile, 256));
  std::vector<std::vector<int>> chunk_g(static_cast<size_t>(nchunk));
  std::vector<int> tcount(static_cast<size_t>(ntile), 0);
  parallel_range(nchunk, [&](int c0, int c1) {
    for (int c = c0; c < c1; ++c) {
      const int64_t t0 = ntile * c / nchunk, t1 = ntile * (c + 1) / nchunk;
      for (int64_t t = t0; t < t1; ++t) {
        int list[64];
        int nl = 0;
        for (int64_t j = t * 16; j < std::min<int64_t>(ctx->N, t * 16 + 16); ++j) {
          uint64_t hw = out.hits[size_t(j)];
          while (hw) {
            const int gg = int(hw & 0x3fffu) - 1;
            hw >>= 16;
            bool seen = false;
            for (int i = 0; i < nl; ++i) seen |= list[i] == gg;
            if (!seen) list[nl++] = gg;
          }
        }
        std::sort(list, list + nl);
        chunk_g[size_t(c)].insert(chunk_g[size_t(c)].end(), list, list + nl);
        tcount[size_t(t)] = nl;
      }
    }
  });
  for (int64_t t = 0; t < ntile; ++t) out.tile_off[size_t(t) + 1] = out.tile_off[size_t(t)] + tcount[size_t(t)];
  out.tile_g.reserve(size_t(out.tile_off[size_t(ntile)]));
  for (auto &v : chunk_g) out.tile_g.insert(out.tile_g.end(), v.begin(), v.end());
  build_row_tiles(ng, out);
  return RSG_OK;
}

namespace {

template <typename T> int upload_analytic_t(rsg_ctx *ctx, AnalyticDev<T> &d, const HostAnalytic &h) {
  // stream-ordered copies (a plain hipMemcpy waits for every stream of the
  // device), synchronised before the host vectors go out of scope
  auto up = [&](DevBuf &b, const void *src, size_t bytes) -> int {
    if (b.ensure(bytes) != hipSuccess) return RSG_ERR_HIP;
    if (bytes && hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
      return RSG_ERR_HIP;
    return RSG_OK;
  };
  auto conv = [](const std::vector<double> &v) { return std::vector<T>(v.begin(), v.end()); };
  const std::vector<T> es = conv(h.es), ep = conv(h.ep), dd = conv(h.dd), amp = conv(h.amp), sc = conv(h.scale);
  d.ngroup = h.ngroup;
  d.nseg = int(h.seg_beg.size());
  d.S = h.S;
  int rc = 0;
  rc |= up(d.es, es.data(), es.size() * sizeof(T));
  rc |= up(d.ep, ep.data(), ep.size() * sizeof(T));
  rc |= up(d.dd, dd.data(), dd.size() * sizeof(T));
  rc |= up(d.amp, amp.data(), amp.size() * sizeof(T));
  rc |= up(d.scale, sc.data(), sc.size() * sizeof(T));
  rc |= up(d.pts, h.pts.data(), h.pts.size() * sizeof(uint32_t));
  rc |= up(d.seg_beg, h.seg_beg.data(), h.seg_beg.size() * sizeof(int));
  rc |= up(d.seg_end, h.seg_end.data(), h.seg_end.size() * sizeof(int));
  rc |= up(d.seg_grp, h.seg_grp.data(), h.seg_grp.size() * sizeof(int));
  rc |= up(d.grp_seg_off, h.grp_seg_off.data(), h.grp_seg_off.size() * sizeof(int));
  rc |= up(d.seg_order, h.seg_order.data(), h.seg_order.size() * sizeof(int));
  rc |= up(d.hits, h.hits.data(), h.hits.size() * sizeof(uint64_t));
  rc |= up(d.tile_off, h.tile_off.data(), h.tile_off.size() * sizeof(int));
  rc |= up(d.tile_g, h.tile_g.data(), h.tile_g.size() * sizeof(int));
  d.ntile = int(h.tile_off.size()) - 1;
  rc |= up(d.rt_piece, h.rt_piece.data(), h.rt_piece.size() * sizeof(int));
  rc |= up(d.rt_piece_off, h.rt_piece_off.data(), h.rt_piece_off.size() * sizeof(int));
  rc |= up(d.rt_slot_off, h.rt_slot_off.data(), h.rt_slot_off.size() * sizeof(int));
  rc |= up(d.rt_slot_gid, h.rt_slot_gid.data(), h.rt_slot_gid.size() * sizeof(int));
  rc |= up(d.grp_slot_off, h.grp_slot_off.data(), h.grp_slot_off.size() * sizeof(int));
  d.rt_ty = h.rt_ty;
  d.rt_zr = h.rt_zr;
  d.rt_ny = h.rt_ny;
  d.rt_nz = h.rt_nz;
  d.rt_nslot_max = h.rt_nslot_max;
  d.rt_nslot = h.grp_slot_off.empty() ? 0 : h.grp_slot_off.back();
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = RSG_ERR_HIP;
  return rc ? ctx->fail(RSG_ERR_HIP, "upload of analytic projector data failed") : RSG_OK;
}

}  // namespace

int upload_analytic(rsg_ctx *ctx, const HostAnalytic &h) {
  int rc = upload_analytic_t<double>(ctx, ctx->ana64, h);
  if (!rc) rc = upload_analytic_t<float>(ctx, ctx->ana32, h);
  ctx->have_ana = rc == RSG_OK;
  ctx->ana_rows_pending = false;
  return rc;
}

int ensure_row_tiles(rsg_ctx *ctx) {
  if (!ctx->ana_rows_pending || !ctx->have_ana) return RSG_OK;
  const AnalyticDev<double> &d = ctx->ana64;
  HostAnalytic h;
  h.ngroup = d.ngroup;
  h.grp_seg_off.assign(size_t(d.ngroup) + 1, 0);
  h.seg_beg.assign(size_t(d.nseg), 0);
  h.seg_end.assign(size_t(d.nseg), 0);
  hipStream_t st = ctx->stream;
  RSG_HIP(hipMemcpyAsync(h.grp_seg_off.data(), d.grp_seg_off.p, h.grp_seg_off.size() * 4, hipMemcpyDeviceToHost, st));
  RSG_HIP(hipMemcpyAsync(h.seg_beg.data(), d.seg_beg.p, h.seg_beg.size() * 4, hipMemcpyDeviceToHost, st));
  RSG_HIP(hipMemcpyAsync(h.seg_end.data(), d.seg_end.p, h.seg_end.size() * 4, hipMemcpyDeviceToHost, st));
  RSG_HIP(hipStreamSynchronize(st));
  const size_t E = d.nseg ? size_t(h.seg_end.back()) : 0;
  h.pts.assign(E, 0);
  if (E) RSG_HIP(hipMemcpyAsync(h.pts.data(), d.pts.p, E * 4, hipMemcpyDeviceToHost, st));
  RSG_HIP(hipStreamSynchronize(st));
  build_row_tiles(ctx->n, h);
  auto up_rows = [&](auto &dv) -> int {
    auto up = [&](DevBuf &b, const std::vector<int> &v) -> int {
      if (b.ensure(std::max<size_t>(v.size() * 4, 4)) != hipSuccess) return RSG_ERR_HIP;
      if (!v.empty() && hipMemcpyAsync(b.p, v.data(), v.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess)
        return RSG_ERR_HIP;
      return RSG_OK;
    };
    int rc = up(dv.rt_piece, h.rt_piece) | up(dv.rt_piece_off, h.rt_piece_off) | up(dv.rt_slot_off, h.rt_slot_off) |
             up(dv.rt_slot_gid, h.rt_slot_gid) | up(dv.grp_slot_off, h.grp_slot_off);
    dv.rt_ty = h.rt_ty;
    dv.rt_zr = h.rt_zr;
    dv.rt_ny = h.rt_ny;
    dv.rt_nz = h.rt_nz;
    dv.rt_nslot_max = h.rt_nslot_max;
    dv.rt_nslot = h.grp_slot_off.empty() ? 0 : h.grp_slot_off.back();
    return rc;
  };
  int rc = up_rows(ctx->ana64);
  rc |= up_rows(ctx->ana32);
  RSG_HIP(hipStreamSynchronize(st));
  if (rc) return ctx->fail(RSG_ERR_HIP, "upload of the projection row tiles failed");
  ctx->ana_rows_pending = false;
  return RSG_OK;
}

}  // namespace rsg
