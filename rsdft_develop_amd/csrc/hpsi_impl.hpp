// hpsi.hip -- Hamiltonian application on gfx950:
//   hpsi = (T + Vloc) psi + V_NL psi  over a block of bands.
//
// Bands are processed in chunks (sized by the autotuner; one chunk of psi
// and H psi can stay in the 256 MB Infinity Cache between launches).
// Nonlocal part, analytic projectors (the default for l <= 1, n = 1 GTH):
//  1. k_proj_m: partial projections sum_j beta_k(j) psi(j,b) per segment
//     (one atom's sphere points on one plane) on the matrix cores
//     (d_ps_nloc2_op_module.f90:182-191 with the projector values formed on
//     the fly instead of read from uVk tables);
//  2. k_proj_reduce: C(g,k,b) = iuV dV * sum over the atom's segments;
//  3. k_apply_m: H psi(j,b) (=|+=) sum_{g,k} beta_k(j) C(g,k,b) per 16-point
//     tile on the matrix cores (d_ps_nloc2_op_module.f90:274-281).
// Nonlocal part, projector tables (any l; nonlocal mode 1):
//  1. k_nl_gather: per (atom-group segment, band tile) partial projections
//     from the uVk/JJP tables, reading psi at each sphere point once;
//  2. k_nl_reduce: uVunk(lma,b) = iuV dV * sum over the group's segments;
//  3. k_nl_scatter: H psi(j,b) = sum_lma uV(j,lma) uVunk(lma,b) from a
//     per-point sliced-ELL table, no atomics.
// Local part:
//  4. k_stencil: H psi (+)= (c + V) psi + 13-point-per-axis FD Laplacian
//     (d_kinetic_sol_module.f90:244-292).  A workgroup owns ty full x-rows
//     of a z-run; each plane is staged in double-buffered LDS with its
//     +-MD halo rows (periodic in x inside the row, in y by row wrap: the
//     bcset_3 halo of bc/bcset_3_module.f90:15-151), the z-neighbours stay
//     in a per-thread register queue, and each thread computes ry rows so
//     neighbouring rows share their y-window reads.
#include <algorithm>
#include <type_traits>
#include <utility>

#include "context.hpp"

namespace rsg {
namespace {

constexpr int kMaxMd = 8;

template <typename T> struct StencilCoef {
  T c[3][kMaxMd];  // coef_lap(axis, m), m = 1..md
  T c0;            // coef_lap0 (+ const_k2 = 0 at Gamma)
};

__host__ __device__ inline int wrapi(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// (b + u) mod n for b in [0, n) and u >= 0: selects for the usual u < 2n
__host__ __device__ inline int add_wrap(int b, int u, int n) {
  const int t = b + u;
  return t < n ? t : (t < 2 * n ? t - n : t % n);
}

// Software pipeline of the z-stream.  Plane zz of the thread's own columns
// lives in register slot (zz - z0 + MD) % R of a ring of R = 2*MD + 1 + PF
// planes (PF planes fetched ahead, no register moves).  The halo rows, Vloc
// and the accumulated H psi of plane z+HS-1 are fetched into one of HS
// rotating register sets while plane z computes.  R is a multiple of HS so
// every slot index is a compile-time constant in the unrolled loop.
//
// The loop body is branch-free: every load goes to a valid (wrapped)
// address whether or not its value is used, padding lanes store into a
// dead buffer, and unused halo lanes write a dead LDS slot.  A conditional
// load or store makes the wait-count pass drain all outstanding memory
// operations at the join, which serialises the whole pipeline.  The LDS
// exchange synchronises with s_waitcnt lgkmcnt(0) + s_barrier only;
// __syncthreads() would also wait for every in-flight global load.
constexpr int kHS = 4;
__host__ __device__ constexpr int ring_pf(int md) { return (md % 2 == 0) ? 3 : 5; }

__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Workgroups of a launch are dealt to the 8 XCDs in turn (block b runs on
// XCD b % 8), and each XCD has its own L2.  xcd_block maps the hardware
// block index to a logical one such that every XCD works through one
// contiguous range of logical blocks: neighbouring tiles, which share halo
// rows or sphere rows, then meet in the same L2.  A bijection on [0, nblk).
__device__ inline unsigned xcd_block(unsigned b, unsigned nblk) {
  const unsigned q = nblk / 8, r = nblk % 8, x = b % 8;
  return x * q + min(x, r) + b / 8;
}

template <typename T, int MD, int RY, int HM, bool ACC>
__global__ __launch_bounds__(512) void k_stencil(const T *psi, T *hpsi, int64_t ld, const T *vloc,
                                                 StencilCoef<T> cf, int n1, int n2, int n3,
                                                 int ty, int zlen, T *dead) {
  constexpr int PF = ring_pf(MD);
  constexpr int R = 2 * MD + 1 + PF;
  static_assert(R % kHS == 0 && R % 2 == 0, "ring length must be a multiple of the rotations");
  extern __shared__ unsigned char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int rows = ty + 2 * MD;
  const int plane = rows * n1;  // one LDS buffer
  const int G = ty / RY;
  const int tid = threadIdx.x, nthr = blockDim.x;
  // idle lanes write a dead slot of their own past both buffers: one slot
  // shared by a whole wave would serialise its 64 writes
  const int dslot = 2 * plane + (tid & 63);
  const bool active = tid < n1 * G;
  const int x = active ? tid % n1 : 0;
  const int g = active ? tid / n1 : 0;
  // logical block: y-tile fastest, then z-run, then band
  const int ntile_y = (n2 + ty - 1) / ty, nz = (n3 + zlen - 1) / zlen;
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int by = int(lb % unsigned(ntile_y)), bz = int((lb / unsigned(ntile_y)) % unsigned(nz));
  const int band = int(lb / unsigned(ntile_y * nz));
  const int y0 = by * ty;
  const int z0 = bz * zlen, z1 = min(z0 + zlen, n3);
  // offsets inside one band stay below N < 2^31: 32-bit arithmetic
  const unsigned n12 = unsigned(n1) * unsigned(n2);
  // No __restrict__ on the streams: with it the loads count as invariant
  // and the register allocator re-issues them next to their uses instead of
  // keeping the prefetched values, which turns every plane into a full
  // memory round trip.  Without it no load moves across the H psi stores.
  const T *w = psi + int64_t(band) * ld;
  T *o = hpsi + int64_t(band) * ld;

  unsigned colofs[RY];
  bool own[RY];
  int lown[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int y = y0 + g * RY + r;
    own[r] = active && y < n2;
    colofs[r] = unsigned(x + (y % n2) * n1);  // rows past n2 in the last tile hold wrapped data
    lown[r] = active ? (MD + g * RY + r) * n1 + x : dslot;
  }
  int xm[MD], xp[MD];
#pragma unroll
  for (int m = 0; m < MD; ++m) {
    xm[m] = wrapi(x - m - 1, n1);
    xp[m] = wrapi(x + m + 1, n1);
  }
  // halo elements this thread stages: LDS offset and in-plane offset
  const int nhalo = 2 * MD * n1;
  int hl[HM];
  unsigned hg[HM];
#pragma unroll
  for (int k = 0; k < HM; ++k) {
    const int idx = tid + k * nthr;
    const int i = idx < nhalo ? idx : 0;
    const int hr = i / n1, hx = i - hr * n1;
    const int lrow = hr < MD ? hr : hr + ty;
    hl[k] = idx < nhalo ? lrow * n1 + hx : dslot;
    hg[k] = unsigned(hx + wrapi(y0 - MD + lrow, n2) * n1);
  }
  T q[RY][R];
#pragma unroll
  for (int r = 0; r < RY; ++r)
#pragma unroll
    for (int k = 0; k < R; ++k) q[r][k] = w[colofs[r] + unsigned(wrapi(z0 - MD + k, n3)) * n12];
  // rotating sets: halo h[s][k], Vloc v[s][r], previous H psi a[s][r]
  T h[kHS][HM], v[kHS][RY], a[kHS][RY];
#pragma unroll
  for (int sidx = 0; sidx < kHS - 1; ++sidx) {
    const unsigned zo = unsigned(wrapi(z0 + sidx, n3)) * n12;
#pragma unroll
    for (int k = 0; k < HM; ++k) h[sidx][k] = w[hg[k] + zo];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      v[sidx][r] = vloc[colofs[r] + zo];
      if (ACC) a[sidx][r] = __builtin_nontemporal_load(o + colofs[r] + zo);
    }
  }
  T *dst_dead = dead + (tid & 63);

  for (int zb = z0; zb < z1; zb += R) {
    // planes of the halo / Vloc / H psi prefetch and of the own-column
    // prefetch: one modulo per ring turn, a compare-select per step
    const int zh0 = wrapi(zb + kHS - 1, n3), zq0 = wrapi(zb + MD + PF, n3);
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int z = zb + u;
      if (z >= z1) break;
      T *buf = lds + (u & 1) * plane;
      T *bufw = lds + (u & 1) * plane;
      const int sc = u % kHS;              // set holding plane z
      const int sn = (u + kHS - 1) % kHS;  // set to fill with plane z+kHS-1
#pragma unroll
      for (int r = 0; r < RY; ++r) (lown[r] == dslot ? lds : bufw)[lown[r]] = q[r][(u + MD) % R];
#pragma unroll
      for (int k = 0; k < HM; ++k) (hl[k] == dslot ? lds : bufw)[hl[k]] = h[sc][k];
      {
        const int zh = add_wrap(zh0, u, n3), zq = add_wrap(zq0, u, n3);
        const unsigned zo = unsigned(zh) * n12;
#pragma unroll
        for (int k = 0; k < HM; ++k) h[sn][k] = w[hg[k] + zo];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
          v[sn][r] = vloc[colofs[r] + zo];
          if (ACC) a[sn][r] = __builtin_nontemporal_load(o + colofs[r] + zo);
        }
        const unsigned zf = unsigned(zq) * n12;
#pragma unroll
        for (int r = 0; r < RY; ++r) q[r][(u + R - 1) % R] = w[colofs[r] + zf];
      }
      lds_barrier();
      const unsigned zb2 = unsigned(z) * n12;
      T col[RY + 2 * MD];
#pragma unroll
      for (int k = 0; k < RY + 2 * MD; ++k) col[k] = buf[(g * RY + k) * n1 + x];
#pragma unroll
      for (int r = 0; r < RY; ++r) {
        const T *row = buf + (MD + g * RY + r) * n1;
        // six independent partial sums (axis, parity of m) instead of one
        // chain of 3 MD + 1 dependent FMAs: the chain's latency, not the
        // issue rate, otherwise bounds the step
        T sx[2] = {T(0), T(0)}, sy[2] = {T(0), T(0)};
        T sz[2] = {(cf.c0 + v[sc][r]) * q[r][(u + MD) % R], T(0)};
#pragma unroll
        for (int m = 0; m < MD; ++m) {
          sx[m & 1] += cf.c[0][m] * (row[xm[m]] + row[xp[m]]);
          sy[m & 1] += cf.c[1][m] * (col[r + MD - 1 - m] + col[r + MD + 1 + m]);
          sz[m & 1] += cf.c[2][m] * (q[r][(u + MD - 1 - m) % R] + q[r][(u + MD + 1 + m) % R]);
        }
        T acc = ((sx[0] + sx[1]) + (sy[0] + sy[1])) + (sz[0] + sz[1]);
        if (ACC) acc += a[sc][r];
        T *dst = own[r] ? o + colofs[r] + zb2 : dst_dead;
        __builtin_nontemporal_store(acc, dst);
      }
    }
  }
}

// ---- pair variant: two adjacent x-points per thread ----
// Same operator and streaming as k_stencil.  Each LDS row carries MD wrapped
// copies of the row's ends on both sides (row length n1 + 2 MD), so the
// x-window of a pair is MD + 1 aligned pair reads and every y-column read
// serves both points: about 40% fewer LDS bytes per point, at twice the
// z-ring registers.  Needs n1 even and full y-tiles (ty | n2); the step loop
// is unrolled by an index-sequence fold so every ring slot is a compile-time
// register.
template <typename T> struct PairOf;
template <> struct PairOf<float> { typedef float2 V; };
template <> struct PairOf<double> { typedef double2 V; };

// H psi is streamed: its accumulated input is read once and the result is
// not read again by this launch, so both go past the caches with
// non-temporal hints, which leaves L2 to the psi rows that neighbouring
// tiles share as halo.
template <typename T> struct NativePair;
template <> struct NativePair<float> { typedef float V __attribute__((ext_vector_type(2))); };
template <> struct NativePair<double> { typedef double V __attribute__((ext_vector_type(2))); };
template <typename T> __device__ inline typename PairOf<T>::V nt_load_pair(const T *p) {
  const typename NativePair<T>::V n = __builtin_nontemporal_load(reinterpret_cast<const typename NativePair<T>::V *>(p));
  typename PairOf<T>::V v;
  v.x = n.x;
  v.y = n.y;
  return v;
}
template <typename T> __device__ inline void nt_store_pair(T *p, const typename PairOf<T>::V &v) {
  typename NativePair<T>::V n;
  n.x = v.x;
  n.y = v.y;
  __builtin_nontemporal_store(n, reinterpret_cast<typename NativePair<T>::V *>(p));
}

template <class F, int... U> __device__ inline void run_steps(F &f, int zb, std::integer_sequence<int, U...>) {
  (void)(f(zb, std::integral_constant<int, U>{}) && ...);
}

template <typename T, int MD, int RY, int HM, bool ACC>
__global__ __launch_bounds__(512) void k_stencil2(const T *psi, T *hpsi, int64_t ld, const T *vloc,
                                                  StencilCoef<T> cf, int n1, int n2, int n3, int ty, int zlen,
                                                  T *dead) {
  typedef typename PairOf<T>::V V;
  constexpr int PF = ring_pf(MD);
  constexpr int R = 2 * MD + 1 + PF;
  static_assert(R % kHS == 0 && R % 2 == 0, "ring length must be a multiple of the rotations");
  extern __shared__ unsigned char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int LS = n1 + 2 * MD;
  const int plane = (ty + 2 * MD) * LS;
  const int hp = n1 / 2;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int deadslot = 2 * plane + 2 * (tid & 63);  // a pair of this lane's own past both buffers
  const bool active = tid < hp * (ty / RY);
  const int px = active ? tid % hp : 0, g = active ? tid / hp : 0;
  const int x = 2 * px;
  const int ntile_y = n2 / ty, nz = (n3 + zlen - 1) / zlen;
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int by = int(lb % unsigned(ntile_y)), bz = int((lb / unsigned(ntile_y)) % unsigned(nz));
  const int band = int(lb / unsigned(ntile_y * nz));
  const int y0 = by * ty;
  const int z0 = bz * zlen, z1 = min(z0 + zlen, n3);
  const unsigned n12 = unsigned(n1) * unsigned(n2);  // in-band offsets < N < 2^31
  const T *w = psi + int64_t(band) * ld;
  T *o = hpsi + int64_t(band) * ld;
  unsigned colofs[RY];
  int lown[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    colofs[r] = unsigned(x + (y0 + g * RY + r) * n1);
    lown[r] = active ? (MD + g * RY + r) * LS + MD + x : deadslot;
  }
  // the pair's wrapped copy: x < MD -> right ghost, x >= n1 - MD -> left ghost
  const int gofs = !active ? 0 : (x < MD ? n1 : (x >= n1 - MD ? -n1 : 0));
  const int nhalo = 2 * MD * hp;
  int hl[HM], hgh[HM];
  unsigned hg[HM];
#pragma unroll
  for (int k = 0; k < HM; ++k) {
    const int idx = tid + k * nthr;
    const int i = idx < nhalo ? idx : 0;
    const int hr = i / hp, hx = 2 * (i - hr * hp);
    const int lrow = hr < MD ? hr : hr + ty;
    hl[k] = idx < nhalo ? lrow * LS + MD + hx : deadslot;
    hgh[k] = (idx < nhalo && (hx < MD || hx >= n1 - MD)) ? hl[k] + (hx < MD ? n1 : -n1) : deadslot;
    hg[k] = unsigned(hx + wrapi(y0 - MD + lrow, n2) * n1);
  }
  V q[RY][R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const T *wp = w + unsigned(wrapi(z0 - MD + k, n3)) * n12;
#pragma unroll
    for (int r = 0; r < RY; ++r) q[r][k] = *reinterpret_cast<const V *>(wp + colofs[r]);
  }
  V h[kHS][HM], v[kHS][RY], a[kHS][RY];
#pragma unroll
  for (int sidx = 0; sidx < kHS - 1; ++sidx) {
    const unsigned zo = unsigned(wrapi(z0 + sidx, n3)) * n12;
#pragma unroll
    for (int k = 0; k < HM; ++k) h[sidx][k] = *reinterpret_cast<const V *>(w + zo + hg[k]);
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      v[sidx][r] = *reinterpret_cast<const V *>(vloc + zo + colofs[r]);
      if (ACC) a[sidx][r] = nt_load_pair(o + zo + colofs[r]);
    }
  }
  V *odead = reinterpret_cast<V *>(dead) + (tid & 63);
  int zh = wrapi(z0 + kHS - 1, n3);  // plane of the halo / Vloc / H psi prefetch
  int zq = wrapi(z0 + MD + PF, n3);  // plane of the own-column prefetch
  auto step = [&](int zb, auto uc) -> bool {
    constexpr int u = decltype(uc)::value;
    const int z = zb + u;
    if (z >= z1) return false;
    T *buf = lds + (u & 1) * plane;
    constexpr int sc = u % kHS, sn = (u + kHS - 1) % kHS;
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const V val = q[r][(u + MD) % R];
      *reinterpret_cast<V *>((lown[r] == deadslot ? lds : buf) + lown[r]) = val;
      *reinterpret_cast<V *>(gofs ? buf + lown[r] + gofs : lds + deadslot) = val;
    }
#pragma unroll
    for (int k = 0; k < HM; ++k) {
      *reinterpret_cast<V *>((hl[k] == deadslot ? lds : buf) + hl[k]) = h[sc][k];
      *reinterpret_cast<V *>((hgh[k] == deadslot ? lds : buf) + hgh[k]) = h[sc][k];
    }
    {
      const unsigned zo = unsigned(zh) * n12;
#pragma unroll
      for (int k = 0; k < HM; ++k) h[sn][k] = *reinterpret_cast<const V *>(w + zo + hg[k]);
#pragma unroll
      for (int r = 0; r < RY; ++r) {
        v[sn][r] = *reinterpret_cast<const V *>(vloc + zo + colofs[r]);
        if (ACC) a[sn][r] = nt_load_pair(o + zo + colofs[r]);
      }
      const T *wq = w + unsigned(zq) * n12;
#pragma unroll
      for (int r = 0; r < RY; ++r) q[r][(u + R - 1) % R] = *reinterpret_cast<const V *>(wq + colofs[r]);
      zh = (zh + 1 == n3) ? 0 : zh + 1;
      zq = (zq + 1 == n3) ? 0 : zq + 1;
    }
    lds_barrier();
    T *oz = o + unsigned(z) * n12;
    V col[RY + 2 * MD];
#pragma unroll
    for (int k = 0; k < RY + 2 * MD; ++k) col[k] = *reinterpret_cast<const V *>(buf + (g * RY + k) * LS + MD + x);
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const T *row = buf + (MD + g * RY + r) * LS + x;  // ghost-padded x - MD .. x + MD + 1
      T xw[2 * MD + 2];
#pragma unroll
      for (int k = 0; k < MD + 1; ++k) {
        const V p2 = *reinterpret_cast<const V *>(row + 2 * k);
        xw[2 * k] = p2.x;
        xw[2 * k + 1] = p2.y;
      }
      const V c = q[r][(u + MD) % R];
      // per point three independent partial sums (one per axis)
      T x0 = T(0), x1 = T(0), y0s = T(0), y1s = T(0);
      T z0 = (cf.c0 + v[sc][r].x) * c.x;
      T z1 = (cf.c0 + v[sc][r].y) * c.y;
#pragma unroll
      for (int m = 0; m < MD; ++m) {
        x0 += cf.c[0][m] * (xw[MD - 1 - m] + xw[MD + 1 + m]);
        x1 += cf.c[0][m] * (xw[MD - m] + xw[MD + 2 + m]);
        y0s += cf.c[1][m] * (col[r + MD - 1 - m].x + col[r + MD + 1 + m].x);
        y1s += cf.c[1][m] * (col[r + MD - 1 - m].y + col[r + MD + 1 + m].y);
        const V lo = q[r][(u + MD - 1 - m) % R], hi = q[r][(u + MD + 1 + m) % R];
        z0 += cf.c[2][m] * (lo.x + hi.x);
        z1 += cf.c[2][m] * (lo.y + hi.y);
      }
      T a0 = (x0 + y0s) + z0;
      T a1 = (x1 + y1s) + z1;
      if (ACC) {
        a0 += a[sc][r].x;
        a1 += a[sc][r].y;
      }
      V res;
      res.x = a0;
      res.y = a1;
      nt_store_pair(active ? reinterpret_cast<T *>(oz + colofs[r]) : reinterpret_cast<T *>(odead), res);
    }
    return true;
  };
  for (int zb = z0; zb < z1; zb += R) run_steps(step, zb, std::make_integer_sequence<int, R>{});
}

template <typename T> __device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

constexpr int kGatherThreads = 256;

template <typename T, int BB>
__global__ __launch_bounds__(kGatherThreads) void k_nl_gather(
    const T *__restrict__ psi, int64_t ld, int nbc, const int *__restrict__ seg_beg,
    const int *__restrict__ seg_end, const int *__restrict__ grp_pts,
    const T *__restrict__ grp_coef, T *__restrict__ part) {
  __shared__ T red[kGatherThreads / 64][4 * BB];
  const int s = blockIdx.x;
  const int b0 = blockIdx.y * BB;
  const int nbb = min(BB, nbc - b0);
  const int beg = seg_beg[s], end = seg_end[s];
  T acc[4][BB];
#pragma unroll
  for (int l = 0; l < 4; ++l)
#pragma unroll
    for (int b = 0; b < BB; ++b) acc[l][b] = T(0);
  const T *__restrict__ w = psi + int64_t(b0) * ld;
  for (int i = beg + threadIdx.x; i < end; i += kGatherThreads) {
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
      const T v = wave_sum(acc[l][b]);
      if (lane == 0) red[wid][l * BB + b] = v;
    }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < 4 * BB) {
    const int l = t / BB, b = t % BB;
    if (b < nbb) {
      T v = T(0);
#pragma unroll
      for (int k = 0; k < kGatherThreads / 64; ++k) v += red[k][t];
      part[(int64_t(s) * 4 + l) * nbc + b0 + b] = v;
    }
  }
}

template <typename T>
__global__ void k_nl_reduce(const T *__restrict__ part, int nbc, int nz, const int *__restrict__ lma_grp,
                            const int *__restrict__ grp_lma0, const int *__restrict__ grp_seg_off,
                            const T *__restrict__ scale, T *__restrict__ uv) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nz * nbc) return;
  const int lma = t / nbc, b = t - lma * nbc;
  const int g = lma_grp[lma], l = lma - grp_lma0[g];
  T v = T(0);
  for (int s = grp_seg_off[g]; s < grp_seg_off[g + 1]; ++s) v += part[(int64_t(s) * 4 + l) * nbc + b];
  uv[int64_t(lma) * nbc + b] = scale[lma] * v;
}

// Sliced-ELL scatter: slice c holds the entries of points 64c..64c+63 as
// slot-major arrays (slot k of lane i at off[c] + 64k + i), so the slot
// loop is uniform across the wave and every table load is coalesced.
template <typename T, int BT>
__global__ __launch_bounds__(256) void k_nl_scatter(T *__restrict__ hpsi, int64_t ld, int nbc, int64_t N,
                                                    const int *__restrict__ sl_off,
                                                    const int *__restrict__ sl_len,
                                                    const int *__restrict__ sl_lma,
                                                    const T *__restrict__ sl_uv,
                                                    const T *__restrict__ uv) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t c = j >> 6;
  const int lane = threadIdx.x & 63;
  if (c * 64 >= N) return;  // whole slice past the grid (wave-uniform)
  const int b0 = blockIdx.y * BT;
  const int nbt = min(BT, nbc - b0);
  T acc[BT];
#pragma unroll
  for (int b = 0; b < BT; ++b) acc[b] = T(0);
  const int base = sl_off[c], K = sl_len[c];
  for (int k = 0; k < K; ++k) {
    const int e = base + k * 64 + lane;
    const T cv = sl_uv[e];
    const T *__restrict__ u = uv + int64_t(sl_lma[e]) * nbc + b0;
#pragma unroll
    for (int b = 0; b < BT; ++b)
      if (b < nbt) acc[b] += cv * u[b];
  }
  if (j < N) {
#pragma unroll
    for (int b = 0; b < BT; ++b)
      if (b < nbt) hpsi[j + int64_t(b0 + b) * ld] = acc[b];
  }
}

// C[(g*4 + l)*nbc + b] = iuV dV * sum over the group's segments (npc
// partial slots per segment, contiguous).  A block owns one (g, l) row and
// 64 bands; each of its four waves sums every fourth slot (four short load
// chains instead of one long one), and the four sums are added in a fixed
// order.
template <typename T>
__global__ __launch_bounds__(256) void k_proj_reduce(const T *__restrict__ part, int nbc, int ngroup,
                                                     const int *__restrict__ grp_seg_off, int npc,
                                                     const T *__restrict__ scale, T *__restrict__ cproj) {
  __shared__ T red[4][64];
  const int nbb = (nbc + 63) / 64;
  const int gl = int(blockIdx.x) / nbb, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int b = (int(blockIdx.x) % nbb) * 64 + lane;
  const int g = gl >> 2, l = gl & 3;
  T v = T(0);
  if (b < nbc)
    for (int s = grp_seg_off[g] * npc + w; s < grp_seg_off[g + 1] * npc; s += 4)
      v += part[(int64_t(s) * 4 + l) * nbc + b];
  red[w][lane] = v;
  __syncthreads();
  if (w == 0 && b < nbc)
    cproj[int64_t(gl) * nbc + b] = scale[gl] * ((red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]));
}

// ---- analytic projectors on the matrix cores ----
//
// Both products are tiny sparse ones (about 5.6 projector values per grid
// point), so they run as dense 16x16x4 MFMA tiles over a few zero-padded
// projectors: the point data is read once per tile and the band dimension
// rides on the matrix-core columns.  Lane maps (verified on gfx950):
//   A (16 x 4): lane l holds A[l % 16][l / 16]
//   B (4 x 16): lane l holds B[l / 16][l % 16]
//   D (16 x 16), register r of lane l:
//     f32: D[4 (l / 16) + r][l % 16]      f64: D[4 r + l / 16][l % 16]
typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef mf_d4 V;
  __device__ static V mac(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int l, int r) { return 4 * r + (l >> 4); }
};
template <> struct Mfma<float> {
  typedef mf_f4 V;
  __device__ static V mac(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
};

// Projector k (0: s, 1..3: p_x, p_y, p_z) of group g at grid point (x, y, z)
// with sphere flags fs, fp.
template <typename T>
__device__ inline T ana_value_k(const T *__restrict__ es, const T *__restrict__ ep, const T *__restrict__ dd,
                                int64_t gS, int oy, int oz, int x, int y, int z, bool fs, bool fp, T amp_s,
                                T amp_p, int k) {
  const int64_t ix = gS + x, iy = gS + oy + y, iz = gS + oz + z;
  if (k == 0) return fs ? amp_s * es[ix] * es[iy] * es[iz] : T(0);
  const int64_t ia = k == 1 ? ix : (k == 2 ? iy : iz);
  return fp ? amp_p * ep[ix] * ep[iy] * ep[iz] * dd[ia] : T(0);
}

// Partial projections of one piece of a segment (one atom, one plane run;
// npc pieces of whole 32-point passes per segment, slot s*npc + q):
//   part[((s*npc + q)*4 + k)*nbc + b] = sum_{j in piece} beta_k(j) psi(j, b)
// D[band][k] += A[band][point] B[point][k]: one wave, NBT tiles of 16 bands.
// Short pieces shorten the time between the reads of two spheres that
// share a chord of the plane, so the second read finds it in L2.
// The A operand puts 16 bands of one point on neighbouring lanes, so psi is
// first staged through LDS in passes of 32 points, loaded with neighbouring
// lanes on neighbouring points (rows padded by one element against bank
// conflicts).
template <typename T, int NBT>
__global__ __launch_bounds__(64) void k_proj_m(
    const T *__restrict__ psi, int64_t ld, int nbc, int n1, int n2, int nseg, int npc,
    const int *__restrict__ seg_order, const int *__restrict__ seg_beg, const int *__restrict__ seg_end,
    const int *__restrict__ seg_grp, const uint32_t *__restrict__ pts, const T *__restrict__ es,
    const T *__restrict__ ep, const T *__restrict__ dd, const T *__restrict__ amp, int S, T *__restrict__ part) {
  using M = Mfma<T>;
  constexpr int P = 32;  // points per pass
  __shared__ T tile[NBT][16][P + 1];
  __shared__ T bsh[4][P + 1];  // projector values of the pass
  // logical block: piece fastest, then segment (in plane order), then band group
  const unsigned lblk = xcd_block(blockIdx.x, gridDim.x);
  const unsigned nsp = unsigned(nseg) * unsigned(npc), sp = lblk % nsp;
  const int s = seg_order[sp / unsigned(npc)], q = int(sp % unsigned(npc));
  const int l = threadIdx.x;
  const int lo = l & 15, hi = l >> 4;
  const int lp = l & (P - 1), lb = l / P;  // staging: point, band parity
  const int b0 = int(lblk / nsp) * 16 * NBT;
  const int sbeg = seg_beg[s], send = seg_end[s], g = seg_grp[s];
  const int npass = (send - sbeg + P - 1) / P;
  const int beg = sbeg + P * (npass * q / npc), end = min(send, sbeg + P * (npass * (q + 1) / npc));
  const int64_t gS = int64_t(g) * S;
  const T amp_s = amp[4 * g], amp_p = amp[4 * g + 1];
  const int64_t n12 = int64_t(n1) * n2;
  typename M::V acc[NBT];
#pragma unroll
  for (int t = 0; t < NBT; ++t) acc[t] = typename M::V{};
  uint32_t wd_next = beg < end ? pts[beg + lp < end ? beg + lp : beg] : 0u;  // a piece may be empty
  for (int i = beg; i < end; i += P) {
    {
      const int ip = i + lp;
      const bool valid = ip < end;
      const uint32_t wd = wd_next;
      wd_next = pts[ip + P < end ? ip + P : beg];  // next pass's point, in flight during this one
      const int x = int(wd & 1023u), y = int((wd >> 10) & 1023u), z = int((wd >> 20) & 1023u);
      const int64_t lin = x + int64_t(y) * n1 + int64_t(z) * n12;
      // this lane's two projector values: slots 2 lb, 2 lb + 1
      const bool fs = valid && ((wd >> 30) & 1u), fp = valid && (wd >> 31);
      const T pv0 = ana_value_k<T>(es, ep, dd, gS, n1, n1 + n2, x, y, z, fs, fp, amp_s, amp_p, 2 * lb);
      const T pv1 = ana_value_k<T>(es, ep, dd, gS, n1, n1 + n2, x, y, z, fs, fp, amp_s, amp_p, 2 * lb + 1);
      T v[NBT][8];
#pragma unroll
      for (int t = 0; t < NBT; ++t)
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          const int b = b0 + 16 * t + 2 * m + lb;
          v[t][m] = (valid && b < nbc) ? psi[lin + int64_t(b) * ld] : T(0);
        }
      __syncthreads();  // previous pass done with the staging tile
#pragma unroll
      for (int t = 0; t < NBT; ++t)
#pragma unroll
        for (int m = 0; m < 8; ++m) tile[t][2 * m + lb][lp] = v[t][m];
      bsh[2 * lb][lp] = pv0;
      bsh[2 * lb + 1][lp] = pv1;
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < P / 4; ++u) {
      const int q = 4 * u + hi;  // K index: point of the pass
      const T bv = lo < 4 ? bsh[lo][q] : T(0);
#pragma unroll
      for (int t = 0; t < NBT; ++t) acc[t] = M::mac(tile[t][lo][q], bv, acc[t]);
    }
  }
  // D[band][k]: keep k < 4
#pragma unroll
  for (int t = 0; t < NBT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int band = b0 + 16 * t + M::row(l, r);
      if (lo < 4 && band < nbc) part[((int64_t(s) * npc + q) * 4 + lo) * nbc + band] = acc[t][r];
    }
}

// Partial projections by row tiles, reading each psi value once.  A block
// owns rt_ty grid rows (all of x, at most kRowTileMax points) of rt_zr
// planes and 16 bands.  Plane by plane it stages those rows of psi in LDS
// with coalesced loads; each wave then takes pieces (one atom's sphere
// points on the plane and in the rows) and forms
// D[band][k] = sum_points psi(point, band) beta_k(point) on the matrix
// cores, in passes of 32 points whose projector values the wave first
// forms into LDS (one dependent load chain per pass, not per K step).  The
// result is added into the block's per-atom sums in LDS (one piece per atom
// and plane, so the waves of a plane never share a sum).  At the end each
// per-atom sum goes to its global slot; an atom's slots are contiguous, and
// k_proj_reduce sums them.
constexpr int kRowTileMax = 288;

template <typename T>
__global__ __launch_bounds__(256) void k_proj_rows(
    const T *__restrict__ psi, int64_t ld, int nbc, int n1, int n2, int ty, int ny, int nrt,
    const int4 *__restrict__ piece, const int *__restrict__ piece_off, const int *__restrict__ slot_off,
    const int *__restrict__ slot_gid, const uint32_t *__restrict__ pts, const T *__restrict__ es,
    const T *__restrict__ ep, const T *__restrict__ dd, const T *__restrict__ amp, int S, T *__restrict__ part) {
  using M = Mfma<T>;
  constexpr int NB = 16, P = 32, QT = (kRowTileMax + 255) / 256;
  extern __shared__ unsigned char smem_raw[];
  __shared__ T bsh[4][4][P + 1];  // per wave: projector values of the pass
  __shared__ int lsh[4][P];        // per wave: the pass's points in the tile (-1: none)
  const unsigned lblk = xcd_block(blockIdx.x, gridDim.x);
  const int blk = int(lblk % unsigned(nrt)), b0 = int(lblk / unsigned(nrt)) * NB;
  const int y0 = (blk % ny) * ty, rows = min(ty, n2 - y0);
  const int np = rows * n1;  // <= kRowTileMax (checked by the launcher)
  const int LS = np | 1;     // odd band stride: the 16 bands of an A operand fall in distinct banks
  T *tile = reinterpret_cast<T *>(smem_raw);  // [NB][LS]
  const int s0 = slot_off[blk], nslot = slot_off[blk + 1] - s0;
  T *acc = tile + NB * LS;  // [nslot][4][NB]
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int lo = l & 15, hi = l >> 4, lp = l & (P - 1), lb = l / P;
  for (int e = tid; e < nslot * 4 * NB; e += 256) acc[e] = T(0);
  const int64_t n12 = int64_t(n1) * n2;
  const int pb = piece_off[blk], pe = piece_off[blk + 1];
  for (int i = pb; i < pe;) {
    const int z = piece[i].z >> 8;
    int j = i + 1;
    while (j < pe && (piece[j].z >> 8) == z) ++j;
    {
      // all loads of the plane's rows first, then the LDS stores
      const T *src = psi + int64_t(z) * n12 + int64_t(y0) * n1;
      T v[NB][QT];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const T *sb = src + int64_t(b0 + b < nbc ? b0 + b : 0) * ld;
#pragma unroll
        for (int k = 0; k < QT; ++k) {
          const int q = tid + 256 * k;
          v[b][k] = (q < np && b0 + b < nbc) ? sb[q] : T(0);
        }
      }
      __syncthreads();  // the previous plane's tile is no longer read
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int k = 0; k < QT; ++k) {
          const int q = tid + 256 * k;
          if (q < np) tile[b * LS + q] = v[b][k];
        }
      __syncthreads();
    }
    for (int p = i + wave; p < j; p += 4) {
      const int4 pc = piece[p];
      const int slot = pc.z & 255, g = pc.w;
      const int64_t gS = int64_t(g) * S;
      const T amp_s = amp[4 * g], amp_p = amp[4 * g + 1];
      // two accumulators (even and odd K steps): two independent chains of
      // matrix-core operations
      typename M::V d = typename M::V{}, d2 = typename M::V{};
      for (int q0 = pc.x; q0 < pc.y; q0 += P) {
        {
          const int q = q0 + lp;
          const bool valid = q < pc.y;
          const uint32_t w = pts[valid ? q : pc.x];
          const int x = int(w & 1023u), y = int((w >> 10) & 1023u);
          const bool fs = valid && ((w >> 30) & 1u), fp = valid && (w >> 31);
          const T pv0 = ana_value_k<T>(es, ep, dd, gS, n1, n1 + n2, x, y, z, fs, fp, amp_s, amp_p, 2 * lb);
          const T pv1 = ana_value_k<T>(es, ep, dd, gS, n1, n1 + n2, x, y, z, fs, fp, amp_s, amp_p, 2 * lb + 1);
          // the previous pass's reads of this wave's slots are done (LDS
          // operations of one wave complete in order)
          __builtin_amdgcn_wave_barrier();
          bsh[wave][2 * lb][lp] = pv0;
          bsh[wave][2 * lb + 1][lp] = pv1;
          if (lb == 0) lsh[wave][lp] = valid ? (y - y0) * n1 + x : -1;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
#pragma unroll
        for (int u = 0; u < P / 4; ++u) {
          const int kk = 4 * u + hi;
          const int loc = lsh[wave][kk];
          const T av = loc >= 0 ? tile[lo * LS + loc] : T(0);
          const T bv = lo < 4 ? bsh[wave][lo][kk] : T(0);
          if (u & 1)
            d2 = M::mac(av, bv, d2);
          else
            d = M::mac(av, bv, d);
        }
      }
      if (lo < 4) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[(slot * 4 + lo) * NB + M::row(l, r)] += d[r] + d2[r];
      }
    }
    i = j;
  }
  __syncthreads();
  for (int e = tid; e < nslot * 4 * NB; e += 256) {
    const int slot = e / (4 * NB), rem = e - slot * 4 * NB, k = rem / NB, b = rem - k * NB;
    if (b0 + b < nbc) part[(int64_t(slot_gid[s0 + slot]) * 4 + k) * nbc + b0 + b] = acc[e];
  }
}

// H psi(j, b) (=|+=) sum over the atoms g meeting j's tile of
//   sum_k beta_k(j) C(g, k, b)
// One wave per 16-point tile: D[band][point] += A[band][k] B[k][point] per
// atom, NBT tiles of 16 bands.
template <typename T, int NBT, bool ACC>
__global__ __launch_bounds__(256) void k_apply_m(
    T *__restrict__ hpsi, int64_t ld, int nbc, int64_t N, int n1, int n2, int ntile,
    const int *__restrict__ tile_off, const int *__restrict__ tile_g, const uint64_t *__restrict__ hits,
    const T *__restrict__ es, const T *__restrict__ ep, const T *__restrict__ dd, const T *__restrict__ amp,
    int S, const T *__restrict__ cproj) {
  using M = Mfma<T>;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= ntile) return;  // whole wave
  // accumulating onto a tile no sphere meets adds nothing: leave it alone
  if (ACC && tile_off[tile] == tile_off[tile + 1]) return;  // whole wave
  const int l = threadIdx.x & 63;
  const int lo = l & 15, hi = l >> 4;
  const int b0 = blockIdx.y * 16 * NBT;
  const int64_t j = int64_t(tile) * 16 + lo;
  const bool inside = j < N;
  const int64_t jj = inside ? j : 0;
  const int x = int(jj % n1), y = int((jj / n1) % n2), z = int(jj / (int64_t(n1) * n2));
  const uint64_t hw = inside ? hits[jj] : 0;
  typename M::V acc[NBT];
#pragma unroll
  for (int t = 0; t < NBT; ++t) acc[t] = typename M::V{};
  // accumulating: the H psi values this lane adds to, loaded before the
  // atom loop so that their latency overlaps the matrix-core work
  T old[NBT][4];
#pragma unroll
  for (int t = 0; t < NBT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int band = b0 + 16 * t + M::row(l, r);
      old[t][r] = (ACC && inside && band < nbc) ? __builtin_nontemporal_load(hpsi + j + int64_t(band) * ld) : T(0);
    }
  const int a0 = tile_off[tile], a1 = tile_off[tile + 1];
  for (int a = a0; a < a1; ++a) {
    const int g = tile_g[a];
    bool fs = false, fp = false;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t e = uint32_t(hw >> (16 * q)) & 0xffffu;
      const bool me = int(e & 0x3fffu) == g + 1;
      fs |= me && ((e >> 14) & 1u);
      fp |= me && ((e >> 15) & 1u);
    }
    const T bv = ana_value_k<T>(es, ep, dd, int64_t(g) * S, n1, n1 + n2, x, y, z, fs, fp, amp[4 * g],
                                amp[4 * g + 1], hi);
    const T *__restrict__ c = cproj + (int64_t(g) * 4 + hi) * nbc;
#pragma unroll
    for (int t = 0; t < NBT; ++t) {
      const int b = b0 + 16 * t + lo;
      const T av = b < nbc ? c[b] : T(0);
      acc[t] = M::mac(av, bv, acc[t]);
    }
  }
  if (!inside) return;
#pragma unroll
  for (int t = 0; t < NBT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int band = b0 + 16 * t + M::row(l, r);
      if (band < nbc) __builtin_nontemporal_store(ACC ? old[t][r] + acc[t][r] : acc[t][r], hpsi + j + int64_t(band) * ld);
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
template <typename T> AnalyticDev<T> &ana_of(rsg_ctx *ctx);
template <> AnalyticDev<double> &ana_of<double>(rsg_ctx *ctx) { return ctx->ana64; }
template <> AnalyticDev<float> &ana_of<float>(rsg_ctx *ctx) { return ctx->ana32; }
template <typename T> const T *vloc_of(rsg_ctx *ctx);
template <> const double *vloc_of<double>(rsg_ctx *ctx) { return ctx->vloc64.as<double>(); }
template <> const float *vloc_of<float>(rsg_ctx *ctx) { return ctx->vloc32.as<float>(); }
template <typename T> HpsiShape &shape_of(rsg_ctx *ctx);
template <> HpsiShape &shape_of<double>(rsg_ctx *ctx) { return ctx->shape64; }
template <> HpsiShape &shape_of<float>(rsg_ctx *ctx) { return ctx->shape32; }

constexpr size_t kMaxDynLds = 64 * 1024;

inline int round64(int v) { return (v + 63) / 64 * 64; }

// Tile rows per workgroup for ry rows per thread: balance idle lanes, the
// partial last tile and the halo rows each plane re-reads.
int pick_ty(int n1, int n2, int md, int ry, size_t s) {
  int best_ty = 0;
  double best = 1e300;
  for (int G = 1; n1 * G <= 512; ++G) {
    const int ty = G * ry;
    if (ty > ((n2 + ry - 1) / ry) * ry) break;
    if ((2 * size_t(ty + 2 * md) * n1 + 64) * s > kMaxDynLds) break;
    const int th = n1 * G;
    if (2 * md * n1 > 4 * round64(th)) continue;
    const int tiles = (n2 + ty - 1) / ty;
    const double cost = double(round64(th)) / th * double(tiles * ty) / n2 *
                        (1.0 + 0.5 * 2.0 * md / ty) * (th < 128 ? 2.0 : 1.0);
    if (cost < best) {
      best = cost;
      best_ty = ty;
    }
  }
  return best_ty;
}

// Pair kernel (two x-points per thread): tile rows for ry rows per
// thread among the divisors of n2, same cost model as pick_ty.
constexpr int kPairMd = 6;

inline bool pair_fits(int n1, int n2, int md, int ry, int ty, size_t s) {
  if (md != kPairMd || n1 % 2 || n1 < 2 * md || ty <= 0 || ty % ry || n2 % ty) return false;
  const int th = round64((n1 / 2) * (ty / ry));
  const size_t lds = (2 * size_t(ty + 2 * md) * (n1 + 2 * md) + 128) * s;
  return th <= 512 && lds <= kMaxDynLds && 2 * md * (n1 / 2) <= 4 * th;
}

int pick_ty_pair(int n1, int n2, int md, int ry, size_t s) {
  int best_ty = 0;
  double best = 1e300;
  for (int ty = ry; ty <= n2; ty += ry) {
    if (!pair_fits(n1, n2, md, ry, ty, s)) continue;
    const int th = (n1 / 2) * (ty / ry);
    const double cost = double(round64(th)) / th * (1.0 + 0.5 * 2.0 * md / ty) * (th < 128 ? 2.0 : 1.0);
    if (cost < best) {
      best = cost;
      best_ty = ty;
    }
  }
  return best_ty;
}

template <int RY, int HM>
int launch_stencil2_t(rsg_ctx *ctx, const float *psi, float *hpsi, int64_t ld, int nb, const HpsiShape &sh,
                      int accumulate) {
  constexpr int MD = kPairMd;
  const int n1 = ctx->n[0], n2 = ctx->n[1], n3 = ctx->n[2];
  if (!pair_fits(n1, n2, ctx->md, RY, sh.ty, sizeof(float)))
    return ctx->fail(RSG_ERR_UNSUPPORTED, "hpsi: pair stencil tile does not fit");
  const int threads = round64((n1 / 2) * (sh.ty / RY));
  if (2 * MD * (n1 / 2) > HM * threads) return ctx->fail(RSG_ERR_UNSUPPORTED, "hpsi: pair stencil halo");
  const size_t lds = (2 * size_t(sh.ty + 2 * MD) * (n1 + 2 * MD) + 128) * sizeof(float);
  const int64_t nblk = int64_t(n2 / sh.ty) * ((n3 + sh.zlen - 1) / sh.zlen) * nb;
  if (nblk > 0x7fffffff) return ctx->fail(RSG_ERR_UNSUPPORTED, "hpsi: too many stencil workgroups");
  const dim3 grid{unsigned(nblk), 1u, 1u};
  if (accumulate)
    hipLaunchKernelGGL((k_stencil2<float, MD, RY, HM, true>), grid, dim3(threads), lds, ctx->stream, psi, hpsi, ld,
                       vloc_of<float>(ctx), make_coef<float>(ctx), n1, n2, n3, sh.ty, sh.zlen,
                       ctx->dead.template as<float>());
  else
    hipLaunchKernelGGL((k_stencil2<float, MD, RY, HM, false>), grid, dim3(threads), lds, ctx->stream, psi, hpsi,
                       ld, vloc_of<float>(ctx), make_coef<float>(ctx), n1, n2, n3, sh.ty, sh.zlen,
                       ctx->dead.template as<float>());
  return RSG_OK;
}

template <int RY>
int launch_stencil2_hm(rsg_ctx *ctx, const float *psi, float *hpsi, int64_t ld, int nb, const HpsiShape &sh,
                       int accumulate) {
  const int hp = ctx->n[0] / 2;
  const int threads = round64(hp * (sh.ty / RY));
  const int need = (2 * kPairMd * hp + threads - 1) / threads;
  if (need <= 1) return launch_stencil2_t<RY, 1>(ctx, psi, hpsi, ld, nb, sh, accumulate);
  if (need <= 2) return launch_stencil2_t<RY, 2>(ctx, psi, hpsi, ld, nb, sh, accumulate);
  return launch_stencil2_t<RY, 4>(ctx, psi, hpsi, ld, nb, sh, accumulate);
}

template <typename T, int MD, int RY, int HM>
int launch_stencil_t(rsg_ctx *ctx, const T *psi, T *hpsi, int64_t ld, int nb, const HpsiShape &sh,
                     int accumulate) {
  const int n1 = ctx->n[0], n2 = ctx->n[1], n3 = ctx->n[2];
  const int G = sh.ty / RY;
  const int threads = round64(n1 * G);
  const size_t lds = (2 * size_t(sh.ty + 2 * MD) * n1 + 64) * sizeof(T);
  if (threads > 512 || lds > kMaxDynLds || 2 * MD * n1 > HM * threads)
    return ctx->fail(RSG_ERR_UNSUPPORTED, "hpsi: stencil tile does not fit");
  const int64_t nblk = int64_t((n2 + sh.ty - 1) / sh.ty) * ((n3 + sh.zlen - 1) / sh.zlen) * nb;
  if (nblk > 0x7fffffff) return ctx->fail(RSG_ERR_UNSUPPORTED, "hpsi: too many stencil workgroups");
  const dim3 grid{unsigned(nblk), 1u, 1u};
  if (accumulate)
    hipLaunchKernelGGL((k_stencil<T, MD, RY, HM, true>), grid, dim3(threads), lds, ctx->stream, psi, hpsi, ld,
                       vloc_of<T>(ctx), make_coef<T>(ctx), n1, n2, n3, sh.ty, sh.zlen, ctx->dead.template as<T>());
  else
    hipLaunchKernelGGL((k_stencil<T, MD, RY, HM, false>), grid, dim3(threads), lds, ctx->stream, psi, hpsi, ld,
                       vloc_of<T>(ctx), make_coef<T>(ctx), n1, n2, n3, sh.ty, sh.zlen, ctx->dead.template as<T>());
  return RSG_OK;
}

template <typename T, int MD, int RY>
int launch_stencil_hm(rsg_ctx *ctx, const T *psi, T *hpsi, int64_t ld, int nb, const HpsiShape &sh,
                      int accumulate) {
  const int threads = round64(ctx->n[0] * (sh.ty / RY));
  const int need = (2 * MD * ctx->n[0] + threads - 1) / threads;
  if (need <= 1) return launch_stencil_t<T, MD, RY, 1>(ctx, psi, hpsi, ld, nb, sh, accumulate);
  if (need <= 2) return launch_stencil_t<T, MD, RY, 2>(ctx, psi, hpsi, ld, nb, sh, accumulate);
  return launch_stencil_t<T, MD, RY, 4>(ctx, psi, hpsi, ld, nb, sh, accumulate);
}

template <typename T, int MD>
int launch_stencil_md(rsg_ctx *ctx, const T *psi, T *hpsi, int64_t ld, int nb, const HpsiShape &sh,
                      int accumulate) {
  switch (sh.ry) {
    case 1: return launch_stencil_hm<T, MD, 1>(ctx, psi, hpsi, ld, nb, sh, accumulate);
    case 2: return launch_stencil_hm<T, MD, 2>(ctx, psi, hpsi, ld, nb, sh, accumulate);
    default: return ctx->fail(RSG_ERR_UNSUPPORTED, "hpsi: ry must be 1 or 2");
  }
}

template <typename T>
int launch_stencil(rsg_ctx *ctx, const T *psi, T *hpsi, int64_t ld, int nb, const HpsiShape &sh,
                   int accumulate) {
  if constexpr (std::is_same<T, float>::value) {
    if (sh.rx == 2) {
      // pair loads and stores are 8-byte vectors: every band column must
      // start 8-byte aligned, else the one-point kernel takes the call
      const bool aligned = ld % 2 == 0 && reinterpret_cast<uintptr_t>(psi) % 8 == 0 &&
                           reinterpret_cast<uintptr_t>(hpsi) % 8 == 0;
      if (aligned) {
        if (sh.ry == 1) return launch_stencil2_hm<1>(ctx, psi, hpsi, ld, nb, sh, accumulate);
        if (sh.ry == 2) return launch_stencil2_hm<2>(ctx, psi, hpsi, ld, nb, sh, accumulate);
        return ctx->fail(RSG_ERR_UNSUPPORTED, "hpsi: ry must be 1 or 2");
      }
      HpsiShape one = sh;
      one.rx = 1;
      one.ty = pick_ty(ctx->n[0], ctx->n[1], ctx->md, one.ry, sizeof(T));
      if (one.ty == 0) {
        one.ry = 1;
        one.ty = pick_ty(ctx->n[0], ctx->n[1], ctx->md, 1, sizeof(T));
      }
      return launch_stencil<T>(ctx, psi, hpsi, ld, nb, one, accumulate);
    }
  }
  switch (ctx->md) {
#define RSG_MD_CASE(M) \
  case M: return launch_stencil_md<T, M>(ctx, psi, hpsi, ld, nb, sh, accumulate);
    RSG_MD_CASE(1)
    RSG_MD_CASE(2)
    RSG_MD_CASE(3)
    RSG_MD_CASE(4)
    RSG_MD_CASE(5)
    RSG_MD_CASE(6)
    RSG_MD_CASE(7)
    RSG_MD_CASE(8)
#undef RSG_MD_CASE
    default: return ctx->fail(RSG_ERR_UNSUPPORTED, "hpsi: md outside 1..8");
  }
}

// Complete a launch shape: fields left at zero are derived from the grid,
// the precision and the number of bands of the call.
HpsiShape resolve_shape(const rsg_ctx *ctx, const HpsiShape &in, size_t s, int nb, bool chunked) {
  HpsiShape sh = in;
  const int n1 = ctx->n[0], n2 = ctx->n[1], n3 = ctx->n[2];
  if (sh.ry < 1 || sh.ry > 2) sh.ry = 2;
  if (sh.rx == 2) {  // pair kernel: FP32 only; else the one-point kernel
    if (s != sizeof(float)) sh.rx = 1;
    else if (!pair_fits(n1, n2, ctx->md, sh.ry, sh.ty, s)) {
      sh.ty = pick_ty_pair(n1, n2, ctx->md, sh.ry, s);
      if (sh.ty == 0) sh.rx = 1;
    }
  } else {
    sh.rx = 1;
  }
  if (sh.rx == 1 && (sh.ty <= 0 || sh.ty % sh.ry || round64(n1 * (sh.ty / sh.ry)) > 512 ||
      2 * ctx->md * n1 > 4 * round64(n1 * (sh.ty / sh.ry)) ||
      (2 * size_t(sh.ty + 2 * ctx->md) * n1 + 64) * s > kMaxDynLds)) {
    sh.ty = pick_ty(n1, n2, ctx->md, sh.ry, s);
    if (sh.ty == 0) {
      sh.ry = 1;
      sh.ty = pick_ty(n1, n2, ctx->md, 1, s);
    }
  }
  if (sh.chunk <= 0) {
    // psi + H psi of one chunk within ~128 MB of Infinity Cache
    const size_t per_band = size_t(ctx->N) * s;
    sh.chunk = chunked ? int(std::max<size_t>(1, (size_t(64) << 20) / per_band)) : nb;
  }
  sh.chunk = std::max(1, std::min(sh.chunk, nb));
  if (sh.zlen <= 0) {
    const int tiles = (n2 + sh.ty - 1) / sh.ty;
    const int64_t want = 3 * 256;
    int zr = int(std::max<int64_t>(1, (want + int64_t(tiles) * sh.chunk - 1) / (int64_t(tiles) * sh.chunk)));
    zr = std::min(zr, std::max(1, n3 / 8));
    sh.zlen = (n3 + zr - 1) / zr;
  }
  sh.zlen = std::max(1, std::min(sh.zlen, n3));
  if (sh.bt != 4 && sh.bt != 8 && sh.bt != 16 && sh.bt != 20 && sh.bt != 32) sh.bt = 8;
  if (sh.order != 0 && sh.order != 1) sh.order = 0;
  if (sh.proj < 0 || sh.proj > 2) sh.proj = 0;
  if (sh.plds < 0 || sh.plds > 64) sh.plds = 0;
  if (sh.psplit != 1 && sh.psplit != 2 && sh.psplit != 4 && sh.psplit != 8) sh.psplit = 1;
  return sh;
}

// dynamic LDS of k_proj_rows<T>: the staged rows of 16 bands (odd stride)
// and the per-atom sums of the block
template <typename T> size_t row_tiles_lds(const rsg_ctx *ctx, const AnalyticDev<T> &an) {
  const int np = std::min(an.rt_ty, ctx->n[1]) * ctx->n[0];
  return (size_t(16) * size_t(np | 1) + size_t(an.rt_nslot_max) * 4 * 16) * sizeof(T);
}

template <typename T> bool row_tiles_fit(const rsg_ctx *ctx, const AnalyticDev<T> &an) {
  // the kernel's own static LDS (pass values and points of 4 waves) comes on top
  constexpr size_t kStatic = 4 * 4 * 33 * sizeof(T) + 4 * 32 * sizeof(int);
  return an.rt_ty > 0 && an.rt_nslot > 0 && an.rt_nslot_max <= 255 &&
         std::min(an.rt_ty, ctx->n[1]) * ctx->n[0] <= kRowTileMax && row_tiles_lds<T>(ctx, an) + kStatic <= kMaxDynLds;
}

// bands per wave = 16 NBT
template <typename T, int NBT>
void launch_apply_m(rsg_ctx *ctx, AnalyticDev<T> &an, T *hpsi, int64_t ld, int nbc, const T *cproj, bool acc) {
  dim3 grid(unsigned((an.ntile + 3) / 4), (nbc + 16 * NBT - 1) / (16 * NBT));
  if (acc)
    hipLaunchKernelGGL((k_apply_m<T, NBT, true>), grid, dim3(256), 0, ctx->stream, hpsi, ld, nbc, ctx->N, ctx->n[0],
                       ctx->n[1], an.ntile, an.tile_off.template as<int>(), an.tile_g.template as<int>(),
                       an.hits.template as<uint64_t>(), an.es.template as<T>(), an.ep.template as<T>(),
                       an.dd.template as<T>(), an.amp.template as<T>(), an.S, cproj);
  else
    hipLaunchKernelGGL((k_apply_m<T, NBT, false>), grid, dim3(256), 0, ctx->stream, hpsi, ld, nbc, ctx->N,
                       ctx->n[0], ctx->n[1], an.ntile, an.tile_off.template as<int>(), an.tile_g.template as<int>(),
                       an.hits.template as<uint64_t>(), an.es.template as<T>(), an.ep.template as<T>(),
                       an.dd.template as<T>(), an.amp.template as<T>(), an.S, cproj);
}

template <typename T, int BT>
void launch_scatter_t(rsg_ctx *ctx, NonlocalDev<T> &nl, T *hpsi, int64_t ld, int nbc, const T *uv) {
  dim3 grid(unsigned((ctx->N + 255) / 256), (nbc + BT - 1) / BT);
  hipLaunchKernelGGL((k_nl_scatter<T, BT>), grid, dim3(256), 0, ctx->stream, hpsi, ld, nbc, ctx->N,
                     nl.sl_off.template as<int>(), nl.sl_len.template as<int>(),
                     nl.sl_lma.template as<int>(), nl.sl_uv.template as<T>(), uv);
}

}  // namespace

// Local part only: hpsi (+)= (T + Vloc) psi with the stencil kernel and the
// launch shape hpsi uses.
template <typename T>
int launch_kinetic(rsg_ctx *ctx, const T *psi, T *hpsi, int64_t ld, int nb, int accumulate) {
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "op_kinetic: grid not set");
  if (nb < 0 || ld < ctx->N) return ctx->fail(RSG_ERR_ARG, "op_kinetic: bad nb or ld");
  if (nb == 0) return RSG_OK;
  if (nb > 65535) return ctx->fail(RSG_ERR_ARG, "op_kinetic: nb above 65535");
  if (!psi || !hpsi) return ctx->fail(RSG_ERR_ARG, "op_kinetic: null buffer");
  const HpsiShape sh = resolve_shape(ctx, shape_of<T>(ctx), sizeof(T), nb, false);
  int rc = launch_stencil<T>(ctx, psi, hpsi, ld, nb, sh, accumulate ? 1 : 0);
  if (rc) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ctx->fail(RSG_ERR_HIP, std::string("op_kinetic launch: ") + hipGetErrorString(e));
  return RSG_OK;
}

// The two projection buffers of a chunk of bands: `nred` reduced and `npart`
// partial values per band.  From the caller's scratch (the chunk shrinks
// until both fit; RSG_ERR_ARG when one band does not) or from the
// context's own buffers, grown on demand.
template <typename T>
int projection_buffers(rsg_ctx *ctx, const HpsiScratch *scr, size_t nred, size_t npart, HpsiShape &sh,
                       T **red, T **part) {
  if (scr) {
    const size_t per_band = (nred + npart) * sizeof(T) + 256;
    if (per_band > scr->bytes)
      return ctx->fail(RSG_ERR_ARG, "hpsi: the projections of one band need " + std::to_string(per_band) +
                                        " bytes of scratch, more than the " + std::to_string(scr->bytes) +
                                        " available");
    const int fit = int(std::min<size_t>(65535, (scr->bytes - 256) / ((nred + npart) * sizeof(T))));
    if (sh.chunk > fit) sh.chunk = fit;
    const uintptr_t base = reinterpret_cast<uintptr_t>(scr->p);
    *red = reinterpret_cast<T *>(base);
    *part = reinterpret_cast<T *>((base + nred * sh.chunk * sizeof(T) + 255) & ~uintptr_t(255));
    return RSG_OK;
  }
  if (ctx->cproj.ensure(nred * sh.chunk * sizeof(T)) != hipSuccess ||
      ctx->uvpart.ensure(npart * sh.chunk * sizeof(T)) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "hpsi: projection scratch allocation failed");
  *red = ctx->cproj.as<T>();
  *part = ctx->uvpart.as<T>();
  return RSG_OK;
}

template <typename T>
int launch_hpsi(rsg_ctx *ctx, const T *psi, T *hpsi, int64_t ld, int nb, const HpsiScratch *scr) {
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "hpsi: grid not set");
  if (nb < 0 || ld < ctx->N) return ctx->fail(RSG_ERR_ARG, "hpsi: bad nb or ld");
  if (nb == 0) return RSG_OK;
  if (nb > 65535) return ctx->fail(RSG_ERR_ARG, "hpsi: nb above 65535");
  if (!psi || !hpsi) return ctx->fail(RSG_ERR_ARG, "hpsi: null buffer");
  HpsiShape sh = resolve_shape(ctx, shape_of<T>(ctx), sizeof(T), nb, ctx->have_nl);
  if (!ctx->have_nl) {
    int rc = launch_stencil<T>(ctx, psi, hpsi, ld, nb, sh, 0);
    if (rc) return rc;
  } else if (ctx->nl_mode != 1 && ctx->have_ana) {
    AnalyticDev<T> &an = ana_of<T>(ctx);
    T *cp = nullptr, *part = nullptr;
    const bool rows = sh.proj == 1 && row_tiles_fit<T>(ctx, an);
    const int npc = rows ? 1 : sh.psplit;  // partial slots per segment
    const int npart = rows ? an.rt_nslot : an.nseg * npc;
    int rc0 = projection_buffers<T>(ctx, scr, size_t(an.ngroup) * 4, size_t(npart) * 4, sh, &cp, &part);
    if (rc0) return rc0;
    for (int c0 = 0; c0 < nb; c0 += sh.chunk) {
      const int nbc = std::min(sh.chunk, nb - c0);
      const T *pc = psi + int64_t(c0) * ld;
      T *hc = hpsi + int64_t(c0) * ld;
      if (rows) {
        const int nrt = an.rt_ny * an.rt_nz;
        hipLaunchKernelGGL((k_proj_rows<T>), dim3(unsigned(nrt * ((nbc + 15) / 16))), dim3(256),
                           row_tiles_lds<T>(ctx, an), ctx->stream, pc, ld, nbc, ctx->n[0], ctx->n[1], an.rt_ty,
                           an.rt_ny, nrt, an.rt_piece.template as<int4>(), an.rt_piece_off.template as<int>(),
                           an.rt_slot_off.template as<int>(), an.rt_slot_gid.template as<int>(),
                           an.pts.template as<uint32_t>(), an.es.template as<T>(), an.ep.template as<T>(),
                           an.dd.template as<T>(), an.amp.template as<T>(), an.S, part);
      } else if (sh.proj == 2) {  // 64 bands per wave: the segment's points and values once per 64 bands
        hipLaunchKernelGGL((k_proj_m<T, 4>), dim3(unsigned(an.nseg * npc * ((nbc + 63) / 64))), dim3(64),
                           size_t(sh.plds) << 10, ctx->stream,
                           pc, ld, nbc, ctx->n[0], ctx->n[1], an.nseg, npc, an.seg_order.template as<int>(),
                           an.seg_beg.template as<int>(), an.seg_end.template as<int>(),
                           an.seg_grp.template as<int>(), an.pts.template as<uint32_t>(), an.es.template as<T>(),
                           an.ep.template as<T>(), an.dd.template as<T>(), an.amp.template as<T>(), an.S, part);
      } else {
        hipLaunchKernelGGL((k_proj_m<T, 2>), dim3(unsigned(an.nseg * npc * ((nbc + 31) / 32))), dim3(64),
                           size_t(sh.plds) << 10, ctx->stream,
                           pc, ld, nbc, ctx->n[0], ctx->n[1], an.nseg, npc, an.seg_order.template as<int>(),
                           an.seg_beg.template as<int>(), an.seg_end.template as<int>(),
                           an.seg_grp.template as<int>(), an.pts.template as<uint32_t>(), an.es.template as<T>(),
                           an.ep.template as<T>(), an.dd.template as<T>(), an.amp.template as<T>(), an.S, part);
      }
      const int64_t nrb = int64_t(an.ngroup) * 4 * ((nbc + 63) / 64);  // blocks: (g, l) rows x 64 bands
      hipLaunchKernelGGL((k_proj_reduce<T>), dim3(unsigned(nrb)), dim3(256), 0, ctx->stream, part,
                         nbc, an.ngroup, (rows ? an.grp_slot_off : an.grp_seg_off).template as<int>(), npc,
                         an.scale.template as<T>(), cp);
      const bool nl_after = sh.order == 1;
      if (nl_after) {
        int rc = launch_stencil<T>(ctx, pc, hc, ld, nbc, sh, 0);
        if (rc) return rc;
      }
      // bands per apply wave: 16 NBT (bt 4, 8, 16, 20, 32 -> NBT 1, 2, 4, 5, 8)
      switch (sh.bt) {
        case 4: launch_apply_m<T, 1>(ctx, an, hc, ld, nbc, cp, nl_after); break;
        case 16: launch_apply_m<T, 4>(ctx, an, hc, ld, nbc, cp, nl_after); break;
        case 20: launch_apply_m<T, 5>(ctx, an, hc, ld, nbc, cp, nl_after); break;
        case 32: launch_apply_m<T, 8>(ctx, an, hc, ld, nbc, cp, nl_after); break;
        default: launch_apply_m<T, 2>(ctx, an, hc, ld, nbc, cp, nl_after); break;
      }
      if (!nl_after) {
        int rc = launch_stencil<T>(ctx, pc, hc, ld, nbc, sh, 1);
        if (rc) return rc;
      }
    }
  } else {
    if (ctx->nl_mode == 2) return ctx->fail(RSG_ERR_STATE, "hpsi: analytic projectors requested but not available");
    NonlocalDev<T> &nl = nl_of<T>(ctx);
    const int nz = ctx->tables.nzlma;
    constexpr int BB = 8;
    T *uv = nullptr, *part = nullptr;
    if (scr) {
      int rc0 = projection_buffers<T>(ctx, scr, size_t(nz), size_t(nl.nseg) * 4, sh, &uv, &part);
      if (rc0) return rc0;
    } else {
      if (ctx->uvunk.ensure(size_t(nz) * sh.chunk * sizeof(T)) != hipSuccess ||
          ctx->uvpart.ensure(size_t(nl.nseg) * 4 * sh.chunk * sizeof(T)) != hipSuccess)
        return ctx->fail(RSG_ERR_HIP, "hpsi: projection scratch allocation failed");
      uv = ctx->uvunk.as<T>();
      part = ctx->uvpart.as<T>();
    }
    for (int c0 = 0; c0 < nb; c0 += sh.chunk) {
      const int nbc = std::min(sh.chunk, nb - c0);
      const T *pc = psi + int64_t(c0) * ld;
      T *hc = hpsi + int64_t(c0) * ld;
      hipLaunchKernelGGL((k_nl_gather<T, BB>), dim3(nl.nseg, (nbc + BB - 1) / BB), dim3(kGatherThreads), 0,
                         ctx->stream, pc, ld, nbc, nl.seg_beg.template as<int>(),
                         nl.seg_end.template as<int>(), nl.grp_pts.template as<int>(),
                         nl.grp_coef.template as<T>(), part);
      const int nred = nz * nbc;
      hipLaunchKernelGGL((k_nl_reduce<T>), dim3((nred + 255) / 256), dim3(256), 0, ctx->stream, part, nbc, nz,
                         nl.lma_grp.template as<int>(), nl.grp_lma0.template as<int>(),
                         nl.grp_seg_off.template as<int>(), nl.scale.template as<T>(), uv);
      switch (sh.bt) {
        case 4: launch_scatter_t<T, 4>(ctx, nl, hc, ld, nbc, uv); break;
        case 16: launch_scatter_t<T, 16>(ctx, nl, hc, ld, nbc, uv); break;
        default: launch_scatter_t<T, 8>(ctx, nl, hc, ld, nbc, uv); break;
      }
      int rc = launch_stencil<T>(ctx, pc, hc, ld, nbc, sh, 1);
      if (rc) return rc;
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ctx->fail(RSG_ERR_HIP, std::string("hpsi launch: ") + hipGetErrorString(e));
  return RSG_OK;
}

#if RSG_HPSI_TYPE_IS_DOUBLE
int effective_proj(rsg_ctx *ctx, int precision) {
  if (!ctx->have_ana || ctx->nl_mode == 1) return 0;
  const int p = (precision == 64 ? ctx->shape64 : ctx->shape32).proj;
  if (p != 1) return p;
  if (precision == 64) return row_tiles_fit<double>(ctx, ctx->ana64) ? 1 : 0;
  return row_tiles_fit<float>(ctx, ctx->ana32) ? 1 : 0;
}

int resolved_shape(rsg_ctx *ctx, int precision, int nb, HpsiShape *out) {
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "hpsi shape: grid not set");
  if (precision != 64 && precision != 32) return ctx->fail(RSG_ERR_ARG, "hpsi shape: precision must be 64 or 32");
  if (nb < 1) return ctx->fail(RSG_ERR_ARG, "hpsi shape: nb < 1");
  const size_t s = precision == 64 ? 8 : 4;
  *out = resolve_shape(ctx, precision == 64 ? ctx->shape64 : ctx->shape32, s, nb, ctx->have_nl);
  return RSG_OK;
}

namespace {

template <typename T>
int time_shape(rsg_ctx *ctx, const T *psi, T *hpsi, int nb, const HpsiShape &sh, int reps, float *ms) {
  HpsiShape saved = shape_of<T>(ctx);
  shape_of<T>(ctx) = sh;
  int rc = launch_hpsi<T>(ctx, psi, hpsi, ctx->N, nb);  // warm-up
  hipEvent_t a = nullptr, b = nullptr;
  if (!rc && (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)) rc = RSG_ERR_HIP;
  // the faster of two timed rounds: one round's average still moves by about
  // ten per cent from run to run
  float best = 0.f;
  for (int round = 0; round < 2 && !rc; ++round) {
    if (hipEventRecord(a, ctx->stream) != hipSuccess) rc = RSG_ERR_HIP;
    for (int i = 0; i < reps && !rc; ++i) rc = launch_hpsi<T>(ctx, psi, hpsi, ctx->N, nb);
    if (!rc && (hipEventRecord(b, ctx->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess)) rc = RSG_ERR_HIP;
    float t = 0.f;
    if (!rc && hipEventElapsedTime(&t, a, b) != hipSuccess) rc = RSG_ERR_HIP;
    if (!rc && (round == 0 || t < best)) best = t;
  }
  if (!rc) *ms = best / reps;
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  shape_of<T>(ctx) = saved;
  return rc;
}

// Coordinate search over the launch shape on seeded scratch bands; keeps
// the fastest shape in the context.
template <typename T> int autotune_t(rsg_ctx *ctx, int nb, int reps, float *best_ms) {
  DeviceOpGuard guard;
  const size_t bytes = size_t(ctx->N) * nb * sizeof(T);
  T *psi = nullptr, *hpsi = nullptr;
  if (hipMalloc(&psi, bytes) != hipSuccess) return ctx->fail(RSG_ERR_HIP, "autotune: allocation failed");
  if (hipMalloc(&hpsi, bytes) != hipSuccess) {
    (void)hipFree(psi);
    return ctx->fail(RSG_ERR_HIP, "autotune: allocation failed");
  }
  // deterministic non-trivial content: timing does not depend on values
  (void)hipMemsetAsync(psi, 0x3f, bytes, ctx->stream);
  const size_t s = sizeof(T);
  HpsiShape best = resolve_shape(ctx, HpsiShape(), s, nb, ctx->have_nl);
  float tbest = 0.f;
  int rc = time_shape<T>(ctx, psi, hpsi, nb, best, reps, &tbest);
  auto consider = [&](HpsiShape cand) {
    if (rc) return;
    cand = resolve_shape(ctx, cand, s, nb, ctx->have_nl);
    float t = 0.f;
    int r = time_shape<T>(ctx, psi, hpsi, nb, cand, reps, &t);
    if (r == RSG_ERR_UNSUPPORTED) return;  // shape does not fit this grid
    if (r) {
      rc = r;
      return;
    }
    if (t < tbest) {
      tbest = t;
      best = cand;
    }
  };
  for (int ry = 1; ry <= 2; ++ry) {
    HpsiShape c = best;
    c.ry = ry;
    c.ty = 0;
    c.zlen = 0;
    consider(c);
  }
  if (std::is_same<T, float>::value) {
    for (int ry = 1; ry <= 2; ++ry) {
      HpsiShape c = best;
      c.rx = 2;
      c.ry = ry;
      c.ty = 0;
      c.zlen = 0;
      consider(c);
    }
  }
  if (best.rx == 2) {  // neighbouring divisors of n2
    const int n2 = ctx->n[1];
    int lo = 0, hi = 0;
    for (int t = best.ty - best.ry; t > 0 && !lo; t -= best.ry)
      if (pair_fits(ctx->n[0], n2, ctx->md, best.ry, t, s)) lo = t;
    for (int t = best.ty + best.ry; t <= n2 && !hi; t += best.ry)
      if (pair_fits(ctx->n[0], n2, ctx->md, best.ry, t, s)) hi = t;
    for (int t : {lo, hi}) {
      HpsiShape c = best;
      c.ty = t;
      c.zlen = 0;
      if (t > 0) consider(c);
    }
  } else {
    for (int f : {-2, -1, 1, 2}) {
      HpsiShape c = best;
      c.ty = best.ty + f * best.ry;
      c.zlen = 0;
      if (c.ty > 0) consider(c);
    }
  }
  if (ctx->have_nl) {
    for (int f : {2, 4}) {
      HpsiShape c = best;
      c.chunk = std::min(nb, best.chunk * f);
      c.zlen = 0;
      consider(c);
    }
    HpsiShape ch = best;
    ch.chunk = std::max(1, best.chunk / 2);
    ch.zlen = 0;
    consider(ch);
    for (int bt : {4, 8, 16, 20, 32}) {
      HpsiShape c = best;
      c.bt = bt;
      consider(c);
    }
    if (ctx->have_ana && ctx->nl_mode != 1) {
      HpsiShape c = best;
      c.order = 1 - best.order;
      consider(c);
      // the row tiles of a description built on the device come on first use
      if (ensure_row_tiles(ctx) == RSG_OK) {
        HpsiShape pj = best;
        pj.proj = best.proj == 1 ? 0 : 1;
        consider(pj);
      }
      {  // segment projections with 64 bands per wave
        HpsiShape pj = best;
        pj.proj = best.proj == 2 ? 0 : 2;
        consider(pj);
      }
      if (best.proj != 1)  // fewer segment waves in flight per CU
        for (int kib : {4, 8, 16, 32}) {
          HpsiShape pl = best;
          pl.plds = kib;
          consider(pl);
        }
      if (best.proj != 1)  // segments in pieces: chords shared by two spheres read close in time
        for (int np : {2, 4, 8}) {
          HpsiShape ps = best;
          ps.psplit = np;
          consider(ps);
        }
      for (int f : {4, 8}) {  // smaller chunks keep psi and H psi in the Infinity Cache
        HpsiShape d = best;
        d.chunk = std::max(1, nb / f);
        d.zlen = 0;
        consider(d);
      }
    }
  }
  for (int zf : {2, 4}) {
    HpsiShape c = best;
    c.zlen = std::max(1, best.zlen / zf);
    consider(c);
  }
  {
    HpsiShape c = best;
    c.zlen = ctx->n[2];
    consider(c);
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(psi);
  (void)hipFree(hpsi);
  if (rc) return rc;
  shape_of<T>(ctx) = best;
  if (best_ms) *best_ms = tbest;
  return RSG_OK;
}

}  // namespace

int autotune_hpsi(rsg_ctx *ctx, int precision, int nb, int reps, float *best_ms) {
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "autotune: grid not set");
  if (nb < 1 || reps < 1) return ctx->fail(RSG_ERR_ARG, "autotune: nb and reps must be >= 1");
  if (precision == 64) return autotune_t<double>(ctx, nb, reps, best_ms);
  if (precision == 32) return autotune_t<float>(ctx, nb, reps, best_ms);
  return ctx->fail(RSG_ERR_ARG, "autotune: precision must be 64 or 32");
}

#endif  // RSG_HPSI_TYPE_IS_DOUBLE

template int launch_hpsi<RSG_HPSI_TYPE>(rsg_ctx *, const RSG_HPSI_TYPE *, RSG_HPSI_TYPE *, int64_t, int,
                                        const HpsiScratch *);
template int launch_kinetic<RSG_HPSI_TYPE>(rsg_ctx *, const RSG_HPSI_TYPE *, RSG_HPSI_TYPE *, int64_t, int, int);

}  // namespace rsg
