// capi.cpp -- the extern "C" entry points declared in include/rsdft_gpu.h.
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <condition_variable>

#include "context.hpp"
#include "lpc.hpp"
#include "solver.hpp"

using rsg::DevBuf;

namespace {

// Device scratch of rsg_hpsi_f64_noalloc: the partial and reduced nonlocal
// projections of a chunk of bands (16 MiB; the chunk shrinks to fit).  Every
// GPU the library runs on holds its own copy until the process ends, and
// every call on that GPU shares it.  Host code never passes the array's
// name to a kernel (on the host it is only a stand-in for the symbol): it
// uses the device address resolved once per device at context creation.
constexpr size_t kNoallocScratchBytes = size_t(16) << 20;
__device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// With `device` current: resolve (once) the address of this device's copy.
int resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return RSG_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return RSG_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RSG_ERR_HIP;
  g_scratch_ptr[device] = p;
  return RSG_OK;
}

void *scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

// Process-wide gate between graph captures and the calls that may
// invalidate them in other threads: any number of guarded calls at once, or
// one capture.  A waiting capture holds back new guarded calls, so that a
// steady stream of them cannot starve it.  t_gate_depth counts this
// thread's nested guards (and a capture in progress), so that only the
// outermost one takes the gate.
class CaptureGateLock {
 public:
  void lock_shared() {
    std::unique_lock<std::mutex> l(m_);
    cv_.wait(l, [&] { return !writer_ && writers_waiting_ == 0; });
    ++readers_;
  }
  void unlock_shared() {
    std::lock_guard<std::mutex> l(m_);
    if (--readers_ == 0) cv_.notify_all();
  }
  void lock() {
    std::unique_lock<std::mutex> l(m_);
    ++writers_waiting_;
    cv_.wait(l, [&] { return !writer_ && readers_ == 0; });
    --writers_waiting_;
    writer_ = true;
  }
  void unlock() {
    std::lock_guard<std::mutex> l(m_);
    writer_ = false;
    cv_.notify_all();
  }

 private:
  std::mutex m_;
  std::condition_variable cv_;
  int readers_ = 0, writers_waiting_ = 0;
  bool writer_ = false;
};
CaptureGateLock g_capture_gate;
thread_local int t_gate_depth = 0;

}  // namespace

namespace rsg {

DeviceOpGuard::DeviceOpGuard() {
  if (t_gate_depth++ == 0) {
    g_capture_gate.lock_shared();
    locked_ = true;
  }
}

DeviceOpGuard::~DeviceOpGuard() {
  --t_gate_depth;
  if (locked_) g_capture_gate.unlock_shared();
}

CaptureGate::CaptureGate() {
  if (t_gate_depth != 0) return;  // this thread holds a guard: capturing would deadlock
  g_capture_gate.lock();
  ++t_gate_depth;
  usable_ = true;
}

CaptureGate::~CaptureGate() {
  if (!usable_) return;
  --t_gate_depth;
  g_capture_gate.unlock();
}

}  // namespace rsg

extern "C" {

int rsg_create(rsg_ctx **out, int device) {
  if (!out) return RSG_ERR_ARG;
  *out = nullptr;
  rsg::DeviceOpGuard guard;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RSG_ERR_HIP;
  if (device < 0 || device >= ndev) return RSG_ERR_ARG;
  rsg_ctx *ctx = new (std::nothrow) rsg_ctx();
  if (!ctx) return RSG_ERR_HIP;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
    delete ctx;
    return RSG_ERR_HIP;
  }
  if (ctx->dead.ensure(64 * sizeof(double)) != hipSuccess || resolve_scratch(device) != RSG_OK) {
    ctx->dead.release();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RSG_ERR_HIP;
  }
  if (rocblas_create_handle(&ctx->blas) != rocblas_status_success ||
      rocblas_set_stream(ctx->blas, ctx->stream) != rocblas_status_success) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RSG_ERR_HIP;
  }
  *out = ctx;
  return RSG_OK;
}

int rsg_destroy(rsg_ctx *ctx) {
  if (!ctx) return RSG_OK;
  rsg::DeviceOpGuard guard;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  DevBuf *bufs[] = {&ctx->vloc64, &ctx->vloc32, &ctx->uvunk, &ctx->unk, &ctx->work,
                    &ctx->hsub, &ctx->esp_dev, &ctx->info, &ctx->cgbuf, &ctx->scal,
                    &ctx->hostio, &ctx->dead, &ctx->cgcoef};
  for (DevBuf *b : bufs) b->release();
  ctx->uvpart.release();
  ctx->cproj.release();
  ctx->setup_tmp.release();
  ctx->setup_scan.release();
  ctx->ana64.release_all();
  ctx->ana32.release_all();
  ctx->nl64.release_all();
  ctx->nl32.release_all();
  rsg::solver_release(ctx);
  rsg::scf_release(ctx);
  if (ctx->blas) rocblas_destroy_handle(ctx->blas);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RSG_OK;
}

const char *rsg_last_error(const rsg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rsg_synchronize(rsg_ctx *ctx) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

void *rsg_stream(rsg_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int rsg_set_grid(rsg_ctx *ctx, int n1, int n2, int n3, const double *aa, int md) {
  if (!ctx || !aa) return RSG_ERR_ARG;
  if (n1 < 1 || n2 < 1 || n3 < 1 || md < 1 || md > 8)
    return ctx->fail(RSG_ERR_ARG, "set_grid: grid sizes must be >= 1 and md in 1..8");
  const int64_t N = int64_t(n1) * n2 * n3;
  if (N > (int64_t(1) << 31) - 1) return ctx->fail(RSG_ERR_ARG, "set_grid: grid above 2^31 points");
  RSG_HIP(hipSetDevice(ctx->device));
  double len[3];
  for (int i = 0; i < 3; ++i)
    len[i] = std::sqrt(aa[3 * i] * aa[3 * i] + aa[3 * i + 1] * aa[3 * i + 1] + aa[3 * i + 2] * aa[3 * i + 2]);
  // orthorhombic only: the lattice vectors must be mutually orthogonal
  for (int i = 0; i < 3; ++i)
    for (int k = i + 1; k < 3; ++k) {
      double d = aa[3 * i] * aa[3 * k] + aa[3 * i + 1] * aa[3 * k + 1] + aa[3 * i + 2] * aa[3 * k + 2];
      if (std::fabs(d) > 1e-12 * len[i] * len[k])
        return ctx->fail(RSG_ERR_UNSUPPORTED, "set_grid: non-orthogonal cells are not supported");
    }
  const double det = aa[0] * (aa[4] * aa[8] - aa[5] * aa[7]) - aa[1] * (aa[3] * aa[8] - aa[5] * aa[6]) +
                     aa[2] * (aa[3] * aa[7] - aa[4] * aa[6]);
  if (!(std::fabs(det) > 0)) return ctx->fail(RSG_ERR_ARG, "set_grid: singular lattice");
  // the SCF fields and FFT plan survive a new cell on the same grid dimensions
  const bool same_dims = ctx->have_grid && ctx->n[0] == n1 && ctx->n[1] == n2 && ctx->n[2] == n3;
  ctx->n[0] = n1;
  ctx->n[1] = n2;
  ctx->n[2] = n3;
  ctx->N = N;
  ctx->md = md;
  std::memcpy(ctx->aa, aa, sizeof(ctx->aa));
  const int ng[3] = {n1, n2, n3};
  for (int i = 0; i < 3; ++i) ctx->hgrid[i] = len[i] / ng[i];
  ctx->dv = std::fabs(det) / double(N);
  for (int i = 0; i < 6; ++i) ctx->ggg[i] = i < 3 ? 1.0 : 0.0;  // orthorhombic metric
  double lap[2 * 8 + 1];
  rsg::fd_lap_closed_form(md, lap);
  ctx->coef_lap.assign(size_t(3) * md, 0.0);
  double csum = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double c = -0.5 * ctx->ggg[a] / (ctx->hgrid[a] * ctx->hgrid[a]);
    csum += c;
    for (int m = 1; m <= md; ++m) ctx->coef_lap[size_t(a) * md + m - 1] = lap[md + m] * c;
  }
  ctx->coef_lap0 = lap[md] * csum;
  if (ctx->vloc64.ensure(size_t(N) * 8) != hipSuccess || ctx->vloc32.ensure(size_t(N) * 4) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "set_grid: allocation failed");
  RSG_HIP(hipMemsetAsync(ctx->vloc64.p, 0, size_t(N) * 8, ctx->stream));
  RSG_HIP(hipMemsetAsync(ctx->vloc32.p, 0, size_t(N) * 4, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  if (same_dims)
    rsg::scf_reset(ctx);
  else
    rsg::scf_release(ctx);  // fields and FFT plan belong to the old grid
  ctx->have_grid = true;
  ctx->have_vloc = false;
  ctx->have_nl = false;
  ctx->have_ana = false;
  ctx->tables = rsg::HostTables();
  ctx->tables_state = 0;
  ctx->mb = 0;
  ctx->band32 = false;
  ctx->unk_fresh = false;
  return RSG_OK;
}

int rsg_get_grid(rsg_ctx *ctx, double *dv, double *hgrid, double *c0, double *cl) {
  if (!ctx) return RSG_ERR_ARG;
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "get_grid: grid not set");
  if (dv) *dv = ctx->dv;
  if (hgrid)
    for (int i = 0; i < 3; ++i) hgrid[i] = ctx->hgrid[i];
  if (c0) *c0 = ctx->coef_lap0;
  if (cl) std::memcpy(cl, ctx->coef_lap.data(), ctx->coef_lap.size() * sizeof(double));
  return RSG_OK;
}

int rsg_set_vloc(rsg_ctx *ctx, const double *vloc) {
  if (!ctx || !vloc) return RSG_ERR_ARG;
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "set_vloc: grid not set");
  RSG_HIP(hipSetDevice(ctx->device));
  std::vector<float> v32(vloc, vloc + ctx->N);
  RSG_HIP(hipMemcpyAsync(ctx->vloc64.p, vloc, size_t(ctx->N) * 8, hipMemcpyHostToDevice, ctx->stream));
  RSG_HIP(hipMemcpyAsync(ctx->vloc32.p, v32.data(), size_t(ctx->N) * 4, hipMemcpyHostToDevice, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_vloc = true;
  return RSG_OK;
}

int rsg_set_projectors(rsg_ctx *ctx, int nzlma, int maxmjj, const int *mjj, const int *jjp,
                       const double *uvk, const int *iuv, const int *amap) {
  if (!ctx) return RSG_ERR_ARG;
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "set_projectors: grid not set");
  if (nzlma < 0 || maxmjj < 0 || (nzlma > 0 && (!mjj || !jjp || !uvk || !iuv)))
    return ctx->fail(RSG_ERR_ARG, "set_projectors: bad sizes or null arrays");
  rsg::HostTables t;
  t.nzlma = nzlma;
  t.maxmjj = maxmjj;
  t.mjj.assign(mjj, mjj + nzlma);
  t.iuv.assign(iuv, iuv + nzlma);
  t.jjp.assign(jjp, jjp + size_t(nzlma) * maxmjj);
  t.uvk.assign(uvk, uvk + size_t(nzlma) * maxmjj);
  if (amap) t.amap.assign(amap, amap + nzlma);
  for (int l = 0; l < nzlma; ++l) {
    if (t.mjj[l] < 0 || t.mjj[l] > maxmjj) return ctx->fail(RSG_ERR_ARG, "set_projectors: MJJ out of range");
    for (int j = 0; j < t.mjj[l]; ++j) {
      int p = t.jjp[size_t(l) * maxmjj + j];
      if (p < 0 || p >= ctx->N) return ctx->fail(RSG_ERR_ARG, "set_projectors: JJP outside the grid");
      if (j > 0 && p <= t.jjp[size_t(l) * maxmjj + j - 1])
        return ctx->fail(RSG_ERR_ARG, "set_projectors: JJP not strictly ascending within an lma");
    }
  }
  RSG_HIP(hipSetDevice(ctx->device));
  ctx->tables = std::move(t);
  ctx->tables_state = 0;
  ctx->have_nl = false;
  ctx->have_ana = false;  // caller tables carry no analytic form
  return rsg::upload_nonlocal(ctx);
}

namespace {

// Reference-layout tables of the deferred GTH input on the host.
int ensure_host_tables(rsg_ctx *ctx) {
  if (ctx->tables_state != 1) return RSG_OK;
  const rsg_ctx::GthInput &g = ctx->gth;
  rsg::HostTables t;
  int rc = rsg::build_gth_tables(ctx, g.natom, g.frac.data(), g.kind.data(), g.nkind, g.norb.data(), g.lo.data(),
                                 g.no.data(), g.rps.data(), g.hnl.data(), t);
  if (rc) return ctx->fail(rc, "build_gth: invalid species parameters");
  ctx->tables = std::move(t);
  ctx->tables_state = 2;
  return RSG_OK;
}

// ... and their device layouts (the table form of the nonlocal operator).
int ensure_device_tables(rsg_ctx *ctx) {
  if (ctx->tables_state == 0) return RSG_OK;
  int rc = ensure_host_tables(ctx);
  if (rc) return rc;
  if ((rc = rsg::upload_nonlocal(ctx))) return rc;
  ctx->tables_state = 0;
  return RSG_OK;
}

}  // namespace

int rsg_build_gth_projectors(rsg_ctx *ctx, int natom, const double *frac, const int *kind, int nkind,
                             const int *norb, const int *lo, const int *no, const double *rps,
                             const double *hnl) {
  if (!ctx) return RSG_ERR_ARG;
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "build_gth: grid not set");
  if (natom < 0 || nkind < 1 || (natom > 0 && (!frac || !kind)) || !norb || !lo || !no || !rps || !hnl)
    return ctx->fail(RSG_ERR_ARG, "build_gth: bad arguments");
  // the checks of build_gth_tables, made before either form is built
  for (int ik = 0; ik < nkind; ++ik) {
    if (norb[ik] < 0 || norb[ik] > 4) return ctx->fail(RSG_ERR_ARG, "build_gth: invalid species parameters");
    for (int io = 0; io < norb[ik]; ++io) {
      const int q = ik * 4 + io;
      if (lo[q] < 0 || lo[q] > 3 || no[q] < 1 || !(rps[q] > 0))
        return ctx->fail(RSG_ERR_ARG, "build_gth: invalid species parameters");
    }
  }
  for (int a = 0; a < natom; ++a)
    if (kind[a] < 0 || kind[a] >= nkind) return ctx->fail(RSG_ERR_ARG, "build_gth: invalid species parameters");
  RSG_HIP(hipSetDevice(ctx->device));
  rsg_ctx::GthInput &g = ctx->gth;
  g.natom = natom;
  g.nkind = nkind;
  g.frac.assign(frac, frac + 3 * size_t(natom));
  g.kind.assign(kind, kind + natom);
  g.norb.assign(norb, norb + nkind);
  g.lo.assign(lo, lo + 4 * size_t(nkind));
  g.no.assign(no, no + 4 * size_t(nkind));
  g.rps.assign(rps, rps + 4 * size_t(nkind));
  g.hnl.assign(hnl, hnl + 4 * size_t(nkind));
  ctx->tables = rsg::HostTables();
  ctx->have_nl = false;
  ctx->have_ana = false;
  ctx->tables_state = 1;
  int rc;
  if (ctx->ana_on_host) {
    rsg::HostAnalytic h;
    rc = rsg::build_gth_analytic(ctx, natom, frac, kind, nkind, norb, lo, no, rps, hnl, h);
    if (rc == RSG_ERR_UNSUPPORTED) return ensure_device_tables(ctx);  // the table form covers this set
    if (rc) return ctx->fail(rc, "build_gth: invalid atom kinds");
    if ((rc = rsg::upload_analytic(ctx, h))) return rc;
  } else {
    rc = rsg::build_gth_analytic_device(ctx, natom, frac, kind, nkind, norb, lo, no, rps, hnl);
    if (rc == RSG_ERR_UNSUPPORTED) {
      ctx->have_ana = false;
      return ensure_device_tables(ctx);  // the table form covers this set
    }
    if (rc == RSG_ERR_ARG) return ctx->fail(rc, "build_gth: invalid atom kinds");
    if (rc) return rc;
  }
  ctx->have_nl = ctx->ana64.ngroup > 0;
  // the table form is built when a caller asks for it (tables, mode 1)
  if (ctx->nl_mode == 1) return ensure_device_tables(ctx);
  return RSG_OK;
}

int rsg_set_nonlocal_mode(rsg_ctx *ctx, int mode) {
  if (!ctx) return RSG_ERR_ARG;
  if (mode < 0 || mode > 2) return ctx->fail(RSG_ERR_ARG, "set_nonlocal_mode: mode must be 0, 1 or 2");
  if (mode == 1) {
    RSG_HIP(hipSetDevice(ctx->device));
    int rc = ensure_device_tables(ctx);
    if (rc) return rc;
  }
  ctx->nl_mode = mode;
  return RSG_OK;
}

int rsg_get_nonlocal_mode(rsg_ctx *ctx, int *mode, int *analytic_available) {
  if (!ctx) return RSG_ERR_ARG;
  if (mode) *mode = ctx->nl_mode;
  if (analytic_available) *analytic_available = ctx->have_ana ? 1 : 0;
  return RSG_OK;
}

int rsg_get_projector_sizes(rsg_ctx *ctx, int *nzlma, int *maxmjj, int64_t *nent) {
  if (!ctx) return RSG_ERR_ARG;
  int rc = ensure_host_tables(ctx);
  if (rc) return rc;
  if (nzlma) *nzlma = ctx->tables.nzlma;
  if (maxmjj) *maxmjj = ctx->tables.maxmjj;
  if (nent) *nent = ctx->tables.nentries();
  return RSG_OK;
}

int rsg_get_projectors(rsg_ctx *ctx, int *mjj, int *jjp, double *uvk, int *iuv) {
  if (!ctx) return RSG_ERR_ARG;
  int rc = ensure_host_tables(ctx);
  if (rc) return rc;
  const rsg::HostTables &t = ctx->tables;
  if (mjj) std::memcpy(mjj, t.mjj.data(), t.mjj.size() * sizeof(int));
  if (jjp) std::memcpy(jjp, t.jjp.data(), t.jjp.size() * sizeof(int));
  if (uvk) std::memcpy(uvk, t.uvk.data(), t.uvk.size() * sizeof(double));
  if (iuv) std::memcpy(iuv, t.iuv.data(), t.iuv.size() * sizeof(int));
  return RSG_OK;
}

int rsg_hpsi_f64(rsg_ctx *ctx, const double *psi, double *hpsi, int64_t ld, int nb) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::launch_hpsi<double>(ctx, psi, hpsi, ld, nb);
}

int rsg_hpsi_f64_noalloc(rsg_ctx *ctx, const double *psi, double *hpsi, int64_t ld, int nb) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  void *p = scratch_of(ctx->device);
  if (!p) return ctx->fail(RSG_ERR_STATE, "hpsi_f64_noalloc: no scratch address for this device");
  const rsg::HpsiScratch s{p, kNoallocScratchBytes};
  return rsg::launch_hpsi<double>(ctx, psi, hpsi, ld, nb, &s);
}

int rsg_hpsi_f32(rsg_ctx *ctx, const float *psi, float *hpsi, int64_t ld, int nb) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::launch_hpsi<float>(ctx, psi, hpsi, ld, nb);
}

int rsg_op_kinetic_f64(rsg_ctx *ctx, const double *psi, double *hpsi, int64_t ld, int nb, int accumulate) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::launch_kinetic<double>(ctx, psi, hpsi, ld, nb, accumulate);
}

int rsg_op_kinetic_f32(rsg_ctx *ctx, const float *psi, float *hpsi, int64_t ld, int nb, int accumulate) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::launch_kinetic<float>(ctx, psi, hpsi, ld, nb, accumulate);
}

int rsg_hpsi_f64_host(rsg_ctx *ctx, const double *psi, double *hpsi, int64_t ld, int nb) {
  if (!ctx || !psi || !hpsi) return RSG_ERR_ARG;
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "hpsi: grid not set");
  if (nb <= 0) return nb == 0 ? RSG_OK : ctx->fail(RSG_ERR_ARG, "hpsi: nb < 0");
  if (ld < ctx->N) return ctx->fail(RSG_ERR_ARG, "hpsi: ld < N");
  RSG_HIP(hipSetDevice(ctx->device));
  const size_t blk = size_t(ctx->N) * nb * sizeof(double);
  if (ctx->hostio.ensure(2 * blk) != hipSuccess) return ctx->fail(RSG_ERR_HIP, "hpsi: staging allocation failed");
  double *dp = ctx->hostio.as<double>(), *dh = dp + size_t(ctx->N) * nb;
  RSG_HIP(hipMemcpy2DAsync(dp, ctx->N * 8, psi, ld * 8, ctx->N * 8, nb, hipMemcpyHostToDevice, ctx->stream));
  int rc = rsg::launch_hpsi<double>(ctx, dp, dh, ctx->N, nb);
  if (rc) return rc;
  RSG_HIP(hipMemcpy2DAsync(hpsi, ld * 8, dh, ctx->N * 8, ctx->N * 8, nb, hipMemcpyDeviceToHost, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

int rsg_set_hpsi_shape(rsg_ctx *ctx, int precision, int ry, int ty, int zlen, int chunk, int bt) {
  if (!ctx) return RSG_ERR_ARG;
  if (precision != 64 && precision != 32) return ctx->fail(RSG_ERR_ARG, "set_hpsi_shape: precision must be 64 or 32");
  if (ry < 0 || ry > 2 || ty < 0 || zlen < 0 || chunk < 0 ||
      (bt != 0 && bt != 4 && bt != 8 && bt != 16 && bt != 20 && bt != 32))
    return ctx->fail(RSG_ERR_ARG, "set_hpsi_shape: ry in 0..2, ty/zlen/chunk >= 0, bt in {0,4,8,16,20,32}");
  rsg::HpsiShape &sh = precision == 64 ? ctx->shape64 : ctx->shape32;
  sh.ry = ry ? ry : 2;
  sh.ty = ty;
  sh.zlen = zlen;
  sh.chunk = chunk;
  sh.bt = bt ? bt : 8;
  return RSG_OK;
}

int rsg_set_hpsi_order(rsg_ctx *ctx, int precision, int order) {
  if (!ctx) return RSG_ERR_ARG;
  if ((precision != 64 && precision != 32) || (order != 0 && order != 1))
    return ctx->fail(RSG_ERR_ARG, "set_hpsi_order: precision 64/32 and order 0/1");
  (precision == 64 ? ctx->shape64 : ctx->shape32).order = order;
  return RSG_OK;
}

int rsg_get_hpsi_order(rsg_ctx *ctx, int precision, int *order) {
  if (!ctx || !order) return RSG_ERR_ARG;
  if (precision != 64 && precision != 32) return ctx->fail(RSG_ERR_ARG, "get_hpsi_order: precision 64/32");
  *order = (precision == 64 ? ctx->shape64 : ctx->shape32).order;
  return RSG_OK;
}

int rsg_set_hpsi_proj(rsg_ctx *ctx, int precision, int proj) {
  if (!ctx) return RSG_ERR_ARG;
  if ((precision != 64 && precision != 32) || proj < 0 || proj > 2)
    return ctx->fail(RSG_ERR_ARG, "set_hpsi_proj: precision 64/32 and proj 0/1/2");
  if (proj == 1) {
    RSG_HIP(hipSetDevice(ctx->device));
    int rc = rsg::ensure_row_tiles(ctx);
    if (rc) return rc;
  }
  (precision == 64 ? ctx->shape64 : ctx->shape32).proj = proj;
  return RSG_OK;
}

int rsg_set_projector_setup(rsg_ctx *ctx, int on_host) {
  if (!ctx) return RSG_ERR_ARG;
  ctx->ana_on_host = on_host != 0;
  return RSG_OK;
}

int rsg_get_analytic_array(rsg_ctx *ctx, int which, void *out, int64_t *count) {
  if (!ctx || !count) return RSG_ERR_ARG;
  if (!ctx->have_ana) return ctx->fail(RSG_ERR_STATE, "get_analytic_array: no analytic projector description");
  const rsg::AnalyticDev<double> &d = ctx->ana64;
  const int ng = d.ngroup;
  int64_t E = 0, ntg = 0;
  rsg::DeviceOpGuard guard;  // synchronous copies: kept apart from other threads' captures
  RSG_HIP(hipSetDevice(ctx->device));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  if (d.nseg) RSG_HIP(hipMemcpy(&E, static_cast<const int *>(d.seg_end.p) + d.nseg - 1, 4, hipMemcpyDeviceToHost));
  {
    int t = 0;
    RSG_HIP(hipMemcpy(&t, static_cast<const int *>(d.tile_off.p) + d.ntile, 4, hipMemcpyDeviceToHost));
    ntg = t;
  }
  struct Item {
    const rsg::DevBuf *b;
    int64_t n;
    size_t sz;
  };
  const Item items[] = {
      {&d.pts, E, 4},          {&d.seg_beg, d.nseg, 4},     {&d.seg_end, d.nseg, 4},
      {&d.seg_grp, d.nseg, 4}, {&d.seg_order, d.nseg, 4},   {&d.grp_seg_off, ng + 1, 4},
      {&d.hits, ctx->N, 8},    {&d.tile_off, d.ntile + 1, 4}, {&d.tile_g, ntg, 4},
      {&d.es, int64_t(ng) * d.S, 8}, {&d.ep, int64_t(ng) * d.S, 8}, {&d.dd, int64_t(ng) * d.S, 8},
      {&d.amp, int64_t(ng) * 4, 8},  {&d.scale, int64_t(ng) * 4, 8}};
  if (which < 0 || which >= int(sizeof(items) / sizeof(items[0])))
    return ctx->fail(RSG_ERR_ARG, "get_analytic_array: which outside 0..13");
  const Item &it = items[which];
  *count = it.n;
  if (out && it.n) RSG_HIP(hipMemcpy(out, it.b->p, size_t(it.n) * it.sz, hipMemcpyDeviceToHost));
  return RSG_OK;
}

int rsg_set_hpsi_proj_lds(rsg_ctx *ctx, int precision, int kib) {
  if (!ctx) return RSG_ERR_ARG;
  if ((precision != 64 && precision != 32) || kib < 0 || kib > 64)
    return ctx->fail(RSG_ERR_ARG, "set_hpsi_proj_lds: precision 64/32 and 0 <= KiB <= 64");
  (precision == 64 ? ctx->shape64 : ctx->shape32).plds = kib;
  return RSG_OK;
}

int rsg_get_hpsi_proj_lds(rsg_ctx *ctx, int precision, int *kib) {
  if (!ctx || !kib) return RSG_ERR_ARG;
  if (precision != 64 && precision != 32) return ctx->fail(RSG_ERR_ARG, "get_hpsi_proj_lds: precision 64/32");
  *kib = (precision == 64 ? ctx->shape64 : ctx->shape32).plds;
  return RSG_OK;
}

int rsg_set_hpsi_proj_split(rsg_ctx *ctx, int precision, int pieces) {
  if (!ctx) return RSG_ERR_ARG;
  if ((precision != 64 && precision != 32) || (pieces != 1 && pieces != 2 && pieces != 4 && pieces != 8))
    return ctx->fail(RSG_ERR_ARG, "set_hpsi_proj_split: precision 64/32 and pieces 1, 2, 4 or 8");
  (precision == 64 ? ctx->shape64 : ctx->shape32).psplit = pieces;
  return RSG_OK;
}

int rsg_get_hpsi_proj_split(rsg_ctx *ctx, int precision, int *pieces) {
  if (!ctx || !pieces) return RSG_ERR_ARG;
  if (precision != 64 && precision != 32) return ctx->fail(RSG_ERR_ARG, "get_hpsi_proj_split: precision 64/32");
  *pieces = (precision == 64 ? ctx->shape64 : ctx->shape32).psplit;
  return RSG_OK;
}

int rsg_get_hpsi_proj(rsg_ctx *ctx, int precision, int *proj) {
  if (!ctx || !proj) return RSG_ERR_ARG;
  if (precision != 64 && precision != 32) return ctx->fail(RSG_ERR_ARG, "get_hpsi_proj: precision 64/32");
  *proj = rsg::effective_proj(ctx, precision);
  return RSG_OK;
}

int rsg_set_hpsi_xpoints(rsg_ctx *ctx, int precision, int rx) {
  if (!ctx) return RSG_ERR_ARG;
  if ((precision != 64 && precision != 32) || (rx != 1 && rx != 2))
    return ctx->fail(RSG_ERR_ARG, "set_hpsi_xpoints: precision 64/32 and rx 1/2");
  (precision == 64 ? ctx->shape64 : ctx->shape32).rx = rx;
  return RSG_OK;
}

int rsg_get_hpsi_xpoints(rsg_ctx *ctx, int precision, int nb, int *rx) {
  if (!ctx || !rx) return RSG_ERR_ARG;
  rsg::HpsiShape sh;
  int rc = rsg::resolved_shape(ctx, precision, nb, &sh);
  if (rc) return rc;
  *rx = sh.rx;
  return RSG_OK;
}

int rsg_get_hpsi_shape(rsg_ctx *ctx, int precision, int nb, int *ry, int *ty, int *zlen, int *chunk, int *bt) {
  if (!ctx) return RSG_ERR_ARG;
  rsg::HpsiShape sh;
  int rc = rsg::resolved_shape(ctx, precision, nb, &sh);
  if (rc) return rc;
  if (ry) *ry = sh.ry;
  if (ty) *ty = sh.ty;
  if (zlen) *zlen = sh.zlen;
  if (chunk) *chunk = sh.chunk;
  if (bt) *bt = sh.bt;
  return RSG_OK;
}

int rsg_autotune_hpsi(rsg_ctx *ctx, int precision, int nb, int reps, float *best_ms) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::autotune_hpsi(ctx, precision, nb, reps, best_ms);
}

int rsg_alloc_bands(rsg_ctx *ctx, int mb) {
  if (!ctx) return RSG_ERR_ARG;
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "alloc_bands: grid not set");
  if (mb < 1) return ctx->fail(RSG_ERR_ARG, "alloc_bands: mb < 1");
  RSG_HIP(hipSetDevice(ctx->device));
  const size_t bytes = size_t(ctx->N) * mb * sizeof(double);
  if (ctx->unk.ensure(bytes) != hipSuccess || ctx->work.ensure(bytes) != hipSuccess)
    return ctx->fail(RSG_ERR_HIP, "alloc_bands: out of device memory");
  RSG_HIP(hipMemsetAsync(ctx->unk.p, 0, bytes, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  ctx->mb = mb;
  ctx->band32 = false;  // a new store starts in FP64
  return RSG_OK;
}

int rsg_set_band_precision(rsg_ctx *ctx, int bits) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::set_band_precision(ctx, bits);
}

int rsg_get_band_precision(rsg_ctx *ctx, int *bits) {
  if (!ctx || !bits) return RSG_ERR_ARG;
  *bits = ctx->band32 ? 32 : 64;
  return RSG_OK;
}

int rsg_upload_bands(rsg_ctx *ctx, const double *u, int64_t ld, int n0, int nb) {
  if (!ctx || !u) return RSG_ERR_ARG;
  if (ctx->mb == 0) return ctx->fail(RSG_ERR_STATE, "upload_bands: no band store");
  if (ctx->band32) return ctx->fail(RSG_ERR_STATE, "upload_bands: the bands are in the FP32 store");
  if (n0 < 0 || nb < 0 || n0 + nb > ctx->mb || ld < ctx->N)
    return ctx->fail(RSG_ERR_ARG, "upload_bands: band range or ld out of bounds");
  RSG_HIP(hipSetDevice(ctx->device));
  double *d = ctx->unk.as<double>() + size_t(n0) * ctx->N;
  RSG_HIP(hipMemcpy2DAsync(d, ctx->N * 8, u, ld * 8, ctx->N * 8, nb, hipMemcpyHostToDevice, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

int rsg_random_bands(rsg_ctx *ctx, int n0, int nb, uint64_t seed) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::random_bands(ctx, n0, nb, seed);
}

int rsg_download_bands(rsg_ctx *ctx, double *u, int64_t ld, int n0, int nb) {
  if (!ctx || !u) return RSG_ERR_ARG;
  if (ctx->mb == 0) return ctx->fail(RSG_ERR_STATE, "download_bands: no band store");
  if (ctx->band32) return ctx->fail(RSG_ERR_STATE, "download_bands: the bands are in the FP32 store");
  if (n0 < 0 || nb < 0 || n0 + nb > ctx->mb || ld < ctx->N)
    return ctx->fail(RSG_ERR_ARG, "download_bands: band range or ld out of bounds");
  RSG_HIP(hipSetDevice(ctx->device));
  const double *d = ctx->unk.as<double>() + size_t(n0) * ctx->N;
  RSG_HIP(hipMemcpy2DAsync(u, ld * 8, d, ctx->N * 8, ctx->N * 8, nb, hipMemcpyDeviceToHost, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return RSG_OK;
}

int rsg_bands_device_ptr(rsg_ctx *ctx, double **d_u, int64_t *ld) {
  if (!ctx || !d_u) return RSG_ERR_ARG;
  if (ctx->mb == 0) return ctx->fail(RSG_ERR_STATE, "bands_device_ptr: no band store");
  if (ctx->band32) return ctx->fail(RSG_ERR_STATE, "bands_device_ptr: the bands are in the FP32 store");
  *d_u = ctx->unk.as<double>();
  if (ld) *ld = ctx->N;
  return RSG_OK;
}

int rsg_gram_schmidt(rsg_ctx *ctx, int n0, int n1) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::gram_schmidt(ctx, n0, n1);
}

int rsg_subspace_diag(rsg_ctx *ctx, double *esp) {
  if (!ctx || !esp) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::subspace_diag(ctx, esp);
}

int rsg_conjugate_gradient(rsg_ctx *ctx, int ncg, int pc_kind, int mloop, double *esp, double *res) {
  if (!ctx || !esp || !res) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::conjugate_gradient(ctx, ncg, pc_kind, mloop, esp, res);
}

int rsg_set_cg_graphs(rsg_ctx *ctx, int enable) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  rsg::DeviceOpGuard guard;
  for (auto &g : ctx->cg_graph_cache)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  ctx->cg_graph_cache.clear();
  ctx->cg_graphs = enable != 0;
  return RSG_OK;
}

int rsg_cg_graph_launches(rsg_ctx *ctx, int64_t *launches) {
  if (!ctx || !launches) return RSG_ERR_ARG;
  *launches = ctx->cg_graph_launches;
  return RSG_OK;
}

int rsg_cg_graph_failures(rsg_ctx *ctx, int64_t *failures) {
  if (!ctx || !failures) return RSG_ERR_ARG;
  *failures = ctx->cg_graph_failures;
  return RSG_OK;
}

int rsg_set_cg_swgs(rsg_ctx *ctx, int swgs) {
  if (!ctx) return RSG_ERR_ARG;
  if (swgs < 0 || swgs > 2) return ctx->fail(RSG_ERR_ARG, "set_cg_swgs: SWGS must be 0, 1 or 2");
  ctx->cg_swgs = swgs;
  return RSG_OK;
}

int rsg_set_precond_storage(rsg_ctx *ctx, int bits) {
  if (!ctx) return RSG_ERR_ARG;
  if (bits != 16 && bits != 32) return ctx->fail(RSG_ERR_ARG, "set_precond_storage: bits must be 16 or 32");
  ctx->pc_bf16 = bits == 16;
  return RSG_OK;
}

int rsg_autotune_gram(rsg_ctx *ctx, int nb, int reps, float *best_ms, int64_t *chunk) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::autotune_gram(ctx, nb, reps, best_ms, chunk);
}

int rsg_precond_hprsdft_f64(rsg_ctx *ctx, int mloop, const double *g, double *pg, int64_t ld, int nb) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::precond_hprsdft_dev(ctx, mloop, g, pg, ld, nb);
}

int rsg_precond_hprsdft_f32(rsg_ctx *ctx, int mloop, const float *g, float *pg, int64_t ld, int nb) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::precond_hprsdft_dev(ctx, mloop, g, pg, ld, nb);
}

int rsg_conjugate_gradient_f32(rsg_ctx *ctx, int ncg, int pc_kind, int mloop, double *esp, double *res) {
  if (!ctx || !esp || !res) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::conjugate_gradient_f32(ctx, ncg, pc_kind, mloop, esp, res);
}

int rsg_esp_calc(rsg_ctx *ctx, double *esp) {
  if (!ctx || !esp) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::esp_calc(ctx, esp);
}

int rsg_set_local_gth(rsg_ctx *ctx, int natom, const double *frac, const int *kind, int nkind,
                      const double *zps, const double *rloc, const double *cloc) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::scf_set_local_gth(ctx, natom, frac, kind, nkind, zps, rloc, cloc);
}

int rsg_get_scf_constants(rsg_ctx *ctx, double *eewald, double *nelectron) {
  if (!ctx) return RSG_ERR_ARG;
  return rsg::scf_constants(ctx, eewald, nelectron);
}

int rsg_get_field(rsg_ctx *ctx, int field, double *out) {
  if (!ctx || !out) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::scf_field(ctx, field, out, true);
}

int rsg_set_field(rsg_ctx *ctx, int field, const double *in) {
  if (!ctx || !in) return RSG_ERR_ARG;
  if (field != RSG_FIELD_RHO && field != RSG_FIELD_RHO_OUT && field != RSG_FIELD_VLOC)
    return ctx->fail(RSG_ERR_ARG, "set_field: only RHO, RHO_OUT and VLOC can be set");
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::scf_field(ctx, field, const_cast<double *>(in), false);
}

int rsg_init_density(rsg_ctx *ctx, double width) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::scf_init_density(ctx, width);
}

int rsg_normalize_density(rsg_ctx *ctx) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::scf_normalize(ctx);
}

int rsg_calc_density(rsg_ctx *ctx, const double *occ) {
  if (!ctx || !occ) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::scf_calc_density(ctx, occ);
}

int rsg_hartree_xc(rsg_ctx *ctx, int field, double *e_hartree, double *e_xc) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::scf_hartree_xc(ctx, field, e_hartree, e_xc);
}

int rsg_update_vloc(rsg_ctx *ctx) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::scf_update_vloc(ctx);
}

int rsg_density_integrals(rsg_ctx *ctx, int field, double *e_loc, double *e_ion) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::scf_density_integrals(ctx, field, e_loc, e_ion);
}

int rsg_mix_init(rsg_ctx *ctx, int mmix, double beta) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::scf_mix_init(ctx, mmix, beta, 10);
}

int rsg_mix_init_imix(rsg_ctx *ctx, int mmix, double beta, int imix) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::scf_mix_init(ctx, mmix, beta, imix);
}

int rsg_mix_pulay(rsg_ctx *ctx) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::scf_mix_pulay(ctx);
}

int rsg_lpc_param_count(void) { return rsg::lpc_param_count(); }

int rsg_lpc_set_params(rsg_ctx *ctx, const float *params, int n) {
  if (!ctx || !params) return RSG_ERR_ARG;
  if (n != rsg::lpc_param_count()) return ctx->fail(RSG_ERR_ARG, "lpc_set_params: wrong parameter count");
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::lpc_set_params(ctx, params);
}

int rsg_lpc_get_params(rsg_ctx *ctx, float *params, int n) {
  if (!ctx || !params) return RSG_ERR_ARG;
  if (n != rsg::lpc_param_count()) return ctx->fail(RSG_ERR_ARG, "lpc_get_params: wrong parameter count");
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::lpc_get_params(ctx, params);
}

int rsg_lpc_set_potential(rsg_ctx *ctx, const double *vloc) {
  if (!ctx) return RSG_ERR_ARG;
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "lpc_set_potential: grid not set");
  RSG_HIP(hipSetDevice(ctx->device));
  if (vloc) return rsg::lpc_set_potential(ctx, vloc);
  std::vector<double> v(static_cast<size_t>(ctx->N));
  RSG_HIP(hipMemcpyAsync(v.data(), ctx->vloc64.p, size_t(ctx->N) * 8, hipMemcpyDeviceToHost, ctx->stream));
  RSG_HIP(hipStreamSynchronize(ctx->stream));
  return rsg::lpc_set_potential(ctx, v.data());
}

int rsg_lpc_apply_f64(rsg_ctx *ctx, const double *g, double *pg, int64_t ld, int nb) {
  if (!ctx || !g || !pg) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::lpc_apply(ctx, g, pg, ld, nb, 64);
}

int rsg_lpc_apply_f32(rsg_ctx *ctx, const float *g, float *pg, int64_t ld, int nb) {
  if (!ctx || !g || !pg) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::lpc_apply(ctx, g, pg, ld, nb, 32);
}

int rsg_lpc_target_f32(rsg_ctx *ctx, const float *g, float *t, int nb, int steps, double shift, double vweight) {
  if (!ctx || !g || !t) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::lpc_target(ctx, g, t, nb, steps, shift, vweight);
}

int rsg_lpc_train_step(rsg_ctx *ctx, const float *g, const float *t, int nsamp, float kappa, float lr, double *loss,
                       float *grads) {
  if (!ctx || !g || !t || nsamp < 1 || !(kappa > 0.f)) return RSG_ERR_ARG;
  if (!ctx->have_grid) return ctx->fail(RSG_ERR_STATE, "lpc_train_step: grid not set");
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::lpc_train_step(ctx, g, t, nsamp, kappa, lr, loss, grads);
}

int rsg_lpc_reset_optimizer(rsg_ctx *ctx) {
  if (!ctx) return RSG_ERR_ARG;
  return rsg::lpc_reset_optimizer(ctx);
}

int rsg_band_residuals_f32(rsg_ctx *ctx, float *d_g) {
  if (!ctx || !d_g) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::band_residuals_f32(ctx, d_g);
}

int rsg_lpc_fallbacks(rsg_ctx *ctx, int64_t *count) {
  if (!ctx || !count) return RSG_ERR_ARG;
  *count = ctx->lpc.fallbacks;
  return RSG_OK;
}

int rsg_lpc_set_reject_limit(rsg_ctx *ctx, int limit) {
  if (!ctx) return RSG_ERR_ARG;
  if (limit < 0) return ctx->fail(RSG_ERR_ARG, "lpc_set_reject_limit: limit < 0");
  ctx->lpc.reject_limit = limit;
  ctx->lpc.reject_streak = 0;
  return RSG_OK;
}

int rsg_lpc_reject_state(rsg_ctx *ctx, int *streak, int64_t *skipped_calls) {
  if (!ctx || !streak || !skipped_calls) return RSG_ERR_ARG;
  *streak = ctx->lpc.reject_streak;
  *skipped_calls = ctx->lpc.skipped_calls;
  return RSG_OK;
}

int rsg_lpc_set_hybrid_scale(rsg_ctx *ctx, double scale) {
  if (!ctx) return RSG_ERR_ARG;
  if (!std::isfinite(scale)) return ctx->fail(RSG_ERR_ARG, "lpc_set_hybrid_scale: scale not finite");
  ctx->lpc.hybrid_scale = scale;
  return RSG_OK;
}

int rsg_axpby_f32(rsg_ctx *ctx, float *d_y, const float *d_x, float a, float b, int64_t n) {
  if (!ctx) return RSG_ERR_ARG;
  RSG_HIP(hipSetDevice(ctx->device));
  return rsg::axpby_f32(ctx, d_y, d_x, a, b, n);
}

}  // extern "C"
