// context.hpp -- internal state of an rsg_ctx (one GPU, one rank).
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <array>
#include <string>
#include <vector>

#include "rsdft_gpu.h"

namespace rsg {

struct ScfState;

// A HIP graph capture on one context's stream must not overlap device
// allocation, frees, stream creation or library calls that may allocate
// (rocBLAS, rocSOLVER, hipFFT) made by other threads of the process: the
// runtime can invalidate the capture in progress.  Such calls hold a
// DeviceOpGuard (shared, re-entrant per thread); a capture holds the
// CaptureGate (exclusive).  Inside a capture the calling thread's own
// guards are no-ops.
struct DeviceOpGuard {
  DeviceOpGuard();
  ~DeviceOpGuard();
  DeviceOpGuard(const DeviceOpGuard &) = delete;
  DeviceOpGuard &operator=(const DeviceOpGuard &) = delete;

 private:
  bool locked_ = false;
};

struct CaptureGate {
  CaptureGate();  // blocks until no other thread holds a DeviceOpGuard
  ~CaptureGate();
  bool usable() const { return usable_; }  // false when this thread already holds a guard
  CaptureGate(const CaptureGate &) = delete;
  CaptureGate &operator=(const CaptureGate &) = delete;

 private:
  bool usable_ = false;
};

// Device buffer owned by the context (hipMalloc/hipFree).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t need);
  void release();
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Nonlocal projector operator in the two layouts the kernels stream.
//  gather: groups of <= 4 projectors that share one point list (all
//          projectors of one atom), coef[pt*4 + l] (zero-padded); each
//          group's points are split into segments of bounded length so a
//          launch has enough workgroups;
//  scatter: per-grid-point entries (lma, coefficient) in sliced ELL: slice c
//          covers points 64c..64c+63, padded to its longest row, stored
//          slot-major so a wavefront reads each slot coalesced.
template <typename T> struct NonlocalDev {
  int ngroup = 0, nseg = 0;
  DevBuf grp_lma0;     // int[ngroup]    first lma of the group
  DevBuf grp_nl;       // int[ngroup]    projectors in the group (<= 4)
  DevBuf grp_seg_off;  // int[ngroup+1]  segment range of each group
  DevBuf seg_beg;      // int[nseg]      first point of the segment
  DevBuf seg_end;      // int[nseg]      one past its last point
  DevBuf seg_grp;      // int[nseg]      owning group
  DevBuf grp_pts;      // int[npts]      grid index per group point
  DevBuf grp_coef;     // T[npts*4]
  DevBuf lma_grp;      // int[nzlma]     group of each projector
  DevBuf scale;        // T[nzlma]       iuV * dV
  DevBuf sl_off;       // int[nslice]    first slot of each 64-point slice
  DevBuf sl_len;       // int[nslice]    slots per point in the slice
  DevBuf sl_lma;       // int[slots]     projector index (0 in padding)
  DevBuf sl_uv;        // T[slots]       coefficient (0 in padding)
  void release_all() {
    for (DevBuf *b : {&grp_lma0, &grp_nl, &grp_seg_off, &seg_beg, &seg_end, &seg_grp, &grp_pts,
                      &grp_coef, &lma_grp, &scale, &sl_off, &sl_len, &sl_lma, &sl_uv})
      b->release();
  }
};

// GTH projectors of the form (l = 0, n = 1) and (l = 1, n = 1) evaluated on
// the fly instead of streamed from tables.  For an atom at R and a grid
// point at displacement d (the image nearest the atom; unique because every
// cell length exceeds twice the largest cutoff), the four projector values
// are
//   s:       amp[0] * es_x(d_x) es_y(d_y) es_z(d_z)
//   p_x,y,z: amp[1] * ep_x ep_y ep_z * d_{x,y,z}
// with e?_axis = exp(-d_axis^2 / (2 r_l^2)) tabulated per atom and axis
// over the grid coordinates.  This equals v_l(r) Y_lm(r^) of the tables up
// to rounding and up to an orthogonal change of basis inside l = 1, which
// leaves sum_m |p_m><p_m| unchanged.  Which points belong to which sphere
// is decided on the host with the table builder's own arithmetic and
// stored as flags.
//
// Work lists: a segment is one atom's sphere points on one grid plane (at
// most kPlaneSegMax of them, in grid order); segments are numbered atom by
// atom (grp_seg_off) and launched plane by plane (seg_order), so the atoms
// that share a plane read it while it is cached.  A tile is 16 consecutive
// grid points; tile_g lists the atoms whose spheres meet it.
template <typename T> struct AnalyticDev {
  int ngroup = 0, nseg = 0, S = 0, ntile = 0;
  DevBuf es, ep, dd;    // T[ngroup * S]: axis x at [0, n1), y at [n1, n1+n2), z after
  DevBuf amp, scale;    // T[ngroup * 4]: amplitudes (s, p, p, p), iuV * dV
  DevBuf pts;           // uint32 per group point: x | y << 10 | z << 20 | s << 30 | p << 31
  DevBuf seg_beg, seg_end, seg_grp;  // int[nseg]
  DevBuf seg_order;     // int[nseg]      launch order (by plane, then atom)
  DevBuf grp_seg_off;   // int[ngroup + 1]
  DevBuf hits;          // uint64 per grid point: 4 x uint16 (group + 1 | s << 14 | p << 15)
  DevBuf tile_off;      // int[ntile + 1]
  DevBuf tile_g;        // int[tile_off[ntile]]
  // row-tile projection: blocks of rt_ty grid rows x rt_zr planes; per block
  // its pieces (one atom's sphere points on one plane and in the block's
  // rows: int4 beg, end, z * 256 + local slot, group), and per local slot
  // the global slot of the partial projections (grouped by atom)
  int rt_ty = 0, rt_zr = 0, rt_ny = 0, rt_nz = 0, rt_nslot_max = 0, rt_nslot = 0;
  DevBuf rt_piece, rt_piece_off, rt_slot_off, rt_slot_gid, grp_slot_off;
  void release_all() {
    for (DevBuf *b : {&es, &ep, &dd, &amp, &scale, &pts, &seg_beg, &seg_end, &seg_grp, &seg_order, &grp_seg_off,
                      &hits, &tile_off, &tile_g, &rt_piece, &rt_piece_off, &rt_slot_off, &rt_slot_gid,
                      &grp_slot_off})
      b->release();
  }
};

struct HostAnalytic {
  int ngroup = 0, S = 0;
  std::vector<double> es, ep, dd, amp, scale;
  std::vector<uint32_t> pts;
  std::vector<int> seg_beg, seg_end, seg_grp, seg_order, grp_seg_off{0};
  std::vector<uint64_t> hits;
  std::vector<int> tile_off, tile_g;
  int rt_ty = 0, rt_zr = 0, rt_ny = 0, rt_nz = 0, rt_nslot_max = 0;
  std::vector<int> rt_piece, rt_piece_off, rt_slot_off, rt_slot_gid, grp_slot_off;
};

// Per-atom parameters of the analytic form, shared by the host and the
// device builders: atom position, nearest grid point, candidate box
// half-widths, cutoffs, Gaussian exponents, amplitudes and signs.
struct AnalyticGroup {
  double R[3];
  int ic[3], mm[3];
  double rs = 0, rp = 0, c2s = 0, c2p = 0, amp_s = 0, amp_p = 0, sg_s = 1, sg_p = 1;
  int has_s = 0, has_p = 0;
};
struct AnalyticGroups {
  int S = 0;  // n1 + n2 + n3
  std::vector<AnalyticGroup> g;  // atoms with projectors, in atom order
};

// Launch shape of the hpsi kernels for one grid and precision.
struct HpsiShape {
  int ry = 2;     // rows per thread (y register blocking)
  int ty = 0;     // rows per workgroup tile
  int zlen = 0;   // z-planes per workgroup
  int chunk = 0;  // bands per chunk (gather -> scatter -> stencil)
  int bt = 8;     // bands per scatter thread
  int order = 0;  // analytic projectors: 0 = nonlocal writes, stencil adds; 1 = stencil writes, nonlocal adds
  int rx = 1;     // x-points per stencil thread: 1, or 2 (pair kernel: FP32, Md 6, even n1, ty | n2)
  int proj = 0;   // analytic projections: 0 = per sphere segment (k_proj_m), 1 = by row tiles (k_proj_rows)
  int plds = 0;   // KiB of LDS reserved per segment-projection wave (fewer waves in flight: a smaller L2 set)
  int psplit = 1; // segment projections: pieces per segment (1, 2, 4, 8), each its own wave and partial slot
};

// Learned preconditioner state: bf16 weight matrices for the kernels, the
// FP32 master parameters with Adam moments, the normalised potential and
// activation buffers.
struct LpcState {
  DevBuf wbf, bias, vnorm, x0, a1, a2, y, dyb, dz1, dz2, sval, gg, part, grad, master, mom1, mom2, lsum;
  DevBuf diag, tbuf;  // training-target solve
  bool have_params = false, have_v = false;
  int64_t vN = 0;
  int adam_t = 0;
  int64_t fallbacks = 0;  // CG blocks that fell back to hprsdft
  // CG calls in a row whose network output failed the descent check; at
  // reject_limit (> 0) the following calls run hprsdft without trying the
  // network (counted in skipped_calls) until the limit is set again
  int reject_streak = 0, reject_limit = 0;
  int64_t skipped_calls = 0;
  double hybrid_scale = 1.0;  // pc_kind 6: P g = hybrid_scale * P_hprsdft g + network(g)
  LpcState() = default;
  LpcState(const LpcState &) = delete;
  LpcState &operator=(const LpcState &) = delete;
  ~LpcState();
  void release();
};

// One captured CG band block: the key bytes (every size, pointer and launch
// parameter the block's kernels are built from) and the executable graph;
// seen counts eager runs of the key before its capture.
struct CgGraph {
  std::vector<unsigned char> key;
  hipGraphExec_t exec = nullptr;
  int seen = 0;
  bool failed = false;  // capture not possible here: stay eager
};

struct HostTables {
  int nzlma = 0, maxmjj = 0;
  std::vector<int> mjj, jjp, iuv, amap;  // jjp/uvk column-major (maxmjj,nzlma)
  std::vector<double> uvk;
  int64_t nentries() const {
    int64_t e = 0;
    for (int v : mjj) e += v;
    return e;
  }
};

}  // namespace rsg

struct rsg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  std::string err;

  // grid
  bool have_grid = false;
  int n[3] = {0, 0, 0};
  int64_t N = 0;
  int md = 0;
  double aa[9];
  double hgrid[3], dv = 0.0, ggg[6];
  double coef_lap0 = 0.0;
  std::vector<double> coef_lap;  // [axis*md + m-1]

  rsg::DevBuf vloc64, vloc32;
  bool have_vloc = false;

  rsg::HostTables tables;
  // GTH input of the last build_gth_projectors call.  With the analytic
  // form in use the reference-layout tables and their device layouts are
  // built on first use only: tables_state 1 = neither built, 2 = host
  // tables built, 0 = nothing deferred.
  struct GthInput {
    int natom = 0, nkind = 0;
    std::vector<double> frac, rps, hnl;
    std::vector<int> kind, norb, lo, no;
  } gth;
  int tables_state = 0;
  rsg::NonlocalDev<double> nl64;
  rsg::NonlocalDev<float> nl32;
  bool have_nl = false;
  rsg::AnalyticDev<double> ana64;
  rsg::AnalyticDev<float> ana32;
  bool have_ana = false;  // analytic projector description available
  bool ana_on_host = false;       // build the analytic description on the host (rsg_set_projector_setup)
  bool ana_rows_pending = false;  // row tiles of a device-built description not built yet
  rsg::DevBuf setup_tmp, setup_scan;  // device set-up scratch
  int nl_mode = 0;        // 0 auto, 1 tables, 2 analytic
  rsg::DevBuf cproj;      // analytic path: projections C[(g*4 + l) * chunk + b]
  rsg::DevBuf uvunk;    // nzlma x chunk projections uVunk, [lma][band]
  rsg::DevBuf uvpart;   // per-segment partial projections
  rsg::HpsiShape shape64, shape32;  // zero fields: chosen from the grid

  // wavefunction store and work arrays
  int mb = 0;
  rsg::DevBuf unk, work;           // N x mb each (work: rotation target)
  // FP32 band store of the mixed-precision phase (rsg_set_band_precision):
  // while band32 is set the bands live in unk32 (work32: rotation target
  // and H psi), and unk is stale until the store returns to FP64
  rsg::DevBuf unk32, work32, hsub32, grampart32;
  bool band32 = false;
  // with band32: unk holds the same bands in FP64 (Gram-Schmidt just wrote
  // them there), so esp_calc and the density read unk
  bool unk_fresh = false;
  rsg::DevBuf hsub, esp_dev, info; // mb x mb dense, mb, int
  rsg::DevBuf cgbuf;               // CG work vectors
  rsg::DevBuf scal;                // small per-band scalars
  rsg::DevBuf hostio;              // staging for the host hpsi entry
  rsg::DevBuf dotpart;             // multidot partial sums
  rsg::DevBuf grampart;            // per-chunk Gram matrices of the K-split U^T X products
  rsg::DevBuf pcbuf;               // work vectors of the stand-alone preconditioner call
  rsg::DevBuf cgcoef;              // cggs coefficients dV U^T Pg (mb x block)
  int cg_swgs = 0;                 // SWGS keyword: 0 none, 1/2 cggs projections in CG
  bool cg_graphs = false;          // replay CG band blocks from captured HIP graphs
  bool pc_bf16 = false;            // FP32 hprsdft calls keep r, p, the new direction in bfloat16
  int64_t cg_graph_launches = 0;   // band blocks run from a captured graph
  int64_t cg_graph_failures = 0;   // captures the runtime refused
  std::vector<rsg::CgGraph> cg_graph_cache;
  // K-split of the U^T X Gram products chosen by the autotuner, per
  // (N, nb): points per chunk, 0 = one GEMM over all of K
  std::vector<std::array<int64_t, 3>> gram_split;
  rsg::DevBuf dead;                // store target of padding lanes (64 doubles)
  rsg::ScfState *scf = nullptr;    // density/potential state of the SCF step
  rsg::LpcState lpc;               // learned CG preconditioner

  int fail(int code, const std::string &msg) {
    err = msg;
    return code;
  }
};

#define RSG_HIP(call)                                                      \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      return ctx->fail(RSG_ERR_HIP, std::string(#call ": ") +              \
                                        hipGetErrorString(e_));            \
  } while (0)

#define RSG_BLAS(call)                                                     \
  do {                                                                     \
    rocblas_status s_ = (call);                                            \
    if (s_ != rocblas_status_success)                                      \
      return ctx->fail(RSG_ERR_HIP, std::string(#call ": ") +              \
                                        rocblas_status_to_string(s_));     \
  } while (0)

// A rocBLAS / rocSOLVER call that may allocate its workspace, held apart
// from graph captures in other threads (DeviceOpGuard)
#define RSG_BLAS_GUARDED(call)                                             \
  do {                                                                     \
    rocblas_status s_;                                                     \
    {                                                                      \
      DeviceOpGuard guard_;                                                \
      s_ = (call);                                                         \
    }                                                                      \
    if (s_ != rocblas_status_success)                                      \
      return ctx->fail(RSG_ERR_HIP, std::string(#call ": ") +              \
                                        rocblas_status_to_string(s_));     \
  } while (0)

// tables.cpp
namespace rsg {
void fd_lap_closed_form(int md, double *lap);
int build_gth_tables(const rsg_ctx *ctx, int natom, const double *frac,
                     const int *kind, int nkind, const int *norb,
                     const int *lo, const int *no, const double *rps,
                     const double *hnl, HostTables &out);
int upload_nonlocal(rsg_ctx *ctx);
int build_gth_analytic(const rsg_ctx *ctx, int natom, const double *frac,
                       const int *kind, int nkind, const int *norb,
                       const int *lo, const int *no, const double *rps,
                       const double *hnl, HostAnalytic &out);
int upload_analytic(rsg_ctx *ctx, const HostAnalytic &h);
int analytic_groups(const rsg_ctx *ctx, int natom, const double *frac, const int *kind, int nkind, const int *norb,
                    const int *lo, const int *no, const double *rps, const double *hnl, AnalyticGroups &out);
void build_row_tiles(const int *ng, HostAnalytic &out);
// the row tiles of the projection for a description built on the device
// (from its point list and segments), on first use
int ensure_row_tiles(rsg_ctx *ctx);
}  // namespace rsg

// setup.hip
namespace rsg {
int build_gth_analytic_device(rsg_ctx *ctx, int natom, const double *frac, const int *kind, int nkind,
                              const int *norb, const int *lo, const int *no, const double *rps, const double *hnl);
}  // namespace rsg

// hpsi.hip
namespace rsg {
// Caller-provided device scratch for the projections of one hpsi call:
// with it, launch_hpsi allocates nothing and runs the bands in chunks small
// enough for the partial and reduced projections to fit in `bytes`.
struct HpsiScratch {
  void *p = nullptr;
  size_t bytes = 0;
};
template <typename T>
int launch_hpsi(rsg_ctx *ctx, const T *psi, T *hpsi, int64_t ld, int nb, const HpsiScratch *scr = nullptr);
template <typename T>
int launch_kinetic(rsg_ctx *ctx, const T *psi, T *hpsi, int64_t ld, int nb, int accumulate);
int resolved_shape(rsg_ctx *ctx, int precision, int nb, HpsiShape *out);
// the projection form an hpsi call uses (HpsiShape::proj where it fits)
int effective_proj(rsg_ctx *ctx, int precision);
int autotune_hpsi(rsg_ctx *ctx, int precision, int nb, int reps, float *best_ms);
}
