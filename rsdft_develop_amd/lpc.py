"""Learned CG preconditioner: parameters, training data and training loop.

The network and its kernels live in the library (csrc/lpc.hip); this module
drives them through the C ABI:

- `init_params` / `pack` / `unpack`: the flat FP32 parameter vector in the
  library's order W1[16][2][27] b1[16] W2[16][16][27] b2[16] W3[1][16][27]
  b3[1] (W[co][ci][tap], tap = d3*9 + d2*3 + d1);
- `make_dataset`: residuals g = -2 (H u - e u) of the plain method's own
  bands, taken from a few plain SCF iterations on perturbed cells, paired
  with targets t = A^-1 g, A = -1/2 lap_7 + vweight (Vloc - min Vloc) +
  shift, solved far past the mloop steps of the hprsdft smoother;
- `train`: Adam over mini-batches, all on the GPU.

No counterpart in the reference: it replaces precond_hprsdft
(cgpc_hprsdft_module.f90:133-361) when CG runs with pc_kind = 5.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

CH = 16
SHAPES = [("W1", (CH, 2, 27)), ("b1", (CH,)), ("W2", (CH, CH, 27)), ("b2", (CH,)), ("W3", (1, CH, 27)), ("b3", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def init_params(seed: int = 0) -> np.ndarray:
    """He-scaled conv weights, zero biases, a small last layer."""
    rng = np.random.default_rng(seed)
    d = {
        "W1": rng.standard_normal((CH, 2, 27)) * np.sqrt(2.0 / (2 * 27)),
        "b1": np.zeros(CH),
        "W2": rng.standard_normal((CH, CH, 27)) * np.sqrt(2.0 / (CH * 27)),
        "b2": np.zeros(CH),
        "W3": rng.standard_normal((1, CH, 27)) * 1e-2,
        "b3": np.zeros(1),
    }
    return pack(d)


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[k], dtype=np.float64).reshape(-1) for k, _ in SHAPES]).astype(np.float32)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for k, s in SHAPES:
        n = int(np.prod(s))
        out[k] = np.asarray(p[o:o + n], dtype=np.float64).reshape(s)
        o += n
    return out


def to_oracle(p: np.ndarray) -> dict:
    """Parameters in the (co, ci, 3, 3, 3) layout of the CPU restatement."""
    d = unpack(p)
    return {k: (v.reshape(v.shape[0], v.shape[1], 3, 3, 3) if k.startswith("W") else v) for k, v in d.items()}


@dataclasses.dataclass
class Dataset:
    g: object       # hip.DeviceArray (N, nsamp) FP32 residuals
    t: object       # hip.DeviceArray (N, nsamp) FP32 targets
    nsamp: int
    npoints: int
    cells: int
    seconds: float
    blocks: list    # (first sample, count, Vloc of that cell and iteration)
    h: object = None  # hip.DeviceArray (N, nsamp) FP32 hprsdft outputs P_h g (hybrid training), or None


def make_dataset(ctx_factory, cells, scf_iters: int = 2, target_steps: int = 40, shift: float = 0.5,
                 vweight: float = 1.0, max_samples: int | None = None, log=None, hybrid_mloop: int = 0) -> Dataset:
    """Residual/target pairs from plain SCF runs on `cells` (all on one grid
    shape).  ctx_factory() returns a fresh gpu.Context.  Each cell gives
    nband samples per SCF iteration.  hybrid_mloop > 0 also keeps the
    hprsdft output of every residual (mloop steps), for hybrid training."""
    from . import gpu, hip, scf

    N = cells[0].npoints
    per_cell = cells[0].nband * scf_iters
    total = per_cell * len(cells)
    if max_samples:
        total = min(total, max_samples)
    g_all = hip.DeviceArray((N, total), dtype=np.float32)
    t_all = hip.DeviceArray((N, total), dtype=np.float32)
    h_all = hip.DeviceArray((N, total), dtype=np.float32) if hybrid_mloop > 0 else None
    t0 = time.perf_counter()
    filled = 0
    blocks = []
    for ic, cell in enumerate(cells):
        if filled >= total:
            break
        assert cell.npoints == N
        ctx = ctx_factory()
        run = scf.SCF(cell, scf.ScfParams(ekbt=0.01), ctx=ctx)
        e_prev = None
        for it in range(scf_iters):
            rec = run.step(it + 1, e_prev)
            e_prev = rec.etot
            run.update_potential()
            nb = min(cell.nband, total - filled)
            if nb <= 0:
                break
            # residuals of the current bands (FP32) into a staging block, then
            # targets with the current potential
            stage = hip.DeviceArray((N, cell.nband), dtype=np.float32)
            ctx.band_residuals_device(stage.ptr)
            g_ptr = g_all.ptr + filled * N * 4
            t_ptr = t_all.ptr + filled * N * 4
            hip.copy_d2d(g_ptr, stage.ptr, nb * N * 4)
            ctx.lpc_target_device(g_ptr, t_ptr, nb, steps=target_steps, shift=shift, vweight=vweight)
            if h_all is not None:
                ctx.precond_device(g_ptr, h_all.ptr + filled * N * 4, N, nb, mloop=hybrid_mloop, dtype=np.float32)
            stage.free()
            blocks.append((filled, nb, ctx.get_field(gpu.FIELD_VLOC)))
            filled += nb
        ctx.close()
        if log:
            log(f"dataset: cell {ic + 1}/{len(cells)}, {filled} samples, {time.perf_counter() - t0:.1f} s")
    return Dataset(g_all, t_all, filled, N, len(cells), time.perf_counter() - t0, blocks, h_all)


def _probe_columns(data: Dataset, nprobe: int):
    """Host copies (FP64) of g, t and h (if kept) at nprobe evenly spaced samples."""
    import ctypes

    from . import hip

    N = data.npoints
    idx = np.linspace(0, data.nsamp - 1, min(nprobe, data.nsamp)).astype(int)
    out = []
    for i in idx:
        cols = []
        for arr in (data.g, data.t, data.h):
            if arr is None:
                cols.append(None)
                continue
            v = np.empty(N, dtype=np.float32)
            hip.runtime().hipMemcpy(v.ctypes.data, ctypes.c_void_p(arr.ptr + int(i) * N * 4), N * 4, 2)
            cols.append(v.astype(np.float64))
        out.append(cols)
    return out


def hybrid_targets(ctx, data: Dataset, nprobe: int = 16) -> tuple:
    """Turn the targets into corrections of the hprsdft output for hybrid
    training: kappa from the original targets, mu = <t, h> / <h, h> over
    probe samples (the least-squares factor on h = P_h g), then t <- t - mu h
    on the device.  The network then learns kappa t - kappa mu h, and CG's
    pc_kind 6 applies P g = (kappa mu) P_h g + net(g).  Returns (kappa, mu)."""
    if data.h is None:
        raise ValueError("dataset was made without hprsdft outputs (hybrid_mloop = 0)")
    kappa = target_scale(data)
    th = hh = 0.0
    for g, t, h in _probe_columns(data, nprobe):
        th += float(np.dot(t, h))
        hh += float(np.dot(h, h))
    mu = th / hh
    ctx.axpby_f32(data.t.ptr, data.h.ptr, 1.0, -mu, data.npoints * data.nsamp)
    ctx.synchronize()
    return kappa, mu


def train(ctx, data: Dataset, params: np.ndarray, epochs: int = 4, batch: int = 16, lr: float = 2e-3,
          seed: int = 0, log=None, kappa: float | None = None) -> tuple:
    """Adam over shuffled mini-batches; a batch never mixes cells, so the
    network's potential input is that cell's Vloc.  Returns the trained
    parameters and the per-epoch mean loss."""
    ctx.lpc_set_params(params)
    ctx.lpc_reset_optimizer()
    rng = np.random.default_rng(seed)
    N = data.npoints
    if kappa is None:
        kappa = target_scale(data)
    batches = [(ib, s) for ib, (first, count, _) in enumerate(data.blocks)
               for s in range(first, first + count - batch + 1, batch)]
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(batches))
        tot = 0.0
        t0 = time.perf_counter()
        current = -1
        for k in order:
            ib, s = batches[int(k)]
            if ib != current:
                ctx.lpc_set_potential(data.blocks[ib][2])
                current = ib
            off = s * N * 4
            loss, _ = ctx.lpc_train_step(data.g.ptr + off, data.t.ptr + off, batch, lr=lr, kappa=kappa)
            tot += loss
        hist.append(tot / max(len(batches), 1))
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.4e} ({time.perf_counter() - t0:.1f} s)")
    return ctx.lpc_get_params(), hist


def target_scale(data: Dataset, nprobe: int = 8) -> float:
    """kappa = 1 / mean over probe samples of rms(t)/rms(g): the network then
    predicts corrections of unit size, and P g = s y stays one fixed linear
    scale of the targets across bands and CG steps."""
    import ctypes

    from . import hip

    N = data.npoints
    idx = np.linspace(0, data.nsamp - 1, min(nprobe, data.nsamp)).astype(int)
    ratios = []
    for i in idx:
        gh = np.empty(N, dtype=np.float32)
        th = np.empty(N, dtype=np.float32)
        hip.runtime().hipMemcpy(gh.ctypes.data, ctypes.c_void_p(data.g.ptr + int(i) * N * 4), N * 4, 2)
        hip.runtime().hipMemcpy(th.ctypes.data, ctypes.c_void_p(data.t.ptr + int(i) * N * 4), N * 4, 2)
        ratios.append(np.sqrt(np.mean(th.astype(np.float64) ** 2) / np.mean(gh.astype(np.float64) ** 2)))
    return float(1.0 / np.mean(ratios))


def save(path: str, params: np.ndarray, meta: dict | None = None):
    np.savez(path, params=np.asarray(params, dtype=np.float32), **{k: np.asarray(v) for k, v in (meta or {}).items()})


def load_hybrid_scale(path: str):
    """The pc_kind 6 scale saved with hybrid-trained weights, else None."""
    with np.load(path, allow_pickle=False) as z:
        return float(z["hybrid_scale"]) if "hybrid_scale" in z.files else None


def load(path: str) -> np.ndarray:
    with np.load(path, allow_pickle=False) as z:
        p = np.asarray(z["params"], dtype=np.float32)
    if p.size != NPARAM:
        raise ValueError(f"{path}: {p.size} parameters, expected {NPARAM}")
    return p
