"""Train the learned CG preconditioner on the GPU and evaluate it.

    python -m rsdft_develop_amd.train_lpc --cells 32 --scf-iters 4 --epochs 4 \
        --out rsdft_develop_amd/data/lpc_si64.npz
    python -m rsdft_develop_amd.train_lpc ... --hybrid 3 \
        --out rsdft_develop_amd/data/lpc_si64_hybrid.npz

With --hybrid the network is trained as a correction of the hprsdft output
(CG pc_kind 6, P g = scale P_hprsdft g + net(g)); the scale is saved with
the weights.

Data: residuals of the plain method's own bands from `scf-iters` plain SCF
iterations on `cells` perturbed Si64 cells (2x2x2, 72^3, 160 bands, atoms
displaced N(0, 0.05 bohr)), with targets from the shifted potential-aware
solve (rsdft_develop_amd.lpc).  Training: Adam on the library's own BF16
forward/backward kernels.  Evaluation (`evaluate`): CG rounds at a fixed
potential until every occupied band's residual is below the tolerance the
plain hprsdft method reaches after `ref_rounds` rounds.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def log(msg):
    print(msg, flush=True)


def si64_cells(n, seed0=100):
    from . import systems

    return [systems.silicon_supercell(2, displacement=0.05, seed=seed0 + i) for i in range(n)]


def evaluate(params, cell, ref_rounds=8, max_rounds=40, scf_warm=3, precision=64, log=None, hybrid_scale=None):
    """Rounds (subspace_diag -> CG(4) -> GS) to reach the plain method's
    residual tolerance with pc_kind 4 (hprsdft) and 5 (learned), at the
    potential of `scf_warm` plain SCF iterations; same start bands.  With
    hybrid_scale (weights trained as a correction of hprsdft) the learned
    runs use pc_kind 6, P g = hybrid_scale P_hprsdft g + net(g)."""
    from . import gpu, scf, systems

    ctx = gpu.Context(0)
    run = scf.SCF(cell, scf.ScfParams(ekbt=0.01), ctx=ctx)
    e_prev = None
    for it in range(scf_warm):
        e_prev = run.step(it + 1, e_prev).etot
        run.update_potential()
    nocc = int(round(run.nelectron / 2))
    u0 = ctx.download_bands()
    ctx.lpc_set_params(params)
    ctx.lpc_set_potential()
    pc_learned = 5
    if hybrid_scale is not None:
        ctx.lpc_set_hybrid_scale(hybrid_scale)
        pc_learned = 6

    def rounds(pc, tol=None, nmax=max_rounds):
        ctx.upload_bands(u0)
        hist = []
        t0 = time.perf_counter()
        for r in range(nmax):
            ctx.subspace_diag()
            _, res = ctx.conjugate_gradient(4, pc, 3, precision=precision)
            ctx.gram_schmidt()
            worst = float(np.max(res[:nocc]))
            hist.append(worst)
            if tol is not None and worst <= tol:
                break
        return hist, time.perf_counter() - t0

    ref_hist, _ = rounds(4, nmax=ref_rounds)
    tol = ref_hist[-1]
    plain, t_plain = rounds(4, tol)
    learned, t_learned = rounds(pc_learned, tol)
    none, t_none = rounds(0, tol)
    out = dict(pc_kind_learned=pc_learned, tolerance=tol, rounds_hprsdft=len(plain), rounds_learned=len(learned),
               rounds_none=len(none),
               reached_learned=learned[-1] <= tol, reached_none=none[-1] <= tol,
               seconds_hprsdft=round(t_plain, 3), seconds_learned=round(t_learned, 3), seconds_none=round(t_none, 3),
               fallbacks=ctx.lpc_fallbacks(), residuals_hprsdft=plain, residuals_learned=learned)
    ctx.close()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, default=32)
    ap.add_argument("--scf-iters", type=int, default=4)
    ap.add_argument("--max-samples", type=int, default=20480)
    ap.add_argument("--target-steps", type=int, default=40)
    ap.add_argument("--shift", type=float, default=0.1)
    ap.add_argument("--vweight", type=float, default=0.0)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default="rsdft_develop_amd/data/lpc_si64.npz")
    ap.add_argument("--eval", action="store_true", help="evaluate on a held-out Si64 cell afterwards")
    ap.add_argument("--hybrid", type=int, default=0,
                    help="mloop > 0: train the network as a correction of hprsdft(mloop) (pc_kind 6)")
    args = ap.parse_args(argv)
    from . import gpu, lpc

    cells = si64_cells(args.cells)
    t0 = time.perf_counter()
    data = lpc.make_dataset(lambda: gpu.Context(0), cells, scf_iters=args.scf_iters, target_steps=args.target_steps,
                            shift=args.shift, vweight=args.vweight, max_samples=args.max_samples, log=log,
                            hybrid_mloop=args.hybrid)
    log(f"dataset: {data.nsamp} samples of {data.npoints} points in {data.seconds:.1f} s")
    ctx = gpu.Context(0)
    c0 = cells[0]
    ctx.set_grid(c0.ngrid, c0.aa, c0.md)
    log(f"training: {len(data.blocks)} blocks, batch {args.batch}, {args.epochs} epochs")
    kappa, hybrid_scale = None, None
    if args.hybrid > 0:
        kappa, mu = lpc.hybrid_targets(ctx, data)
        hybrid_scale = kappa * mu
        log(f"hybrid: kappa {kappa:.4e}, mu {mu:.4e}, scale {hybrid_scale:.4e}")
    p0 = lpc.init_params(args.seed)
    t1 = time.perf_counter()
    params, hist = lpc.train(ctx, data, p0, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed, log=log,
                             kappa=kappa)
    train_s = time.perf_counter() - t1
    data.g.free()
    data.t.free()
    if data.h is not None:
        data.h.free()
    ctx.close()
    meta = dict(loss_history=np.array(hist), samples=data.nsamp, cells=args.cells, epochs=args.epochs,
                shift=args.shift, vweight=args.vweight, target_steps=args.target_steps)
    if hybrid_scale is not None:
        meta.update(hybrid_scale=hybrid_scale, hybrid_mloop=args.hybrid)
    lpc.save(args.out, params, meta)
    summary = dict(samples=data.nsamp, dataset_s=round(data.seconds, 1), train_s=round(train_s, 1),
                   loss_history=hist, out=args.out, shift=args.shift, vweight=args.vweight,
                   target_steps=args.target_steps)
    if args.eval:
        from . import systems

        held_out = systems.silicon_supercell(2, displacement=0.05, seed=999)
        summary["evaluation"] = evaluate(params, held_out, hybrid_scale=hybrid_scale)
    summary["total_s"] = round(time.perf_counter() - t0, 1)
    print(json.dumps(summary))
    return 0


if __name__ == "__main__":
    sys.exit(main())
