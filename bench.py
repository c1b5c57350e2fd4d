"""Benchmark: hpsi HBM throughput on one MI355X (Si64, 72^3 grid, 160 bands).

A step is one Hamiltonian application H psi over all 160 bands of the Si64
cell (2x2x2 diamond supercell, a = 20.52 bohr, atoms displaced by N(0, 0.05
bohr), 13-point-per-axis FD kinetic, GTH Si projectors), with psi and H psi
resident in HBM.  `value` is algorithmic GB/s of the whole call:

    B = 2 s N nb  (psi read, H psi written)  + s N (Vloc)  + (s_uv + 4) E

with E the projector-table entries (uVk value + JJP index), divided by the
measured time per call (HIP events on the library's stream).  The FP32 twin
is reported beside it.

`roofline` is for the dominant kernel of the call, the accumulating
kinetic + Vloc stencil (k_stencil), timed live here through rsg_op_kinetic
(the same kernel and launch shape hpsi uses) with HIP events on the
library's stream; its algorithmic bytes are 2 s N nb + s N.  `traffic` is
that kernel's HBM-side bytes per launch from the committed rocprofv3
FETCH_SIZE / WRITE_SIZE summary (scaled by a device-to-device copy of known
size), or null when the summary is absent.  The SCF leg times whole SCF runs
of Si216 in FP64 and in mixed precision; the step leg times two SCF steps of
Si512 (144^3, 1100 bands) in each precision, with lower bounds on the GEMM
rates of the Gram-Schmidt and subspace phases; the batch leg runs 64 uneven
Si8 cells as concurrent SCFs and one at a time; the learned-preconditioner
leg counts CG rounds to hprsdft's residual tolerance with the committed
weights.  The CPU baseline is the FP64
oracle hpsi on the host's cores, on a bounded band sample of the same cell.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "s to SCF conv (ΔE<1e-6 Ha/atom) Si216, mixed+learned vs FP64; hpsi HBM GB/s"


def hpsi_bytes(N: int, nb: int, E: int, s: int) -> float:
    return 2.0 * s * N * nb + s * N + (s + 4) * E


def gpu_leg(cell, vloc, steps: int, warmup: int, dtype, pin=None):
    from rsdft_develop_amd import gpu, hip, systems

    ctx = gpu.Context(0).setup(cell, vloc)
    prec = 64 if dtype == np.float64 else 32
    if pin:  # launch shape given (ry, ty, zlen, chunk, bt, order[, rx, proj, plds, psplit]): a counter run
        # repeats a timed run's shape
        keys = ("ry", "ty", "zlen", "chunk", "bt", "order", "rx", "proj", "plds", "psplit")
        ctx.apply_hpsi_shape(prec, dict(zip(keys, pin)))
    else:
        ctx.autotune_hpsi(prec, cell.nband, reps=6)  # launch shape from the library's autotuner
    shape = ctx.hpsi_shape(prec, cell.nband)
    tabs = ctx.projector_tables()
    E = int(tabs["nentries"])
    N, nb = cell.npoints, cell.nband
    psi = hip.DeviceArray.from_numpy(systems.random_bands(N, nb, seed=2024), dtype=dtype)
    hpsi = hip.DeviceArray((N, nb), dtype=dtype)

    def timed(fn):
        for _ in range(warmup):
            fn()
        ctx.synchronize()
        t0, t1 = hip.Event(), hip.Event()
        wall0 = time.perf_counter()
        t0.record(ctx.stream)
        for _ in range(steps):
            fn()
        t1.record(ctx.stream)
        t1.synchronize()
        return t0.elapsed_ms(t1) / steps, time.perf_counter() - wall0

    ms, wall = timed(lambda: ctx.hpsi_device(psi.ptr, hpsi.ptr, N, nb, dtype=dtype))
    # the dominant kernel alone, as hpsi launches it: accumulating onto the
    # nonlocal part (order 0) or writing H psi first (order 1)
    acc = shape.get("order", 0) == 0
    ms_k, _ = timed(lambda: ctx.op_kinetic_device(psi.ptr, hpsi.ptr, N, nb, dtype=dtype, accumulate=acc))
    s = 8 if dtype == np.float64 else 4
    B = hpsi_bytes(N, nb, E, s)
    Bk = 2.0 * s * N * nb + s * N
    # one device-to-device copy of s N nb bytes: the known-size reference a
    # counter run of this script scales FETCH_SIZE / WRITE_SIZE by
    hip.copy_d2d(hpsi.ptr, psi.ptr, s * N * nb)
    psi.free()
    hpsi.free()
    ctx.close()
    return dict(ms=ms, gbs=B / (ms * 1e-3) / 1e9, bytes=B, E=E, wall=wall, shape=shape, kernel_acc=acc,
                kernel_us=ms_k * 1e3, kernel_bytes=Bk, kernel_gbs=Bk / (ms_k * 1e-3) / 1e9)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_stencil_si64_72cubed_160bands_pmc.md")


def kernel_traffic(tag: str):
    """HBM-side bytes per launch of the stencil from the committed PMC
    summary (line 'TRAFFIC_BYTES_PER_LAUNCH <tag> <bytes>'), else None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "TRAFFIC_BYTES_PER_LAUNCH" and parts[1] == tag:
                    return float(parts[2])
    except OSError:
        return None
    return None


def physical_cores() -> int:
    """Physical cores this process may run on: lscpu's cores per socket times
    sockets, capped by the CPU affinity mask (a shared host gives a share)."""
    import subprocess

    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        info = dict(l.split(":", 1) for l in subprocess.run(["lscpu"], capture_output=True, text=True,
                                                           timeout=10).stdout.splitlines() if ":" in l)
        phys = int(info["Core(s) per socket"].strip()) * int(info["Socket(s)"].strip())
        threads_per_core = int(info.get("Thread(s) per core", "1").strip())
        return max(1, min(phys, allowed // max(1, threads_per_core) if allowed >= threads_per_core else allowed))
    except Exception:
        return max(1, allowed)


def cpu_leg(cell, vloc):
    """FP64 oracle hpsi over all 160 bands of the same cell on the host: the
    median of 10 calls on all physical cores and of 3 calls on one thread."""
    from oracle import pyoracle
    from rsdft_develop_amd import systems

    ham = pyoracle.Ham(cell, vloc=vloc)
    E = int(ham.tables()["mjj"].sum())
    psi = systems.random_bands(cell.npoints, cell.nband, seed=7)
    B = hpsi_bytes(cell.npoints, cell.nband, E, 8)

    def median_call(threads: int, calls: int) -> float:
        pyoracle.set_threads(threads)
        ham.hpsi(psi)  # warm-up
        ts = []
        for _ in range(calls):
            t = time.perf_counter()
            ham.hpsi(psi)
            ts.append(time.perf_counter() - t)
        return float(np.median(ts))

    cores = physical_cores()
    t_all = median_call(cores, 10)
    t_one = median_call(1, 3)
    pyoracle.set_threads(cores)
    return dict(value=B / t_all / 1e9, unit="GB/s", cores=cores, kind="port",
                sample=f"oracle hpsi FP64 over all {cell.nband} bands of the same Si64 cell, median of 10 calls on "
                       f"{cores} physical cores ({t_all * 1e3:.1f} ms/call)",
                one_thread={"value": B / t_one / 1e9, "unit": "GB/s", "cores": 1,
                            "sample": f"median of 3 calls on one thread ({t_one * 1e3:.1f} ms/call)"})


NBANDS = {1: 20, 2: 160, 3: 480, 4: 1100}


def progress(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def scf_leg(nrep: int, mixed: bool, etol: float = 1e-6, max_iter: int = 80, pc_kind: int = 4, weights=None,
            reject_limit: int = 0, precond_bits: int = 32, continue_iters: int = 0):
    """Wall seconds of a whole SCF run to dE < etol Ha/atom (the run's own
    set-up on the device included: tables, initial density and GS).  Pulay
    mixing in G space with Kerker weights (IMIX 30), EKBT 0.01 Ha.  With
    pc_kind 5 or 6 the learned preconditioner's weights come from `weights`,
    and after `reject_limit` CG calls in a row whose network output failed
    the descent check the run keeps hprsdft (0: the network is tried in
    every call).  precond_bits 16 keeps the internal vectors of the FP32
    hprsdft calls in bfloat16.  With continue_iters > 0 the run then goes
    on, untimed, for that many FP64 iterations (`esp_continued`,
    `etot_continued`).
    After the timed run, untimed, the band energies are converged at the
    run's final potential (FP64 eigensolver passes, SCF.refine_levels) and
    returned beside the run's own as `esp_levels`."""
    from rsdft_develop_amd import gpu, lpc, scf, systems

    cell = systems.silicon_supercell(nrep, nband=NBANDS[nrep], displacement=0.05, seed=0)
    ctx = gpu.Context(0)
    ctx.set_grid(cell.ngrid, cell.aa, cell.md)
    ctx.build_gth_projectors(cell)
    ctx.autotune_hpsi(64, cell.nband, reps=2)
    ctx.autotune_hpsi(32, cell.nband, reps=2)
    gram = ctx.autotune_gram(cell.nband)
    if weights is not None:
        ctx.lpc_set_params(lpc.load(weights))
        if pc_kind == 6:
            ctx.lpc_set_hybrid_scale(lpc.load_hybrid_scale(weights))
    tag = f"scf Si{cell.natom} {'mixed' if mixed else 'fp64'} pc{pc_kind}"
    progress(f"{tag}: start")
    fb0 = ctx.lpc_fallbacks()
    t0 = time.perf_counter()
    run = scf.SCF(cell, scf.ScfParams(ekbt=0.01, etol_per_atom=etol, max_iter=max_iter, mixed=mixed, imix=30,
                                      pc_kind=pc_kind, start="uniform", lpc_reject_limit=reject_limit,
                                      precond_bits=precond_bits), ctx=ctx)
    progress(f"{tag}: set-up {run.setup_s:.2f} s")
    res = run.run(callback=lambda r: progress(f"{tag}: iter {r.iter} prec {r.precision} dE/atom {r.de_per_atom:.2e}"))
    wall = time.perf_counter() - t0
    its = res.iterations
    out = dict(seconds=round(wall, 3), setup_s=round(res.setup, 3), iterations=len(its), converged=bool(res.converged),
               pc_kind=pc_kind, fp32_phase_precond_storage_bits=precond_bits if mixed else None,
               gram_chunk=gram[1], gram_ms=round(gram[0], 3),
               etot=res.etot,
               fp32_iterations=sum(1 for r in its if r.precision == 32),
               per_phase_s={k: round(sum(r.times.get(k, 0.0) for r in its), 3) for k in PHASES},
               esp=np.sort(res.esp).tolist())
    if pc_kind in (5, 6):
        out["descent_check_fallbacks"] = ctx.lpc_fallbacks() - fb0
        out["reject_limit"] = reject_limit
        out["cg_calls_without_network"] = ctx.lpc_reject_state()[1]
    if len(its) > 1:  # how far the band energies still move per iteration at the stop
        nocc = int(round(cell.nelectron / 2))
        out["last_iteration_eig_change_occupied_Ha"] = float(
            np.abs(np.sort(its[-1].esp)[:nocc] - np.sort(its[-2].esp)[:nocc]).max())
    run.p.pc_kind = 4  # the converged levels do not depend on the preconditioner
    levels, passes = run.refine_levels(tol=1e-9, max_passes=12)
    out["esp_levels"] = np.sort(levels).tolist()
    out["level_passes"] = passes
    if continue_iters > 0:
        progress(f"{tag}: {continue_iters} more FP64 iterations (untimed)")
        run.p.mixed, run.p.etol_per_atom, run.p.max_iter = False, 0.0, continue_iters
        rc = run.run()
        out["esp_continued"] = np.sort(rc.esp).tolist()
        out["etot_continued"] = rc.etot
        out["continued_last_dE_per_atom"] = rc.iterations[-1].de_per_atom
    ctx.close()
    return out


PHASES = ("sd", "cg", "gs", "esp", "eto", "pot")


def mean_phases(times: list) -> dict:
    """Mean seconds per SCF step in the reference log's categories."""
    return {k: round(sum(t.get(k, 0.0) for t in times) / max(1, len(times)), 5) for k in PHASES}


def scf_config0_leg():
    """The Si8 configuration (36^3, 20 bands, EKBT 0.01 Ha, IMIX 30, device
    start bands) solved to dE < 1e-8 Ha/atom twice: by the library on the
    GPU and by the oracle's FP64 CPU restatement on all host cores, with the
    same start bands.  Wall times of whole SCF runs, the mean time per step
    in the reference log's categories (sd/cg/gs/esp/eto/pot), and the two
    energies' difference."""
    from oracle import pyoracle, scf_oracle
    from rsdft_develop_amd import gpu, scf, systems

    cell = systems.silicon_supercell(1, nband=20, displacement=0.0, seed=0)
    params = dict(ekbt=0.01, etol_per_atom=1e-8, max_iter=80, imix=30, start="uniform")
    ctx = gpu.Context(0)
    scf.SCF(cell, scf.ScfParams(**params), ctx=ctx).run()  # first use of this size, untimed
    t0 = time.perf_counter()
    g = scf.SCF(cell, scf.ScfParams(**params), ctx=ctx).run()
    gpu_s = time.perf_counter() - t0
    ctx.close()
    cores = physical_cores()
    pyoracle.set_threads(cores)
    t0 = time.perf_counter()
    c = scf_oracle.run_scf(cell, **params)
    cpu_s = time.perf_counter() - t0
    return {"workload": "Si8 36^3, 20 bands, EKBT 0.01 Ha, IMIX 30, to dE < 1e-8 Ha/atom",
            "gpu_seconds": round(gpu_s, 3), "gpu_iterations": g.niter, "gpu_converged": bool(g.converged),
            "gpu_per_step_s": mean_phases([r.times for r in g.iterations]),
            "cpu_seconds": round(cpu_s, 3), "cpu_cores": cores, "cpu_kind": "port",
            "cpu_iterations": len(c.get("history", [])), "cpu_converged": bool(c.get("converged")),
            "cpu_per_step_s": mean_phases([h["times"] for h in c["history"]]),
            "etot_diff_Ha_per_atom": abs(float(g.etot) - float(c["etot"])) / cell.natom}


CONTINUE_ITERS = 40


def precision_si64_leg():
    """Configuration 1's precision report: the Si64 cell (72^3, 160 bands,
    displacements N(0, 0.05 bohr)) solved to dE < 1e-6 Ha/atom in plain FP64
    and in mixed precision (FP32 hpsi and CG, FP64 Gram-Schmidt, subspace
    eigensolve, density and energies), IMIX 30, EKBT 0.01 Ha.  Energy and
    band-energy differences of the mixed run against the FP64 one, as the
    runs leave them and after 12 fixed-potential passes, with the FP64
    run's own last-iteration level change for scale.  Then, untimed, both
    runs continue for CONTINUE_ITERS FP64 iterations and are compared again
    (`continued`): the stop test bounds the energy change only, and the
    levels of either run still move by about 1e-5 Ha after it."""
    from rsdft_develop_amd import gpu, scf, systems

    cell = systems.silicon_supercell(2, nband=NBANDS[2], displacement=0.05, seed=0)
    nocc = int(round(cell.nelectron / 2))
    out = {"workload": "Si64 72^3, 160 bands, EKBT 0.01 Ha, IMIX 30, to dE < 1e-6 Ha/atom"}
    res = {}
    # untimed: the first use of the kernels and library paths at this size
    ctx = gpu.Context(0)
    scf.SCF(cell, scf.ScfParams(ekbt=0.01, max_iter=2, mixed=True, fp64_below=1e9, imix=30, start="uniform"),
            ctx=ctx).run()
    ctx.close()
    for mixed in (False, True):
        ctx = gpu.Context(0)
        ctx.set_grid(cell.ngrid, cell.aa, cell.md)
        ctx.build_gth_projectors(cell)
        ctx.autotune_hpsi(64, cell.nband, reps=1)
        ctx.autotune_hpsi(32, cell.nband, reps=1)
        t0 = time.perf_counter()
        run = scf.SCF(cell, scf.ScfParams(ekbt=0.01, etol_per_atom=1e-6, max_iter=80, mixed=mixed, imix=30,
                                          start="uniform"), ctx=ctx)
        r = run.run()
        sec = time.perf_counter() - t0
        levels, _ = run.refine_levels(tol=1e-9, max_passes=12)
        its = r.iterations
        # untimed: the same run continued by CONTINUE_ITERS FP64 iterations,
        # towards the SCF fixed point both runs approach
        run.p.mixed, run.p.etol_per_atom, run.p.max_iter = False, 0.0, CONTINUE_ITERS
        rc = run.run()
        res[mixed] = dict(esp=np.sort(r.esp), levels=np.sort(levels), etot=r.etot, esp_c=np.sort(rc.esp),
                          etot_c=rc.etot)
        key = "mixed" if mixed else "fp64"
        out[key] = dict(seconds=round(sec, 3), iterations=r.niter, converged=bool(r.converged),
                        fp32_iterations=sum(1 for x in its if x.precision == 32))
        if len(its) > 1:
            out[key]["last_iteration_eig_change_occupied_Ha"] = float(
                np.abs(np.sort(its[-1].esp)[:nocc] - np.sort(its[-2].esp)[:nocc]).max())
        out[key]["continued_last_dE_per_atom"] = rc.iterations[-1].de_per_atom
        ctx.close()
    f, m = res[False], res[True]
    out["continued"] = dict(
        fp64_iterations_after_stop=CONTINUE_ITERS,
        max_eig_diff_vs_fp64_occupied_Ha=float(np.abs(m["esp_c"] - f["esp_c"])[:nocc].max()),
        max_eig_diff_vs_fp64_Ha=float(np.abs(m["esp_c"] - f["esp_c"]).max()),
        etot_diff_vs_fp64_Ha_per_atom=abs(m["etot_c"] - f["etot_c"]) / cell.natom,
        fp64_own_eig_change_from_stop_occupied_Ha=float(np.abs(f["esp_c"] - f["esp"])[:nocc].max()))
    out["mixed"].update(
        etot_diff_vs_fp64_Ha_per_atom=abs(m["etot"] - f["etot"]) / cell.natom,
        max_eig_diff_vs_fp64_occupied_Ha=float(np.abs(m["esp"] - f["esp"])[:nocc].max()),
        max_eig_diff_vs_fp64_Ha=float(np.abs(m["esp"] - f["esp"]).max()),
        max_level_diff_vs_fp64_occupied_Ha=float(np.abs(m["levels"] - f["levels"])[:nocc].max()),
        max_level_diff_vs_fp64_Ha=float(np.abs(m["levels"] - f["levels"]).max()),
        speedup_vs_fp64=round(out["fp64"]["seconds"] / out["mixed"]["seconds"], 3))
    return out


def cpu_si216_step_leg(steps: int):
    """Per-SCF-step time of the oracle's FP64 restatement at Si216 (108^3,
    480 bands, EKBT 0.01 Ha, IMIX 30) on all host cores, in the reference
    log's categories: `steps` steps only, no extrapolation to convergence."""
    from oracle import pyoracle, scf_oracle
    from rsdft_develop_amd import systems

    cell = systems.silicon_supercell(3, nband=NBANDS[3], displacement=0.05, seed=0)
    cores = physical_cores()
    pyoracle.set_threads(cores)
    t0 = time.perf_counter()
    progress("cpu Si216: set-up (tables, start bands, Gram-Schmidt)")
    c = scf_oracle.run_scf(cell, ekbt=0.01, etol_per_atom=0.0, max_iter=steps, imix=30, start="uniform",
                           phase_callback=lambda name, sec: progress(f"cpu Si216 {name} done ({sec:.1f} s)"),
                           callback=lambda h: progress(f"cpu Si216 step {h['iter']}: "
                                                       + " ".join(f"{k} {v:.1f}" for k, v in h["times"].items())))
    return {"workload": f"Si216 108^3, 480 bands, {steps} SCF steps of the oracle (FP64) on {cores} cores",
            "seconds_with_setup": round(time.perf_counter() - t0, 2), "cores": cores, "kind": "port",
            "per_step_s": [{k: round(v, 3) for k, v in h["times"].items()} for h in c["history"]]}


def scf_warmup(nrep: int):
    """One FP32 and one FP64 SCF iteration of the same cell, untimed, so that
    neither timed run pays the process's first use of the kernels and
    library paths at this size (rocSOLVER's eigensolver for its band count
    among them)."""
    from rsdft_develop_amd import gpu, lpc, scf, systems

    cell = systems.silicon_supercell(nrep, nband=NBANDS[nrep], displacement=0.05, seed=0)
    ctx = gpu.Context(0)
    ctx.set_grid(cell.ngrid, cell.aa, cell.md)
    ctx.build_gth_projectors(cell)
    for pc_kind, weights in ((4, None), (6, LPC_HYBRID)):
        if weights is not None:
            ctx.lpc_set_params(lpc.load(weights))
            ctx.lpc_set_hybrid_scale(lpc.load_hybrid_scale(weights))
        scf.SCF(cell, scf.ScfParams(ekbt=0.01, max_iter=2, mixed=True, fp64_below=1e9, imix=30, start="uniform",
                                    pc_kind=pc_kind), ctx=ctx).run()
    ctx.close()


def step_leg(nrep: int, steps: int, precond_bits: int = 16, reject_limit: int = 2):
    """Time per SCF step of Si(8 nrep^3) in FP64 and in mixed precision (FP32
    band store): with hprsdft in FP32 (`mixed`, the same algorithm as FP64
    up to rounding), with its internal vectors stored with `precond_bits`
    (`mixed_bf16pc`), and with the learned preconditioner (`mixed_learned`,
    configuration 4: tried until `reject_limit` CG calls in a row rejected
    it), for the same number of steps (IMIX 30, EKBT 0.01 Ha, device start
    bands); the energy difference to FP64 after each step (from random start
    bands the first steps' energies follow the CG's progress, so only
    `mixed` measures rounding; a different preconditioner moves them by its
    different progress); and the FP64 GEMM rates of Gram-Schmidt
    (S = dV U^T U: 2 MB^2 N FLOP) and subspace_diag (U^T HU and the
    rotation: 4 MB^2 N FLOP) as lower bounds: whole phase time, including
    its hpsi, Cholesky and eigensolve."""
    from rsdft_develop_amd import gpu, lpc, scf, systems

    cell = systems.silicon_supercell(nrep, nband=NBANDS[nrep], displacement=0.05, seed=0)
    N, mb = cell.npoints, cell.nband
    out = {"workload": f"Si{cell.natom} {cell.ngrid[0]}^3 grid, {mb} bands, {steps} SCF steps"}
    etots = {}
    variants = [("mixed", True, 4, 32), ("mixed_bf16pc", True, 4, precond_bits),
                ("mixed_learned", True, 5, precond_bits), ("fp64", False, 4, 32)]
    for key, mixed, pc, bits in variants:
        if pc == 5 and not os.path.exists(LPC_WEIGHTS):
            continue
        ctx = gpu.Context(0)
        ctx.set_grid(cell.ngrid, cell.aa, cell.md)
        ctx.build_gth_projectors(cell)
        ctx.autotune_hpsi(64, mb, reps=1)
        ctx.autotune_hpsi(32, mb, reps=1)
        ctx.autotune_gram(mb, reps=1)
        if pc == 5:
            ctx.lpc_set_params(lpc.load(LPC_WEIGHTS))
        tag = f"step Si{cell.natom} {key}"
        progress(f"{tag}: start")
        fb0 = ctx.lpc_fallbacks()
        run = scf.SCF(cell, scf.ScfParams(ekbt=0.01, etol_per_atom=0.0, max_iter=steps, mixed=mixed, imix=30,
                                          start="uniform", pc_kind=pc, precond_bits=bits,
                                          lpc_reject_limit=reject_limit), ctx=ctx)
        res = run.run(callback=lambda r: progress(f"{tag}: iter {r.iter} prec {r.precision} "
                                                  + " ".join(f"{k} {v:.2f}" for k, v in r.times.items())))
        its = res.iterations
        per = {k: round(sum(r.times.get(k, 0.0) for r in its) / len(its), 3)
               for k in ("sd", "cg", "gs", "esp", "eto", "pot")}
        out[key] = dict(seconds_per_step=round(sum(sum(r.times.values()) for r in its) / len(its), 3),
                        setup_s=round(res.setup, 3), setup_parts_s={k: round(v, 4) for k, v in res.setup_parts.items()},
                        per_phase_s=per, pc_kind=pc,
                        precisions=[r.precision for r in its])
        if mixed:
            out[key]["fp32_phase_precond_storage_bits"] = bits
        if pc == 5:
            out[key]["descent_check_fallbacks"] = ctx.lpc_fallbacks() - fb0
            out[key]["cg_calls_without_network"] = ctx.lpc_reject_state()[1]
        etots[key] = [r.etot for r in its]
        if not mixed:
            out["fp64"]["gs_gemm_tflops_lower_bound"] = round(2.0 * mb * mb * N / per["gs"] / 1e12, 2)
            out["fp64"]["sd_gemm_tflops_lower_bound"] = round(4.0 * mb * mb * N / per["sd"] / 1e12, 2)
        ctx.close()
    for key in ("mixed", "mixed_bf16pc", "mixed_learned"):
        if key in out:
            out[key]["etot_diff_vs_fp64_Ha_per_atom"] = [abs(a - b) / cell.natom
                                                         for a, b in zip(etots[key], etots["fp64"])]
            out[key]["speedup_vs_fp64_per_step"] = round(out["fp64"]["seconds_per_step"]
                                                         / out[key]["seconds_per_step"], 3)
    out["mixed_vs_fp64_etot_diff_Ha_per_atom"] = out["mixed"]["etot_diff_vs_fp64_Ha_per_atom"]
    out["speedup_mixed_vs_fp64_per_step"] = out["mixed"]["speedup_vs_fp64_per_step"]
    return out


def batch_leg(ncells: int = 64, workers: int = 8, seq_cells: int | None = None):
    """Config-2 throughput: ncells uneven Si8 cells (scale U(0.95, 1.10),
    displacements N(0, 0.1), grids {24, 32, 40, 48}^3, U{20..40} bands), plain
    FP64 SCF to dE < 1e-6 Ha/atom, `workers` concurrent contexts, against the
    same solves one at a time (all of them unless seq_cells is given).  Launch
    shapes are tuned per grid before either timing.  The process runs with
    16 hardware queues (GPU_MAX_HW_QUEUES, set before HIP starts) so that
    the workers' streams do not share 4."""
    from rsdft_develop_amd import batch

    cells = batch.config2_cells(ncells, seed=2024)
    seq_cells = ncells if seq_cells is None else seq_cells
    shapes = batch.ShapeCache(reps=2)
    shapes.warm(cells)
    # untimed pass over one cell per grid: loads the library kernels (rocBLAS,
    # rocSOLVER, hipFFT code objects load on first use) before either timing
    firsts = list({c.cell.ngrid: c for c in cells}.values())
    batch.run_batch(firsts, workers=1, shapes=shapes)
    progress(f"batch: {ncells} cells, {workers} workers, native SCF runs")
    par = batch.run_batch(cells, workers=workers, shapes=shapes, graphs=False, native=True)
    progress(f"batch: {par.wall:.2f} s; native with CG from captured graphs")
    par_native_graphs = batch.run_batch(cells, workers=workers, shapes=shapes, graphs=True, native=True)
    progress(f"batch: {par_native_graphs.wall:.2f} s; Python driver with graphs")
    par_py = batch.run_batch(cells, workers=workers, shapes=shapes, graphs=True)
    progress(f"batch: {par_py.wall:.2f} s; Python driver without graphs")
    par_direct = batch.run_batch(cells, workers=workers, shapes=shapes, graphs=False)
    progress(f"batch: {par_direct.wall:.2f} s; sequential plain solves of {seq_cells}")
    seq = batch.run_batch(cells[:seq_cells], workers=1, shapes=shapes, graphs=False)
    its = [r.iterations for r in par.results]
    same = all(a.iterations == b.iterations and a.etot == b.etot for a, b in zip(par_py.results, par_direct.results))
    native_vs_python = max(max(abs(x - y) for x, y in zip(a.history, b.history)) / c.cell.natom
                           for a, b, c in zip(par.results, par_direct.results, cells))
    return dict(workload=f"{ncells} Si8 cells, scale U(0.95,1.10), displacement N(0,0.1), grids {{24,32,40,48}}^3, "
                         f"U{{20..40}} bands, FP64 SCF to dE < 1e-6 Ha/atom",
                gpu_max_hw_queues=os.environ.get("GPU_MAX_HW_QUEUES"),
                workers=workers, batched_seconds=round(par.wall, 3), batched_cells_per_s=round(par.cells_per_s, 3),
                batched_mode="native SCF runs (rsg_scf_run), direct launches",
                native_graphs_cells_per_s=round(par_native_graphs.cells_per_s, 3),
                python_driver_graphs_cells_per_s=round(par_py.cells_per_s, 3),
                python_driver_direct_launch_cells_per_s=round(par_direct.cells_per_s, 3),
                graph_results_equal_direct=same, graph_launches=par_native_graphs.graph_launches,
                graph_capture_failures=par_native_graphs.graph_failures + par_py.graph_failures,
                native_vs_python_max_etot_diff_Ha_per_atom=native_vs_python,
                sequential_cells=seq_cells, sequential_seconds=round(seq.wall, 3),
                sequential_cells_per_s=round(seq.cells_per_s, 3),
                speedup=round(par.cells_per_s / seq.cells_per_s, 3),
                all_converged=all(r.converged for r in par.results) and all(r.converged for r in seq.results),
                iterations_min_max=[min(its), max(its)])


LPC_WEIGHTS = os.path.join(ROOT, "rsdft_develop_amd", "data", "lpc_si64.npz")
LPC_HYBRID = os.path.join(ROOT, "rsdft_develop_amd", "data", "lpc_si64_hybrid.npz")


def lpc_leg():
    """The learned CG preconditioner (weights trained by train_lpc on 20480
    residuals of 32 Si64 cells) judged by CG rounds to the residual
    tolerance the plain hprsdft method reaches in 8 rounds, on a held-out
    Si64 cell, against hprsdft and no preconditioner."""
    if not os.path.exists(LPC_WEIGHTS):
        return None
    from rsdft_develop_amd import lpc, systems, train_lpc

    params = lpc.load(LPC_WEIGHTS)
    cell = systems.silicon_supercell(2, displacement=0.05, seed=999)
    ev = train_lpc.evaluate(params, cell)
    keep = ("tolerance", "rounds_hprsdft", "rounds_learned", "rounds_none", "reached_learned", "fallbacks",
            "seconds_hprsdft", "seconds_learned", "seconds_none")
    out = {"workload": "held-out Si64 cell, FP64 CG(4) rounds at a fixed potential", **{k: ev[k] for k in keep}}
    if os.path.exists(LPC_HYBRID):  # the network as a correction of hprsdft (CG pc_kind 6)
        hy = train_lpc.evaluate(lpc.load(LPC_HYBRID), cell, hybrid_scale=lpc.load_hybrid_scale(LPC_HYBRID))
        out.update(rounds_hybrid=hy["rounds_learned"], reached_hybrid=hy["reached_learned"],
                   fallbacks_hybrid=hy["fallbacks"], seconds_hybrid=hy["seconds_learned"])
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--scf", type=int, default=3, help="supercell repeat for the SCF-time leg (3 = Si216); 0 skips")
    ap.add_argument("--si512", type=int, default=2, help="SCF steps of the Si512 per-step leg (0 skips)")
    ap.add_argument("--batch", type=int, default=64, help="cells of the batched-SCF leg (0 skips)")
    ap.add_argument("--batch-workers", type=int, default=8, help="concurrent contexts of the batched-SCF leg")
    ap.add_argument("--lpc", type=int, default=1, help="learned-preconditioner leg on the held-out Si64 cell (0 skips)")
    ap.add_argument("--precision64", type=int, default=1, help="Si64 FP64-vs-mixed SCF precision leg (0 skips)")
    ap.add_argument("--mixed-precond-bits", type=int, default=16, choices=(16, 32),
                    help="Si216 mixed legs: storage of the FP32 hprsdft calls' internal vectors (16 = bfloat16)")
    ap.add_argument("--scf-continue", type=int, default=40,
                    help="Si216 FP64 and mixed legs: FP64 iterations after the stop, untimed, before a second "
                         "comparison (0 skips)")
    ap.add_argument("--lpc-reject-limit", type=int, default=2,
                    help="Si216 learned/hybrid legs: CG calls in a row with a rejected network output after which "
                         "the run keeps hprsdft (0: try the network in every call)")
    ap.add_argument("--cpu-si216-steps", type=int, default=0,
                    help="SCF steps of the oracle at Si216 on the host cores (0 skips; about a minute per step)")
    ap.add_argument("--shape64", default=None, help="pin the FP64 hpsi launch shape: ry,ty,zlen,chunk,bt,order[,rx,proj,plds,psplit]")
    ap.add_argument("--shape32", default=None, help="pin the FP32 hpsi launch shape: ry,ty,zlen,chunk,bt,order[,rx,proj,plds,psplit]")
    args = ap.parse_args()
    # more hardware queues for the batched leg's concurrent streams; read by
    # HIP when it starts, which has not happened yet in this process
    os.environ["GPU_MAX_HW_QUEUES"] = "16"
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the path is not sharded in this build); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2
    from rsdft_develop_amd import systems

    cell = systems.silicon_supercell(2, nband=160, displacement=0.05, seed=0)
    vloc = systems.model_vloc(cell)
    progress("hpsi FP64")
    pin = lambda t: [int(v) for v in t.split(",")] if t else None  # noqa: E731
    r64 = gpu_leg(cell, vloc, args.steps, args.warmup, np.float64, pin(args.shape64))
    progress("hpsi FP32")
    r32 = gpu_leg(cell, vloc, args.steps, args.warmup, np.float32, pin(args.shape32))
    batch_out = batch_leg(args.batch, args.batch_workers) if args.batch > 0 else None
    progress("learned preconditioner")
    lpc_out = lpc_leg() if args.lpc else None
    progress("CPU baseline")
    cpu = None if args.no_cpu else cpu_leg(cell, vloc)
    progress("Si8 SCF on the GPU and on the host")
    scf0 = None if args.no_cpu else scf_config0_leg()
    cpu216 = cpu_si216_step_leg(args.cpu_si216_steps) if args.cpu_si216_steps > 0 else None
    if args.precision64:
        progress("Si64 precision: FP64 and mixed SCF")
    prec64 = precision_si64_leg() if args.precision64 else None
    scf_out = None
    if args.scf > 0:
        progress("scf warm-up")
        scf_warmup(args.scf)
        natom = 8 * args.scf ** 3
        nocc = 2 * natom  # 4 valence electrons per Si: the lowest 2 natom bands are the occupied ones
        # plain FP64, then mixed precision with hprsdft, with the learned
        # preconditioner (pc_kind 5) and with the hybrid (pc_kind 6); errors
        # against the plain FP64 run (the FP64 path equals the oracle's SCF
        # at Si8 and its operators at this size), from the runs' own band
        # energies and from the levels converged at each run's final potential
        plain = scf_leg(args.scf, mixed=False, continue_iters=args.scf_continue)
        runs = {"fp64": plain,
                "mixed": scf_leg(args.scf, mixed=True, precond_bits=args.mixed_precond_bits,
                                 continue_iters=args.scf_continue),
                "mixed_learned": scf_leg(args.scf, mixed=True, pc_kind=5, weights=LPC_WEIGHTS,
                                         reject_limit=args.lpc_reject_limit, precond_bits=args.mixed_precond_bits),
                "mixed_hybrid": scf_leg(args.scf, mixed=True, pc_kind=6, weights=LPC_HYBRID,
                                        reject_limit=args.lpc_reject_limit, precond_bits=args.mixed_precond_bits)}
        e_ref = np.array(plain["esp"])
        l_ref = np.array(plain["esp_levels"])
        continued = None
        if args.scf_continue > 0:
            f, m = runs["fp64"], runs["mixed"]
            ef, em = np.array(f.pop("esp_continued")), np.array(m.pop("esp_continued"))
            de = abs(m.pop("etot_continued") - f.pop("etot_continued")) / natom
            continued = dict(fp64_iterations_after_stop=args.scf_continue,
                             max_eig_diff_vs_fp64_occupied_Ha=float(np.abs(em - ef)[:nocc].max()),
                             max_eig_diff_vs_fp64_Ha=float(np.abs(em - ef).max()),
                             etot_diff_vs_fp64_Ha_per_atom=de,
                             fp64_own_eig_change_from_stop_occupied_Ha=float(np.abs(ef - e_ref)[:nocc].max()))
        for name, r in runs.items():
            d = np.abs(np.array(r.pop("esp")) - e_ref)
            dl = np.abs(np.array(r.pop("esp_levels")) - l_ref)
            if name == "fp64":
                continue
            r["max_eig_diff_vs_fp64_Ha"] = float(d.max())
            r["max_eig_diff_vs_fp64_occupied_Ha"] = float(d[:nocc].max())
            r["max_level_diff_vs_fp64_Ha"] = float(dl.max())
            r["max_level_diff_vs_fp64_occupied_Ha"] = float(dl[:nocc].max())
            r["etot_diff_vs_fp64_Ha_per_atom"] = abs(r["etot"] - plain["etot"]) / natom
            r["speedup_vs_fp64"] = round(plain["seconds"] / r["seconds"], 3)
        scf_out = {"workload": f"Si{natom} {36 * args.scf}^3 grid, {NBANDS[args.scf]} bands, EKBT 0.01 Ha, "
                               f"IMIX 30, uniform random start bands drawn on the device, to dE < 1e-6 Ha/atom",
                   "occupied_bands": nocc, **runs, "mixed_vs_fp64_continued": continued,
                   "speedup_mixed_learned_vs_fp64": runs["mixed_learned"]["speedup_vs_fp64"],
                   "speedup_mixed_vs_fp64": runs["mixed"]["speedup_vs_fp64"]}
    step_out = (step_leg(4, args.si512, precond_bits=args.mixed_precond_bits, reject_limit=args.lpc_reject_limit)
                if args.si512 > 0 else None)
    out = {
        "metric": METRIC,
        "value": round(r64["gbs"], 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r64["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded N(0,1) bands, model Vloc, GTH Si tables)",
        "config": {"workload": "hpsi Si64 2x2x2 a=20.52 bohr, 72^3 grid, 160 bands, Md=6",
                   "grid": [72, 72, 72], "bands": 160, "projector_entries": r64["E"]},
        "roofline": {"bound": "hbm", "achieved": round(r64["kernel_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r64["kernel_gbs"] / HBM_PEAK_GBS, 4), "traffic": kernel_traffic("f64"),
                     "kernel": f"k_stencil<double, 6, ry, hm, {'true' if r64['kernel_acc'] else 'false'}> "
                               "(13-point FD kinetic + Vloc)",
                     "kernel_avg_us": round(r64["kernel_us"], 1),
                     "algorithmic_bytes": r64["kernel_bytes"]},
        "hpsi_call": {"algorithmic_bytes": r64["bytes"], "frac": round(r64["gbs"] / HBM_PEAK_GBS, 4)},
        "hpsi_shape": r64["shape"],
        "f32": {"value": round(r32["gbs"], 1), "ms_per_step": round(r32["ms"], 4),
                "frac": round(r32["gbs"] / HBM_PEAK_GBS, 4), "hpsi_shape": r32["shape"],
                "kernel_avg_us": round(r32["kernel_us"], 1), "kernel_gbs": round(r32["kernel_gbs"], 1),
                "kernel_traffic": kernel_traffic("f32")},
        "cpu_baseline": cpu,
        "scf_si8_gpu_vs_cpu": scf0,
        "cpu_si216_steps": cpu216,
        "precision_si64": prec64,
        "scf": scf_out,
        "si512_step": step_out,
        "batch": batch_out,
        "learned_preconditioner": lpc_out,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
