"""Batched SCF over many uneven cells on one GPU.

Each cell runs the plain FP64 SCF on a library context of its own, with
its own HIP stream and rocBLAS/rocSOLVER/FFT state.  Worker threads pull
cells from a queue.  The library calls release the GIL, so the streams'
small launches overlap on the GPU, where one cell alone leaves most of the
chip idle.

Each context replays the CG band blocks of its SCF steps from captured HIP
graphs once their buffers and launch parameters have settled (two calls
with the same block), so a small cell's CG step costs one graph launch
instead of a hundred kernel launches.

The hpsi launch shape is tuned once per grid size by the library's
autotuner (`rsg_autotune_hpsi`) and handed to every context that meets that
grid.  The band chunk is left to the library, so it follows each cell's
band count.

No counterpart in the reference, which solves one cell per run.
"""
from __future__ import annotations

import dataclasses
import queue
import threading
import time

import numpy as np

from . import gpu, scf, systems


@dataclasses.dataclass
class BatchCell:
    cell: systems.Cell
    scale: float
    seed: int


def config2_cells(n: int = 64, seed: int = 2024) -> list:
    """n Si8 cells with lattice scale U(0.95, 1.10), atom displacements
    N(0, 0.1 bohr), grids drawn from {24, 32, 40, 48}^3 and U{20..40} bands."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        scale = float(rng.uniform(0.95, 1.10))
        ng = int(rng.choice([24, 32, 40, 48]))
        nb = int(rng.integers(20, 41))
        cseed = int(rng.integers(0, 2 ** 31))
        cell = systems.silicon_supercell(1, nband=nb, displacement=0.1, seed=cseed, scale=scale, ngrid=ng)
        out.append(BatchCell(cell, scale, cseed))
    return out


@dataclasses.dataclass
class CellResult:
    index: int
    converged: bool
    iterations: int
    etot: float
    seconds: float
    history: list = dataclasses.field(default_factory=list)  # Etot after each SCF iteration


@dataclasses.dataclass
class BatchResult:
    results: list
    wall: float
    workers: int
    graph_launches: int = 0  # CG band blocks replayed from captured graphs, all workers
    graph_failures: int = 0  # captures the runtime refused (those blocks ran directly)

    @property
    def cells_per_s(self) -> float:
        return len(self.results) / self.wall


class ShapeCache:
    """Autotuned hpsi launch shapes per (grid, precision), shared by workers."""

    def __init__(self, reps: int = 2):
        self.reps = reps
        self.lock = threading.Lock()
        self.shapes = {}

    def warm(self, cells: list, precision: int = 64):
        """Tune every distinct grid of `cells` now (one scratch context each)."""
        seen = {}
        for bc in cells:
            seen.setdefault(tuple(bc.cell.ngrid), bc.cell)
        for cell in seen.values():
            ctx = gpu.Context(0)
            try:
                ctx.set_grid(cell.ngrid, cell.aa, cell.md)
                ctx.build_gth_projectors(cell)
                self.apply(ctx, cell, precision)
            finally:
                ctx.close()

    def apply(self, ctx: gpu.Context, cell: systems.Cell, precision: int = 64):
        key = (tuple(cell.ngrid), precision)
        with self.lock:
            sh = self.shapes.get(key)
            if sh is None:  # tune on this context (grid and projectors already set)
                ctx.autotune_hpsi(precision, cell.nband, reps=self.reps)
                sh = ctx.hpsi_shape(precision, cell.nband)
                self.shapes[key] = sh
                return
        ctx.apply_hpsi_shape(precision, sh, chunk=0)


def run_batch(cells: list, workers: int = 8, params: scf.ScfParams | None = None, shapes: ShapeCache | None = None,
              log=None, graphs: bool = True, native: bool = False) -> BatchResult:
    """SCF of every cell, `workers` at a time.  Results in input order.
    With `graphs` every context replays its CG band blocks from captured
    HIP graphs (rsg_set_cg_graphs), which removes the per-kernel launch cost
    that dominates small cells.  With `native` each cell's whole SCF is one
    library call (rsg_scf_run), so a worker holds the interpreter only
    between cells; otherwise the Python driver (scf.SCF) runs it."""
    params = params or scf.ScfParams(ekbt=0.01, etol_per_atom=1e-6, max_iter=80, imix=30, start="uniform")
    shapes = shapes or ShapeCache()
    todo = queue.Queue()
    # largest work first, cells of one grid together
    order = sorted(range(len(cells)), key=lambda i: (-cells[i].cell.npoints, -cells[i].cell.nband))
    for i in order:
        todo.put(i)
    results = [None] * len(cells)
    errors = []
    graph_stats = []

    def worker():
        # one context per worker for all its cells: device buffers and, on
        # equal grid dimensions, the FFT plan are reused instead of freed
        # (a free waits for the whole device, stalling every other worker)
        ctx = gpu.Context(0)
        try:
            ctx.set_cg_graphs(graphs)
            while True:
                try:
                    i = todo.get_nowait()
                except queue.Empty:
                    return
                bc = cells[i]
                t0 = time.perf_counter()
                try:
                    ctx.set_grid(bc.cell.ngrid, bc.cell.aa, bc.cell.md)
                    ctx.build_gth_projectors(bc.cell)
                    shapes.apply(ctx, bc.cell, 64)
                    p = dataclasses.replace(params, seed=bc.seed)
                    if native:
                        ctx.set_local_gth(bc.cell)
                        r = ctx.scf_run(bc.cell.nband, p)
                        results[i] = CellResult(i, r["converged"], r["iterations"], float(r["etot"]),
                                                time.perf_counter() - t0, r["history"])
                    else:
                        res = scf.SCF(bc.cell, p, ctx=ctx).run()
                        results[i] = CellResult(i, bool(res.converged), res.niter, float(res.etot),
                                                time.perf_counter() - t0, [float(r.etot) for r in res.iterations])
                    if log:
                        log(f"cell {i}: {bc.cell.ngrid[0]}^3, {bc.cell.nband} bands, {results[i].iterations} "
                            f"iterations, {time.perf_counter() - t0:.2f} s")
                except Exception as e:  # noqa: BLE001 - reported to the caller after the pool drains
                    errors.append((i, e))
                    return
        finally:
            try:
                graph_stats.append((ctx.cg_graph_launches(), ctx.cg_graph_failures()))
            finally:
                ctx.close()

    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker, daemon=True) for _ in range(max(1, workers))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    if errors:
        i, e = errors[0]
        raise RuntimeError(f"batch cell {i} failed: {e}") from e
    return BatchResult(results, wall, workers, sum(g[0] for g in graph_stats), sum(g[1] for g in graph_stats))
