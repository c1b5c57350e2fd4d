"""Kohn-Sham SCF driver over the C ABI, in the order of RSDFT's calc_scf.

One iteration (scf_module.f90:183-481, defaults Ndiag = 1, second_diag = 2):
subspace_diag -> conjugate_gradient -> gram_schmidt -> esp_calc -> Fermi
occupations -> Harris energy with the input density -> output density ->
Hartree and XC of the output density -> Kohn-Sham total energy -> convergence
test -> Pulay mixing of the density -> Hartree/XC of the mixed density ->
new Vloc.  Bands, densities and potentials stay on the GPU; only band
energies, occupations and scalars cross to the host.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import gpu, systems


@dataclasses.dataclass
class ScfParams:
    ncg: int = 4            # NCG
    pc_kind: int = 4        # IPC: 4 = hprsdft, 0 = none, 5 = learned (parameters set on the context)
    mloop: int = 3          # MLOOP
    mmix: int = 4           # MMIX
    imix: int = 10          # IMIX: 10 Pulay on rho, 30 Pulay in G space with Kerker weights
    beta: float = 0.3       # BETA
    ekbt: float = 1e-5      # EKBT (Fermi-Dirac temperature, Ha)
    etol_per_atom: float = 1e-8   # stop when |Etot_k - Etot_(k-1)| / natom < etol
    max_iter: int = 200     # DITER
    init_width: float = 1.0  # width (bohr) of the initial atomic Gaussians
    seed: int = 0           # seed of the initial bands
    start: str = "normal"   # initial bands: "normal" N(0,1) from the host, "uniform" [0, 1) drawn on the device
                            # (random_initial_wf, wf_module.f90:166-191)
    mixed: bool = False     # FP32 CG/hpsi while dE/atom >= fp64_below, FP64 after; the stop test
                            # needs dE/atom < etol between two FP64 iterations
    fp64_below: float = 0.0  # 0: twice etol_per_atom (FP32 energies follow FP64 to well below 1e-6 Ha/atom)
    fp32_stall: int = 3      # ... or after this many FP32 iterations without a new lowest dE/atom
    precond_bits: int = 32  # storage of the FP32 hprsdft calls' internal vectors: 32, or 16 (bfloat16)
    fp32_store: bool = True  # mixed: FP32 iterations keep the bands in the FP32 store (GS through FP64)
    lpc_reject_limit: int = 0  # pc_kind 5/6: after this many CG calls in a row with the network's output
                               # rejected, run hprsdft without trying it (0: try it in every call)


@dataclasses.dataclass
class ScfIteration:
    iter: int
    etot: float
    eharris: float
    fene: float             # Etot - Eentropy
    de_per_atom: float      # |Etot_k - Etot_(k-1)| / natom (stop test)
    dharris_per_atom: float  # |Etot - Eharris| / natom (scf_module.f90:335-339)
    esp: np.ndarray
    times: dict
    precision: int = 64     # CG/hpsi precision of this iteration


@dataclasses.dataclass
class ScfResult:
    converged: bool
    iterations: list
    etot: float
    fene: float
    esp: np.ndarray
    occ: np.ndarray
    eewald: float
    wall: float
    setup: float = 0.0      # set-up seconds before the first iteration (tables, density, start bands, GS)
    setup_parts: dict = dataclasses.field(default_factory=dict)  # the same by part

    @property
    def niter(self) -> int:
        return len(self.iterations)


def fermi_dirac(esp: np.ndarray, nelectron: float, kT: float, nspin: int = 1):
    """Occupations 2/nspin * f((e - mu)/kT) with mu by bisection so that the
    occupations sum to the electron count (fermi_dirac in fermi_module.f90,
    mp_version = -1).  Returns (occ, mu)."""
    esp = np.asarray(esp, dtype=np.float64)
    w = 2.0 / nspin

    def count(mu):
        x = np.clip((esp - mu) / kT, -700.0, 700.0)
        return np.sum(w / (1.0 + np.exp(x)))

    lo, hi = esp.min() - 50.0 * kT - 1.0, esp.max() + 50.0 * kT + 1.0
    if count(hi) < nelectron - 1e-12:
        raise ValueError("not enough bands for the electron count")
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if count(mid) > nelectron:
            hi = mid
        else:
            lo = mid
        if hi - lo < 1e-15 * max(1.0, abs(mid)):
            break
    mu = 0.5 * (lo + hi)
    x = np.clip((esp - mu) / kT, -700.0, 700.0)
    return w / (1.0 + np.exp(x)), mu


def entropy_fd(occ: np.ndarray, kT: float, nspin: int = 1) -> float:
    """Eentropy = (2/nspin) kT sum_n [-f ln f - (1-f) ln(1-f)] with
    f = occ * nspin/2 (entropy_fd in fermi_dirac_module.f90:325-350)."""
    f = np.asarray(occ, dtype=np.float64) * (nspin / 2.0)
    g = 1.0 - f
    s = -np.sum(np.where(f > 0, f * np.log(np.where(f > 0, f, 1.0)), 0.0))
    s -= np.sum(np.where(g > 0, g * np.log(np.where(g > 0, g, 1.0)), 0.0))
    return float((2.0 / nspin) * kT * s)


class SCF:
    """Plain FP64 SCF of one cell on one GPU."""

    def __init__(self, cell: systems.Cell, params: ScfParams | None = None, ctx: gpu.Context | None = None,
                 bands: np.ndarray | None = None, rho: np.ndarray | None = None):
        t_setup = time.perf_counter()
        parts = {}
        mark = [t_setup]

        def lap(name):  # wall seconds of each set-up part, after the device finished it
            self.ctx.synchronize()
            t = time.perf_counter()
            parts[name] = t - mark[0]
            mark[0] = t

        self.cell = cell
        self.p = params or ScfParams()
        self.ctx = ctx or gpu.Context(0)
        c = self.ctx
        c.cell = cell
        c.set_grid(cell.ngrid, cell.aa, cell.md)
        lap("grid")
        c.build_gth_projectors(cell)
        lap("projectors")
        c.set_local_gth(cell)
        self.eewald, self.nelectron = c.scf_constants()
        lap("local_potential_ewald")
        if rho is None:
            c.init_density(self.p.init_width)
        else:  # restart density (IC with vrho.dat1)
            c.set_field(gpu.FIELD_RHO, rho)
        self.eh_in, self.exc_in = c.hartree_xc(gpu.FIELD_RHO)
        c.update_vloc()
        if self.p.pc_kind in (5, 6):
            c.lpc_set_potential()  # the network's potential input follows Vloc
            c.lpc_set_reject_limit(self.p.lpc_reject_limit)  # also starts a new streak
        lap("density_potential")
        c.set_precond_storage(self.p.precond_bits)
        c.alloc_bands(cell.nband)
        if bands is None and self.p.start == "uniform":
            c.random_bands(self.p.seed)
        else:
            if self.p.start not in ("normal", "uniform"):
                raise ValueError(f"start must be 'normal' or 'uniform', got {self.p.start!r}")
            u0 = bands if bands is not None else systems.random_bands(cell.npoints, cell.nband, seed=self.p.seed)
            c.upload_bands(u0)
        lap("start_bands")
        c.gram_schmidt(1, cell.nband)   # main.f90:392
        c.mix_init(self.p.mmix, self.p.beta, self.p.imix)
        lap("first_gram_schmidt")
        self.setup_s = time.perf_counter() - t_setup
        self.setup_parts = parts

    def _band_store(self, bits: int):
        """FP32 iterations of a mixed run on the FP32 band store, FP64 ones
        (and everything after the run) on the FP64 store."""
        if self.p.mixed and self.p.fp32_store and getattr(self, "_store_bits", 64) != bits:
            self.ctx.set_band_precision(bits)
            self._store_bits = bits

    def step(self, it: int, e_prev: float | None, precision: int = 64) -> ScfIteration:
        c, p = self.ctx, self.p
        self._band_store(precision)
        t = {}
        t0 = time.perf_counter()
        c.subspace_diag()
        t["sd"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        c.conjugate_gradient(p.ncg, p.pc_kind, p.mloop, precision=precision)
        t["cg"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        c.gram_schmidt()
        t["gs"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        esp = c.esp_calc()
        t["esp"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        occ, self.efermi = fermi_dirac(esp, self.nelectron, p.ekbt)
        eeig = float(np.sum(occ * esp))
        eloc_in, eion_in = c.density_integrals(gpu.FIELD_RHO)
        eharris = eeig - eloc_in + self.eh_in + self.exc_in + eion_in + self.eewald
        c.calc_density(occ)
        eh, exc = c.hartree_xc(gpu.FIELD_RHO_OUT)
        eloc, eion = c.density_integrals(gpu.FIELD_RHO_OUT)
        etot = eeig - eloc + eh + exc + eion + self.eewald
        t["eto"] = time.perf_counter() - t0
        de = abs(etot - e_prev) / self.cell.natom if e_prev is not None else np.inf
        dh = abs(etot - eharris) / self.cell.natom
        fene = etot - entropy_fd(occ, p.ekbt)
        self.occ = occ
        return ScfIteration(it, etot, eharris, fene, de, dh, esp, t, precision)

    @classmethod
    def from_restart(cls, cell: systems.Cell, params: ScfParams | None = None, ctx: gpu.Context | None = None,
                     wf: str | None = None, vrho: str | None = None) -> "SCF":
        """Start from the reference's restart files: bands from wf.dat1 and/or
        the density from vrho.dat1 (read_data, io_module.f90:267-491)."""
        from . import restart

        bands = rho = None
        if wf is not None:
            w = restart.read_wf(wf)
            if w.ngrid != tuple(cell.ngrid) or w.band_range != (1, cell.nband) or w.mb != cell.nband:
                raise ValueError(f"{wf}: grid {w.ngrid} / bands {w.band_range} do not match the cell")
            bands = w.bands
        if vrho is not None:
            v = restart.read_vrho(vrho)
            if v.ngrid != tuple(cell.ngrid):
                raise ValueError(f"{vrho}: grid {v.ngrid} does not match the cell")
            rho = v.rho
        return cls(cell, params, ctx=ctx, bands=bands, rho=rho)

    def save(self, wf: str | None = None, vrho: str | None = None, oc: int = 3):
        """Write the current bands (wf.dat1) and the density of those bands
        with its Hartree, XC and local potentials (vrho.dat1), in the
        reference's formats (write_data, io_module.f90:59-263)."""
        from . import restart

        c, cell = self.ctx, self.cell
        if wf is not None:
            occ = getattr(self, "occ", np.zeros(cell.nband))
            restart.write_wf(wf, c.download_bands(), occ, cell.aa, cell.ngrid, oc=oc)
        if vrho is not None:
            c.hartree_xc(gpu.FIELD_RHO_OUT)  # VH, VXC of the bands' density
            vion, vh, vxc = (c.get_field(f) for f in (gpu.FIELD_VION, gpu.FIELD_VH, gpu.FIELD_VXC))
            mu = getattr(self, "efermi", 0.0)
            zn = np.array([restart.atomic_number(sp.name) for sp in cell.species])
            restart.write_vrho(vrho, restart.VrhoData(tuple(cell.ngrid), cell.aa, cell.frac, cell.kind, zn, mu,
                                                      c.get_field(gpu.FIELD_RHO_OUT), vion + vh + vxc, vh, vxc))

    def update_potential(self) -> float:
        t0 = time.perf_counter()
        c = self.ctx
        c.mix_pulay()
        c.normalize_density()
        self.eh_in, self.exc_in = c.hartree_xc(gpu.FIELD_RHO)
        c.update_vloc()
        if self.p.pc_kind in (5, 6):
            c.lpc_set_potential()
        return time.perf_counter() - t0

    def refine_levels(self, tol: float = 1e-11, max_passes: int = 60, precision: int = 64) -> tuple:
        """Band energies of the final Hamiltonian converged at a fixed
        potential: passes of subspace_diag, conjugate_gradient, gram_schmidt
        and esp_calc (the eigensolver half of an SCF iteration,
        scf_module.f90:231-268) until no band energy moves by more than tol.
        The SCF's stop test bounds the energy change only; members of a
        degenerate shell stay split by the eigenvectors' remaining error.
        Returns (esp, passes)."""
        c, p = self.ctx, self.p
        esp = c.esp_calc()
        for k in range(1, max_passes + 1):
            c.subspace_diag()
            c.conjugate_gradient(p.ncg, p.pc_kind, p.mloop, precision=precision)
            c.gram_schmidt()
            e = c.esp_calc()
            moved = float(np.abs(e - esp).max())
            esp = e
            if moved < tol:
                return esp, k
        return esp, max_passes

    def run(self, callback=None) -> ScfResult:
        its = []
        e_prev = None
        conv = False
        w0 = time.perf_counter()
        prec = 32 if self.p.mixed else 64
        switch_below = self.p.fp64_below if self.p.fp64_below > 0 else 2.0 * self.p.etol_per_atom
        best, stall = np.inf, 0
        prev_prec = None
        for it in range(1, self.p.max_iter + 1):
            rec = self.step(it, e_prev, prec)
            # FP32 is never trusted for the stop test: switch first, also
            # when FP32 stops making progress (its rounding floor)
            if prec == 32:
                if rec.de_per_atom < best:
                    best, stall = rec.de_per_atom, 0
                else:
                    stall += 1
                if rec.de_per_atom < switch_below or stall >= self.p.fp32_stall:
                    prec = 64
            # the stop test compares two FP64 energies: the first FP64
            # iteration's dE is measured against an FP32 one
            conv = rec.precision == 64 and prev_prec == 64 and rec.de_per_atom < self.p.etol_per_atom
            prev_prec = rec.precision
            if not conv and it < self.p.max_iter:
                rec.times["pot"] = self.update_potential()
            its.append(rec)
            if callback:
                callback(rec)
            e_prev = rec.etot
            if conv:
                break
        if getattr(self, "_store_bits", 64) != 64:
            self._band_store(64)
        wall = time.perf_counter() - w0
        return ScfResult(conv, its, its[-1].etot, its[-1].fene, its[-1].esp, self.occ, self.eewald, wall, self.setup_s,
                         dict(getattr(self, "setup_parts", {})))
