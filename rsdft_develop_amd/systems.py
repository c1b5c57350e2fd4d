"""Periodic test systems for the real8 Gamma-point inner loop.

The reference reads geometry from ``fort.970`` (``lattice_module.f90:38-46``:
``AX``, ``A1..A3``, then ``Nelem Natom Z``, then ``elem x y z``) and a GTH
pseudopotential file (``ps_gth_module.f90:16-160``: ``rloc C1..C4`` then per-l
``r_l h1``).  This module builds the same information in memory for the
silicon cells the benchmarks use: the 8-atom diamond cubic cell and its
supercells, with optional Gaussian atom displacements from a seed.
"""
from __future__ import annotations

import dataclasses

import numpy as np

BOHR_PER_SI8 = 10.26  # conventional diamond-cubic Si lattice constant (bohr)

# Fractional coordinates of the 8 diamond sites in the cubic cell.
DIAMOND_SITES = np.array(
    [
        [0.00, 0.00, 0.00],
        [0.00, 0.50, 0.50],
        [0.50, 0.00, 0.50],
        [0.50, 0.50, 0.00],
        [0.25, 0.25, 0.25],
        [0.25, 0.75, 0.75],
        [0.75, 0.25, 0.75],
        [0.75, 0.75, 0.25],
    ],
    dtype=np.float64,
)


@dataclasses.dataclass(frozen=True)
class GTHSpecies:
    """GTH/HGH pseudopotential in the diagonal (GTH) form.

    Mirrors what ``read_ps_gth`` (``ps_gth_module.f90:16-160``) stores:
    ``Zps`` (valence charge), ``rloc`` and ``C1..C4`` (local part), and per
    projector ``(l, n, r_l, h_n^l)``.  A negative ``h`` flips the projector
    sign ``iuV`` (``ps_gth_module.f90:129-138``).
    """

    name: str
    zps: float
    rloc: float
    cloc: tuple
    projectors: tuple  # ((l, n, r_l, h), ...)
    znuc: float = 0.0  # nuclear charge (Znuc of the file's first line)


# Silicon GTH (LDA) parameters as given for the benchmark cells.
SI_GTH = GTHSpecies(
    name="Si",
    zps=4.0,
    rloc=0.44,
    cloc=(-7.336103, 0.0, 0.0, 0.0),
    projectors=((0, 1, 0.422738, 5.906928), (1, 1, 0.484278, 2.727013)),
    znuc=14.0,
)


@dataclasses.dataclass
class Cell:
    """A periodic cell on a real-space grid.

    ``aa[:, i]`` is lattice vector i (bohr), ``frac[a]`` the fractional
    position of atom a, ``kind[a]`` its species index, ``ngrid`` the grid.
    """

    aa: np.ndarray
    frac: np.ndarray
    kind: np.ndarray
    species: tuple
    ngrid: tuple
    nband: int
    md: int = 6

    @property
    def natom(self) -> int:
        return int(self.frac.shape[0])

    @property
    def npoints(self) -> int:
        n1, n2, n3 = self.ngrid
        return n1 * n2 * n3

    @property
    def dv(self) -> float:
        return abs(float(np.linalg.det(self.aa))) / self.npoints

    @property
    def hgrid(self) -> np.ndarray:
        return np.linalg.norm(self.aa, axis=0) / np.array(self.ngrid, dtype=np.float64)

    @property
    def nelectron(self) -> float:
        return float(sum(self.species[k].zps for k in self.kind))

    def species_arrays(self, maxorb: int = 4):
        """Flattened per-species projector arrays ``[kind*maxorb + iorb]``."""
        nk = len(self.species)
        norb = np.zeros(nk, dtype=np.int32)
        lo = np.zeros(nk * maxorb, dtype=np.int32)
        no = np.zeros(nk * maxorb, dtype=np.int32)
        rps = np.zeros(nk * maxorb, dtype=np.float64)
        hnl = np.zeros(nk * maxorb, dtype=np.float64)
        for k, sp in enumerate(self.species):
            norb[k] = len(sp.projectors)
            for i, (l, n, r, h) in enumerate(sp.projectors):
                lo[k * maxorb + i] = l
                no[k * maxorb + i] = n
                rps[k * maxorb + i] = r
                hnl[k * maxorb + i] = h
        return norb, lo, no, rps, hnl


def silicon_supercell(
    nrep: int,
    ngrid_per_cell: int = 36,
    nband: int | None = None,
    displacement: float = 0.0,
    seed: int = 0,
    scale: float = 1.0,
    ngrid: int | None = None,
    species: GTHSpecies | None = None,
) -> Cell:
    """``nrep^3`` supercell of the 8-atom diamond Si cell.

    ``displacement`` is the standard deviation (bohr) of i.i.d. Gaussian atom
    displacements drawn from ``seed``.  ``nband`` defaults to 2.5 bands per
    atom (20 for Si8, 160 for Si64, 480 for Si216).  ``species`` replaces
    the built-in Si GTH parameters (e.g. ``pseudo.read_ps_gth(path)``).
    """
    a = BOHR_PER_SI8 * scale * nrep
    aa = np.eye(3) * a
    sites = []
    for i in range(nrep):
        for j in range(nrep):
            for k in range(nrep):
                sites.append((DIAMOND_SITES + np.array([i, j, k])) / nrep)
    frac = np.concatenate(sites, axis=0)
    if displacement > 0.0:
        rng = np.random.default_rng(seed)
        frac = frac + rng.normal(0.0, displacement, size=frac.shape) / a
    natom = frac.shape[0]
    if nband is None:
        nband = (natom * 5) // 2
    n = ngrid if ngrid is not None else ngrid_per_cell * nrep
    return Cell(
        aa=aa,
        frac=frac,
        kind=np.zeros(natom, dtype=np.int32),
        species=(species or SI_GTH,),
        ngrid=(n, n, n),
        nband=nband,
    )


def counter_bands(npoints: int, nband: int, seed: int, n0: int = 0) -> np.ndarray:
    """Uniform [0, 1) band block, column-major (npoints, nband), band index
    n0 + b: the host twin of the library's device generator
    (rsg_random_bands), SplitMix64 of seed + (band * npoints + j + 1) *
    0x9E3779B97F4A7C15 with the top 53 bits kept."""
    out = np.empty((npoints, nband), dtype=np.float64, order="F")
    j = np.arange(npoints, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for b in range(nband):
            z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (np.uint64(n0 + b) * np.uint64(npoints) + j + np.uint64(1)) \
                * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z ^= z >> np.uint64(31)
            out[:, b] = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return out


def random_bands(npoints: int, nband: int, seed: int, dtype=np.float64) -> np.ndarray:
    """Seeded i.i.d. N(0,1) band block, column-major ``(npoints, nband)``."""
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((npoints, nband)).astype(dtype))


def model_vloc(cell: Cell, seed: int = 0, amplitude: float = 0.5) -> np.ndarray:
    """A smooth periodic local potential for operator tests.

    Sum of Gaussians of depth ``amplitude`` Ha at the atom sites, folded over
    the nearest images; stands in for Vion+VH+Vxc where only the operator's
    arithmetic is being tested.
    """
    n = np.array(cell.ngrid)
    if np.any(np.abs(cell.aa - np.diag(np.diag(cell.aa))) > 0):
        raise ValueError("model_vloc: orthorhombic cells only")
    # exp(-|r|^2/2) factorises over the axes of an orthorhombic cell; each
    # atom touches only the points within 9 bohr along every axis (the
    # Gaussian is below 1e-17 of its depth beyond), nearest image per axis
    v = np.zeros(tuple(n[::-1]))  # (i3, i2, i1): x fastest
    side = np.diag(cell.aa)
    for f in cell.frac:
        facs = []
        for ax in range(3):
            h = side[ax] / n[ax]
            w = min(int(np.ceil(9.0 / h)), n[ax] // 2)
            c = int(np.floor(f[ax] * n[ax]))
            idx = np.arange(c - w, c + w + 1)
            d = idx / n[ax] - f[ax]
            d -= np.round(d)
            keep = np.unique(np.mod(idx, n[ax]), return_index=True)[1]  # windows wider than the cell
            facs.append((np.mod(idx, n[ax])[keep], np.exp(-(d[keep] * side[ax]) ** 2 / 2.0)))
        (ix, gx), (iy, gy), (iz, gz) = facs
        v[np.ix_(iz, iy, ix)] -= amplitude * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return v.ravel().copy()
