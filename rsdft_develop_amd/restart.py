"""RSDFT restart files: `wf.dat1` (bands) and `vrho.dat1` (density and
potentials), read and written in the reference's record layout.

Both are Fortran unformatted sequential files: every `write(unit) ...` is
one record, framed by a 4-byte little-endian length before and after.
Integers are 4 bytes, reals 8 (or 4 for single-precision bands).

wf.dat1, new format (`write_wf_simple`, io_write_wf_simple_module.f90:153-172
for the header, :194-278 for the bands; one grid rank, IO_ctrl = 0):

    -1
    ML, ML1, ML2, ML3
    MB, MBwr1, MBwr2
    MK, MS, MMBZ
    IO_ctrl, OC, TYPE_WF              (TYPE_WF = 1: real bands)
    LL(3, ML)                          (0-based i1, i2, i3 of each point)
    occ(MB, MK, MS)
    aa(3,3), bb(3,3), kbb(3, MMBZ)
    np(0:7)
    grid%id, grid%ir ; band%id, band%ir ; bzsm%id, bzsm%ir ; spin%id, spin%ir
    one record of ML values per band MBwr1..MBwr2 (per k and spin):
    real(8), or real(4) when OC is 4 or 5

The old format (OC >= 10) has no leading -1 and stops after occ
(:174-179).

vrho.dat1 (`write_data`, io_module.f90:117-224):

    version                            (character(64))
    ML, ML1, ML2, ML3
    LL(3, ML)
    comment ; aa(3,3)
    comment ; Nelement, Natom ; aa_atom(3, Natom) ; ki_atom(Natom) ; zn_atom(Nelement)
    comment ; efermi
    per spin: rho(ML), Vloc(ML), [Vh(ML) for spin 1], Vxc(ML)

`LL` maps the file's point order to grid coordinates. With one grid rank it
is the natural order, x fastest, which is this library's grid index. The
reader places values by `LL`, so files from other rank layouts read
correctly.
"""
from __future__ import annotations

import dataclasses
import struct

import numpy as np

VERSION = "version3.0, comment_length=64"

_SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br "
            "Kr").split()


def atomic_number(symbol: str) -> int:
    """zn_atom of an element symbol (0 when unknown)."""
    return _SYMBOLS.index(symbol) + 1 if symbol in _SYMBOLS else 0


# ---------------------------------------------------------------- records
class RecordWriter:
    def __init__(self, path: str):
        self.f = open(path, "wb")

    def write(self, *arrays):
        """One record holding the given items (numpy arrays or scalars with
        their dtypes), concatenated in Fortran order."""
        payload = b"".join(np.asarray(a).tobytes(order="F") for a in arrays)
        n = len(payload)
        if n >= 2 ** 31:
            raise ValueError("record above 2 GiB")
        self.f.write(struct.pack("<i", n))
        self.f.write(payload)
        self.f.write(struct.pack("<i", n))

    def close(self):
        self.f.close()


class RecordReader:
    def __init__(self, path: str):
        self.f = open(path, "rb")

    def read(self) -> bytes:
        head = self.f.read(4)
        if len(head) < 4:
            raise EOFError("end of file")
        (n,) = struct.unpack("<i", head)
        if n < 0:
            raise ValueError("continued records (above 2 GiB) are not supported")
        data = self.f.read(n)
        (m,) = struct.unpack("<i", self.f.read(4))
        if m != n or len(data) != n:
            raise ValueError("corrupt record framing")
        return data

    def ints(self, count=None) -> np.ndarray:
        a = np.frombuffer(self.read(), dtype="<i4")
        if count is not None and a.size != count:
            raise ValueError(f"expected {count} integers, got {a.size}")
        return a

    def reals(self, dtype="<f8") -> np.ndarray:
        return np.frombuffer(self.read(), dtype=dtype)

    def close(self):
        self.f.close()


def grid_map(ngrid) -> np.ndarray:
    """LL(3, ML): 0-based (i1, i2, i3) of each point, x fastest."""
    n1, n2, n3 = ngrid
    i3, i2, i1 = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    return np.stack([i1.ravel(), i2.ravel(), i3.ravel()]).astype("<i4")


def recip(aa: np.ndarray) -> np.ndarray:
    """bb(3,3) = 2 pi (a2 x a3, a3 x a1, a1 x a2) / V, columns as in aa."""
    a1, a2, a3 = aa[:, 0], aa[:, 1], aa[:, 2]
    vol = float(np.dot(a1, np.cross(a2, a3)))
    return 2.0 * np.pi / vol * np.stack([np.cross(a2, a3), np.cross(a3, a1), np.cross(a1, a2)], axis=1)


def _to_natural(values: np.ndarray, LL: np.ndarray, ngrid) -> np.ndarray:
    """Values in file order -> natural grid order (x fastest)."""
    n1, n2, _ = ngrid
    j = LL[0].astype(np.int64) + n1 * (LL[1].astype(np.int64) + n2 * LL[2].astype(np.int64))
    out = np.empty_like(values)
    out[j] = values
    return out


# ---------------------------------------------------------------- wf.dat1
@dataclasses.dataclass
class WfData:
    ngrid: tuple
    mb: int
    band_range: tuple      # (MBwr1, MBwr2), 1-based inclusive
    occ: np.ndarray        # (MB,)
    aa: np.ndarray         # (3, 3), columns = lattice vectors
    bands: np.ndarray      # (ML, MBwr2 - MBwr1 + 1) float64, natural grid order
    oc: int = 3


def write_wf(path: str, bands: np.ndarray, occ, aa, ngrid, band_range=None, oc: int = 3):
    """Write real Gamma-point bands (ML x nb, natural order) as wf.dat1.
    oc 4 or 5 store the bands in single precision, as the reference does."""
    bands = np.asarray(bands, dtype=np.float64)
    ML = int(np.prod(ngrid))
    if bands.shape[0] != ML:
        raise ValueError("bands must have ML rows")
    mb = int(len(occ))
    lo, hi = band_range if band_range is not None else (1, bands.shape[1])
    if hi - lo + 1 != bands.shape[1] or lo < 1 or hi > mb:
        raise ValueError("band_range does not match the bands given")
    if oc >= 10:
        raise ValueError("only the new format (OC < 10) is written")
    aa = np.asarray(aa, dtype="<f8").reshape(3, 3)
    w = RecordWriter(path)
    try:
        w.write(np.int32(-1))
        w.write(np.array([ML, *ngrid], dtype="<i4"))
        w.write(np.array([mb, lo, hi], dtype="<i4"))
        w.write(np.array([1, 1, 1], dtype="<i4"))          # MK, MS, MMBZ
        w.write(np.array([0, oc, 1], dtype="<i4"))         # IO_ctrl, OC, TYPE_WF (real)
        w.write(grid_map(ngrid))
        w.write(np.asarray(occ, dtype="<f8"))
        w.write(aa, recip(aa), np.zeros(3, dtype="<f8"))   # aa, bb, kbb(3, 1)
        w.write(np.ones(8, dtype="<i4"))                   # np(0:7): one rank
        w.write(np.array([0, ML], dtype="<i4"))            # grid id, ir
        w.write(np.array([0, mb], dtype="<i4"))            # band id, ir
        w.write(np.array([0, 1], dtype="<i4"))             # bzsm id, ir
        w.write(np.array([0, 1], dtype="<i4"))             # spin id, ir
        dt = "<f4" if oc in (4, 5) else "<f8"
        for b in range(bands.shape[1]):
            w.write(bands[:, b].astype(dt))
    finally:
        w.close()


def read_wf(path: str) -> WfData:
    r = RecordReader(path)
    try:
        first = r.ints()
        if first.size == 1 and first[0] == -1:
            ML, n1, n2, n3 = r.ints(4)
            mb, lo, hi = r.ints(3)
            mk, ms, mmbz = r.ints(3)
            io_ctrl, oc, type_wf = r.ints(3)
            LL = r.ints(3 * ML).reshape(3, ML, order="F")
            occ = r.reals().reshape(mb, mk, ms, order="F")
            hdr = r.reals()
            aa = hdr[:9].reshape(3, 3, order="F").copy()
            for _ in range(5):  # np(0:7) and the four partitions
                r.read()
        else:  # old format
            if first.size != 4:
                raise ValueError("not a wf.dat1 file")
            ML, n1, n2, n3 = first
            mb, lo, hi = r.ints(3)
            mk = ms = 1
            io_ctrl, oc, type_wf = 0, 10, 1
            LL = r.ints(3 * ML).reshape(3, ML, order="F")
            occ = r.reals().reshape(mb, mk, ms, order="F")
            aa = np.zeros((3, 3))
        if type_wf != 1:
            raise ValueError("complex bands (TYPE_WF 0) are not supported: Gamma-point real bands only")
        if mk != 1 or ms != 1 or io_ctrl != 0:
            raise ValueError("one k-point, one spin and IO_ctrl 0 only")
        dt = "<f4" if oc in (4, 5) else "<f8"
        ng = (int(n1), int(n2), int(n3))
        bands = np.empty((int(ML), int(hi - lo + 1)), dtype=np.float64, order="F")
        for b in range(bands.shape[1]):
            bands[:, b] = _to_natural(r.reals(dt).astype(np.float64), LL, ng)
        return WfData(ng, int(mb), (int(lo), int(hi)), occ[:, 0, 0].copy(), aa, bands, int(oc))
    finally:
        r.close()


# ---------------------------------------------------------------- vrho.dat1
@dataclasses.dataclass
class VrhoData:
    ngrid: tuple
    aa: np.ndarray
    frac: np.ndarray       # (Natom, 3) fractional positions
    kind: np.ndarray       # (Natom,) 0-based species index
    zn: np.ndarray         # (Nelement,) atomic numbers
    efermi: float
    rho: np.ndarray        # (ML,) natural grid order
    vloc: np.ndarray
    vh: np.ndarray
    vxc: np.ndarray


def _comment(text: str) -> np.ndarray:
    return np.frombuffer(text.ljust(64)[:64].encode("ascii"), dtype="S1")


def write_vrho(path: str, d: VrhoData):
    ML = int(np.prod(d.ngrid))
    w = RecordWriter(path)
    try:
        w.write(_comment(VERSION))
        w.write(np.array([ML, *d.ngrid], dtype="<i4"))
        w.write(grid_map(d.ngrid))
        w.write(_comment("lattice aa(3,3)"))
        w.write(np.asarray(d.aa, dtype="<f8").reshape(3, 3))
        w.write(_comment("atom: Nelement Natom / aa_atom / ki_atom /zn_atom"))
        w.write(np.array([len(d.zn), len(d.kind)], dtype="<i4"))
        w.write(np.asarray(d.frac, dtype="<f8").T)          # aa_atom(3, Natom)
        w.write(np.asarray(d.kind, dtype="<i4") + 1)        # ki_atom, 1-based
        w.write(np.asarray(d.zn, dtype="<i4"))
        w.write(_comment("Fermi energy"))
        w.write(np.float64(d.efermi))
        for a in (d.rho, d.vloc, d.vh, d.vxc):                # spin 1: rho, Vloc, Vh, Vxc
            w.write(np.asarray(a, dtype="<f8"))
    finally:
        w.close()


def read_vrho(path: str) -> VrhoData:
    r = RecordReader(path)
    try:
        version = r.read().decode("ascii", "replace")
        if not version.startswith("version"):
            raise ValueError("only the versioned vrho.dat1 format is read")
        ML, n1, n2, n3 = r.ints(4)
        LL = r.ints(3 * ML).reshape(3, ML, order="F")
        r.read()
        aa = r.reals().reshape(3, 3, order="F").copy()
        r.read()
        nelem, natom = r.ints(2)
        frac = r.reals().reshape(3, natom, order="F").T.copy()
        kind = r.ints(natom) - 1
        zn = r.ints(nelem).copy()
        r.read()
        (efermi,) = r.reals()
        ng = (int(n1), int(n2), int(n3))
        fields = [_to_natural(r.reals().copy(), LL, ng) for _ in range(4)]
        return VrhoData(ng, aa, frac, kind.copy(), zn, float(efermi), *fields)
    finally:
        r.close()
