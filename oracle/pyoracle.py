"""ctypes wrapper for the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_dpc = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_fd_lap.argtypes = [_int, _dpc]
        L.orc_kinetic_coef.argtypes = [_dpc, _ip, _int, ctypes.POINTER(ctypes.c_double), _dpc, _dpc]
        L.orc_ham_create.argtypes = [_int, _int, _int, _dpc, _int]
        L.orc_ham_create.restype = _vp
        L.orc_ham_destroy.argtypes = [_vp]
        L.orc_ham_dv.argtypes = [_vp]
        L.orc_ham_dv.restype = ctypes.c_double
        L.orc_ham_set_vloc.argtypes = [_vp, _dpc]
        L.orc_ham_set_proj.argtypes = [_vp, _int, _int, _ip, _ip, _dpc, _ip]
        L.orc_ham_build_gth.argtypes = [_vp, _int, _dpc, _ip, _int, _ip, _ip, _ip, _dpc, _dpc]
        L.orc_ham_proj_sizes.argtypes = [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int)]
        L.orc_ham_proj_get.argtypes = [_vp, _ip, _ip, _dpc, _ip]
        L.orc_gth_cutoff.argtypes = [_int, _int, ctypes.c_double]
        L.orc_gth_cutoff.restype = ctypes.c_double
        L.orc_ylm.argtypes = [ctypes.c_double] * 3 + [_int, _int]
        L.orc_ylm.restype = ctypes.c_double
        for name in ("orc_op_kinetic", "orc_op_nonlocal", "orc_hpsi"):
            getattr(L, name).argtypes = [_vp, _dp, _dp, _i64, _int]
        L.orc_dot.argtypes = [_dpc, _dpc, _i64, ctypes.c_double]
        L.orc_dot.restype = ctypes.c_double
        L.orc_gram_schmidt_t.argtypes = [_dp, _i64, _i64, _int, ctypes.c_double]
        L.orc_gram_schmidt_lusl.argtypes = [_dp, _i64, _i64, _int, ctypes.c_double]
        L.orc_syev_jacobi.argtypes = [_int, _dp, _dpc]
        L.orc_potrf_upper.argtypes = [_int, _dp]
        L.orc_trtri_upper.argtypes = [_int, _dp]
        L.orc_sygv2.argtypes = [_dp, _dp, _dpc]
        L.orc_subspace_diag.argtypes = [_vp, _dp, _i64, _int, _dpc]
        L.orc_precond_hprsdft.argtypes = [_vp, _int, _dp, _dp, _i64, _int]
        L.orc_conjugate_gradient.argtypes = [_vp, _dp, _i64, _int, _int, _int, _int, _int, _dpc, _dpc]
        L.orc_conjugate_gradient_gs.argtypes = [_vp, _dp, _i64, _int, _int, _int, _int, _int, _int, _dpc, _dpc]
        L.orc_esp_calc.argtypes = [_vp, _dp, _i64, _int, _dpc]
        L.orc_dense_h.argtypes = [_vp, _dp]
        L.orc_set_threads.argtypes = [_int]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def fd_lap(md: int) -> np.ndarray:
    out = np.zeros(2 * md + 1)
    lib().orc_fd_lap(md, out)
    return out


def kinetic_coef(aa: np.ndarray, ngrid, md: int):
    c0 = ctypes.c_double()
    cl = np.zeros(3 * md)
    ggg = np.zeros(6)
    rc = lib().orc_kinetic_coef(
        np.ascontiguousarray(aa.T, dtype=np.float64), np.asarray(ngrid, dtype=np.int32), md,
        ctypes.byref(c0), cl, ggg,
    )
    return rc, c0.value, cl.reshape(3, md), ggg


class Ham:
    """Oracle Hamiltonian on one cell (grid, FD coefficients, Vloc, tables)."""

    def __init__(self, cell, vloc=None, projectors=True):
        L = lib()
        self.cell = cell
        n1, n2, n3 = cell.ngrid
        self.N = n1 * n2 * n3
        self.h = L.orc_ham_create(n1, n2, n3, np.ascontiguousarray(cell.aa.T, dtype=np.float64), cell.md)
        if not self.h:
            raise ValueError("oracle: unsupported cell (non-orthogonal or bad md)")
        if vloc is not None:
            L.orc_ham_set_vloc(self.h, np.ascontiguousarray(vloc, dtype=np.float64))
        if projectors:
            norb, lo, no, rps, hnl = cell.species_arrays()
            L.orc_ham_build_gth(
                self.h, cell.natom, np.ascontiguousarray(cell.frac, dtype=np.float64),
                np.ascontiguousarray(cell.kind, dtype=np.int32), len(cell.species), norb, lo, no, rps, hnl,
            )

    def __del__(self):
        try:
            if self.h:
                lib().orc_ham_destroy(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def dv(self) -> float:
        return lib().orc_ham_dv(self.h)

    def set_vloc(self, vloc):
        lib().orc_ham_set_vloc(self.h, np.ascontiguousarray(vloc, dtype=np.float64))

    def tables(self):
        L = lib()
        nz, mx = _int(), _int()
        L.orc_ham_proj_sizes(self.h, ctypes.byref(nz), ctypes.byref(mx))
        nz, mx = nz.value, mx.value
        mjj = np.zeros(max(nz, 1), dtype=np.int32)
        iuv = np.zeros(max(nz, 1), dtype=np.int32)
        jjp = np.zeros(max(nz * mx, 1), dtype=np.int32)
        uvk = np.zeros(max(nz * mx, 1), dtype=np.float64)
        L.orc_ham_proj_get(self.h, mjj, jjp, uvk, iuv)
        return dict(nzlma=nz, maxmjj=mx, mjj=mjj[:nz], iuv=iuv[:nz],
                    jjp=jjp[: nz * mx].reshape(nz, mx), uvk=uvk[: nz * mx].reshape(nz, mx))

    def set_tables(self, t):
        lib().orc_ham_set_proj(
            self.h, int(t["nzlma"]), int(t["maxmjj"]), np.ascontiguousarray(t["mjj"], dtype=np.int32),
            np.ascontiguousarray(t["jjp"], dtype=np.int32).ravel(),
            np.ascontiguousarray(t["uvk"], dtype=np.float64).ravel(),
            np.ascontiguousarray(t["iuv"], dtype=np.int32),
        )

    def _op(self, name, psi):
        psi = np.asfortranarray(psi, dtype=np.float64)
        if psi.ndim == 1:
            psi = psi.reshape(-1, 1, order="F")
        out = np.zeros_like(psi, order="F")
        if name == "orc_op_nonlocal":
            out[...] = 0.0
        getattr(lib(), name)(self.h, psi, out, psi.shape[0], psi.shape[1])
        return out

    def hpsi(self, psi):
        return self._op("orc_hpsi", psi)

    def kinetic_local(self, psi):
        return self._op("orc_op_kinetic", psi)

    def nonlocal_(self, psi):
        return self._op("orc_op_nonlocal", psi)

    def dense(self):
        H = np.zeros((self.N, self.N), order="F")
        lib().orc_dense_h(self.h, H)
        return H

    def gram_schmidt(self, u, variant="t"):
        u = np.array(u, dtype=np.float64, order="F", copy=True)
        fn = lib().orc_gram_schmidt_t if variant == "t" else lib().orc_gram_schmidt_lusl
        rc = fn(u, u.shape[0], u.shape[0], u.shape[1], self.dv)
        if rc:
            raise np.linalg.LinAlgError(f"oracle GS failed ({rc})")
        return u

    def subspace_diag(self, u):
        u = np.array(u, dtype=np.float64, order="F", copy=True)
        esp = np.zeros(u.shape[1])
        lib().orc_subspace_diag(self.h, u, u.shape[0], u.shape[1], esp)
        return u, esp

    def precond(self, gk, mloop=3):
        gk = np.asfortranarray(gk, dtype=np.float64)
        pgk = np.array(gk, order="F", copy=True)
        lib().orc_precond_hprsdft(self.h, mloop, gk, pgk, gk.shape[0], gk.shape[1])
        return pgk

    def conjugate_gradient(self, u, ncg=4, mbd=None, pc_kind=4, mloop=3, swgs=0):
        u = np.array(u, dtype=np.float64, order="F", copy=True)
        mb = u.shape[1]
        esp = np.zeros(mb)
        res = np.zeros(mb)
        rc = lib().orc_conjugate_gradient_gs(self.h, u, u.shape[0], mb, mbd or mb, ncg, pc_kind, mloop, swgs,
                                             esp, res)
        if rc:
            raise ValueError(f"oracle CG failed ({rc})")
        return u, esp, res

    def esp_calc(self, u):
        u = np.asfortranarray(u, dtype=np.float64)
        esp = np.zeros(u.shape[1])
        lib().orc_esp_calc(self.h, u, u.shape[0], u.shape[1], esp)
        return esp


def syev(a):
    a = np.array(a, dtype=np.float64, order="F", copy=True)
    w = np.zeros(a.shape[0])
    lib().orc_syev_jacobi(a.shape[0], a, w)
    return w, a


def potrf(a):
    a = np.array(a, dtype=np.float64, order="F", copy=True)
    rc = lib().orc_potrf_upper(a.shape[0], a)
    return rc, a


def trtri(a):
    a = np.array(a, dtype=np.float64, order="F", copy=True)
    rc = lib().orc_trtri_upper(a.shape[0], a)
    return rc, a


def sygv2(a, b):
    a = np.array(a, dtype=np.float64, order="F", copy=True)
    b = np.array(b, dtype=np.float64, order="F", copy=True)
    w = np.zeros(2)
    rc = lib().orc_sygv2(a, b, w)
    return rc, w, a
