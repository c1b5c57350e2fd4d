"""Python mirror of RSDFT's inner-loop module interfaces over the gfx950 C ABI.

The method names and argument meanings follow the reference's Fortran
interfaces: ``hamiltonian(tpsi, htpsi, n, k, s)`` (hamiltonian_module.f90:
118-144), ``gram_schmidt(n0, n1, k, s)`` (gram_schmidt_module.f90:35-79),
``subspace_diag(k, s)`` (subspace_diag_module.f90:25-65),
``conjugate_gradient(...)`` (cg_module.f90:53-109) and ``esp_calc``
(esp_calc_module.f90:17-68).  The reference stops the program on fatal
errors (stop_program -> MPI_Abort); here every nonzero status of the C ABI
raises ``RSDFTGPUError`` carrying the library's message.  Only Gamma-point,
single-spin calculations exist on this path: ``k`` and ``s`` must be 1.

There is no CPU fallback: if ``lib/librsdft_gpu.so`` is missing or cannot be
loaded, importing the context raises.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "librsdft_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "rsdft_gpu.h")

RSG_ERR_NOTPD = 4
FIELD_RHO, FIELD_RHO_OUT, FIELD_VION, FIELD_VH, FIELD_VXC, FIELD_VLOC = range(6)


class RSDFTGPUError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[rsg error {code}] {msg}")
        self.code = code


_lib = None
_c_int, _c_i64, _c_dbl, _vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p


class ScfRunParams(ctypes.Structure):
    """rsg_scf_params of include/rsdft_gpu.h."""
    _fields_ = [("nband", ctypes.c_int), ("ncg", ctypes.c_int), ("pc_kind", ctypes.c_int), ("mloop", ctypes.c_int),
                ("mmix", ctypes.c_int), ("imix", ctypes.c_int), ("beta", ctypes.c_double), ("ekbt", ctypes.c_double),
                ("etol_per_atom", ctypes.c_double), ("max_iter", ctypes.c_int), ("init_width", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("mixed", ctypes.c_int), ("fp64_below", ctypes.c_double),
                ("fp32_stall", ctypes.c_int), ("fp32_store", ctypes.c_int), ("precond_bits", ctypes.c_int)]


class ScfRunSummary(ctypes.Structure):
    """rsg_scf_summary of include/rsdft_gpu.h."""
    _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int), ("fp32_iterations", ctypes.c_int),
                ("etot", ctypes.c_double), ("fene", ctypes.c_double), ("efermi", ctypes.c_double),
                ("eewald", ctypes.c_double), ("wall_s", ctypes.c_double)]
_P = ctypes.POINTER


def header_symbols() -> list:
    """Entry points declared in include/rsdft_gpu.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(rsg_[a-z0-9_]+)\s*\(", txt)))


def load_library():
    """Load the HIP library; raises (never falls back) if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run rsdft_develop_amd/build.py (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "rsg_create": (_c_int, [_P(_vp), _c_int]),
        "rsg_destroy": (_c_int, [_vp]),
        "rsg_last_error": (ctypes.c_char_p, [_vp]),
        "rsg_synchronize": (_c_int, [_vp]),
        "rsg_stream": (_vp, [_vp]),
        "rsg_set_grid": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _c_int]),
        "rsg_get_grid": (_c_int, [_vp, _vp, _vp, _vp, _vp]),
        "rsg_set_vloc": (_c_int, [_vp, _vp]),
        "rsg_set_projectors": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
        "rsg_build_gth_projectors": (_c_int, [_vp, _c_int, _vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp]),
        "rsg_set_nonlocal_mode": (_c_int, [_vp, _c_int]),
        "rsg_get_nonlocal_mode": (_c_int, [_vp, _P(_c_int), _P(_c_int)]),
        "rsg_get_projector_sizes": (_c_int, [_vp, _P(_c_int), _P(_c_int), _P(_c_i64)]),
        "rsg_get_projectors": (_c_int, [_vp, _vp, _vp, _vp, _vp]),
        "rsg_hpsi_f64": (_c_int, [_vp, _vp, _vp, _c_i64, _c_int]),
        "rsg_hpsi_f32": (_c_int, [_vp, _vp, _vp, _c_i64, _c_int]),
        "rsg_hpsi_f64_noalloc": (_c_int, [_vp, _vp, _vp, _c_i64, _c_int]),
        "rsg_random_bands": (_c_int, [_vp, _c_int, _c_int, ctypes.c_uint64]),
        "rsg_op_kinetic_f64": (_c_int, [_vp, _vp, _vp, ctypes.c_int64, _c_int, _c_int]),
        "rsg_op_kinetic_f32": (_c_int, [_vp, _vp, _vp, ctypes.c_int64, _c_int, _c_int]),
        "rsg_hpsi_f64_host": (_c_int, [_vp, _vp, _vp, _c_i64, _c_int]),
        "rsg_set_hpsi_shape": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int]),
        "rsg_get_hpsi_shape": (_c_int, [_vp, _c_int, _c_int, _P(_c_int), _P(_c_int), _P(_c_int), _P(_c_int), _P(_c_int)]),
        "rsg_autotune_hpsi": (_c_int, [_vp, _c_int, _c_int, _c_int, _P(ctypes.c_float)]),
        "rsg_set_hpsi_order": (_c_int, [_vp, _c_int, _c_int]),
        "rsg_get_hpsi_order": (_c_int, [_vp, _c_int, _P(_c_int)]),
        "rsg_set_hpsi_proj": (_c_int, [_vp, _c_int, _c_int]),
        "rsg_set_projector_setup": (_c_int, [_vp, _c_int]),
        "rsg_scf_run": (_c_int, [_vp, _vp, _vp, _vp, _vp]),
        "rsg_get_analytic_array": (_c_int, [_vp, _c_int, _vp, _P(_c_i64)]),
        "rsg_get_hpsi_proj": (_c_int, [_vp, _c_int, _P(_c_int)]),
        "rsg_set_hpsi_xpoints": (_c_int, [_vp, _c_int, _c_int]),
        "rsg_precond_hprsdft_f64": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _c_int]),
        "rsg_lpc_set_hybrid_scale": (_c_int, [_vp, _c_dbl]),
        "rsg_set_cg_swgs": (_c_int, [_vp, _c_int]),
        "rsg_set_cg_graphs": (_c_int, [_vp, _c_int]),
        "rsg_cg_graph_launches": (_c_int, [_vp, _P(_c_i64)]),
        "rsg_cg_graph_failures": (_c_int, [_vp, _P(_c_i64)]),
        "rsg_autotune_gram": (_c_int, [_vp, _c_int, _c_int, _P(ctypes.c_float), _P(_c_i64)]),
        "rsg_axpby_f32": (_c_int, [_vp, _vp, _vp, ctypes.c_float, ctypes.c_float, _c_i64]),
        "rsg_precond_hprsdft_f32": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _c_int]),
        "rsg_get_hpsi_xpoints": (_c_int, [_vp, _c_int, _c_int, _P(_c_int)]),
        "rsg_alloc_bands": (_c_int, [_vp, _c_int]),
        "rsg_upload_bands": (_c_int, [_vp, _vp, _c_i64, _c_int, _c_int]),
        "rsg_download_bands": (_c_int, [_vp, _vp, _c_i64, _c_int, _c_int]),
        "rsg_bands_device_ptr": (_c_int, [_vp, _P(_vp), _P(_c_i64)]),
        "rsg_gram_schmidt": (_c_int, [_vp, _c_int, _c_int]),
        "rsg_subspace_diag": (_c_int, [_vp, _vp]),
        "rsg_conjugate_gradient": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _vp]),
        "rsg_conjugate_gradient_f32": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _vp]),
        "rsg_esp_calc": (_c_int, [_vp, _vp]),
        "rsg_set_local_gth": (_c_int, [_vp, _c_int, _vp, _vp, _c_int, _vp, _vp, _vp]),
        "rsg_get_scf_constants": (_c_int, [_vp, _P(_c_dbl), _P(_c_dbl)]),
        "rsg_get_field": (_c_int, [_vp, _c_int, _vp]),
        "rsg_set_field": (_c_int, [_vp, _c_int, _vp]),
        "rsg_init_density": (_c_int, [_vp, _c_dbl]),
        "rsg_normalize_density": (_c_int, [_vp]),
        "rsg_calc_density": (_c_int, [_vp, _vp]),
        "rsg_hartree_xc": (_c_int, [_vp, _c_int, _P(_c_dbl), _P(_c_dbl)]),
        "rsg_update_vloc": (_c_int, [_vp]),
        "rsg_density_integrals": (_c_int, [_vp, _c_int, _P(_c_dbl), _P(_c_dbl)]),
        "rsg_mix_init": (_c_int, [_vp, _c_int, _c_dbl]),
        "rsg_mix_init_imix": (_c_int, [_vp, _c_int, _c_dbl, _c_int]),
        "rsg_mix_pulay": (_c_int, [_vp]),
        "rsg_lpc_param_count": (_c_int, []),
        "rsg_lpc_set_params": (_c_int, [_vp, _vp, _c_int]),
        "rsg_lpc_get_params": (_c_int, [_vp, _vp, _c_int]),
        "rsg_lpc_set_potential": (_c_int, [_vp, _vp]),
        "rsg_lpc_apply_f64": (_c_int, [_vp, _vp, _vp, _c_i64, _c_int]),
        "rsg_lpc_apply_f32": (_c_int, [_vp, _vp, _vp, _c_i64, _c_int]),
        "rsg_lpc_target_f32": (_c_int, [_vp, _vp, _vp, _c_int, _c_int, _c_dbl, _c_dbl]),
        "rsg_lpc_train_step": (_c_int, [_vp, _vp, _vp, _c_int, ctypes.c_float, ctypes.c_float, _P(_c_dbl), _vp]),
        "rsg_lpc_reset_optimizer": (_c_int, [_vp]),
        "rsg_band_residuals_f32": (_c_int, [_vp, _vp]),
        "rsg_lpc_fallbacks": (_c_int, [_vp, _P(_c_i64)]),
        "rsg_lpc_set_reject_limit": (_c_int, [_vp, _c_int]),
        "rsg_set_precond_storage": (_c_int, [_vp, _c_int]),
        "rsg_set_hpsi_proj_lds": (_c_int, [_vp, _c_int, _c_int]),
        "rsg_get_hpsi_proj_lds": (_c_int, [_vp, _c_int, _P(_c_int)]),
        "rsg_set_hpsi_proj_split": (_c_int, [_vp, _c_int, _c_int]),
        "rsg_get_hpsi_proj_split": (_c_int, [_vp, _c_int, _P(_c_int)]),
        "rsg_set_band_precision": (_c_int, [_vp, _c_int]),
        "rsg_get_band_precision": (_c_int, [_vp, _P(_c_int)]),
        "rsg_lpc_reject_state": (_c_int, [_vp, _P(_c_int), _P(_c_i64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class Context:
    """One GPU's operator state and band store (an ``rsg_ctx``)."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        h = _vp()
        rc = self.L.rsg_create(ctypes.byref(h), device)
        if rc != 0:
            raise RSDFTGPUError(rc, f"rsg_create(device={device}) failed")
        self.h = h
        self.N = 0
        self.mb = 0
        self.cell = None

    def close(self):
        if getattr(self, "h", None):
            self.L.rsg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            msg = self.L.rsg_last_error(self.h)
            raise RSDFTGPUError(rc, msg.decode() if msg else "")

    @staticmethod
    def _gamma_only(k: int, s: int):
        if k != 1 or s != 1:
            raise ValueError("only the Gamma point (k=1) and one spin (s=1) exist on this path")

    # ------------------------------------------------------------ setup
    def set_grid(self, ngrid, aa, md: int = 6):
        n1, n2, n3 = (int(v) for v in ngrid)
        self._check(self.L.rsg_set_grid(self.h, n1, n2, n3, _ptr(_f64(np.asarray(aa).T)), int(md)))
        self.N = n1 * n2 * n3
        self.ngrid = (n1, n2, n3)

    def grid_info(self):
        dv = _c_dbl()
        hg = np.zeros(3)
        c0 = _c_dbl()
        md = self.cell.md if self.cell else 6
        cl = np.zeros(3 * 8)
        self._check(self.L.rsg_get_grid(self.h, ctypes.byref(dv), _ptr(hg), ctypes.byref(c0), _ptr(cl)))
        return dv.value, hg, c0.value, cl[: 3 * md].reshape(3, md)

    @property
    def dv(self) -> float:
        return self.grid_info()[0]

    def set_vloc(self, vloc):
        v = _f64(vloc)
        if v.size != self.N:
            raise ValueError("Vloc must have N entries")
        self._check(self.L.rsg_set_vloc(self.h, _ptr(v)))

    def build_gth_projectors(self, cell):
        norb, lo, no, rps, hnl = cell.species_arrays()
        frac, kind = _f64(cell.frac), _i32(cell.kind)
        self._check(self.L.rsg_build_gth_projectors(
            self.h, cell.natom, _ptr(frac), _ptr(kind), len(cell.species),
            _ptr(norb), _ptr(lo), _ptr(no), _ptr(rps), _ptr(hnl)))

    def set_projectors(self, t, amap=None):
        mjj, jjp, uvk, iuv = _i32(t["mjj"]), _i32(t["jjp"]), _f64(t["uvk"]), _i32(t["iuv"])
        am = _i32(amap) if amap is not None else None
        self._check(self.L.rsg_set_projectors(
            self.h, int(t["nzlma"]), int(t["maxmjj"]), _ptr(mjj), _ptr(jjp), _ptr(uvk), _ptr(iuv),
            _ptr(am) if am is not None else None))

    def set_nonlocal_mode(self, mode: int):
        """0 automatic, 1 tables, 2 analytic GTH projectors."""
        self._check(self.L.rsg_set_nonlocal_mode(self.h, int(mode)))

    def nonlocal_mode(self) -> tuple:
        """(mode, analytic_available)."""
        m, a = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.rsg_get_nonlocal_mode(self.h, ctypes.byref(m), ctypes.byref(a)))
        return m.value, bool(a.value)

    def projector_tables(self):
        nz, mx, ne = _c_int(), _c_int(), _c_i64()
        self._check(self.L.rsg_get_projector_sizes(self.h, ctypes.byref(nz), ctypes.byref(mx), ctypes.byref(ne)))
        nz, mx = nz.value, mx.value
        mjj = np.zeros(max(nz, 1), np.int32)
        iuv = np.zeros(max(nz, 1), np.int32)
        jjp = np.zeros(max(nz * mx, 1), np.int32)
        uvk = np.zeros(max(nz * mx, 1), np.float64)
        self._check(self.L.rsg_get_projectors(self.h, _ptr(mjj), _ptr(jjp), _ptr(uvk), _ptr(iuv)))
        return dict(nzlma=nz, maxmjj=mx, nentries=ne.value, mjj=mjj[:nz], iuv=iuv[:nz],
                    jjp=jjp[: nz * mx].reshape(nz, mx), uvk=uvk[: nz * mx].reshape(nz, mx))

    def setup(self, cell, vloc=None):
        """Grid, Vloc and GTH projector tables for ``cell``."""
        self.cell = cell
        self.set_grid(cell.ngrid, cell.aa, cell.md)
        if vloc is not None:
            self.set_vloc(vloc)
        self.build_gth_projectors(cell)
        return self

    # -------------------------------------------------------------- hpsi
    def hamiltonian(self, tpsi, htpsi=None, n: int = 1, k: int = 1, s: int = 1):
        """htpsi = H tpsi for a host block ``tpsi(ML0, nb)`` (column-major)."""
        self._gamma_only(k, s)
        t = np.asfortranarray(tpsi, dtype=np.float64)
        if t.ndim == 1:
            t = t.reshape(-1, 1, order="F")
        if t.shape[0] != self.N:
            raise ValueError("tpsi must have N rows")
        out = htpsi if htpsi is not None else np.empty_like(t, order="F")
        if not (out.flags.f_contiguous and out.dtype == np.float64 and out.shape == t.shape):
            raise ValueError("htpsi must be a Fortran-ordered float64 array shaped like tpsi")
        self._check(self.L.rsg_hpsi_f64_host(self.h, _ptr(t), _ptr(out), t.shape[0], t.shape[1]))
        return out

    def hpsi_device(self, d_psi: int, d_hpsi: int, ld: int, nb: int, dtype=np.float64):
        fn = self.L.rsg_hpsi_f64 if np.dtype(dtype) == np.float64 else self.L.rsg_hpsi_f32
        self._check(fn(self.h, _vp(d_psi), _vp(d_hpsi), int(ld), int(nb)))

    def op_kinetic_device(self, d_psi: int, d_hpsi: int, ld: int, nb: int, dtype=np.float64, accumulate=False):
        """hpsi (+)= (T + Vloc) psi on device arrays (the local part of hpsi)."""
        fn = self.L.rsg_op_kinetic_f64 if dtype == np.float64 else self.L.rsg_op_kinetic_f32
        self._check(fn(self.h, d_psi, d_hpsi, int(ld), int(nb), int(bool(accumulate))))

    def precond_device(self, d_g: int, d_pg: int, ld: int, nb: int, mloop: int = 3, dtype=np.float64):
        """pg = P g, the hprsdft preconditioner on device arrays (precond_hprsdft)."""
        fn = self.L.rsg_precond_hprsdft_f64 if np.dtype(dtype) == np.float64 else self.L.rsg_precond_hprsdft_f32
        self._check(fn(self.h, int(mloop), _vp(d_g), _vp(d_pg), int(ld), int(nb)))

    def autotune_gram(self, nb: int, reps: int = 3) -> tuple:
        """Pick the K-split of the nb-band Gram products on this grid; returns
        (ms per product, points per chunk, 0 = one GEMM)."""
        ms, kc = ctypes.c_float(), _c_i64()
        self._check(self.L.rsg_autotune_gram(self.h, int(nb), int(reps), ctypes.byref(ms), ctypes.byref(kc)))
        return float(ms.value), int(kc.value)

    def lpc_set_hybrid_scale(self, scale: float):
        """pc_kind 6: P g = scale * P_hprsdft g + network(g)."""
        self._check(self.L.rsg_lpc_set_hybrid_scale(self.h, float(scale)))

    def set_cg_graphs(self, enable: bool = True):
        """Replay CG band blocks (pc_kind 0/4, SWGS 0) from captured HIP graphs."""
        self._check(self.L.rsg_set_cg_graphs(self.h, int(bool(enable))))

    def cg_graph_launches(self) -> int:
        n = _c_i64()
        self._check(self.L.rsg_cg_graph_launches(self.h, ctypes.byref(n)))
        return n.value

    def cg_graph_failures(self) -> int:
        """Graph captures the runtime refused (those blocks ran directly)."""
        n = _c_i64()
        self._check(self.L.rsg_cg_graph_failures(self.h, ctypes.byref(n)))
        return n.value

    def set_cg_swgs(self, swgs: int):
        """SWGS keyword of the CG: 1/2 project each preconditioned gradient
        against the bands before it / all other bands (cggs)."""
        self._check(self.L.rsg_set_cg_swgs(self.h, int(swgs)))

    def axpby_f32(self, d_y: int, d_x: int, a: float, b: float, n: int):
        """y = a y + b x on n device floats (context stream)."""
        self._check(self.L.rsg_axpby_f32(self.h, _vp(d_y), _vp(d_x), float(a), float(b), int(n)))

    def set_hpsi_shape(self, precision=64, ry=0, ty=0, zlen=0, chunk=0, bt=0):
        self._check(self.L.rsg_set_hpsi_shape(self.h, int(precision), int(ry), int(ty), int(zlen), int(chunk), int(bt)))

    def set_hpsi_order(self, precision=64, order=0):
        self._check(self.L.rsg_set_hpsi_order(self.h, int(precision), int(order)))

    def hpsi_order(self, precision=64) -> int:
        o = _c_int()
        self._check(self.L.rsg_get_hpsi_order(self.h, int(precision), ctypes.byref(o)))
        return o.value

    def set_hpsi_proj(self, precision=64, proj=0):
        """Analytic projections per sphere segment, 32 (0) or 64 (2) bands per
        wave, or by row tiles (1)."""
        self._check(self.L.rsg_set_hpsi_proj(self.h, int(precision), int(proj)))

    def scf_run(self, nband: int, p) -> dict:
        """A whole SCF run in native code (rsg_scf_run) with the settings of
        an scf.ScfParams (start bands drawn on the device); grid, projectors
        and local part set before.  Returns the summary, the energy after
        each iteration and the final band energies."""
        if p.start != "uniform":
            raise ValueError("the native SCF run draws its start bands on the device (start='uniform')")
        prm = ScfRunParams(int(nband), int(p.ncg), int(p.pc_kind), int(p.mloop), int(p.mmix), int(p.imix),
                           float(p.beta), float(p.ekbt), float(p.etol_per_atom), int(p.max_iter), float(p.init_width),
                           int(p.seed) & 0xFFFFFFFFFFFFFFFF, int(bool(p.mixed)), float(p.fp64_below), int(p.fp32_stall),
                           int(bool(p.fp32_store)), int(p.precond_bits))
        out = ScfRunSummary()
        hist = np.zeros(int(p.max_iter))
        esp = np.zeros(int(nband))
        self._check(self.L.rsg_scf_run(self.h, ctypes.addressof(prm), ctypes.addressof(out), _ptr(hist), _ptr(esp)))
        self.mb = int(nband)
        n = out.iterations
        return dict(converged=bool(out.converged), iterations=n, fp32_iterations=out.fp32_iterations, etot=out.etot,
                    fene=out.fene, efermi=out.efermi, eewald=out.eewald, wall_s=out.wall_s,
                    history=hist[:n].tolist(), esp=esp)

    def set_projector_setup(self, on_host: bool = False):
        """Build the analytic projector description on the host instead of
        with device kernels (next build_gth_projectors)."""
        self._check(self.L.rsg_set_projector_setup(self.h, int(bool(on_host))))

    ANALYTIC_ARRAYS = ("pts", "seg_beg", "seg_end", "seg_grp", "seg_order", "grp_seg_off", "hits", "tile_off",
                       "tile_g", "es", "ep", "dd", "amp", "scale")

    def analytic_array(self, name: str) -> np.ndarray:
        """One array of the analytic projector description (FP64 form)."""
        which = self.ANALYTIC_ARRAYS.index(name)
        n = _c_i64()
        self._check(self.L.rsg_get_analytic_array(self.h, which, None, ctypes.byref(n)))
        dt = {"pts": np.uint32, "hits": np.uint64}.get(name, np.float64 if which >= 9 else np.int32)
        out = np.zeros(max(n.value, 0), dtype=dt)
        if n.value:
            self._check(self.L.rsg_get_analytic_array(self.h, which, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)))
        return out

    def set_hpsi_proj_lds(self, precision=64, kib=0):
        """LDS (KiB) each segment-projection wave reserves: fewer waves per CU."""
        self._check(self.L.rsg_set_hpsi_proj_lds(self.h, int(precision), int(kib)))

    def set_hpsi_proj_split(self, precision=64, pieces=1):
        """Sphere segments projected in 1, 2, 4 or 8 pieces (one wave each)."""
        self._check(self.L.rsg_set_hpsi_proj_split(self.h, int(precision), int(pieces)))

    def hpsi_proj(self, precision=64) -> int:
        o = _c_int()
        self._check(self.L.rsg_get_hpsi_proj(self.h, int(precision), ctypes.byref(o)))
        return o.value

    def set_hpsi_xpoints(self, precision=64, rx=1):
        """x-points per stencil thread: 2 selects the FP32 pair kernel where it fits."""
        self._check(self.L.rsg_set_hpsi_xpoints(self.h, int(precision), int(rx)))

    def hpsi_shape(self, precision=64, nb=1) -> dict:
        v = [_c_int() for _ in range(5)]
        self._check(self.L.rsg_get_hpsi_shape(self.h, int(precision), int(nb), *[ctypes.byref(x) for x in v]))
        d = dict(zip(("ry", "ty", "zlen", "chunk", "bt"), (x.value for x in v)))
        d["order"] = self.hpsi_order(precision)
        rx = _c_int()
        self._check(self.L.rsg_get_hpsi_xpoints(self.h, int(precision), int(nb), ctypes.byref(rx)))
        d["rx"] = rx.value
        d["proj"] = self.hpsi_proj(precision)
        k = _c_int()
        self._check(self.L.rsg_get_hpsi_proj_lds(self.h, int(precision), ctypes.byref(k)))
        d["plds"] = k.value
        self._check(self.L.rsg_get_hpsi_proj_split(self.h, int(precision), ctypes.byref(k)))
        d["psplit"] = k.value
        return d

    def apply_hpsi_shape(self, precision, sh: dict, chunk=None):
        """Set a shape dict as returned by hpsi_shape (chunk None: keep sh's)."""
        self.set_hpsi_shape(precision, ry=sh["ry"], ty=sh["ty"], zlen=sh["zlen"],
                            chunk=sh["chunk"] if chunk is None else chunk, bt=sh["bt"])
        self.set_hpsi_order(precision, sh.get("order", 0))
        self.set_hpsi_xpoints(precision, sh.get("rx", 1))
        self.set_hpsi_proj(precision, sh.get("proj", 0))
        self.set_hpsi_proj_lds(precision, sh.get("plds", 0))
        self.set_hpsi_proj_split(precision, sh.get("psplit", 1))

    def autotune_hpsi(self, precision=64, nb=1, reps=5) -> float:
        ms = ctypes.c_float()
        self._check(self.L.rsg_autotune_hpsi(self.h, int(precision), int(nb), int(reps), ctypes.byref(ms)))
        return float(ms.value)

    def synchronize(self):
        self._check(self.L.rsg_synchronize(self.h))

    @property
    def stream(self) -> int:
        return int(self.L.rsg_stream(self.h) or 0)

    # -------------------------------------------------------- band store
    def alloc_bands(self, mb: int):
        self._check(self.L.rsg_alloc_bands(self.h, int(mb)))
        self.mb = int(mb)

    def upload_bands(self, u, n0: int = 0):
        u = np.asfortranarray(u, dtype=np.float64)
        self._check(self.L.rsg_upload_bands(self.h, _ptr(u), u.shape[0], int(n0), u.shape[1]))

    def random_bands(self, seed: int, n0: int = 0, nb: int | None = None):
        """Uniform [0, 1) start bands generated on the device (values equal to
        systems.counter_bands(N, nb, seed) columns n0 ..)."""
        nb = self.mb - n0 if nb is None else nb
        self._check(self.L.rsg_random_bands(self.h, int(n0), int(nb), int(seed) & 0xFFFFFFFFFFFFFFFF))

    def download_bands(self, n0: int = 0, nb: int | None = None):
        nb = self.mb - n0 if nb is None else nb
        out = np.zeros((self.N, nb), order="F")
        self._check(self.L.rsg_download_bands(self.h, _ptr(out), self.N, int(n0), int(nb)))
        return out

    def bands_device_ptr(self):
        p, ld = _vp(), _c_i64()
        self._check(self.L.rsg_bands_device_ptr(self.h, ctypes.byref(p), ctypes.byref(ld)))
        return int(p.value), int(ld.value)

    # ------------------------------------------------------ eigensolver
    def gram_schmidt(self, n0: int = 1, n1: int | None = None, k: int = 1, s: int = 1):
        self._gamma_only(k, s)
        self._check(self.L.rsg_gram_schmidt(self.h, int(n0), int(self.mb if n1 is None else n1)))

    def subspace_diag(self, k: int = 1, s: int = 1) -> np.ndarray:
        self._gamma_only(k, s)
        esp = np.zeros(self.mb)
        self._check(self.L.rsg_subspace_diag(self.h, _ptr(esp)))
        return esp

    def conjugate_gradient(self, ncg: int = 3, pc_kind: int = 4, mloop: int = 3, k: int = 1, s: int = 1,
                           precision: int = 64):
        """Band CG update; precision 32 runs the FP32-storage variant."""
        self._gamma_only(k, s)
        esp = np.zeros(self.mb)
        res = np.zeros(self.mb)
        f = {64: self.L.rsg_conjugate_gradient, 32: self.L.rsg_conjugate_gradient_f32}[precision]
        self._check(f(self.h, int(ncg), int(pc_kind), int(mloop), _ptr(esp), _ptr(res)))
        return esp, res

    def esp_calc(self, k: int = 1, s: int = 1) -> np.ndarray:
        self._gamma_only(k, s)
        esp = np.zeros(self.mb)
        self._check(self.L.rsg_esp_calc(self.h, _ptr(esp)))
        return esp

    # ------------------------------------------- density/potential half
    def set_local_gth(self, cell):
        """Vion (GTH local part), Ewald energy and electron count for ``cell``."""
        nk = len(cell.species)
        zps = np.array([sp.zps for sp in cell.species], dtype=np.float64)
        rloc = np.array([sp.rloc for sp in cell.species], dtype=np.float64)
        cloc = np.array([c for sp in cell.species for c in sp.cloc], dtype=np.float64)
        frac, kind = _f64(cell.frac), _i32(cell.kind)
        self._check(self.L.rsg_set_local_gth(self.h, cell.natom, _ptr(frac), _ptr(kind), nk,
                                             _ptr(zps), _ptr(rloc), _ptr(cloc)))

    def scf_constants(self):
        e, n = _c_dbl(), _c_dbl()
        self._check(self.L.rsg_get_scf_constants(self.h, ctypes.byref(e), ctypes.byref(n)))
        return e.value, n.value

    def get_field(self, field: int) -> np.ndarray:
        out = np.zeros(self.N)
        self._check(self.L.rsg_get_field(self.h, int(field), _ptr(out)))
        return out

    def set_field(self, field: int, values):
        v = _f64(values)
        self._check(self.L.rsg_set_field(self.h, int(field), _ptr(v)))

    def init_density(self, width: float = 1.0):
        self._check(self.L.rsg_init_density(self.h, float(width)))

    def normalize_density(self):
        self._check(self.L.rsg_normalize_density(self.h))

    def calc_density(self, occ):
        o = _f64(occ)
        self._check(self.L.rsg_calc_density(self.h, _ptr(o)))

    def hartree_xc(self, field: int):
        eh, exc = _c_dbl(), _c_dbl()
        self._check(self.L.rsg_hartree_xc(self.h, int(field), ctypes.byref(eh), ctypes.byref(exc)))
        return eh.value, exc.value

    def update_vloc(self):
        self._check(self.L.rsg_update_vloc(self.h))

    def density_integrals(self, field: int):
        a, b = _c_dbl(), _c_dbl()
        self._check(self.L.rsg_density_integrals(self.h, int(field), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def mix_init(self, mmix: int = 4, beta: float = 0.3, imix: int = 10):
        """IMIX 10: Pulay on the density; 30: Pulay on its Fourier
        coefficients with Kerker weights."""
        self._check(self.L.rsg_mix_init_imix(self.h, int(mmix), float(beta), int(imix)))

    # ---- learned CG preconditioner
    def lpc_param_count(self) -> int:
        return int(self.L.rsg_lpc_param_count())

    def lpc_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.L.rsg_lpc_set_params(self.h, p.ctypes.data, int(p.size)))

    def lpc_get_params(self) -> np.ndarray:
        p = np.zeros(self.lpc_param_count(), dtype=np.float32)
        self._check(self.L.rsg_lpc_get_params(self.h, p.ctypes.data, int(p.size)))
        return p

    def lpc_set_potential(self, vloc=None):
        """Normalised potential input (None: the context's current Vloc)."""
        if vloc is None:
            self._check(self.L.rsg_lpc_set_potential(self.h, None))
        else:
            v = _f64(vloc)
            self._check(self.L.rsg_lpc_set_potential(self.h, _ptr(v)))

    def lpc_apply_device(self, d_g: int, d_pg: int, ld: int, nb: int, dtype=np.float64):
        f = self.L.rsg_lpc_apply_f64 if dtype == np.float64 else self.L.rsg_lpc_apply_f32
        self._check(f(self.h, d_g, d_pg, int(ld), int(nb)))

    def lpc_target_device(self, d_g: int, d_t: int, nb: int, steps: int = 40, shift: float = 0.5,
                          vweight: float = 1.0):
        self._check(self.L.rsg_lpc_target_f32(self.h, d_g, d_t, int(nb), int(steps), float(shift), float(vweight)))

    def lpc_train_step(self, d_g: int, d_t: int, nsamp: int, lr: float = 0.0, want_grads: bool = False,
                       kappa: float = 1.0):
        loss = _c_dbl()
        grads = np.zeros(self.lpc_param_count(), dtype=np.float32) if want_grads else None
        self._check(self.L.rsg_lpc_train_step(self.h, d_g, d_t, int(nsamp), float(kappa), float(lr), ctypes.byref(loss),
                                              grads.ctypes.data if grads is not None else None))
        return loss.value, grads

    def lpc_reset_optimizer(self):
        self._check(self.L.rsg_lpc_reset_optimizer(self.h))

    def band_residuals_device(self, d_g: int):
        self._check(self.L.rsg_band_residuals_f32(self.h, d_g))

    def lpc_fallbacks(self) -> int:
        c = _c_i64()
        self._check(self.L.rsg_lpc_fallbacks(self.h, ctypes.byref(c)))
        return c.value

    def set_band_precision(self, bits: int):
        """Move the band store to FP32 (32) for the mixed-precision phase, or
        back to FP64 (64)."""
        self._check(self.L.rsg_set_band_precision(self.h, int(bits)))

    def band_precision(self) -> int:
        b = _c_int()
        self._check(self.L.rsg_get_band_precision(self.h, ctypes.byref(b)))
        return b.value

    def set_precond_storage(self, bits: int):
        """Internal vectors of FP32 hprsdft calls in bfloat16 (16) or FP32
        (32, the default)."""
        self._check(self.L.rsg_set_precond_storage(self.h, int(bits)))

    def lpc_set_reject_limit(self, limit: int):
        """After `limit` CG calls in a row with a rejected network output,
        pc_kind 5/6 calls run hprsdft without trying the network (0: try it
        in every call).  Setting the limit clears the streak."""
        self._check(self.L.rsg_lpc_set_reject_limit(self.h, int(limit)))

    def lpc_reject_state(self) -> tuple:
        """(current streak of rejected calls, calls run without the network)."""
        st, sk = _c_int(), _c_i64()
        self._check(self.L.rsg_lpc_reject_state(self.h, ctypes.byref(st), ctypes.byref(sk)))
        return st.value, sk.value

    def mix_pulay(self):
        self._check(self.L.rsg_mix_pulay(self.h))
