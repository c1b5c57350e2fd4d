"""Analytic projector set-up on the device: every array equals the host
builder's (point lists, hit words, segments and tiles exactly; the Gaussian
factor tables to the rounding of exp), and hpsi over the device-built form
equals hpsi over the host-built one and the oracle's table operator."""
import dataclasses

import numpy as np
import pytest

from rsdft_develop_amd import gpu, systems

pytestmark = pytest.mark.gpu

EXACT = ("pts", "seg_beg", "seg_end", "seg_grp", "seg_order", "grp_seg_off", "hits", "tile_off", "tile_g",
         "dd", "amp", "scale")


def _cells():
    si8 = systems.silicon_supercell(1, nband=12, displacement=0.1, seed=4)
    # non-cubic grid and cell: atoms near the cell faces wrap in every axis
    odd = dataclasses.replace(systems.silicon_supercell(1, nband=12, displacement=0.2, seed=9),
                              ngrid=(40, 36, 44))
    odd = dataclasses.replace(odd, aa=np.diag([10.26, 10.9, 11.4]))
    si64 = systems.silicon_supercell(2, nband=16, displacement=0.05, seed=0)
    return {"si8": si8, "si8_noncubic": odd, "si64": si64}


@pytest.mark.parametrize("name", ["si8", "si8_noncubic", "si64"])
def test_device_setup_equals_host_setup(name):
    cell = _cells()[name]
    print(f"case: {name} grid {cell.ngrid}", flush=True)
    got = {}
    for on_host in (True, False):
        ctx = gpu.Context(0)
        ctx.set_projector_setup(on_host)
        ctx.set_grid(cell.ngrid, cell.aa, cell.md)
        ctx.build_gth_projectors(cell)
        assert ctx.nonlocal_mode()[1], "analytic form expected"
        got[on_host] = {k: ctx.analytic_array(k) for k in gpu.Context.ANALYTIC_ARRAYS}
        ctx.close()
    h, d = got[True], got[False]
    for k in EXACT:
        assert h[k].shape == d[k].shape, k
        assert np.array_equal(h[k], d[k]), k
    for k in ("es", "ep"):
        assert np.allclose(d[k], h[k], rtol=4e-16, atol=1e-300), k


@pytest.mark.parametrize("name", ["si8_noncubic", "si64"])
def test_hpsi_over_device_setup(name, orc):
    cell = _cells()[name]
    print(f"case: hpsi {name}, device vs host set-up vs oracle", flush=True)
    v = systems.model_vloc(cell)
    psi = systems.random_bands(cell.npoints, cell.nband, seed=5)
    out = {}
    for on_host in (True, False):
        ctx = gpu.Context(0)
        ctx.set_projector_setup(on_host)
        ctx.setup(cell, v)
        out[on_host] = ctx.hamiltonian(psi)
        if not on_host:
            ctx.set_hpsi_proj(64, 1)  # row tiles built from the device arrays on first use
            out["rows"] = ctx.hamiltonian(psi)
        ctx.close()
    ref = orc.Ham(cell, vloc=v).hpsi(psi)
    scale = np.abs(ref).max()
    assert np.abs(out[False] - out[True]).max() / scale < 1e-14
    assert np.abs(out["rows"] - out[True]).max() / scale < 1e-13
    assert np.abs(out[False] - ref).max() / scale < 1e-12
