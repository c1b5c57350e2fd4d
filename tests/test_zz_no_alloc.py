"""rsg_hpsi_f64_noalloc (the hpsi entry point that allocates nothing)
against rsg_hpsi_f64 through the C ABI: two calls in a row, the second
with fewer bands, for the analytic projector path (Si8, and Si216 where the
projections of 160 bands overflow the fixed scratch, so the call runs in
chunks) and for the table path.  The per-band arithmetic does not depend on
the chunking, so the results agree bit for bit."""
import numpy as np
import pytest

from rsdft_develop_amd import gpu, hip, systems

pytestmark = pytest.mark.gpu


def _hpsi(ctx, fn, psi, nb):
    out = hip.DeviceArray((ctx.N, nb), dtype=np.float64)
    out.fill_bytes(0xFF)  # NaN pattern: every output value must be written
    ctx._check(fn(ctx.h, gpu._vp(psi.ptr), gpu._vp(out.ptr), ctx.N, nb))
    ctx.synchronize()
    got = out.to_numpy()
    out.free()
    return got


@pytest.mark.parametrize("case", ["si8_analytic", "si8_tables", "si216_chunked"])
def test_noalloc_hpsi_equals_hpsi(case):
    nrep, nb1, nb2 = (3, 160, 100) if case == "si216_chunked" else (1, 20, 7)
    print(f"case {case}: nrep {nrep} bands {nb1} then {nb2}", flush=True)
    cell = systems.silicon_supercell(nrep, displacement=0.05, seed=5)
    ctx = gpu.Context(0).setup(cell, systems.model_vloc(cell))
    try:
        if case == "si8_tables":
            ctx.set_nonlocal_mode(1)
        psi = hip.DeviceArray.from_numpy(systems.random_bands(cell.npoints, nb1, seed=11))
        for nb in (nb1, nb2):
            want = _hpsi(ctx, ctx.L.rsg_hpsi_f64, psi, nb)
            got = _hpsi(ctx, ctx.L.rsg_hpsi_f64_noalloc, psi, nb)
            assert np.all(np.isfinite(got))
            np.testing.assert_array_equal(got, want)
        psi.free()
    finally:
        ctx.close()
