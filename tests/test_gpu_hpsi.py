"""GPU parity of the projector tables and hpsi against the CPU oracle.

Tolerances: FP64 hpsi agrees with the oracle to 1e-12 of max|H psi|
(different summation order only); FP32 hpsi (FP32 storage, accumulation and
tables) to 2e-5 of max|H psi|.  Integer tables (MJJ, JJP, iuV) are exact.
"""
import numpy as np
import pytest

from rsdft_develop_amd import systems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_mod():
    from rsdft_develop_amd import gpu

    return gpu


def _pair(orc, gpu_mod, cell, vloc_seed=0, projectors=True):
    v = systems.model_vloc(cell, seed=vloc_seed)
    ham = orc.Ham(cell, vloc=v, projectors=projectors)
    ctx = gpu_mod.Context()
    ctx.cell = cell
    ctx.set_grid(cell.ngrid, cell.aa, cell.md)
    ctx.set_vloc(v)
    if projectors:
        ctx.build_gth_projectors(cell)
    return ham, ctx


def _close(got, want, rel):
    scale = max(np.abs(want).max(), 1e-300)
    err = np.abs(got - want).max() / scale
    assert err <= rel, f"max rel error {err:.3e} > {rel:.1e}"


@pytest.mark.parametrize("ng,disp", [(36, 0.0), (24, 0.1), (30, 0.05)])
def test_gth_tables_match_oracle(orc, gpu_mod, ng, disp):
    cell = systems.silicon_supercell(1, ngrid_per_cell=ng, displacement=disp, seed=ng)
    ham, ctx = _pair(orc, gpu_mod, cell)
    a, b = ham.tables(), ctx.projector_tables()
    assert a["nzlma"] == b["nzlma"] and a["maxmjj"] == b["maxmjj"]
    np.testing.assert_array_equal(a["mjj"], b["mjj"])
    np.testing.assert_array_equal(a["iuv"], b["iuv"])
    np.testing.assert_array_equal(a["jjp"], b["jjp"])
    np.testing.assert_allclose(b["uvk"], a["uvk"], rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("nb", [1, 7, 20])
def test_hpsi_f64_si8(orc, gpu_mod, nb):
    cell = systems.silicon_supercell(1)
    ham, ctx = _pair(orc, gpu_mod, cell)
    psi = systems.random_bands(cell.npoints, nb, seed=100 + nb)
    _close(ctx.hamiltonian(psi), ham.hpsi(psi), 1e-12)


def test_hpsi_f64_si64_block(orc, gpu_mod):
    cell = systems.silicon_supercell(2, displacement=0.05, seed=1)
    ham, ctx = _pair(orc, gpu_mod, cell)
    psi = systems.random_bands(cell.npoints, 12, seed=7)
    _close(ctx.hamiltonian(psi), ham.hpsi(psi), 1e-12)


@pytest.mark.parametrize("md", [1, 4, 6, 8])
def test_hpsi_anisotropic_grid_and_md(orc, gpu_mod, md):
    aa = np.diag([9.7, 11.3, 10.1])
    base = systems.silicon_supercell(1, displacement=0.1, seed=3)
    cell = systems.Cell(aa=aa, frac=base.frac, kind=base.kind, species=base.species,
                        ngrid=(20, 27, 23), nband=4, md=md)
    ham, ctx = _pair(orc, gpu_mod, cell)
    psi = systems.random_bands(cell.npoints, 5, seed=md)
    _close(ctx.hamiltonian(psi), ham.hpsi(psi), 1e-12)


def test_hpsi_without_projectors(orc, gpu_mod):
    cell = systems.silicon_supercell(1, ngrid_per_cell=24)
    ham, ctx = _pair(orc, gpu_mod, cell, projectors=False)
    psi = systems.random_bands(cell.npoints, 3, seed=9)
    _close(ctx.hamiltonian(psi), ham.hpsi(psi), 1e-12)


def test_hpsi_caller_tables_without_atom_map(orc, gpu_mod):
    """Tables handed over in the reference layout, one projector per group."""
    cell = systems.silicon_supercell(1, ngrid_per_cell=28, displacement=0.1, seed=5)
    ham, ctx = _pair(orc, gpu_mod, cell, projectors=True)
    ctx.set_projectors(ham.tables(), amap=None)
    psi = systems.random_bands(cell.npoints, 9, seed=11)
    _close(ctx.hamiltonian(psi), ham.hpsi(psi), 1e-12)


def test_hpsi_device_f64_and_f32_with_padded_ld(orc, gpu_mod):
    from rsdft_develop_amd import hip

    cell = systems.silicon_supercell(1)
    ham, ctx = _pair(orc, gpu_mod, cell)
    nb, N = 10, cell.npoints
    ld = N + 37
    psi = systems.random_bands(N, nb, seed=21)
    want = ham.hpsi(psi)
    host = np.zeros((ld, nb), order="F")
    host[:N] = psi
    for dt, rel in ((np.float64, 1e-12), (np.float32, 2e-5)):
        d_in = hip.DeviceArray.from_numpy(host, dtype=dt)
        d_out = hip.DeviceArray((ld, nb), dtype=dt)
        d_out.fill_bytes(0xFF)  # NaN pattern in both precisions
        ctx.hpsi_device(d_in.ptr, d_out.ptr, ld, nb, dtype=dt)
        ctx.synchronize()
        got = d_out.to_numpy().astype(np.float64)
        _close(got[:N], want, rel)
        assert np.all(np.isnan(got[N:]))  # padding rows untouched


def test_op_kinetic_is_the_local_part(orc, gpu_mod):
    """op_kinetic alone gives (T + Vloc) psi; accumulated onto V_NL psi it
    gives the whole H psi."""
    from rsdft_develop_amd import hip

    cell = systems.silicon_supercell(1, displacement=0.05, seed=3)
    ham, ctx = _pair(orc, gpu_mod, cell)
    nb, N = 7, cell.npoints
    psi = systems.random_bands(N, nb, seed=5)
    loc = ham.kinetic_local(psi)
    nl = ham.nonlocal_(psi)
    for dt, rel in ((np.float64, 1e-12), (np.float32, 2e-5)):
        d_in = hip.DeviceArray.from_numpy(psi, dtype=dt)
        d_out = hip.DeviceArray((N, nb), dtype=dt)
        d_out.fill_bytes(0xFF)
        ctx.op_kinetic_device(d_in.ptr, d_out.ptr, N, nb, dtype=dt)
        ctx.synchronize()
        _close(d_out.to_numpy().astype(np.float64), loc, rel)
        d_acc = hip.DeviceArray.from_numpy(nl, dtype=dt)
        ctx.op_kinetic_device(d_in.ptr, d_acc.ptr, N, nb, dtype=dt, accumulate=True)
        ctx.synchronize()
        _close(d_acc.to_numpy().astype(np.float64), loc + nl, rel)
        for a in (d_in, d_out, d_acc):
            a.free()


def _f32_local_and_full(ham, ctx, psi, ld, rel=2e-5):
    """op_kinetic (write and accumulate) and hpsi in FP32 against the oracle."""
    from rsdft_develop_amd import hip

    N, nb = psi.shape
    loc, full = ham.kinetic_local(psi), ham.hpsi(psi)
    host = np.zeros((ld, nb), order="F")
    host[:N] = psi
    d_in = hip.DeviceArray.from_numpy(host, dtype=np.float32)
    d_out = hip.DeviceArray((ld, nb), dtype=np.float32)
    d_out.fill_bytes(0xFF)
    ctx.op_kinetic_device(d_in.ptr, d_out.ptr, ld, nb, dtype=np.float32)
    ctx.synchronize()
    got = d_out.to_numpy().astype(np.float64)
    _close(got[:N], loc, rel)
    assert np.all(np.isnan(got[N:]))
    ctx.op_kinetic_device(d_in.ptr, d_out.ptr, ld, nb, dtype=np.float32, accumulate=True)
    ctx.synchronize()
    _close(d_out.to_numpy().astype(np.float64)[:N], 2 * loc, rel)
    ctx.hpsi_device(d_in.ptr, d_out.ptr, ld, nb, dtype=np.float32)
    ctx.synchronize()
    _close(d_out.to_numpy().astype(np.float64)[:N], full, rel)
    d_in.free()
    d_out.free()


@pytest.mark.parametrize("ng,ry,ty,zlen", [(36, 2, 0, 0), (36, 1, 0, 0), (36, 2, 12, 7), (30, 2, 0, 0),
                                           (24, 1, 8, 5), (24, 2, 24, 24)])
def test_pair_stencil_f32(orc, gpu_mod, ng, ry, ty, zlen):
    """The FP32 kernel with two x-points per thread, over tile shapes,
    z-runs that do not divide n3, and both passes of hpsi."""
    cell = systems.silicon_supercell(1, ngrid_per_cell=ng, displacement=0.05, seed=ng)
    ham, ctx = _pair(orc, gpu_mod, cell)
    ctx.set_hpsi_shape(32, ry=ry, ty=ty, zlen=zlen)
    ctx.set_hpsi_xpoints(32, 2)
    nb = 5
    sh = ctx.hpsi_shape(32, nb)
    print(f"case: pair stencil {ng}^3, shape {sh}", flush=True)
    assert sh["rx"] == 2 and sh["ry"] == ry and ng % sh["ty"] == 0
    if ty:
        assert sh["ty"] == ty
    _f32_local_and_full(ham, ctx, systems.random_bands(cell.npoints, nb, seed=ng + ry), cell.npoints)


def test_pair_stencil_falls_back_where_it_does_not_apply(orc, gpu_mod):
    """Odd n1, md other than 6, FP64, and an odd leading dimension all run
    the one-point kernel and stay exact."""
    base = systems.silicon_supercell(1, displacement=0.1, seed=3)
    for ngrid, md in (((21, 24, 22), 6), ((24, 24, 20), 4)):
        cell = systems.Cell(aa=np.diag([9.7, 11.3, 10.1]), frac=base.frac, kind=base.kind, species=base.species,
                            ngrid=ngrid, nband=4, md=md)
        ham, ctx = _pair(orc, gpu_mod, cell)
        ctx.set_hpsi_xpoints(32, 2)
        ctx.set_hpsi_xpoints(64, 2)
        assert ctx.hpsi_shape(32, 4)["rx"] == 1 and ctx.hpsi_shape(64, 4)["rx"] == 1
        psi = systems.random_bands(cell.npoints, 4, seed=md)
        _f32_local_and_full(ham, ctx, psi, cell.npoints)
        _close(ctx.hamiltonian(psi), ham.hpsi(psi), 1e-12)
    cell = systems.silicon_supercell(1, ngrid_per_cell=24, displacement=0.05, seed=8)
    ham, ctx = _pair(orc, gpu_mod, cell)
    ctx.set_hpsi_xpoints(32, 2)
    assert ctx.hpsi_shape(32, 3)["rx"] == 2
    _f32_local_and_full(ham, ctx, systems.random_bands(cell.npoints, 3, seed=4), cell.npoints + 37)


def test_hpsi_errors(gpu_mod):
    ctx = gpu_mod.Context()
    a = np.zeros((8, 1), order="F")
    with pytest.raises(gpu_mod.RSDFTGPUError):  # no grid yet -> state error
        ctx._check(ctx.L.rsg_hpsi_f64_host(ctx.h, a.ctypes.data, a.ctypes.data, 8, 1))
    with pytest.raises(gpu_mod.RSDFTGPUError):
        ctx.set_grid((8, 8, 8), np.array([[10, 1, 0], [0, 10, 0], [0, 0, 10.0]]), 6)
    with pytest.raises(ValueError):
        ctx.hamiltonian(np.zeros((8, 1)), k=2)


@pytest.mark.parametrize("nrep,ng,disp,nb", [(1, 36, 0.1, 13), (2, 36, 0.05, 19), (1, 30, 0.1, 21)])
def test_nonlocal_modes_agree_with_oracle(orc, gpu_mod, nrep, ng, disp, nb):
    """Tables (mode 1) and on-the-fly GTH projectors (mode 2), FP64 and FP32;
    the analytic projections per sphere segment with 32 or 64 bands per wave
    (proj 0, 2; the band counts leave partial waves) and by row tiles
    (proj 1, whose last tile of rows is partial on these grids)."""
    from rsdft_develop_amd import hip

    cell = systems.silicon_supercell(nrep, ngrid_per_cell=ng, displacement=disp, seed=nb)
    ham, ctx = _pair(orc, gpu_mod, cell)
    assert ctx.nonlocal_mode() == (0, True)
    psi = systems.random_bands(cell.npoints, nb, seed=300 + nb)
    want = ham.hpsi(psi)
    for mode, order, chunk, proj in ((1, 0, 0, 0), (2, 0, 0, 0), (2, 1, 0, 0), (2, 0, 4, 0), (2, 1, 4, 0),
                                     (2, 0, 0, 1), (2, 1, 4, 1), (2, 0, 17, 1), (2, 0, 0, 2), (2, 1, 4, 2)):
        ctx.set_nonlocal_mode(mode)
        for prec in (64, 32):
            ctx.set_hpsi_shape(prec, chunk=chunk)
            ctx.set_hpsi_order(prec, order)
            ctx.set_hpsi_proj(prec, proj)
            if mode == 2:
                assert ctx.hpsi_shape(prec, nb)["proj"] == proj
        print(f"case: mode {mode} order {order} chunk {chunk} proj {proj}, Si{8 * nrep ** 3}, {nb} bands", flush=True)
        _close(ctx.hamiltonian(psi), want, 1e-12)
        d_in = hip.DeviceArray.from_numpy(psi, dtype=np.float32)
        d_out = hip.DeviceArray(psi.shape, dtype=np.float32)
        ctx.hpsi_device(d_in.ptr, d_out.ptr, cell.npoints, nb, dtype=np.float32)
        ctx.synchronize()
        _close(d_out.to_numpy().astype(np.float64), want, 2e-5)
        d_in.free()
        d_out.free()


def test_projection_lds_reservation_leaves_results_unchanged(orc, gpu_mod):
    """The LDS a segment-projection wave reserves only limits how many waves
    share a CU: H psi is bit for bit the same for 0 and 32 KiB, in both
    projection forms and precisions, and equals the oracle."""
    cell = systems.silicon_supercell(1, ngrid_per_cell=36, displacement=0.1, seed=5)
    ham, ctx = _pair(orc, gpu_mod, cell)
    psi = systems.random_bands(cell.npoints, 40, seed=77)
    want = ham.hpsi(psi)
    for proj in (0, 2):
        got = {}
        for kib in (0, 32):
            for prec in (64, 32):
                ctx.set_hpsi_proj(prec, proj)
                ctx.set_hpsi_proj_lds(prec, kib)
            assert ctx.hpsi_shape(64, 40)["plds"] == kib
            print(f"case: projection form {proj}, {kib} KiB reserved", flush=True)
            got[kib] = ctx.hamiltonian(psi)
        assert np.array_equal(got[0], got[32])
        _close(got[32], want, 1e-12)
    for prec in (64, 32):
        ctx.set_hpsi_proj_lds(prec, 0)
        ctx.set_hpsi_proj(prec, 0)


def test_projection_segment_pieces_agree_with_oracle(orc, gpu_mod):
    """Sphere segments projected in 2, 4 or 8 pieces (the segments at the
    poles of a sphere span one 32-point pass, so some pieces are empty) give
    the oracle's H psi in both segment forms and precisions, and the same
    result on a second call."""
    cell = systems.silicon_supercell(1, ngrid_per_cell=36, displacement=0.1, seed=5)
    ham, ctx = _pair(orc, gpu_mod, cell)
    from rsdft_develop_amd import hip

    psi = systems.random_bands(cell.npoints, 40, seed=78)
    want = ham.hpsi(psi)
    d_in = hip.DeviceArray.from_numpy(psi, dtype=np.float32)
    d_out = hip.DeviceArray(psi.shape, dtype=np.float32)
    for proj in (0, 2):
        for pieces in (1, 2, 4, 8):
            for prec in (64, 32):
                ctx.set_hpsi_proj(prec, proj)
                ctx.set_hpsi_proj_split(prec, pieces)
            assert ctx.hpsi_shape(64, 40)["psplit"] == pieces
            print(f"case: projection form {proj}, {pieces} pieces per segment", flush=True)
            got = ctx.hamiltonian(psi)
            _close(got, want, 1e-12)
            assert np.array_equal(ctx.hamiltonian(psi), got)
            ctx.hpsi_device(d_in.ptr, d_out.ptr, cell.npoints, 40, dtype=np.float32)
            ctx.synchronize()
            _close(d_out.to_numpy().astype(np.float64), want, 2e-5)
    d_in.free()
    d_out.free()
    for prec in (64, 32):
        ctx.set_hpsi_proj_split(prec, 1)
        ctx.set_hpsi_proj(prec, 0)


@pytest.mark.parametrize("bt", [4, 8, 16, 20, 32])
def test_apply_band_tiles_agree_with_oracle(orc, gpu_mod, bt):
    """The analytic apply with 16, 32, 64, 80 and 128 bands per wave, writing
    H psi (order 0) and accumulating onto the stencil's (order 1), on a band
    count that leaves the last band tile ragged."""
    from rsdft_develop_amd import hip

    cell = systems.silicon_supercell(1, ngrid_per_cell=36, displacement=0.1, seed=bt)
    ham, ctx = _pair(orc, gpu_mod, cell)
    nb = 90
    psi = systems.random_bands(cell.npoints, nb, seed=700 + bt)
    want = ham.hpsi(psi)
    for order in (0, 1):
        for prec in (64, 32):
            ctx.set_hpsi_shape(prec, bt=bt)
            ctx.set_hpsi_order(prec, order)
        print(f"case: apply band tile bt {bt} order {order}, Si8, {nb} bands", flush=True)
        _close(ctx.hamiltonian(psi), want, 1e-12)
        d_in = hip.DeviceArray.from_numpy(psi, dtype=np.float32)
        d_out = hip.DeviceArray(psi.shape, dtype=np.float32)
        ctx.hpsi_device(d_in.ptr, d_out.ptr, cell.npoints, nb, dtype=np.float32)
        ctx.synchronize()
        _close(d_out.to_numpy().astype(np.float64), want, 2e-5)
        d_in.free()
        d_out.free()


def test_analytic_form_needs_a_cell_longer_than_two_cutoffs(orc, gpu_mod):
    # one Si atom in a 6.5 bohr cube: the p sphere (3.52 bohr) meets its own
    # periodic images, so only the tables (which fold images) apply
    sp = systems.SI_GTH
    cell = systems.Cell(aa=np.eye(3) * 6.5, frac=np.array([[0.1, 0.2, 0.3]]), kind=np.zeros(1, np.int32),
                        species=(sp,), ngrid=(24, 24, 24), nband=4)
    ham, ctx = _pair(orc, gpu_mod, cell)
    assert ctx.nonlocal_mode() == (0, False)
    psi = systems.random_bands(cell.npoints, 3, seed=5)
    _close(ctx.hamiltonian(psi), ham.hpsi(psi), 1e-12)
    ctx.set_nonlocal_mode(2)
    with pytest.raises(gpu_mod.RSDFTGPUError):
        ctx.hamiltonian(psi)
    with pytest.raises(gpu_mod.RSDFTGPUError):
        ctx.set_nonlocal_mode(3)
