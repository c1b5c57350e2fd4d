"""The benchmark configuration at full size: Si64 (2x2x2, 72^3 grid),
160 bands, Md = 6, analytic GTH projectors on the matrix cores.

Direct parity on a sample of bands (the oracle's hpsi costs a few ms per
band at this grid) and size-independent properties over all 160:

- hpsi of sampled bands against the oracle: 1e-12 of max|H psi| in FP64,
  2e-5 in FP32 (FP32 storage, accumulation and tables);
- X^T H X over all bands is symmetric to 1e-12 of its largest entry (FP64);
- hpsi is exactly linear under a power-of-two scale (bitwise equal);
- the hprsdft preconditioner on sampled bands against the oracle, FP64
  1e-12 / FP32 5e-5 of each band's largest value.
"""
import numpy as np
import pytest

from rsdft_develop_amd import systems

pytestmark = pytest.mark.gpu

SAMPLE = [0, 57, 113, 159]


@pytest.fixture(scope="module")
def full(orc):
    from rsdft_develop_amd import gpu

    cell = systems.silicon_supercell(2, nband=160, displacement=0.05, seed=0)
    vloc = systems.model_vloc(cell)
    ctx = gpu.Context().setup(cell, vloc)
    ham = orc.Ham(cell, vloc=vloc)
    psi = systems.random_bands(cell.npoints, cell.nband, seed=2024)
    return cell, ctx, ham, psi


def _hpsi(ctx, psi, dtype):
    from rsdft_develop_amd import hip

    N, nb = psi.shape
    d_in = hip.DeviceArray.from_numpy(psi, dtype=dtype)
    d_out = hip.DeviceArray((N, nb), dtype=dtype)
    ctx.hpsi_device(d_in.ptr, d_out.ptr, N, nb, dtype=dtype)
    ctx.synchronize()
    out = d_out.to_numpy().astype(np.float64)
    d_in.free()
    d_out.free()
    return out


def test_hpsi_sampled_bands_match_oracle(full):
    cell, ctx, ham, psi = full
    want = ham.hpsi(psi[:, SAMPLE])
    for dt, rel in ((np.float64, 1e-12), (np.float32, 2e-5)):
        print(f"case: full-size hpsi {np.dtype(dt).name}", flush=True)
        got = _hpsi(ctx, psi, dt)[:, SAMPLE]
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= rel, err


def test_subspace_matrix_is_symmetric(full):
    cell, ctx, ham, psi = full
    print("case: full-size X^T H X symmetry", flush=True)
    hx = _hpsi(ctx, psi, np.float64)
    s = psi.T @ hx
    assert np.abs(s - s.T).max() <= 1e-12 * np.abs(s).max()


def test_hpsi_is_exactly_linear_under_power_of_two_scale(full):
    cell, ctx, ham, psi = full
    print("case: full-size hpsi scaling", flush=True)
    cols = psi[:, :16]
    a = _hpsi(ctx, cols, np.float64)
    b = _hpsi(ctx, 4.0 * cols, np.float64)
    np.testing.assert_array_equal(b, 4.0 * a)


def test_preconditioner_sampled_bands_match_oracle(full):
    from rsdft_develop_amd import hip

    cell, ctx, ham, psi = full
    N = cell.npoints
    g = psi[:, SAMPLE] - 0.5
    want = ham.precond(g, mloop=3)
    for dt, rel in ((np.float64, 1e-12), (np.float32, 5e-5)):
        print(f"case: full-size preconditioner {np.dtype(dt).name}", flush=True)
        d_g = hip.DeviceArray.from_numpy(g, dtype=dt)
        d_p = hip.DeviceArray(g.shape, dtype=dt)
        ctx.precond_device(d_g.ptr, d_p.ptr, N, len(SAMPLE), mloop=3, dtype=dt)
        ctx.synchronize()
        got = d_p.to_numpy().astype(np.float64)
        d_g.free()
        d_p.free()
        for b in range(len(SAMPLE)):
            err = np.abs(got[:, b] - want[:, b]).max() / np.abs(want[:, b]).max()
            assert err <= rel, (b, err)
