"""Pins of the oracle's SCF pieces against closed forms and independent
numerics."""
import math

import numpy as np
import pytest
from scipy import integrate, special

from oracle import scf_oracle as so
from rsdft_develop_amd import scf as product_scf
from rsdft_develop_amd import systems

SI = systems.SI_GTH


def gth_local_r(r, zps, rloc, c):
    """GTH local potential in real space (GTH 1996, eq. 1)."""
    x = r / rloc
    return (-zps / r * special.erf(x / math.sqrt(2.0))
            + np.exp(-0.5 * x * x) * (c[0] + c[1] * x ** 2 + c[2] * x ** 4 + c[3] * x ** 6))


@pytest.mark.parametrize("G", [0.3, 1.0, 2.5, 4.0])
def test_gth_local_form_factor_is_fourier_transform(G):
    # 4 pi int r^2 sin(Gr)/(Gr) V(r) dr, with the Coulomb tail -Z/r handled
    # analytically: FT[-Z erf(r/(sqrt2 rl))/r] = -4 pi Z exp(-G^2 rl^2/2) / G^2
    c = (-7.336103, 0.5, -0.3, 0.1)
    short = lambda r: np.exp(-0.5 * (r / SI.rloc) ** 2) * (c[0] + c[1] * (r / SI.rloc) ** 2
                                                           + c[2] * (r / SI.rloc) ** 4 + c[3] * (r / SI.rloc) ** 6)
    num = 4 * np.pi * integrate.quad(lambda r: r * r * np.sinc(G * r / np.pi) * short(r), 0, 20, limit=400)[0]
    num += -4 * np.pi * SI.zps * math.exp(-0.5 * (G * SI.rloc) ** 2) / G ** 2
    assert so.gth_local_g(G * G, SI.zps, SI.rloc, c, 1.0) == pytest.approx(num, rel=1e-9, abs=1e-10)


def test_gth_local_g0_is_limit():
    c = SI.cloc
    g = 1e-4
    lim = so.gth_local_g(g * g, SI.zps, SI.rloc, c, 1.0) + 4 * np.pi * SI.zps / g ** 2
    assert so.gth_local_g(0.0, SI.zps, SI.rloc, c, 1.0) == pytest.approx(lim, rel=1e-6)


def test_hartree_of_plane_wave():
    cell = systems.silicon_supercell(1, ngrid_per_cell=24)
    n = 24
    i1, i2, i3 = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    k = (2, 1, 3)
    rho = so.to_flat(np.cos(2 * np.pi * (k[0] * i1 + k[1] * i2 + k[2] * i3) / n))
    vh, eh = so.hartree(cell, rho)
    G2 = (2 * np.pi / cell.aa[0, 0]) ** 2 * sum(v * v for v in k)
    np.testing.assert_allclose(vh, 4 * np.pi / G2 * rho, atol=1e-12)
    assert eh == pytest.approx(0.5 * cell.dv * rho @ (4 * np.pi / G2 * rho), rel=1e-12)


def test_pz81_potential_is_density_derivative():
    rho = np.array([1e-4, 3e-3, 0.02, 0.1, 0.5, 2.0, 20.0])
    v, ed = so.ldapz81(rho)
    h = 1e-6 * rho
    _, ep = so.ldapz81(rho + h)
    _, em = so.ldapz81(rho - h)
    np.testing.assert_allclose(v, (ep - em) / (2 * h), rtol=2e-7)
    # high-density limit: correlation is A ln rs + B + ..., exchange exact:
    # e_x = -(3/4)(3 rho/pi)^(1/3); at rho = 20, rs = 0.228
    r = np.array([20.0])
    _, e = so.ldapz81(r)
    rs = (3 / (4 * np.pi * r[0])) ** (1 / 3)
    ec = 0.0311 * math.log(rs) - 0.048 + 0.002 * rs * math.log(rs) - 0.0116 * rs
    assert e[0] / r[0] == pytest.approx(-0.75 * (3 * r[0] / np.pi) ** (1 / 3) + ec, rel=1e-12)
    # the two PZ81 branches meet at rs = 1 (to the 4-digit fit's ~6e-5)
    r1 = 3 / (4 * np.pi)
    _, a = so.ldapz81(np.array([r1 * (1 + 1e-12)]))
    _, b = so.ldapz81(np.array([r1 * (1 - 1e-12)]))
    assert a[0] == pytest.approx(b[0], rel=1e-4)


def test_ewald_simple_cubic_madelung():
    # point charges Z = 1 on a simple cubic lattice in a uniform background:
    # E / ion = -0.880059 / r_s  (Wigner-Seitz radius r_s)
    sp = systems.GTHSpecies("H", 1.0, 0.2, (0, 0, 0, 0), ())
    a = 7.0
    cell = systems.Cell(aa=np.eye(3) * a, frac=np.zeros((1, 3)), kind=np.zeros(1, np.int32),
                        species=(sp,), ngrid=(8, 8, 8), nband=1)
    rs = (3 * a ** 3 / (4 * np.pi)) ** (1 / 3)
    assert so.ewald(cell) == pytest.approx(-0.8800589 / rs, rel=1e-6)


def test_ewald_is_eta_independent():
    cell = systems.silicon_supercell(1, displacement=0.1, seed=4)
    e0 = so.ewald(cell)
    for f in (0.5, 2.0):
        assert so.ewald(cell, eta=f * np.pi / abs(np.linalg.det(cell.aa)) ** (2 / 3)) == pytest.approx(e0, abs=1e-9)


def test_pulay_first_steps_are_simple_mixing_and_finds_fixed_point():
    rng = np.random.default_rng(0)
    n = 30
    M = rng.standard_normal((n, n))
    M = 0.4 * (M + M.T) / np.linalg.norm(M + M.T, 2)
    c = rng.standard_normal(n)
    x = np.zeros(n)
    mix = so.Pulay(x, mmix=4, beta=0.3)
    x1 = mix.mix(M @ x + c)
    np.testing.assert_allclose(x1, x + 0.3 * (M @ x + c - x))
    for _ in range(40):
        x1 = mix.mix(M @ x1 + c)
    np.testing.assert_allclose(x1, np.linalg.solve(np.eye(n) - M, c), atol=1e-8)


def test_fermi_gapped_and_count():
    esp = np.array([-0.5, -0.2, 0.1, 0.3, 0.6])
    occ = so.fermi_dirac(esp, 6.0, 1e-5)
    np.testing.assert_allclose(occ, [2, 2, 2, 0, 0], atol=1e-12)
    occ2, _ = product_scf.fermi_dirac(esp, 6.0, 1e-5)
    np.testing.assert_allclose(occ2, occ, atol=1e-12)
    occ3 = so.fermi_dirac(esp, 5.0, 0.05)
    assert occ3.sum() == pytest.approx(5.0, abs=1e-9)


def test_entropy_closed_forms():
    kT = 0.01
    assert so.entropy_fd(np.array([2.0, 2.0, 0.0]), kT) == 0.0
    # one half-filled level: 2 kT [-(1/2)ln(1/2) * 2] = 2 kT ln 2
    assert so.entropy_fd(np.array([1.0]), kT) == pytest.approx(2 * kT * math.log(2), rel=1e-15)
    occ = so.fermi_dirac(np.linspace(-0.3, 0.2, 9), 9.0, kT)
    assert product_scf.entropy_fd(occ, kT) == pytest.approx(so.entropy_fd(occ, kT), rel=1e-14)


@pytest.fixture(scope="module")
def si8_scf():
    # Si8 on its 36^3 grid.  With these GTH parameters the 16th and 17th
    # Gamma levels lie in one partly filled six-fold shell, so the density
    # only settles with a finite electronic temperature (EKBT = 0.01 Ha).
    cell = systems.silicon_supercell(1, ngrid_per_cell=36)
    return cell, so.run_scf(cell, etol_per_atom=1e-8, ekbt=0.01, max_iter=60)


def test_si8_scf_converges(si8_scf):
    cell, out = si8_scf
    assert out["converged"]
    h = out["history"][-1]
    assert h["de"] < 1e-8
    assert h["dharris"] < 1e-6
    assert cell.dv * out["rho"].sum() == pytest.approx(32.0, rel=1e-12)
    assert out["occ"].sum() == pytest.approx(32.0, abs=1e-9)
    # free energy lies below the Kohn-Sham energy by the smearing entropy
    assert 0 < out["etot"] - out["fene"] < 0.01 * 20
    # Gamma-point levels of diamond: one Gamma_1 below the six X_1 states
    esp = np.sort(out["esp"])
    assert esp[1] - esp[0] > 0.05
    np.testing.assert_allclose(esp[1:7], esp[1], atol=1e-4)
    # the 16th and 17th levels (the highest occupied at zero temperature and
    # the one above it) lie in one six-fold shell, levels 15-20, which holds
    # the last 4 of the 32 electrons: the cell is metallic at Gamma, each
    # member partly filled (about 0.68 at EKBT 0.01 Ha)
    occ = out["occ"][np.argsort(out["esp"])]
    np.testing.assert_allclose(esp[14:20], esp[15], atol=1e-5)
    assert esp[14] - esp[13] > 0.05 and esp.size == 20
    assert np.all((occ[14:20] > 0.5) & (occ[14:20] < 0.9))
    assert occ[14:20].sum() == pytest.approx(32.0 - occ[:14].sum(), abs=1e-9)


def test_kerker_weights_closed_form():
    cell = systems.silicon_supercell(1, ngrid_per_cell=12)
    wq, wq2 = so.kerker_weights(cell)
    g1 = (2 * np.pi / cell.aa[0, 0]) ** 2   # |G|^2 of (1,0,0)
    b2 = ((1.5 * 0.529177) ** 2) ** 2
    assert wq2[0, 0, 0] == 1.0 and wq[0, 0, 0] == 1.0
    assert wq2[1, 0, 0] == pytest.approx(g1 / (g1 + b2), rel=1e-14)
    gmax = g1 * 3 * 6 ** 2  # corner (6,6,6) on a 12^3 grid
    b1 = 19 * g1 * gmax / (gmax - 20 * g1)
    assert wq[0, 0, 1] == pytest.approx((g1 + b1) / g1, rel=1e-14)


def test_pulay_g_reduces_to_weighted_linear_mixing_and_fixed_point():
    cell = systems.silicon_supercell(1, ngrid_per_cell=10)
    rng = np.random.default_rng(3)
    n = cell.npoints
    x0 = 0.03 + 0.01 * rng.random(n)
    mix = so.PulayG(cell, x0, 4, 0.3)
    f = 0.02 + 0.01 * rng.random(n)
    x1 = mix.mix(f)
    _, wq2 = so.kerker_weights(cell)
    want = so.to_flat(np.real(np.fft.ifftn(np.fft.fftn(so.to_3d(x0, cell.ngrid))
                                           + 0.3 * wq2 * np.fft.fftn(so.to_3d(f - x0, cell.ngrid)))))
    np.testing.assert_allclose(x1, want, atol=1e-14)
    # a contraction x -> M x + c converges to its fixed point
    c = 0.01 * rng.random(n)
    Mop = lambda x: 0.3 * x + 0.2 * np.roll(x, 1) + c
    x = x1
    for _ in range(60):
        x = mix.mix(Mop(x))
    fixed = np.linalg.solve(np.eye(n) - 0.3 * np.eye(n) - 0.2 * np.roll(np.eye(n), 1, axis=0), c)
    np.testing.assert_allclose(x, fixed, atol=1e-9)
