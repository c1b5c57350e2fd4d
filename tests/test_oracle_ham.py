"""Pins of the oracle's operator pieces against closed forms and independent
constructions (no reference run is possible: the reference needs MPI,
ScaLAPACK and FFTW, none of which exist here)."""
import math

import numpy as np
import pytest
from scipy import special

from rsdft_develop_amd import systems


def closed_form_lap(md):
    lap = np.zeros(2 * md + 1)
    for m in range(1, md + 1):
        v = 2.0 * (-1) ** (m + 1) * math.factorial(md) ** 2 / (
            m * m * math.factorial(md - m) * math.factorial(md + m))
        lap[md + m] = lap[md - m] = v
    lap[md] = -2.0 * sum(1.0 / (m * m) for m in range(1, md + 1))
    return lap


@pytest.mark.parametrize("md", [1, 2, 3, 4, 6, 8])
def test_fd_weights_closed_form(orc, md):
    lap = orc.fd_lap(md)
    np.testing.assert_allclose(lap, closed_form_lap(md), rtol=1e-13, atol=1e-15)
    assert abs(lap.sum()) < 1e-12


def test_fd_weights_known_tables(orc):
    np.testing.assert_allclose(orc.fd_lap(1), [1.0, -2.0, 1.0], atol=1e-15)
    np.testing.assert_allclose(orc.fd_lap(2), [-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12], atol=1e-15)


def test_kinetic_coef_scaling(orc):
    aa = np.diag([10.0, 12.0, 9.0])
    ng = (20, 24, 18)
    rc, c0, cl, ggg = orc.kinetic_coef(aa, ng, 6)
    assert rc == 0
    np.testing.assert_allclose(ggg, [1, 1, 1, 0, 0, 0], atol=1e-14)
    h = np.array([10.0 / 20, 12.0 / 24, 9.0 / 18])
    lap = closed_form_lap(6)
    np.testing.assert_allclose(cl, np.outer(-0.5 / h ** 2, lap[7:]), rtol=1e-13)
    assert c0 == pytest.approx(lap[6] * np.sum(-0.5 / h ** 2), rel=1e-13)


def test_kinetic_rejects_non_orthogonal(orc):
    aa = np.array([[10.0, 1.0, 0.0], [0.0, 10.0, 0.0], [0.0, 0.0, 10.0]])
    rc, *_ = orc.kinetic_coef(aa, (16, 16, 16), 6)
    assert rc == 1


def _cell(ng, aa=None, md=6, nband=4):
    aa = np.diag([10.26, 10.26, 10.26]) if aa is None else aa
    return systems.Cell(aa=aa, frac=np.zeros((0, 3)), kind=np.zeros(0, dtype=np.int32),
                        species=(systems.SI_GTH,), ngrid=ng, nband=nband, md=md)


@pytest.mark.parametrize("md", [4, 6])
def test_kinetic_plane_wave_symbol(orc, md):
    """T cos(G.r) = lambda cos(G.r) with the FD symbol of each axis."""
    ng = (16, 20, 12)
    aa = np.diag([9.0, 11.0, 7.5])
    cell = _cell(ng, aa, md)
    ham = orc.Ham(cell, vloc=np.zeros(cell.npoints), projectors=False)
    h = aa.diagonal() / np.array(ng)
    lap = closed_form_lap(md)
    i1, i2, i3 = np.meshgrid(*[np.arange(n) for n in ng], indexing="ij")
    for k in [(1, 0, 0), (0, 3, 0), (2, 1, 5), (8, 10, 6)]:
        th = [2 * np.pi * k[a] / ng[a] for a in range(3)]
        psi = np.cos(th[0] * i1 + th[1] * i2 + th[2] * i3).ravel(order="F")
        lam = 0.0
        for a in range(3):
            sym = lap[md] + 2 * sum(lap[md + m] * np.cos(m * th[a]) for m in range(1, md + 1))
            lam += -0.5 * sym / h[a] ** 2
        out = ham.hpsi(psi)[:, 0]
        np.testing.assert_allclose(out, lam * psi, atol=1e-10 * max(1.0, abs(lam)))


def test_local_potential_is_diagonal(orc):
    cell = _cell((12, 12, 12))
    rng = np.random.default_rng(3)
    v = rng.standard_normal(cell.npoints)
    psi = rng.standard_normal((cell.npoints, 3))
    h0 = orc.Ham(cell, vloc=np.zeros(cell.npoints), projectors=False)
    h1 = orc.Ham(cell, vloc=v, projectors=False)
    np.testing.assert_allclose(h1.hpsi(psi) - h0.hpsi(psi), v[:, None] * psi, atol=1e-12)


def gth_radial(r, l, n, rl):
    """GTH projector radial part (Goedecker-Teter-Hutter 1996, eq. 3)."""
    e = l + (4 * n - 1) / 2.0
    return np.sqrt(2.0) * r ** (l + 2 * (n - 1)) * np.exp(-r * r / (2 * rl * rl)) / (
        rl ** e * np.sqrt(special.gamma(e)))


def test_gth_cutoff_radius(orc):
    for (l, n, rl, h) in systems.SI_GTH.projectors:
        r = np.arange(1, 3001) * 0.01
        v = gth_radial(r, l, n, rl)
        want = r[np.argmax(np.abs(v) < 1e-10)]
        assert orc.lib().orc_gth_cutoff(l, n, rl) == pytest.approx(want, abs=1e-12)
    assert orc.lib().orc_gth_cutoff(0, 1, 0.422738) == pytest.approx(2.98, abs=1e-9)
    assert orc.lib().orc_gth_cutoff(1, 1, 0.484278) == pytest.approx(3.52, abs=1e-9)


def test_ylm_addition_theorem(orc):
    rng = np.random.default_rng(7)
    for _ in range(20):
        a, b = rng.standard_normal(3), rng.standard_normal(3)
        cg = a @ b / np.linalg.norm(a) / np.linalg.norm(b)
        for L in range(4):
            s = sum(orc.lib().orc_ylm(*a, L, m) * orc.lib().orc_ylm(*b, L, m) for m in range(-L, L + 1))
            assert s == pytest.approx((2 * L + 1) / (4 * np.pi) * special.eval_legendre(L, cg), abs=1e-12)


def independent_vnl(cell, psi):
    """V_NL psi from the GTH closed form and the addition theorem:
    V(r,r') = dV sum_l h_l p_l(r) p_l(r') (2l+1)/(4pi) P_l(cos gamma), summed
    over periodic images, restricted to the projector cutoff spheres."""
    from oracle.pyoracle import lib
    n1, n2, n3 = cell.ngrid
    g = np.stack(np.meshgrid(np.arange(n1) / n1, np.arange(n2) / n2, np.arange(n3) / n3,
                             indexing="ij"), -1).reshape(-1, 3, order="F")
    out = np.zeros_like(psi)
    for f in cell.frac:
        for (l, n, rl, h) in cell.species[0].projectors:
            rc = lib().orc_gth_cutoff(l, n, rl)
            rows, vecs = [], []
            for s in np.ndindex(3, 3, 3):
                d = (g - f + np.array(s) - 1) @ cell.aa.T
                r2 = np.sum(d * d, axis=1)
                idx = np.nonzero(r2 <= rc * rc + 1e-10)[0]
                if idx.size:
                    rows.append(idx)
                    vecs.append(d[idx])
            if not rows:
                continue
            idx = np.concatenate(rows)
            d = np.concatenate(vecs)
            r = np.linalg.norm(d, axis=1)
            p = gth_radial(r, l, n, rl)
            if l == 0:
                B = (p / np.sqrt(4 * np.pi))[:, None]
            else:
                with np.errstate(invalid="ignore", divide="ignore"):
                    rh = np.where(r[:, None] > 0, d / r[:, None], 0.0)
                B = np.sqrt(3 / (4 * np.pi)) * p[:, None] * rh
            # fold images onto grid points
            Bg = np.zeros((cell.npoints, B.shape[1]))
            np.add.at(Bg, idx, B)
            out += cell.dv * h * Bg @ (Bg.T @ psi)
    return out


def test_nonlocal_matches_addition_theorem(orc):
    cell = systems.silicon_supercell(1, ngrid_per_cell=20, displacement=0.1, seed=11)
    ham = orc.Ham(cell, vloc=np.zeros(cell.npoints))
    psi = systems.random_bands(cell.npoints, 3, seed=5)
    want = independent_vnl(cell, psi)
    got = ham.nonlocal_(psi)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_tables_sorted_counts_and_size(orc):
    cell = systems.silicon_supercell(1)  # Si8 on 36^3
    ham = orc.Ham(cell, vloc=np.zeros(cell.npoints))
    t = ham.tables()
    assert t["nzlma"] == 4 * cell.natom
    assert t["maxmjj"] == t["mjj"].max()  # leading dim = maxval(MJJ) after the drop
    E = int(t["mjj"].sum())
    assert 4.5 < E / cell.npoints < 4.8  # ~27k entries per Si atom at h = 0.285
    for lma in range(t["nzlma"]):
        j = t["jjp"][lma, : t["mjj"][lma]]
        assert np.all(np.diff(j) > 0) and j.min() >= 0 and j.max() < cell.npoints
    assert np.all(t["iuv"] == 1)
    # s channel: every grid point inside the cutoff sphere is present
    g = np.stack(np.meshgrid(*[np.arange(36) / 36] * 3, indexing="ij"), -1).reshape(-1, 3, order="F")
    d = g - cell.frac[0]
    d -= np.round(d)
    r2 = np.sum((d @ cell.aa.T) ** 2, axis=1)
    assert t["mjj"][0] == np.count_nonzero(r2 <= 2.98 ** 2 + 1e-10)


def test_dense_h_symmetric_and_vnl_psd(orc):
    cell = systems.silicon_supercell(1, ngrid_per_cell=12)
    v = systems.model_vloc(cell)
    ham = orc.Ham(cell, vloc=v)
    H = ham.dense()
    assert np.abs(H - H.T).max() < 1e-10 * np.abs(H).max()
    h0 = orc.Ham(cell, vloc=v, projectors=False)
    vnl = H - h0.dense()
    w = np.linalg.eigvalsh(0.5 * (vnl + vnl.T))
    assert w.min() > -1e-10 * w.max()
    assert np.count_nonzero(w > 1e-8 * w.max()) <= 4 * cell.natom
