"""Pins of the oracle's dense helpers, Gram-Schmidt, Rayleigh-Ritz, the
hprsdft preconditioner and band CG against numpy/scipy and invariants."""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp

from rsdft_develop_amd import systems


def test_jacobi_matches_eigh(orc):
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 40):
        a = rng.standard_normal((n, n))
        a = a + a.T
        w, v = orc.syev(a)
        w0, v0 = np.linalg.eigh(a)
        np.testing.assert_allclose(w, w0, atol=1e-12 * max(1, np.abs(w0).max()))
        np.testing.assert_allclose(np.abs(np.sum(v * v0, axis=0)), 1.0, atol=1e-10)


def test_potrf_trtri(orc):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((30, 12))
    s = x.T @ x
    rc, r = orc.potrf(s)
    assert rc == 0
    np.testing.assert_allclose(r, np.linalg.cholesky(s).T, rtol=1e-12, atol=1e-12)
    rc, ri = orc.trtri(r)
    np.testing.assert_allclose(ri, np.linalg.inv(r), rtol=1e-10, atol=1e-12)
    rc, _ = orc.potrf(-np.eye(3))
    assert rc == 1


def test_sygv2_matches_scipy(orc):
    rng = np.random.default_rng(2)
    for _ in range(50):
        a = rng.standard_normal((2, 2))
        a = a + a.T
        m = rng.standard_normal((2, 2))
        b = m @ m.T + 0.1 * np.eye(2)
        rc, w, x = orc.sygv2(a, b)
        w0, x0 = scipy.linalg.eigh(a, b)
        assert rc == 0
        np.testing.assert_allclose(w, w0, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(x.T @ b @ x, np.eye(2), atol=1e-11)
        np.testing.assert_allclose(np.abs(np.sum(x * (b @ x0), axis=0)), 1.0, atol=1e-10)


def _qr_reference(u, dv):
    q, r = np.linalg.qr(u)
    s = np.sign(np.diag(r))
    return q * s / np.sqrt(dv)


@pytest.mark.parametrize("variant", ["t", "lusl"])
@pytest.mark.parametrize("mb", [1, 5, 20, 37])
def test_gram_schmidt_matches_qr(orc, variant, mb):
    cell = systems.silicon_supercell(1, ngrid_per_cell=16)
    ham = orc.Ham(cell, projectors=False)
    u = systems.random_bands(cell.npoints, mb, seed=mb)
    q = ham.gram_schmidt(u, variant)
    if variant == "t" and mb == 1:
        # gram_schmidt_sub(1,1,1,1,1) skips every block (ne = min(ne, me-1)
        # = 0), so the recursive variant leaves a single band untouched.
        np.testing.assert_array_equal(q, u)
        return
    np.testing.assert_allclose(q, _qr_reference(u, ham.dv), atol=1e-10)
    np.testing.assert_allclose(q.T @ q * ham.dv, np.eye(mb), atol=1e-12)


def _tiny_system(orc, ng=12, nb=8, seed=0):
    cell = systems.silicon_supercell(1, ngrid_per_cell=ng, nband=nb, displacement=0.05, seed=seed)
    v = systems.model_vloc(cell)
    ham = orc.Ham(cell, vloc=v)
    u = ham.gram_schmidt(systems.random_bands(cell.npoints, nb, seed=seed + 1))
    return cell, ham, u


def test_subspace_diag_matches_dense(orc):
    cell, ham, u = _tiny_system(orc)
    H = ham.dense()
    hsub = ham.dv * u.T @ H @ u
    w0, v0 = np.linalg.eigh(0.5 * (hsub + hsub.T))
    u2, esp = ham.subspace_diag(u)
    np.testing.assert_allclose(esp, w0, atol=1e-10)
    np.testing.assert_allclose(np.abs(np.sum(u2 * (u @ v0), axis=0)) * ham.dv, 1.0, atol=1e-9)
    # Ritz residual is orthogonal to the subspace
    r = H @ u2 - u2 * esp
    assert np.abs(u2.T @ r).max() * ham.dv < 1e-9


def test_precond_is_three_step_cg(orc):
    cell = systems.silicon_supercell(1, ngrid_per_cell=12)
    ham = orc.Ham(cell, projectors=False)
    n1 = n2 = n3 = 12
    h = cell.hgrid

    def lap1(n, hh):
        e = np.ones(n)
        m = sp.diags([e, -2 * e, e], [-1, 0, 1], shape=(n, n), format="lil")
        m[0, n - 1] = m[n - 1, 0] = 1.0
        return m.tocsr() / hh ** 2

    I = lambda n: sp.identity(n, format="csr")
    A = -0.5 * (sp.kron(I(n3), sp.kron(I(n2), lap1(n1, h[0]))) +
                sp.kron(I(n3), sp.kron(lap1(n2, h[1]), I(n1))) +
                sp.kron(lap1(n3, h[2]), sp.kron(I(n2), I(n1))))
    rng = np.random.default_rng(4)
    g = rng.standard_normal((cell.npoints, 3))
    got = ham.precond(g, mloop=3)
    for b in range(3):
        x = g[:, b].copy()
        r = g[:, b] - A @ x
        p = r.copy()
        rr = r @ r
        for it in range(3):
            f = A @ p
            a = rr / (p @ f)
            x += a * p
            r -= a * f
            rr1 = r @ r
            p = r + rr1 / rr * p
            rr = rr1
        np.testing.assert_allclose(got[:, b], x, rtol=1e-11, atol=1e-11 * np.abs(x).max())


def test_cg_lowers_rayleigh_quotients(orc):
    cell, ham, u = _tiny_system(orc)
    e0 = ham.esp_calc(u)
    u1, esp, res = ham.conjugate_gradient(u, ncg=4)
    e1 = ham.esp_calc(u1)
    assert np.all(esp <= e0 + 1e-12)
    np.testing.assert_allclose(esp, e1, rtol=1e-10, atol=1e-10)
    assert np.all(res > 0)


@pytest.mark.parametrize("mbd", [1, 3, 8])
def test_cg_band_block_size_independent(orc, mbd):
    """Bands are updated independently, so MB_d only changes the blocking."""
    cell, ham, u = _tiny_system(orc)
    _, e_ref, _ = ham.conjugate_gradient(u, ncg=3, mbd=8)
    u1, e1, _ = ham.conjugate_gradient(u, ncg=3, mbd=mbd)
    np.testing.assert_allclose(e1, e_ref, rtol=1e-12, atol=1e-12)


def test_inner_loop_converges_to_dense_spectrum(orc):
    """Repeated CG + GS + subspace_diag at fixed V reaches the lowest
    eigenvalues of the dense Hamiltonian."""
    cell, ham, u = _tiny_system(orc, ng=12, nb=8)
    H = ham.dense()
    w = np.linalg.eigvalsh(0.5 * (H + H.T))[:8]
    for _ in range(40):
        u, esp = ham.subspace_diag(u)
        u, esp, res = ham.conjugate_gradient(u, ncg=4)
        u = ham.gram_schmidt(u)
    u, esp = ham.subspace_diag(u)
    np.testing.assert_allclose(esp[:6], w[:6], atol=1e-7)


def _exact_free_cell(n=8, nb=3):
    """A cubic cell with h = 0.5 bohr and md = 1: the FD weights
    (1, -2, 1) * (-1/2)/h^2 are exact binary numbers, so the kinetic operator
    maps a constant band to exactly zero and the constant band has an exactly
    zero CG gradient."""
    return systems.Cell(aa=np.eye(3) * 0.5 * n, frac=np.zeros((1, 3)), kind=np.zeros(1, dtype=np.int32),
                        species=(systems.SI_GTH,), ngrid=(n, n, n), nband=nb, md=1)


def zero_gradient_bands(ham, nb, seed=0):
    """Band 0 constant (an exact eigenvector of T with eigenvalue 0), the rest
    random, orthonormalised with band 0 first."""
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((ham.N, nb))
    u[:, 0] = 1.0
    return ham.gram_schmidt(u, "lusl")


def test_cg_keeps_a_band_with_zero_gradient(orc):
    """A band whose gradient is exactly zero takes no step: no 0/0 in the
    preconditioner, the direction update or the 2x2 Ritz step, and the
    other bands still converge (their energies drop)."""
    cell = _exact_free_cell()
    ham = orc.Ham(cell, projectors=False)
    u = zero_gradient_bands(ham, 3)
    c = u[:, 0].copy()
    assert np.abs(ham.hpsi(u[:, :1])).max() == 0.0
    e0 = np.array([u[:, k] @ ham.hpsi(u[:, k:k + 1])[:, 0] * ham.dv for k in range(3)])
    for pc in (0, 4):
        for mbd in (1, 3):
            v, esp, res = ham.conjugate_gradient(u, ncg=4, mbd=mbd, pc_kind=pc)
            assert np.all(np.isfinite(v)) and np.all(np.isfinite(esp)) and np.all(np.isfinite(res))
            np.testing.assert_array_equal(v[:, 0], c)
            assert esp[0] == 0.0 and res[0] == 0.0
            assert np.all(esp[1:] < e0[1:])


def _overlaps(a, b, dv):
    return a.T @ b * dv


def test_cggs_swgs2_one_step_keeps_bands_orthogonal_to_the_others(orc):
    """SWGS 2 (cggs_2) removes every other band from each preconditioned
    gradient, so after one CG step of a whole block each new band is
    orthogonal to every other band's old vector; without it they are not."""
    _, ham, u = _tiny_system(orc, nb=6)
    v2, _, _ = ham.conjugate_gradient(u, ncg=1, pc_kind=4, swgs=2)
    v0, _, _ = ham.conjugate_gradient(u, ncg=1, pc_kind=4, swgs=0)
    s2 = _overlaps(v2, u, ham.dv)
    s0 = _overlaps(v0, u, ham.dv)
    off = ~np.eye(6, dtype=bool)
    assert np.abs(s2[off]).max() < 1e-12
    assert np.abs(s0[off]).max() > 1e-4


def test_cggs_swgs1_one_step_projects_out_the_earlier_bands(orc):
    """SWGS 1 (cggs_1) removes only the bands before each one: after one CG
    step of a whole block, new band j is orthogonal to old bands i < j but not
    to later ones, and band 0 (nothing before it) moves exactly as without
    the projection."""
    _, ham, u = _tiny_system(orc, nb=6)
    v1, e1, _ = ham.conjugate_gradient(u, ncg=1, pc_kind=4, swgs=1)
    v0, e0, _ = ham.conjugate_gradient(u, ncg=1, pc_kind=4, swgs=0)
    s1 = _overlaps(v1, u, ham.dv)  # s1[j, i] = dV <new j, old i>
    below = np.array([[i < j for i in range(6)] for j in range(6)])
    assert np.abs(s1[below]).max() < 1e-12
    above = np.array([[i > j for i in range(6)] for j in range(6)])
    assert np.abs(s1[above]).max() > 1e-4
    np.testing.assert_array_equal(v1[:, 0], v0[:, 0])
    assert e1[0] == e0[0]


def test_cggs_swgs2_band_by_band_keeps_the_set_orthonormal(orc):
    """With one band per block (MB_d = 1) and SWGS 2, each band's CG moves
    only within the complement of all the others, so an orthonormal start
    stays orthonormal through four CG steps."""
    _, ham, u = _tiny_system(orc, nb=6)
    v2, _, _ = ham.conjugate_gradient(u, ncg=4, mbd=1, pc_kind=4, swgs=2)
    v0, _, _ = ham.conjugate_gradient(u, ncg=4, mbd=1, pc_kind=4, swgs=0)
    off = ~np.eye(6, dtype=bool)
    assert np.abs(_overlaps(v2, v2, ham.dv)[off]).max() < 1e-11
    assert np.abs(_overlaps(v0, v0, ham.dv)[off]).max() > 1e-4
