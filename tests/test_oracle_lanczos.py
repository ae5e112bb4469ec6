"""Pins the oracle's Lanczos, tridiagonal eigensolver and SIL step
(reference multiconfig_lanczos.c, coefficients_integration.c:59-160)
against dense exponentials, numpy's eigh and closed-form special cases."""
import math

import numpy as np
import pytest
import scipy.linalg

import bruteforce as bf
import inputs


@pytest.mark.parametrize("n", [1, 2, 3, 6, 8, 15])
def test_tridiag_eig_matches_eigh(orc, n):
    rng = np.random.default_rng(n)
    d = rng.standard_normal(n)
    e = rng.standard_normal(max(n - 1, 0))
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w, Z = orc.tridiag_eig(d, e)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(T), atol=1e-13)
    assert np.all(np.diff(w) >= 0)
    np.testing.assert_allclose(Z.T @ Z, np.eye(n), atol=1e-13)
    np.testing.assert_allclose(T @ Z, Z * w, atol=1e-12)


def test_tridiag_eig_degenerate_and_split(orc):
    d = np.array([1.0, 1.0, 1.0, 2.0, 2.0])
    e = np.array([0.0, 1e-30, 0.0, 0.0])
    w, Z = orc.tridiag_eig(d, e)
    np.testing.assert_allclose(w, np.sort(d), atol=1e-15)


def _setup(npar, norb, seed=23, gscale=0.3):
    dim = math.comb(npar + norb - 1, norb - 1)
    return inputs.coefficients(dim, seed), inputs.onebody(norb, seed), inputs.twobody(norb, seed, gscale)


def test_lanczos_basis_is_orthonormal_and_tridiagonalises(orc):
    npar, norb, lm = 6, 4, 8
    S = orc.Space(npar, norb)
    c, h, g = _setup(npar, norb)
    done, diag, offd, V = S.lanczos(c, h, g, lm)
    assert done == lm
    np.testing.assert_allclose(V.conj() @ V.T, np.eye(lm), atol=1e-12)
    H = bf.hamiltonian(npar, norb, h, g)
    T = V.conj() @ H @ V.T
    Tw = np.diag(diag.real) + np.diag(offd[: lm - 1].real, 1) + np.diag(offd[: lm - 1].real, -1)
    np.testing.assert_allclose(T, Tw, atol=1e-11)
    assert np.max(np.abs(diag.imag)) < 1e-12


@pytest.mark.parametrize("npar,norb,lm", [(2, 3, 6), (3, 2, 4), (2, 2, 3)])
def test_sil_exact_when_krylov_space_is_exhausted(orc, npar, norb, lm):
    S = orc.Space(npar, norb)
    assert lm >= S.dim
    c, h, g = _setup(npar, norb, 29)
    H = bf.hamiltonian(npar, norb, h, g)
    dt = 0.37
    done, got = S.sil_step(c, h, g, lm, dt, imagtime=False)
    assert done <= S.dim
    np.testing.assert_allclose(got, scipy.linalg.expm(-1j * H * dt) @ c, atol=1e-11)


@pytest.mark.parametrize("lm,tol", [(4, 1e-5), (6, 1e-8), (8, 1e-11)])
def test_sil_converges_to_dense_exponential(orc, lm, tol):
    npar, norb = 6, 4  # dim 84 >> lm
    S = orc.Space(npar, norb)
    c, h, g = _setup(npar, norb, 31)
    H = bf.hamiltonian(npar, norb, h, g)
    dt = 0.01
    _, got = S.sil_step(c, h, g, lm, dt, imagtime=False)
    err = np.linalg.norm(got - scipy.linalg.expm(-1j * H * dt) @ c)
    assert err < tol
    assert abs(np.linalg.norm(got) - 1) < 1e-12  # unitary to Krylov accuracy


def test_sil_imaginary_time_reaches_ground_state(orc):
    npar, norb = 5, 3
    S = orc.Space(npar, norb)
    c, h, g = _setup(npar, norb, 37)
    H = bf.hamiltonian(npar, norb, h, g)
    w, U = np.linalg.eigh(H)
    dt = 0.05
    for _ in range(600):
        _, c = S.sil_step(c, h, g, 5, -1j * dt, imagtime=True)
    assert abs(np.linalg.norm(c) - 1) < 1e-13
    assert abs(abs(np.vdot(U[:, 0], c)) - 1) < 1e-9
    assert abs(np.vdot(c, H @ c).real - w[0]) < 1e-9


def test_lanczos_breakdown_in_invariant_subspace(orc):
    # N = 1: configurations are orbitals.  H couples 0 <-> 1 only, so the
    # Krylov space of e_0 is 2-dimensional and the second beta vanishes.
    npar, norb = 1, 3
    S = orc.Space(npar, norb)
    t = 0.8
    h = np.array([[0, t, 0], [t, 0, 0], [0, 0, 1.0]], complex)
    g = np.zeros(norb ** 4, complex)
    c = np.array([1, 0, 0], complex)
    done, diag, offd, V = S.lanczos(c, h, g, 5)
    assert done == 2
    assert abs(offd[0] - t) < 1e-15
    dt = 0.3
    done, got = S.sil_step(c, h, g, 5, dt, imagtime=False)
    assert done == 2
    np.testing.assert_allclose(got, scipy.linalg.expm(-1j * h * dt) @ c, atol=1e-14)


def test_lanczos_first_iteration_breakdown_is_not_detected(orc):
    # The reference tests beta_i / max(beta_0..i) < 1e-14; at i = 0 an exact
    # zero gives 0/0, so an eigenvector start is not caught and the next
    # vector is 0 * inf.  The oracle reproduces that behaviour.
    npar, norb = 4, 3
    S = orc.Space(npar, norb)
    c = np.zeros(S.dim, complex)
    c[2] = 1.0
    h = np.diag([0.1, 0.7, 1.3]).astype(complex)
    g = np.zeros(norb ** 4, complex)
    done, diag, offd, V = S.lanczos(c, h, g, 3)
    assert done == 3
    assert offd[0] == 0
    assert np.isnan(V[1]).all()


def test_lowest_state_lanczos_matches_eigh(orc):
    npar, norb = 3, 3
    S = orc.Space(npar, norb)
    c, h, g = _setup(npar, norb, 41)
    H = bf.hamiltonian(npar, norb, h, g)
    w, U = np.linalg.eigh(H)
    e, v = S.lowest_state_lanczos(c, h, g, S.dim)
    assert abs(e - w[0]) < 1e-10
    assert abs(abs(np.vdot(U[:, 0], v)) - 1) < 1e-9
