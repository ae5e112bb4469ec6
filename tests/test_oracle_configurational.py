"""Pins the oracle's H.C, rho1, rho2 and total energy (reference
hamiltonian.c, density_matrices.c, integrator_monitor.c:55-88) against a
dense brute-force Fock-space build and closed-form identities."""
import numpy as np
import pytest

import bruteforce as bf
import inputs

SHAPES = [(1, 2), (1, 4), (2, 2), (2, 3), (3, 3), (4, 2), (3, 4), (5, 3), (4, 5), (6, 4), (5, 5)]


def _problem(npar, norb, seed=11):
    return (inputs.coefficients(__import__("math").comb(npar + norb - 1, norb - 1), seed),
            inputs.onebody(norb, seed), inputs.twobody(norb, seed, scale=0.3))


def test_twobody_input_symmetries():
    M = 4
    g = inputs.twobody(M, 5).reshape((M,) * 4, order="F")  # g[k, s, q, l]
    assert np.allclose(g, g.transpose(1, 0, 2, 3))
    assert np.allclose(g, g.transpose(0, 1, 3, 2))
    assert np.allclose(g, np.conj(g.transpose(3, 2, 1, 0)))


@pytest.mark.parametrize("npar,norb", SHAPES)
def test_hc_matches_dense(orc, npar, norb):
    c, h, g = _problem(npar, norb)
    S = orc.Space(npar, norb)
    H = bf.hamiltonian(npar, norb, h, g)
    assert np.allclose(H, H.conj().T, atol=1e-13)  # Hermitian for symmetric g
    got = S.apply_hamiltonian(c, h, g)
    np.testing.assert_allclose(got, H @ c, rtol=0, atol=1e-12)


def test_hc_noninteracting_diagonal(orc):
    npar, norb = 7, 4
    S = orc.Space(npar, norb)
    c = inputs.coefficients(S.dim, 3)
    h = np.diag([0.5, -1.25, 2.0, 3.5]).astype(complex)
    g = np.zeros(norb ** 4, complex)
    occ = S.hash_table.reshape(S.dim, norb).astype(float)
    np.testing.assert_allclose(S.apply_hamiltonian(c, h, g), (occ @ np.diag(h).real) * c,
                               atol=1e-13)


def test_hc_single_particle_is_h(orc):
    norb = 5
    S = orc.Space(1, norb)
    c = inputs.coefficients(norb, 4)
    h = inputs.onebody(norb, 4)
    g = inputs.twobody(norb, 4)
    np.testing.assert_allclose(S.apply_hamiltonian(c, h, g), h @ c, atol=1e-13)


def test_hc_two_mode_closed_form(orc):
    # M = 2: <n0-1, n1+1| a+_1 a_0 |n0, n1> = sqrt(n0 (n1+1)), one-body only
    npar = 9
    S = orc.Space(npar, 2)
    c = inputs.coefficients(S.dim, 5)
    t = 0.7 - 0.2j
    h = np.array([[0, np.conj(t)], [t, 0]])
    g = np.zeros(16, complex)
    want = np.zeros_like(c)
    for i in range(S.dim):  # state i: n0 = N - i, n1 = i
        n0, n1 = npar - i, i
        if i + 1 < S.dim:  # a+_0 a_1 maps state i+1 to i
            want[i] += h[0, 1] * np.sqrt(n0 * (n1 + 1)) * c[i + 1]
        if i > 0:  # a+_1 a_0 maps state i-1 to i
            want[i] += h[1, 0] * np.sqrt((n0 + 1) * n1) * c[i - 1]
    np.testing.assert_allclose(S.apply_hamiltonian(c, h, g), want, atol=1e-13)


@pytest.mark.parametrize("npar,norb", SHAPES)
def test_density_matrices_match_dense(orc, npar, norb):
    c, h, g = _problem(npar, norb, 13)
    S = orc.Space(npar, norb)
    np.testing.assert_allclose(S.onebody_dm(c), bf.onebody_dm(npar, norb, c), atol=1e-12)
    np.testing.assert_allclose(S.twobody_dm(c), bf.twobody_dm(npar, norb, c), atol=1e-12)


@pytest.mark.parametrize("npar,norb", [(8, 3), (12, 5), (20, 4), (9, 7)])
def test_density_matrix_identities(orc, npar, norb):
    S = orc.Space(npar, norb)
    c = inputs.coefficients(S.dim, 17)
    r1 = S.onebody_dm(c)
    r2 = S.twobody_dm(c).reshape((norb,) * 4, order="F")  # r2[k, s, q, l]
    assert abs(np.trace(r1) - npar) < 1e-12
    assert np.allclose(r1, r1.conj().T, atol=1e-13)
    # sum_s <a+_k a+_s a_s a_l> = (N - 1) <a+_k a_l>
    contracted = np.einsum("kssl->kl", r2)
    np.testing.assert_allclose(contracted, (npar - 1) * r1, atol=1e-11)
    # <a+_k a+_s a_q a_l> = conj <a+_l a+_q a_s a_k>
    np.testing.assert_allclose(r2, np.conj(r2.transpose(3, 2, 1, 0)), atol=1e-13)


def test_density_matrices_single_fock_state(orc):
    npar, norb = 6, 4
    S = orc.Space(npar, norb)
    i = S.dim // 3
    occ = S.hash_table.reshape(S.dim, norb)[i].astype(float)
    c = np.zeros(S.dim, complex)
    c[i] = 1
    np.testing.assert_allclose(S.onebody_dm(c), np.diag(occ), atol=1e-15)
    r2 = S.twobody_dm(c).reshape((norb,) * 4, order="F")
    for k in range(norb):
        assert abs(r2[k, k, k, k] - occ[k] * (occ[k] - 1)) < 1e-15
        for s in range(norb):
            if s != k:
                assert abs(r2[k, s, k, s] - occ[k] * occ[s]) < 1e-15
                assert abs(r2[k, s, s, k] - occ[k] * occ[s]) < 1e-15


@pytest.mark.parametrize("npar,norb", [(3, 3), (5, 4), (4, 5), (10, 3)])
def test_energy_identity(orc, npar, norb):
    S = orc.Space(npar, norb)
    c, h, g = _problem(npar, norb, 19)
    e = orc.total_energy(norb, S.onebody_dm(c), S.twobody_dm(c), h, g)
    want = np.vdot(c, S.apply_hamiltonian(c, h, g))
    assert abs(e - want) < 1e-11
    assert abs(want - np.vdot(c, bf.hamiltonian(npar, norb, h, g) @ c)) < 1e-11
