"""Pins of the oracle's full-equation orbital step (oracle/fullstep.c,
reference orbital_integration.c:22-348, split_linear_orbitals.c:209-286,
split_nonlinear_orbitals.c:4-60, orbital_matrices.c:7-20): plane-wave
symbols of the linear parts, the interacting part against the pinned
nonlinear term, the overlap pattern of quirk Q2, the projector as the
Simpson-weighted orthogonal projection, the right-hand side orthogonal to
the orbital space, and the convergence order of each Runge-Kutta tableau."""
import math

import numpy as np
import pytest

import oracle
from mctdhb_amd import inputs


def _eq(norb, grid, L=10.0, bounds=1001, g=1.0, dt=0.01, imagtime=False, d1=0j, pot=None, rk=4):
    x = np.linspace(-L / 2, L / 2, grid)
    if pot is None:
        pot = 0.5 * x ** 2
    return x, oracle.make_orbeq(norb, grid, x[0], x[-1], dt, g, pot, imagtime=imagtime, bounds=bounds, d1coef=d1,
                                rk_order=rk)


def test_fd_linear_part_plane_wave():
    """Periodic grid: e^{ikx} is an eigenvector of the stencil with
    a2 (2cos(k dx) - 2)/dx^2 + i a1 sin(k dx)/dx + V."""
    grid = 65
    n = grid - 1
    for d1 in (0j, 0.3 + 0.2j):
        x, eq = _eq(1, grid, pot=np.full(grid, 0.7), d1=d1)
        dx = x[1] - x[0]
        for m in (1, 3, 7):
            kk = 2 * np.pi * m / (n * dx)
            f = np.exp(1j * kk * (x - x[0]))
            out = np.zeros(grid, complex)
            oracle.lib().orc_linear_horb_fd(eq, f, out)
            lam = -0.5 * (2 * np.cos(kk * dx) - 2) / dx ** 2 + 1j * d1 * np.sin(kk * dx) / dx + 0.7
            assert np.allclose(out, lam * f, rtol=0, atol=1e-11 * abs(lam))


def test_fd_linear_part_hard_wall():
    """Zero boundaries: interior points follow the 3-point stencil; the end
    points drop the neighbour outside the box."""
    grid = 33
    x, eq = _eq(1, grid, bounds=1000, pot=np.zeros(grid))
    dx = x[1] - x[0]
    f = np.sin(np.pi * (x - x[0]) / (x[-1] - x[0])) + 0.1j * x
    out = np.zeros(grid, complex)
    oracle.lib().orc_linear_horb_fd(eq, f, out)
    c = -0.5 / dx ** 2
    assert np.allclose(out[1:-1], c * (f[2:] - 2 * f[1:-1] + f[:-2]), atol=1e-10)
    assert out[0] == pytest.approx(c * (f[1] - 2 * f[0]))
    assert out[-1] == pytest.approx(c * (f[-2] - 2 * f[-1]))


def test_fft_linear_part_plane_wave():
    """e^{ikx} has symbol i k d1 - k^2 d2 (+ V); the last point repeats the first."""
    for grid in (65, 61):  # FFT and direct-DFT lengths
        n = grid - 1
        x, eq = _eq(1, grid, pot=np.full(grid, -0.2), d1=0.4j)
        dx = x[1] - x[0]
        for m in (2, -3):
            kk = 2 * np.pi * m / (n * dx)
            f = np.exp(1j * kk * (x - x[0]))
            out = np.zeros(grid, complex)
            oracle.lib().orc_linear_horb_fft(eq, f, out)
            lam = 1j * kk * 0.4j + 0.5 * kk ** 2 - 0.2
            assert np.allclose(out[:n], lam * f[:n], atol=1e-10)
            assert out[n] == out[0]


def _simpson_weights(grid, dx):
    return np.array([oracle.csimps(np.eye(grid)[p].astype(complex), dx).real for p in range(grid)])


def test_dvr_linear_part_plane_wave():
    """The exponential DVR (split_linear_orbitals.c:82-165) is exact on the
    grid's plane waves: symbol i k d1 - k^2 d2 (+ V), as the FFT form."""
    grid = 49
    n = grid - 1
    x, eq = _eq(1, grid, pot=np.full(grid, 0.4), d1=0.3 + 0.1j)
    dx = x[1] - x[0]
    mat = np.zeros(n * n, complex)
    oracle.lib().orc_expdvr_mat(eq, mat)
    D1 = mat.reshape(n, n).copy()
    for m in (0, 1, -5, 11):
        kk = 2 * np.pi * m / (n * dx)
        f = np.exp(1j * kk * (x - x[0]))
        out = np.zeros(grid, complex)
        oracle.lib().orc_linear_horb_dvr(eq, mat, f, out)
        lam = 1j * kk * (0.3 + 0.1j) + 0.5 * kk ** 2 + 0.4
        assert np.allclose(out[:n], lam * f[:n], atol=1e-10 * max(1, abs(lam)))
        assert out[n] == out[0]
    # without d1 the matrix is Hermitian (the second derivative); with it, not
    x, eq = _eq(1, grid, d1=0.0j)
    oracle.lib().orc_expdvr_mat(eq, mat)
    D = mat.reshape(n, n)
    assert np.allclose(D, D.conj().T, atol=1e-12 * np.max(abs(D)))
    assert not np.allclose(D1, D1.conj().T)


def _state(norb, npar, grid, seed, L=10.0):
    """Orbitals orthonormal in the Simpson inner product (classical
    Gram-Schmidt in numpy: the reference's point-by-point variant leaves
    small overlaps), densities of seeded coefficients."""
    x = np.linspace(-L / 2, L / 2, grid)
    w = _simpson_weights(grid, x[1] - x[0])
    F = inputs.hermite_gauss(norb, x, seed=seed, noise=0.05)
    for i in range(norb):
        for j in range(i):
            F[i] -= (np.sum(w * np.conj(F[j]) * F[i])) * F[j]
        F[i] /= math.sqrt(np.sum(w * abs(F[i]) ** 2))
    orb = np.ascontiguousarray(F.ravel())
    S = oracle.Space(npar, norb)
    r1, r2, rinv = S.sync_density(inputs.coefficients(S.dim, seed))
    return x, orb, np.ascontiguousarray(r2), np.ascontiguousarray(rinv.ravel())


def test_interacting_part_is_the_nonlinear_term_without_projections():
    """With h = 0 and g_ksql = 0 the split-step nonlinear term keeps only
    the interacting part (split_nonlinear_orbitals.c:62-151 vs :4-60), up
    to the order of the products."""
    M, grid = 3, 33
    x, orb, r2, rinv = _state(M, 4, grid, 2)
    zh, zg = np.zeros(M * M, complex), np.zeros(M ** 4, complex)
    for k in range(M):
        for n in (0, 7, 20):
            a = oracle.lib().orc_orb_interacting_part(k, n, 1.7, M, grid, orb, r2, rinv)
            b = oracle.lib().orc_orb_full_nonlinear(k, n, 1.7, M, grid, orb, zh, zg, r2, rinv)
            # the same terms; g multiplies before rather than after the products
            assert abs(complex(a.re, a.im) - complex(b.re, b.im)) <= 1e-14 * abs(complex(b.re, b.im))


def test_interacting_part_gross_pitaevskii():
    """M = 1: rho2 = N(N-1), rinv = 1/N, so the term is g (N-1) |f|^2 f."""
    grid, N, g = 17, 5, 0.8
    f = np.exp(-np.linspace(-2, 2, grid) ** 2) * (1 + 0.3j)
    r2 = np.array([N * (N - 1)], complex)
    rinv = np.array([1.0 / N], complex)
    for n in range(grid):
        v = oracle.lib().orc_orb_interacting_part(0, n, g, 1, grid, f, r2, rinv)
        assert complex(v.re, v.im) == pytest.approx(g * (N - 1) * abs(f[n]) ** 2 * f[n], rel=1e-14)


def test_overlap_matrix_pattern_q2():
    """The j loop starts at 2 (orbital_matrices.c:13): entries [0][1] and
    [1][0] are never set, everything else is <f_i, f_j> (Hermitian)."""
    M, grid = 4, 41
    x = np.linspace(-5, 5, grid)
    dx = x[1] - x[0]
    orb = np.ascontiguousarray(inputs.hermite_gauss(M, x, seed=3, noise=0.3).ravel())
    S = np.zeros(M * M, complex)
    oracle.lib().orc_overlap_matrix(M, grid, dx, orb, S)
    S = S.reshape(M, M)
    F = orb.reshape(M, grid)
    for i in range(M):
        for j in range(M):
            ref = oracle.csimps(np.conj(F[i]) * F[j], dx)
            if {i, j} == {0, 1}:
                assert S[i, j] == 0
            else:
                assert abs(S[i, j] - ref) <= 1e-14 * abs(ref) + 1e-16


def test_projectors_are_the_orthogonal_projection():
    """Simple projector on an orthonormal set and robust projector on a set
    whose only unset overlap (orbitals 0, 1: even and odd) is exactly zero:
    both equal the Simpson-weighted least-squares projection onto the span."""
    grid = 81
    x, eq = _eq(3, grid)
    dx = x[1] - x[0]
    w = _simpson_weights(grid, dx)
    g = np.exp(-x ** 2 / 2)
    F = np.array([g, x * g, (x ** 2 + 0.3) * g], complex)  # <f0, f1> = 0 by parity
    rng = np.random.default_rng(1)
    Hm = rng.standard_normal((3, grid)) + 1j * rng.standard_normal((3, grid))
    Gm = (np.conj(F) * w) @ F.T          # Gram matrix <f_a, f_b>
    Pm = (np.conj(F) * w) @ Hm.T         # <f_a, h_k>
    ref = (np.linalg.solve(Gm, Pm)).T @ F
    out = np.zeros(3 * grid, complex)
    oracle.lib().orc_multiorb_projector(eq, 1, np.ascontiguousarray(F.ravel()), np.ascontiguousarray(Hm.ravel()),
                                        out)
    assert np.allclose(out.reshape(3, grid), ref, atol=1e-11)
    Q = np.ascontiguousarray(F.ravel())
    oracle.lib().orc_orthonormalize(grid, dx, 3, Q)
    Qm = Q.reshape(3, grid)
    ref = ((np.conj(Qm) * w) @ Hm.T).T @ Qm
    oracle.lib().orc_multiorb_projector(eq, 0, Q, np.ascontiguousarray(Hm.ravel()), out)
    assert np.allclose(out.reshape(3, grid), ref, atol=1e-11)


@pytest.mark.parametrize("use_fd", [1, 0, 2])
@pytest.mark.parametrize("robust", [0, 1])
def test_rhs_is_orthogonal_to_the_orbitals(use_fd, robust):
    """der_k = -i(H f_k - P H f_k) has no component along the orbitals."""
    M, grid = 3, 65
    x, orb, r2, rinv = _state(M, 4, grid, 5)
    _, eq = _eq(M, grid, g=1.3)
    der = np.zeros(M * grid, complex)
    oracle.lib().orc_fullstep_rhs(eq, use_fd, robust, orb, r2, rinv, der)
    F, D = orb.reshape(M, grid), der.reshape(M, grid)
    dx = x[1] - x[0]
    scale = np.max(abs(D))
    for s in range(M):
        for k in range(M):
            assert abs(oracle.csimps(np.conj(F[s]) * D[k], dx)) < 1e-11 * scale


@pytest.mark.parametrize("rk,order", [(2, 2), (4, 4), (5, 5)])
def test_fullstep_convergence_order(rk, order):
    """One step of dt against two of dt/2 (rho2, rinv fixed): the error
    ratio for halving dt approaches 2^order."""
    M, grid = 2, 33
    x, orb0, r2, rinv = _state(M, 3, grid, 7, L=8.0)

    def run(dt, steps):
        _, eq = _eq(M, grid, L=8.0, g=1.0, dt=dt, rk=rk)
        o = orb0.copy()
        for _ in range(steps):
            oracle.lib().orc_fullstep_orb(eq, 0, 1, 0, o, r2, rinv)
        return o

    h0 = 0.01  # the RK5 ratio is pre-asymptotic (5.6) at 0.04
    ref = run(h0 / 64, 64)
    e1 = np.max(abs(run(h0, 1) - ref))
    e2 = np.max(abs(run(h0 / 2, 2) - ref))
    assert math.log2(e1 / e2) == pytest.approx(order, abs=0.25)
