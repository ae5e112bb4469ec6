"""GPU parity of the full-equation Runge-Kutta orbital step against the
FP64 oracle (reference orbital_integration.c:22-348), through the C ABI.
Bar: 1e-10 relative (max-abs error over max-abs value) for the right-hand
side, the projectors and a few steps; 1e-9 for whole MCTDHB steps."""
import numpy as np
import pytest

import gpu_util as gu
import oracle
from mctdhb_amd import inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import mctdhb_amd

    mctdhb_amd.lib()
    return mctdhb_amd


def _setup(amd, M, grid, N=4, seed=3, bounds=1001, imagtime=False, rk=4, L=10.0, g=1.3, dt=0.004, d1=0j):
    x = np.linspace(-L / 2, L / 2, grid)
    pot = 0.5 * x ** 2 + 0.3 * np.cos(x)
    eq = oracle.make_orbeq(M, grid, x[0], x[-1], dt, g, pot, imagtime=imagtime, bounds=bounds, d1coef=d1,
                           rk_order=rk)
    w = amd.DeviceOrbital(M, grid, x[0], x[-1], dt, g, pot, imagtime=imagtime, bounds=bounds, d1coef=d1,
                          rk_order=rk)
    orb = inputs.hermite_gauss(M, x, seed=seed, noise=0.05)  # not orthonormal: the projectors differ
    if bounds == 1001:
        orb[:, -1] = orb[:, 0]
    S = oracle.Space(N, M)
    _, r2, rinv = S.sync_density(inputs.coefficients(S.dim, seed))
    return eq, w, np.ascontiguousarray(orb.ravel()), np.ascontiguousarray(r2), np.ascontiguousarray(rinv.ravel())


@pytest.mark.parametrize("M,grid", [(3, 129), (9, 65), (4, 101)])
@pytest.mark.parametrize("use_fd", [1, 0, 2])  # finite differences, FFT, DVR
@pytest.mark.parametrize("robust", [0, 1])
def test_fullstep_rhs_matches_oracle(amd, M, grid, use_fd, robust):
    print(f"fullstep rhs M={M} grid={grid} fd={use_fd} robust={robust}", flush=True)
    eq, w, orb, r2, rinv = _setup(amd, M, grid, N=3 if M == 9 else 4, d1=0.2j)
    ref = np.zeros_like(orb)
    oracle.lib().orc_fullstep_rhs(eq, use_fd, robust, orb, r2, rinv, ref)
    d_der = gu.empty(orb.size)
    w.fullstep_rhs(gu.dev(orb), gu.dev(r2), gu.dev(rinv), d_der, use_fd=use_fd, robust=robust)
    assert gu.rel_err(gu.host(d_der), ref) < 1e-10


@pytest.mark.parametrize("robust", [0, 1])
def test_projector_matches_oracle(amd, robust):
    print(f"projector robust={robust}", flush=True)
    M, grid = 5, 97
    eq, w, orb, _, _ = _setup(amd, M, grid, N=2)
    rng = np.random.default_rng(4)
    h = rng.standard_normal(M * grid) + 1j * rng.standard_normal(M * grid)
    ref = np.zeros_like(orb)
    oracle.lib().orc_multiorb_projector(eq, robust, orb, h, ref)
    d_p = gu.empty(orb.size)
    w.projector(gu.dev(orb), gu.dev(h), d_p, robust=robust)
    assert gu.rel_err(gu.host(d_p), ref) < 1e-10


@pytest.mark.parametrize("rk", [2, 4, 5])
@pytest.mark.parametrize("imagtime", [False, True])
@pytest.mark.parametrize("use_fd,bounds", [(1, 1001), (0, 1001), (1, 1000), (2, 1001)])
def test_fullstep_matches_oracle(amd, rk, imagtime, use_fd, bounds):
    print(f"fullstep rk={rk} imag={imagtime} fd={use_fd} bounds={bounds}", flush=True)
    M, grid = 3, 129
    eq, w, orb, r2, rinv = _setup(amd, M, grid, imagtime=imagtime, rk=rk, bounds=bounds)
    ref = orb.copy()
    d = gu.dev(orb)
    dr2, dri = gu.dev(r2), gu.dev(rinv)
    for _ in range(3):
        oracle.lib().orc_fullstep_orb(eq, int(imagtime), use_fd, 1, ref, r2, rinv)
        w.fullstep(d, dr2, dri, use_fd=use_fd, robust=True)
    assert gu.rel_err(gu.host(d), ref) < 1e-10


@pytest.mark.parametrize("imagtime", [True, False])
def test_propagate_step_with_fullstep_integrator(amd, imagtime):
    """mctdhb_propagate_step_dev after mctdhb_orbital_set_integrator(30):
    full step, periodic copy, h/g, SIL, density sync (driver.c:158-202)."""
    print(f"propagate fullstep imag={imagtime}", flush=True)
    npar, M, grid, lm, dt = 5, 3, 129, 5, 0.004
    x = np.linspace(-8, 8, grid)
    pot = 0.5 * x ** 2
    eq = oracle.make_orbeq(M, grid, x[0], x[-1], dt, 1.0, pot, imagtime=imagtime)
    orb = np.ascontiguousarray(inputs.hermite_gauss(M, x, seed=1, noise=1e-3).ravel())
    oracle.lib().orc_orthonormalize(grid, x[1] - x[0], M, orb)
    S = oracle.Space(npar, M)
    coef = inputs.coefficients(S.dim, 1)
    r1, r2, rinv = S.sync_density(coef)
    h = dict(orb=orb, coef=coef, r1=np.ascontiguousarray(r1.ravel()), r2=np.ascontiguousarray(r2),
             rinv=np.ascontiguousarray(rinv.ravel()), hob=np.zeros(M * M, complex), hint=np.zeros(M ** 4, complex))
    oracle.lib().orc_orbital_hob(eq, h["orb"], h["hob"])
    oracle.lib().orc_orbital_hint(eq, h["orb"], h["hint"])
    D = amd.DeviceSpace(npar, M, tables=True)
    sil = amd.DeviceSil(D, lm)
    w = amd.DeviceOrbital(M, grid, x[0], x[-1], dt, 1.0, pot, imagtime=imagtime)
    w.set_integrator(fullstep=True, robust=False)
    d = {k: gu.dev(v) for k, v in h.items()}
    for _ in range(4):
        oracle.lib().orc_fullstep_orb(eq, int(imagtime), 0, 0, h["orb"], h["r2"], h["rinv"])
        h["orb"].reshape(M, grid)[:, -1] = h["orb"].reshape(M, grid)[:, 0]
        oracle.lib().orc_orbital_hob(eq, h["orb"], h["hob"])
        oracle.lib().orc_orbital_hint(eq, h["orb"], h["hint"])
        _, h["coef"] = S.sil_step(h["coef"], h["hob"], h["hint"], lm, -1j * dt if imagtime else dt, imagtime)
        r1, r2, ri = S.sync_density(h["coef"])
        h["r1"], h["r2"], h["rinv"] = r1.ravel(), np.ascontiguousarray(r2), np.ascontiguousarray(ri.ravel())
        amd.propagate_step(D, sil, w, d["orb"], d["coef"], d["hob"], d["hint"], d["r1"], d["r2"], d["rinv"],
                           use_fft=True)
    for k in ("orb", "coef", "hob", "hint", "r1", "r2", "rinv"):
        assert gu.rel_err(gu.host(d[k]), h[k]) < 1e-9, k
