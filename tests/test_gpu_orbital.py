"""GPU parity of the orbital split-step against the FP64 oracle (through the
C ABI).  Bar: 1e-10 relative (max-abs error over max-abs value) for single
operations; a multi-step propagation is compared at 1e-9; the ground-state
run reproduces benchmark.dat's known answer."""
import math

import numpy as np
import pytest

import gpu_util as gu
from mctdhb_amd import inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import mctdhb_amd

    mctdhb_amd.lib()
    return mctdhb_amd


def _setup(orc, amd, norb, grid, imagtime, bounds=1001, L=16.0, g=1.3, dt=0.01, seed=5, d1coef=0j, rk_order=4):
    x = np.linspace(-L / 2, L / 2, grid)
    pot = 0.5 * x ** 2
    eq = orc.make_orbeq(norb, grid, x[0], x[-1], dt, g, pot, imagtime=imagtime, bounds=bounds, d1coef=d1coef,
                        rk_order=rk_order)
    w = amd.DeviceOrbital(norb, grid, x[0], x[-1], dt, g, pot, imagtime=imagtime, bounds=bounds, d1coef=d1coef,
                          rk_order=rk_order)
    orb = np.ascontiguousarray(inputs.hermite_gauss(norb, x, seed=seed, noise=0.02).ravel())
    if bounds == 1001:
        orb.reshape(norb, grid)[:, -1] = orb.reshape(norb, grid)[:, 0]
    return x, eq, w, orb


@pytest.mark.parametrize("grid", [257, 101, 513])
@pytest.mark.parametrize("imagtime", [False, True])
def test_linear_fft_step(orc, amd, grid, imagtime):
    print(f"fft grid={grid} imag={imagtime}", flush=True)
    x, eq, w, orb = _setup(orc, amd, 3, grid, imagtime)
    ref = orb.copy()
    orc.lib().orc_advance_linear_fft(eq, ref)
    d = gu.dev(orb)
    w.linear_fft(d)
    assert gu.rel_err(gu.host(d), ref) < 1e-12


# grid 2049: the largest the parallel (PCR) solve takes, 2 or 3 rows per
# thread; beyond it the sequential sweeps run
@pytest.mark.parametrize("grid", [129, 257, 2049, 2305])
@pytest.mark.parametrize("bounds", [1001, 1000])
@pytest.mark.parametrize("imagtime", [False, True])
def test_linear_cn_step(orc, amd, bounds, imagtime, grid):
    print(f"cn bounds={bounds} imag={imagtime} grid={grid}", flush=True)
    x, eq, w, orb = _setup(orc, amd, 3, grid, imagtime, bounds=bounds, d1coef=0.3 + 0.1j)
    ref = orb.copy()
    orc.lib().orc_advance_linear_cn(eq, ref)
    d = gu.dev(orb)
    w.linear_cn(d)
    assert gu.rel_err(gu.host(d), ref) < 1e-12


@pytest.mark.parametrize("grid", [257, 258])
def test_orbital_matrices(orc, amd, grid):
    norb = 4
    x, eq, w, orb = _setup(orc, amd, norb, grid, True)
    hob = np.zeros(norb * norb, complex)
    hint = np.zeros(norb ** 4, complex)
    orc.lib().orc_orbital_hob(eq, orb, hob)
    orc.lib().orc_orbital_hint(eq, orb, hint)
    dh, dg = gu.empty(norb * norb), gu.empty(norb ** 4)
    w.orbital_matrices(gu.dev(orb), dh, dg)
    assert gu.rel_err(gu.host(dh), hob) < 1e-12
    assert gu.rel_err(gu.host(dg), hint) < 1e-12


def _densities(orc, npar, norb, seed=3):
    S = orc.Space(npar, norb)
    c = inputs.coefficients(S.dim, seed)
    r1, r2, rinv = S.sync_density(c)
    return S, c, np.ascontiguousarray(r1.ravel()), r2, np.ascontiguousarray(rinv.ravel())


@pytest.mark.parametrize("imagtime", [False, True])
def test_nonlinear_rhs(orc, amd, imagtime):
    npar, norb = 6, 4
    x, eq, w, orb = _setup(orc, amd, norb, 129, imagtime)
    S, c, r1, r2, rinv = _densities(orc, npar, norb)
    hob = np.zeros(norb * norb, complex)
    hint = np.zeros(norb ** 4, complex)
    orc.lib().orc_orbital_hob(eq, orb, hob)
    orc.lib().orc_orbital_hint(eq, orb, hint)
    ref = np.zeros_like(orb)
    orc.lib().orc_nonlinear_rhs(eq, orb, hob, hint, r2, rinv, ref)
    der = gu.empty(orb.size)
    w.nonlinear_rhs(gu.dev(orb), gu.dev(hob), gu.dev(hint), gu.dev(r2), gu.dev(rinv), der)
    assert gu.rel_err(gu.host(der), ref) < 1e-12


# 5 orbitals: staged in LDS up to 160 KiB (1025, 2000 points), not at 2049
@pytest.mark.parametrize("grid", [257, 258, 1025, 2000, 2049])
def test_orthonormalize(orc, amd, grid):
    norb = 5
    x, eq, w, orb = _setup(orc, amd, norb, grid, True)
    ref = orb.copy()
    orc.lib().orc_orthonormalize(grid, eq.dx, norb, ref)
    d = gu.dev(orb)
    w.orthonormalize(d)
    assert gu.rel_err(gu.host(d), ref) < 1e-10


def test_density_sync(orc, amd):
    npar, norb = 8, 4
    S, c, r1, r2, rinv = _densities(orc, npar, norb, seed=7)
    D = amd.DeviceSpace(npar, norb, tables=False)
    d1, d2, di = gu.empty(norb * norb), gu.empty(norb ** 4), gu.empty(norb * norb)
    eig = gu.empty(norb, "float64")
    amd.density_sync(D, gu.dev(c), d1, d2, di, eig)
    assert gu.rel_err(gu.host(d1), r1) < 1e-12
    assert gu.rel_err(gu.host(di), rinv) < 1e-10
    np.testing.assert_allclose(gu.host(eig), np.linalg.eigvalsh(r1.reshape(norb, norb)), atol=1e-12)


def test_density_sync_regularises_a_condensate(orc, amd):
    # a pure condensate has a singular rho1: the reference regularises it
    npar, norb = 5, 3
    S = orc.Space(npar, norb)
    c = np.zeros(S.dim, complex)
    c[0] = 1.0
    r1, r2, rinv = S.sync_density(c)
    D = amd.DeviceSpace(npar, norb, tables=False)
    d1, d2, di = gu.empty(norb * norb), gu.empty(norb ** 4), gu.empty(norb * norb)
    amd.density_sync(D, gu.dev(c), d1, d2, di)
    assert gu.rel_err(gu.host(d1), r1.ravel()) < 1e-12
    assert gu.rel_err(gu.host(di), rinv.ravel()) < 1e-8


@pytest.mark.parametrize("use_fft,imagtime,rk", [(True, False, 4), (True, True, 4), (False, False, 4),
                                                 (False, True, 4), (True, False, 2), (True, True, 5),
                                                 (False, False, 5)])
def test_splitstep(orc, amd, use_fft, imagtime, rk):
    print(f"splitstep fft={use_fft} imag={imagtime} rk={rk}", flush=True)
    npar, norb = 6, 3
    x, eq, w, orb = _setup(orc, amd, norb, 129, imagtime, rk_order=rk)
    S, c, r1, r2, rinv = _densities(orc, npar, norb)
    hob = np.zeros(norb * norb, complex)
    hint = np.zeros(norb ** 4, complex)
    ref = orb.copy()
    orc.lib().orc_splitstep_orb(eq, int(imagtime), int(use_fft), ref, hob, hint, r2, rinv)
    d = gu.dev(orb)
    dh, dg = gu.empty(norb * norb), gu.empty(norb ** 4)
    w.splitstep(d, dh, dg, gu.dev(r2), gu.dev(rinv), use_fft=use_fft)
    assert gu.rel_err(gu.host(d), ref) < 1e-10
    assert gu.rel_err(gu.host(dh), hob) < 1e-10
    assert gu.rel_err(gu.host(dg), hint) < 1e-10


def _full_setup(orc, amd, npar, norb, grid, imagtime, g, dt, L=16.0, lm=5, prec=0):
    x = np.linspace(-L / 2, L / 2, grid)
    dx = x[1] - x[0]
    pot = 0.5 * x ** 2
    eq = orc.make_orbeq(norb, grid, x[0], x[-1], dt, g, pot, imagtime=imagtime)
    orb = np.ascontiguousarray(inputs.hermite_gauss(norb, x, seed=1, noise=1e-3).ravel())
    orc.lib().orc_orthonormalize(grid, dx, norb, orb)
    S = orc.Space(npar, norb)
    coef = inputs.coefficients(S.dim, 1)
    r1, r2, rinv = S.sync_density(coef)
    r1 = np.ascontiguousarray(r1.ravel())
    rinv = np.ascontiguousarray(rinv.ravel())
    hob = np.zeros(norb * norb, complex)
    hint = np.zeros(norb ** 4, complex)
    orc.lib().orc_orbital_hob(eq, orb, hob)
    orc.lib().orc_orbital_hint(eq, orb, hint)
    D = amd.DeviceSpace(npar, norb, tables=True)
    sil = amd.DeviceSil(D, lm, prec=prec)
    w = amd.DeviceOrbital(norb, grid, x[0], x[-1], dt, g, pot, imagtime=imagtime)
    host = dict(orb=orb, coef=coef, hob=hob, hint=hint, r1=r1, r2=r2, rinv=rinv)
    dev = {k: gu.dev(v) for k, v in host.items()}
    return S, eq, D, sil, w, host, dev


@pytest.mark.parametrize("imagtime", [True, False])
def test_propagate_step_matches_oracle(orc, amd, imagtime):
    print(f"propagate imag={imagtime}", flush=True)
    npar, norb, lm = 5, 3, 5
    S, eq, D, sil, w, h, d = _full_setup(orc, amd, npar, norb, 129, imagtime, 1.0, 0.01, lm=lm)
    lvec = np.zeros(lm * S.dim, complex)
    for _ in range(5):
        orc.lib().orc_propagate_step(S._p, eq, int(imagtime), 1, lm, h["orb"], h["coef"], h["hob"], h["hint"],
                                     h["r1"], h["r2"], h["rinv"], lvec)
        amd.propagate_step(D, sil, w, d["orb"], d["coef"], d["hob"], d["hint"], d["r1"], d["r2"], d["rinv"])
    for k in ("orb", "coef", "hob", "hint", "r1", "r2", "rinv"):
        assert gu.rel_err(gu.host(d[k]), h[k]) < 1e-9, k


def test_ground_state_known_answer_on_device(orc, amd):
    """benchmark.dat:23: harmonic trap N=4, M=3, g=1, E/N = 0.9685859."""
    npar, norb = 4, 3
    S, eq, D, sil, w, h, d = _full_setup(orc, amd, npar, norb, 257, True, 1.0, 0.01)
    for _ in range(1500):
        amd.propagate_step(D, sil, w, d["orb"], d["coef"], d["hob"], d["hint"], d["r1"], d["r2"], d["rinv"])
    r1 = gu.host(d["r1"])
    e = orc.total_energy(norb, r1, gu.host(d["r2"]), gu.host(d["hob"]), gu.host(d["hint"])).real / npar
    assert abs(e - 0.9685859) < 3e-4, e


@pytest.mark.parametrize("imagtime,use_fft,prec", [(True, True, 0), (False, True, 0), (False, False, 0),
                                                   (False, True, 1), (True, False, 1)])
def test_step_graph_replays_match_eager_steps(orc, amd, imagtime, use_fft, prec):
    """The captured step, replayed, gives after every replay the same arrays
    as the eager step on a second set of buffers: the same kernels run in the
    same order on the same data, so the bar is bit-exact.  Eager steps, copies
    to the host and a trap change run between replays."""
    import torch

    print(f"step graph imag={imagtime} fft={use_fft} prec={prec}", flush=True)
    npar, norb, grid, lm, dt, g = 6, 4, 129, 5, 0.01, 1.0
    S, eq, D, sil, w, h, d = _full_setup(orc, amd, npar, norb, grid, imagtime, g, dt, lm=lm, prec=prec)
    if prec:
        D.hc_tiling(0)  # both step forms then run the same H.C tiling
    x = np.linspace(-8.0, 8.0, grid)
    sil_g = amd.DeviceSil(D, lm, prec=prec)
    w_g = amd.DeviceOrbital(norb, grid, x[0], x[-1], dt, g, 0.5 * x ** 2, imagtime=imagtime)
    d_g = {k: gu.dev(v) for k, v in h.items()}
    graph = amd.StepGraph(D, sil_g, w_g, d_g["orb"], d_g["coef"], d_g["hob"], d_g["hint"], d_g["r1"], d_g["r2"],
                          d_g["rinv"], use_fft=use_fft)
    for r in range(6):
        if r == 3:  # the trap is read at each launch
            pot = 0.5 * x ** 2 + 0.3 * np.cos(x)
            w.set_potential(pot)
            w_g.set_potential(pot)
        amd.propagate_step(D, sil, w, d["orb"], d["coef"], d["hob"], d["hint"], d["r1"], d["r2"], d["rinv"],
                           use_fft=use_fft)
        graph.launch()
        torch.cuda.synchronize()
        for k in ("orb", "coef", "hob", "hint", "r1", "r2", "rinv"):
            a, b = gu.host(d_g[k]), gu.host(d[k])
            assert np.array_equal(a, b), (r, k, gu.rel_err(a, b))
    graph.close()
