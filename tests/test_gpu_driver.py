"""The job loop end to end on the GPU: reference-format input files ->
full_setup_mctdhb_current_dir -> integration_driver -> output files
(reference dataio.c:295-378, driver.c:203-333, dataio.c:836-866).

Bars: the ground states reproduce benchmark.dat's known answers (energy
per particle to 5e-4, condensate fraction to 1.2 percent points); a fixed
number of driver steps matches the oracle's steps from the same files at
1e-9 relative, in imaginary time and in real time with a time-dependent
trap and interaction (evaluated between the orbital step and the h/g
rebuild, the reference's order, driver.c:175-186)."""
import math

import numpy as np
import pytest

import oracle
from mctdhb_amd import inputs

pytestmark = pytest.mark.gpu

INTEGRATOR = """# integrator description in the reference's format
# time: 10 imaginary, 11 real
{time}
# coefficients: 20 Lanczos
20
# orbitals: 30 full-equation Runge-Kutta, 31 split-step
{orb}
# derivatives: 301 FFT, 302 finite differences
{der}
# Runge-Kutta order
{rk}
# boundaries
1001
# Lanczos iterations
5
"""


@pytest.fixture(scope="module")
def dropin():
    from mctdhb_amd import dropin as d

    d.lib()
    return d


def _write_inputs(tmp_path, prefix, line, npar, norb, grid, xi, xf, time=10, der=301, seed=1, orb_method=31, rk=4):
    inp, out = tmp_path / "in", tmp_path / "out"
    inp.mkdir(exist_ok=True)
    out.mkdir(exist_ok=True)
    (inp / f"{prefix}_mctdhb_parameters.dat").write_text(line + "\n")
    conf = tmp_path / "integrator.conf"
    conf.write_text(INTEGRATOR.format(time=time, der=der, orb=orb_method, rk=rk))
    x = np.linspace(xi, xf, grid)
    orb = inputs.hermite_gauss(norb, x, seed=seed, noise=1e-3)
    orb[:, -1] = orb[:, 0]
    coef = inputs.coefficients(math.comb(npar + norb - 1, norb - 1), seed)
    np.savetxt(inp / f"{prefix}_job1_orb.dat", orb.T, fmt="%.14E")  # scripts/initial_condition.py:177-178
    np.savetxt(inp / f"{prefix}_job1_coef.dat", coef.T, fmt="%.14E")
    return inp, out, conf


@pytest.fixture()
def job_dirs(dropin, tmp_path):
    yield tmp_path
    dropin.set_directories("input/", "output/", "mctdhb_integrator.conf")
    L = dropin.lib()
    L.set_final_diagonalization(1)
    L.set_autoconvergence_check(0)


@pytest.mark.parametrize("case", [
    # benchmark.dat:23: harmonic trap, N=4, M=3, g=1: E/N 0.9685859, 97 %
    ("4 3 257 -8.0 8.0 0.01 12.0 -0.5 0.0 linear_ramp 1.0 1.0 0.0 0.0 0.0 harmonic 1.0 1.0 0.0 0.0 0.0",
     0.9685859, 0.97),
    # benchmark.dat:25: N=4, M=3, g=10: E/N 2.7420095, 65 %
    ("4 3 257 -8.0 8.0 0.01 12.0 -0.5 0.0 linear_ramp 10.0 10.0 0.0 0.0 0.0 harmonic 1.0 1.0 0.0 0.0 0.0",
     2.7420095, 0.65),
    # benchmark.dat:20: N=5, M=5, g=1: E/N 1.0889904, 96 %
    ("5 5 257 -8.0 8.0 0.01 30.0 -0.5 0.0 linear_ramp 1.0 1.0 0.0 0.0 0.0 harmonic 1.0 1.0 0.0 0.0 0.0",
     1.0889904, 0.96),
    # benchmark.dat:43: 0.5 x^2 + 9 exp(-x^2 / 0.5), N=10, M=4, g=2: E/N 4.1038147, 68 %
    ("10 4 257 -8.0 8.0 0.01 30.0 -0.5 0.0 linear_ramp 2.0 2.0 0.0 0.0 0.0 "
     "harmonicgauss 1.0 0.7071067811865476 9.0 9.0 0.0", 4.1038147, 0.68),
])
def test_job_from_reference_files_reaches_benchmark(dropin, job_dirs, case):
    line, e_ref, frac_ref = case
    print(f"driver job {line[:12]}", flush=True)
    npar, norb, grid = (int(v) for v in line.split()[:3])
    inp, out, conf = _write_inputs(job_dirs, "bench", line, npar, norb, grid, -8.0, 8.0)
    dropin.set_directories(str(inp), str(out), str(conf))
    m = dropin.MCTDHB.from_files("bench")
    m.run("bench_job1", rec_nsteps=10, monitor_rate=200)
    m.append_parameters("bench")
    par = np.loadtxt(out / "bench_mctdhb_parameters.dat")
    assert par.shape == (11,) and int(par[0]) == npar and int(par[1]) == norb
    assert abs(par[10] - e_ref) < 5e-4, par[10]
    occ = np.linalg.eigvalsh(m.get("ob_denmat"))
    assert abs(occ[-1] / npar - frac_ref) < 0.012, occ / npar  # the file rounds to whole percent
    # outputs numpy reads back: the final coefficients are normalised
    c = np.loadtxt(out / "bench_job1_coef.dat", dtype=np.complex128)
    assert c.shape == (m.dim,) and abs(np.linalg.norm(c) - 1) < 1e-10
    assert np.loadtxt(out / "bench_job1_orb.dat", dtype=np.complex128).shape == (grid, norb)
    assert np.loadtxt(out / "bench_job1_init_coef.dat", dtype=np.complex128).shape == (m.dim,)
    assert np.loadtxt(out / "bench_job1_obpotential.dat").shape == (grid,)
    m.close()


def _oracle_start(inp, prefix, npar, norb, grid, xi, xf, dt, g, pot, imagtime):
    eq = oracle.make_orbeq(norb, grid, xi, xf, dt, g, pot, imagtime=imagtime)
    orb = np.ascontiguousarray(np.loadtxt(inp / f"{prefix}_job1_orb.dat", dtype=np.complex128).T.ravel())
    coef = np.ascontiguousarray(np.loadtxt(inp / f"{prefix}_job1_coef.dat", dtype=np.complex128))
    S = oracle.Space(npar, norb)
    r1, r2, rinv = S.sync_density(coef)
    h = dict(orb=orb, coef=coef, r1=np.ascontiguousarray(r1.ravel()), r2=np.ascontiguousarray(r2),
             rinv=np.ascontiguousarray(rinv.ravel()), hob=np.zeros(norb * norb, complex),
             hint=np.zeros(norb ** 4, complex))
    oracle.lib().orc_orbital_hob(eq, h["orb"], h["hob"])
    oracle.lib().orc_orbital_hint(eq, h["orb"], h["hint"])
    return S, eq, h


@pytest.mark.parametrize("der", [301, 302])
def test_imaginary_time_driver_matches_oracle_steps(dropin, job_dirs, der):
    print(f"driver imag der={der}", flush=True)
    npar, norb, grid, dt, steps = 5, 3, 129, 0.01, 12
    line = (f"{npar} {norb} {grid} -6.0 6.0 {dt} {steps * dt - dt / 2} -0.5 0.0 linear_ramp 1.5 1.5 0.0 0.0 0.0 "
            "harmonic 1.0 1.0 0.0 0.0 0.0")
    inp, out, conf = _write_inputs(job_dirs, "im", line, npar, norb, grid, -6.0, 6.0, der=der, seed=4)
    dropin.set_directories(str(inp), str(out), str(conf))
    dropin.lib().set_final_diagonalization(0)
    m = dropin.MCTDHB.from_files("im")
    m.run("im_job1", rec_nsteps=5, monitor_rate=5)
    x = np.linspace(-6, 6, grid)
    S, eq, h = _oracle_start(inp, "im", npar, norb, grid, -6.0, 6.0, dt, 1.5, 0.5 * x ** 2, True)
    lvec = np.zeros(5 * S.dim, complex)
    for _ in range(steps):
        oracle.lib().orc_propagate_step(S._p, eq, 1, int(der == 301), 5, h["orb"], h["coef"], h["hob"], h["hint"],
                                        h["r1"], h["r2"], h["rinv"], lvec)
    c = np.loadtxt(out / "im_job1_coef.dat", dtype=np.complex128)
    o = np.loadtxt(out / "im_job1_orb.dat", dtype=np.complex128).T.ravel()
    for got, ref in ((c, h["coef"]), (o, h["orb"]), (m.get("ob_denmat").ravel(), h["r1"]),
                     (m.get("hint"), h["hint"])):
        assert np.max(abs(got - ref)) / np.max(abs(ref)) < 1e-9
    m.close()


def test_real_time_driver_with_time_dependent_trap_matches_oracle(dropin, job_dirs):
    """omega 1 -> 1.3 and g 1 -> 2 over t = 0.06 (harmonic / linear_ramp),
    10 real-time steps of 0.01 recorded every 5."""
    print("driver real time", flush=True)
    npar, norb, grid, dt, steps, rec = 5, 3, 129, 0.01, 10, 5
    line = (f"{npar} {norb} {grid} -6.0 6.0 {dt} {steps * dt - dt / 2} -0.5 0.0 linear_ramp 1.0 2.0 0.06 0.0 0.0 "
            "harmonic 1.0 1.3 0.06 0.0 0.0")
    inp, out, conf = _write_inputs(job_dirs, "rt", line, npar, norb, grid, -6.0, 6.0, time=11, seed=6)
    dropin.set_directories(str(inp), str(out), str(conf))
    m = dropin.MCTDHB.from_files("rt")
    m.run("rt_job1", rec_nsteps=rec, monitor_rate=5)
    x = np.linspace(-6, 6, grid)

    def omega(t):
        return 1.0 + 0.3 * t / 0.06

    def g_of(t):
        return 2.0 if t >= 0.06 else 1.0 + t / 0.06

    S, eq, h = _oracle_start(inp, "rt", npar, norb, grid, -6.0, 6.0, dt, g_of(0), 0.5 * omega(0) ** 2 * x ** 2,
                             False)
    for n in range(steps):  # driver.c:158-202 for REALTIME
        oracle.lib().orc_splitstep_orb(eq, 0, 1, h["orb"], h["hob"], h["hint"], h["r2"], h["rinv"])
        t = (n + 1) * dt
        eq._pot_keep[:] = 0.5 * omega(t) ** 2 * x ** 2
        eq.g = g_of(t)
        h["orb"].reshape(norb, grid)[:, -1] = h["orb"].reshape(norb, grid)[:, 0]
        oracle.lib().orc_orbital_hob(eq, h["orb"], h["hob"])
        oracle.lib().orc_orbital_hint(eq, h["orb"], h["hint"])
        _, h["coef"] = S.sil_step(h["coef"], h["hob"], h["hint"], 5, dt, False)
        r1, r2, ri = S.sync_density(h["coef"])
        h["r1"], h["r2"], h["rinv"] = r1.ravel(), np.ascontiguousarray(r2), np.ascontiguousarray(ri.ravel())
    for got, ref in ((m.get("coef"), h["coef"]), (m.get("orbitals").ravel(), h["orb"]),
                     (m.get("ob_denmat").ravel(), h["r1"]), (m.get("hob").ravel(), h["hob"])):
        assert np.max(abs(got - ref)) / np.max(abs(ref)) < 1e-9
    # records: the initial state plus every 5th step; g(t) and the trap
    assert np.loadtxt(out / "rt_job1_obmat.dat", dtype=np.complex128).shape == (3, norb * norb)
    pots = np.loadtxt(out / "rt_job1_obpotential.dat")
    assert pots.shape == (3, grid) and np.allclose(pots[2], 0.5 * omega(0.1) ** 2 * x ** 2)
    assert np.allclose(np.loadtxt(out / "rt_job1_timesteps.dat"), [0.0, 0.05, 0.1])
    m.close()


def test_reference_example_integrator_full_step_reaches_benchmark(dropin, job_dirs):
    """The integrator of the reference's example file
    (config_files_example/mctdhb_integrator.txt: full-equation RK, finite
    differences, RK5) in imaginary time: benchmark.dat:23, N=4, M=3, g=1."""
    print("driver full step", flush=True)
    line = "4 3 257 -8.0 8.0 0.005 12.0 -0.5 0.0 linear_ramp 1.0 1.0 0.0 0.0 0.0 harmonic 1.0 1.0 0.0 0.0 0.0"
    inp, out, conf = _write_inputs(job_dirs, "fs", line, 4, 3, 257, -8.0, 8.0, der=302, orb_method=30, rk=5)
    dropin.set_directories(str(inp), str(out), str(conf))
    m = dropin.MCTDHB.from_files("fs")
    m.run("fs_job1", rec_nsteps=10, monitor_rate=400)
    m.append_parameters("fs")
    par = np.loadtxt(out / "fs_mctdhb_parameters.dat")
    assert abs(par[10] - 0.9685859) < 5e-4, par[10]
    occ = np.linalg.eigvalsh(m.get("ob_denmat"))
    assert abs(occ[-1] / 4 - 0.97) < 0.012
    m.close()


def test_reference_default_integrator_full_step_dvr(dropin, job_dirs, tmp_path):
    """No integrator file: the reference's defaults (mctdhb_types.h:93-99:
    imaginary time, Lanczos, full-equation RK5 with the DVR linear part,
    periodic, 5 Lanczos vectors) reach benchmark.dat:23 (N=4, M=3, g=1)."""
    print("driver defaults (DVR)", flush=True)
    line = "4 3 129 -8.0 8.0 0.005 12.0 -0.5 0.0 linear_ramp 1.0 1.0 0.0 0.0 0.0 harmonic 1.0 1.0 0.0 0.0 0.0"
    inp, out, conf = _write_inputs(job_dirs, "df", line, 4, 3, 129, -8.0, 8.0)
    dropin.set_directories(str(inp), str(out), str(tmp_path / "no_such_integrator.conf"))
    m = dropin.MCTDHB.from_files("df")
    m.run("df_job1", rec_nsteps=10, monitor_rate=400)
    m.append_parameters("df")
    par = np.loadtxt(out / "df_mctdhb_parameters.dat")
    assert abs(par[10] - 0.9685859) < 5e-4, par[10]
    m.close()
