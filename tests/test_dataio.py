"""Text formats, builtin trap/interaction functions and the host monitors
of libmctdhb_hip.so (reference dataio.c, builtin_potential.c,
builtin_time_parameter.c, integrator_monitor.c:30-53).  Host code only: no
GPU is touched.  Pinned by closed forms, by files written the way the
reference's own input generator writes them (scripts/initial_condition.py:
177-178, numpy.savetxt with fmt "%.14E"), by numpy.loadtxt (how the
reference's Python package reads outputs, multiconfpy/stationary.py:87-88)
and by the oracle's Simpson rule."""
import ctypes as C
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from mctdhb_amd import inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def dropin():
    from mctdhb_amd import dropin as d

    d.lib()
    return d


@pytest.fixture()
def dirs(dropin, tmp_path):
    inp, out = tmp_path / "in", tmp_path / "out"
    inp.mkdir()
    out.mkdir()
    dropin.set_directories(str(inp), str(out))
    yield inp, out
    dropin.set_directories("input/", "output/")


# ------------------------------------------------------------ builtins
def test_harmonic_sweep(dropin):
    x = np.linspace(-3, 3, 61)
    assert np.allclose(dropin.builtin_potential("harmonic", 0.0, x, [1.0, 2.0, 4.0]), 0.5 * x ** 2, rtol=0,
                       atol=1e-15)
    # halfway through the sweep omega = 1.5; after it the value keeps growing linearly
    assert np.allclose(dropin.builtin_potential("harmonic", 2.0, x, [1.0, 2.0, 4.0]), 0.5 * 2.25 * x ** 2)
    assert np.allclose(dropin.builtin_potential("harmonic", 8.0, x, [1.0, 2.0, 4.0]), 0.5 * 9.0 * x ** 2)
    # period 0: the final value at once
    assert np.allclose(dropin.builtin_potential("harmonic", 0.0, x, [1.0, 2.0, 0.0]), 2.0 * x ** 2)


def test_doublewell_is_a_x2_minus_b2_squared(dropin):
    """V = a (x^2 - b^2)^2 is doublewell with p0 = sqrt(a), p1 = p2 = b sqrt(a)."""
    x = np.linspace(-4, 4, 81)
    for a, b in ((0.5, 1.0), (2.0, 3.0), (1.3, 1.7)):
        v = dropin.builtin_potential("doublewell", 0.7, x, [math.sqrt(a), b * math.sqrt(a), b * math.sqrt(a), 0.0])
        assert np.allclose(v, a * (x ** 2 - b ** 2) ** 2, rtol=1e-13, atol=1e-12)


def test_harmonic_gauss_barrier_of_the_benchmark(dropin):
    """benchmark.dat:33-35: 0.5 x^2 + b^2 exp(-x^2 / 2c^2), b = 3, c = 0.5."""
    x = np.linspace(-8, 8, 257)
    v = dropin.builtin_potential("harmonicgauss", 0.0, x, [1.0, math.sqrt(0.5), 9.0, 9.0, 0.0])
    assert np.allclose(v, 0.5 * x ** 2 + 9.0 * np.exp(-x ** 2 / 0.5), rtol=1e-14, atol=1e-14)


def test_barriers_lattices_steps(dropin):
    x = np.linspace(-2, 2, 401)
    w = 1.0
    v = dropin.builtin_potential("barrier", 0.0, x, [w, 4.0, 4.0, 0.0, 0.25])
    assert v[200] == pytest.approx(4.25)                       # x = 0: top + background
    assert v[np.argmin(abs(x - w / 4))] == pytest.approx(2.25)  # cos^2(pi/4) = 1/2
    assert np.all(v[abs(x) > w / 2 + 1e-12] == 0.25)
    # trapezoid in time: up over 1, hold 2, down over 1, then 0
    tb = [w, 6.0, 1.0, 2.0]
    assert dropin.builtin_potential("trapezoid_barrier", 0.5, x, tb)[200] == pytest.approx(3.0)
    assert dropin.builtin_potential("trapezoid_barrier", 2.0, x, tb)[200] == pytest.approx(6.0)
    assert dropin.builtin_potential("trapezoid_barrier", 3.5, x, tb)[200] == pytest.approx(3.0)
    assert dropin.builtin_potential("trapezoid_barrier", 5.0, x, tb)[200] == 0.0
    L = x[-1] - x[0]
    v = dropin.builtin_potential("opticallattice", 0.0, x, [2.0, 5.0, 5.0, 0.0])
    assert np.allclose(v, 5.0 * np.sin(np.pi * 2 * x / L) ** 2, atol=1e-13)
    v = dropin.builtin_potential("trapezoid_opticallattice", 0.5, x, [2.0, 5.0, 1.0, 1.0])
    assert np.allclose(v, 2.5 * np.sin(np.pi * 2 * x / L) ** 2, atol=1e-13)
    v = dropin.builtin_potential("potfunc_step", 1.0, x, [0.5, 0.0, 4.0, 2.0])
    assert np.all(v[x >= 0.5] == 2.0) and np.all(v[x < 0.5] == 0.0)
    # the reference's square-well mask: zero wherever x > xi or x < xf
    # (builtin_potential.c:211), kept as written
    v = dropin.builtin_potential("potfunc_square_well", 0.0, x, [-1.0, 1.0, 7.0, 7.0, 0.0])
    assert np.all(v == 0.0)
    v = dropin.builtin_potential("potfunc_square_well", 0.0, x, [1.0, -1.0, 7.0, 7.0, 0.0])
    assert np.all(v[(x <= 1.0) & (x >= -1.0)] == 7.0) and np.all(v[(x > 1.0) | (x < -1.0)] == 0.0)


def test_linear_ramp_and_lookup(dropin):
    L = dropin.lib()
    par = np.array([1.0, 3.0, 2.0])
    f = lambda t: L.timeparam_linear_ramp(t, par.ctypes.data)  # noqa: E731
    assert f(0.0) == 1.0 and f(1.0) == 2.0 and f(2.0) == 3.0 and f(9.0) == 3.0
    assert L.get_builtin_param_func(b"linear_ramp") == C.cast(L.timeparam_linear_ramp, C.c_void_p).value
    assert L.get_builtin_param_func(b"custom") is None
    assert L.get_builtin_pot(b"custom") is None


def _run_child(code):
    env = dict(os.environ, PYTHONPATH=os.pathsep.join([os.path.join(ROOT, "1d-mctdhb_amd"), ROOT]))
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=300)


def test_fail_stop_messages():
    """Unknown names and out-of-range settings stop the process with the
    reference's message and EXIT_FAILURE (builtin_potential.c:258-259,
    driver.c:119-127)."""
    r = _run_child("import mctdhb_amd.dropin as d; L = d.lib(); d._declare(L); L.get_builtin_pot(b'nosuch')")
    assert r.returncode == 1 and "Potential 'nosuch' not valid" in r.stdout
    r = _run_child("import mctdhb_amd.dropin as d; L = d.lib(); d._declare(L); L.set_energy_convergence_digits(14)")
    assert r.returncode == 1 and "too much due to roundoff" in r.stdout


# ------------------------------------------------------------ text files
def _fill(state, seed):
    x = np.linspace(-5, 5, state.grid_size)
    orb = inputs.hermite_gauss(state.norb, x, seed=seed, noise=0.1)
    coef = inputs.coefficients(state.dim, seed)
    state.set("orbitals", orb)
    state.set("coef", coef)
    return orb, coef


def test_raw_state_round_trip_is_exact(dropin, dirs):
    inp, out = dirs
    a = dropin.HostState(5, 3, 65)
    orb, coef = _fill(a, 3)
    dropin.lib().record_raw_state(b"rt", a._p)
    c_file, o_file = out / "rt_coef.dat", out / "rt_orb.dat"
    # numpy reads what is written, bit for bit
    assert np.array_equal(np.loadtxt(c_file, dtype=np.complex128), coef)
    assert np.array_equal(np.loadtxt(o_file, dtype=np.complex128).T, orb)
    # and so do the reference-format readers
    b = dropin.HostState(5, 3, 65)
    dropin.lib().set_orbitals_from_file(str(o_file).encode(), b._p)
    dropin.lib().set_coef_from_file(str(c_file).encode(), b._p)
    assert np.array_equal(b.get("orbitals"), orb) and np.array_equal(b.get("coef"), coef)


def test_reads_files_written_like_the_reference_generator(dropin, tmp_path):
    """scripts/initial_condition.py:177-178 writes orbitals.T and C with
    numpy.savetxt(fmt="%.14E"); a '#' comment line in front is skipped."""
    s = dropin.HostState(4, 3, 33)
    x = np.linspace(-4, 4, 33)
    orb = inputs.hermite_gauss(3, x, seed=1, noise=1e-3)
    coef = inputs.coefficients(s.dim, 1)
    np.savetxt(tmp_path / "o.dat", orb.T, fmt="%.14E")
    np.savetxt(tmp_path / "c.dat", coef.T, fmt="%.14E", header="coefficients")
    dropin.lib().set_orbitals_from_file(str(tmp_path / "o.dat").encode(), s._p)
    dropin.lib().set_coef_from_file(str(tmp_path / "c.dat").encode(), s._p)
    assert np.max(abs(s.get("orbitals") - orb)) <= 1e-14 * np.max(abs(orb))
    assert np.max(abs(s.get("coef") - coef)) <= 1e-14 * np.max(abs(coef))


def test_append_processed_state_lines(dropin, dirs):
    inp, out = dirs
    s = dropin.HostState(3, 2, 17)
    _fill(s, 2)
    M = 2
    rho1 = np.arange(M * M).reshape(M, M) * (1 + 0.5j)
    rho2 = np.arange(M ** 4) * (0.25 - 1j)
    s.set("ob_denmat", rho1)
    s.set("tb_denmat", rho2)
    L = dropin.lib()
    for _ in range(3):
        L.append_processed_state(b"ap", s._p)
    ob = np.loadtxt(out / "ap_obmat.dat", dtype=np.complex128)
    tb = np.loadtxt(out / "ap_tbmat.dat", dtype=np.complex128)
    orbs = np.loadtxt(out / "ap_orb.dat", dtype=np.complex128)
    assert ob.shape == (3, 4) and np.array_equal(ob[2], rho1.ravel())
    assert tb.shape == (3, 16) and np.array_equal(tb[0], rho2)
    assert orbs.shape == (3, 2 * 17) and np.array_equal(orbs[1], s.get("orbitals").ravel())
    # toggle_new_append_files: the next append starts the file afresh
    L.toggle_new_append_files()
    L.append_processed_state(b"ap", s._p)
    L.toggle_new_append_files()
    assert np.loadtxt(out / "ap_obmat.dat", dtype=np.complex128).shape == (4,)


def test_time_records(dropin, dirs):
    """record_time_array writes 0 and every recn-th step time; the
    interaction record writes g(0), then g at the steps that are NOT
    multiples of recn (the inverted test of dataio.c:804, kept)."""
    inp, out = dirs
    L = dropin.lib()
    L.record_time_array(b"tr", 1.0, 0.25, 2)
    assert np.allclose(np.loadtxt(out / "tr_timesteps.dat"), [0.0, 0.5, 1.0])
    par = np.array([1.0, 3.0, 1.0])
    eq = dropin._OrbitalEquation()
    eq.tend, eq.tstep = 1.0, 0.25
    eq.inter_param = dropin.INTER_FUNC(C.cast(L.timeparam_linear_ramp, C.c_void_p).value)
    eq.inter_extra_args = par.ctypes.data
    L.record_time_interaction(b"tr", C.byref(eq), 2)
    # steps 1, 3, 5 (t = 0.25, 0.75, 1.25), then t >= tend stops the loop
    assert np.allclose(np.loadtxt(out / "tr_interaction.dat"), [1.0, 1.5, 2.5, 3.0])


def test_output_names(dropin, dirs):
    inp, out = dirs
    L = dropin.lib()
    buf = C.create_string_buffer(128)
    for rec, suffix in ((dropin.ORBITALS_REC, "_orb.dat"), (dropin.COEFFICIENTS_REC, "_coef.dat"),
                        (dropin.PARAMETERS_REC, "_mctdhb_parameters.dat"), (dropin.ONE_BODY_MATRIX_REC, "_obmat.dat"),
                        (dropin.TWO_BODY_MATRIX_REC, "_tbmat.dat"),
                        (dropin.ONE_BODY_POTENTIAL_REC, "_obpotential.dat")):
        L.set_output_fname(b"job", rec, buf)
        assert buf.value.decode() == str(out) + "/job" + suffix
    (inp / "p_mctdhb_parameters.dat").write_text("# comment\n1 2 3\n\n4 5 6\n# tail\n")
    assert L.auto_number_of_jobs(b"p") == 2


# ------------------------------------------------------------ monitors
def test_overlap_residual_and_orbital_norm(dropin):
    """integrator_monitor.c:30-53 against the oracle's Simpson rule."""
    for grid in (64, 65):  # even: 3/8 tail; odd: plain composite rule
        s = dropin.HostState(3, 4, grid)
        orb, _ = _fill(s, 7)
        x = np.linspace(-5, 5, grid)
        dx = x[1] - x[0]
        ov = sum(abs(oracle.csimps(np.conj(orb[k]) * orb[l], dx)) for k in range(4) for l in range(k + 1, 4)) / 4
        nrm = np.mean([math.sqrt(oracle.csimps(abs(orb[k]) ** 2 + 0j, dx).real) for k in range(4)])
        st = s._p.contents
        L = dropin.lib()
        assert L.overlap_residual(4, grid, dx, st.orbitals) == pytest.approx(ov, rel=1e-13)
        assert L.avg_orbitals_norm(4, grid, dx, st.orbitals) == pytest.approx(nrm, rel=1e-13)


def test_set_overlap_matrix_writes_what_the_reference_writes(dropin):
    """orbital_matrices.c:7-20: the j loop starts at 2, so [0][1] and [1][0]
    keep the caller's values (quirk Q2); the rest is <f_i, f_j>."""
    M, grid = 4, 41
    s = dropin.HostState(3, M, grid)
    orb, _ = _fill(s, 9)
    sentinel = 7.0 + 7.0j
    s.set("ob_denmat", np.full((M, M), sentinel))
    dx = 10.0 / (grid - 1)
    st = s._p.contents
    dropin.lib().set_overlap_matrix(M, grid, dx, st.orbitals, st.ob_denmat)
    S = s.get("ob_denmat")
    for i in range(M):
        for j in range(M):
            if {i, j} == {0, 1}:
                assert S[i, j] == sentinel
            else:
                ref = oracle.csimps(np.conj(orb[i]) * orb[j], dx)
                assert abs(S[i, j] - ref) <= 1e-13 * abs(ref) + 1e-16
