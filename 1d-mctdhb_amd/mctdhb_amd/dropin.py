"""ctypes mirror of the reference's whole-problem structs and the
reference-named time-step entry points of libmctdhb_hip.so.

``MCTDHB`` owns a reference ``_MCTDHBDataStruct`` built by the library's
``get_mctdhb_struct`` (reference types_definition.c:246-314) and drives it
with ``sync_density_matrices`` / ``sync_orbital_matrices`` (synchronize.c:
34-56) and ``mctdhb_propagate_step`` (driver.c:158-202): host arrays in the
reference's layout, the work done on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _MultiConfiguration, lib

IMAGTIME, REALTIME = 10, 11
LANCZOS, RUNGEKUTTA = 20, 21
FULLSTEP_RUNGEKUTTA, SPLITSTEP = 30, 31
DVR, SPECTRAL, FINITEDIFF = 300, 301, 302
ZERO_BOUNDS, PERIODIC_BOUNDS = 1000, 1001

STR_BUFF_SIZE = 128


class Complex(C.Structure):
    """C99 double complex (passed and returned like two doubles)."""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


POT_FUNC = C.CFUNCTYPE(None, C.c_double, C.c_uint16, C.POINTER(C.c_double), C.c_void_p, C.POINTER(C.c_double))
INTER_FUNC = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)

_cmat = C.POINTER(C.POINTER(Complex))


class _OrbitalEquation(C.Structure):
    """reference _OrbitalEquation (mctdhb_types.h:136-149)."""
    _fields_ = [("eq_name", C.c_char * STR_BUFF_SIZE), ("norb", C.c_uint16), ("grid_size", C.c_uint16),
                ("bounds", C.c_int), ("t", C.c_double), ("xi", C.c_double), ("xf", C.c_double),
                ("dx", C.c_double), ("tstep", C.c_double), ("tend", C.c_double), ("g", C.c_double),
                ("d2coef", C.c_double), ("d1coef", Complex), ("prop_dt", Complex), ("time_fac", Complex),
                ("grid_pts", C.POINTER(C.c_double)), ("pot_grid", C.POINTER(C.c_double)),
                ("pot_extra_args", C.c_void_p), ("inter_extra_args", C.c_void_p), ("pot_func", POT_FUNC),
                ("inter_param", INTER_FUNC)]


class _ManyBodyState(C.Structure):
    """reference _ManyBodyState (mctdhb_types.h:175-181)."""
    _fields_ = [("npar", C.c_uint16), ("norb", C.c_uint16), ("grid_size", C.c_uint16), ("space_dim", C.c_uint32),
                ("coef", C.POINTER(Complex)), ("hint", C.POINTER(Complex)), ("tb_denmat", C.POINTER(Complex)),
                ("hob", _cmat), ("orbitals", _cmat), ("ob_denmat", _cmat), ("inv_ob_denmat", _cmat)]


class _MCTDHBDataStruct(C.Structure):
    """reference _MCTDHBDataStruct (mctdhb_types.h:228-240)."""
    _fields_ = [("integ_type", C.c_int), ("coef_integ_method", C.c_int), ("orb_integ_method", C.c_int),
                ("orb_der_method", C.c_int), ("rk_order", C.c_int),
                ("multiconfig_space", C.POINTER(_MultiConfiguration)),
                ("orb_eq", C.POINTER(_OrbitalEquation)), ("coef_workspace", C.c_void_p),
                ("orb_workspace", C.c_void_p), ("state", C.POINTER(_ManyBodyState))]


def _declare(L):
    if getattr(L, "_dropin_declared", False):
        return
    P = C.POINTER(_MCTDHBDataStruct)
    L.get_mctdhb_struct.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint16, C.c_uint16,
                                    C.c_uint16, C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_uint16,
                                    C.c_double, C.c_double, C.c_double, Complex, C.c_void_p, C.c_void_p,
                                    POT_FUNC, INTER_FUNC]
    L.get_mctdhb_struct.restype = P
    L.destroy_mctdhb_struct.argtypes = [P]
    L.sync_orbital_matrices.argtypes = [C.POINTER(_OrbitalEquation), C.POINTER(_ManyBodyState)]
    L.sync_density_matrices.argtypes = [C.POINTER(_MultiConfiguration), C.POINTER(_ManyBodyState)]
    L.mctdhb_propagate_step.argtypes = [P, C.c_void_p, C.c_void_p, C.c_uint32]
    L.propagate_splitstep_orb.argtypes = [P, C.c_void_p]
    L.total_energy.argtypes = [C.POINTER(_ManyBodyState)]
    L.total_energy.restype = Complex
    L.set_regulatization_factor.argtypes = [C.c_double]
    # host state, text formats, builtin functions and the job loop
    L.get_manybody_state.argtypes = [C.c_uint16, C.c_uint16, C.c_uint16]
    L.get_manybody_state.restype = C.POINTER(_ManyBodyState)
    L.destroy_manybody_sate.argtypes = [C.POINTER(_ManyBodyState)]
    for name in ("record_raw_state", "append_processed_state", "set_orbitals_from_file", "set_coef_from_file"):
        getattr(L, name).argtypes = [C.c_char_p, C.POINTER(_ManyBodyState)]
    L.append_timestep_potential.argtypes = [C.c_char_p, C.POINTER(_OrbitalEquation)]
    L.record_time_interaction.argtypes = [C.c_char_p, C.POINTER(_OrbitalEquation), C.c_uint16]
    L.record_time_array.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_uint16]
    L.append_mctdhb_parameters.argtypes = [C.c_char_p, P]
    L.set_output_fname.argtypes = [C.c_char_p, C.c_int, C.c_char_p]
    L.auto_number_of_jobs.argtypes = [C.c_char_p]
    L.auto_number_of_jobs.restype = C.c_uint32
    L.get_mctdhb_from_files.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]
    L.get_mctdhb_from_files.restype = P
    L.full_setup_mctdhb_current_dir.argtypes = [C.c_char_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_void_p]
    L.full_setup_mctdhb_current_dir.restype = P
    L.integration_driver.argtypes = [P, C.c_uint16, C.c_char_p, C.c_double, C.c_uint8]
    for name in ("set_final_diagonalization", "set_autoconvergence_check"):
        getattr(L, name).argtypes = [C.c_int]
    L.set_energy_convergence_digits.argtypes = [C.c_uint8]
    L.set_energy_convergence_eig_residual.argtypes = [C.c_double]
    L.set_overlap_residue_threshold.argtypes = [C.c_double]
    L.set_overlap_matrix.argtypes = [C.c_uint16, C.c_uint16, C.c_double, _cmat, _cmat]
    L.overlap_residual.argtypes = [C.c_uint16, C.c_uint16, C.c_double, _cmat]
    L.overlap_residual.restype = C.c_double
    L.avg_orbitals_norm.argtypes = [C.c_uint16, C.c_uint16, C.c_double, _cmat]
    L.avg_orbitals_norm.restype = C.c_double
    L.get_builtin_pot.argtypes = [C.c_char_p]
    L.get_builtin_pot.restype = C.c_void_p
    L.get_builtin_param_func.argtypes = [C.c_char_p]
    L.get_builtin_param_func.restype = C.c_void_p
    L.timeparam_linear_ramp.argtypes = [C.c_double, C.c_void_p]
    L.timeparam_linear_ramp.restype = C.c_double
    L._dropin_declared = True


COMMON_INP, MULTIPLE_INP, LAST_JOB_OUT = 0, 1, 2
(ORBITALS_REC, COEFFICIENTS_REC, ONE_BODY_MATRIX_REC, TWO_BODY_MATRIX_REC, PARAMETERS_REC,
 ONE_BODY_POTENTIAL_REC) = range(6)


def set_directories(inp: str, out: str, integrator_conf: str | None = None) -> None:
    """The reference's global input/output directory names and integrator
    description file (dataio.c:16-18); directories end with '/'."""
    L = lib()
    _declare(L)
    for name, v in (("inp_dirname", inp), ("out_dirname", out), ("integrator_desc_fname", integrator_conf)):
        if v is None:
            continue
        if name != "integrator_desc_fname" and not v.endswith("/"):
            v += "/"
        b = v.encode()
        if len(b) >= STR_BUFF_SIZE:
            raise ValueError(f"{name} longer than {STR_BUFF_SIZE - 1} bytes")
        (C.c_char * STR_BUFF_SIZE).in_dll(L, name).value = b


def builtin_potential(name: str, t: float, x, params) -> np.ndarray:
    """V(x) at time t from the library's builtin trap ``name``."""
    L = lib()
    _declare(L)
    f = L.get_builtin_pot(name.encode())
    if not f:
        raise ValueError(f"{name} is not a builtin potential")
    x = np.ascontiguousarray(x, np.float64)
    par = np.ascontiguousarray(params, np.float64)
    out = np.zeros_like(x)
    POT_FUNC(f)(t, x.size, x.ctypes.data_as(C.POINTER(C.c_double)), par.ctypes.data, out.ctypes.data_as(
        C.POINTER(C.c_double)))
    return out


class _StateView:
    """get/set of the reference state arrays through ``self._state``."""

    def _row(self, ptr, n):
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), (2 * n,)).view(np.complex128)

    def _matrix(self, rows_ptr, nrows, ncols):
        return [self._row(rows_ptr[r], ncols) for r in range(nrows)]

    def get(self, name):
        s, M = self._state, self.norb
        if name == "coef":
            return self._row(s.coef, self.dim).copy()
        if name in ("hint", "tb_denmat"):
            return self._row(getattr(s, name), M ** 4).copy()
        if name == "orbitals":
            return np.array(self._matrix(s.orbitals, M, self.grid_size))
        return np.array(self._matrix(getattr(s, name), M, M))

    def set(self, name, value):
        s, M = self._state, self.norb
        value = np.asarray(value, np.complex128)
        if name == "coef":
            self._row(s.coef, self.dim)[:] = value
        elif name in ("hint", "tb_denmat"):
            self._row(getattr(s, name), M ** 4)[:] = value.ravel()
        elif name == "orbitals":
            for r, row in enumerate(self._matrix(s.orbitals, M, self.grid_size)):
                row[:] = value.reshape(M, self.grid_size)[r]
        else:
            for r, row in enumerate(self._matrix(getattr(s, name), M, M)):
                row[:] = value.reshape(M, M)[r]


class HostState(_StateView):
    """A reference ``_ManyBodyState`` from ``get_manybody_state`` (host
    arrays only; no device work), for the text-format entry points."""

    def __init__(self, npar, norb, grid_size):
        L = lib()
        _declare(L)
        self._p = L.get_manybody_state(npar, norb, grid_size)
        self.npar, self.norb, self.grid_size = npar, norb, grid_size
        self.dim = int(self._p.contents.space_dim)

    @property
    def _state(self):
        return self._p.contents

    def close(self):
        if getattr(self, "_p", None):
            lib().destroy_manybody_sate(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except TypeError:  # interpreter shutdown: module globals already cleared
            pass


class MCTDHB(_StateView):
    """One reference-shaped MCTDHB problem.

    ``pot(t, x) -> V(x)`` and ``g(t) -> float`` are called back from C the way
    the reference calls ``single_particle_pot`` / ``time_dependent_parameter``.
    """

    def __init__(self, npar, norb, grid_size, xi, xf, tstep, tend, pot, g, imagtime=True, orb_der=SPECTRAL,
                 rk_order=4, lanczos_iter=5, bounds=PERIODIC_BOUNDS):
        L = lib()
        _declare(L)

        def pot_cb(t, npts, pts, _args, out):
            x = np.ctypeslib.as_array(pts, (npts,))
            np.ctypeslib.as_array(out, (npts,))[:] = pot(t, x)

        def g_cb(t, _args):
            return float(g(t))

        self._pot_cb, self._g_cb = POT_FUNC(pot_cb), INTER_FUNC(g_cb)
        self._p = L.get_mctdhb_struct(IMAGTIME if imagtime else REALTIME, LANCZOS, SPLITSTEP, orb_der, rk_order,
                                      lanczos_iter, npar, norb, b"mctdhb", bounds, xi, xf, grid_size, tstep, tend,
                                      -0.5, Complex(0.0, 0.0), None, None, self._pot_cb, self._g_cb)
        st = self._p.contents.state.contents
        self.npar, self.norb, self.grid_size, self.dim = npar, norb, grid_size, int(st.space_dim)
        self._orb_work = np.zeros(norb * grid_size, np.complex128)
        self._coef_work = np.zeros(self.dim, np.complex128)
        self.step_count = 0

    def close(self):
        if getattr(self, "_p", None):
            lib().destroy_mctdhb_struct(self._p)
            self._p = None

    __del__ = close

    @classmethod
    def from_files(cls, prefix: str, job: int = 1, which_inp: int = COMMON_INP) -> "MCTDHB":
        """reference full_setup_mctdhb_current_dir (dataio.c:295-378): the
        parameters line ``job`` of <inp>/<prefix>_mctdhb_parameters.dat, the
        integrator description, orbitals and coefficients from files."""
        L = lib()
        _declare(L)
        self = cls.__new__(cls)
        self._p = L.full_setup_mctdhb_current_dir(prefix.encode(), job, which_inp, None, None, None, None)
        st = self._p.contents.state.contents
        self.npar, self.norb, self.grid_size, self.dim = int(st.npar), int(st.norb), int(st.grid_size), \
            int(st.space_dim)
        self._orb_work = np.zeros(self.norb * self.grid_size, np.complex128)
        self._coef_work = np.zeros(self.dim, np.complex128)
        self.step_count = 0
        return self

    # ---- host views of the state arrays (reference layout)
    @property
    def _state(self):
        return self._p.contents.state.contents

    def run(self, prefix: str, tend: float | None = None, rec_nsteps: int = 10, monitor_rate: int = 10):
        """reference integration_driver (driver.c:203-333)."""
        if tend is None:
            tend = self._p.contents.orb_eq.contents.tend
        lib().integration_driver(self._p, rec_nsteps, prefix.encode(), tend, monitor_rate)

    def append_parameters(self, prefix: str):
        """reference append_mctdhb_parameters (dataio.c:836-866)."""
        lib().append_mctdhb_parameters(prefix.encode(), self._p)

    # ---- reference entry points
    def sync(self):
        """sync_density_matrices, then sync_orbital_matrices (the order of
        reference sync_integration_new_step, synchronize.c:70-77)."""
        p = self._p.contents
        lib().sync_density_matrices(p.multiconfig_space, p.state)
        lib().sync_orbital_matrices(p.orb_eq, p.state)

    def step(self):
        lib().mctdhb_propagate_step(self._p, self._orb_work.ctypes.data, self._coef_work.ctypes.data,
                                    self.step_count)
        self.step_count += 1

    def energy(self) -> complex:
        e = lib().total_energy(self._p.contents.state)
        return complex(e.re, e.im)
