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
    L._dropin_declared = True


class MCTDHB:
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

    # ---- host views of the state arrays (reference layout)
    @property
    def _state(self):
        return self._p.contents.state.contents

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
