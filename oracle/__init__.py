"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the GPU library is compared against.
Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this module.  Each wrapper names the reference function its C code
restates (see the C sources for file:line citations).

Complex arrays are numpy complex128; matrices are contiguous row-major.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(np.complex128, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class _Space(C.Structure):
    _fields_ = [
        ("npar", C.c_uint16), ("norb", C.c_uint16), ("dim", C.c_uint32),
        ("hash_table", C.POINTER(C.c_uint16)), ("subsdim", C.POINTER(C.c_uint32)),
        ("map", C.POINTER(C.c_uint32)), ("mapot", C.POINTER(C.c_uint32)),
        ("maptt", C.POINTER(C.c_uint32)), ("strideot", C.POINTER(C.c_uint32)),
        ("stridett", C.POINTER(C.c_uint32)),
        ("mapot_len", C.c_uint64), ("maptt_len", C.c_uint64),
    ]


class _CD(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class OrbEq(C.Structure):
    """Mirror of ``orc_orbeq`` (the fields of reference ``_OrbitalEquation``
    the orbital step reads, mctdhb_types.h:136-149)."""
    _fields_ = [
        ("norb", C.c_int), ("grid_size", C.c_int), ("bounds", C.c_int),
        ("xi", C.c_double), ("xf", C.c_double), ("dx", C.c_double),
        ("tstep", C.c_double), ("g", C.c_double), ("d2coef", C.c_double),
        ("d1coef_re", C.c_double), ("d1coef_im", C.c_double),
        ("prop_dt_re", C.c_double), ("prop_dt_im", C.c_double),
        ("time_fac_re", C.c_double), ("time_fac_im", C.c_double),
        ("pot", C.POINTER(C.c_double)), ("rk_order", C.c_int),
    ]


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built; run oracle.build()")
        L = C.CDLL(_LIB_PATH)
        L.orc_space_dimension.argtypes = [C.c_uint16, C.c_uint16, C.POINTER(C.c_uint32)]
        L.orc_space_dimension.restype = C.c_int
        L.orc_subspaces_dim.argtypes = [C.c_uint16, C.c_uint16, _u32p]
        L.orc_set_config.argtypes = [C.c_uint32, C.c_uint16, C.c_uint16, _u16p]
        L.orc_hash_index.argtypes = [C.c_uint16, C.c_uint16, _u32p, _u16p]
        L.orc_hash_index.restype = C.c_uint32
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_space_create.argtypes = [C.c_uint16, C.c_uint16]
        L.orc_space_create.restype = C.POINTER(_Space)
        L.orc_space_destroy.argtypes = [C.POINTER(_Space)]
        sp = C.POINTER(_Space)
        L.orc_apply_hamiltonian.argtypes = [sp, _cp, _cp, _cp, _cp]
        L.orc_onebody_dm.argtypes = [sp, _cp, _cp]
        L.orc_twobody_dm.argtypes = [sp, _cp, _cp]
        L.orc_total_energy.argtypes = [C.c_int, _cp, _cp, _cp, _cp, _cp]
        L.orc_tridiag_eig.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_tridiag_eig.restype = C.c_int
        L.orc_lanczos.argtypes = [sp, _cp, _cp, C.c_int, _cp, _cp, _cp]
        L.orc_lanczos.restype = C.c_int
        L.orc_sil_step.argtypes = [sp, _cp, _cp, C.c_int, C.c_double, C.c_double, C.c_int,
                                   _cp, _cp]
        L.orc_sil_step.restype = C.c_int
        L.orc_lowest_state_lanczos.argtypes = [sp, _cp, _cp, C.c_int, _cp]
        L.orc_lowest_state_lanczos.restype = C.c_double
        L.orc_csimps.argtypes = [C.c_uint32, C.c_double, _cp]
        L.orc_csimps.restype = _CD
        L.orc_rsimps.argtypes = [C.c_uint32, C.c_double, _dp]
        L.orc_rsimps.restype = C.c_double
        L.orc_dxfd.argtypes = [C.c_uint32, C.c_double, _cp, _cp]
        L.orc_renormalize.argtypes = [C.c_uint32, C.c_double, C.c_double, _cp]
        L.orc_orthonormalize.argtypes = [C.c_uint32, C.c_double, C.c_uint32, _cp]
        oe = C.POINTER(OrbEq)
        L.orc_orbital_hob.argtypes = [oe, _cp, _cp]
        L.orc_orbital_hint.argtypes = [oe, _cp, _cp]
        L.orc_cn_tridiagonal.argtypes = [oe, _cp, _cp, _cp]
        L.orc_cn_rhs.argtypes = [oe, _cp, _cp]
        L.orc_solve_tridiag.argtypes = [C.c_uint32, _cp, _cp, _cp, _cp, _cp]
        L.orc_solve_cyclic_tridiag_sm.argtypes = [C.c_uint32, _cp, _cp, _cp, _cp, _cp]
        L.orc_advance_linear_cn.argtypes = [oe, _cp]
        L.orc_hder_fftexp.argtypes = [oe, _cp]
        L.orc_dft.argtypes = [C.c_int, C.c_int, C.c_double, _cp, _cp]
        L.orc_advance_linear_fft.argtypes = [oe, _cp]
        L.orc_nonlinear_rhs.argtypes = [oe, _cp, _cp, _cp, _cp, _cp, _cp]
        L.orc_herm_eig.argtypes = [C.c_int, _cp, _dp, _cp]
        L.orc_herm_eig.restype = C.c_int
        L.orc_herm_inverse.argtypes = [C.c_int, _cp, _cp]
        L.orc_herm_inverse.restype = C.c_int
        L.orc_regularization.argtypes = [C.c_int, C.c_double, _cp]
        L.orc_sync_density.argtypes = [sp, _cp, _cp, _cp, _cp, C.c_double]
        L.orc_rk_nonlinear.argtypes = [oe, _cp, _cp, _cp, _cp, _cp]
        L.orc_splitstep_orb.argtypes = [oe, C.c_int, C.c_int, _cp, _cp, _cp, _cp, _cp]
        L.orc_propagate_step.argtypes = [sp, oe, C.c_int, C.c_int, C.c_int, _cp, _cp, _cp,
                                         _cp, _cp, _cp, _cp, _cp]
        L.orc_orb_full_nonlinear.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _cp, _cp, _cp, _cp,
                                             _cp]
        L.orc_orb_full_nonlinear.restype = _CD
        L.orc_orb_interacting_part.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _cp, _cp, _cp]
        L.orc_orb_interacting_part.restype = _CD
        L.orc_linear_horb_fd.argtypes = [oe, _cp, _cp]
        L.orc_linear_horb_fft.argtypes = [oe, _cp, _cp]
        L.orc_expdvr_mat.argtypes = [oe, _cp]
        L.orc_linear_horb_dvr.argtypes = [oe, _cp, _cp, _cp]
        L.orc_overlap_matrix.argtypes = [C.c_int, C.c_int, C.c_double, _cp, _cp]
        L.orc_multiorb_projector.argtypes = [oe, C.c_int, _cp, _cp, _cp]
        L.orc_fullstep_rhs.argtypes = [oe, C.c_int, C.c_int, _cp, _cp, _cp, _cp]
        L.orc_fullstep_orb.argtypes = [oe, C.c_int, C.c_int, C.c_int, _cp, _cp, _cp]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's later calls."""
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def space_dimension(npar: int, norb: int) -> int:
    """Reference space.c:84-117; raises OverflowError where it would abort."""
    d = C.c_uint32()
    if lib().orc_space_dimension(npar, norb, C.byref(d)) != 0:
        raise OverflowError("space dimension overflows 32 bits")
    return int(d.value)


def subspaces_dim(npar: int, norb: int) -> np.ndarray:
    out = np.zeros((npar + 1) * (norb + 1), dtype=np.uint32)
    lib().orc_subspaces_dim(npar, norb, out)
    return out


def set_config(hash_i: int, npar: int, norb: int) -> np.ndarray:
    out = np.zeros(norb, dtype=np.uint16)
    lib().orc_set_config(hash_i, npar, norb, out)
    return out


def hash_index(npar: int, norb: int, conf) -> int:
    sd = subspaces_dim(npar, norb)
    return int(lib().orc_hash_index(npar, norb, sd, np.ascontiguousarray(conf, np.uint16)))


class Space:
    """Owns an ``orc_space`` (tables built by the restated reference code)."""

    def __init__(self, npar: int, norb: int):
        p = lib().orc_space_create(npar, norb)
        if not p:
            raise ValueError(f"invalid space N={npar} M={norb}")
        self._p = p
        s = p.contents
        self.npar, self.norb, self.dim = int(s.npar), int(s.norb), int(s.dim)
        m2 = self.norb * self.norb
        self.hash_table = np.ctypeslib.as_array(s.hash_table, (self.dim * self.norb,))
        self.subsdim = np.ctypeslib.as_array(s.subsdim, ((npar + 1) * (norb + 1),))
        self.map = np.ctypeslib.as_array(s.map, (self.dim * m2,))
        self.strideot = np.ctypeslib.as_array(s.strideot, (self.dim,))
        self.stridett = np.ctypeslib.as_array(s.stridett, (self.dim,))
        self.mapot = np.ctypeslib.as_array(s.mapot, (max(int(s.mapot_len), 1),))[: int(s.mapot_len)]
        self.maptt = np.ctypeslib.as_array(s.maptt, (max(int(s.maptt_len), 1),))[: int(s.maptt_len)]

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_space_destroy(self._p)
            self._p = None

    def apply_hamiltonian(self, coef, hob, hint):
        out = np.zeros(self.dim, np.complex128)
        lib().orc_apply_hamiltonian(self._p, _c(coef), _c(hob), _c(hint), out)
        return out

    def onebody_dm(self, coef):
        out = np.zeros(self.norb * self.norb, np.complex128)
        lib().orc_onebody_dm(self._p, _c(coef), out)
        return out.reshape(self.norb, self.norb)

    def twobody_dm(self, coef):
        out = np.zeros(self.norb ** 4, np.complex128)
        lib().orc_twobody_dm(self._p, _c(coef), out)
        return out

    def lanczos(self, coef, hob, hint, lm):
        lvec = np.zeros(lm * self.dim, np.complex128)
        lvec[: self.dim] = coef
        diag = np.zeros(lm, np.complex128)
        offd = np.zeros(lm, np.complex128)
        done = lib().orc_lanczos(self._p, _c(hob), _c(hint), lm, diag, offd, lvec)
        return done, diag, offd, lvec.reshape(lm, self.dim)

    def sil_step(self, coef, hob, hint, lm, prop_dt, imagtime):
        c = np.array(coef, np.complex128, copy=True)
        work = np.zeros(lm * self.dim, np.complex128)
        done = lib().orc_sil_step(self._p, _c(hob), _c(hint), lm, float(np.real(prop_dt)),
                                  float(np.imag(prop_dt)), int(bool(imagtime)), c, work)
        return done, c

    def lowest_state_lanczos(self, coef, hob, hint, niter):
        c = np.array(coef, np.complex128, copy=True)
        e = lib().orc_lowest_state_lanczos(self._p, _c(hob), _c(hint), niter, c)
        return e, c

    def sync_density(self, coef, reg_factor=1e-9):
        M = self.norb
        r1 = np.zeros(M * M, np.complex128)
        r2 = np.zeros(M ** 4, np.complex128)
        ri = np.zeros(M * M, np.complex128)
        lib().orc_sync_density(self._p, _c(coef), r1, r2, ri, reg_factor)
        return r1.reshape(M, M), r2, ri.reshape(M, M)


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128).ravel())


def total_energy(norb, rho1, rho2, hob, hint) -> complex:
    out = np.zeros(1, np.complex128)
    lib().orc_total_energy(norb, _c(rho1), _c(rho2), _c(hob), _c(hint), out)
    return complex(out[0])


def tridiag_eig(d, e):
    d = np.array(d, np.float64, copy=True)
    n = d.size
    ee = np.zeros(max(n, 1), np.float64)
    ee[: n - 1] = np.asarray(e, np.float64)[: n - 1]
    z = np.zeros(n * n, np.float64)
    rc = lib().orc_tridiag_eig(n, d, ee, z)
    if rc != 0:
        raise RuntimeError("tridiagonal eigensolver did not converge")
    return d, z.reshape(n, n)


def make_orbeq(norb, grid_size, xi, xf, tstep, g, pot, imagtime=True, bounds=1001,
               d2coef=-0.5, d1coef=0j, rk_order=4):
    """Fill an ``OrbEq`` as reference get_orbital_equation does
    (types_definition.c:30-108: dx = (xf - xi)/(grid_size - 1); imaginary
    time uses prop_dt = -i dt and time_fac = -i)."""
    pot = np.ascontiguousarray(pot, np.float64)
    eq = OrbEq()
    eq.norb, eq.grid_size, eq.bounds = norb, grid_size, bounds
    eq.xi, eq.xf = xi, xf
    eq.dx = (xf - xi) / (grid_size - 1)
    eq.tstep, eq.g, eq.d2coef = tstep, g, d2coef
    eq.d1coef_re, eq.d1coef_im = float(np.real(d1coef)), float(np.imag(d1coef))
    if imagtime:
        eq.prop_dt_re, eq.prop_dt_im = 0.0, -tstep
        eq.time_fac_re, eq.time_fac_im = 0.0, -1.0
    else:
        eq.prop_dt_re, eq.prop_dt_im = tstep, 0.0
        eq.time_fac_re, eq.time_fac_im = 1.0, 0.0
    eq.pot = pot.ctypes.data_as(C.POINTER(C.c_double))
    eq.rk_order = rk_order
    eq._pot_keep = pot  # keep the buffer alive with the struct
    return eq


def csimps(f, h) -> complex:
    r = lib().orc_csimps(len(f), h, _c(f))
    return complex(r.re, r.im)
