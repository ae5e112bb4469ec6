"""Python front end of libmctdhb_hip.so, the MI355X MCTDHB configurational core.

Mirrors the reference library's C interface (reference include/mctdhb_types.h
and the headers under include/configurational, include/linalg,
include/integrator) over ctypes:

* ``Multiconf`` wraps the reference-named drop-in entry points
  (``get_multiconf_struct``, ``apply_hamiltonian``, ``set_onebody_dm``,
  ``set_twobody_dm``, ``lanczos``) on host numpy arrays;
* ``DeviceSpace`` / ``DeviceSil`` wrap the device-resident ``mctdhb_*``
  entry points on device pointers (torch tensors are plumbing only: they
  own device memory and streams).

There is no CPU fallback: loading fails loudly when the HIP library is not
built, and every call raises ``MctdhbError`` with the library's message on a
non-zero status.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libmctdhb_hip.so")

FP64, MIXED = 0, 1
INDEX_TABLES, INDEX_FLY = 0, 1
SPACE_TABLES = 1
IMAGTIME, REALTIME = 10, 11

_lib = None


class MctdhbError(RuntimeError):
    pass


def build(jobs: int = 8) -> str:
    """Compile every HIP source for gfx950 with the package Makefile."""
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", PKG_ROOT], check=True)
    return LIB_PATH


class _OperatorMappings(C.Structure):
    _fields_ = [("map", C.POINTER(C.c_uint32)), ("mapot", C.POINTER(C.c_uint32)),
                ("maptt", C.POINTER(C.c_uint32)), ("strideot", C.POINTER(C.c_uint32)),
                ("stridett", C.POINTER(C.c_uint32))]


class _MultiConfiguration(C.Structure):
    """reference _MultiConfiguration (mctdhb_types.h:162-169) + appended dev."""
    _fields_ = [("npar", C.c_uint16), ("norb", C.c_uint16), ("dim", C.c_uint32),
                ("hash_table", C.POINTER(C.c_uint16)), ("subspaces_dim", C.POINTER(C.c_uint32)),
                ("op_maps", C.POINTER(_OperatorMappings)), ("dev", C.c_void_p)]


_vp = C.c_void_p
_i = C.c_int
_cpp = C.POINTER(C.c_void_p)


def lib():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MctdhbError(f"{LIB_PATH} is not built; run mctdhb_amd.build() (hipcc, gfx950)")
    # torch ships its own libamdhip64 with the same soname (libamdhip64.so.7)
    # but finds it by another name and path.  Loaded first, it is the one HIP
    # runtime of the process: this library's dependency then resolves to it
    # by soname.  Loaded after this library, torch would bring up a second
    # runtime, which cannot open the GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    L.mctdhb_last_error.restype = C.c_char_p
    L.mctdhb_device_count.argtypes = [C.POINTER(_i)]
    L.mctdhb_space_create.argtypes = [C.c_uint16, C.c_uint16, _i, _cpp]
    L.mctdhb_space_destroy.argtypes = [_vp]
    L.mctdhb_space_info.argtypes = [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64), C.POINTER(_i)]
    L.mctdhb_space_tables.argtypes = [_vp] + [_cpp] * 7
    L.mctdhb_apply_hamiltonian_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp]
    L.mctdhb_apply_hamiltonian_noalloc_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp]
    L.mctdhb_hc_coef_bytes.argtypes = [_vp, _i]
    L.mctdhb_hc_coef_bytes.restype = C.c_size_t
    L.mctdhb_hc_coefficients_dev.argtypes = [_vp, _vp, _vp, _vp, _i, _vp]
    L.mctdhb_apply_hamiltonian_cf_dev.argtypes = [_vp, _vp, _vp, _vp, _i, _i, _vp]
    L.mctdhb_density_dev.argtypes = [_vp, _vp, _vp, _vp, _i, _i, _vp]
    L.mctdhb_hc_tiling.argtypes = [_vp, _i, C.POINTER(_i)]
    L.mctdhb_sil_create.argtypes = [_vp, _i, _i, _i, _cpp]
    L.mctdhb_sil_destroy.argtypes = [_vp]
    L.mctdhb_sil_device_bytes.argtypes = [_vp, _i, _i]
    L.mctdhb_sil_device_bytes.restype = C.c_size_t
    L.mctdhb_sil_step_dev.argtypes = [_vp, _vp, _vp, _vp, C.c_double, C.c_double, _i, _vp]
    L.mctdhb_sil_set_small.argtypes = [_vp, _i]
    L.mctdhb_lanczos_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
    L.mctdhb_sil_scalars.argtypes = [_vp, C.POINTER(_i), _vp, _vp, _vp]
    L.mctdhb_lowest_state_dev.argtypes = [_vp, _vp, _vp, _vp, C.POINTER(C.c_double), _vp]
    L.mctdhb_eig_residual_dev.argtypes = [_vp, _vp, _vp, _vp, C.c_double, C.POINTER(C.c_double), _vp]
    L.eig_residual.argtypes = [C.POINTER(_MultiConfiguration), _vp, C.POINTER(_vp), _vp, C.c_double]
    L.eig_residual.restype = C.c_double
    L.mctdhb_orbital_create.argtypes = [_i, _i, _i, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, _i, _vp, _cpp]
    L.mctdhb_orbital_destroy.argtypes = [_vp]
    L.mctdhb_orbital_set_potential.argtypes = [_vp, _vp]
    L.mctdhb_orbital_set_interaction.argtypes = [_vp, C.c_double]
    L.mctdhb_orbital_set_rk_order.argtypes = [_vp, _i]
    L.mctdhb_linear_fft_dev.argtypes = [_vp, _vp, _vp]
    L.mctdhb_linear_cn_dev.argtypes = [_vp, _vp, _vp]
    L.mctdhb_orbital_matrices_dev.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.mctdhb_nonlinear_rhs_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.mctdhb_orthonormalize_dev.argtypes = [_vp, _vp, _vp]
    L.mctdhb_splitstep_dev.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]
    L.mctdhb_fullstep_dev.argtypes = [_vp, _i, _i, _vp, _vp, _vp, _vp]
    L.mctdhb_fullstep_rhs_dev.argtypes = [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp]
    L.mctdhb_projector_dev.argtypes = [_vp, _i, _vp, _vp, _vp, _vp]
    L.mctdhb_orbital_set_integrator.argtypes = [_vp, _i, _i, _i]
    L.mctdhb_density_sync_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, C.c_double, _i, _vp]
    L.mctdhb_propagate_step_dev.argtypes = [_vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.mctdhb_step_graph_create.argtypes = [_vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _cpp]
    L.mctdhb_step_graph_launch.argtypes = [_vp, _vp]
    L.mctdhb_step_graph_launch_many.argtypes = [_cpp, _i, _cpp, _i, _i]
    L.mctdhb_step_graph_destroy.argtypes = [_vp]
    L.mctdhb_pack_eligible.argtypes = [_vp, _vp, _vp, _i, C.POINTER(_i)]
    L.mctdhb_pack_create.argtypes = [_i, _cpp, _cpp, _cpp, _i, _cpp, _cpp, _cpp, _cpp, _cpp, _cpp, _cpp, _cpp]
    L.mctdhb_pack_groups.argtypes = [_vp, C.POINTER(_i)]
    L.mctdhb_pack_step.argtypes = [_vp, _cpp, _i, _i]
    L.mctdhb_pack_destroy.argtypes = [_vp]
    # reference-named drop-in layer
    L.space_dimension.argtypes = [C.c_uint16, C.c_uint16]
    L.space_dimension.restype = C.c_uint32
    L.get_multiconf_struct.argtypes = [C.c_uint16, C.c_uint16]
    L.get_multiconf_struct.restype = C.POINTER(_MultiConfiguration)
    L.destroy_multiconf_struct.argtypes = [C.POINTER(_MultiConfiguration)]
    L.set_config.argtypes = [C.c_uint32, C.c_uint16, C.c_uint16, C.POINTER(C.c_uint16)]
    L.hash_index.argtypes = [C.c_uint16, C.c_uint16, C.POINTER(C.c_uint32), C.POINTER(C.c_uint16)]
    L.hash_index.restype = C.c_uint32
    L.apply_hamiltonian.argtypes = [C.POINTER(_MultiConfiguration), _vp, C.POINTER(_vp), _vp, _vp]
    L.set_onebody_dm.argtypes = [C.POINTER(_MultiConfiguration), _vp, C.POINTER(_vp)]
    L.set_twobody_dm.argtypes = [C.POINTER(_MultiConfiguration), _vp, _vp]
    L.lanczos.argtypes = [C.POINTER(_MultiConfiguration), C.POINTER(_vp), _vp, _i, _vp, _vp, C.POINTER(_vp)]
    L.lanczos.restype = _i
    L.lowest_state_lanczos.argtypes = [C.c_uint16, C.POINTER(_MultiConfiguration), C.POINTER(_vp), _vp, _vp]
    L.lowest_state_lanczos.restype = C.c_double
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        raise MctdhbError(f"libmctdhb_hip error {rc}: {lib().mctdhb_last_error().decode()}")


def device_count() -> int:
    n = _i()
    check(lib().mctdhb_device_count(C.byref(n)))
    return n.value


def _ptr(x) -> int:
    """Device address of a torch tensor (or an int address)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _stream(s):
    if s is None:
        return None
    return s.cuda_stream if hasattr(s, "cuda_stream") else int(s)


class DeviceSpace:
    """Configurational space (hash table, optional jump tables) on the GPU."""

    def __init__(self, npar: int, norb: int, tables: bool = True):
        h = _vp()
        check(lib().mctdhb_space_create(npar, norb, SPACE_TABLES if tables else 0, C.byref(h)))
        self._h = h
        self.npar, self.norb = npar, norb
        dim, ot, tt, ht = C.c_uint32(), C.c_uint64(), C.c_uint64(), _i()
        check(lib().mctdhb_space_info(h, C.byref(dim), C.byref(ot), C.byref(tt), C.byref(ht)))
        self.dim, self.mapot_len, self.maptt_len, self.has_tables = dim.value, ot.value, tt.value, bool(ht.value)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().mctdhb_space_destroy(self._h)
            self._h = None

    __del__ = close

    def table_pointers(self) -> dict:
        p = [_vp() for _ in range(7)]
        check(lib().mctdhb_space_tables(self._h, *[C.byref(x) for x in p]))
        names = ["hash_table", "subspaces_dim", "map", "mapot", "maptt", "strideot", "stridett"]
        return {n: x.value for n, x in zip(names, p)}

    def apply_hamiltonian(self, coef, hob, hint, out, prec=FP64, index=INDEX_FLY, stream=None):
        check(lib().mctdhb_apply_hamiltonian_dev(self._h, _ptr(coef), _ptr(hob), _ptr(hint), _ptr(out),
                                                 prec, index, _stream(stream)))

    def apply_hamiltonian_noalloc(self, coef, hob, hint, out, prec=FP64, index=INDEX_FLY, stream=None):
        """apply_hamiltonian with no device allocation (the library's fixed
        per-GPU scratch; calls that overlap on one GPU race)."""
        check(lib().mctdhb_apply_hamiltonian_noalloc_dev(self._h, _ptr(coef), _ptr(hob), _ptr(hint), _ptr(out),
                                                         prec, index, _stream(stream)))

    def coef_bytes(self, prec=FP64) -> int:
        return int(lib().mctdhb_hc_coef_bytes(self._h, prec))

    def hc_coefficients(self, hob, hint, cf, prec=FP64, stream=None):
        check(lib().mctdhb_hc_coefficients_dev(self._h, _ptr(hob), _ptr(hint), _ptr(cf), prec, _stream(stream)))

    def apply_hamiltonian_cf(self, coef, cf, out, prec=FP64, index=INDEX_FLY, stream=None):
        check(lib().mctdhb_apply_hamiltonian_cf_dev(self._h, _ptr(coef), _ptr(cf), _ptr(out), prec, index,
                                                    _stream(stream)))

    def density(self, coef, rho1, rho2=None, prec=FP64, index=INDEX_FLY, stream=None):
        check(lib().mctdhb_density_dev(self._h, _ptr(coef), _ptr(rho1), _ptr(rho2), prec, index,
                                       _stream(stream)))

    def hc_tiling(self, set: int = -2) -> int:
        """Tiling of the mixed H.C kernel (0 per lane, 1 configuration
        pairs, 2 coefficient windows in LDS); set -1 to time the possible
        ones again, 0/1/2 to force."""
        t = _i()
        check(lib().mctdhb_hc_tiling(self._h, set, C.byref(t)))
        return t.value

    def eig_residual(self, coef, hob, hint, e, stream=None) -> float:
        """max_i |e c_i - (H c)_i| / N (reference eig_residual,
        integrator_monitor.c:9-28); synchronises."""
        r = C.c_double()
        check(lib().mctdhb_eig_residual_dev(self._h, _ptr(coef), _ptr(hob), _ptr(hint), float(e), C.byref(r),
                                            _stream(stream)))
        return r.value


class DeviceSil:
    """Krylov workspace of the short-iterative-Lanczos step."""

    def __init__(self, space: DeviceSpace, lanczos_iter: int, prec=FP64, index=INDEX_FLY):
        h = _vp()
        check(lib().mctdhb_sil_create(space.handle, lanczos_iter, prec, index, C.byref(h)))
        self._h, self.space, self.lm, self.prec = h, space, lanczos_iter, prec

    def close(self):
        if getattr(self, "_h", None):
            lib().mctdhb_sil_destroy(self._h)
            self._h = None

    __del__ = close

    def set_small(self, on: bool):
        """One-workgroup steps for small FP64 spaces (default on)."""
        check(lib().mctdhb_sil_set_small(self._h, int(bool(on))))

    def step(self, coef, hob, hint, prop_dt: complex, imagtime: bool, stream=None):
        check(lib().mctdhb_sil_step_dev(self._h, _ptr(coef), _ptr(hob), _ptr(hint), float(np.real(prop_dt)),
                                        float(np.imag(prop_dt)), int(bool(imagtime)), _stream(stream)))

    def lanczos(self, coef, hob, hint, lvec, stream=None):
        check(lib().mctdhb_lanczos_dev(self._h, _ptr(coef), _ptr(hob), _ptr(hint), _ptr(lvec), _stream(stream)))

    def lowest_state(self, coef, hob, hint, stream=None):
        """coef <- lowest Ritz vector (reference lowest_state_lanczos,
        multiconfig_lanczos.c:93-172); returns the lowest Ritz value."""
        e = C.c_double()
        check(lib().mctdhb_lowest_state_dev(self._h, _ptr(coef), _ptr(hob), _ptr(hint), C.byref(e),
                                            _stream(stream)))
        return e.value

    def scalars(self, stream=None):
        done = _i()
        diag = np.zeros(2 * self.lm)
        offd = np.zeros(self.lm)
        check(lib().mctdhb_sil_scalars(self._h, C.byref(done), diag.ctypes.data, offd.ctypes.data,
                                       _stream(stream)))
        return done.value, diag.view(np.complex128), offd


class DeviceOrbital:
    """Orbital-equation workspace (reference _OrbitalEquation +
    _OrbitalWorkspace, mctdhb_types.h:136-149, :199-212) on the GPU."""

    def __init__(self, norb, grid_size, xi, xf, tstep, g, pot, imagtime=True, bounds=1001, d2coef=-0.5,
                 d1coef=0j, rk_order=4):
        pot = np.ascontiguousarray(pot, np.float64)
        h = _vp()
        check(lib().mctdhb_orbital_create(norb, grid_size, bounds, xi, xf, tstep, g, d2coef,
                                          float(np.real(d1coef)), float(np.imag(d1coef)), int(bool(imagtime)),
                                          pot.ctypes.data, C.byref(h)))
        self._h, self.norb, self.grid_size, self.imagtime = h, norb, grid_size, imagtime
        if rk_order != 4:
            check(lib().mctdhb_orbital_set_rk_order(self._h, rk_order))

    def close(self):
        if getattr(self, "_h", None):
            lib().mctdhb_orbital_destroy(self._h)
            self._h = None

    __del__ = close

    @property
    def handle(self):
        return self._h

    def set_interaction(self, g):
        check(lib().mctdhb_orbital_set_interaction(self._h, float(g)))

    def set_potential(self, pot):
        pot = np.ascontiguousarray(pot, np.float64)
        check(lib().mctdhb_orbital_set_potential(self._h, pot.ctypes.data))

    def linear_fft(self, orb, stream=None):
        check(lib().mctdhb_linear_fft_dev(self._h, _ptr(orb), _stream(stream)))

    def linear_cn(self, orb, stream=None):
        check(lib().mctdhb_linear_cn_dev(self._h, _ptr(orb), _stream(stream)))

    def orbital_matrices(self, orb, hob, hint, stream=None):
        check(lib().mctdhb_orbital_matrices_dev(self._h, _ptr(orb), _ptr(hob), _ptr(hint), _stream(stream)))

    def nonlinear_rhs(self, orb, hob, hint, rho2, rinv, der, stream=None):
        check(lib().mctdhb_nonlinear_rhs_dev(self._h, _ptr(orb), _ptr(hob), _ptr(hint), _ptr(rho2), _ptr(rinv),
                                             _ptr(der), _stream(stream)))

    def orthonormalize(self, orb, stream=None):
        check(lib().mctdhb_orthonormalize_dev(self._h, _ptr(orb), _stream(stream)))

    def fullstep(self, orb, rho2, rinv, use_fd=1, robust=False, stream=None):
        """use_fd: 0 FFT, 1 finite differences, 2 DVR."""
        check(lib().mctdhb_fullstep_dev(self._h, int(use_fd), int(bool(robust)), _ptr(orb), _ptr(rho2),
                                        _ptr(rinv), _stream(stream)))

    def fullstep_rhs(self, orb, rho2, rinv, der, use_fd=1, robust=False, stream=None):
        check(lib().mctdhb_fullstep_rhs_dev(self._h, int(use_fd), int(bool(robust)), _ptr(orb), _ptr(rho2),
                                            _ptr(rinv), _ptr(der), _stream(stream)))

    def projector(self, orb, hact, proj, robust=False, stream=None):
        check(lib().mctdhb_projector_dev(self._h, int(bool(robust)), _ptr(orb), _ptr(hact), _ptr(proj),
                                         _stream(stream)))

    def set_integrator(self, fullstep: bool, robust: bool = False, dvr: bool = False):
        """Orbital integrator of propagate_step / StepGraph: FULLSTEP_RUNGEKUTTA
        (30) or SPLITSTEP (31); dvr selects the DVR linear part of the full step."""
        check(lib().mctdhb_orbital_set_integrator(self._h, 30 if fullstep else 31, 300 if dvr else 0,
                                                  int(bool(robust))))

    def splitstep(self, orb, hob, hint, rho2, rinv, use_fft=True, stream=None):
        check(lib().mctdhb_splitstep_dev(self._h, int(bool(use_fft)), _ptr(orb), _ptr(hob), _ptr(hint),
                                         _ptr(rho2), _ptr(rinv), _stream(stream)))


def density_sync(space: DeviceSpace, coef, rho1, rho2, rinv, eigvals=None, reg_factor=1e-9, prec=FP64,
                 stream=None):
    check(lib().mctdhb_density_sync_dev(space.handle, _ptr(coef), _ptr(rho1), _ptr(rho2), _ptr(rinv),
                                        _ptr(eigvals), reg_factor, prec, _stream(stream)))


def propagate_step(space: DeviceSpace, sil: DeviceSil, orbital: DeviceOrbital, orb, coef, hob, hint, rho1, rho2,
                   rinv, use_fft=True, stream=None):
    check(lib().mctdhb_propagate_step_dev(space.handle, sil._h, orbital.handle, int(bool(use_fft)), _ptr(orb),
                                          _ptr(coef), _ptr(hob), _ptr(hint), _ptr(rho1), _ptr(rho2), _ptr(rinv),
                                          _stream(stream)))


class StepGraph:
    """propagate_step recorded once as a HIP graph; launch() replays it on
    the same buffers."""

    def __init__(self, space: DeviceSpace, sil: DeviceSil, orbital: DeviceOrbital, orb, coef, hob, hint, rho1,
                 rho2, rinv, use_fft=True):
        h = _vp()
        check(lib().mctdhb_step_graph_create(space.handle, sil._h, orbital.handle, int(bool(use_fft)), _ptr(orb),
                                             _ptr(coef), _ptr(hob), _ptr(hint), _ptr(rho1), _ptr(rho2),
                                             _ptr(rinv), C.byref(h)))
        self._h = h
        self._keep = (space, sil, orbital)

    def launch(self, stream=None):
        check(lib().mctdhb_step_graph_launch(self._h, _stream(stream)))

    def close(self):
        if getattr(self, "_h", None):
            lib().mctdhb_step_graph_destroy(self._h)
            self._h = None

    __del__ = close


def launch_graphs(graphs, streams, rounds: int = 1):
    """`rounds` replays of every StepGraph in `graphs`, graph i on
    streams[i % len(streams)], in one library call."""
    n = len(graphs)
    if n == 0 or rounds <= 0:
        return
    ga = (_vp * n)(*[g._h for g in graphs])
    sa = (_vp * len(streams))(*[_stream(s) for s in streams])
    check(lib().mctdhb_step_graph_launch_many(ga, n, sa, len(streams), int(rounds)))


def pack_eligible(space: DeviceSpace, sil: DeviceSil, orbital: DeviceOrbital, use_fft=True) -> bool:
    """Whether this solve's step can run in a Pack (FFT linear step of a
    power-of-two length, split step, 2 <= M <= 7, the one-workgroup FP64 SIL
    step)."""
    ok = _i(0)
    check(lib().mctdhb_pack_eligible(space.handle, sil._h, orbital.handle, int(bool(use_fft)), C.byref(ok)))
    return bool(ok.value)


class Pack:
    """The steps of many small solves as one kernel launch per stage
    (mctdhb_pack_*).  `solves` holds, per solve, (space, sil, orbital, orb,
    coef, hob, hint, rho1, rho2, rinv) as for StepGraph; each solve's
    results are those of its own step graph, bit for bit."""

    def __init__(self, solves, use_fft=True):
        n = len(solves)
        if n == 0:
            raise ValueError("a pack needs at least one solve")
        cols = list(zip(*solves))
        sp = (_vp * n)(*[x.handle for x in cols[0]])
        sl = (_vp * n)(*[x._h for x in cols[1]])
        ob = (_vp * n)(*[x.handle for x in cols[2]])
        bufs = [(_vp * n)(*[_ptr(t) for t in cols[k]]) for k in range(3, 10)]
        h = _vp()
        check(lib().mctdhb_pack_create(n, sp, sl, ob, int(bool(use_fft)), *bufs, C.byref(h)))
        self._h = h
        self._keep = solves

    @property
    def groups(self) -> int:
        g = _i(0)
        check(lib().mctdhb_pack_groups(self._h, C.byref(g)))
        return g.value

    def step(self, streams, rounds: int = 1):
        """`rounds` steps of every solve, group g on streams[g % len(streams)]."""
        if rounds <= 0:
            return
        sa = (_vp * len(streams))(*[_stream(s) for s in streams])
        check(lib().mctdhb_pack_step(self._h, sa, len(streams), int(rounds)))

    def close(self):
        if getattr(self, "_h", None):
            lib().mctdhb_pack_destroy(self._h)
            self._h = None
            self._keep = None

    __del__ = close


class Multiconf:
    """The reference-named drop-in layer on host numpy arrays
    (get_multiconf_struct / apply_hamiltonian / set_onebody_dm /
    set_twobody_dm / lanczos, reference hamiltonian.h:7-13,
    density_matrices.h:7-12, multiconfig_lanczos.h:31-39)."""

    def __init__(self, npar: int, norb: int):
        self._p = lib().get_multiconf_struct(npar, norb)
        s = self._p.contents
        self.npar, self.norb, self.dim = int(s.npar), int(s.norb), int(s.dim)
        m2 = self.norb * self.norb
        om = s.op_maps.contents
        self.hash_table = np.ctypeslib.as_array(s.hash_table, (self.dim * self.norb,))
        self.map = np.ctypeslib.as_array(om.map, (self.dim * m2,))
        self.strideot = np.ctypeslib.as_array(om.strideot, (self.dim,))
        self.stridett = np.ctypeslib.as_array(om.stridett, (self.dim,))

    def close(self):
        if getattr(self, "_p", None):
            lib().destroy_multiconf_struct(self._p)
            self._p = None

    __del__ = close

    @staticmethod
    def _rows(mat: np.ndarray):
        rows = (_vp * mat.shape[0])(*[mat[r].ctypes.data for r in range(mat.shape[0])])
        return rows

    def apply_hamiltonian(self, coef, hob, hint):
        coef = np.ascontiguousarray(coef, np.complex128)
        hob = np.ascontiguousarray(hob, np.complex128).reshape(self.norb, self.norb)
        hint = np.ascontiguousarray(hint, np.complex128)
        out = np.zeros(self.dim, np.complex128)
        lib().apply_hamiltonian(self._p, coef.ctypes.data, self._rows(hob), hint.ctypes.data, out.ctypes.data)
        return out

    def onebody_dm(self, coef):
        coef = np.ascontiguousarray(coef, np.complex128)
        rho = np.zeros((self.norb, self.norb), np.complex128)
        lib().set_onebody_dm(self._p, coef.ctypes.data, self._rows(rho))
        return rho

    def twobody_dm(self, coef):
        coef = np.ascontiguousarray(coef, np.complex128)
        rho = np.zeros(self.norb ** 4, np.complex128)
        lib().set_twobody_dm(self._p, coef.ctypes.data, rho.ctypes.data)
        return rho

    def lanczos(self, coef, hob, hint, lm):
        hob = np.ascontiguousarray(hob, np.complex128).reshape(self.norb, self.norb)
        hint = np.ascontiguousarray(hint, np.complex128)
        lvec = np.zeros((lm, self.dim), np.complex128)
        lvec[0] = coef
        diag = np.zeros(lm, np.complex128)
        offd = np.zeros(lm, np.complex128)
        done = lib().lanczos(self._p, self._rows(hob), hint.ctypes.data, lm, diag.ctypes.data,
                             offd.ctypes.data, self._rows(lvec))
        return done, diag, offd, lvec

    def lowest_state_lanczos(self, coef, hob, hint, niter):
        """Returns (lowest Ritz value, Ritz vector) like reference
        lowest_state_lanczos (multiconfig_lanczos.c:93-172)."""
        c = np.array(coef, np.complex128, copy=True)
        hob = np.ascontiguousarray(hob, np.complex128).reshape(self.norb, self.norb)
        hint = np.ascontiguousarray(hint, np.complex128)
        e = lib().lowest_state_lanczos(niter, self._p, self._rows(hob), hint.ctypes.data, c.ctypes.data)
        return e, c

    def eig_residual(self, coef, hob, hint, e) -> float:
        coef = np.ascontiguousarray(coef, np.complex128)
        hob = np.ascontiguousarray(hob, np.complex128).reshape(self.norb, self.norb)
        hint = np.ascontiguousarray(hint, np.complex128)
        return lib().eig_residual(self._p, coef.ctypes.data, self._rows(hob), hint.ctypes.data, float(e))
