"""GPU parity of the configurational core against the FP64 oracle.

Every test calls the HIP library through its C ABI (libmctdhb_hip.so via
ctypes); inputs are seeded (tests/inputs.py) and identical on both sides.
Bars: tables bit-exact; FP64 H.C / density matrices / Lanczos within 1e-10
relative (max-abs error over max-abs value); mixed precision within 2e-5.
"""
import math

import numpy as np
import pytest

import gpu_util as gu
import inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import mctdhb_amd

    mctdhb_amd.lib()  # fails loudly if the HIP library is missing
    return mctdhb_amd


def _problem(npar, norb, seed=11, gscale=0.3):
    dim = math.comb(npar + norb - 1, norb - 1)
    return inputs.coefficients(dim, seed), inputs.onebody(norb, seed), inputs.twobody(norb, seed, gscale)


TABLE_SHAPES = [(1, 1), (4, 1), (4, 3), (6, 4), (5, 5), (10, 4), (3, 7), (20, 5), (12, 6), (2, 9), (5, 11)]


@pytest.mark.parametrize("npar,norb", TABLE_SHAPES)
def test_device_tables_bit_exact(amd, orc, npar, norb):
    print(f"tables N={npar} M={norb}", flush=True)
    S = orc.Space(npar, norb)
    D = amd.DeviceSpace(npar, norb, tables=True)
    assert D.dim == S.dim
    assert D.mapot_len == S.mapot.size and D.maptt_len == S.maptt.size
    p = D.table_pointers()
    dim, m2 = S.dim, norb * norb
    assert np.array_equal(gu.d2h(p["hash_table"], dim * norb, np.uint16), S.hash_table)
    assert np.array_equal(gu.d2h(p["subspaces_dim"], (npar + 1) * (norb + 1), np.uint32), S.subsdim)
    assert np.array_equal(gu.d2h(p["map"], dim * m2, np.uint32), S.map)
    assert np.array_equal(gu.d2h(p["strideot"], dim, np.uint32), S.strideot)
    assert np.array_equal(gu.d2h(p["stridett"], dim, np.uint32), S.stridett)
    assert np.array_equal(gu.d2h(p["mapot"], S.mapot.size, np.uint32), S.mapot)
    assert np.array_equal(gu.d2h(p["maptt"], S.maptt.size, np.uint32), S.maptt)


HC_SHAPES = [(5, 1), (9, 2), (6, 3), (8, 4), (10, 5), (6, 6), (5, 7), (3, 9), (3, 11), (30, 3), (40, 4)]


@pytest.mark.parametrize("npar,norb", HC_SHAPES)
def test_hc_fp64_matches_oracle(amd, orc, npar, norb):
    print(f"hc fp64 N={npar} M={norb}", flush=True)
    S = orc.Space(npar, norb)
    c, h, g = _problem(npar, norb)
    ref = S.apply_hamiltonian(c, h, g)
    D = amd.DeviceSpace(npar, norb, tables=True)
    dc, dh, dg = gu.dev(c), gu.dev(h), gu.dev(g)
    out_t = gu.empty(S.dim)
    D.apply_hamiltonian(dc, dh, dg, out_t, prec=amd.FP64, index=amd.INDEX_TABLES)
    got_t = gu.host(out_t)
    assert gu.rel_err(got_t, ref) < 1e-12
    if 2 <= norb <= 7:
        out_f = gu.empty(S.dim)
        D.apply_hamiltonian(dc, dh, dg, out_f, prec=amd.FP64, index=amd.INDEX_FLY)
        got_f = gu.host(out_f)
        # same arithmetic on the same jump targets: bit-identical
        assert np.array_equal(got_f, got_t)


# small spaces put many configurations next to the first and last ranks and
# at the ends of odd-length runs (the single slots of the paired kernel)
@pytest.mark.parametrize("npar,norb", [(9, 2), (10, 4), (10, 5), (5, 7), (40, 4), (1, 2), (2, 2), (1, 3),
                                       (2, 3), (3, 4), (2, 7), (1, 5), (7, 6), (60, 3)])
@pytest.mark.parametrize("tiling", [0, 1, 2])
def test_hc_mixed_matches_oracle(amd, orc, npar, norb, tiling):
    if tiling == 2 and norb > 5:
        pytest.skip("the windowed tiling covers 2 <= M <= 5")
    print(f"hc mixed N={npar} M={norb} tiling={tiling}", flush=True)
    S = orc.Space(npar, norb)
    c, h, g = _problem(npar, norb)
    ref = S.apply_hamiltonian(c, h, g)
    D = amd.DeviceSpace(npar, norb, tables=False)
    assert D.hc_tiling(tiling) == tiling
    out = gu.empty(S.dim, "complex64")
    D.apply_hamiltonian(gu.dev(c, np.complex64), gu.dev(h), gu.dev(g), out, prec=amd.MIXED,
                        index=amd.INDEX_FLY)
    assert gu.rel_err(gu.host(out), ref) < 2e-5


@pytest.mark.parametrize("npar,norb", [(5, 1), (9, 2), (6, 3), (8, 4), (10, 5), (6, 6), (5, 7), (3, 9)])
@pytest.mark.parametrize("index", [0, 1])
def test_density_fp64_matches_oracle(amd, orc, npar, norb, index):
    if index == 1 and not 2 <= norb <= 7:
        pytest.skip("on-the-fly indexing covers 2 <= M <= 7")
    print(f"density fp64 N={npar} M={norb} index={index}", flush=True)
    S = orc.Space(npar, norb)
    c, _, _ = _problem(npar, norb, 13)
    D = amd.DeviceSpace(npar, norb, tables=True)
    r1 = gu.empty(norb * norb)
    r2 = gu.empty(norb ** 4)
    D.density(gu.dev(c), r1, r2, prec=amd.FP64, index=index)
    assert gu.rel_err(gu.host(r1).reshape(norb, norb), S.onebody_dm(c)) < 1e-12
    assert gu.rel_err(gu.host(r2), S.twobody_dm(c)) < 1e-12


@pytest.mark.parametrize("npar,norb", [(9, 2), (10, 4), (10, 5), (5, 7)])
def test_density_mixed_matches_oracle(amd, orc, npar, norb):
    print(f"density mixed N={npar} M={norb}", flush=True)
    S = orc.Space(npar, norb)
    c, _, _ = _problem(npar, norb, 13)
    D = amd.DeviceSpace(npar, norb, tables=False)
    r1 = gu.empty(norb * norb)
    r2 = gu.empty(norb ** 4)
    D.density(gu.dev(c, np.complex64), r1, r2, prec=amd.MIXED)
    assert gu.rel_err(gu.host(r1).reshape(norb, norb), S.onebody_dm(c)) < 2e-6
    assert gu.rel_err(gu.host(r2), S.twobody_dm(c)) < 2e-6


@pytest.mark.parametrize("npar,norb,lm", [(6, 4, 6), (10, 5, 8), (3, 9, 5), (20, 3, 15)])
def test_lanczos_matches_oracle(amd, orc, npar, norb, lm):
    print(f"lanczos N={npar} M={norb} lm={lm}", flush=True)
    S = orc.Space(npar, norb)
    c, h, g = _problem(npar, norb, 23)
    done, diag, offd, V = S.lanczos(c, h, g, lm)
    D = amd.DeviceSpace(npar, norb, tables=True)
    sil = amd.DeviceSil(D, lm, prec=amd.FP64, index=amd.INDEX_FLY)
    lvec = gu.empty(lm * S.dim)
    sil.lanczos(gu.dev(c), gu.dev(h), gu.dev(g), lvec)
    gdone, gdiag, goffd = sil.scalars()
    assert gdone == done
    assert gu.rel_err(gdiag[:done], diag[:done]) < 1e-10
    assert gu.rel_err(goffd[: done - 1], offd[: done - 1].real) < 1e-10
    gV = gu.host(lvec).reshape(lm, S.dim)
    assert gu.rel_err(gV[:done], V[:done]) < 1e-9


@pytest.mark.parametrize("npar,norb,lm,imag", [(6, 4, 6, False), (10, 5, 8, True), (3, 9, 5, False),
                                               (12, 3, 5, True), (10, 4, 20, False), (10, 4, 18, True)])
def test_sil_step_matches_oracle(amd, orc, npar, norb, lm, imag):
    print(f"sil N={npar} M={norb} lm={lm} imag={imag}", flush=True)
    S = orc.Space(npar, norb)
    c, h, g = _problem(npar, norb, 29)
    dt = -0.01j if imag else 0.01
    _, ref = S.sil_step(c, h, g, lm, dt, imag)
    D = amd.DeviceSpace(npar, norb, tables=True)
    sil = amd.DeviceSil(D, lm, prec=amd.FP64, index=amd.INDEX_FLY)
    dc = gu.dev(c)
    sil.step(dc, gu.dev(h), gu.dev(g), dt, imag)
    assert gu.rel_err(gu.host(dc), ref) < 1e-10


@pytest.mark.parametrize("npar,norb,lm,index", [(10, 4, 5, 1), (10, 4, 5, 0), (2, 2, 5, 1), (30, 2, 15, 1),
                                                (6, 7, 8, 1), (12, 5, 6, 0), (3, 3, 12, 1),
                                                # staged in LDS above 64 KiB: NC 4845, 3003, 5456
                                                (16, 5, 5, 1), (8, 7, 5, 1), (30, 4, 6, 0)])
@pytest.mark.parametrize("imag", [False, True])
def test_one_workgroup_sil_step_matches_multikernel_and_oracle(amd, orc, npar, norb, lm, index, imag):
    """Small FP64 spaces take the one-workgroup step (mctdhb_sil_set_small);
    several steps agree with the multi-kernel step and with the oracle,
    including Krylov breakdown (N=2, M=2: three configurations, 5 vectors)
    and the table-driven indices."""
    print(f"small sil N={npar} M={norb} lm={lm} index={index} imag={imag}", flush=True)
    S = orc.Space(npar, norb)
    c, h, g = _problem(npar, norb, 41)
    dt = -0.01j if imag else 0.01
    ref = c
    for _ in range(3):
        _, ref = S.sil_step(ref, h, g, lm, dt, imag)
    D = amd.DeviceSpace(npar, norb, tables=True)
    small = amd.DeviceSil(D, lm, prec=amd.FP64, index=index)
    multi = amd.DeviceSil(D, lm, prec=amd.FP64, index=index)
    multi.set_small(False)
    dh, dg = gu.dev(h), gu.dev(g)
    a, b = gu.dev(c), gu.dev(c)
    for _ in range(3):
        small.step(a, dh, dg, dt, imag)
        multi.step(b, dh, dg, dt, imag)
    assert small.scalars()[0] == multi.scalars()[0]  # Krylov vectors used
    assert gu.rel_err(gu.host(a), gu.host(b)) < 1e-12
    assert gu.rel_err(gu.host(a), ref) < 1e-10


@pytest.mark.parametrize("npar,norb,small", [(10, 4, True), (10, 4, False), (40, 5, False)])
def test_imaginary_time_sil_invariant_under_energy_shift(amd, orc, npar, norb, small):
    """In imaginary time the renormalised SIL step is unchanged by H -> H + c
    (h -> h + c/N I adds c to every diagonal element).  With c dt = 900 the
    plain exponentials exp(-E dt) underflow for every Krylov component and
    the oracle's step (the reference's formula) gives no finite result; the
    device takes them relative to the dominant one and must still agree with
    the unshifted step, which agrees with the oracle."""
    print(f"sil shift N={npar} M={norb} small={small}", flush=True)
    lm, dt = 6, -0.005j
    c, h, g = _problem(npar, norb, 43)
    S = orc.Space(npar, norb)
    _, ref = S.sil_step(c, h, g, lm, dt, True)
    shift = 900.0 / 0.005  # total energy shift
    hs = h + np.eye(norb) * (shift / npar)
    _, bad = S.sil_step(c, hs, g, lm, dt, True)
    assert not np.all(np.isfinite(bad))  # the reference formula breaks down here
    D = amd.DeviceSpace(npar, norb, tables=False)
    out = {}
    for name, hh in (("plain", h), ("shifted", hs)):
        sil = amd.DeviceSil(D, lm, prec=amd.FP64, index=amd.INDEX_FLY)
        sil.set_small(small)
        a = gu.dev(c)
        sil.step(a, gu.dev(np.ascontiguousarray(hh)), gu.dev(g), dt, True)
        out[name] = gu.host(a)
        sil.close()
    assert gu.rel_err(out["plain"], ref) < 1e-10
    assert np.all(np.isfinite(out["shifted"]))
    assert gu.rel_err(out["shifted"], out["plain"]) < 1e-8


@pytest.mark.parametrize("npar,norb,lm,prec", [(6, 4, 8, 0), (10, 5, 12, 0), (10, 5, 12, 1), (3, 9, 6, 0),
                                               (12, 5, 40, 0), (12, 5, 24, 1)])
def test_lowest_state_matches_oracle(amd, orc, npar, norb, lm, prec):
    print(f"lowest N={npar} M={norb} lm={lm} prec={prec}", flush=True)
    S = orc.Space(npar, norb)
    c, h, g = _problem(npar, norb, 37)
    e_ref, ref = S.lowest_state_lanczos(c, h, g, lm)
    D = amd.DeviceSpace(npar, norb, tables=norb > 7)  # M > 7 runs the table-driven generic H.C
    sil = amd.DeviceSil(D, lm, prec=prec, index=amd.INDEX_FLY)
    dc = gu.dev(c)
    e = sil.lowest_state(dc, gu.dev(h), gu.dev(g))
    tol = 1e-10 if prec == amd.FP64 else 1e-5
    assert abs(e - e_ref) <= tol * abs(e_ref)  # mixed: complex64 Krylov vectors
    assert gu.rel_err(gu.host(dc), ref) < tol
    # the Ritz value is the Rayleigh quotient of the (unit) Ritz vector
    v = gu.host(dc)
    hv = S.apply_hamiltonian(v, h, g)
    assert abs(np.vdot(v, hv).real / np.vdot(v, v).real - e_ref) < 100 * tol * abs(e_ref)


@pytest.mark.parametrize("npar,norb", [(8, 4), (20, 5), (3, 9)])
def test_eig_residual_matches_definition(amd, orc, npar, norb):
    print(f"eig_residual N={npar} M={norb}", flush=True)
    S = orc.Space(npar, norb)
    c, h, g = _problem(npar, norb, 41)
    hc = S.apply_hamiltonian(c, h, g)
    e = np.vdot(c, hc).real
    ref = np.max(np.abs(e * c - hc)) / npar
    D = amd.DeviceSpace(npar, norb, tables=True)
    r = D.eig_residual(gu.dev(c), gu.dev(h), gu.dev(g), e)
    assert abs(r - ref) <= 1e-12 * max(ref, 1.0)
    mc = amd.Multiconf(npar, norb)
    assert abs(mc.eig_residual(c, h, g, e) - ref) <= 1e-12 * max(ref, 1.0)
    if S.dim <= 400:  # an eigenvector has a vanishing residual
        w, v = np.linalg.eigh(np.array([S.apply_hamiltonian(np.eye(S.dim)[k], h, g) for k in range(S.dim)]).T)
        r0 = D.eig_residual(gu.dev(np.ascontiguousarray(v[:, 0])), gu.dev(h), gu.dev(g), w[0])
        assert r0 < 1e-10


@pytest.mark.parametrize("npar,norb,lm,imag", [(10, 5, 8, False), (9, 2, 5, True), (2, 3, 3, False),
                                               (6, 7, 6, True), (30, 4, 18, False)])
@pytest.mark.parametrize("tiling", [0, 1, 2])
def test_sil_step_mixed_matches_oracle(amd, orc, npar, norb, lm, imag, tiling):
    if tiling == 2 and norb > 5:
        pytest.skip("the windowed tiling covers 2 <= M <= 5")
    print(f"sil mixed N={npar} M={norb} lm={lm} imag={imag} tiling={tiling}", flush=True)
    S = orc.Space(npar, norb)
    c, h, g = _problem(npar, norb, 43)
    dt = -0.01j if imag else 0.01
    _, ref = S.sil_step(c, h, g, lm, dt, imag)
    D = amd.DeviceSpace(npar, norb, tables=False)
    D.hc_tiling(tiling)
    sil = amd.DeviceSil(D, lm, prec=amd.MIXED, index=amd.INDEX_FLY)
    dc = gu.dev(c)
    sil.step(dc, gu.dev(h), gu.dev(g), dt, imag)
    # complex64 Krylov vectors and H.C, FP64 scalars and recombination
    assert gu.rel_err(gu.host(dc), ref) < 2e-5


def test_lanczos_breakdown_on_device(amd, orc):
    # H couples orbitals 0 and 1 only; e_0 spans a 2-dimensional Krylov space
    npar, norb = 1, 3
    h = np.array([[0, 0.8, 0], [0.8, 0, 0], [0, 0, 1.0]], complex)
    g = np.zeros(norb ** 4, complex)
    c = np.array([1, 0, 0], complex)
    S = orc.Space(npar, norb)
    _, ref = S.sil_step(c, h, g, 5, 0.3, False)
    D = amd.DeviceSpace(npar, norb, tables=True)
    sil = amd.DeviceSil(D, 5, prec=amd.FP64)
    dc = gu.dev(c)
    sil.step(dc, gu.dev(h), gu.dev(g), 0.3, False)
    done, _, _ = sil.scalars()
    assert done == 2
    assert gu.rel_err(gu.host(dc), ref) < 1e-13


def test_dropin_reference_entry_points(amd, orc):
    npar, norb = 8, 4
    S = orc.Space(npar, norb)
    c, h, g = _problem(npar, norb, 31)
    mc = amd.Multiconf(npar, norb)
    assert mc.dim == S.dim
    assert np.array_equal(mc.hash_table, S.hash_table)
    assert np.array_equal(mc.map, S.map)
    assert gu.rel_err(mc.apply_hamiltonian(c, h, g), S.apply_hamiltonian(c, h, g)) < 1e-12
    assert gu.rel_err(mc.onebody_dm(c), S.onebody_dm(c)) < 1e-12
    assert gu.rel_err(mc.twobody_dm(c), S.twobody_dm(c)) < 1e-12
    e, v = mc.lowest_state_lanczos(c, h, g, 7)
    re, rv = S.lowest_state_lanczos(c, h, g, 7)
    assert abs(e - re) < 1e-10 * abs(re)
    assert gu.rel_err(v, rv) < 1e-10
    done, diag, offd, V = mc.lanczos(c, h, g, 6)
    rdone, rdiag, roffd, rV = S.lanczos(c, h, g, 6)
    assert done == rdone
    assert gu.rel_err(diag, rdiag) < 1e-10
    assert gu.rel_err(offd[:5], roffd[:5]) < 1e-10
    assert gu.rel_err(V, rV) < 1e-9
    mc.close()
