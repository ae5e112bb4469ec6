"""GPU parity at the benchmark size: N=100 bosons, M=5 orbitals
(NC = C(104, 4) = 4,598,126), against the FP64 oracle on the same seeded
inputs.

Inputs follow the benchmark workload: complex-normal coefficients normalised
to 1, a random Hermitian one-body matrix of scale 1 and a symmetrised
two-body tensor of scale 0.1 (tests/inputs.py).

What is checked, through the C ABI:
* every device table (hash table, subspace LUT, map, mapot, maptt, strides)
  bit-exact against the oracle's restatement of reference space.c:181-452;
* FP64 H.C from the tables and from on-the-fly rank deltas within 1e-10
  relative of the oracle's apply_hamiltonian (hamiltonian.c:7-406), and the
  two device forms bit-identical;
* FP64 rho1 / rho2 (density_matrices.c:9-443) within 1e-10 relative;
* two FP64 real-time SIL steps (m = 6, dt = 0.01;
  coefficients_integration.c:59-160) within 1e-10 relative;
* size-independent properties on the device results: Hermiticity
  <u, H v> = conj <v, H u>, the energy identity <C, H C> = total_energy(rho1,
  rho2) (integrator_monitor.c:56-88), tr rho1 = N, the partial trace
  sum_s rho2[k,s,s,l] = (N - 1) rho1[k,l], and the norm kept in real time;
* the mixed form (complex64 H.C and rho) against the same oracle, at the
  mixed bars (2e-5 H.C, 2e-6 rho).

The oracle's tables take about 10-20 s to build on the host and its SIL step
a few seconds; everything is built once per module.
"""
import math

import numpy as np
import pytest

import gpu_util as gu
import inputs

pytestmark = pytest.mark.gpu

NPAR, NORB, LM, DT = 100, 5, 6, 0.01
DIM = math.comb(NPAR + NORB - 1, NORB - 1)
SEED = 11


@pytest.fixture(scope="module")
def amd():
    import mctdhb_amd

    mctdhb_amd.lib()  # fails loudly if the HIP library is missing
    return mctdhb_amd


@pytest.fixture(scope="module")
def prob():
    return (inputs.coefficients(DIM, SEED), inputs.onebody(NORB, SEED), inputs.twobody(NORB, SEED, 0.1))


@pytest.fixture(scope="module")
def ospace(orc):
    print(f"oracle tables N={NPAR} M={NORB} dim={DIM}", flush=True)
    S = orc.Space(NPAR, NORB)
    assert S.dim == DIM
    return S


@pytest.fixture(scope="module")
def dspace(amd):
    print(f"device tables N={NPAR} M={NORB}", flush=True)
    D = amd.DeviceSpace(NPAR, NORB, tables=True)
    yield D
    D.close()


@pytest.fixture(scope="module")
def ref_hc(ospace, prob):
    c, h, g = prob
    return ospace.apply_hamiltonian(c, h, g)


def test_tables_bit_exact_full_size(ospace, dspace):
    print("config2 tables", flush=True)
    S, D = ospace, dspace
    assert D.dim == S.dim and D.mapot_len == S.mapot.size and D.maptt_len == S.maptt.size
    p = D.table_pointers()
    m2 = NORB * NORB
    assert np.array_equal(gu.d2h(p["hash_table"], DIM * NORB, np.uint16), S.hash_table)
    assert np.array_equal(gu.d2h(p["subspaces_dim"], (NPAR + 1) * (NORB + 1), np.uint32), S.subsdim)
    assert np.array_equal(gu.d2h(p["strideot"], DIM, np.uint32), S.strideot)
    assert np.array_equal(gu.d2h(p["stridett"], DIM, np.uint32), S.stridett)
    assert np.array_equal(gu.d2h(p["map"], DIM * m2, np.uint32), S.map)
    assert np.array_equal(gu.d2h(p["mapot"], S.mapot.size, np.uint32), S.mapot)
    assert np.array_equal(gu.d2h(p["maptt"], S.maptt.size, np.uint32), S.maptt)


def test_hc_fp64_full_size(amd, dspace, prob, ref_hc):
    print("config2 hc fp64", flush=True)
    c, h, g = prob
    dc, dh, dg = gu.dev(c), gu.dev(h), gu.dev(g)
    out_t, out_f = gu.empty(DIM), gu.empty(DIM)
    dspace.apply_hamiltonian(dc, dh, dg, out_t, prec=amd.FP64, index=amd.INDEX_TABLES)
    dspace.apply_hamiltonian(dc, dh, dg, out_f, prec=amd.FP64, index=amd.INDEX_FLY)
    got_t, got_f = gu.host(out_t), gu.host(out_f)
    assert gu.rel_err(got_t, ref_hc) < 1e-10
    assert np.array_equal(got_t, got_f)  # same arithmetic on the same targets


def test_hc_hermitian_and_energy_identity_full_size(amd, dspace, prob, orc):
    """<u, H v> = conj <v, H u> for two independent vectors, and
    <C, H C> = sum rho1_kl h_kl + 1/2 sum rho2 g (the reference's
    total_energy) with every factor computed on the device."""
    print("config2 hermiticity / energy identity", flush=True)
    c, h, g = prob
    u = inputs.coefficients(DIM, SEED + 1)
    dh, dg = gu.dev(h), gu.dev(g)
    hu, hv = gu.empty(DIM), gu.empty(DIM)
    dspace.apply_hamiltonian(gu.dev(u), dh, dg, hu, prec=amd.FP64, index=amd.INDEX_FLY)
    dspace.apply_hamiltonian(gu.dev(c), dh, dg, hv, prec=amd.FP64, index=amd.INDEX_FLY)
    hu, hv = gu.host(hu), gu.host(hv)
    a, b = np.vdot(u, hv), np.vdot(c, hu)
    assert abs(a - np.conj(b)) < 1e-12 * max(abs(a), 1.0)
    r1, r2 = gu.empty(NORB * NORB), gu.empty(NORB ** 4)
    dspace.density(gu.dev(c), r1, r2, prec=amd.FP64, index=amd.INDEX_FLY)
    r1, r2 = gu.host(r1), gu.host(r2)
    e_dm = orc.total_energy(NORB, r1, r2, h, g)
    e_hc = np.vdot(c, hv)
    assert abs(e_hc.imag) < 1e-12 * abs(e_hc)
    assert abs(e_dm - e_hc) < 1e-11 * abs(e_hc)


@pytest.mark.parametrize("index", [0, 1])
def test_density_fp64_full_size(amd, ospace, dspace, prob, index):
    print(f"config2 density fp64 index={index}", flush=True)
    c = prob[0]
    r1, r2 = gu.empty(NORB * NORB), gu.empty(NORB ** 4)
    dspace.density(gu.dev(c), r1, r2, prec=amd.FP64, index=index)
    g1 = gu.host(r1).reshape(NORB, NORB)
    g2 = gu.host(r2)
    ref1, ref2 = ospace.onebody_dm(c), ospace.twobody_dm(c)
    assert gu.rel_err(g1, ref1) < 1e-10
    assert gu.rel_err(g2, ref2) < 1e-10
    # trace identities (reference conventions: rho2 index k + sM + qM^2 + lM^3)
    assert abs(np.trace(g1) - NPAR) < 1e-9 * NPAR
    R2 = g2.reshape((NORB,) * 4, order="F")  # R2[k, s, q, l]
    part = np.einsum("kssl->kl", R2)
    assert np.max(np.abs(part - (NPAR - 1) * g1)) < 1e-9 * NPAR * NPAR


def test_sil_fp64_two_steps_full_size(amd, ospace, dspace, prob):
    print("config2 sil fp64 x2 (oracle on the host)", flush=True)
    c, h, g = prob
    ref = c
    for _ in range(2):
        _, ref = ospace.sil_step(ref, h, g, LM, DT, False)
    sil = amd.DeviceSil(dspace, LM, prec=amd.FP64, index=amd.INDEX_FLY)
    dc = gu.dev(c)
    dh, dg = gu.dev(h), gu.dev(g)
    for _ in range(2):
        sil.step(dc, dh, dg, DT, False)
    got = gu.host(dc)
    sil.close()
    assert gu.rel_err(got, ref) < 1e-10
    assert abs(np.linalg.norm(got) - 1.0) < 1e-12  # unitary in real time


def test_mixed_full_size(amd, dspace, prob, ref_hc, ospace):
    print("config2 mixed hc / density", flush=True)
    c, h, g = prob
    out = gu.empty(DIM, "complex64")
    dc32 = gu.dev(c, np.complex64)
    dspace.apply_hamiltonian(dc32, gu.dev(h), gu.dev(g), out, prec=amd.MIXED, index=amd.INDEX_FLY)
    assert gu.rel_err(gu.host(out), ref_hc) < 2e-5
    r1, r2 = gu.empty(NORB * NORB), gu.empty(NORB ** 4)
    dspace.density(dc32, r1, r2, prec=amd.MIXED)
    assert gu.rel_err(gu.host(r1).reshape(NORB, NORB), ospace.onebody_dm(c)) < 2e-6
    assert gu.rel_err(gu.host(r2), ospace.twobody_dm(c)) < 2e-6


@pytest.mark.parametrize("tiling", [0, 1, 2])
def test_mixed_tilings_full_size(amd, dspace, prob, ref_hc, tiling):
    """Each mixed H.C tiling (one configuration per lane, pairs, coefficient
    windows in LDS) at the benchmark size against the oracle; the windowed
    one at this size copies 19 far-group windows per block and sends the
    lanes whose targets leave them (blocks crossing a change of P_2 or P_3,
    and the rows next to the last configuration) through the gathering
    path, so both of its paths are exercised."""
    print(f"config2 mixed hc tiling {tiling}", flush=True)
    c, h, g = prob
    before = dspace.hc_tiling()
    try:
        dspace.hc_tiling(tiling)
        out = gu.empty(DIM, "complex64")
        dspace.apply_hamiltonian(gu.dev(c, np.complex64), gu.dev(h), gu.dev(g), out, prec=amd.MIXED,
                                 index=amd.INDEX_FLY)
        got = gu.host(out)
        assert gu.rel_err(got, ref_hc) < 2e-5
        # entry-wise too: a wrong window entry would show in a few rows only
        assert np.max(np.abs(got - ref_hc)) <= 2e-5 * np.max(np.abs(ref_hc))
    finally:
        dspace.hc_tiling(before if before >= 0 else -1)
