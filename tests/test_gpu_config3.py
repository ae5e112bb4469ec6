"""Mixed precision at its stated error bar, on the benchmark shapes.

From the same seeded coefficients, one-body matrix (scale 1) and two-body
tensor (scale 0.1), 50 real-time SIL steps (Lanczos dimension 8, full
reorthogonalisation, dt = 0.01; reference coefficients_integration.c:59-160
with multiconfig_lanczos.c:12-90) run once in the mixed form (complex64 H.C
with FP64 Lanczos scalars) and once in FP64.  The FP64 path is itself pinned
to the oracle at these shapes (one step, and the tables, below; the N=100,
M=5 shape in test_gpu_config2.py).

Bars, at the end of the 50 steps:
* energy per particle, E = <C, H C> / N: relative error <= 1e-6;
* natural occupations (eigenvalues of rho1 / N): relative error <= 1e-5;
* rho1: relative Frobenius error reported and held to <= 1e-5.
"""
import math

import numpy as np
import pytest

import gpu_util as gu
import inputs

pytestmark = pytest.mark.gpu

SEED = 7
STEPS, LM, DT = 50, 8, 0.01


@pytest.fixture(scope="module")
def amd():
    import mctdhb_amd

    mctdhb_amd.lib()
    return mctdhb_amd


def _problem(npar, norb):
    dim = math.comb(npar + norb - 1, norb - 1)
    return inputs.coefficients(dim, SEED), inputs.onebody(norb, SEED), inputs.twobody(norb, SEED, 0.1)


def _observables(amd, D, C, H, G, npar, norb):
    import torch

    hc = torch.empty_like(C)
    D.apply_hamiltonian(C, H, G, hc, amd.FP64, amd.INDEX_FLY)
    r1 = gu.empty(norb * norb)
    D.density(C, r1, None, amd.FP64, amd.INDEX_FLY)
    torch.cuda.synchronize()
    e = torch.vdot(C, hc).real.item() / npar
    rho1 = gu.host(r1).reshape(norb, norb)
    return e, rho1, np.sort(np.linalg.eigvalsh(rho1)) / npar


@pytest.mark.parametrize("npar,norb", [(10, 4), (200, 4), (30, 7), (100, 5)])
def test_mixed_within_error_bar(amd, npar, norb):
    print(f"config3 mixed vs fp64 N={npar} M={norb} ({STEPS} steps, m={LM})", flush=True)
    c, h, g = _problem(npar, norb)
    D = amd.DeviceSpace(npar, norb, tables=False)
    H, G = gu.dev(h), gu.dev(g)
    res = {}
    for name, prec in (("mixed", amd.MIXED), ("fp64", amd.FP64)):
        sil = amd.DeviceSil(D, LM, prec, amd.INDEX_FLY)
        C = gu.dev(c)
        for _ in range(STEPS):
            sil.step(C, H, G, DT, False)
        sil.close()
        res[name] = _observables(amd, D, C, H, G, npar, norb)
        assert abs(np.linalg.norm(gu.host(C)) - 1.0) < 1e-9
    (e32, r32, o32), (e64, r64, o64) = res["mixed"], res["fp64"]
    e_rel = abs(e32 - e64) / abs(e64)
    o_rel = float(np.max(np.abs(o32 - o64) / np.abs(o64)))
    r_rel = float(np.linalg.norm(r32 - r64) / np.linalg.norm(r64))
    print(f"  E rel {e_rel:.3e}  occupations rel {o_rel:.3e}  rho1 rel {r_rel:.3e}", flush=True)
    assert e_rel <= 1e-6
    assert o_rel <= 1e-5
    assert r_rel <= 1e-5


@pytest.mark.parametrize("npar,norb", [(200, 4), (30, 7)])
def test_fp64_step_and_tables_match_oracle(amd, orc, npar, norb):
    """The FP64 path the bar is measured against: device tables bit-exact
    and one FP64 SIL step (m = 8) within 1e-10 of the oracle."""
    print(f"config3 fp64 vs oracle N={npar} M={norb}", flush=True)
    c, h, g = _problem(npar, norb)
    S = orc.Space(npar, norb)
    D = amd.DeviceSpace(npar, norb, tables=True)
    p = D.table_pointers()
    assert np.array_equal(gu.d2h(p["hash_table"], S.dim * norb, np.uint16), S.hash_table)
    assert np.array_equal(gu.d2h(p["map"], S.dim * norb * norb, np.uint32), S.map)
    assert np.array_equal(gu.d2h(p["strideot"], S.dim, np.uint32), S.strideot)
    assert np.array_equal(gu.d2h(p["stridett"], S.dim, np.uint32), S.stridett)
    assert np.array_equal(gu.d2h(p["mapot"], S.mapot.size, np.uint32), S.mapot)
    assert np.array_equal(gu.d2h(p["maptt"], S.maptt.size, np.uint32), S.maptt)
    _, ref = S.sil_step(c, h, g, LM, DT, False)
    for index in (amd.INDEX_FLY, amd.INDEX_TABLES):
        sil = amd.DeviceSil(D, LM, amd.FP64, index)
        C = gu.dev(c)
        sil.step(C, gu.dev(h), gu.dev(g), DT, False)
        sil.close()
        assert gu.rel_err(gu.host(C), ref) < 1e-10
    D.close()
