"""The no-allocation H.C entry point (mctdhb_apply_hamiltonian_noalloc_dev)
against the allocating one (mctdhb_apply_hamiltonian_dev), through the C
ABI: two calls in a row, the second on a smaller space and a smaller
coefficient block (the fixed scratch keeps the first call's contents), in
FP64 (bit-identical: same kernels, same block) and in the mixed form (same
kernels when the tiling agrees; within 1e-6 relative when it does not), and
both within the H.C bars of the oracle (1e-12 FP64, 2e-5 mixed).  A third
call covers the reference-order kernel (M = 9, which reads h and g
directly)."""
import math

import numpy as np
import pytest

import gpu_util as gu
import inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import mctdhb_amd

    mctdhb_amd.lib()
    return mctdhb_amd


def _run(amd, orc, npar, norb, prec, tables=False):
    dim = math.comb(npar + norb - 1, norb - 1)
    c, h, g = inputs.coefficients(dim, 5), inputs.onebody(norb, 5), inputs.twobody(norb, 5, 0.2)
    D = amd.DeviceSpace(npar, norb, tables=tables)
    cdt = np.complex128 if prec == amd.FP64 else np.complex64
    dc, dh, dg = gu.dev(c, cdt), gu.dev(h), gu.dev(g)
    ref_out = gu.empty(dim, "complex128" if prec == amd.FP64 else "complex64")
    new_out = gu.empty(dim, "complex128" if prec == amd.FP64 else "complex64")
    D.apply_hamiltonian_noalloc(dc, dh, dg, new_out, prec=prec, index=amd.INDEX_FLY)
    got_new = gu.host(new_out)
    D.apply_hamiltonian(dc, dh, dg, ref_out, prec=prec, index=amd.INDEX_FLY)
    got_ref = gu.host(ref_out)
    oracle = orc.Space(npar, norb).apply_hamiltonian(c, h, g)
    D.close()
    return got_new, got_ref, oracle


@pytest.mark.parametrize("prec", [0, 1])
def test_no_alloc_matches_allocating_entry_point(amd, orc, prec):
    # first call on the larger space, the second on a smaller one
    for npar, norb in ((20, 5), (8, 3)):
        print(f"no-alloc H.C N={npar} M={norb} prec={prec}", flush=True)
        got_new, got_ref, oracle = _run(amd, orc, npar, norb, prec)
        if prec == amd.FP64:
            assert np.array_equal(got_new, got_ref)
            assert gu.rel_err(got_new, oracle) < 1e-12
        else:
            assert gu.rel_err(got_new, got_ref) < 1e-6
            assert gu.rel_err(got_new, oracle) < 2e-5


def test_no_alloc_reference_order_kernel(amd, orc):
    print("no-alloc H.C N=3 M=9 (tables)", flush=True)
    got_new, got_ref, oracle = _run(amd, orc, 3, 9, amd.FP64, tables=True)
    assert np.array_equal(got_new, got_ref)
    assert gu.rel_err(got_new, oracle) < 1e-12


def test_no_alloc_rejects_bad_arguments(amd):
    D = amd.DeviceSpace(6, 3, tables=False)
    out = gu.empty(D.dim)
    with pytest.raises(amd.MctdhbError):
        D.apply_hamiltonian_noalloc(gu.empty(D.dim), gu.empty(9), gu.empty(81), out, prec=7)
    with pytest.raises(amd.MctdhbError):
        D.apply_hamiltonian_noalloc(None, gu.empty(9), gu.empty(81), out)
    D.close()
