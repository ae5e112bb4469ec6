"""The density kernels of large spaces, against the FP64 oracle (reference
density_matrices.c:9-443) for every orbital count of the fused forms, each
shape above 65536 configurations: the element-family passes (one family per
block) for M <= 5 and the one-pass wave-sum kernel for M >= 6.  FP64 with
on-the-fly and table indices within 1e-10 relative, complex64 products
within 2e-6, and the trace identities tr rho1 = N and
sum_s rho2[k,s,s,l] = (N - 1) rho1[k,l]; the wave form is also deterministic
from call to call."""
import math

import numpy as np
import pytest

import gpu_util as gu
import inputs

pytestmark = pytest.mark.gpu

WAVE_MIN_DIM = 1 << 16


@pytest.fixture(scope="module")
def amd():
    import mctdhb_amd

    mctdhb_amd.lib()
    return mctdhb_amd


@pytest.mark.parametrize("npar,norb", [(400, 3), (80, 4), (40, 5), (25, 6), (18, 7)])
def test_wave_density_matches_oracle(amd, orc, npar, norb):
    dim = math.comb(npar + norb - 1, norb - 1)
    assert dim >= WAVE_MIN_DIM
    print(f"wave density N={npar} M={norb} dim={dim}", flush=True)
    S = orc.Space(npar, norb)
    c = inputs.coefficients(dim, 19)
    ref1, ref2 = S.onebody_dm(c), S.twobody_dm(c)
    D = amd.DeviceSpace(npar, norb, tables=True)
    for prec, index, tol in ((amd.FP64, amd.INDEX_FLY, 1e-10), (amd.FP64, amd.INDEX_TABLES, 1e-10),
                             (amd.MIXED, amd.INDEX_FLY, 2e-6)):
        r1, r2 = gu.empty(norb * norb), gu.empty(norb ** 4)
        src = gu.dev(c) if prec == amd.FP64 else gu.dev(c, np.complex64)
        D.density(src, r1, r2, prec=prec, index=index)
        g1, g2 = gu.host(r1).reshape(norb, norb), gu.host(r2)
        assert gu.rel_err(g1, ref1) < tol, (prec, index)
        assert gu.rel_err(g2, ref2) < tol, (prec, index)
        assert abs(np.trace(g1) - npar) < 10 * tol * npar
        part = np.einsum("kssl->kl", g2.reshape((norb,) * 4, order="F"))
        assert np.max(np.abs(part - (npar - 1) * g1)) < 10 * tol * npar * npar
    D.close()


def test_wave_density_is_deterministic(amd):
    npar, norb = 25, 6
    dim = math.comb(npar + norb - 1, norb - 1)
    D = amd.DeviceSpace(npar, norb, tables=False)
    c = gu.dev(inputs.coefficients(dim, 23))
    outs = []
    for _ in range(2):
        r1, r2 = gu.empty(norb * norb), gu.empty(norb ** 4)
        D.density(c, r1, r2, prec=amd.FP64, index=amd.INDEX_FLY)
        outs.append((gu.host(r1), gu.host(r2)))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    D.close()
