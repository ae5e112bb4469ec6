"""CPU check of the closed-form ranking the device kernels use (fock.hpp):

    rank(n)  = NC - 1 - sum_{m=0}^{M-2} W_m(P_m),   W_m(P) = C(P + m, m + 1)
    rank(n') = rank(n) - sum_m [W_m(P_m + delta_m) - W_m(P_m)]

restated here in integer numpy arithmetic and compared, entry by entry,
with the oracle's restatement of the reference tables (space.c).  Every jump
target the GPU gathers from comes from this formula, so it is checked on the
CPU before any kernel runs.
"""
import math

import numpy as np
import pytest


def w_table(npar, norb):
    ld = npar + 5
    W = np.zeros((max(norb - 1, 1), ld), np.int64)
    for m in range(norb - 1):
        for P in range(-2, npar + 3):
            W[m, P + 2] = math.comb(P + m, m + 1) if P >= 0 else 0
    return W


def delta(m, r1, r2, a1, a2):
    return sum(1 for a in (a1, a2) if a >= 0 and a <= m) - sum(1 for r in (r1, r2) if r >= 0 and r <= m)


def unrank(i, npar, norb, W, dim):
    R = dim - 1 - i
    hi = npar
    n = [0] * norb
    for m in range(norb - 2, -1, -1):
        lo, up = 0, hi
        while lo < up:
            mid = (lo + up + 1) // 2
            if W[m, mid + 2] <= R:
                lo = mid
            else:
                up = mid - 1
        R -= W[m, lo + 2]
        n[m + 1] = hi - lo
        hi = lo
    n[0] = hi
    return n


def jump(i, n, W, norb, r1, r2, a1, a2):
    P = np.cumsum(n)
    j = i
    for m in range(norb - 1):
        d = delta(m, r1, r2, a1, a2)
        if d:
            j -= int(W[m, P[m] + d + 2] - W[m, P[m] + 2])
    return j


SHAPES = [(1, 2), (3, 3), (5, 2), (4, 4), (6, 3), (7, 4), (5, 5), (4, 6), (3, 7), (12, 3), (2, 8)]


@pytest.mark.parametrize("npar,norb", SHAPES)
def test_unrank_and_jumps_match_reference_tables(orc, npar, norb):
    S = orc.Space(npar, norb)
    dim = S.dim
    W = w_table(npar, norb)
    ht = S.hash_table.reshape(dim, norb)
    M2 = norb * norb
    for i in range(dim):
        n = unrank(i, npar, norb, W, dim)
        assert n == list(ht[i])
        for k in range(norb):
            for l in range(norb):
                want = S.map[i + k * dim + l * norb * dim]
                got = jump(i, n, W, norb, k, -1, l, -1) if n[k] >= 1 else 0
                assert got == want
        c = 0
        for k in range(norb):
            if n[k] < 2:
                continue
            for q in range(norb):
                for l in range(norb):
                    j = jump(i, n, W, norb, k, k, l, q)
                    assert 0 <= j < dim
                    assert j == S.mapot[S.strideot[i] + c * M2 + l + q * norb]
            c += 1
        c = 0
        for k in range(norb):
            for s in range(k + 1, norb):
                if n[k] < 1 or n[s] < 1:
                    continue
                for q in range(norb):
                    for l in range(norb):
                        j = jump(i, n, W, norb, k, s, l, q)
                        assert 0 <= j < dim
                        assert j == S.maptt[S.stridett[i] + c * M2 + l + q * norb]
                c += 1


def test_w_table_entries_fit_uint32_for_valid_jumps():
    # at the largest benchmark shapes every W value a valid jump touches is < 2^32
    for npar, norb in [(100, 5), (200, 4), (30, 7), (10000, 2)]:
        dim = math.comb(npar + norb - 1, norb - 1)
        for m in range(norb - 1):
            assert math.comb(npar + m, m + 1) <= dim
