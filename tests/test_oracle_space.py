"""Pins the oracle's Fock-space tables (reference src/configurational/space.c)
against an independent itertools enumeration and closed forms."""
import math

import numpy as np
import pytest

import bruteforce as bf

SMALL = [(1, 1), (1, 4), (4, 1), (2, 3), (3, 3), (5, 2), (4, 4), (6, 3), (3, 6), (7, 4), (5, 5)]


@pytest.mark.parametrize("npar,norb", [(1, 1), (10, 4), (100, 5), (200, 4), (30, 7), (44, 5),
                                       (10000, 2), (5, 11), (6, 8)])
def test_space_dimension_closed_form(orc, npar, norb):
    assert orc.space_dimension(npar, norb) == math.comb(npar + norb - 1, norb - 1)


def test_space_dimension_overflow_guard(orc):
    # C(40+12-1, 11) > 2^32: the reference aborts, the oracle reports it
    assert math.comb(51, 11) > 2 ** 32
    with pytest.raises(OverflowError):
        orc.space_dimension(40, 12)


@pytest.mark.parametrize("npar,norb", SMALL)
def test_hash_table_matches_enumeration(orc, npar, norb):
    S = orc.Space(npar, norb)
    states = bf.basis(npar, norb)
    assert S.dim == len(states)
    assert np.array_equal(S.hash_table.reshape(S.dim, norb), np.array(states, np.uint16))
    # extremes documented by the reference: 0 = all in orbital 0, last = all in last orbital
    assert S.hash_table[0] == npar
    assert S.hash_table[(S.dim - 1) * norb + norb - 1] == npar


@pytest.mark.parametrize("npar,norb", SMALL)
def test_rank_unrank_bijection(orc, npar, norb):
    states = bf.basis(npar, norb)
    for i, s in enumerate(states):
        assert tuple(orc.set_config(i, npar, norb)) == s
        assert orc.hash_index(npar, norb, s) == i


def test_rank_unrank_sampled_large(orc):
    npar, norb = 100, 5
    dim = math.comb(104, 4)
    rng = np.random.default_rng(3)
    for i in list(rng.integers(0, dim, 300)) + [0, 1, dim - 2, dim - 1]:
        conf = orc.set_config(int(i), npar, norb)
        assert int(conf.sum()) == npar
        assert orc.hash_index(npar, norb, conf) == i


def test_subspaces_dim_closed_form(orc):
    npar, norb = 12, 5
    sd = orc.subspaces_dim(npar, norb)
    for m in range(norb + 1):
        for n in range(npar + 1):
            want = 0 if m == 0 else (1 if m == 1 else math.comb(n + m - 1, m - 1))
            assert sd[n + (npar + 1) * m] == want


@pytest.mark.parametrize("npar,norb", SMALL)
def test_single_jump_map(orc, npar, norb):
    S = orc.Space(npar, norb)
    states = bf.basis(npar, norb)
    index = {s: i for i, s in enumerate(states)}
    dim = S.dim
    for i, s in enumerate(states):
        for k in range(norb):
            for l in range(norb):
                got = S.map[i + k * dim + l * norb * dim]
                if s[k] < 1:
                    assert got == 0
                    continue
                occ = list(s)
                occ[k] -= 1
                occ[l] += 1
                assert got == index[tuple(occ)]


@pytest.mark.parametrize("npar,norb", SMALL)
def test_double_jump_maps(orc, npar, norb):
    S = orc.Space(npar, norb)
    states = bf.basis(npar, norb)
    index = {s: i for i, s in enumerate(states)}
    M2 = norb * norb
    ot_total = tt_total = 0
    for i, s in enumerate(states):
        assert S.strideot[i] == ot_total * M2
        assert S.stridett[i] == tt_total * M2
        # same-orbital removals: one chunk per k with n_k >= 2, in k order
        for k in [k for k in range(norb) if s[k] >= 2]:
            base = ot_total * M2
            for l in range(norb):
                for q in range(norb):
                    occ = list(s)
                    occ[k] -= 2
                    occ[l] += 1
                    occ[q] += 1
                    assert S.mapot[base + l + q * norb] == index[tuple(occ)]
            ot_total += 1
        # different-orbital removals: one chunk per occupied pair k < s
        occd = [k for k in range(norb) if s[k] >= 1]
        for a in range(len(occd)):
            for b in range(a + 1, len(occd)):
                k, t = occd[a], occd[b]
                base = tt_total * M2
                for l in range(norb):
                    for q in range(norb):
                        occ = list(s)
                        occ[k] -= 1
                        occ[t] -= 1
                        occ[l] += 1
                        occ[q] += 1
                        assert S.maptt[base + l + q * norb] == index[tuple(occ)]
                tt_total += 1
    assert S.mapot.size == ot_total * M2
    assert S.maptt.size == tt_total * M2


@pytest.mark.parametrize("npar,norb", [(12, 5), (9, 7), (40, 4)])
def test_double_jump_table_sizes_closed_form(orc, npar, norb):
    # chunks: same-orbital M*C(N+M-3, M-1); different-orbital C(M,2)*C(N+M-3, M-1)
    S = orc.Space(npar, norb)
    nc3 = math.comb(npar + norb - 3, norb - 1)
    assert S.mapot.size == norb * nc3 * norb * norb
    assert S.maptt.size == math.comb(norb, 2) * nc3 * norb * norb
