"""Independent dense Fock-space checker for small (N, M).

Nothing here shares code or formulas with the oracle or the GPU library:
the basis is enumerated with itertools and ordered by sorting, and every
operator is applied to explicit occupation tuples with the textbook bosonic
factors a|n> = sqrt(n)|n-1>, a+|n> = sqrt(n+1)|n+1>.  It pins the oracle's
tables, H.C and density matrices on shapes small enough to build densely.

Ordering: the reference ranks configurations with the last orbital most
significant (index 0 = all particles in orbital 0), i.e. ascending
lexicographic order of the reversed occupation tuple.
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def basis(npar: int, norb: int):
    """All occupation tuples, in the reference's rank order."""
    states = [c for c in itertools.product(range(npar + 1), repeat=norb) if sum(c) == npar]
    states.sort(key=lambda c: tuple(reversed(c)))
    return states


def _apply(ops, state):
    """Apply a sequence of (dagger, mode) right-to-left; returns (factor, state)."""
    occ = list(state)
    fac = 1.0
    for dagger, mode in reversed(ops):
        if dagger:
            fac *= math.sqrt(occ[mode] + 1)
            occ[mode] += 1
        else:
            if occ[mode] == 0:
                return 0.0, None
            fac *= math.sqrt(occ[mode])
            occ[mode] -= 1
    return fac, tuple(occ)


def operator(states, ops):
    """Dense matrix <i| ops |j> on the number-conserving basis."""
    index = {s: i for i, s in enumerate(states)}
    n = len(states)
    mat = np.zeros((n, n))
    for j, s in enumerate(states):
        fac, t = _apply(ops, s)
        if t is not None:
            mat[index[t], j] += fac
    return mat


def hamiltonian(npar, norb, hob, hint):
    """H = sum h_kl a+_k a_l + 1/2 sum g[k + sM + qM^2 + lM^3] a+_k a+_s a_q a_l."""
    states = basis(npar, norb)
    M = norb
    hob = np.asarray(hob).reshape(M, M)
    hint = np.asarray(hint).ravel()
    H = np.zeros((len(states), len(states)), complex)
    for k in range(M):
        for l in range(M):
            if hob[k, l] != 0:
                H += hob[k, l] * operator(states, [(1, k), (0, l)])
    for k, s, q, l in itertools.product(range(M), repeat=4):
        g = hint[k + s * M + q * M * M + l * M ** 3]
        if g != 0:
            H += 0.5 * g * operator(states, [(1, k), (1, s), (0, q), (0, l)])
    return H


def onebody_dm(npar, norb, coef):
    states = basis(npar, norb)
    c = np.asarray(coef)
    rho = np.zeros((norb, norb), complex)
    for k in range(norb):
        for l in range(norb):
            rho[k, l] = np.vdot(c, operator(states, [(1, k), (0, l)]) @ c)
    return rho


def twobody_dm(npar, norb, coef):
    """rho2[k + sM + qM^2 + lM^3] = <a+_k a+_s a_q a_l>."""
    states = basis(npar, norb)
    c = np.asarray(coef)
    M = norb
    rho = np.zeros(M ** 4, complex)
    for k, s, q, l in itertools.product(range(M), repeat=4):
        rho[k + s * M + q * M * M + l * M ** 3] = np.vdot(
            c, operator(states, [(1, k), (1, s), (0, q), (0, l)]) @ c)
    return rho


def jump_target(states, i, remove, add):
    """Index of the configuration reached from states[i] by removing one
    particle per entry of ``remove`` and adding one per entry of ``add``."""
    occ = list(states[i])
    for r in remove:
        occ[r] -= 1
    if min(occ) < 0:
        return None
    for a in add:
        occ[a] += 1
    return {s: n for n, s in enumerate(states)}[tuple(occ)]
