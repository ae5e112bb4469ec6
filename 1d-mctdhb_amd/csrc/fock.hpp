// fock.hpp - combinatorial ranking of occupation vectors on the device.
//
// Ranking convention (the reference's, space.c:138-179): the last orbital is
// most significant, index 0 holds every particle in orbital 0.  With prefix
// sums P_m = n_0 + ... + n_m the rank has the closed form
//
//     rank(n) = NC - 1 - sum_{m=0}^{M-2} C(P_m + m, m + 1),
//
// i.e. NC - 1 minus the colex rank of the strictly increasing set
// {P_m + m}.  A jump that moves particles between orbitals shifts each P_m
// by a small delta (+-1, +-2), so the rank of any jump target is
//
//     rank(n') = rank(n) - sum_m [W_m(P_m + delta_m) - W_m(P_m)],
//     W_m(P) = C(P + m, m + 1),
//
// which needs only a small (M-1) x (N+5) table of W (kept in LDS) and no
// jump tables at all.  The same formula builds the reference-layout tables.
#pragma once

#include "common.hpp"

namespace mctdhb {

// W table row stride and offset: W[m * ldw + P + kWOff] for P in [-2, N+2].
constexpr int kWOff = 2;
inline int w_ld(int npar) { return npar + 5; }

// Host: exact binomial (0 when k > n), saturating at 2^64 - 1.
uint64_t binom_u64(uint64_t n, uint64_t k);
// Host: fill the W table (values modulo 2^32; entries that could exceed it
// are only reached by jumps out of empty orbitals, which are never used).
void fill_w_table(int npar, int norb, uint32_t* w);

// Delta in P_m for a jump removing one particle from r1 and r2 (-1: none)
// and adding one to a1 and a2 (-1: none).
__host__ __device__ constexpr int jump_delta(int m, int r1, int r2, int a1, int a2) {
    return (a1 >= 0 && a1 <= m) + (a2 >= 0 && a2 <= m) - (r1 >= 0 && r1 <= m) - (r2 >= 0 && r2 <= m);
}

// Runtime-M rank of a jump target (generic kernels and table builds).
__device__ __forceinline__ uint32_t jump_rank_rt(uint32_t i, int M, const uint32_t* P, const uint32_t* W,
                                                 int ldw, int r1, int r2, int a1, int a2) {
    uint32_t j = i;
    for (int m = 0; m < M - 1; m++) {
        const int d = jump_delta(m, r1, r2, a1, a2);
        if (d != 0) j -= W[m * ldw + (int) P[m] + d + kWOff] - W[m * ldw + (int) P[m] + kWOff];
    }
    return j;
}

// Runtime-M unranking: occupations n[0..M) of configuration i.
__device__ __forceinline__ void unrank_rt(uint32_t i, int N, int M, uint32_t dim, const uint32_t* W, int ldw,
                                          uint32_t* n) {
    uint32_t R = dim - 1u - i;
    uint32_t hi = (uint32_t) N;  // P_{m+1}
    for (int m = M - 2; m >= 0; m--) {
        const uint32_t* Wm = W + m * ldw + kWOff;
        // largest P in [0, hi] with W_m(P) <= R (W_m is non-decreasing in P)
        uint32_t lo = 0, up = hi;
        while (lo < up) {
            uint32_t mid = (lo + up + 1) >> 1;
            if (Wm[mid] <= R) lo = mid;
            else up = mid - 1;
        }
        R -= Wm[lo];
        n[m + 1] = hi - lo;
        hi = lo;
    }
    n[0] = hi;
}

}  // namespace mctdhb
