// rowctx.hpp - per-configuration context shared by the H.C and density
// kernels: occupations, bosonic square roots, and the targets of single and
// double jumps either from the reference tables or from rank deltas.
#pragma once

#include "fock.hpp"

namespace mctdhb {

struct RowCtx {
    uint32_t dim;
    int npar;
    const uint16_t* ht;
    const uint8_t* occ8;  // occupations as bytes, row stride 8 (M <= 8, N <= 255), or null
    const uint32_t* map;
    const uint32_t* mapot;
    const uint32_t* maptt;
    const uint32_t* sot;
    const uint32_t* stt;
};

enum : int { kIdxTables = 0, kIdxFly = 1 };

// SHIFT scales every index the context returns by 2^SHIFT, so the H.C
// kernel gets byte offsets (coefficient size 2^SHIFT) with no extra
// arithmetic per gather; SHIFT = 0 returns plain configuration indices.
template <int M, typename R, int IDX, int SHIFT = 0> struct Cfg {
    static constexpr int MD = M > 1 ? M - 1 : 1;
    uint32_t i;   // configuration index
    uint32_t is;  // i << SHIFT
    int n[M];
    R sm[M], s0[M], s1[M], s2[M];  // sqrt(n-1), sqrt(n), sqrt(n+1), sqrt(n+2)
    uint32_t Dm2[MD], Dm1[MD], Dp1[MD], Dp2[MD];
    uint32_t sot, stt;
    const RowCtx* ctx;

    // sql: table of sqrt(0..N+2) in LDS, or null to compute the roots
    __device__ __forceinline__ void load(uint32_t idx, const RowCtx& c, const uint32_t* __restrict__ Wl, int ldw,
                                         const R* __restrict__ sql) {
        if (sql) load_impl<true>(idx, c, Wl, ldw, sql);
        else load_impl<false>(idx, c, Wl, ldw, sql);
    }

    template <bool LUT>
    __device__ __forceinline__ void load_impl(uint32_t idx, const RowCtx& c, const uint32_t* __restrict__ Wl,
                                              int ldw, const R* __restrict__ sql) {
        i = idx;
        is = idx << SHIFT;
        ctx = &c;
        if (M <= 8 && c.occ8) {
            // one 8-byte load instead of M 2-byte loads
            const uint2 w = *(const uint2*) (c.occ8 + (size_t) i * 8);
#pragma unroll
            for (int k = 0; k < M; k++) n[k] = ((k < 4 ? w.x : w.y) >> (8 * (k & 3))) & 255u;
        } else {
#pragma unroll
            for (int k = 0; k < M; k++) n[k] = c.ht[(size_t) i * M + k];
        }
#pragma unroll
        for (int k = 0; k < M; k++) {
            if constexpr (LUT) {
                sm[k] = n[k] >= 1 ? sql[n[k] - 1] : R(0);
                s0[k] = sql[n[k]];
                s1[k] = sql[n[k] + 1];
                s2[k] = sql[n[k] + 2];
            } else {
                sm[k] = n[k] >= 1 ? sqrt(R(n[k] - 1)) : R(0);
                s0[k] = sqrt(R(n[k]));
                s1[k] = sqrt(R(n[k] + 1));
                s2[k] = sqrt(R(n[k] + 2));
            }
        }
        if constexpr (IDX == kIdxFly) {
            int P = 0;
#pragma unroll
            for (int m = 0; m < M - 1; m++) {
                P += n[m];
                const uint32_t* Wm = Wl + m * ldw + kWOff + P;
                const uint32_t w0 = Wm[0];
                Dm2[m] = (Wm[-2] - w0) << SHIFT;
                Dm1[m] = (Wm[-1] - w0) << SHIFT;
                Dp1[m] = (Wm[1] - w0) << SHIFT;
                Dp2[m] = (Wm[2] - w0) << SHIFT;
            }
        } else {
            sot = c.sot[i];
            stt = c.stt[i];
        }
    }

    // rank (<< SHIFT) of the configuration reached by removing r1, r2 and
    // adding a1, a2
    __device__ __forceinline__ uint32_t fly(int r1, int r2, int a1, int a2) const {
        uint32_t j = is;
#pragma unroll
        for (int m = 0; m < M - 1; m++) {
            const int d = jump_delta(m, r1, r2, a1, a2);
            if (d == -2) j -= Dm2[m];
            else if (d == -1) j -= Dm1[m];
            else if (d == 1) j -= Dp1[m];
            else if (d == 2) j -= Dp2[m];
        }
        return j;
    }
    // #{h < k : n_h >= 2}
    __device__ __forceinline__ uint32_t ot_chunk(int k) const {
        uint32_t c = 0;
#pragma unroll
        for (int h = 0; h < M; h++)
            if (h < k) c += n[h] >= 2;
        return c;
    }
    // occupied pairs (h < g) with h < k, plus occupied g in (k, s)
    __device__ __forceinline__ uint32_t tt_chunk(int k, int s) const {
        uint32_t c = 0;
#pragma unroll
        for (int h = 0; h < M; h++)
#pragma unroll
            for (int g = h + 1; g < M; g++)
                if (h < k) c += (n[h] >= 1) & (n[g] >= 1);
#pragma unroll
        for (int g = 0; g < M; g++)
            if (g > k && g < s) c += n[g] >= 1;
        return c;
    }
    // i - e_k + e_l
    __device__ __forceinline__ uint32_t single(int k, int l) const {
        if constexpr (IDX == kIdxFly) return fly(k, -1, l, -1);
        else return ctx->map[i + (uint32_t) k * ctx->dim + (uint32_t) l * M * ctx->dim] << SHIFT;
    }
    // i - 2 e_k + e_q + e_l
    __device__ __forceinline__ uint32_t same(int k, int q, int l) const {
        if constexpr (IDX == kIdxFly) return fly(k, k, q, l);
        else return ctx->mapot[sot + ot_chunk(k) * (M * M) + q + l * M] << SHIFT;
    }
    // i - e_k - e_s + e_q + e_l (k < s)
    __device__ __forceinline__ uint32_t diff(int k, int s, int q, int l) const {
        if constexpr (IDX == kIdxFly) return fly(k, s, q, l);
        else return ctx->maptt[stt + tt_chunk(k, s) * (M * M) + q + l * M] << SHIFT;
    }
};

inline RowCtx make_row_ctx(uint32_t dim, int npar, const uint16_t* ht, const uint32_t* map, const uint32_t* mapot,
                           const uint32_t* maptt, const uint32_t* sot, const uint32_t* stt,
                           const uint8_t* occ8 = nullptr) {
    RowCtx c;
    c.dim = dim;
    c.npar = npar;
    c.ht = ht;
    c.occ8 = occ8;
    c.map = map;
    c.mapot = mapot;
    c.maptt = maptt;
    c.sot = sot;
    c.stt = stt;
    return c;
}

}  // namespace mctdhb
