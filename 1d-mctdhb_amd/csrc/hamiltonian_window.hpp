// hamiltonian_window.hpp - the mixed H.C with the coefficient windows of a
// block staged in LDS (the third mixed tiling, `kTilingWindow`, for the
// plain product).
//
// A jump changes the prefix occupations P_m (m = 0 .. M-2) by d_m in
// [-2, 2], and the target's rank differs from the source's by the sum over
// levels of W_m(P_m + d_m) - W_m(P_m).  Levels 0 and 1 move the rank by at
// most 2 P_1 + 5 (W_0(P) = P, W_1(P) = P (P + 1) / 2); the higher levels by
// far more.  Jump types with the same changes at levels >= 2 (one "far
// group") therefore send a block of 256 consecutive configurations into a
// range of a little over 256 + 4N ranks, unless one of the higher prefixes
// changes inside the block.  The kernel walks the far groups in turn: the
// block copies the group's window of the coefficient vector into LDS with
// one or two 16-byte loads per thread (double-buffered: the next group's copy
// is in flight while this group is summed), and every lane reads its targets
// of the group from LDS instead of gathering them through the texture path.
// At N = 100, M = 5 that is 19 groups; the vector-memory instructions per
// wave drop from 114 to 36 and the texture addresser from 88 % to 33 % busy,
// but the kernel is then bound by vector ALU work (about 1700 instructions
// per wave against 1600 for the one-lane kernel, which was 76 % ALU-busy):
// 265 us against 283 us per product.
//
// A lane checks once per group that the targets its valid jumps can reach,
// [base - (2 P_1 + 5), base + (2 P_1 + 5)] around the group's base rank, lie
// in the window; one that fails for some group that can hold a valid jump
// (1.9 % of the lanes at N = 100, M = 5, mostly in blocks that cross a change
// of P_2 or P_3) recomputes its row with the gathering kernel's code.
// Invalid jumps (out of an empty orbital) are summed too, from an address
// masked into the window and with a zero annihilation factor, which spares a
// branch per jump type.
//
// Arithmetic is that of the complex64 lane kernel (hc_row_fused, R = float):
// FP32 products and sums, the single-jump coefficient F_kl(n) and the
// diagonal term in FP64.  Only the order of the off-diagonal sum differs
// (by far group, then by removed orbitals).
#pragma once
#include "hamiltonian_impl.hpp"

#include <utility>

namespace mctdhb {

// Jump types of one row in far-group order.
struct WinTab {
    static constexpr int kMax = 130;  // types of a full row at M = 5
    int nt, ng;
    int8_t r1[kMax], r2[kMax], a1[kMax], a2[kMax];  // r2 = -1: single jump; r2 = r1: two out of r1
    int16_t gs[kMax + 1];                            // group g holds types gs[g] .. gs[g+1]-1
};

// The enumeration and the half-row rule are those of hc_row_fused: singles
// k -> l, then two out of k into q <= l, then one out of each of k < s into
// q <= l; with HALF only l < k (singles) and q < k (doubles).  Types are then
// sorted by their changes at levels 2 .. M-2, stably.
template <int M, bool HALF> __host__ __device__ constexpr WinTab win_tab() {
    WinTab T{};
    int8_t r1[WinTab::kMax] = {}, r2[WinTab::kMax] = {}, a1[WinTab::kMax] = {}, a2[WinTab::kMax] = {};
    int key[WinTab::kMax] = {};
    int n = 0;
    auto add = [&](int x1, int x2, int y1, int y2) {
        r1[n] = (int8_t) x1;
        r2[n] = (int8_t) x2;
        a1[n] = (int8_t) y1;
        a2[n] = (int8_t) y2;
        int kv = 0;
        for (int m = 2; m < M - 1; m++) kv = kv * 5 + jump_delta(m, x1, x2, y1, y2) + 2;
        key[n] = kv;
        n++;
    };
    for (int k = 0; k < M; k++)
        for (int l = 0; l < M; l++)
            if (l != k && !(HALF && l > k)) add(k, -1, l, -1);
    for (int k = 0; k < M; k++)
        for (int q = 0; q < M; q++) {
            if (q == k || (HALF && q > k)) continue;
            for (int l = q; l < M; l++)
                if (l != k) add(k, k, q, l);
        }
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++)
            for (int q = 0; q < M; q++) {
                if (q == k || q == s || (HALF && q > k)) continue;
                for (int l = q; l < M; l++)
                    if (l != k && l != s) add(k, s, q, l);
            }
    T.nt = n;
    int out = 0, last = -1;
    while (out < n) {
        int best = 1 << 30;
        for (int t = 0; t < n; t++)
            if (key[t] > last && key[t] < best) best = key[t];
        T.gs[T.ng++] = (int16_t) out;
        for (int t = 0; t < n; t++)
            if (key[t] == best) {
                T.r1[out] = r1[t];
                T.r2[out] = r2[t];
                T.a1[out] = a1[t];
                T.a2[out] = a2[t];
                out++;
            }
        last = best;
    }
    T.gs[T.ng] = (int16_t) out;
    return T;
}

constexpr int kWinMaxGroups = 32;  // far groups of M <= 5 (19 at M = 5)
constexpr int kWinMaxOrb = 5;      // more orbitals: too many groups of too few jumps

// Window geometry of a launch.  A window is WB = FILL * 4096 bytes (a power
// of two: every thread copies FILL 16-byte chunks, and a lane's LDS address
// is its offset masked to the window, one instruction); the two window
// buffers are the kernel's static LDS, so their base folds into the
// instruction offsets.  The block's window
// of group g starts at glo[g] = (the first configuration's group base) - m0b,
// rounded down to 16 bytes, which centres the block's 256 bases in it.
struct WinArgs {
    uint32_t m0b, glo_off;  // glo_off: in the dynamic LDS
};

// FILL of a space (0: the tiling is not available) and the dynamic LDS
// bytes of its launch (the windows, 2 FILL * 4096 bytes, are static).
inline size_t window_geometry(const Space* S, WinArgs& wa, int& fill) {
    wa = WinArgs{0, 0};
    fill = 0;
    if (S->norb < 2 || S->norb > kWinMaxOrb || (uint64_t) S->dim * 8 >= (1ull << 32)) return 0;
    // the block's bases span 256 entries; levels 0 and 1 reach 2 P_1 + 5
    const uint32_t need = 8u * ((uint32_t) kHcBlock + 2u * (2u * (uint32_t) S->npar + 6u));
    const uint32_t wb = need <= 4096u ? 4096u : need <= 8192u ? 8192u : 0u;
    if (!wb) return 0;
    fill = (int) (wb / 4096u);
    wa.m0b = (wb - 8u * kHcBlock) / 2u;
    const size_t wsz = (size_t) (S->norb - 1) * S->ldw * 4;
    const size_t off = kRedBytes + ((wsz + 15) & ~(size_t) 15) + ((4 * (size_t) (S->npar + 3) + 15) & ~(size_t) 15);
    wa.glo_off = (uint32_t) off;
    return off + 4 * kWinMaxGroups;
}

inline bool window_possible(const Space* S) {
    WinArgs wa;
    int fill;
    const size_t smem = window_geometry(S, wa, fill);
    return fill > 0 && smem + 2 * 4096 * (size_t) fill <= 64 * 1024;
}

// The table once per (M, HALF); every jump type below is a template
// instance, so the compiler sees its orbitals as constants.
template <int M, bool HALF> struct WinTabOf {
    static constexpr WinTab T = win_tab<M, HALF>();
    // same removal as type t - 1 / t + 1 within group g
    static constexpr bool same_removal(int t, int u) { return T.r1[t] == T.r1[u] && T.r2[t] == T.r2[u]; }
};

// Per-lane state of the windowed row.
template <int M, int FILL> struct WinRow {
    Cfg<M, float, kIdxFly, 3> x;
    float P[M][M];  // creation factors of two added particles
    int pfx[M];     // prefix occupations P_m
    float2 acc;     // off-diagonal sum
    float2 g;       // sum of the current run of jumps with one removal
    uint32_t miss;
    uint32_t brel;  // this group's base rank (bytes) relative to its window
    uint32_t reach, span;
    const float2* cf;
    const double2* d64;  // FP64 A', B', d1, d2, d3
    float4 stage[FILL];  // next window's copy in flight
};

// rank change (bytes) of levels lo <= m < hi of jump type t
template <int M, bool HALF, int t, int lo, int hi, int FILL>
__device__ __forceinline__ uint32_t win_shift(const WinRow<M, FILL>& r) {
    constexpr int r1 = WinTabOf<M, HALF>::T.r1[t], r2 = WinTabOf<M, HALF>::T.r2[t];
    constexpr int a1 = WinTabOf<M, HALF>::T.a1[t], a2 = WinTabOf<M, HALF>::T.a2[t];
    uint32_t f = 0;
#pragma unroll
    for (int m = lo; m < hi; m++) {
        const int d = jump_delta(m, r1, r2, a1, a2);
        if (d == -2) f += r.x.Dm2[m];
        else if (d == -1) f += r.x.Dm1[m];
        else if (d == 1) f += r.x.Dp1[m];
        else if (d == 2) f += r.x.Dp2[m];
    }
    return f;
}

// One jump type of group G: its target from the window, times its
// coefficient and creation factor, summed into the run of jumps that share
// its removal; the run's annihilation factor is applied once at its end.
template <int M, bool HALF, int G, int t, int FILL>
__device__ __forceinline__ void win_jump(WinRow<M, FILL>& r, const unsigned char* win) {
    using Tab = WinTabOf<M, HALF>;
    constexpr int k = Tab::T.r1[t], s = Tab::T.r2[t], q = Tab::T.a1[t], l = Tab::T.a2[t];
    constexpr bool first = t == Tab::T.gs[G] || !Tab::same_removal(t, t - 1);
    constexpr bool last = t + 1 == Tab::T.gs[G + 1] || !Tab::same_removal(t, t + 1);
    constexpr uint32_t wb = FILL * 4096u;
    const uint32_t rel = r.brel - win_shift<M, HALF, t, 0, (M - 1 < 2 ? M - 1 : 2)>(r);
    // masked into this group's buffer: out-of-window targets (invalid jumps,
    // or a lane that misses) read some in-window entry
    const float2 v = *(const float2*) (win + ((rel & (wb - 8u)) | (G & 1) * wb));
    float2 coef, sv;
    if constexpr (s < 0) {
        // F_kl(n) = A'_kl + sum_{m < M-1} n_m B'_klm in FP64, rounded once
        double2 Fd = r.d64[cf_off_A(M) + k * M + q];
#pragma unroll
        for (int m = 0; m < M - 1; m++) {
            const double2 B = r.d64[cf_off_B(M) + (k * M + q) * M + m];
            Fd.x += (double) r.x.n[m] * B.x;
            Fd.y += (double) r.x.n[m] * B.y;
        }
        coef = mk<float>((float) Fd.x, (float) Fd.y);
        sv = cscale(v, r.x.s1[q]);
    } else if constexpr (s == k) {
        coef = r.cf[cf_off_ot(M) + (k * M + q) * M + l];
        sv = cscale(v, r.P[q][l]);
    } else {
        coef = r.cf[cf_off_tt(M) + ((k * M + s) * M + q) * M + l];
        sv = cscale(v, r.P[q][l]);
    }
    if constexpr (first) r.g = cmul(coef, sv);
    else cmac<float>(r.g, coef, sv);
    if constexpr (last) {
        const float a = s < 0 ? r.x.s0[k] : (s == k ? r.x.sm[k] * r.x.s0[k] : r.x.s0[k] * r.x.s0[s]);
        r.acc.x += a * r.g.x;
        r.acc.y += a * r.g.y;
    }
}

template <int M, bool HALF, int G, int T0, int FILL, int... js>
__device__ __forceinline__ void win_jumps(WinRow<M, FILL>& r, const unsigned char* win,
                                          std::integer_sequence<int, js...>) {
    (win_jump<M, HALF, G, T0 + js>(r, win), ...);
}

// Window base of group G: the first configuration's group base, less m0b.
template <int M, bool HALF, int G, int FILL>
__device__ __forceinline__ uint32_t win_base(const WinRow<M, FILL>& r, uint32_t m0b) {
    const uint32_t b0 = r.x.is - win_shift<M, HALF, WinTabOf<M, HALF>::T.gs[G], 2, M - 1>(r);
    return (b0 < m0b ? 0u : b0 - m0b) & ~15u;
}
template <int M, bool HALF, int FILL, int... gs>
__device__ __forceinline__ void win_bases(const WinRow<M, FILL>& r, uint32_t* glo, uint32_t m0b,
                                          std::integer_sequence<int, gs...>) {
    ((glo[gs] = win_base<M, HALF, gs>(r, m0b)), ...);
}

// The window copies: coefficient bytes [lo, lo + FILL * 4096) through
// registers into one LDS buffer.  Loads past the vector return zero (the
// buffer resource's range check).
template <int M, int FILL>
__device__ __forceinline__ void win_fetch(WinRow<M, FILL>& r, __amdgpu_buffer_rsrc_t rs, uint32_t lo) {
#pragma unroll
    for (int j = 0; j < FILL; j++) r.stage[j] = ld16(rs, lo + (threadIdx.x + (uint32_t) j * kHcBlock) * 16u);
}
template <int M, int FILL>
__device__ __forceinline__ void win_store(const WinRow<M, FILL>& r, unsigned char* buf) {
#pragma unroll
    for (int j = 0; j < FILL; j++) *(float4*) (buf + (threadIdx.x + (uint32_t) j * kHcBlock) * 16u) = r.stage[j];
}

// Groups G, G+1, ...: copy group G+1's window while summing group G.
template <int M, bool HALF, int G, int FILL>
__device__ __forceinline__ void win_groups(WinRow<M, FILL>& r, const uint32_t* glo, unsigned char* win,
                                           __amdgpu_buffer_rsrc_t rs) {
    constexpr WinTab T = WinTabOf<M, HALF>::T;
    constexpr uint32_t wb = FILL * 4096u;
    if constexpr (G < T.ng) {
        if constexpr (G + 1 < T.ng) win_fetch<M, FILL>(r, rs, glo[G + 1]);
        r.brel = r.x.is - win_shift<M, HALF, T.gs[G], 2, M - 1>(r) - glo[G];
        // a jump of the group can be valid only if every level m >= 2 keeps
        // 0 <= P_m + d_m <= P_{m+1} + d_{m+1} (the last level: <= N)
        bool live = true;
        constexpr int t0 = T.gs[G];
        constexpr int r1 = T.r1[t0], r2 = T.r2[t0], a1 = T.a1[t0], a2 = T.a2[t0];
#pragma unroll
        for (int m = 2; m < M - 1; m++) {
            const int d = jump_delta(m, r1, r2, a1, a2);
            const int dn = m + 1 < M - 1 ? jump_delta(m + 1, r1, r2, a1, a2) : 0;
            if (d < 0) live = live && r.pfx[m] >= -d;
            if (d - dn > 0) live = live && r.x.n[m + 1] >= d - dn;
        }
        win_jumps<M, HALF, G, T.gs[G]>(r, win, std::make_integer_sequence<int, T.gs[G + 1] - T.gs[G]>{});
        r.miss |= (live && r.brel - r.reach > r.span) ? 1u : 0u;
        // keep this group's products here: without it the compiler moves
        // the whole sum behind the last barrier (it is only used on the
        // in-window path) and holds every window value until then
        asm volatile("" : "+v"(r.acc.x), "+v"(r.acc.y), "+v"(r.miss));
        if constexpr (G + 1 < T.ng) win_store<M, FILL>(r, win + ((G + 1) & 1) * wb);
        __syncthreads();
        win_groups<M, HALF, G + 1, FILL>(r, glo, win, rs);
    }
}

// The full row (the table's half-row rule is not used by a kernel).
template <int M, int FILL, typename Epi>
__global__ __launch_bounds__(kHcBlock) void k_hc_window(RowCtx ctx, const float2* __restrict__ c,
                                                        const float2* __restrict__ cf,
                                                        const uint32_t* __restrict__ Wg, int ldw, int stop_check,
                                                        const SilState* st, Epi epi, WinArgs wa) {
    constexpr bool HALF = false;
    constexpr WinTab T = WinTabOf<M, HALF>::T;
    constexpr uint32_t wb = FILL * 4096u;
    // a window's buffer is addressed by masking: it sits on a multiple of
    // its size
    __shared__ __align__(2 * wb) unsigned char win[2 * wb];
    extern __shared__ __align__(16) unsigned char smem[];
    if (stop_check && st->stop) return;
    double* red = (double*) smem;
    uint32_t* Wl = (uint32_t*) (smem + kRedBytes);
    const int wsz = (M - 1) * ldw;
    float* sql = (float*) (smem + kRedBytes + ((wsz * 4 + 15) & ~15));
    uint32_t* glo = (uint32_t*) (smem + wa.glo_off);
    for (int t = threadIdx.x; t < wsz; t += blockDim.x) Wl[t] = Wg[t];
    for (int t = threadIdx.x; t < ctx.npar + 3; t += blockDim.x) sql[t] = sqrtf((float) t);
    __syncthreads();
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = b * kHcBlock + threadIdx.x;
    const bool valid = i < ctx.dim;
    WinRow<M, FILL> r;
    // lanes past the end shadow the block's first configuration
    r.x.template load_impl<true>(valid ? i : b * kHcBlock, ctx, Wl, ldw, sql);
    const int* n = r.x.n;
    if (threadIdx.x == 0) win_bases<M, HALF>(r, glo, wa.m0b, std::make_integer_sequence<int, T.ng>{});
    __syncthreads();

    const __amdgpu_buffer_rsrc_t rs = coef_rsrc(c, ctx.dim);
    win_fetch<M, FILL>(r, rs, glo[0]);
    win_store<M, FILL>(r, win);
    __syncthreads();

#pragma unroll
    for (int q = 0; q < M; q++)
#pragma unroll
        for (int l = q; l < M; l++) r.P[q][l] = r.x.s1[q] * (q == l ? r.x.s2[q] : r.x.s1[l]);
    r.pfx[0] = n[0];
#pragma unroll
    for (int m = 1; m < M; m++) r.pfx[m] = r.pfx[m - 1] + n[m];
    r.cf = cf;
    r.d64 = (const double2*) (cf + cf_f64_off(M));
    // valid jumps of levels 0 and 1 move the rank by at most 2 P_1 + 5
    r.reach = 8u * (M > 2 ? 2u * (uint32_t) (n[0] + n[1]) + 5u : 2u);
    r.span = wb - 8u - 2u * r.reach;  // 2 reach <= 2 m0b < wb: no wrap
    r.acc = mk<float>(0.f, 0.f);
    r.g = mk<float>(0.f, 0.f);
    // a copy may return a whole 16-byte chunk as zero when only its first
    // entry lies in the vector (odd NC): rows that can reach the last
    // configuration (all particles in orbital M-1) take the gathering path
    r.miss = n[M - 1] + 2 >= ctx.npar ? 1u : 0u;
    win_groups<M, HALF, 0, FILL>(r, glo, win, rs);

    double2 row = make_double2(0, 0), ci = make_double2(0, 0);
    if (valid) {
        const float2 own = gather_b<float, float2>(c, r.x.is);
        ci = make_double2(own.x, own.y);
        if (r.miss) {
            float2 dummy;
            row = hc_row_fused<M, float, kIdxFly, float2, HALF>(i, ctx, c, cf, Wl, ldw, sql, dummy);
        } else {
            const double2 e = diag_energy<M, double, double2>(n, r.d64 + cf_off_d1(M));
            row = cmul(e, ci) + make_double2(r.acc.x, r.acc.y);
        }
    }
    epi(i, valid, row, ci, red);
}

template <int M, typename Epi>
static inline int launch_window(const Space* S, const float2* c, const float2* cf, int stop_check,
                                const SilState* st, Epi epi, hipStream_t stream) {
    if constexpr (M > kWinMaxOrb) {
        set_error("the windowed H.C covers 2 <= norb <= 5");
        return kInvalidArgument;
    } else {
        WinArgs wa;
        int fill;
        const size_t smem = window_geometry(S, wa, fill);
        if (!window_possible(S)) {
            set_error("coefficient windows of this (N, M) exceed the LDS budget of the windowed H.C kernel");
            return kInvalidArgument;
        }
        const unsigned nb = hc_grid_blocks(S->dim);
        const RowCtx ctx = row_ctx(S);
        if (fill == 1)
            k_hc_window<M, 1, Epi><<<nb, kHcBlock, smem, stream>>>(ctx, c, cf, S->d_W, S->ldw, stop_check, st,
                                                                         epi, wa);
        else
            k_hc_window<M, 2, Epi><<<nb, kHcBlock, smem, stream>>>(ctx, c, cf, S->d_W, S->ldw, stop_check, st,
                                                                         epi, wa);
        MCTDHB_HIP_CHECK(hipGetLastError());
        return kOk;
    }
}

}  // namespace mctdhb
