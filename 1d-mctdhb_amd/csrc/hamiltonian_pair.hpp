// hamiltonian_pair.hpp - the mixed-precision H.C kernel on configuration
// pairs.
//
// The gathers of the fused kernel are bound by the vector-memory address
// path (each 64-lane gather instruction costs the texture addresser about
// 16 cycles however wide it is).  Here a lane takes two configurations
// i0, i0 + 1 of one run (the work list Space::d_slots): they share every
// rank delta, so every jump target of i0 + 1 is the target of i0 plus one
// and a single 16-byte load fetches the complex64 coefficients of both.
// Occupations of i0 + 1 are those of i0 with one particle moved from orbital
// 0 to orbital 1; only jumps out of orbitals 0 and 1 can be allowed for one
// of the two and not the other, and the bosonic factors (zero for a jump out
// of an empty orbital) discard that half.  Configurations left single at the
// end of odd-length runs take the one-configuration path.
#pragma once

#include "hamiltonian_impl.hpp"
#include "hamiltonian_window.hpp"
#include "sil_small.hpp"

namespace mctdhb {

struct PairCtx {
    const uint32_t* slots;
    uint32_t n_pairs, n_slots;
};

// Gathers are raw buffer loads (one 16-byte load per pair that the compiler
// cannot split, with the hardware range check as a safety net).  The
// coefficient vector carries one zero element before and after it
// (c[-1], c[dim]), so the window (j0, j0 + 1) of a pair, j0 in [-1, dim - 1],
// is always readable and no gather needs a bounds test.
// Resource over c[-1 .. dim]; byte offset 8 (j + 1) addresses c[j].
__device__ __forceinline__ __amdgpu_buffer_rsrc_t padded_rsrc(const float2* c, uint32_t dim) {
    return __builtin_amdgcn_make_buffer_rsrc((void*) (c - 1), (short) 0, (int) ((dim + 2) * 8u), 0x00020000);
}

__device__ __forceinline__ float2 lo2(const float4& v) { return make_float2(v.x, v.y); }
__device__ __forceinline__ float2 hi2(const float4& v) { return make_float2(v.z, v.w); }

template <int M>
__device__ __forceinline__ void hc_pair_fused(uint32_t i0, const RowCtx& ctx, __amdgpu_buffer_rsrc_t c,
                                              const float2* __restrict__ cf, const uint32_t* __restrict__ Wl,
                                              int ldw, const float* __restrict__ sql, double2& out0,
                                              double2& out1) {
    using Z = float2;
    // occupations of i0; those of i0 + 1 follow
    int na[M], nb[M];
    if (M <= 8 && ctx.occ8) {
        const uint2 w = *(const uint2*) (ctx.occ8 + (size_t) i0 * 8);
#pragma unroll
        for (int k = 0; k < M; k++) na[k] = ((k < 4 ? w.x : w.y) >> (8 * (k & 3))) & 255u;
    } else {
#pragma unroll
        for (int k = 0; k < M; k++) na[k] = ctx.ht[(size_t) i0 * M + k];
    }
#pragma unroll
    for (int k = 0; k < M; k++) nb[k] = na[k];
    nb[0] -= 1;
    nb[1] += 1;
    // square roots sqrt(n-1), sqrt(n), sqrt(n+1), sqrt(n+2) (LDS table)
    float am[M], a0[M], a1[M], a2[M], bm[M], b0[M], b1[M], b2[M];
#pragma unroll
    for (int k = 0; k < M; k++) {
        am[k] = na[k] >= 1 ? sql[na[k] - 1] : 0.f;
        a0[k] = sql[na[k]];
        a1[k] = sql[na[k] + 1];
        a2[k] = sql[na[k] + 2];
        if (k < 2) {
            bm[k] = nb[k] >= 1 ? sql[nb[k] - 1] : 0.f;
            b0[k] = sql[nb[k]];
            b1[k] = sql[nb[k] + 1];
            b2[k] = sql[nb[k] + 2];
        } else {
            bm[k] = am[k];
            b0[k] = a0[k];
            b1[k] = a1[k];
            b2[k] = a2[k];
        }
    }
    // byte-offset rank deltas of i0 (shared by i0 + 1)
    constexpr int MD = M > 1 ? M - 1 : 1;
    uint32_t Dm2[MD], Dm1[MD], Dp1[MD], Dp2[MD];
    {
        int P = 0;
#pragma unroll
        for (int m = 0; m < M - 1; m++) {
            P += na[m];
            const uint32_t* Wm = Wl + m * ldw + kWOff + P;
            const uint32_t w0 = Wm[0];
            Dm2[m] = (Wm[-2] - w0) << 3;
            Dm1[m] = (Wm[-1] - w0) << 3;
            Dp1[m] = (Wm[1] - w0) << 3;
            Dp2[m] = (Wm[2] - w0) << 3;
        }
    }
    const uint32_t is = (i0 + 1) << 3;  // byte offset of c[i0] in the padded resource
    auto fly = [&](int r1, int r2, int q1, int q2) {
        uint32_t j = is;
#pragma unroll
        for (int m = 0; m < M - 1; m++) {
            const int d = jump_delta(m, r1, r2, q1, q2);
            if (d == -2) j -= Dm2[m];
            else if (d == -1) j -= Dm1[m];
            else if (d == 1) j -= Dp1[m];
            else if (d == 2) j -= Dp2[m];
        }
        return j;
    };

    // the off-diagonal sums in FP32; the diagonal terms are added in FP64 at
    // the end (as in hc_row_fused)
    Z acc0 = mk<float>(0, 0), acc1 = mk<float>(0, 0);
    const double2* f64 = (const double2*) (cf + cf_f64_off(M));  // FP64 A, B, d1, d2, d3

    // single jumps k -> l; F_kl(i0 + 1) = F_kl(i0) - B'_kl0 + B'_kl1
#pragma unroll
    for (int k = 0; k < M; k++) {
        const bool va = na[k] >= 1, vb = nb[k] >= 1;
        if (va || vb) {
#pragma unroll
            for (int l = 0; l < M; l++) {
                if (l == k) continue;
                // F in FP64, rounded once (as in hc_row_fused)
                double2 Fd = f64[cf_off_A(M) + k * M + l];
#pragma unroll
                for (int m = 0; m < M - 1; m++) {
                    const double2 B = f64[cf_off_B(M) + (k * M + l) * M + m];
                    Fd.x += (double) na[m] * B.x;
                    Fd.y += (double) na[m] * B.y;
                }
                double2 Fbd = Fd;
                {
                    const double2 B0 = f64[cf_off_B(M) + (k * M + l) * M + 0];
                    Fbd.x -= B0.x;
                    Fbd.y -= B0.y;
                    if (M > 2) {
                        const double2 B1 = f64[cf_off_B(M) + (k * M + l) * M + 1];
                        Fbd.x += B1.x;
                        Fbd.y += B1.y;
                    }
                }
                const Z F = mk<float>((float) Fd.x, (float) Fd.y), Fb = mk<float>((float) Fbd.x, (float) Fbd.y);
                const uint32_t off = fly(k, -1, l, -1);
                const float4 v = ld16(c, off);
                cmac<float>(acc0, F, cscale(lo2(v), a0[k] * a1[l]));
                cmac<float>(acc1, Fb, cscale(hi2(v), b0[k] * b1[l]));
            }
        }
    }
    // creation factors of the two added particles (formed at each use: the
    // vector ALU has room, the register file does not)
    auto Pa = [&](int q, int l) { return a1[q] * (q == l ? a2[q] : a1[l]); };
    auto Pb = [&](int q, int l) { return b1[q] * (q == l ? b2[q] : b1[l]); };

    // two particles out of orbital k
#pragma unroll
    for (int k = 0; k < M; k++) {
        const bool va = na[k] >= 2, vb = nb[k] >= 2;
        if (va || vb) {
            Z ga = mk<float>(0, 0), gb = mk<float>(0, 0);
#pragma unroll
            for (int q = 0; q < M; q++) {
                if (q == k) continue;
#pragma unroll
                for (int l = q; l < M; l++) {
                    if (l == k) continue;
                    const Z G = cf[cf_off_ot(M) + (k * M + q) * M + l];
                    const uint32_t off = fly(k, k, q, l);
                    const float4 v = ld16(c, off);
                    cmac<float>(ga, G, cscale(lo2(v), Pa(q, l)));
                    cmac<float>(gb, G, cscale(hi2(v), Pb(q, l)));
                }
            }
            const float fa = am[k] * a0[k], fb = bm[k] * b0[k];
            acc0.x += fa * ga.x;
            acc0.y += fa * ga.y;
            acc1.x += fb * gb.x;
            acc1.y += fb * gb.y;
        }
    }
    // one particle out of each of k < s
#pragma unroll
    for (int k = 0; k < M; k++) {
#pragma unroll
        for (int s = k + 1; s < M; s++) {
            const bool va = na[k] >= 1 && na[s] >= 1, vb = nb[k] >= 1 && nb[s] >= 1;
            if (va || vb) {
                Z ga = mk<float>(0, 0), gb = mk<float>(0, 0);
#pragma unroll
                for (int q = 0; q < M; q++) {
                    if (q == k || q == s) continue;
#pragma unroll
                    for (int l = q; l < M; l++) {
                        if (l == k || l == s) continue;
                        const Z G = cf[cf_off_tt(M) + ((k * M + s) * M + q) * M + l];
                        const uint32_t off = fly(k, s, q, l);
                        const float4 v = ld16(c, off);
                        cmac<float>(ga, G, cscale(lo2(v), Pa(q, l)));
                        cmac<float>(gb, G, cscale(hi2(v), Pb(q, l)));
                    }
                }
                const float fa = a0[k] * a0[s], fb = b0[k] * b0[s];
                acc0.x += fa * ga.x;
                acc0.y += fa * ga.y;
                acc1.x += fb * gb.x;
                acc1.y += fb * gb.y;
            }
        }
    }
    const float4 cc = ld16(c, is);
    const double2* d64 = f64 + cf_off_d1(M);
    const double2 ea = diag_energy<M, double, double2>(na, d64), eb = diag_energy<M, double, double2>(nb, d64);
    out0 = cmul(ea, make_double2(cc.x, cc.y)) + make_double2(acc0.x, acc0.y);
    out1 = cmul(eb, make_double2(cc.z, cc.w)) + make_double2(acc1.x, acc1.y);
}

// ------------------------------------------------------------- epilogues
// Called by every thread of the block (the Lanczos form reduces over it);
// has1 marks a lane that holds a pair, has0 a lane with any configuration.

template <typename Zout> struct EpiPlainPair {
    Zout* out;
    __device__ __forceinline__ void operator()(uint32_t i0, bool has0, bool has1, double2 v0, double2 v1,
                                               double* lds) const {
        if (has0) out[i0] = from_d2<Zout>(v0);
        if (has1) out[i0 + 1] = from_d2<Zout>(v1);
    }
};

// The Lanczos epilogue on a pair (complex64 Krylov vectors; the two
// entries of a vector are read with one 16-byte load).
struct EpiLanczosPair {
    LanczosHcArgs a;
    uint32_t dim;
    __device__ __forceinline__ void operator()(uint32_t i0, bool has0, bool has1, double2 v0, double2 v1,
                                               double* lds) const {
        const SilState* st = a.st;
        const int it = a.it;
        // entries i0 and i0 + 1 of a complex64 vector (the second zero
        // unless this lane holds a pair)
        auto pair_of = [&](const void* w, double2& x0, double2& x1) {
            const __amdgpu_buffer_rsrc_t r = coef_rsrc((const float2*) w, dim);
            if (has1) {
                const float4 q = ld16(r, i0 << 3);
                x0 = make_double2(q.x, q.y);
                x1 = make_double2(q.z, q.w);
            } else {
                const float2 q = ld8(r, i0 << 3);
                x0 = make_double2(q.x, q.y);
                x1 = make_double2(0, 0);
            }
        };
        double2 o0 = make_double2(0, 0), o1 = make_double2(0, 0);
        const double s_it = st->s[it];
        if (has0) {
            o0 = cscale(v0, s_it);
            if (has1) o1 = cscale(v1, s_it);
            if (it > 0) {
                const double bp = st->beta[it - 1] * st->s[it - 1];
                double2 x0, x1;
                pair_of(a.w[it - 1], x0, x1);
                o0 = o0 - cscale(x0, bp);
                o1 = o1 - cscale(x1, bp);
            }
            if (!a.last) {
                a.out[i0] = o0;
                if (has1) a.out[i0 + 1] = o1;
            }
        }
        const SlotSums ss(lds);
        const int k0 = a.last ? it : 0;
        for (int k = k0; k <= it; k++) {
            double2 p = make_double2(0, 0);
            if (has0) {
                double2 x0, x1;
                pair_of(a.w[k], x0, x1);
                p = cscale(cconjmul(x0, o0) + cconjmul(x1, o1), st->s[k]);
            }
            ss.put(2 * k, p.x);
            ss.put(2 * k + 1, p.y);
        }
        if (!a.last) {
            double n2 = 0;
            if (has0 && it == 0) {
                double2 x0, x1;
                pair_of(a.w[0], x0, x1);
                n2 = x0.x * x0.x + x0.y * x0.y + x1.x * x1.x + x1.y * x1.y;
            }
            ss.put(2 * (it + 1), n2);
            ss.put(2 * (it + 1) + 1, 0.0);
        }
        ss.finish(2 * k0, a.last ? 2 : 2 * (it + 1) + 2, a.part + blockIdx.x, gridDim.x);
    }
};

// One lane per slot: pairs through hc_pair_fused, singles through the
// one-configuration row.
template <int M, typename Epi>
__global__ __launch_bounds__(kHcBlock) void k_hc_pair(RowCtx ctx, PairCtx pc, const float2* __restrict__ c,
                                                      const float2* __restrict__ cf, const uint32_t* __restrict__ Wg,
                                                      int ldw, int stop_check, const SilState* st, Epi epi) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (stop_check && st->stop) return;
    double* red = (double*) smem;
    uint32_t* Wl = (uint32_t*) (smem + kRedBytes);
    const int wsz = (M - 1) * ldw;
    float* sql = (float*) (smem + kRedBytes + ((wsz * 4 + 15) & ~15));
    for (int t = threadIdx.x; t < wsz; t += blockDim.x) Wl[t] = Wg[t];
    for (int t = threadIdx.x; t < ctx.npar + 3; t += blockDim.x) sql[t] = sqrtf((float) t);
    __syncthreads();
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t slot = b * kHcBlock + threadIdx.x;
    const bool has0 = slot < pc.n_slots;
    const bool has1 = slot < pc.n_pairs;
    const uint32_t i0 = has0 ? pc.slots[slot] : 0;
    double2 r0 = make_double2(0, 0), r1 = make_double2(0, 0);
    if (has1) {
        hc_pair_fused<M>(i0, ctx, padded_rsrc(c, ctx.dim), cf, Wl, ldw, sql, r0, r1);
    } else if (has0) {
        float2 own;
        r0 = hc_row_fused<M, float, kIdxFly, float2>(i0, ctx, c, cf, Wl, ldw, sql, own);
    }
    epi(i0, has0, has1, r0, r1, red);
}

template <int M, typename Epi>
static inline int launch_pair(const Space* S, const float2* c, const float2* cf, int stop_check, const SilState* st,
                              Epi epi, hipStream_t stream) {
    const size_t smem = fused_smem(S, sizeof(float));
    if (smem > 64 * 1024) {
        set_error("rank and square-root tables exceed the LDS budget of the fused H.C kernel for this (N, M)");
        return kInvalidArgument;
    }
    if ((uint64_t) S->dim * sizeof(float2) >= (1ull << 32)) {
        set_error("coefficient vector above 4 GiB: the fused H.C kernel needs 32-bit byte offsets");
        return kInvalidArgument;
    }
    PairCtx pc;
    pc.slots = S->d_slots;
    pc.n_pairs = S->n_pairs;
    pc.n_slots = S->n_pairs + S->n_singles;
    k_hc_pair<M, Epi><<<hc_lanczos_blocks(S, kPrecMixed), kHcBlock, smem, stream>>>(row_ctx(S), pc, c, cf, S->d_W,
                                                                                   S->ldw, stop_check, st, epi);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

template <int M>
int HcFused<M>::plain(const Space* S, const void* c, const void* cf, void* out, int prec, int index,
                      hipStream_t st) {
    if (prec == kPrecFP64)
        return launch_fused<M, double, double2>(S, (const double2*) c, (const double2*) cf, index, 0, nullptr,
                                                EpiPlain<double2>{(double2*) out}, st);
    if (use_pair(S, prec)) {
        const float2* cp = hc_padded_copy(S, (const float2*) c, st);
        if (!cp) {
            set_error("padded staging of the coefficient vector failed");
            return kDeviceError;
        }
        return launch_pair<M>(S, cp, (const float2*) cf, 0, nullptr, EpiPlainPair<float2>{(float2*) out}, st);
    }
    if (use_window(S, prec))
        return launch_window<M>(S, (const float2*) c, (const float2*) cf, 0, nullptr, EpiPlain<float2>{(float2*) out},
                                st);
    return launch_fused<M, float, float2>(S, (const float2*) c, (const float2*) cf, kIndexOnTheFly, 0, nullptr,
                                          EpiPlain<float2>{(float2*) out}, st);
}

template <int M>
int HcFused<M>::plain_tiled(const Space* S, const void* c, const void* cf, void* out, int tiling, hipStream_t st) {
    if (tiling == kTilingPair)
        return launch_pair<M>(S, (const float2*) c, (const float2*) cf, 0, nullptr, EpiPlainPair<float2>{(float2*) out},
                              st);
    if (tiling == kTilingWindow)
        return launch_window<M>(S, (const float2*) c, (const float2*) cf, 0, nullptr, EpiPlain<float2>{(float2*) out},
                                st);
    return launch_fused<M, float, float2>(S, (const float2*) c, (const float2*) cf, kIndexOnTheFly, 0, nullptr,
                                          EpiPlain<float2>{(float2*) out}, st);
}

// The Lanczos forms run the one-lane kernel under the window tiling too: the
// windowed kernel is bound by vector ALU work, so the epilogue's dot products
// and sums cost it their full time (measured at N = 100, M = 5: 317-416 us
// per iteration against 295-396 us, and 188 against 149 us for the half row).
template <int M>
int HcFused<M>::lanczos(const Space* S, const void* cf, const LanczosHcArgs& a, int prec, int index,
                        hipStream_t st) {
    if (a.last && hc_lanczos_half(S, prec)) {
        if (prec == kPrecFP64)
            return launch_fused<M, double, double2, EpiRayleigh, true>(S, (const double2*) a.w[a.it],
                                                                       (const double2*) cf, index, 1, a.st,
                                                                       EpiRayleigh{a}, st);
        return launch_fused<M, float, float2, EpiRayleigh, true>(S, (const float2*) a.w[a.it], (const float2*) cf,
                                                                 kIndexOnTheFly, 1, a.st, EpiRayleigh{a}, st);
    }
    if (prec == kPrecFP64)
        return launch_fused<M, double, double2>(S, (const double2*) a.w[a.it], (const double2*) cf, index, 1, a.st,
                                                EpiLanczos<double2>{a}, st);
    if (use_pair(S, prec))
        return launch_pair<M>(S, (const float2*) a.w[a.it], (const float2*) cf, 1, a.st, EpiLanczosPair{a, S->dim},
                              st);
    return launch_fused<M, float, float2>(S, (const float2*) a.w[a.it], (const float2*) cf, kIndexOnTheFly, 1, a.st,
                                          EpiLanczos<float2>{a}, st);
}

}  // namespace mctdhb
