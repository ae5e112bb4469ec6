// hamiltonian_impl.hpp - templates of the fused H.C kernel, instantiated
// once per orbital count in hc_m<M>.hip (split so they compile in parallel).
#pragma once

#include "hamiltonian.hpp"
#include "rowctx.hpp"

#include <type_traits>

namespace mctdhb {

constexpr int kHcBlock = 256;

// ----------------------------------------------------------- fused row

// log2 of the coefficient size: gathers use 32-bit byte offsets from a
// uniform base (one VALU op per gather for the offset, none for a 64-bit
// address), which needs dim * sizeof(Zin) < 2^32 (checked at launch).
template <typename Zin> constexpr int coef_shift() { return sizeof(Zin) == 16 ? 4 : 3; }

template <typename R, typename Zin>
__device__ __forceinline__ cplx<R> gather_b(const Zin* __restrict__ c, uint32_t off) {
    const Zin v = *(const Zin*) ((const char*) c + off);
    return mk<R>((R) v.x, (R) v.y);
}

// Raw buffer loads over a complex64 vector c[0 .. dim): a 16-byte load that
// the compiler keeps whole, with the hardware range check behind it.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t coef_rsrc(const float2* c, uint32_t dim) {
    return __builtin_amdgcn_make_buffer_rsrc((void*) c, (short) 0, (int) (dim * 8u), 0x00020000);
}
__device__ __forceinline__ float4 ld16(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const u32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
}
__device__ __forceinline__ float2 ld8(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const u32x2_t v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}

// acc += (G * t) for complex G (uniform) and t.
template <typename R> __device__ __forceinline__ void cmac(cplx<R>& acc, const cplx<R>& G, const cplx<R>& t) {
    acc.x += G.x * t.x;
    acc.y += G.x * t.y;
    acc.x -= G.y * t.y;
    acc.y += G.y * t.x;
}

// Diagonal coefficient E(n) = sum_k n_k (d1_k + n_k d2_k) + sum_{k<s} n_k n_s d3_ks
// in the compute type T, from a block whose d1, d2, d3 sections start at
// d (complex T entries, the layout of cf_off_d1 onwards).
template <int M, typename T, typename ZT>
__device__ __forceinline__ ZT diag_energy(const int* n, const ZT* __restrict__ d) {
    ZT e = mk<T>(0, 0);
#pragma unroll
    for (int k = 0; k < M; k++) {
        const T rk = (T) n[k];
        const ZT d1 = d[k], d2 = d[M + k];
        e.x += rk * (d1.x + rk * d2.x);
        e.y += rk * (d1.y + rk * d2.y);
#pragma unroll
        for (int s = k + 1; s < M; s++) {
            const ZT d3 = d[2 * M + k * M + s];
            const T ns = rk * (T) n[s];
            e.x += ns * d3.x;
            e.y += ns * d3.y;
        }
    }
    return e;
}

// Row i of H.C.  With HALF, the row's share of the Rayleigh quotient
// <c, H c> instead: H is Hermitian and the jump types come in inverse pairs
// (remove R, add A) / (remove A, add R), so sum_i conj(c_i) (H c)_i =
// sum_i E_i |c_i|^2 + 2 Re sum_i conj(c_i) sum_{T in half} H(i, T(i)) c_T(i),
// keeping of each pair the type whose sorted addition list comes first
// (l < k for k -> l, q < k for the double jumps: this keeps the jumps into
// orbitals 0 and 1, whose targets pair up for 16-byte loads).  The real
// result is returned in .x.
//
// The row value is returned in FP64.  In the complex64 form (R = float) the
// gathers, products and off-diagonal sums are FP32; the single-jump
// coefficients F_kl(n) and the diagonal term E(n) c_i are formed in FP64 from
// the block's FP64 sections (F rounded to FP32 once).
template <int M, typename R, int IDX, typename Zin, bool HALF = false>
__device__ __forceinline__ double2 hc_row_fused(uint32_t i, const RowCtx& ctx, const Zin* __restrict__ c,
                                                const cplx<R>* __restrict__ cf, const uint32_t* __restrict__ Wl,
                                                int ldw, const R* __restrict__ sql, cplx<R>& ci_out) {
    using Z = cplx<R>;
    constexpr bool F32 = std::is_same<R, float>::value;
    Cfg<M, R, IDX, coef_shift<Zin>()> x;
    x.template load_impl<true>(i, ctx, Wl, ldw, sql);
    const int* n = x.n;
    R rn[M];
#pragma unroll
    for (int k = 0; k < M; k++) rn[k] = (R) n[k];

    const Z ci = gather_b<R, Zin>(c, x.is);
    ci_out = ci;
    Z acc = mk<R>(0, 0);
    if constexpr (!F32 && !HALF) acc = cmul(diag_energy<M, R, Z>(n, cf + cf_off_d1(M)), ci);
    const double2* d64 = F32 ? (const double2*) (cf + cf_f64_off(M)) : nullptr;  // FP64 A, B, d1, d2, d3

    // Targets that differ only in whether one added particle lands in orbital
    // 0 or 1 have consecutive ranks (rank(n + e_1) = rank(n + e_0) + 1 when
    // the rest agrees): in the complex64 form one 16-byte load fetches both.
    // Such pairs exist when the removed orbitals are >= 2.
    constexpr bool ADJ = std::is_same<Zin, float2>::value;
    const __amdgpu_buffer_rsrc_t rs = coef_rsrc((const float2*) c, ctx.dim);  // used only when ADJ
    auto pair_at = [&](uint32_t off, Z& lo, Z& hi) {
        const float4 v = ld16(rs, off);
        lo = mk<R>(v.x, v.y);
        hi = mk<R>(v.z, v.w);
    };

    // single jumps k -> l: F_kl = A'_kl + sum_{m < M-1} n_m B'_klm (the last
    // occupation folded in through sum_m n_m = N when the block is built)
#pragma unroll
    for (int k = 0; k < M; k++) {
        if (n[k] >= 1) {
            Z c_to0{}, c_to1{};
            if (ADJ && k >= 2) pair_at(x.single(k, 0), c_to0, c_to1);
#pragma unroll
            for (int l = 0; l < M; l++) {
                if (l == k || (HALF && l > k)) continue;
                Z cj;
                if (ADJ && k >= 2 && l <= 1) cj = l == 0 ? c_to0 : c_to1;
                else cj = gather_b<R, Zin>(c, x.single(k, l));
                Z F;
                if constexpr (F32) {
                    // F_kl(n) in FP64, rounded once: in FP32 the folded sum
                    // cancels terms of size N |B| and its rounding (~1e-7 of
                    // F, worse where F is small) moves the energy of a long
                    // real-time run by ~1e-6 relative
                    double2 Fd = d64[cf_off_A(M) + k * M + l];
#pragma unroll
                    for (int m = 0; m < M - 1; m++) {
                        const double2 B = d64[cf_off_B(M) + (k * M + l) * M + m];
                        Fd.x += (double) n[m] * B.x;
                        Fd.y += (double) n[m] * B.y;
                    }
                    F = mk<R>((R) Fd.x, (R) Fd.y);
                } else {
                    F = cf[cf_off_A(M) + k * M + l];
#pragma unroll
                    for (int m = 0; m < M - 1; m++) {
                        const Z B = cf[cf_off_B(M) + (k * M + l) * M + m];
                        F.x += rn[m] * B.x;
                        F.y += rn[m] * B.y;
                    }
                }
                cmac<R>(acc, F, cscale(cj, x.s0[k] * x.s1[l]));
            }
        }
    }
    // creation factors of the two added particles, shared by every double
    // jump into (q, l): sqrt((n_q+1)(n_l+1)), or sqrt((n_q+1)(n_q+2)) at q = l
    R P[M][M];
#pragma unroll
    for (int q = 0; q < M; q++)
#pragma unroll
        for (int l = q; l < M; l++) P[q][l] = x.s1[q] * (q == l ? x.s2[q] : x.s1[l]);

    // two particles out of the same orbital k, summed per k and scaled once
    // by the annihilation factor sqrt(n_k (n_k - 1))
#pragma unroll
    for (int k = 0; k < M; k++) {
        if (n[k] >= 2) {
            Z g = mk<R>(0, 0);
            // with k >= 2: (0,0),(0,1) share a load, and (0,l),(1,l) for l >= 2
            Z c01{}, c1l[M];
#pragma unroll
            for (int q = 0; q < M; q++) {
                if (q == k || (HALF && q > k)) continue;
#pragma unroll
                for (int l = q; l < M; l++) {
                    if (l == k) continue;
                    const Z G = cf[cf_off_ot(M) + (k * M + q) * M + l];
                    Z cj;
                    if (ADJ && k >= 2 && q == 0 && l == 0) pair_at(x.same(k, 0, 0), cj, c01);
                    else if (ADJ && k >= 2 && q == 0 && l == 1) cj = c01;
                    else if (ADJ && k >= 2 && q == 0) pair_at(x.same(k, 0, l), cj, c1l[l]);
                    else if (ADJ && k >= 2 && q == 1 && l >= 2) cj = c1l[l];
                    else cj = gather_b<R, Zin>(c, x.same(k, q, l));
                    cmac<R>(g, G, cscale(cj, P[q][l]));
                }
            }
            const R a = x.sm[k] * x.s0[k];
            acc.x += a * g.x;
            acc.y += a * g.y;
        }
    }
    // one particle out of each of k < s, scaled once by sqrt(n_k n_s)
#pragma unroll
    for (int k = 0; k < M; k++) {
#pragma unroll
        for (int s = k + 1; s < M; s++) {
            if (n[k] >= 1 && n[s] >= 1) {
                Z g = mk<R>(0, 0);
                // with 2 <= k < s: the same sharing as above
                Z c01{}, c1l[M];
#pragma unroll
                for (int q = 0; q < M; q++) {
                    if (q == k || q == s || (HALF && q > k)) continue;
#pragma unroll
                    for (int l = q; l < M; l++) {
                        if (l == k || l == s) continue;
                        const Z G = cf[cf_off_tt(M) + ((k * M + s) * M + q) * M + l];
                        Z cj;
                        if (ADJ && k >= 2 && q == 0 && l == 0) pair_at(x.diff(k, s, 0, 0), cj, c01);
                        else if (ADJ && k >= 2 && q == 0 && l == 1) cj = c01;
                        else if (ADJ && k >= 2 && q == 0) pair_at(x.diff(k, s, 0, l), cj, c1l[l]);
                        else if (ADJ && k >= 2 && q == 1 && l >= 2) cj = c1l[l];
                        else cj = gather_b<R, Zin>(c, x.diff(k, s, q, l));
                        cmac<R>(g, G, cscale(cj, P[q][l]));
                    }
                }
                const R a = x.s0[k] * x.s0[s];
                acc.x += a * g.x;
                acc.y += a * g.y;
            }
        }
    }
    if constexpr (F32) {
        const double2 e = diag_energy<M, double, double2>(n, d64 + cf_off_d1(M));
        const double2 cd = make_double2(ci.x, ci.y);
        const double2 off = make_double2(acc.x, acc.y);
        if constexpr (HALF) {
            // E_i |c_i|^2 + 2 Re(conj(c_i) acc)
            const double n2 = cd.x * cd.x + cd.y * cd.y;
            return make_double2(e.x * n2 + 2 * (cd.x * off.x + cd.y * off.y), 0);
        }
        return cmul(e, cd) + off;
    } else {
        if constexpr (HALF) {
            const Z e = diag_energy<M, R, Z>(n, cf + cf_off_d1(M));
            const R n2 = ci.x * ci.x + ci.y * ci.y;
            return make_double2(e.x * n2 + 2 * (ci.x * acc.x + ci.y * acc.y), 0);
        }
        return acc;
    }
}

// --------------------------------------------------------------- epilogues

// Epilogues receive the row value and the row's own coefficient c_i (as
// gathered by the row), and are called by every thread of the block.
// An epilogue may load what it needs before the row is formed
// (prefetch(i, valid), handed back as the last argument).
struct NoPre {};
template <typename Zout> struct EpiPlain {
    Zout* out;
    __device__ __forceinline__ NoPre prefetch(uint32_t, bool) const { return {}; }
    __device__ __forceinline__ void operator()(uint32_t i, bool valid, double2 v, double2 ci, double* lds,
                                               NoPre = {}) const {
        if (valid) out[i] = from_d2<Zout>(v);
    }
};

// LDS bytes in front of the W table of the fused kernels: per-row partial
// sums (4 rows of 16 lanes per wave) of up to 2 kMaxKrylov + 4 slots.
constexpr int kRedBytes = 8 * (kHcBlock / 16) * (2 * kMaxKrylov + 4);

// Per-row partial sums into LDS, one barrier for any number of slots: put()
// sums a value over each 16-lane row with DPP shifts (vector ALU only; the
// cross-lane shuffles of a whole-wave sum are LDS permutes whose latency
// chain ended every wave), finish() adds the block's rows in a fixed order
// and stores slot t of this block at part[t * ps].  Every thread of the
// block calls both.
struct SlotSums {
    double* red;
    int nr, row, lane;
    __device__ __forceinline__ SlotSums(double* lds)
        : red(lds), nr(blockDim.x >> 4), row(threadIdx.x >> 4), lane(threadIdx.x & 15) {}
    __device__ __forceinline__ void put(int slot, double v) const {
        v = row_sum(v);
        if (lane == 15) red[slot * nr + row] = v;
    }
    __device__ __forceinline__ void finish(int slot0, int nslot, double* part, size_t ps) const {
        __syncthreads();
        for (int t = threadIdx.x; t < nslot; t += blockDim.x) {
            double s = 0;
            for (int r = 0; r < nr; r++) s += red[(slot0 + t) * nr + r];
            part[(size_t) (slot0 + t) * ps] = s;
        }
    }
};

template <typename Zw> struct EpiLanczos {
    static constexpr bool kReuse = std::is_same<Zw, float2>::value;
    // complex64: w_{it-1} .. w_{it-kPre} at this lane's row are loaded before
    // the row is formed, so their HBM latency overlaps the gathers (loaded in
    // the epilogue, it stalled every wave at its end); older vectors are read
    // there.  kPre complex64 values are 2 kPre VGPRs, within the kernel's
    // occupancy step.  complex128: everything is read in the epilogue (an
    // early read slowed the FP64 step by 8 %).
    static constexpr int kPre = kReuse ? 4 : 1;
    struct Pre {
        Zw w[kPre];
    };
    LanczosHcArgs a;
    __device__ __forceinline__ Pre prefetch(uint32_t i, bool valid) const {
        Pre p;
        const int it = a.it;
        // the last iteration needs w_{it-1} only (the recurrence)
        const int lo = a.last ? it - 1 : 0;
#pragma unroll
        for (int j = 0; j < kPre; j++) {
            const int k = it - 1 - j;
            p.w[j] = Zw{};
            if (kReuse && valid && k >= 0 && k >= lo) p.w[j] = ((const Zw*) a.w[k])[i];
        }
        return p;
    }
    __device__ __forceinline__ void operator()(uint32_t i, bool valid, double2 row, double2 ci, double* lds) const {
        (*this)(i, valid, row, ci, lds, prefetch(i, valid));
    }
    __device__ __forceinline__ void operator()(uint32_t i, bool valid, double2 row, double2 ci, double* lds,
                                               const Pre& pre) const {
        const SilState* st = a.st;
        const int it = a.it;
        // w_k[i] for it - kPre <= k < it: from the prefetch (complex64)
        auto earlier = [&](int j) {
            return kReuse ? to_d2(pre.w[j]) : to_d2(((const Zw*) a.w[it - 1 - j])[i]);
        };
        // w_it[i] is the row's own coefficient (complex64) or read here
        // (complex128); w_{it-1}[i] is earlier(0)
        double2 o = make_double2(0, 0), wprev = make_double2(0, 0);
        if (valid) {
            o = cscale(row, st->s[it]);
            if (it > 0) {
                wprev = earlier(0);
                o = o - cscale(wprev, st->beta[it - 1] * st->s[it - 1]);
            }
            if (!a.last) a.out[i] = o;
        }
        // slot-major partials: slot t of block b at part[t * gridDim.x + b]
        const SlotSums ss(lds);
        const int k0 = a.last ? it : 0;
        auto put_dot = [&](int k, double2 wk) {
            double2 p = make_double2(0, 0);
            if (valid) p = cscale(cconjmul(wk, o), st->s[k]);
            ss.put(2 * k, p.x);
            ss.put(2 * k + 1, p.y);
        };
        // k = it
        put_dot(it, valid ? (kReuse ? ci : to_d2(((const Zw*) a.w[it])[i])) : make_double2(0, 0));
        if (!a.last) {
            // it - kPre <= k < it from the registers, older ones from memory
#pragma unroll
            for (int j = 0; j < kPre; j++)
                if (it - 1 - j >= 0) put_dot(it - 1 - j, !valid ? make_double2(0, 0) : j == 0 ? wprev : earlier(j));
            for (int k = 0; k < it - kPre; k++)
                put_dot(k, valid ? to_d2(((const Zw*) a.w[k])[i]) : make_double2(0, 0));
            double n2 = 0;  // |w_0|^2
            if (valid && it == 0) {
                const double2 w0 = kReuse ? ci : to_d2(((const Zw*) a.w[0])[i]);
                n2 = w0.x * w0.x + w0.y * w0.y;
            }
            ss.put(2 * (it + 1), n2);
            ss.put(2 * (it + 1) + 1, 0.0);
        }
        ss.finish(2 * k0, a.last ? 2 : 2 * (it + 1) + 2, a.part + blockIdx.x, gridDim.x);
    }
};

// Last Lanczos iteration with the half-row form: only <w_it, H w_it> (real)
// into slot 2 it; alpha is formed from it in the scalar update.
struct EpiRayleigh {
    LanczosHcArgs a;
    __device__ __forceinline__ NoPre prefetch(uint32_t, bool) const { return {}; }
    __device__ __forceinline__ void operator()(uint32_t i, bool valid, double2 row, double2 ci, double* lds,
                                               NoPre = {}) const {
        double v[1] = {valid ? row.x : 0.0};
        block_sum<1>(v, lds);
        if (threadIdx.x == 0) {
            const int it = a.it;
            a.part[(size_t) (2 * it) * gridDim.x + blockIdx.x] = v[0];
            a.part[(size_t) (2 * it + 1) * gridDim.x + blockIdx.x] = 0;
        }
    }
};

// ----------------------------------------------------------------- kernels

template <int M, typename R, int IDX, typename Zin, typename Epi, bool HALF = false>
__global__ __launch_bounds__(kHcBlock) void k_hc_fused(RowCtx ctx, const Zin* __restrict__ c,
                                                       const cplx<R>* __restrict__ cf,
                                                       const uint32_t* __restrict__ Wg, int ldw,
                                                       int stop_check, const SilState* st, Epi epi) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (stop_check && st->stop) return;
    double* red = (double*) smem;                         // partial sums (kRedBytes)
    uint32_t* Wl = (uint32_t*) (smem + kRedBytes);        // W table
    const int wsz = (M - 1) * ldw;
    R* sql = (R*) (smem + kRedBytes + ((wsz * 4 + 15) & ~15));  // sqrt LUT
    if constexpr (IDX == kIdxFly) {
        for (int t = threadIdx.x; t < wsz; t += blockDim.x) Wl[t] = Wg[t];
    }
    for (int t = threadIdx.x; t < ctx.npar + 3; t += blockDim.x) sql[t] = sqrt((R) t);
    __syncthreads();
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = b * kHcBlock + threadIdx.x;
    const bool valid = i < ctx.dim;
    const auto pre = epi.prefetch(i, valid);
    double2 v = make_double2(0, 0), ci = make_double2(0, 0);
    if (valid) {
        cplx<R> own;
        v = hc_row_fused<M, R, IDX, Zin, HALF>(i, ctx, c, cf, Wl, ldw, sql, own);
        if constexpr (std::is_same<Zin, float2>::value) ci = to_d2(own);
    }
    epi(i, valid, v, ci, red, pre);
}

static inline RowCtx row_ctx(const Space* S) {
    return make_row_ctx(S->dim, S->npar, S->d_ht, S->d_map, S->d_mapot, S->d_maptt, S->d_strideot, S->d_stridett,
                        S->d_occ8);
}

// Layout matches k_hc_fused: partial-sum scratch, the W table region
// (reserved in both index modes so the sqrt LUT offset is the same), sqrt LUT.
static inline size_t fused_smem(const Space* S, size_t rsize) {
    size_t wsz = (size_t) (S->norb - 1) * S->ldw * 4;
    return kRedBytes + ((wsz + 15) & ~(size_t) 15) + rsize * (S->npar + 3);
}

template <int M, typename R, typename Zin, typename Epi, bool HALF = false>
static inline int launch_fused(const Space* S, const Zin* c, const cplx<R>* cf, int index, int stop_check,
                        const SilState* st, Epi epi, hipStream_t stream) {
    const bool fly = index == kIndexOnTheFly;
    const size_t smem = fused_smem(S, sizeof(R));
    if (smem > 64 * 1024) {
        set_error("rank and square-root tables exceed the LDS budget of the fused H.C kernel for this (N, M)");
        return kInvalidArgument;
    }
    if ((uint64_t) S->dim * sizeof(Zin) >= (1ull << 32)) {
        set_error("coefficient vector above 4 GiB: the fused H.C kernel needs 32-bit byte offsets");
        return kInvalidArgument;
    }
    const unsigned nb = hc_grid_blocks(S->dim);
    RowCtx ctx = row_ctx(S);
    if (fly) {
        k_hc_fused<M, R, kIdxFly, Zin, Epi, HALF><<<nb, kHcBlock, smem, stream>>>(
            ctx, c, cf, S->d_W, S->ldw, stop_check, st, epi);
    } else {
        if constexpr (std::is_same<R, double>::value)
            k_hc_fused<M, R, kIdxTables, Zin, Epi, HALF><<<nb, kHcBlock, smem, stream>>>(
                ctx, c, cf, S->d_W, S->ldw, stop_check, st, epi);
    }
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

}  // namespace mctdhb
