// sil_small.hpp - the whole SIL step in one workgroup, for spaces small
// enough that the multi-kernel step is a chain of launches whose work is a
// few microseconds each (the ground states of the reference's own example
// problems: NC in the hundreds or thousands).
//
// The recursion is the one of lanczos.hip (reference
// multiconfig_lanczos.c:11-90 and coefficients_integration.c:59-160) with
// the same unnormalised vectors and the same scalar updates
// (sil_device.hpp); the block does every pass in turn with barriers in
// between: H w_it by rows (the fused row of hamiltonian_impl.hpp), the dot
// products as block reductions, the reorthogonalisation pass, the
// tridiagonal exponential on wave 0 and the recombination.  FP64 only.
#pragma once

#include "hamiltonian_impl.hpp"
#include "pack.hpp"
#include "sil_device.hpp"

namespace mctdhb {


constexpr int kSmallSlots = 2 * kMaxLanczos + 2;  // d_k (2 per k) and one norm

// tot[t] = block sum of acc[t] for t < nv (nv uniform): one pass of wave
// shuffles, then thread t adds the wave partials of slot t.
__device__ __forceinline__ void block_sum_slots(const double (&acc)[kSmallSlots], int nv, double* red, double* tot) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int t = 0; t < kSmallSlots; t++) {
        if (t < nv) {
            const double v = wave_sum(acc[t]);
            if (lane == 0) red[t * nw + wid] = v;
        }
    }
    __syncthreads();
    if ((int) threadIdx.x < nv) {
        double v = 0;
        for (int w = 0; w < nw; w++) v += red[threadIdx.x * nw + w];
        tot[threadIdx.x] = v;
    }
    __syncthreads();
}

// STAGE: the vector H acts on (and the occupation rows) are kept in LDS, so
// the row gathers read LDS; w_{it+1} is written there by the pass that
// forms it, while the pass over w_it has finished.
template <int M, int IDX, bool STAGE>
__device__ __forceinline__ void sil_small_block(unsigned char* smem, const RowCtx& ctx, const double2* __restrict__ cf,
                                                const uint32_t* __restrict__ Wg, int ldw, const SmallSilArgs& a) {
    const int wsz = (M - 1) * ldw;
    uint32_t* Wl = (uint32_t*) smem;
    double* sql = (double*) (smem + ((wsz * 4 + 15) & ~15));
    unsigned char* lds_l = smem + ((wsz * 4 + 15) & ~15) + ((sizeof(double) * (ctx.npar + 3) + 15) & ~15);
    SmallSilLds& L = *(SmallSilLds*) lds_l;
    double2* Ws = (double2*) (lds_l + sizeof(SmallSilLds));  // STAGE only: dim entries
    uint2* Os = (uint2*) (Ws + ctx.dim);                       // STAGE only: dim occupation rows
    if constexpr (IDX == kIdxFly)
        for (int t = threadIdx.x; t < wsz; t += blockDim.x) Wl[t] = Wg[t];
    for (int t = threadIdx.x; t < ctx.npar + 3; t += blockDim.x) sql[t] = sqrt((double) t);
    RowCtx rc = ctx;
    if constexpr (STAGE) {
        const double2* c0 = (const double2*) a.w[0];
        for (uint32_t t = threadIdx.x; t < ctx.dim; t += blockDim.x) Ws[t] = c0[t];
        if (ctx.occ8) {
            for (uint32_t t = threadIdx.x; t < ctx.dim; t += blockDim.x) Os[t] = ((const uint2*) ctx.occ8)[t];
            rc.occ8 = (const uint8_t*) Os;
        }
    }
    // the scalars live in LDS (the one-thread updates then make no global
    // round trips) and are copied to a.st once the eigensolve is done
    SilState* st = &L.state;
    for (uint32_t t = threadIdx.x; t < sizeof(SilState) / sizeof(double); t += blockDim.x) ((double*) st)[t] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        st->s[0] = 1.0;
        st->iter_done = a.lm;
#pragma unroll
        for (int k = 0; k < kMaxLanczos; k++) L.w[k] = (double2*) a.w[k];  // constant indices: no private copy
    }
    __syncthreads();
    const uint32_t n = ctx.dim, T = blockDim.x;
    double2* out = a.out;
    double acc[kSmallSlots];
    for (int it = 0; it < a.lm; it++) {
        const double2* wit = L.w[it];
        const double2* wpr = it > 0 ? L.w[it - 1] : nullptr;
        const double sit = st->s[it], bprev = it > 0 ? st->beta[it - 1] * st->s[it - 1] : 0.0;
        const bool last = it == a.lm - 1;
        const int k0 = last ? it : 0;
        // out = s_it H w_it - beta_{it-1} s_{it-1} w_{it-1}, with the sums of
        // conj(w_k) out (k0 <= k <= it) and, on the first pass, |w_0|^2
#pragma unroll
        for (int t = 0; t < kSmallSlots; t++) acc[t] = 0;
        for (uint32_t i = threadIdx.x; i < n; i += T) {
            double2 own;
            double2 o = cscale(
                hc_row_fused<M, double, IDX, double2, false>(i, rc, STAGE ? Ws : wit, cf, Wl, ldw, sql, own), sit);
            if (it > 0) o = o - cscale(wpr[i], bprev);
            out[i] = o;
#pragma unroll
            for (int k = 0; k < kMaxLanczos; k++) {
                if (k >= k0 && k <= it) {
                    const double2 p = cconjmul(L.w[k][i], o);
                    acc[2 * k] += p.x;
                    acc[2 * k + 1] += p.y;
                }
            }
            if (it == 0) acc[2] += wit[i].x * wit[i].x + wit[i].y * wit[i].y;
        }
        block_sum_slots(acc, 2 * (it + 1) + 1, L.red, L.tot);
        if (threadIdx.x == 0) {
            for (int k = k0; k <= it; k++) {  // d_k = s_k <w_k, out>
                L.tot[2 * k] *= st->s[k];
                L.tot[2 * k + 1] *= st->s[k];
            }
            sil_after_hc(st, L.tot, it, last ? 1 : 0);
        }
        __syncthreads();
        if (last) break;
        // w_{it+1} = out - sum_k cre_k w_k, with <w_k, w_{it+1}> and |w_{it+1}|^2
        double2* wn = L.w[it + 1];
        if ((int) threadIdx.x <= it) L.cre[threadIdx.x] = st->cre[threadIdx.x];
#pragma unroll
        for (int t = 0; t < kSmallSlots; t++) acc[t] = 0;
        __syncthreads();
        double nrm = 0;
        for (uint32_t i = threadIdx.x; i < n; i += T) {
            double2 w = out[i];
            for (int k = 0; k <= it; k++) w = w - cmul(L.cre[k], L.w[k][i]);
            wn[i] = w;
            if constexpr (STAGE) Ws[i] = w;
#pragma unroll
            for (int k = 0; k < kMaxLanczos; k++) {
                if (k <= it) {  // the second read of w_k[i] hits the cache
                    const double2 p = cconjmul(L.w[k][i], w);
                    acc[2 * k] += p.x;
                    acc[2 * k + 1] += p.y;
                }
            }
            nrm += w.x * w.x + w.y * w.y;
        }
#pragma unroll
        for (int t = 0; t < kSmallSlots; t++)  // slot 2 (it + 1), by constant indices
            if (t == 2 * (it + 1)) acc[t] = nrm;
        block_sum_slots(acc, 2 * (it + 1) + 1, L.red, L.tot);
        if (threadIdx.x == 0) sil_after_reorth(st, L.tot, it);
        __syncthreads();
        if (st->stop) break;
    }
    // exp(-i T prop_dt) e_1 on wave 0, then C' = sum_k y_k w_k
    if (threadIdx.x < 64) sil_tridiag_wave(st, a.dt_re, a.dt_im, L.z, kMaxLanczos + 2, a.imagtime);
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < sizeof(SilState) / sizeof(double); t += blockDim.x)
        ((double*) a.st)[t] = ((const double*) st)[t];
    if (st->eig_fail) return;
    const int nd = st->iter_done;
    double2* C = L.w[0];
#pragma unroll
    for (int t = 0; t < kSmallSlots; t++) acc[t] = 0;
    for (uint32_t i = threadIdx.x; i < n; i += T) {
        double2 v = cmul(st->y[0], C[i]);
        for (int k = 1; k < nd; k++) v = v + cmul(st->y[k], L.w[k][i]);
        C[i] = v;
        acc[0] += v.x * v.x + v.y * v.y;
    }
    if (!a.imagtime) return;
    block_sum_slots(acc, 1, L.red, L.tot);  // renormalise (coefficients_integration.c:147-158)
    const double sc = 1.0 / sqrt(L.tot[0]);
    for (uint32_t i = threadIdx.x; i < n; i += T) C[i] = cscale(C[i], sc);
}

template <int M, int IDX, bool STAGE>
__global__ __launch_bounds__(kSmallSilBlock) void k_sil_small(RowCtx ctx, const double2* __restrict__ cf,
                                                              const uint32_t* __restrict__ Wg, int ldw,
                                                              SmallSilArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    sil_small_block<M, IDX, STAGE>(smem, ctx, cf, Wg, ldw, a);
}

// Packed form: one workgroup per instance.
template <int M, int IDX, bool STAGE>
__global__ __launch_bounds__(kSmallSilBlock) void k_sil_small_pk(const PkSil* __restrict__ tab) {
    extern __shared__ __align__(16) unsigned char smem[];
    const PkSil& p = tab[blockIdx.x];
    sil_small_block<M, IDX, STAGE>(smem, p.ctx, p.cf, p.W, p.ldw, p.a);
}

template <int M>
int HcFused<M>::sil_small(const Space* S, const void* cf, const SmallSilArgs& a, int index, hipStream_t st) {
    if (!sil_small_fits(S, a.lm)) {
        set_error("space too large for the one-workgroup SIL step");
        return kInvalidArgument;
    }
    const bool stage = sil_small_stage(S);
    const size_t smem = sil_small_smem(S) + (stage ? sil_small_stage_bytes(S) : 0);
    const RowCtx ctx = row_ctx(S);
    if (index == kIndexOnTheFly) {
        if (stage)
            k_sil_small<M, kIdxFly, true><<<1, kSmallSilBlock, smem, st>>>(ctx, (const double2*) cf, S->d_W, S->ldw, a);
        else
            k_sil_small<M, kIdxFly, false><<<1, kSmallSilBlock, smem, st>>>(ctx, (const double2*) cf, S->d_W, S->ldw,
                                                                           a);
    } else {
        if (stage)
            k_sil_small<M, kIdxTables, true><<<1, kSmallSilBlock, smem, st>>>(ctx, (const double2*) cf, S->d_W,
                                                                             S->ldw, a);
        else
            k_sil_small<M, kIdxTables, false><<<1, kSmallSilBlock, smem, st>>>(ctx, (const double2*) cf, S->d_W,
                                                                              S->ldw, a);
    }
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

template <int M>
int HcFused<M>::sil_small_pk(const PkSil* tab, int count, int index, bool stage, size_t smem, hipStream_t st) {
    if (index == kIndexOnTheFly) {
        if (stage) k_sil_small_pk<M, kIdxFly, true><<<count, kSmallSilBlock, smem, st>>>(tab);
        else k_sil_small_pk<M, kIdxFly, false><<<count, kSmallSilBlock, smem, st>>>(tab);
    } else {
        if (stage) k_sil_small_pk<M, kIdxTables, true><<<count, kSmallSilBlock, smem, st>>>(tab);
        else k_sil_small_pk<M, kIdxTables, false><<<count, kSmallSilBlock, smem, st>>>(tab);
    }
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

}  // namespace mctdhb
