// lanczos.hip - the Lanczos tridiagonalisation and the short-iterative-
// Lanczos (SIL) coefficient step on the device.
//
// Replaces reference src/linalg/multiconfig_lanczos.c:11-90 (lanczos) and
// src/integrator/coefficients_integration.c:59-160 (propagate_coef_sil).
//
// Iteration it of the reference computes
//   out = H v_it - beta_{it-1} v_{it-1};  alpha_it = <v_it, out>
//   out -= alpha_it v_it;  ortho_k = <v_k, out> (k <= it);  out -= sum ortho_k v_k
//   beta_it = |out|;  v_{it+1} = out / beta_it
// Here the H.C kernel writes out and, in the same pass, the partial sums of
// d_k = <v_k, out> for k <= it.  Because <v_k, out - alpha v_it> = d_k -
// alpha <v_k, v_it>, the reorthogonalisation coefficients follow from d_k and
// the Gram column <v_k, v_it> that the previous pass produced, so one more
// streaming pass both subtracts every projection and accumulates |w|^2 and
// the next Gram column.  Vectors stay unnormalised (v_k = s_k w_k).  All
// scalars live in a device SilState; the step is a fixed kernel sequence with
// no host synchronisation.
#include "density.hpp"
#include "hamiltonian.hpp"
#include "sil_api.hpp"
#include "sil_device.hpp"

namespace mctdhb {

constexpr int kStreamBlock = 256;
constexpr int kStreamGrid = 1024;

// -------------------------------------------------------------- scalars

__global__ void k_sil_init(SilState* st, int lm) {
    if (threadIdx.x == 0) sil_init_state(st, lm);
}


// Sum per-block partials (slot-major): tot[t] = sum_b part[t * nblocks + b],
// one block per slot.
__global__ __launch_bounds__(256) void k_reduce_slots(const double* __restrict__ part, int nblocks, int nslot,
                                                      double* __restrict__ tot, const SilState* st) {
    __shared__ double red[4];
    if (st && st->stop) return;
    const int t = blockIdx.x;
    double v[1] = {0};
    const double* p = part + (size_t) t * nblocks;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) v[0] += p[b];
    block_sum<1>(v, red);
    if (threadIdx.x == 0) tot[t] = v[0];
}



// k_reduce_slots over slots slot0 + blockIdx.x, then (in the last block to
// finish) the scalar update OP of iteration it: 0 = sil_after_hc, 1 =
// sil_after_reorth.  Blocks count themselves in st->arrive (a vector
// atomic); the last one resets it.
constexpr int kReduceBlock = 1024;

template <int OP>
__global__ __launch_bounds__(kReduceBlock) void k_reduce_scalars(const double* __restrict__ part, int nblocks,
                                                                 double* tot, SilState* st, int it, int slot0,
                                                                 int final_it) {
    __shared__ double red[kReduceBlock / 64];
    __shared__ int last;
    if (st->stop) return;  // set only by a previous kernel: uniform over the grid
    const int t = slot0 + blockIdx.x;
    const double* p = part + (size_t) t * nblocks;
    // four independent partial sums keep four loads in flight per thread
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int b = threadIdx.x;
    for (; b + 3 * kReduceBlock < nblocks; b += 4 * kReduceBlock) {
        a0 += p[b];
        a1 += p[b + kReduceBlock];
        a2 += p[b + 2 * kReduceBlock];
        a3 += p[b + 3 * kReduceBlock];
    }
    for (; b < nblocks; b += kReduceBlock) a0 += p[b];
    double v[1] = {(a0 + a1) + (a2 + a3)};
    block_sum<1>(v, red);
    if (threadIdx.x == 0) {
        tot[t] = v[0];
        __threadfence();
        last = atomicAdd(&st->arrive, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    __threadfence();
    st->arrive = 0;
    if (OP == 0) sil_after_hc(st, tot, it, final_it);
    else sil_after_reorth(st, tot, it);
}



// exp(-i T prop_dt) e_1 (reference coefficients_integration.c:104-145).  One wave.
__global__ __launch_bounds__(64) void k_sil_tridiag(SilState* st, double dt_re, double dt_im, int renorm) {
    __shared__ double z[kMaxKrylov * (kMaxKrylov + 1)];
    sil_tridiag_wave(st, dt_re, dt_im, z, kMaxKrylov + 1, renorm);
}

// Ground-state mode (reference lowest_state_lanczos, multiconfig_lanczos.c:
// 131-169): eigenvector of the smallest Ritz value (the first index found
// when several tie, with the reference's 1e15 sentinel).  One wave.
__global__ __launch_bounds__(64) void k_sil_ground(SilState* st) {
    __shared__ double z[kMaxKrylov * (kMaxKrylov + 1)];
    const int lane = threadIdx.x;
    const int n = st->iter_done;
    double d = lane < n ? st->alpha[lane].x : 0.0;
    double e = lane < n ? st->beta[lane] : 0.0;
    double* zr = z + lane * (kMaxKrylov + 1);
    if (tridiag_eig_wave(n, d, e, zr)) {
        if (lane == 0) st->eig_fail = 1;
        return;
    }
    double lowest = 1e15;
    int j = 0;
    for (int k = 0; k < n; k++) {
        const double dk = rlane(d, k);
        if (lowest > dk) {
            lowest = dk;
            j = k;
        }
    }
    if (lane == 0) st->eground = lowest;
    if (lane < n) st->y[lane] = make_double2(zr[j] * st->s[lane], 0.0);
}

// ---------------------------------------------------------- streaming

template <typename Zw>
__global__ __launch_bounds__(kStreamBlock) void k_to_complex64(const double2* __restrict__ a, Zw* __restrict__ b,
                                                               uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        b[i] = from_d2<Zw>(a[i]);
}

struct WPtrs {
    const void* w[kMaxKrylov];
};

// w_{it+1} = out - sum_{k<=it} cre_k w_k, with partials of <w_k, w_{it+1}>
// (slots 2k, 2k+1) and |w_{it+1}|^2 (slot 2(it+1)).
template <typename Zw, int NK>
__global__ __launch_bounds__(kStreamBlock) void k_sil_reorth(const double2* __restrict__ out, WPtrs wp,
                                                             Zw* __restrict__ wnew, uint32_t n,
                                                             const SilState* __restrict__ st,
                                                             double* __restrict__ part, int nslot) {
    __shared__ double red[2 * NK + 1][kStreamBlock / 64];
    if (st->stop) return;
    double2 cre[NK];
#pragma unroll
    for (int k = 0; k < NK; k++) cre[k] = st->cre[k];
    double acc[2 * NK + 1];
#pragma unroll
    for (int t = 0; t < 2 * NK + 1; t++) acc[t] = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double2 w = out[i];
        double2 wk[NK];
#pragma unroll
        for (int k = 0; k < NK; k++) {
            wk[k] = to_d2(((const Zw*) wp.w[k])[i]);
            w = w - cmul(cre[k], wk[k]);
        }
        const Zw ws = from_d2<Zw>(w);
        wnew[i] = ws;
        const double2 wr = to_d2(ws);  // the stored vector defines v_{it+1}
#pragma unroll
        for (int k = 0; k < NK; k++) {
            const double2 p = cconjmul(wk[k], wr);
            acc[2 * k] += p.x;
            acc[2 * k + 1] += p.y;
        }
        acc[2 * NK] += wr.x * wr.x + wr.y * wr.y;
    }
    block_sum<2 * NK + 1>(acc, &red[0][0]);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int t = 0; t < 2 * NK + 1; t++) part[(size_t) t * gridDim.x + blockIdx.x] = acc[t];
    }
}

// Krylov dimensions past the register-resident kernels: the update pass
// w_{it+1} = out - sum_k cre_k w_k with the partials of |w_{it+1}|^2 (slot
// 2 nk, nk = it + 1), then the Gram column in chunks of eight vectors.
template <typename Zw>
__global__ __launch_bounds__(kStreamBlock) void k_sil_reorth_any(const double2* __restrict__ out, WPtrs wp,
                                                                 Zw* __restrict__ wnew, uint32_t n,
                                                                 const SilState* __restrict__ st, int nk,
                                                                 double* __restrict__ part) {
    __shared__ double red[kStreamBlock / 64];
    if (st->stop) return;
    double acc[1] = {0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double2 w = out[i];
        for (int k = 0; k < nk; k++) w = w - cmul(st->cre[k], to_d2(((const Zw*) wp.w[k])[i]));
        const Zw ws = from_d2<Zw>(w);
        wnew[i] = ws;
        const double2 wr = to_d2(ws);
        acc[0] += wr.x * wr.x + wr.y * wr.y;
    }
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) part[(size_t) (2 * nk) * gridDim.x + blockIdx.x] = acc[0];
}

template <typename Zw>
__global__ __launch_bounds__(kStreamBlock) void k_multidot8(WPtrs wp, const Zw* __restrict__ w, uint32_t n,
                                                            const SilState* __restrict__ st, int k0, int nk,
                                                            double* __restrict__ part) {
    __shared__ double red[16 * (kStreamBlock / 64)];
    if (st->stop) return;
    double acc[16];
#pragma unroll
    for (int t = 0; t < 16; t++) acc[t] = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double2 wr = to_d2(w[i]);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (k0 + j < nk) {
                const double2 p = cconjmul(to_d2(((const Zw*) wp.w[k0 + j])[i]), wr);
                acc[2 * j] += p.x;
                acc[2 * j + 1] += p.y;
            }
        }
    }
    block_sum<16>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (k0 + j < nk) {
                part[(size_t) (2 * (k0 + j)) * gridDim.x + blockIdx.x] = acc[2 * j];
                part[(size_t) (2 * (k0 + j) + 1) * gridDim.x + blockIdx.x] = acc[2 * j + 1];
            }
        }
    }
}

// C' = sum_j y_j w_j for any Krylov dimension (coefficients read per vector).
template <typename Zw>
__global__ __launch_bounds__(kStreamBlock) void k_sil_recombine_any(double2* __restrict__ C, WPtrs wp, uint32_t n,
                                                                    const SilState* __restrict__ st,
                                                                    double* __restrict__ part) {
    __shared__ double red[kStreamBlock / 64];
    const int nd = st->iter_done;
    double acc[1] = {0};
    if (!st->eig_fail) {
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
            double2 v = cmul(st->y[0], C[i]);
            for (int k = 1; k < nd; k++) v = v + cmul(st->y[k], to_d2(((const Zw*) wp.w[k])[i]));
            C[i] = v;
            acc[0] += v.x * v.x + v.y * v.y;
        }
    }
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// C' = sum_j y_j w_j (w_0 is C itself in FP64), in place, plus |C'|^2
// partials.  NK is the Krylov dimension; vectors past a breakdown
// (k >= iter_done) are skipped.
template <typename Zw, int NK>
__global__ __launch_bounds__(kStreamBlock) void k_sil_recombine(double2* __restrict__ C, WPtrs wp, uint32_t n,
                                                                const SilState* __restrict__ st,
                                                                double* __restrict__ part) {
    __shared__ double red[kStreamBlock / 64];
    const int nd = st->iter_done;
    double2 y[NK];
#pragma unroll
    for (int k = 0; k < NK; k++) y[k] = st->y[k];
    double acc[1] = {0};
    if (!st->eig_fail) {
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
            double2 v = cmul(y[0], C[i]);
#pragma unroll
            for (int k = 1; k < NK; k++)
                if (k < nd) v = v + cmul(y[k], to_d2(((const Zw*) wp.w[k])[i]));
            C[i] = v;
            acc[0] += v.x * v.x + v.y * v.y;
        }
    }
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(kStreamBlock) void k_scale_by_inv_norm(double2* __restrict__ C, uint32_t n,
                                                                    const double* __restrict__ nrm2) {
    const double s = 1.0 / sqrt(nrm2[0]);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        C[i] = cscale(C[i], s);
}

// Normalised Lanczos vectors for the reference-shaped lanczos() entry point.
template <typename Zw>
__global__ __launch_bounds__(kStreamBlock) void k_export_vector(const Zw* __restrict__ w, double2* __restrict__ v,
                                                                uint32_t n, const SilState* __restrict__ st, int k) {
    const double s = st->s[k];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        v[i] = cscale(to_d2(w[i]), s);
}

// ------------------------------------------------------------ the driver

int sil_create(const Space* S, int lm, int prec, int index, Sil** out) {
    *out = nullptr;
    if (lm < 2 || lm > kMaxKrylov) {
        set_error("Lanczos iterations must be between 2 and 64");
        return kInvalidArgument;
    }
    if (prec != kPrecFP64 && prec != kPrecMixed) {
        set_error("unknown precision mode");
        return kInvalidArgument;
    }
    const bool generic = S->norb == 1 || S->norb > kFusedMaxOrb;
    if (generic && prec != kPrecFP64) {
        set_error("mixed precision SIL covers 2 <= norb <= 7");
        return kInvalidArgument;
    }
    if ((generic || (index == kIndexTables && prec == kPrecFP64)) && !S->tables) {
        set_error("this SIL form needs the jump tables; build the space with tables");
        return kNotBuilt;
    }
    if (prec == kPrecMixed) {
        const int rc = hc_tune_tiling(S, 0);
        if (rc) return rc;
    }
    Sil* L = new Sil();
    L->S = S;
    L->lm = lm;
    L->prec = prec;
    L->index = prec == kPrecMixed ? kIndexOnTheFly : index;
    L->generic = generic;
    const size_t n = S->dim;
    const size_t zw = prec == kPrecFP64 ? sizeof(double2) : sizeof(float2);
    const int M = S->norb;
    const size_t cfn = generic ? (size_t) M * M + (size_t) M * M * M * M : (size_t) cf_size(M);
    const unsigned nbh = std::max(hc_grid_blocks(S->dim), hc_lanczos_blocks(S, prec));
    L->part_cap = std::max<size_t>((size_t) nbh * lanczos_hc_slots(kMaxKrylov),
                                   (size_t) kStreamGrid * (2 * kMaxKrylov + 2));
    bool ok = hipMalloc((void**) &L->d_st, sizeof(SilState)) == hipSuccess;
    ok = ok && hipMalloc((void**) &L->cf64, sizeof(double2) * cfn) == hipSuccess;
    if (prec == kPrecMixed && !generic)
        ok = ok && hipMalloc((void**) &L->cf32, sizeof(float2) * cf32_size(M)) == hipSuccess;
    ok = ok && hipMalloc((void**) &L->out, sizeof(double2) * n) == hipSuccess;
    ok = ok && hipMalloc((void**) &L->part, sizeof(double) * L->part_cap) == hipSuccess;
    ok = ok && hipMalloc((void**) &L->tot, sizeof(double) * (2 * kMaxKrylov + 4)) == hipSuccess;
    auto padded = [&](void** base, void** vec, size_t esz) {
        if (hipMalloc(base, esz * (n + 2)) != hipSuccess) return false;
        if (hipMemset(*base, 0, esz * (n + 2)) != hipSuccess) return false;
        *vec = (char*) *base + esz;
        return true;
    };
    for (int k = 1; k < lm && ok; k++) ok = padded(&L->wbase[k], &L->w[k], zw);
    if (prec == kPrecMixed) ok = ok && padded(&L->w0fbase, &L->w0f, sizeof(float2));
    if (!ok) {
        sil_destroy(L);
        set_error("device allocation failed for the Lanczos workspace");
        return kOutOfMemory;
    }
    *out = L;
    return kOk;
}

void sil_destroy(Sil* L) {
    if (!L) return;
    hipFree(L->d_st);
    hipFree(L->cf64);
    hipFree(L->cf32);
    hipFree(L->out);
    hipFree(L->part);
    hipFree(L->tot);
    for (int k = 1; k < kMaxKrylov; k++) hipFree(L->wbase[k]);
    hipFree(L->w0fbase);
    delete L;
}

size_t sil_device_bytes(const Space* S, int lm, int prec) {
    const size_t n = S->dim;
    const size_t zw = prec == kPrecFP64 ? sizeof(double2) : sizeof(float2);
    size_t b = (size_t) (lm - 1) * zw * n + sizeof(double2) * n;
    if (prec == kPrecMixed) b += sizeof(float2) * n;
    return b;
}

template <typename Zw, int NK>
static void launch_reorth(Sil* L, int it, hipStream_t st) {
    WPtrs wp{};
    for (int k = 0; k <= it; k++) wp.w[k] = L->w[k];
    const int nslot = 2 * NK + 1;
    k_sil_reorth<Zw, NK><<<kStreamGrid, kStreamBlock, 0, st>>>(L->out, wp, (Zw*) L->w[it + 1], L->S->dim, L->d_st,
                                                               L->part, nslot);
    k_reduce_scalars<1><<<nslot, kReduceBlock, 0, st>>>(L->part, kStreamGrid, L->tot, L->d_st, it, 0, 0);
}

template <typename Zw>
static void recombine_dispatch(Sil* L, double2* C, hipStream_t st) {
    WPtrs wp{};
    for (int k = 0; k < L->lm; k++) wp.w[k] = L->w[k];
    switch (L->lm) {
#define CASE(K)                                                                                        \
    case K:                                                                                            \
        k_sil_recombine<Zw, K><<<kStreamGrid, kStreamBlock, 0, st>>>(C, wp, L->S->dim, L->d_st, L->part); \
        break;
        CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13)
        CASE(14) CASE(15)
#undef CASE
        default:
            k_sil_recombine_any<Zw><<<kStreamGrid, kStreamBlock, 0, st>>>(C, wp, L->S->dim, L->d_st, L->part);
            break;
    }
}

template <typename Zw>
static void reorth_dispatch(Sil* L, int it, hipStream_t st) {
    switch (it + 1) {
#define CASE(K) \
    case K: launch_reorth<Zw, K>(L, it, st); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12)
        CASE(13) CASE(14)
#undef CASE
        default: {
            const int nk = it + 1;
            WPtrs wp{};
            for (int k = 0; k < nk; k++) wp.w[k] = L->w[k];
            Zw* wnew = (Zw*) L->w[it + 1];
            k_sil_reorth_any<Zw><<<kStreamGrid, kStreamBlock, 0, st>>>(L->out, wp, wnew, L->S->dim, L->d_st, nk,
                                                                       L->part);
            for (int k0 = 0; k0 < nk; k0 += 8)
                k_multidot8<Zw><<<kStreamGrid, kStreamBlock, 0, st>>>(wp, wnew, L->S->dim, L->d_st, k0, nk, L->part);
            k_reduce_scalars<1><<<2 * nk + 1, kReduceBlock, 0, st>>>(L->part, kStreamGrid, L->tot, L->d_st, it, 0, 0);
            break;
        }
    }
}

// Enqueue the Lanczos recursion (lm H.C applications).  C is w_0 (FP64).
static int sil_run_lanczos(Sil* L, const double2* C, const double2* hob, const double2* hint, hipStream_t st) {
    const Space* S = L->S;
    const uint32_t n = S->dim;
    int rc;
    k_sil_init<<<1, 64, 0, st>>>(L->d_st, L->lm);
    if (L->generic) {
        const int M = S->norb;
        MCTDHB_HIP_CHECK(hipMemcpyAsync(L->cf64, hob, sizeof(double2) * M * M, hipMemcpyDeviceToDevice, st));
        MCTDHB_HIP_CHECK(hipMemcpyAsync(L->cf64 + M * M, hint, sizeof(double2) * M * M * M * M,
                                        hipMemcpyDeviceToDevice, st));
    } else {
        rc = hc_coefficients(S, hob, hint, L->cf64, L->cf32, st);
        if (rc) return rc;
    }
    if (L->prec == kPrecMixed) {
        k_to_complex64<float2><<<kStreamGrid, kStreamBlock, 0, st>>>(C, (float2*) L->w0f, n);
        L->w[0] = L->w0f;
    } else {
        L->w[0] = (void*) C;
    }
    if (L->prec == kPrecMixed && (rc = hc_tune_tiling(S, st))) return rc;
    const unsigned nbh = hc_lanczos_blocks(S, L->prec);
    for (int it = 0; it < L->lm; it++) {
        LanczosHcArgs a{};
        for (int k = 0; k <= it; k++) a.w[k] = L->w[k];
        a.it = it;
        a.st = L->d_st;
        a.out = L->out;
        a.part = L->part;
        a.last = it == L->lm - 1;
        rc = hc_lanczos(S, L->cf64, L->cf32, a, L->prec, L->index, st);
        if (rc) return rc;
        if (a.last)
            k_reduce_scalars<0><<<2, kReduceBlock, 0, st>>>(L->part, nbh, L->tot, L->d_st, it, 2 * it,
                                                             hc_lanczos_half(S, L->prec) && !L->generic ? 2 : 1);
        else
            k_reduce_scalars<0><<<lanczos_hc_slots(it), kReduceBlock, 0, st>>>(L->part, nbh, L->tot, L->d_st, it, 0, 0);
        if (it == L->lm - 1) break;
        if (L->prec == kPrecFP64) reorth_dispatch<double2>(L, it, st);
        else reorth_dispatch<float2>(L, it, st);
    }
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

int sil_step(Sil* L, double2* C, const double2* hob, const double2* hint, double dt_re, double dt_im, int imagtime,
             hipStream_t st) {
    if (L->prec == kPrecFP64 && !L->generic && sil_small_fits(L->S, L->lm) && L->small) {
        // small spaces: the whole step in one workgroup
        int rc = hc_coefficients(L->S, hob, hint, L->cf64, L->cf32, st);
        if (rc) return rc;
        SmallSilArgs a{};
        a.st = L->d_st;
        a.w[0] = C;
        for (int k = 1; k < L->lm; k++) a.w[k] = L->w[k];
        a.out = L->out;
        a.lm = L->lm;
        a.dt_re = dt_re;
        a.dt_im = dt_im;
        a.imagtime = imagtime;
        return hc_sil_small(L->S, L->cf64, a, L->index, st);
    }
    int rc = sil_run_lanczos(L, C, hob, hint, st);
    if (rc) return rc;
    k_sil_tridiag<<<1, 64, 0, st>>>(L->d_st, dt_re, dt_im, imagtime ? 1 : 0);
    if (L->prec == kPrecFP64) recombine_dispatch<double2>(L, C, st);
    else recombine_dispatch<float2>(L, C, st);
    if (imagtime) {
        k_reduce_slots<<<1, 256, 0, st>>>(L->part, kStreamGrid, 1, L->tot + 48, nullptr);
        k_scale_by_inv_norm<<<kStreamGrid, kStreamBlock, 0, st>>>(C, L->S->dim, L->tot + 48);
    }
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

int sil_ground(Sil* L, double2* C, const double2* hob, const double2* hint, hipStream_t st) {
    int rc = sil_run_lanczos(L, C, hob, hint, st);
    if (rc) return rc;
    k_sil_ground<<<1, 64, 0, st>>>(L->d_st);
    if (L->prec == kPrecFP64) recombine_dispatch<double2>(L, C, st);
    else recombine_dispatch<float2>(L, C, st);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

int sil_lanczos(Sil* L, const double2* C, const double2* hob, const double2* hint, double2* lvec, hipStream_t st) {
    int rc = sil_run_lanczos(L, C, hob, hint, st);
    if (rc) return rc;
    const uint32_t n = L->S->dim;
    MCTDHB_HIP_CHECK(hipMemcpyAsync(lvec, C, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
    for (int k = 1; k < L->lm; k++) {
        if (L->prec == kPrecFP64)
            k_export_vector<double2><<<kStreamGrid, kStreamBlock, 0, st>>>((const double2*) L->w[k],
                                                                           lvec + (size_t) k * n, n, L->d_st, k);
        else
            k_export_vector<float2><<<kStreamGrid, kStreamBlock, 0, st>>>((const float2*) L->w[k],
                                                                          lvec + (size_t) k * n, n, L->d_st, k);
    }
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

int sil_read_state(const Sil* L, SilState* host, hipStream_t st) {
    MCTDHB_HIP_CHECK(hipMemcpyAsync(host, L->d_st, sizeof(SilState), hipMemcpyDeviceToHost, st));
    MCTDHB_HIP_CHECK(hipStreamSynchronize(st));
    return kOk;
}

}  // namespace mctdhb
