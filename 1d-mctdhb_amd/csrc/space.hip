// space.hip - Fock-space hashing and the reference-layout jump tables,
// built on the device.
//
// Replaces reference src/configurational/space.c:84-452 and the table part
// of get_multiconf_struct (src/assistant/types_definition.c:110-141).  The
// tables are bit-exact with the reference's:
//   hash_table[k + M*i]                    = n_k of configuration i
//   map[i + k*NC + l*M*NC]                 = rank(i - e_k + e_l), 0 if n_k = 0
//   mapot[strideot[i] + c*M^2 + l + q*M]   = rank(i - 2e_k + e_l + e_q),
//        c = #{h < k : n_h >= 2}, one chunk per k with n_k >= 2
//   maptt[stridett[i] + c*M^2 + l + q*M]   = rank(i - e_k - e_s + e_l + e_q),
//        one chunk per occupied pair k < s in lexicographic order
// Every entry is computed independently from the closed-form rank (fock.hpp),
// one configuration per thread; strides come from a device prefix scan.
#include "space.hpp"

#include <cstdlib>
#include <cstring>
#include <vector>

namespace mctdhb {

uint64_t binom_u64(uint64_t n, uint64_t k) {
    if (k > n) return 0;
    if (k > n - k) k = n - k;
    unsigned __int128 r = 1;
    for (uint64_t j = 1; j <= k; j++) {
        r = r * (n - k + j) / j;  // exact: r is C(n-k+j, j) after each step
        if (r > (unsigned __int128) UINT64_MAX) return UINT64_MAX;
    }
    return (uint64_t) r;
}

void fill_w_table(int npar, int norb, uint32_t* w) {
    const int ld = w_ld(npar);
    for (int m = 0; m < norb - 1; m++)
        for (int P = -kWOff; P <= npar + 2; P++) {
            uint64_t v = P < 0 ? 0 : binom_u64((uint64_t) P + m, (uint64_t) m + 1);
            w[m * ld + P + kWOff] = (uint32_t) v;  // modulo 2^32 beyond the valid range
        }
}

int space_dimension_exact(int npar, int norb, uint32_t* dim) {
    if (npar < 0 || norb < 1) {
        set_error("invalid number of particles or orbitals");
        return kInvalidArgument;
    }
    uint64_t d = binom_u64((uint64_t) npar + norb - 1, (uint64_t) norb - 1);
    if (d > UINT32_MAX) {
        set_error("space dimension overflows 32-bit integers");
        return kOverflow;
    }
    *dim = (uint32_t) d;
    return kOk;
}

// ------------------------------------------------------------------ kernels

__global__ void k_hash_table(int N, int M, uint32_t dim, const uint32_t* __restrict__ W, int ldw,
                             uint16_t* __restrict__ ht, uint32_t* __restrict__ cnt_ot,
                             uint32_t* __restrict__ cnt_tt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dim) return;
    uint32_t n[kMaxOrb];
    unrank_rt(i, N, M, dim, W, ldw, n);
    uint32_t two = 0, occ = 0;
    for (int k = 0; k < M; k++) {
        ht[(size_t) i * M + k] = (uint16_t) n[k];
        two += n[k] >= 2;
        occ += n[k] >= 1;
    }
    if (cnt_ot) cnt_ot[i] = two;
    if (cnt_tt) cnt_tt[i] = occ * (occ - 1) / 2;
}

static __device__ __forceinline__ void load_prefix(const uint16_t* __restrict__ ht, uint32_t i, int M, uint32_t* n,
                                                   uint32_t* P) {
    uint32_t acc = 0;
    for (int k = 0; k < M; k++) {
        n[k] = ht[(size_t) i * M + k];
        acc += n[k];
        P[k] = acc;
    }
}

__global__ void k_single_map(int M, uint32_t dim, const uint32_t* __restrict__ W, int ldw,
                             const uint16_t* __restrict__ ht, uint32_t* __restrict__ map) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dim) return;
    uint32_t n[kMaxOrb], P[kMaxOrb];
    load_prefix(ht, i, M, n, P);
    for (int k = 0; k < M; k++)
        for (int l = 0; l < M; l++) {
            uint32_t j = 0;
            if (n[k] >= 1) j = jump_rank_rt(i, M, P, W, ldw, k, -1, l, -1);
            map[i + (size_t) k * dim + (size_t) l * M * dim] = j;
        }
}

__global__ void k_double_maps(int M, uint32_t dim, const uint32_t* __restrict__ W, int ldw,
                              const uint16_t* __restrict__ ht, const uint32_t* __restrict__ strideot,
                              const uint32_t* __restrict__ stridett, uint32_t* __restrict__ mapot,
                              uint32_t* __restrict__ maptt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dim) return;
    uint32_t n[kMaxOrb], P[kMaxOrb];
    load_prefix(ht, i, M, n, P);
    const uint32_t M2 = (uint32_t) M * M;
    uint32_t c = 0;
    for (int k = 0; k < M; k++) {
        if (n[k] < 2) continue;
        uint32_t* out = mapot + strideot[i] + c * M2;
        for (int q = 0; q < M; q++)
            for (int l = 0; l < M; l++) out[l + q * M] = jump_rank_rt(i, M, P, W, ldw, k, k, l, q);
        c++;
    }
    c = 0;
    for (int k = 0; k < M; k++) {
        if (n[k] < 1) continue;
        for (int s = k + 1; s < M; s++) {
            if (n[s] < 1) continue;
            uint32_t* out = maptt + stridett[i] + c * M2;
            for (int q = 0; q < M; q++)
                for (int l = 0; l < M; l++) out[l + q * M] = jump_rank_rt(i, M, P, W, ldw, k, s, l, q);
            c++;
        }
    }
}

// Exclusive scan of per-configuration chunk counts -> strides (count * M^2).
constexpr int kScanBlock = 1024;

__global__ void k_scan_block_sums(const uint32_t* __restrict__ cnt, uint32_t n, uint64_t* __restrict__ sums) {
    __shared__ uint64_t sh[kScanBlock / 64];
    const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    uint64_t v = i < n ? cnt[i] : 0;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t s = 0;
        for (int w = 0; w < kScanBlock / 64; w++) s += sh[w];
        sums[blockIdx.x] = s;
    }
}

// One block: exclusive scan of the block sums in place, total into sums[nb].
__global__ void k_scan_sums(uint64_t* __restrict__ sums, uint32_t nb) {
    if (threadIdx.x != 0) return;
    uint64_t acc = 0;
    for (uint32_t b = 0; b < nb; b++) {
        uint64_t v = sums[b];
        sums[b] = acc;
        acc += v;
    }
    sums[nb] = acc;
}

__global__ void k_scan_apply(const uint32_t* __restrict__ cnt, uint32_t n, const uint64_t* __restrict__ sums,
                             uint32_t m2, uint32_t* __restrict__ stride) {
    __shared__ uint64_t sh[kScanBlock];
    const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    sh[threadIdx.x] = i < n ? cnt[i] : 0;
    __syncthreads();
    // Hillis-Steele inclusive scan in LDS
    for (int off = 1; off < kScanBlock; off <<= 1) {
        uint64_t v = threadIdx.x >= (unsigned) off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    if (i < n) {
        uint64_t excl = sums[blockIdx.x] + sh[threadIdx.x] - cnt[i];
        stride[i] = (uint32_t) (excl * m2);
    }
}

__global__ void k_occ8(int M, uint32_t dim, const uint16_t* __restrict__ ht, uint2* __restrict__ occ8) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dim) return;
    uint32_t w[2] = {0, 0};
    for (int k = 0; k < M; k++) w[k >> 2] |= (uint32_t) ht[(size_t) i * M + k] << (8 * (k & 3));
    occ8[i] = make_uint2(w[0], w[1]);
}

// Pair slots.  Within a run (fixed P_1 .. P_{M-2}) the rank grows by one as
// P_0 = n_0 falls by one, i.e. as n_1 grows: run position = n_1.  A
// configuration at an even position starts a pair when the run goes on
// (n_0 >= 1) and is single otherwise.
__global__ void k_slot_flags(int M, uint32_t dim, const uint16_t* __restrict__ ht, uint32_t* __restrict__ pf,
                             uint32_t* __restrict__ sf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dim) return;
    const uint32_t n0 = ht[(size_t) i * M], n1 = ht[(size_t) i * M + 1];
    const bool even = (n1 & 1u) == 0;
    pf[i] = even && n0 >= 1;
    sf[i] = even && n0 == 0;
}

__global__ void k_slot_scatter(uint32_t dim, const uint32_t* __restrict__ pf, const uint32_t* __restrict__ sf,
                               const uint32_t* __restrict__ pp, const uint32_t* __restrict__ sp, uint32_t npairs,
                               uint32_t* __restrict__ slots) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dim) return;
    if (pf[i]) slots[pp[i]] = i;
    if (sf[i]) slots[npairs + sp[i]] = i;
}

// ------------------------------------------------------------ host driver

static int scan_strides(const uint32_t* d_cnt, uint32_t n, uint32_t m2, uint32_t* d_stride, uint64_t* total,
                        hipStream_t st) {
    const uint32_t nb = div_up(n, kScanBlock);
    uint64_t* d_sums = nullptr;
    MCTDHB_HIP_CHECK(hipMallocAsync((void**) &d_sums, sizeof(uint64_t) * (nb + 1), st));
    k_scan_block_sums<<<nb, kScanBlock, 0, st>>>(d_cnt, n, d_sums);
    k_scan_sums<<<1, 64, 0, st>>>(d_sums, nb);
    k_scan_apply<<<nb, kScanBlock, 0, st>>>(d_cnt, n, d_sums, m2, d_stride);
    MCTDHB_HIP_CHECK(hipMemcpyAsync(total, d_sums + nb, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    MCTDHB_HIP_CHECK(hipStreamSynchronize(st));
    MCTDHB_HIP_CHECK(hipFreeAsync(d_sums, st));
    return kOk;
}

// The temporaries are stream-ordered allocations on st: freeing them waits
// for no other stream's work (hipFree would synchronise the device).
static int build_pair_slots(Space* S, hipStream_t st) {
    const uint32_t dim = S->dim;
    uint32_t *pf = nullptr, *sf = nullptr, *pp = nullptr, *sp = nullptr;
    const size_t b = sizeof(uint32_t) * dim;
    if (hipMallocAsync((void**) &pf, b, st) != hipSuccess || hipMallocAsync((void**) &sf, b, st) != hipSuccess ||
        hipMallocAsync((void**) &pp, b, st) != hipSuccess || hipMallocAsync((void**) &sp, b, st) != hipSuccess) {
        if (pf) hipFreeAsync(pf, st);
        if (sf) hipFreeAsync(sf, st);
        if (pp) hipFreeAsync(pp, st);
        hipStreamSynchronize(st);
        set_error("device allocation failed (pair slots)");
        return kOutOfMemory;
    }
    const int tb = 256;
    k_slot_flags<<<div_up(dim, tb), tb, 0, st>>>(S->norb, dim, S->d_ht, pf, sf);
    uint64_t npairs = 0, nsingles = 0;
    int rc = scan_strides(pf, dim, 1, pp, &npairs, st);
    if (!rc) rc = scan_strides(sf, dim, 1, sp, &nsingles, st);
    if (!rc && 2 * npairs + nsingles != dim) {
        set_error("internal error: pair slots do not cover the space");
        rc = kNumerical;
    }
    if (!rc && hipMalloc((void**) &S->d_slots, sizeof(uint32_t) * (npairs + nsingles + 1)) != hipSuccess) {
        set_error("device allocation failed (pair slots)");
        rc = kOutOfMemory;
    }
    if (!rc) {
        S->n_pairs = (uint32_t) npairs;
        S->n_singles = (uint32_t) nsingles;
        k_slot_scatter<<<div_up(dim, tb), tb, 0, st>>>(dim, pf, sf, pp, sp, S->n_pairs, S->d_slots);
        if (hipGetLastError() != hipSuccess) {
            set_error("pair slot scatter failed to launch");
            rc = kDeviceError;
        }
    }
    hipFreeAsync(pf, st);
    hipFreeAsync(sf, st);
    hipFreeAsync(pp, st);
    hipFreeAsync(sp, st);
    hipStreamSynchronize(st);
    return rc;
}

int space_create(int npar, int norb, int flags, Space** out) {
    *out = nullptr;
    if (npar < 1 || norb < 1) {
        set_error("invalid number of particles or orbitals (both must be >= 1)");
        return kInvalidArgument;
    }
    if (norb > kMaxOrb) {
        set_error("exceeded max number of orbitals 30");
        return kInvalidArgument;
    }
    if (npar > kMaxPar) {
        set_error("exceeded max number of particles 10000");
        return kInvalidArgument;
    }
    uint32_t dim;
    int rc = space_dimension_exact(npar, norb, &dim);
    if (rc) return rc;
    if ((uint64_t) dim * norb > UINT32_MAX) {
        set_error("the size of the hashing table cannot be indexed by 32-bit integers");
        return kOverflow;
    }
    const uint64_t m2 = (uint64_t) norb * norb;
    if ((flags & kSpaceTables) && (uint64_t) dim > UINT32_MAX / norb / norb) {
        set_error("the size of 1-body jump-mappings cannot be indexed by 32-bit integers");
        return kOverflow;
    }

    Space* S = new Space();
    S->npar = npar;
    S->norb = norb;
    S->dim = dim;
    S->ldw = w_ld(npar);
    hipStream_t st = 0;
    MCTDHB_HIP_CHECK(hipGetDevice(&S->device));

    std::vector<uint32_t> w((size_t) std::max(norb - 1, 1) * S->ldw, 0);
    fill_w_table(npar, norb, w.data());
    std::vector<uint32_t> sub((size_t) (npar + 1) * (norb + 1));
    for (int n = 0; n <= npar; n++) {
        sub[n] = 0;
        if (norb >= 1) sub[n + (npar + 1)] = 1;
        for (int m = 2; m <= norb; m++) sub[n + (size_t) (npar + 1) * m] = (uint32_t) binom_u64(n + m - 1, m - 1);
    }
    S->w_host = w;
    S->subsdim_host = sub;

#define ALLOC(ptr, bytes)                                                   \
    do {                                                                    \
        if (hipMalloc((void**) &(ptr), (bytes)) != hipSuccess) {            \
            set_error("device allocation failed (" #ptr ")");               \
            space_destroy(S);                                               \
            return kOutOfMemory;                                            \
        }                                                                   \
    } while (0)

    ALLOC(S->d_W, sizeof(uint32_t) * w.size());
    ALLOC(S->d_subsdim, sizeof(uint32_t) * sub.size());
    ALLOC(S->d_ht, sizeof(uint16_t) * (size_t) dim * norb);
    MCTDHB_HIP_CHECK(hipMemcpy(S->d_W, w.data(), sizeof(uint32_t) * w.size(), hipMemcpyHostToDevice));
    MCTDHB_HIP_CHECK(hipMemcpy(S->d_subsdim, sub.data(), sizeof(uint32_t) * sub.size(), hipMemcpyHostToDevice));

    const bool tables = (flags & kSpaceTables) != 0;
    uint32_t *d_cnt_ot = nullptr, *d_cnt_tt = nullptr;
    if (tables) {
        ALLOC(d_cnt_ot, sizeof(uint32_t) * dim);
        ALLOC(d_cnt_tt, sizeof(uint32_t) * dim);
    }
    const int tb = 256;
    k_hash_table<<<div_up(dim, tb), tb, 0, st>>>(npar, norb, dim, S->d_W, S->ldw, S->d_ht, d_cnt_ot, d_cnt_tt);
    MCTDHB_HIP_CHECK(hipGetLastError());
    if (tables) {
        ALLOC(S->d_map, sizeof(uint32_t) * (size_t) dim * m2);
        ALLOC(S->d_strideot, sizeof(uint32_t) * dim);
        ALLOC(S->d_stridett, sizeof(uint32_t) * dim);
        k_single_map<<<div_up(dim, tb), tb, 0, st>>>(norb, dim, S->d_W, S->ldw, S->d_ht, S->d_map);
        MCTDHB_HIP_CHECK(hipGetLastError());
        uint64_t n_ot = 0, n_tt = 0;
        rc = scan_strides(d_cnt_ot, dim, (uint32_t) m2, S->d_strideot, &n_ot, st);
        if (!rc) rc = scan_strides(d_cnt_tt, dim, (uint32_t) m2, S->d_stridett, &n_tt, st);
        hipFree(d_cnt_ot);
        hipFree(d_cnt_tt);
        if (rc) {
            space_destroy(S);
            return rc;
        }
        if (n_tt > UINT32_MAX / m2) {
            set_error("size of the 2-body jump-mappings cannot be indexed by 32-bit integers");
            space_destroy(S);
            return kOverflow;
        }
        if (n_ot > UINT32_MAX / m2) {
            set_error("size of 2-body(same orbital) jump-mappings cannot be indexed by 32-bit integers");
            space_destroy(S);
            return kOverflow;
        }
        S->mapot_len = n_ot * m2;
        S->maptt_len = n_tt * m2;
        ALLOC(S->d_mapot, sizeof(uint32_t) * (S->mapot_len + 1));
        ALLOC(S->d_maptt, sizeof(uint32_t) * (S->maptt_len + 1));
        k_double_maps<<<div_up(dim, tb), tb, 0, st>>>(norb, dim, S->d_W, S->ldw, S->d_ht, S->d_strideot,
                                                      S->d_stridett, S->d_mapot, S->d_maptt);
        MCTDHB_HIP_CHECK(hipGetLastError());
        S->tables = true;
    }
    if (norb <= 8 && npar <= 255) {
        ALLOC(S->d_occ8, (size_t) 8 * dim);
        k_occ8<<<div_up(dim, tb), tb, 0, st>>>(norb, dim, S->d_ht, (uint2*) S->d_occ8);
        MCTDHB_HIP_CHECK(hipGetLastError());
    }
    if (norb >= 2) {
        rc = build_pair_slots(S, st);
        if (rc) {
            space_destroy(S);
            return rc;
        }
    }
#undef ALLOC
    // the build ran on stream st alone: other streams' work is not waited for
    MCTDHB_HIP_CHECK(hipStreamSynchronize(st));
    *out = S;
    return kOk;
}

void space_destroy(Space* S) {
    if (!S) return;
    hipFree(S->d_W);
    hipFree(S->d_subsdim);
    hipFree(S->d_ht);
    hipFree(S->d_map);
    hipFree(S->d_mapot);
    hipFree(S->d_maptt);
    hipFree(S->d_strideot);
    hipFree(S->d_stridett);
    hipFree(S->d_slots);
    hipFree(S->d_pad);
    hipFree(S->d_occ8);
    delete S;
}

// Host-side rank/unrank on the same closed form, for the scalar reference
// entry points (set_config / hash_index).
uint32_t rank_host(int npar, int norb, const uint16_t* conf) {
    uint64_t dim = binom_u64((uint64_t) npar + norb - 1, (uint64_t) norb - 1);
    uint64_t s = 0, P = 0;
    for (int m = 0; m < norb - 1; m++) {
        P += conf[m];
        s += binom_u64(P + m, (uint64_t) m + 1);
    }
    return (uint32_t) (dim - 1 - s);
}

void unrank_host(uint32_t i, int npar, int norb, uint16_t* conf) {
    uint64_t dim = binom_u64((uint64_t) npar + norb - 1, (uint64_t) norb - 1);
    uint64_t R = dim - 1 - i, hi = (uint64_t) npar;
    for (int m = norb - 2; m >= 0; m--) {
        uint64_t lo = 0, up = hi;
        while (lo < up) {
            uint64_t mid = (lo + up + 1) / 2;
            if (binom_u64(mid + m, (uint64_t) m + 1) <= R) lo = mid;
            else up = mid - 1;
        }
        R -= binom_u64(lo + m, (uint64_t) m + 1);
        conf[m + 1] = (uint16_t) (hi - lo);
        hi = lo;
    }
    conf[0] = (uint16_t) hi;
}

}  // namespace mctdhb
