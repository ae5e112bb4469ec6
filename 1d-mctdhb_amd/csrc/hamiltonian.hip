// hamiltonian.hip - many-body Hamiltonian action H.C on the device.
//
// Replaces reference src/configurational/hamiltonian.c:6-406
// (apply_hamiltonian).  One configuration per lane; the occupation row is
// read from the hash table, the coefficient block (M^2 one-body and M^4
// two-body derived factors) is read with uniform addresses, and the jump
// targets come either from the reference tables or from the closed-form rank
// deltas of fock.hpp (W table staged in LDS).  See hamiltonian.hpp for the
// rule grouping.
#include "hamiltonian_impl.hpp"
#include "hamiltonian_window.hpp"
#include "pack.hpp"
#include "rowctx.hpp"

#include <map>
#include <mutex>
#include <utility>

namespace mctdhb {

constexpr int kHcBlockG = 256;

unsigned hc_grid_blocks(uint32_t dim) { return div_up(dim, kHcBlockG); }

const float2* hc_padded_copy(const Space* S, const float2* c, hipStream_t st) {
    const size_t bytes = sizeof(float2) * ((size_t) S->dim + 2);
    if (!S->d_pad) {
        if (hipMalloc(&S->d_pad, bytes) != hipSuccess) return nullptr;
        if (hipMemset(S->d_pad, 0, bytes) != hipSuccess) return nullptr;
    }
    float2* p = (float2*) S->d_pad + 1;
    if (c && hipMemcpyAsync(p, c, sizeof(float2) * S->dim, hipMemcpyDeviceToDevice, st) != hipSuccess) return nullptr;
    return p;
}

// The autotuner's choice per (N, M), shared by every space of the process:
// timings vary from run to run, and two spaces of one shape must run the same
// kernels so that a solve gives the same bits alone and among others.
static std::mutex g_tiling_mu;
static std::map<std::pair<int, int>, int> g_tiling_by_shape;

void hc_forget_tiling(const Space* S) {
    std::lock_guard<std::mutex> lock(g_tiling_mu);
    g_tiling_by_shape.erase(std::make_pair((int) S->npar, (int) S->norb));
    S->hc_tiling = -1;
}

bool hc_window_possible(const Space* S) { return window_possible(S); }

int hc_tune_tiling(const Space* S, hipStream_t st) {
    if (S->hc_tiling >= 0) return kOk;
    const bool cand[3] = {true, pair_possible(S), hc_window_possible(S)};
    if (!cand[kTilingPair] && !cand[kTilingWindow]) {
        S->hc_tiling = kTilingLane;
        return kOk;
    }
    const std::pair<int, int> shape(S->npar, S->norb);
    {
        std::lock_guard<std::mutex> lock(g_tiling_mu);
        auto it = g_tiling_by_shape.find(shape);
        if (it != g_tiling_by_shape.end()) {
            S->hc_tiling = it->second;
            return kOk;
        }
    }
    const int M = S->norb;
    float2 *cf = nullptr, *out = nullptr;
    const size_t cfb = sizeof(float2) * cf32_size(M);
    if (!hc_padded_copy(S, nullptr, st)) {
        set_error("padded staging of the coefficient vector failed");
        return kDeviceError;
    }
    MCTDHB_HIP_CHECK(hipMallocAsync((void**) &cf, cfb, st));
    MCTDHB_HIP_CHECK(hipMallocAsync((void**) &out, sizeof(float2) * S->dim, st));
    MCTDHB_HIP_CHECK(hipMemsetAsync(cf, 0, cfb, st));
    const float2* in = (const float2*) S->d_pad + 1;
    hipEvent_t e0, e1;
    MCTDHB_HIP_CHECK(hipEventCreate(&e0));
    MCTDHB_HIP_CHECK(hipEventCreate(&e1));
    float best[3] = {1e30f, 1e30f, 1e30f};
    int rc = kOk;
    for (int rep = 0; rep < 3 && !rc; rep++) {
        for (int t = 0; t < 3 && !rc; t++) {
            if (!cand[t]) continue;
            hipEventRecord(e0, st);
            switch (M) {
                case 2: rc = HcFused<2>::plain_tiled(S, in, cf, out, t, st); break;
                case 3: rc = HcFused<3>::plain_tiled(S, in, cf, out, t, st); break;
                case 4: rc = HcFused<4>::plain_tiled(S, in, cf, out, t, st); break;
                case 5: rc = HcFused<5>::plain_tiled(S, in, cf, out, t, st); break;
                case 6: rc = HcFused<6>::plain_tiled(S, in, cf, out, t, st); break;
                case 7: rc = HcFused<7>::plain_tiled(S, in, cf, out, t, st); break;
                default: break;
            }
            hipEventRecord(e1, st);
            hipEventSynchronize(e1);
            float ms = 0;
            hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0 && ms < best[t]) best[t] = ms;  // the first round warms up
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFreeAsync(cf, st);
    hipFreeAsync(out, st);
    MCTDHB_HIP_CHECK(hipStreamSynchronize(st));
    if (rc) return rc;
    int pick = kTilingLane;
    for (int t = 1; t < 3; t++)
        if (cand[t] && best[t] < best[pick]) pick = t;
    S->hc_tiling = pick;
    std::lock_guard<std::mutex> lock(g_tiling_mu);
    g_tiling_by_shape[shape] = S->hc_tiling;
    return kOk;
}

unsigned hc_lanczos_blocks(const Space* S, int prec) {
    const bool fused = S->norb >= 2 && S->norb <= kFusedMaxOrb;
    if (fused && use_pair(S, prec)) return div_up((uint64_t) S->n_pairs + S->n_singles, kHcBlock);
    return hc_grid_blocks(S->dim);
}

// ------------------------------------------------------- coefficient block

// With sum_m n_m = N the last occupation is folded into the single-jump
// coefficients: A'_kl = A_kl + N B_kl,M-1 and B'_klm = B_klm - B_kl,M-1, so
// F_kl(n) = A'_kl + sum_{m < M-1} n_m B'_klm.
__device__ __forceinline__ void hc_coefficients_block(int M, int N, const double2* __restrict__ hob,
                                                      const double2* __restrict__ hint, double2* __restrict__ cf64,
                                                      float2* __restrict__ cf32) {
    const int t = threadIdx.x;
    const int M2 = M * M, M3 = M2 * M;
    auto G = [&](int a, int b, int c, int d) { return hint[a + b * M + c * M2 + d * M3]; };
    double2* d64 = cf32 ? (double2*) (cf32 + cf_f64_off(M)) : nullptr;
    auto put = [&](int idx, double2 v) {
        cf64[idx] = v;
        if (cf32) {
            cf32[idx] = make_float2((float) v.x, (float) v.y);
            // the one-body and diagonal sections once more in FP64
            if (idx < cf_f64_count(M)) d64[idx] = v;
        }
    };
    const int total = cf_size(M);
    for (int idx = t; idx < total; idx += blockDim.x) {
        double2 v = make_double2(0, 0);
        // B_klm = g_kkkl (m = k), g_klll (m = l), 2 g_kmml (otherwise)
        auto B = [&](int k, int l, int m) {
            if (m == k) return G(k, k, k, l);
            if (m == l) return G(k, l, l, l);
            return cscale(G(k, m, m, l), 2.0);
        };
        if (idx < cf_off_B(M)) {  // A'_kl = h_kl - g_kkkl + N B_kl,M-1
            int k = idx / M, l = idx % M;
            if (k != l) v = hob[k * M + l] - G(k, k, k, l) + cscale(B(k, l, M - 1), (double) N);
        } else if (idx < cf_off_d1(M)) {  // B'_klm = B_klm - B_kl,M-1
            int r = idx - cf_off_B(M);
            int k = r / M2, l = (r / M) % M, m = r % M;
            if (k != l && m < M - 1) v = B(k, l, m) - B(k, l, M - 1);
        } else if (idx < cf_off_d2(M)) {  // d1_k = h_kk - g_kkkk / 2
            int k = idx - cf_off_d1(M);
            v = hob[k * M + k] - cscale(G(k, k, k, k), 0.5);
        } else if (idx < cf_off_d3(M)) {  // d2_k = g_kkkk / 2
            int k = idx - cf_off_d2(M);
            v = cscale(G(k, k, k, k), 0.5);
        } else if (idx < cf_off_ot(M)) {  // d3_ks = 2 g_ksks (k < s)
            int r = idx - cf_off_d3(M);
            int k = r / M, s = r % M;
            if (k < s) v = cscale(G(k, s, k, s), 2.0);
        } else if (idx < cf_off_tt(M)) {  // Got[k][q][l]
            int r = idx - cf_off_ot(M);
            int k = r / M2, q = (r / M) % M, l = r % M;
            if (q != k && l != k && q <= l) v = (q == l) ? cscale(G(k, k, q, q), 0.5) : G(k, k, q, l);
        } else {  // Gtt[k][s][q][l]
            int r = idx - cf_off_tt(M);
            int k = r / M3, s = (r / M2) % M, q = (r / M) % M, l = r % M;
            if (k < s && q <= l && q != k && q != s && l != k && l != s)
                v = (q == l) ? G(k, s, q, q) : cscale(G(k, s, q, l), 2.0);
        }
        put(idx, v);
    }
}

__global__ void k_hc_coefficients(int M, int N, const double2* __restrict__ hob, const double2* __restrict__ hint,
                                  double2* __restrict__ cf64, float2* __restrict__ cf32) {
    hc_coefficients_block(M, N, hob, hint, cf64, cf32);
}

// Packed form (FP64 block only): one block per instance.
__global__ void k_hc_coefficients_pk(const PkCf* __restrict__ tab, int M) {
    const PkCf a = tab[blockIdx.x];
    hc_coefficients_block(M, a.npar, a.hob, a.hint, a.cf64, nullptr);
}

int pk_hc_coefficients(const PkCf* tab, int count, int M, hipStream_t st) {
    k_hc_coefficients_pk<<<count, 256, 0, st>>>(tab, M);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

int hc_coefficients(const Space* S, const double2* hob, const double2* hint, double2* cf64, float2* cf32,
                    hipStream_t st) {
    k_hc_coefficients<<<1, 256, 0, st>>>(S->norb, S->npar, hob, hint, cf64, cf32);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

// ------------------------------------------- reference-order generic row
// Any M (up to 30), tables required, FP64: the reference's 13 rule blocks in
// its order (hamiltonian.c:44-402).
__device__ double2 hc_row_generic(uint32_t i, int M, const RowCtx& ctx, const double2* __restrict__ c,
                                  const double2* __restrict__ hob, const double2* __restrict__ hint) {
    const uint32_t dim = ctx.dim, M2 = M * M, M3 = M2 * M;
    uint16_t occ[kMaxOrb];
    for (int k = 0; k < M; k++) occ[k] = ctx.ht[(size_t) i * M + k];
    auto G = [&](int a, int b, int cc, int d) { return hint[a + b * M + cc * M2 + d * M3]; };
    auto otc = [&](int k) {
        uint32_t cnt = 0;
        for (int h = 0; h < k; h++) cnt += occ[h] > 1;
        return cnt;
    };
    auto ttc = [&](int k, int s) {
        uint32_t cnt = 0;
        for (int h = 0; h < k; h++)
            for (int g = h + 1; g < M; g++) cnt += (occ[h] > 0 && occ[g] > 0);
        for (int g = k + 1; g < s; g++) cnt += occ[g] > 0;
        return cnt;
    };
    const uint32_t sot = ctx.sot ? ctx.sot[i] : 0, stt = ctx.stt ? ctx.stt[i] : 0;
    const double2 ci = c[i];
    double2 w = make_double2(0, 0), z = make_double2(0, 0);
    double bf;
    uint32_t j;
    for (int k = 0; k < M; k++) {
        if (occ[k] < 1) continue;
        w = w + cmul(cscale(hob[k * M + k], (double) occ[k]), ci);
        for (int l = 0; l < M; l++) {
            if (l == k) continue;
            bf = sqrt((double) occ[k] * (occ[l] + 1));
            j = ctx.map[i + k * dim + l * M * dim];
            w = w + cmul(cscale(hob[k * M + l], bf), c[j]);
        }
    }
    for (int k = 0; k < M; k++) {  // Rule 1
        bf = occ[k] * (occ[k] - 1);
        z = z + cscale(cmul(G(k, k, k, k), ci), bf);
    }
    for (int k = 0; k < M; k++) {  // Rule 2
        if (occ[k] < 1) continue;
        for (int s = k + 1; s < M; s++) {
            bf = occ[k] * occ[s];
            z = z + cmul(cscale(cscale(G(k, s, k, s), 4.0), bf), ci);
        }
    }
    for (int k = 0; k < M; k++) {  // Rule 3
        if (occ[k] < 2) continue;
        for (int q = 0; q < M; q++) {
            if (q == k) continue;
            bf = sqrt((double) (occ[k] - 1) * occ[k] * (occ[q] + 1) * (occ[q] + 2));
            j = ctx.mapot[sot + otc(k) * M2 + q + q * M];
            z = z + cscale(cmul(G(k, k, q, q), c[j]), bf);
        }
    }
    for (int k = 0; k < M; k++) {  // Rule 4
        if (occ[k] < 2) continue;
        for (int l = 0; l < M; l++) {
            if (l == k) continue;
            bf = (occ[k] - 1) * sqrt((double) occ[k] * (occ[l] + 1));
            j = ctx.map[i + k * dim + l * M * dim];
            z = z + cscale(cmul(cscale(G(k, k, k, l), 2.0), c[j]), bf);
        }
    }
    for (int k = 0; k < M; k++) {  // Rule 5
        if (occ[k] < 1) continue;
        for (int s = 0; s < M; s++) {
            if (s == k || occ[s] < 1) continue;
            bf = occ[s] * sqrt((double) occ[k] * (occ[s] + 1));
            j = ctx.map[i + k * dim + s * M * dim];
            z = z + cscale(cmul(cscale(G(k, s, s, s), 2.0), c[j]), bf);
        }
    }
    // Rule 6 (6.0 k<q<l, 6.1 q<k<l, 6.2 q<l<k)
    for (int pass = 0; pass < 3; pass++)
        for (int a = 0; a < M; a++)
            for (int b = a + 1; b < M; b++)
                for (int d = b + 1; d < M; d++) {
                    int k, q, l;
                    if (pass == 0) { k = a; q = b; l = d; }
                    else if (pass == 1) { q = a; k = b; l = d; }
                    else { q = a; l = b; k = d; }
                    if (occ[k] < 2) continue;
                    bf = sqrt((double) occ[k] * (occ[k] - 1) * (occ[q] + 1) * (occ[l] + 1));
                    j = ctx.mapot[sot + otc(k) * M2 + q + l * M];
                    z = z + cscale(cmul(cscale(G(k, k, q, l), 2.0), c[j]), bf);
                }
    // Rule 7 (7.0 q<k<s, 7.1 k<q<s, 7.2 k<s<q)
    for (int pass = 0; pass < 3; pass++)
        for (int a = 0; a < M; a++)
            for (int b = a + 1; b < M; b++)
                for (int d = b + 1; d < M; d++) {
                    int k, s, q;
                    if (pass == 0) { q = a; k = b; s = d; }
                    else if (pass == 1) { k = a; q = b; s = d; }
                    else { k = a; s = b; q = d; }
                    if (occ[k] < 1 || occ[s] < 1) continue;
                    bf = sqrt((double) occ[k] * occ[s] * (occ[q] + 1) * (occ[q] + 2));
                    j = ctx.maptt[stt + ttc(k, s) * M2 + q + q * M];
                    z = z + cscale(cmul(cscale(G(k, s, q, q), 2.0), c[j]), bf);
                }
    for (int s = 0; s < M; s++) {  // Rule 8
        if (occ[s] < 1) continue;
        for (int k = 0; k < M; k++) {
            if (occ[k] < 1 || k == s) continue;
            for (int l = 0; l < M; l++) {
                if (l == k || l == s) continue;
                bf = occ[s] * sqrt((double) occ[k] * (occ[l] + 1));
                j = ctx.map[i + k * dim + l * M * dim];
                z = z + cscale(cmul(cscale(G(k, s, s, l), 4.0), c[j]), bf);
            }
        }
    }
    for (int k = 0; k < M; k++) {  // Rule 9
        if (occ[k] < 1) continue;
        for (int s = k + 1; s < M; s++) {
            if (occ[s] < 1) continue;
            for (int q = 0; q < M; q++) {
                if (q == s || q == k) continue;
                for (int l = q + 1; l < M; l++) {
                    if (l == k || l == s) continue;
                    bf = sqrt((double) occ[k] * occ[s] * (occ[q] + 1) * (occ[l] + 1));
                    j = ctx.maptt[stt + ttc(k, s) * M2 + q + l * M];
                    z = z + cscale(cmul(cscale(G(k, s, q, l), 4.0), c[j]), bf);
                }
            }
        }
    }
    return w + cscale(z, 0.5);
}

template <typename Zout> struct EpiPlainG {
    Zout* out;
    __device__ __forceinline__ void operator()(uint32_t i, bool valid, double2 v, double2 ci, double* lds) const {
        if (valid) out[i] = v;
    }
};

template <typename Epi>
__global__ __launch_bounds__(kHcBlockG) void k_hc_generic(int M, RowCtx ctx, const double2* __restrict__ c,
                                                         const double2* __restrict__ hob,
                                                         const double2* __restrict__ hint, int stop_check,
                                                         const SilState* st, Epi epi) {
    // the Lanczos epilogue's slot partials (SlotSums: up to 2 kMaxKrylov + 4
    // slots of blockDim / 16 row sums)
    static_assert(kHcBlockG == kHcBlock, "the epilogue partials are sized for kHcBlock threads");
    __shared__ double red[kRedBytes / sizeof(double)];
    if (stop_check && st->stop) return;
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = b * kHcBlockG + threadIdx.x;
    const bool valid = i < ctx.dim;
    double2 v = make_double2(0, 0);
    if (valid) v = hc_row_generic(i, M, ctx, c, hob, hint);
    epi(i, valid, v, valid ? c[i] : make_double2(0, 0), red);
}

static RowCtx row_ctx_g(const Space* S) {
    return make_row_ctx(S->dim, S->npar, S->d_ht, S->d_map, S->d_mapot, S->d_maptt, S->d_strideot, S->d_stridett);
}

static int dispatch_fused_plain(const Space* S, const void* c, const void* cf, void* out, int prec, int index,
                                hipStream_t st) {
    switch (S->norb) {
        case 2: return HcFused<2>::plain(S, c, cf, out, prec, index, st);
        case 3: return HcFused<3>::plain(S, c, cf, out, prec, index, st);
        case 4: return HcFused<4>::plain(S, c, cf, out, prec, index, st);
        case 5: return HcFused<5>::plain(S, c, cf, out, prec, index, st);
        case 6: return HcFused<6>::plain(S, c, cf, out, prec, index, st);
        case 7: return HcFused<7>::plain(S, c, cf, out, prec, index, st);
        default: set_error("fused H.C kernels cover 2 <= norb <= 7"); return kInvalidArgument;
    }
}

static int dispatch_fused_lanczos(const Space* S, const void* cf, const LanczosHcArgs& a, int prec, int index,
                                  hipStream_t st) {
    switch (S->norb) {
        case 2: return HcFused<2>::lanczos(S, cf, a, prec, index, st);
        case 3: return HcFused<3>::lanczos(S, cf, a, prec, index, st);
        case 4: return HcFused<4>::lanczos(S, cf, a, prec, index, st);
        case 5: return HcFused<5>::lanczos(S, cf, a, prec, index, st);
        case 6: return HcFused<6>::lanczos(S, cf, a, prec, index, st);
        case 7: return HcFused<7>::lanczos(S, cf, a, prec, index, st);
        default: set_error("fused H.C kernels cover 2 <= norb <= 7"); return kInvalidArgument;
    }
}

static bool use_generic(const Space* S) { return S->norb == 1 || S->norb > kFusedMaxOrb; }

static int check_mode(const Space* S, int prec, int index) {
    if (prec != kPrecFP64 && prec != kPrecMixed) {
        set_error("unknown precision mode");
        return kInvalidArgument;
    }
    if (index != kIndexTables && index != kIndexOnTheFly) {
        set_error("unknown index mode");
        return kInvalidArgument;
    }
    if ((index == kIndexTables || use_generic(S)) && !S->tables) {
        set_error("this H.C form needs the jump tables; build the space with tables");
        return kNotBuilt;
    }
    if (use_generic(S) && prec != kPrecFP64) {
        set_error("mixed precision H.C covers 2 <= norb <= 7");
        return kInvalidArgument;
    }
    return kOk;
}

int hc_apply(const Space* S, const void* coef, const double2* hob, const double2* hint, void* hc, int prec,
             int index, hipStream_t st) {
    int rc = check_mode(S, prec, index);
    if (rc) return rc;
    if (use_generic(S)) {
        EpiPlainG<double2> epi{(double2*) hc};
        k_hc_generic<<<hc_grid_blocks(S->dim), kHcBlockG, 0, st>>>(S->norb, row_ctx_g(S), (const double2*) coef,
                                                                   hob, hint, 0, nullptr, epi);
        MCTDHB_HIP_CHECK(hipGetLastError());
        return kOk;
    }
    if (prec == kPrecMixed && (rc = hc_tune_tiling(S, st))) return rc;
    const int M = S->norb;
    double2* cf64 = nullptr;
    float2* cf32 = nullptr;
    MCTDHB_HIP_CHECK(hipMallocAsync((void**) &cf64, sizeof(double2) * cf_size(M), st));
    if (prec == kPrecMixed) MCTDHB_HIP_CHECK(hipMallocAsync((void**) &cf32, sizeof(float2) * cf32_size(M), st));
    rc = hc_coefficients(S, hob, hint, cf64, cf32, st);
    if (!rc) rc = dispatch_fused_plain(S, coef, prec == kPrecFP64 ? (const void*) cf64 : (const void*) cf32, hc,
                                       prec, index, st);
    hipFreeAsync(cf64, st);
    if (cf32) hipFreeAsync(cf32, st);
    return rc;
}

// ------------------------------------------- H.C without device allocation
// The coefficient block of hc_apply_noalloc lives in this array.  Host code
// never passes its name to a kernel or a copy: each device's copy of it is
// at the address hipGetSymbolAddress returns with that device current, kept
// in g_scratch_addr.
static __device__ __align__(256) unsigned char scratch[kHcScratchBytes];
constexpr int kScratchMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[kScratchMaxDevices] = {};

static int scratch_address(void** out) {
    int dev = -1;
    MCTDHB_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchMaxDevices) {
        set_error("device index outside the fixed-scratch table");
        return kDeviceError;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress of the H.C scratch failed: ") + hipGetErrorString(e));
            return kDeviceError;
        }
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return kOk;
}

int hc_scratch_setup() {
    void* p = nullptr;
    return scratch_address(&p);
}

static int dispatch_plain_tiled(const Space* S, const void* c, const void* cf, void* out, int tiling,
                                hipStream_t st) {
    switch (S->norb) {
        case 2: return HcFused<2>::plain_tiled(S, c, cf, out, tiling, st);
        case 3: return HcFused<3>::plain_tiled(S, c, cf, out, tiling, st);
        case 4: return HcFused<4>::plain_tiled(S, c, cf, out, tiling, st);
        case 5: return HcFused<5>::plain_tiled(S, c, cf, out, tiling, st);
        case 6: return HcFused<6>::plain_tiled(S, c, cf, out, tiling, st);
        case 7: return HcFused<7>::plain_tiled(S, c, cf, out, tiling, st);
        default: set_error("fused H.C kernels cover 2 <= norb <= 7"); return kInvalidArgument;
    }
}

int hc_apply_noalloc(const Space* S, const void* coef, const double2* hob, const double2* hint, void* hc, int prec,
                     int index, hipStream_t st) {
    int rc = check_mode(S, prec, index);
    if (rc) return rc;
    int dev = -1;
    MCTDHB_HIP_CHECK(hipGetDevice(&dev));
    if (dev != S->device) {
        set_error("the space lives on another device: make its device current first");
        return kInvalidArgument;
    }
    if (use_generic(S)) {  // reads hob and hint directly: no block
        EpiPlainG<double2> epi{(double2*) hc};
        k_hc_generic<<<hc_grid_blocks(S->dim), kHcBlockG, 0, st>>>(S->norb, row_ctx_g(S), (const double2*) coef,
                                                                   hob, hint, 0, nullptr, epi);
        MCTDHB_HIP_CHECK(hipGetLastError());
        return kOk;
    }
    const size_t need = hc_coef_bytes(S, prec);
    if (need > kHcScratchBytes) {
        set_error("coefficient block of " + std::to_string(need) + " bytes exceeds the fixed scratch of " +
                  std::to_string(kHcScratchBytes) + " bytes");
        return kInvalidArgument;
    }
    void* blk = nullptr;
    if ((rc = scratch_address(&blk))) return rc;
    // every entry of the block is written here, so nothing needs clearing
    if ((rc = hc_build_coef(S, hob, hint, blk, prec, st))) return rc;
    if (prec == kPrecFP64) return dispatch_fused_plain(S, coef, blk, hc, prec, index, st);
    const void* cf32 = (const double2*) blk + cf_size(S->norb);
    if (S->hc_tiling == 1 && S->d_pad) {  // staging already allocated: a device copy only
        const float2* cp = hc_padded_copy(S, (const float2*) coef, st);
        if (!cp) {
            set_error("padded staging of the coefficient vector failed");
            return kDeviceError;
        }
        return dispatch_plain_tiled(S, cp, cf32, hc, 1, st);
    }
    return dispatch_plain_tiled(S, coef, cf32, hc, 0, st);
}

size_t hc_coef_bytes(const Space* S, int prec) {
    const int M = S->norb;
    if (use_generic(S)) return sizeof(double2) * ((size_t) M * M + (size_t) M * M * M * M);
    // FP64 block, followed by the complex64 copy for the mixed form
    size_t b = sizeof(double2) * cf_size(M);
    if (prec == kPrecMixed) b += sizeof(float2) * cf32_size(M);
    return b;
}

int hc_build_coef(const Space* S, const double2* hob, const double2* hint, void* cf, int prec, hipStream_t st) {
    const int M = S->norb;
    if (use_generic(S)) {
        MCTDHB_HIP_CHECK(hipMemcpyAsync(cf, hob, sizeof(double2) * M * M, hipMemcpyDeviceToDevice, st));
        MCTDHB_HIP_CHECK(hipMemcpyAsync((double2*) cf + M * M, hint, sizeof(double2) * M * M * M * M,
                                        hipMemcpyDeviceToDevice, st));
        return kOk;
    }
    double2* cf64 = (double2*) cf;
    float2* cf32 = prec == kPrecMixed ? (float2*) (cf64 + cf_size(M)) : nullptr;
    return hc_coefficients(S, hob, hint, cf64, cf32, st);
}

int hc_apply_cf(const Space* S, const void* coef, const void* cf, void* hc, int prec, int index, hipStream_t st) {
    int rc = check_mode(S, prec, index);
    if (rc) return rc;
    const int M = S->norb;
    if (use_generic(S)) {
        const double2* hob = (const double2*) cf;
        EpiPlainG<double2> epi{(double2*) hc};
        k_hc_generic<<<hc_grid_blocks(S->dim), kHcBlockG, 0, st>>>(S->norb, row_ctx_g(S), (const double2*) coef,
                                                                   hob, hob + M * M, 0, nullptr, epi);
        MCTDHB_HIP_CHECK(hipGetLastError());
        return kOk;
    }
    if (prec == kPrecMixed && (rc = hc_tune_tiling(S, st))) return rc;
    const void* blk = prec == kPrecFP64 ? cf : (const void*) ((const double2*) cf + cf_size(M));
    return dispatch_fused_plain(S, coef, blk, hc, prec, index, st);
}

int pk_sil_small(const PkSil* tab, int count, int M, int index, bool stage, size_t smem, hipStream_t st) {
    switch (M) {
        case 2: return HcFused<2>::sil_small_pk(tab, count, index, stage, smem, st);
        case 3: return HcFused<3>::sil_small_pk(tab, count, index, stage, smem, st);
        case 4: return HcFused<4>::sil_small_pk(tab, count, index, stage, smem, st);
        case 5: return HcFused<5>::sil_small_pk(tab, count, index, stage, smem, st);
        case 6: return HcFused<6>::sil_small_pk(tab, count, index, stage, smem, st);
        case 7: return HcFused<7>::sil_small_pk(tab, count, index, stage, smem, st);
        default: set_error("the one-workgroup SIL step covers 2 <= norb <= 7"); return kInvalidArgument;
    }
}

int hc_sil_small(const Space* S, const double2* cf64, const SmallSilArgs& a, int index, hipStream_t st) {
    switch (S->norb) {
        case 2: return HcFused<2>::sil_small(S, cf64, a, index, st);
        case 3: return HcFused<3>::sil_small(S, cf64, a, index, st);
        case 4: return HcFused<4>::sil_small(S, cf64, a, index, st);
        case 5: return HcFused<5>::sil_small(S, cf64, a, index, st);
        case 6: return HcFused<6>::sil_small(S, cf64, a, index, st);
        case 7: return HcFused<7>::sil_small(S, cf64, a, index, st);
        default: set_error("the one-workgroup SIL step covers 2 <= norb <= 7"); return kInvalidArgument;
    }
}

int hc_lanczos(const Space* S, const double2* cf64, const float2* cf32, const LanczosHcArgs& a, int prec,
               int index, hipStream_t st) {
    int rc = check_mode(S, prec, index);
    if (rc) return rc;
    if (use_generic(S)) {
        // generic path is FP64: hob/hint are carried in cf64 as [hob | hint]
        const double2* hob = cf64;
        const double2* hint = cf64 + S->norb * S->norb;
        EpiLanczos<double2> epi{a};
        k_hc_generic<<<hc_grid_blocks(S->dim), kHcBlockG, 0, st>>>(S->norb, row_ctx_g(S),
                                                                   (const double2*) a.w[a.it], hob, hint, 1,
                                                                   a.st, epi);
        MCTDHB_HIP_CHECK(hipGetLastError());
        return kOk;
    }
    return dispatch_fused_lanczos(S, prec == kPrecFP64 ? (const void*) cf64 : (const void*) cf32, a, prec, index,
                                  st);
}

}  // namespace mctdhb
