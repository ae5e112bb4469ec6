// hamiltonian.hpp - H.C on the device.
//
// The reference applies H in 13 rule blocks (hamiltonian.c:44-401), many of
// which gather the same coefficient: every single-jump target i - e_k + e_l
// is read by the one-body term, Rule 4, Rule 5 and Rule 8.  Grouping the
// rules by jump target gives, with g[a,b,c,d] = hint[a + bM + cM^2 + dM^3]:
//
//  hc_i = E(n) c_i
//       + sum_{k: n_k>=1} sum_{l!=k} F_kl(n) sqrt(n_k (n_l+1)) c[i - e_k + e_l]
//       + sum_{k: n_k>=2} sum_{q<=l; q,l!=k} Got[k,q,l] b_ot c[i - 2e_k + e_q + e_l]
//       + sum_{k<s: n_k,n_s>=1} sum_{q<=l; q,l!=k,s} Gtt[k,s,q,l] b_tt c[i - e_k - e_s + e_q + e_l]
//
//  E(n)    = sum_k (h_kk - g_kkkk/2) n_k + g_kkkk/2 n_k^2 + sum_{k<s} 2 g_ksks n_k n_s   (one-body diag, Rules 1-2)
//  F_kl(n) = A_kl + sum_m n_m B_klm, A_kl = h_kl - g_kkkl,
//            B_klk = g_kkkl, B_kll = g_klll, B_kls = 2 g_kssl (s != k,l)      (one-body, Rules 4, 5, 8)
//  Got     = g_kkqq / 2 (q = l, Rule 3), g_kkql (q < l, Rule 6)
//  Gtt     = g_ksqq (q = l, Rule 7), 2 g_ksql (q < l, Rule 9)
//  b_ot    = sqrt(n_k (n_k-1) (n_q+1) (n_l+1 + [q=l])),  b_tt = sqrt(n_k n_s (n_q+1) (n_l+1 + [q=l]))
//
// which is the reference's H.C term for term, with at most one gather per
// distinct target (130 instead of 230 at N=100, M=5).
#pragma once

#include "sil.hpp"
#include "space.hpp"

namespace mctdhb {

// Index scheme of the H.C kernels.
enum HcIndex : int {
    kIndexTables = 0,  // reference jump tables (map, mapot, maptt)
    kIndexOnTheFly = 1 // closed-form rank deltas, no table reads
};
// Arithmetic form.
enum HcPrec : int {
    kPrecFP64 = 0,   // complex128 throughout
    kPrecMixed = 1   // complex64 gathers and FMAs (vectors stored as complex64)
};

// Coefficient block layout (complex entries), see the formula above.
__host__ __device__ constexpr int cf_off_A(int M) { return 0; }
__host__ __device__ constexpr int cf_off_B(int M) { return M * M; }
__host__ __device__ constexpr int cf_off_d1(int M) { return M * M + M * M * M; }
__host__ __device__ constexpr int cf_off_d2(int M) { return cf_off_d1(M) + M; }
__host__ __device__ constexpr int cf_off_d3(int M) { return cf_off_d2(M) + M; }
__host__ __device__ constexpr int cf_off_ot(int M) { return cf_off_d3(M) + M * M; }
__host__ __device__ constexpr int cf_off_tt(int M) { return cf_off_ot(M) + M * M * M; }
__host__ __device__ constexpr int cf_size(int M) { return cf_off_tt(M) + M * M * M * M; }
// The mixed-precision block: the complex64 copy of the block above, then
// (16-byte aligned) its one-body and diagonal sections A, B, d1, d2, d3 again
// in complex128, for the terms the mixed rows form in FP64.  Sizes in
// complex64 elements.
__host__ __device__ constexpr int cf_f64_off(int M) { return (cf_size(M) + 1) & ~1; }
__host__ __device__ constexpr int cf_f64_count(int M) { return cf_off_ot(M); }
__host__ __device__ constexpr int cf32_size(int M) { return cf_f64_off(M) + 2 * cf_f64_count(M); }

constexpr int kFusedMaxOrb = 7;  // fused kernels instantiated for 2 <= M <= 7

// Build the coefficient block from hob (M x M row-major) and hint (M^4),
// both complex128 device arrays, into cf64 (and the mixed block cf32, of
// cf32_size(M) complex64 elements, if non-null).
int hc_coefficients(const Space* S, const double2* hob, const double2* hint, double2* cf64, float2* cf32,
                    hipStream_t st);

// hc = H coef.  prec / index select the kernel; coef/hc are complex128 for
// FP64 and complex64 for mixed.  Uses the reference-order generic kernel for
// M > 8 (tables required) and M == 1.
int hc_apply(const Space* S, const void* coef, const double2* hob, const double2* hint, void* hc, int prec,
             int index, hipStream_t st);

// hc_apply with no device allocation: the coefficient block goes to a fixed
// file-scope device array of kHcScratchBytes (one per GPU, shared by every
// call on that GPU, so calls that overlap on one GPU race).  The mixed form
// runs one configuration per lane unless the paired tiling is already chosen
// and its padded staging exists.  hc_scratch_setup resolves the array's
// address on the current device (once per device; idempotent).
constexpr size_t kHcScratchBytes = 128 * 1024;
int hc_scratch_setup();
int hc_apply_noalloc(const Space* S, const void* coef, const double2* hob, const double2* hint, void* hc, int prec,
                     int index, hipStream_t st);

// Coefficient-block entry points: bytes needed, build, and H.C with a
// prebuilt block (the generic kernel's block is [hob | hint] in FP64).
size_t hc_coef_bytes(const Space* S, int prec);
int hc_build_coef(const Space* S, const double2* hob, const double2* hint, void* cf, int prec, hipStream_t st);
int hc_apply_cf(const Space* S, const void* coef, const void* cf, void* hc, int prec, int index, hipStream_t st);

// Lanczos form (iteration it): out = s_it H w_it - beta_{it-1} s_{it-1} w_{it-1}
// (out is complex128) plus per-block partial sums, slot-major: slot t of
// block b at part[t * nblocks + b], slots 2k, 2k+1 = sum conj(s_k w_k) out
// for k <= it and, at it = 0, slot 2(it+1) = |w_0|^2.
// All scales come from the device SilState; the kernel is a no-op after a
// Lanczos breakdown (st->stop).
struct LanczosHcArgs {
    const void* w[kMaxKrylov];
    int it;
    int last;  // final iteration: only alpha_it (slots 2it, 2it+1), out not stored
    const SilState* st;
    double2* out;
    double* part;
};
__host__ __device__ inline int lanczos_hc_slots(int it) { return 2 * (it + 1) + 2; }
int hc_lanczos(const Space* S, const double2* cf64, const float2* cf32, const LanczosHcArgs& a, int prec,
               int index, hipStream_t st);

unsigned hc_grid_blocks(uint32_t dim);
// The mixed form has three tilings: one configuration per lane (0), two
// configurations of one run per lane (1, the pair work list of the space),
// or one configuration per lane with each block's coefficient windows staged
// in LDS, one far group of jump types after another (2, 2 <= norb <= 5 and
// N up to 185; the plain product only: the Lanczos forms keep the one-lane
// kernel under it).
constexpr int kTilingLane = 0, kTilingPair = 1, kTilingWindow = 2;
inline bool pair_possible(const Space* S) { return S->d_slots && S->dim >= 2 && S->norb >= 2 && S->norb <= 7; }
inline bool use_pair(const Space* S, int prec) {
    return prec == kPrecMixed && pair_possible(S) && S->hc_tiling == kTilingPair;
}
// Whether the windows of this (N, M) fit the windowed kernel's LDS budget.
bool hc_window_possible(const Space* S);
inline bool use_window(const Space* S, int prec) {
    return prec == kPrecMixed && S->hc_tiling == kTilingWindow && hc_window_possible(S);
}
// Times the possible mixed tilings once per (N, M) (zero data: the kernels'
// cost does not depend on the values) and records the fastest in S->hc_tiling.
// Synchronises the stream; no-op once done.
int hc_tune_tiling(const Space* S, hipStream_t st);
// The choice is kept per (N, M) for the process (so every space of a shape
// runs the same kernels); this drops it, and the space's, for a new timing.
void hc_forget_tiling(const Space* S);
// Blocks of the H.C launch of the Lanczos form (one partial-sum row each).
unsigned hc_lanczos_blocks(const Space* S, int prec);
// Whether the last Lanczos iteration runs the half-row Rayleigh form (the
// one-configuration-per-lane fused kernel): slot 2 it then holds
// <w_it, H w_it> instead of <v_it, out>.
inline bool hc_lanczos_half(const Space* S, int prec) {
    return S->norb >= 2 && S->norb <= kFusedMaxOrb && !use_pair(S, prec);
}
// The paired kernel reads c[-1] and c[dim] (zero padding).  Copy of c into
// the space's padded staging (allocated on first use); returns element 0.
const float2* hc_padded_copy(const Space* S, const float2* c, hipStream_t st);

// The whole FP64 SIL step in one workgroup (sil_small.hpp): w[0] is C
// (updated in place), w[1..lm-1] Krylov storage, out one more vector.
constexpr uint32_t kSmallSilMaxDim = 8192;
constexpr int kSmallSilBlock = 512;
// LDS of that kernel after the W table and the sqrt LUT.
struct SmallSilLds {
    double red[(2 * kMaxLanczos + 2) * (kSmallSilBlock / 64)];
    double tot[2 * kMaxLanczos + 4];
    double z[64 * (kMaxLanczos + 2)];  // every lane of wave 0 owns a row (odd stride: no bank conflicts)
    double2 cre[kMaxLanczos];
    double2* w[kMaxLanczos];  // the vectors, indexed at run time
    SilState state;           // the scalars during the step, copied out after the eigensolve
};
inline size_t sil_small_smem(const Space* S) {
    const size_t wsz = (size_t) (S->norb - 1) * S->ldw * 4;
    return ((wsz + 15) & ~(size_t) 15) + ((sizeof(double) * (S->npar + 3) + 15) & ~(size_t) 15) + sizeof(SmallSilLds);
}
// The staged form keeps the current vector and the occupation rows in LDS
// when everything fits in the 160 KiB one workgroup may allocate on gfx950
// (NC up to about 6000).
inline size_t sil_small_stage_bytes(const Space* S) { return (size_t) S->dim * (sizeof(double2) + 8); }
inline bool sil_small_stage(const Space* S) {
    return sil_small_smem(S) + sil_small_stage_bytes(S) <= kLdsPerWorkgroup;
}
inline bool sil_small_fits(const Space* S, int lm) {
    return S->norb >= 2 && S->norb <= 7 && S->dim <= kSmallSilMaxDim && lm <= kMaxLanczos &&
           sil_small_smem(S) <= 64 * 1024;
}
struct SmallSilArgs {
    SilState* st;
    void* w[kMaxLanczos];
    double2* out;
    int lm;
    double dt_re, dt_im;
    int imagtime;
};
int hc_sil_small(const Space* S, const double2* cf64, const SmallSilArgs& a, int index, hipStream_t st);

struct PkSil;

// Per-orbital-count entry points of the fused kernel (hc_m<M>.hip).  The
// mixed form always uses on-the-fly indices.
template <int M> struct HcFused {
    static int plain(const Space* S, const void* c, const void* cf, void* out, int prec, int index, hipStream_t st);
    // mixed form with an explicit tiling (0 per lane, 1 pairs, 2 windows; c padded for 1)
    static int plain_tiled(const Space* S, const void* c, const void* cf, void* out, int tiling, hipStream_t st);
    static int lanczos(const Space* S, const void* cf, const LanczosHcArgs& a, int prec, int index, hipStream_t st);
    static int sil_small(const Space* S, const void* cf, const SmallSilArgs& a, int index, hipStream_t st);
    // the same step for count instances, one workgroup each (pack.hpp)
    static int sil_small_pk(const PkSil* tab, int count, int index, bool stage, size_t smem, hipStream_t st);
};

}  // namespace mctdhb
