// density_impl.hpp - templates of the fused density-matrix kernel,
// instantiated once per orbital count in dm_m<M>.hip.
#pragma once

#include <type_traits>
#include <utility>

#include "density.hpp"
#include "pack.hpp"
#include "rowctx.hpp"

namespace mctdhb {

constexpr int kDmBlock = 256;
constexpr int kDmCPT = 16;  // configurations per thread per tile
constexpr int kDmTile = kDmBlock * kDmCPT;

// ---------------------------------------------------------------- layout
__host__ __device__ constexpr int c2(int m) { return m * (m - 1) / 2; }
struct DmLayout {
    int M;
    __host__ __device__ constexpr int d1() const { return 0; }
    __host__ __device__ constexpr int d2() const { return M; }
    __host__ __device__ constexpr int d3() const { return 2 * M; }
    __host__ __device__ constexpr int pair() const { return 2 * M + c2(M); }
    __host__ __device__ constexpr int o3() const { return pair() + c2(M) * (M + 1); }
    __host__ __device__ constexpr int o6() const { return o3() + c2(M); }
    __host__ __device__ constexpr int t8() const { return o6() + M * c2(M - 1); }
    __host__ __device__ constexpr int total() const { return t8() + c2(M) * c2(M > 2 ? M - 2 : 0); }
    // lexicographic index of the pair a < b among 0..M-1
    __host__ __device__ constexpr int pidx(int a, int b) const { return a * M - a * (a + 1) / 2 + (b - a - 1); }
    // index of the pair q < l among the orbitals other than x (lexicographic)
    __host__ __device__ constexpr int pidx_excl1(int x, int q, int l) const {
        const int qq = q - (q > x), ll = l - (l > x);
        const int m = M - 1;
        return qq * m - qq * (qq + 1) / 2 + (ll - qq - 1);
    }
    __host__ __device__ constexpr int pidx_excl2(int x, int y, int q, int l) const {
        const int qq = q - (q > x) - (q > y), ll = l - (l > x) - (l > y);
        const int m = M - 2;
        return qq * m - qq * (qq + 1) / 2 + (ll - qq - 1);
    }
};

// --------------------------------------------------------- accumulators
template <typename R> struct Acc;
template <> struct Acc<double> {
    double2 s;
    __device__ __forceinline__ void zero() { s = make_double2(0, 0); }
    __device__ __forceinline__ void add(double2 t) { s = s + t; }
    __device__ __forceinline__ void add_re(double t) { s.x += t; }
    __device__ __forceinline__ double2 get() const { return s; }
};
// Kahan-compensated complex FP32 sum
template <> struct Acc<float> {
    float2 s, c;
    __device__ __forceinline__ void zero() {
        s = make_float2(0, 0);
        c = make_float2(0, 0);
    }
    __device__ __forceinline__ static void k1(float& sum, float& comp, float t) {
        const float y = t - comp;
        const float u = sum + y;
        comp = (u - sum) - y;
        sum = u;
    }
    __device__ __forceinline__ void add(float2 t) {
        k1(s.x, c.x, t.x);
        k1(s.y, c.y, t.y);
    }
    __device__ __forceinline__ void add_re(float t) { k1(s.x, c.x, t); }
    __device__ __forceinline__ double2 get() const {
        return make_double2((double) s.x - (double) c.x, (double) s.y - (double) c.y);
    }
};

// ------------------------------------------------------------- passes
// kind 0: diagonal moments; 1: pairs (k, l in [a, b)); 2: same-orbital (k);
// 3: different-orbital (k, s in [a, b)).
struct PassDesc {
    int kind, k, a, b;
};
template <int M> struct Passes {
    static constexpr int kPairChunk = M <= 6 ? 8 : 3;
    static constexpr bool kTTPerK = (M - 1) * c2(M > 2 ? M - 2 : 0) <= 32;
    static constexpr int count() {
        int n = 1;
        for (int k = 0; k < M - 1; k++) n += (M - 1 - k + kPairChunk - 1) / kPairChunk;
        n += M;
        if (M >= 4) {
            for (int k = 0; k < M - 1; k++) n += kTTPerK ? 1 : (M - 1 - k);
        }
        return n;
    }
    static constexpr PassDesc get(int p) {
        if (p == 0) return {0, 0, 0, 0};
        int n = 1;
        for (int k = 0; k < M - 1; k++)
            for (int a = k + 1; a < M; a += kPairChunk) {
                if (n == p) return {1, k, a, (a + kPairChunk < M ? a + kPairChunk : M)};
                n++;
            }
        for (int k = 0; k < M; k++) {
            if (n == p) return {2, k, 0, 0};
            n++;
        }
        for (int k = 0; k < M - 1; k++) {
            if (kTTPerK) {
                if (n == p) return {3, k, k + 1, M};
                n++;
            } else {
                for (int s = k + 1; s < M; s++) {
                    if (n == p) return {3, k, s, s + 1};
                    n++;
                }
            }
        }
        return {-1, 0, 0, 0};
    }
    static constexpr int nacc(int p) {
        const PassDesc d = get(p);
        if (d.kind == 0) return 2 * M + c2(M);
        if (d.kind == 1) return (d.b - d.a) * (M + 1);
        if (d.kind == 2) return (M - 1 - d.k) + c2(M - 1);
        if (d.kind == 3) return (d.b - d.a) * c2(M - 2);
        return 0;
    }
};

template <int M, typename R, int IDX, int P>
__device__ __forceinline__ void run_pass(const RowCtx& ctx, const cplx<R>* __restrict__ c, uint32_t base,
                                         const uint32_t* Wl, int ldw, const R* sql, double* red,
                                         double2* __restrict__ part, uint32_t row) {
    using Z = cplx<R>;
    constexpr PassDesc d = Passes<M>::get(P);
    constexpr int NA = Passes<M>::nacc(P) > 0 ? Passes<M>::nacc(P) : 1;
    constexpr DmLayout L{M};
    Acc<R> acc[NA];
#pragma unroll
    for (int t = 0; t < NA; t++) acc[t].zero();

    for (int r = 0; r < kDmCPT; r++) {
        const uint32_t i = base + threadIdx.x + r * kDmBlock;
        if (i >= ctx.dim) break;
        Cfg<M, R, IDX> x;
        x.load(i, ctx, Wl, ldw, sql);
        const int* n = x.n;
        const Z ci = c[i];
        if constexpr (d.kind == 0) {
            const R w = ci.x * ci.x + ci.y * ci.y;
#pragma unroll
            for (int k = 0; k < M; k++) {
                acc[k].add_re(w * (R) n[k]);
                acc[M + k].add_re(w * (R) (n[k] * (n[k] - 1)));
#pragma unroll
                for (int s = k + 1; s < M; s++) acc[2 * M + L.pidx(k, s)].add_re(w * (R) (n[k] * n[s]));
            }
        } else if constexpr (d.kind == 1) {
            constexpr int k = d.k;
            if (n[k] >= 1) {
#pragma unroll
                for (int l = d.a; l < d.b; l++) {
                    const Z cj = c[x.single(k, l)];
                    const Z t = cscale(cconjmul(ci, cj), x.s0[k] * x.s1[l]);
                    const int o = (l - d.a) * (M + 1);
                    acc[o].add(t);
#pragma unroll
                    for (int m = 0; m < M; m++) acc[o + 1 + m].add(cscale(t, (R) n[m]));
                }
            }
        } else if constexpr (d.kind == 2) {
            constexpr int k = d.k;
            if (n[k] >= 2) {
                const R a = x.sm[k] * x.s0[k];
#pragma unroll
                for (int q = k + 1; q < M; q++) {
                    const Z cj = c[x.same(k, q, q)];
                    acc[q - k - 1].add(cscale(cconjmul(ci, cj), a * x.s1[q] * x.s2[q]));
                }
#pragma unroll
                for (int q = 0; q < M; q++) {
                    if (q == k) continue;
#pragma unroll
                    for (int l = q + 1; l < M; l++) {
                        if (l == k) continue;
                        const Z cj = c[x.same(k, q, l)];
                        acc[(M - 1 - k) + L.pidx_excl1(k, q, l)].add(
                            cscale(cconjmul(ci, cj), a * x.s1[q] * x.s1[l]));
                    }
                }
            }
        } else if constexpr (d.kind == 3) {
            constexpr int k = d.k;
#pragma unroll
            for (int s = d.a; s < d.b; s++) {
                if (n[k] >= 1 && n[s] >= 1) {
                    const R a = x.s0[k] * x.s0[s];
#pragma unroll
                    for (int q = 0; q < M; q++) {
                        if (q == k || q == s) continue;
#pragma unroll
                        for (int l = q + 1; l < M; l++) {
                            if (l == k || l == s) continue;
                            const Z cj = c[x.diff(k, s, q, l)];
                            acc[(s - d.a) * c2(M - 2) + L.pidx_excl2(k, s, q, l)].add(
                                cscale(cconjmul(ci, cj), a * x.s1[q] * x.s1[l]));
                        }
                    }
                }
            }
        }
    }
    // block reduction, then one partial per element of this pass
    double v[2 * NA];
#pragma unroll
    for (int t = 0; t < NA; t++) {
        const double2 g = acc[t].get();
        v[2 * t] = g.x;
        v[2 * t + 1] = g.y;
    }
    block_sum<2 * NA>(v, red);
    if (threadIdx.x == 0) {
        const int E = L.total();
        double2* out = part + (size_t) row * E;
#pragma unroll
        for (int t = 0; t < NA; t++) {
            int e;
            if constexpr (d.kind == 0) e = t;  // d1, d2, d3 are contiguous from 0
            else if constexpr (d.kind == 1) {
                const int l = d.a + t / (M + 1);
                e = L.pair() + L.pidx(d.k, l) * (M + 1) + t % (M + 1);
            } else if constexpr (d.kind == 2) {
                if (t < M - 1 - d.k) e = L.o3() + L.pidx(d.k, d.k + 1 + t);
                else e = L.o6() + d.k * c2(M - 1) + (t - (M - 1 - d.k));
            } else {
                const int s = d.a + t / c2(M - 2);
                e = L.t8() + L.pidx(d.k, s) * c2(M - 2) + t % c2(M - 2);
            }
            out[e] = make_double2(v[2 * t], v[2 * t + 1]);
        }
    }
}

template <int M, typename R, int IDX, int... P>
__device__ __forceinline__ void run_one(std::integer_sequence<int, P...>, int pass, const RowCtx& ctx,
                                        const cplx<R>* __restrict__ c, uint32_t base, const uint32_t* Wl, int ldw,
                                        const R* sql, double* red, double2* part, uint32_t row) {
    ((P == pass ? run_pass<M, R, IDX, P>(ctx, c, base, Wl, ldw, sql, red, part, row) : (void) 0), ...);
}

// Block w of the (padded) grid of 8 R blocks runs (tile, family) pair
// v = (w mod 8) R + w / 8, tile-major: XCD w mod 8 gets the contiguous range
// [x R, (x + 1) R), so a tile's families are dispatched one after another on
// one XCD and read its coefficients and occupation rows through that XCD's
// L2 (dispatched family after family over the whole space, every family
// streamed them from memory again: 3.2 GB per FP64 launch at N=100, M=5,
// 0.35 GB now).  Returns false for the padding blocks.
__device__ __forceinline__ bool dm_tile_family(uint32_t bid, uint32_t grid, uint32_t ntiles, uint32_t nfam,
                                               uint32_t& tile, uint32_t& fam) {
    const uint32_t R = grid / 8;
    const uint32_t v = (bid % 8) * R + bid / 8;
    if (v >= ntiles * nfam) return false;
    tile = v / nfam;
    fam = v % nfam;
    return true;
}
inline unsigned dm_family_grid(unsigned ntiles, unsigned nfam) { return 8 * div_up((uint64_t) ntiles * nfam, 8); }

template <int M, typename R, int IDX>
__device__ __forceinline__ void density_block(unsigned char* smem, uint32_t bid, uint32_t grid, const RowCtx& ctx,
                                              const cplx<R>* __restrict__ c, const uint32_t* __restrict__ Wg,
                                              int ldw, int use_sq, double2* __restrict__ part, uint32_t ntiles) {
    uint32_t tile, fam;
    if (!dm_tile_family(bid, grid, ntiles, Passes<M>::count(), tile, fam)) return;  // before any barrier
    constexpr int kRed = 2 * (2 * M + c2(M) + 8 * 9 + 64);  // >= 2 * max nacc
    double* red = (double*) smem;
    uint32_t* Wl = (uint32_t*) (smem + kRed * (kDmBlock / 64) * sizeof(double));
    const int wsz = (M - 1) * ldw;
    R* sql = (R*) ((unsigned char*) Wl + ((wsz * 4 + 15) & ~15));
    if constexpr (IDX == kIdxFly) {
        for (int t = threadIdx.x; t < wsz; t += blockDim.x) Wl[t] = Wg[t];
    }
    if (use_sq)
        for (int t = threadIdx.x; t < ctx.npar + 3; t += blockDim.x) sql[t] = sqrt((R) t);
    __syncthreads();
    run_one<M, R, IDX>(std::make_integer_sequence<int, Passes<M>::count()>{}, (int) fam, ctx, c, tile * kDmTile, Wl,
                       ldw, use_sq ? sql : nullptr, red, part, tile);
}

template <int M, typename R, int IDX>
__global__ __launch_bounds__(kDmBlock) void k_density(RowCtx ctx, const cplx<R>* __restrict__ c,
                                                      const uint32_t* __restrict__ Wg, int ldw, int use_sq,
                                                      double2* __restrict__ part, uint32_t ntiles) {
    extern __shared__ __align__(16) unsigned char smem[];
    density_block<M, R, IDX>(smem, blockIdx.x, gridDim.x, ctx, c, Wg, ldw, use_sq, part, ntiles);
}

// Packed form: instance blockIdx.y over its own family grid; blocks past it
// return.
template <int M>
__global__ __launch_bounds__(kDmBlock) void k_density_pk(const PkDm* __restrict__ tab) {
    extern __shared__ __align__(16) unsigned char smem[];
    const PkDm& p = tab[blockIdx.y];
    if (blockIdx.x >= p.grid) return;
    density_block<M, double, kIdxFly>(smem, blockIdx.x, p.grid, p.ctx, (const double2*) p.c, p.W, p.ldw, p.use_sq,
                                      p.part, p.ntiles);
}

// ------------------------------------------------ one-pass wave form
// (M >= 6, where the element families are many: 17 at M = 6, 38 at M = 7.)
// Every configuration is visited once: its 2M + C(M,2) diagonal moments and
// its jump-target products are formed once each, and each element's values
// are summed over the wave's 64 configurations right away (DPP row shifts
// inside each 16-lane row, the four row totals through readlane, all in
// FP64) into a per-wave accumulator in LDS.  Nothing per configuration is
// recomputed per element family, and no per-thread accumulator array is
// needed.  The mixed form forms the products in complex64 and sums in FP64.

// Sum of v over the 64 lanes of the wave, returned in every lane; every lane
// of the wave must be active.  The order of the additions is fixed.
__device__ __forceinline__ double wave_total(double v) { return wave_sum(v); }

// Per-wave accumulators, spread over the lanes: element e lives in lane
// e % 64 of acc[e / 64], so adding a wave total costs a masked add and no
// memory round trip.
template <int E> struct WaveAcc {
    static constexpr int NR = (E + 63) / 64;
    double2 acc[NR];
    int lane;
    __device__ __forceinline__ void zero(int l) {
        lane = l;
#pragma unroll
        for (int j = 0; j < NR; j++) acc[j] = make_double2(0, 0);
    }
    __device__ __forceinline__ void add(int e, double2 v) {
        const double x = wave_total(v.x), y = wave_total(v.y);
        if (lane == (e & 63)) acc[e >> 6] = acc[e >> 6] + make_double2(x, y);
    }
    __device__ __forceinline__ void add_re(int e, double v) {
        const double x = wave_total(v);
        if (lane == (e & 63)) acc[e >> 6].x += x;
    }
};

template <int M, typename R, int IDX, typename WA>
__device__ __forceinline__ void wave_config(const RowCtx& ctx, const cplx<R>* __restrict__ c, uint32_t i,
                                            const uint32_t* Wl, int ldw, const R* sql, WA& wa) {
    using Z = cplx<R>;
    constexpr DmLayout L{M};
    const bool valid = i < ctx.dim;
    Cfg<M, R, IDX> x;
    Z ci = mk<R>(0, 0);
    if (valid) {
        x.load(i, ctx, Wl, ldw, sql);
        ci = c[i];
    } else {
#pragma unroll
        for (int k = 0; k < M; k++) {  // an empty row: every term is zero
            x.n[k] = 0;
            x.sm[k] = x.s0[k] = x.s1[k] = x.s2[k] = R(0);
        }
    }
    const int* n = x.n;
    // diagonal moments (real)
    {
        const double w = (double) ci.x * ci.x + (double) ci.y * ci.y;
#pragma unroll
        for (int k = 0; k < M; k++) {
            wa.add_re(L.d1() + k, w * n[k]);
            wa.add_re(L.d2() + k, w * (double) (n[k] * (n[k] - 1)));
#pragma unroll
            for (int s = k + 1; s < M; s++) wa.add_re(L.d3() + L.pidx(k, s), w * (double) (n[k] * n[s]));
        }
    }
    // one particle k -> l (k < l): the pair element and its n_m moments
#pragma unroll
    for (int k = 0; k < M - 1; k++) {
#pragma unroll
        for (int l = k + 1; l < M; l++) {
            double2 t = make_double2(0, 0);
            if (n[k] >= 1) {
                const Z cj = c[x.single(k, l)];
                t = to_d2(cscale(cconjmul(ci, cj), x.s0[k] * x.s1[l]));
            }
            const int o = L.pair() + L.pidx(k, l) * (M + 1);
            wa.add(o, t);
#pragma unroll
            for (int m = 0; m < M; m++) wa.add(o + 1 + m, cscale(t, (double) n[m]));
        }
    }
    // two particles out of orbital k
#pragma unroll
    for (int k = 0; k < M; k++) {
        const bool ok = n[k] >= 2;
        const R a = x.sm[k] * x.s0[k];
#pragma unroll
        for (int q = k + 1; q < M; q++) {
            double2 t = make_double2(0, 0);
            if (ok) t = to_d2(cscale(cconjmul(ci, c[x.same(k, q, q)]), a * x.s1[q] * x.s2[q]));
            wa.add(L.o3() + L.pidx(k, q), t);
        }
#pragma unroll
        for (int q = 0; q < M; q++) {
            if (q == k) continue;
#pragma unroll
            for (int l = q + 1; l < M; l++) {
                if (l == k) continue;
                double2 t = make_double2(0, 0);
                if (ok) t = to_d2(cscale(cconjmul(ci, c[x.same(k, q, l)]), a * x.s1[q] * x.s1[l]));
                wa.add(L.o6() + k * c2(M - 1) + L.pidx_excl1(k, q, l), t);
            }
        }
    }
    // one particle out of each of k < s
    if constexpr (M >= 4) {
#pragma unroll
        for (int k = 0; k < M - 1; k++) {
#pragma unroll
            for (int s = k + 1; s < M; s++) {
                const bool ok = n[k] >= 1 && n[s] >= 1;
                const R a = x.s0[k] * x.s0[s];
#pragma unroll
                for (int q = 0; q < M; q++) {
                    if (q == k || q == s) continue;
#pragma unroll
                    for (int l = q + 1; l < M; l++) {
                        if (l == k || l == s) continue;
                        double2 t = make_double2(0, 0);
                        if (ok) t = to_d2(cscale(cconjmul(ci, c[x.diff(k, s, q, l)]), a * x.s1[q] * x.s1[l]));
                        wa.add(L.t8() + L.pidx(k, s) * c2(M - 2) + L.pidx_excl2(k, s, q, l), t);
                    }
                }
            }
        }
    }
}

// One block of 256 threads takes tile b of `cpt` x 256 consecutive
// configurations; its four per-wave accumulator sets are added in a fixed
// order into part[b * E ...].
template <int M, typename R, int IDX>
__global__ __launch_bounds__(kDmBlock) void k_density_wave(RowCtx ctx, const cplx<R>* __restrict__ c,
                                                           const uint32_t* __restrict__ Wg, int ldw, int cpt,
                                                           double2* __restrict__ part) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int E = DmLayout{M}.total();
    constexpr int NW = kDmBlock / 64;
    double2* acc = (double2*) smem;
    uint32_t* Wl = (uint32_t*) (smem + sizeof(double2) * NW * E);
    const int wsz = (M - 1) * ldw;
    R* sql = (R*) ((unsigned char*) Wl + ((wsz * 4 + 15) & ~15));
    if constexpr (IDX == kIdxFly) {
        for (int t = threadIdx.x; t < wsz; t += blockDim.x) Wl[t] = Wg[t];
    }
    for (int t = threadIdx.x; t < ctx.npar + 3; t += blockDim.x) sql[t] = sqrt((R) t);
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    WaveAcc<E> wa;
    wa.zero(lane);
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t base = b * (uint32_t) cpt * kDmBlock + threadIdx.x;
    for (int r = 0; r < cpt; r++) wave_config<M, R, IDX>(ctx, c, base + r * kDmBlock, Wl, ldw, sql, wa);
#pragma unroll
    for (int j = 0; j < WaveAcc<E>::NR; j++)
        if (j * 64 + lane < E) acc[wid * E + j * 64 + lane] = wa.acc[j];
    __syncthreads();
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        double2 v = acc[e];
#pragma unroll
        for (int w = 1; w < NW; w++) v = v + acc[w * E + e];
        part[(size_t) b * E + e] = v;
    }
}

template <int M> static size_t dm_wave_smem(const Space* S, size_t rsz) {
    return sizeof(double2) * (kDmBlock / 64) * DmLayout{M}.total() +
           (((size_t) (M - 1) * S->ldw * 4 + 15) & ~(size_t) 15) + rsz * (S->npar + 3);
}

template <int M>
bool DmFused<M>::wave_fits(const Space* S, int prec, int index) {
    const size_t rsz = prec == kDmMixed ? sizeof(float) : sizeof(double);
    return M >= kDmWaveMinOrb && dm_wave_smem<M>(S, rsz) <= 64 * 1024;
}

template <int M>
int DmFused<M>::run_pk(const PkDm* tab, int count, unsigned max_grid, size_t smem, hipStream_t st) {
    k_density_pk<M><<<dim3(max_grid, count), kDmBlock, smem, st>>>(tab);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

template <int M>
int DmFused<M>::run(const Space* S, const void* c, int prec, int index, double2* part, const DmPlan& plan,
                    hipStream_t st) {
    const bool fly = prec == kDmMixed || index == kIdxFly;
    const size_t rsz = prec == kDmMixed ? sizeof(float) : sizeof(double);
    RowCtx ctx = make_row_ctx(S->dim, S->npar, S->d_ht, S->d_map, S->d_mapot, S->d_maptt, S->d_strideot,
                              S->d_stridett, S->d_occ8);
    if constexpr (M >= kDmWaveMinOrb) {
    if (plan.wave) {
        const size_t smem = dm_wave_smem<M>(S, rsz);
        if (prec == kDmMixed)
            k_density_wave<M, float, kIdxFly><<<plan.nb, kDmBlock, smem, st>>>(ctx, (const float2*) c, S->d_W, S->ldw,
                                                                              plan.cpt, part);
        else if (fly)
            k_density_wave<M, double, kIdxFly><<<plan.nb, kDmBlock, smem, st>>>(ctx, (const double2*) c, S->d_W,
                                                                               S->ldw, plan.cpt, part);
        else
            k_density_wave<M, double, kIdxTables><<<plan.nb, kDmBlock, smem, st>>>(ctx, (const double2*) c, S->d_W,
                                                                                  S->ldw, plan.cpt, part);
        MCTDHB_HIP_CHECK(hipGetLastError());
        return kOk;
    }
    }
    const unsigned nb = plan.nb;
    const bool use_sq = S->npar + 3 <= 2048;
    const size_t smem = dm_family_smem(S, rsz);
    if (!smem) {
        set_error("density kernel LDS budget exceeded for this (N, M); use table indexing");
        return kInvalidArgument;
    }
    // One element family per block: every block's chain of dependent loads
    // is one family long, and 14x more blocks (M = 5) are in flight than
    // with every family in turn in one block; a tile's families run
    // together on one XCD (dm_tile_family).
    const unsigned grid = dm_family_grid(nb, Passes<M>::count());
    if (prec == kDmMixed)
        k_density<M, float, kIdxFly><<<grid, kDmBlock, smem, st>>>(ctx, (const float2*) c, S->d_W, S->ldw, use_sq,
                                                                   part, nb);
    else if (fly)
        k_density<M, double, kIdxFly><<<grid, kDmBlock, smem, st>>>(ctx, (const double2*) c, S->d_W, S->ldw,
                                                                    use_sq, part, nb);
    else
        k_density<M, double, kIdxTables><<<grid, kDmBlock, smem, st>>>(ctx, (const double2*) c, S->d_W, S->ldw,
                                                                       use_sq, part, nb);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

}  // namespace mctdhb
