// density.hip - one- and two-body density matrices on the device.
//
// Replaces reference src/configurational/density_matrices.c:8-443
// (set_onebody_dm, set_twobody_dm).  The reference runs one OpenMP
// reduction over all configurations per unique element (~155 sweeps at
// M = 5).  Here each element family is a block reduction over a tile of
// configurations, every gather serves all the elements that share its jump
// target, and a deterministic second stage sums the per-block partials.
//
// Unique elements (e = layout index, see DmLayout):
//   D1[k]    = sum |c|^2 n_k                        -> rho1[k][k]
//   D2[k]    = sum |c|^2 n_k (n_k - 1)              -> rho2[kkkk]          (Rule 1)
//   D3[k<s]  = sum |c|^2 n_k n_s                    -> rho2[ksks] & perms  (Rule 2)
//   pair (k<l), t = conj(c_i) c[i-e_k+e_l] sqrt(n_k (n_l+1)):
//     T0 = sum t                                    -> rho1[k][l]
//     T_m = sum n_m t                               -> rho2[kkkl] = T_k - T0 (Rule 4),
//                                                      rho2[klll] = T_l (Rule 5),
//                                                      rho2[kssl] = T_s (Rule 7)
//   O3 (k<q)      = sum conj(c_i) c[i-2e_k+2e_q] sqrt((n_k-1) n_k (n_q+1)(n_q+2))   (Rule 3)
//   O6 (k; q<l)   = sum conj(c_i) c[i-2e_k+e_q+e_l] sqrt(n_k (n_k-1)(n_q+1)(n_l+1)) (Rule 6)
//   T8 (k<s; q<l) = sum conj(c_i) c[i-e_k-e_s+e_q+e_l] sqrt(n_k n_s (n_q+1)(n_l+1)) (Rule 8)
// and the reference's symmetric fill assembles rho1 (M x M) and rho2 (M^4,
// element k + sM + qM^2 + lM^3 = <a+_k a+_s a_q a_l>).
//
// The mixed form reads complex64 coefficients, multiplies in FP32 and sums
// each thread's configurations with Kahan compensation; the block and grid
// stages run in FP64.
#include <algorithm>
#include <vector>

#include "density_impl.hpp"
#include "mxm_wave.hpp"
#include "pack.hpp"

namespace mctdhb {

int dm_num_elements(int M) { return DmLayout{M}.total(); }

// -------------------------------------------------- generic (any M, FP64)
// One block per (tile, element); the element is described at run time.
struct ElemDesc {
    int type;  // 0 d1, 1 d2, 2 d3, 3 pair T0, 4 pair Tm, 5 o3, 6 o6, 7 t8
    int k, s, q, l, m;
};

static std::vector<ElemDesc> element_list(int M) {
    DmLayout L{M};
    std::vector<ElemDesc> v(L.total());
    for (int k = 0; k < M; k++) {
        v[L.d1() + k] = {0, k, 0, 0, 0, 0};
        v[L.d2() + k] = {1, k, 0, 0, 0, 0};
    }
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++) {
            v[L.d3() + L.pidx(k, s)] = {2, k, s, 0, 0, 0};
            const int o = L.pair() + L.pidx(k, s) * (M + 1);
            v[o] = {3, k, 0, 0, s, 0};
            for (int m = 0; m < M; m++) v[o + 1 + m] = {4, k, 0, 0, s, m};
            v[L.o3() + L.pidx(k, s)] = {5, k, 0, s, s, 0};
        }
    for (int k = 0; k < M; k++)
        for (int q = 0; q < M; q++)
            for (int l = q + 1; l < M; l++)
                if (q != k && l != k) v[L.o6() + k * c2(M - 1) + L.pidx_excl1(k, q, l)] = {6, k, 0, q, l, 0};
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++)
            for (int q = 0; q < M; q++)
                for (int l = q + 1; l < M; l++)
                    if (q != k && q != s && l != k && l != s)
                        v[L.t8() + L.pidx(k, s) * c2(M - 2) + L.pidx_excl2(k, s, q, l)] = {7, k, s, q, l, 0};
    return v;
}

__global__ __launch_bounds__(kDmBlock) void k_density_generic(int M, RowCtx ctx, const double2* __restrict__ c,
                                                              const ElemDesc* __restrict__ elems, int E,
                                                              double2* __restrict__ part) {
    __shared__ double red[2 * (kDmBlock / 64)];
    const ElemDesc d = elems[blockIdx.y];
    const uint32_t base = blockIdx.x * kDmTile;
    const uint32_t M2 = M * M;
    double2 acc = make_double2(0, 0);
    for (int r = 0; r < kDmCPT; r++) {
        const uint32_t i = base + threadIdx.x + r * kDmBlock;
        if (i >= ctx.dim) break;
        uint16_t occ[kMaxOrb];
        for (int k = 0; k < M; k++) occ[k] = ctx.ht[(size_t) i * M + k];
        const double2 ci = c[i];
        const double w = ci.x * ci.x + ci.y * ci.y;
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
        const int k = d.k;
        switch (d.type) {
            case 0: acc.x += w * occ[k]; break;
            case 1: acc.x += w * (double) (occ[k] * (occ[k] - 1)); break;
            case 2: acc.x += w * (double) (occ[k] * occ[d.s]); break;
            case 3:
            case 4:
                if (occ[k] >= 1) {
                    const uint32_t j = ctx.map[i + k * ctx.dim + d.l * M * ctx.dim];
                    double2 t = cscale(cconjmul(ci, c[j]), sqrt((double) occ[k] * (occ[d.l] + 1)));
                    if (d.type == 4) t = cscale(t, (double) occ[d.m]);
                    acc = acc + t;
                }
                break;
            case 5:
                if (occ[k] >= 2) {
                    const uint32_t j = ctx.mapot[ctx.sot[i] + otc(k) * M2 + d.q + d.q * M];
                    acc = acc + cscale(cconjmul(ci, c[j]),
                                       sqrt((double) (occ[k] - 1) * occ[k] * (occ[d.q] + 1) * (occ[d.q] + 2)));
                }
                break;
            case 6:
                if (occ[k] >= 2) {
                    const uint32_t j = ctx.mapot[ctx.sot[i] + otc(k) * M2 + d.q + d.l * M];
                    acc = acc + cscale(cconjmul(ci, c[j]),
                                       sqrt((double) occ[k] * (occ[k] - 1) * (occ[d.q] + 1) * (occ[d.l] + 1)));
                }
                break;
            case 7:
                if (occ[k] >= 1 && occ[d.s] >= 1) {
                    const uint32_t j = ctx.maptt[ctx.stt[i] + ttc(k, d.s) * M2 + d.q + d.l * M];
                    acc = acc + cscale(cconjmul(ci, c[j]),
                                       sqrt((double) occ[k] * occ[d.s] * (occ[d.q] + 1) * (occ[d.l] + 1)));
                }
                break;
        }
    }
    double v[2] = {acc.x, acc.y};
    block_sum<2>(v, red);
    if (threadIdx.x == 0) part[(size_t) blockIdx.x * E + blockIdx.y] = make_double2(v[0], v[1]);
}

// -------------------------------------------------- second stage + fill

__global__ __launch_bounds__(256) void k_dm_reduce(const double2* __restrict__ part, int nblocks, int E,
                                                   double2* __restrict__ tot) {
    __shared__ double red[2 * 4];
    const int e = blockIdx.x;
    double v[2] = {0, 0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        const double2 p = part[(size_t) b * E + e];
        v[0] += p.x;
        v[1] += p.y;
    }
    block_sum<2>(v, red);
    if (threadIdx.x == 0) tot[e] = make_double2(v[0], v[1]);
}

__device__ __forceinline__ double2 dconj(double2 a) { return make_double2(a.x, -a.y); }

// Symmetric fill of reference density_matrices.c:23-50 and :76-442.  Every
// element of rho1 and rho2 is written exactly once, so the fill is spread
// over the threads of one block with no ordering: work item t is a diagonal
// index k (t < M), an orbital pair k < l (Rules 1-5, 7 and 3), a (k, q < l)
// triple of Rule 6 or a (k < s, q < l) quadruple of Rule 8.
__device__ __forceinline__ void dm_assemble_block(int M, const double2* __restrict__ tot, double2* __restrict__ rho1,
                                                  double2* __restrict__ rho2) {
    const DmLayout L{M};
    const int M2 = M * M, M3 = M2 * M, np = c2(M);
    auto R2 = [&](int a, int b, int cc, int d) -> double2& { return rho2[a + b * M + cc * M2 + d * M3]; };
    auto pair_of = [&](int p, int& a, int& b) {
        a = 0;
        while (p >= M - 1 - a) p -= M - 1 - a++;
        b = a + 1 + p;
    };
    const int n_diag = M, n_pair = np, n_r6 = rho2 ? M * np : 0, n_r8 = rho2 ? np * np : 0;
    for (int t = threadIdx.x; t < n_diag + n_pair + n_r6 + n_r8; t += blockDim.x) {
        if (t < n_diag) {
            const int k = t;
            rho1[k * M + k] = make_double2(tot[L.d1() + k].x, 0);
            if (rho2) R2(k, k, k, k) = make_double2(tot[L.d2() + k].x, 0);
        } else if (t < n_diag + n_pair) {
            const int p = t - n_diag;
            int k, l;
            pair_of(p, k, l);
            const double2* T = tot + L.pair() + p * (M + 1);
            rho1[k * M + l] = T[0];
            rho1[l * M + k] = dconj(T[0]);
            if (!rho2) continue;
            const double2 d3 = make_double2(tot[L.d3() + p].x, 0);  // Rule 2
            R2(k, l, k, l) = d3;
            R2(l, k, k, l) = d3;
            R2(l, k, l, k) = d3;
            R2(k, l, l, k) = d3;
            const double2 r4 = T[1 + k] - T[0];  // Rule 4
            R2(k, k, k, l) = r4;
            R2(k, k, l, k) = r4;
            R2(l, k, k, k) = dconj(r4);
            R2(k, l, k, k) = dconj(r4);
            const double2 r5 = T[1 + l];  // Rule 5
            R2(k, l, l, l) = r5;
            R2(l, k, l, l) = r5;
            R2(l, l, l, k) = dconj(r5);
            R2(l, l, k, l) = dconj(r5);
            for (int s = 0; s < M; s++) {  // Rule 7 (element k s s l)
                if (s == k || s == l) continue;
                const double2 r7 = T[1 + s];
                R2(k, s, s, l) = r7;
                R2(s, k, s, l) = r7;
                R2(s, k, l, s) = r7;
                R2(k, s, l, s) = r7;
                R2(l, s, s, k) = dconj(r7);
                R2(s, l, s, k) = dconj(r7);
                R2(s, l, k, s) = dconj(r7);
                R2(l, s, k, s) = dconj(r7);
            }
            const double2 r3 = tot[L.o3() + p];  // Rule 3 (k < q = l)
            R2(k, k, l, l) = r3;
            R2(l, l, k, k) = dconj(r3);
        } else if (t < n_diag + n_pair + n_r6) {  // Rule 6
            const int u = t - n_diag - n_pair, k = u / np;
            int q, l;
            pair_of(u % np, q, l);
            if (q == k || l == k) continue;
            const double2 r = tot[L.o6() + k * c2(M - 1) + L.pidx_excl1(k, q, l)];
            R2(k, k, q, l) = r;
            R2(k, k, l, q) = r;
            R2(l, q, k, k) = dconj(r);
            R2(q, l, k, k) = dconj(r);
        } else {  // Rule 8
            const int u = t - n_diag - n_pair - n_r6;
            int k, s, q, l;
            pair_of(u / np, k, s);
            pair_of(u % np, q, l);
            if (q == k || q == s || l == k || l == s) continue;
            const double2 r = tot[L.t8() + L.pidx(k, s) * c2(M - 2) + L.pidx_excl2(k, s, q, l)];
            R2(k, s, q, l) = r;
            R2(s, k, q, l) = r;
            R2(k, s, l, q) = r;
            R2(s, k, l, q) = r;
        }
    }
}

__global__ __launch_bounds__(256) void k_dm_assemble(int M, const double2* __restrict__ tot,
                                                     double2* __restrict__ rho1, double2* __restrict__ rho2) {
    dm_assemble_block(M, tot, rho1, rho2);
}

// With fill set (few configuration blocks): the second-stage sums (each
// element by one thread, over the blocks in order) and the fill; then, with
// sy.rinv set, the density synchronisation of rho1, all in one block.
__device__ __forceinline__ void dm_finish_block(double2* work, int M, const double2* __restrict__ part, int nblocks,
                                                int E, double2* __restrict__ tot, double2* __restrict__ rho1,
                                                double2* __restrict__ rho2, const DmSync& sy, int fill) {
    if (fill) {
        for (int e = threadIdx.x; e < E; e += blockDim.x) {
            double2 v = make_double2(0, 0);
            for (int b = 0; b < nblocks; b++) v = v + part[(size_t) b * E + e];
            tot[e] = v;
        }
        __syncthreads();
        dm_assemble_block(M, tot, rho1, rho2);
    }
    if (sy.rinv) {
        __syncthreads();
        density_sync_body(M, rho1, sy.rinv, sy.eig, sy.reg, work);
    }
}

__global__ __launch_bounds__(256) void k_dm_finish(int M, const double2* __restrict__ part, int nblocks, int E,
                                                   double2* __restrict__ tot, double2* __restrict__ rho1,
                                                   double2* __restrict__ rho2, DmSync sy, int fill) {
    extern __shared__ double2 work[];  // 4 M^2 for the synchronisation
    dm_finish_block(work, M, part, nblocks, E, tot, rho1, rho2, sy, fill);
}

// Packed form: one block per instance, the fill and (with sy.rinv) the
// synchronisation.
__global__ __launch_bounds__(256) void k_dm_finish_pk(const PkDmFin* __restrict__ tab, int M, int E) {
    extern __shared__ double2 work[];
    const PkDmFin& a = tab[blockIdx.x];
    dm_finish_block(work, M, a.part, a.nblocks, E, a.tot, a.rho1, a.rho2, a.sy, 1);
}

int pk_dm_finish(const PkDmFin* tab, int count, int M, size_t smem, hipStream_t st) {
    k_dm_finish_pk<<<count, 256, smem, st>>>(tab, M, DmLayout{M}.total());
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

int pk_density(const PkDm* tab, int count, int M, unsigned max_grid, size_t smem, hipStream_t st) {
    switch (M) {
        case 2: return DmFused<2>::run_pk(tab, count, max_grid, smem, st);
        case 3: return DmFused<3>::run_pk(tab, count, max_grid, smem, st);
        case 4: return DmFused<4>::run_pk(tab, count, max_grid, smem, st);
        case 5: return DmFused<5>::run_pk(tab, count, max_grid, smem, st);
        case 6: return DmFused<6>::run_pk(tab, count, max_grid, smem, st);
        case 7: return DmFused<7>::run_pk(tab, count, max_grid, smem, st);
        default: set_error("packed density kernels cover 2 <= norb <= 7"); return kInvalidArgument;
    }
}

unsigned pk_density_grid(int M, unsigned ntiles) {
    switch (M) {
        case 2: return dm_family_grid(ntiles, Passes<2>::count());
        case 3: return dm_family_grid(ntiles, Passes<3>::count());
        case 4: return dm_family_grid(ntiles, Passes<4>::count());
        case 5: return dm_family_grid(ntiles, Passes<5>::count());
        case 6: return dm_family_grid(ntiles, Passes<6>::count());
        case 7: return dm_family_grid(ntiles, Passes<7>::count());
        default: return 0;
    }
}

size_t dm_family_smem(const Space* S, size_t rsz) {
    const int M = S->norb;
    // block sums (as k_density lays them out), the W table, the sqrt table
    size_t smem = (size_t) 2 * (2 * M + c2(M) + 8 * 9 + 64) * (kDmBlock / 64) * sizeof(double);
    smem += ((size_t) (M - 1) * S->ldw * 4 + 15) & ~(size_t) 15;
    if (S->npar + 3 <= 2048) smem += rsz * (S->npar + 3);
    return smem > 64 * 1024 ? 0 : smem;
}

// instantiated once each in dm_m<M>.hip
extern template struct DmFused<2>;
extern template struct DmFused<3>;
extern template struct DmFused<4>;
extern template struct DmFused<5>;
extern template struct DmFused<6>;
extern template struct DmFused<7>;

static int dispatch_dm(const Space* S, const void* c, int prec, int index, double2* part, const DmPlan& plan,
                       hipStream_t st) {
    switch (S->norb) {
        case 2: return DmFused<2>::run(S, c, prec, index, part, plan, st);
        case 3: return DmFused<3>::run(S, c, prec, index, part, plan, st);
        case 4: return DmFused<4>::run(S, c, prec, index, part, plan, st);
        case 5: return DmFused<5>::run(S, c, prec, index, part, plan, st);
        case 6: return DmFused<6>::run(S, c, prec, index, part, plan, st);
        case 7: return DmFused<7>::run(S, c, prec, index, part, plan, st);
        default: set_error("fused density kernels cover 2 <= norb <= 7"); return kInvalidArgument;
    }
}

static bool wave_fits(const Space* S, int prec, int index) {
    switch (S->norb) {
        case 2: return DmFused<2>::wave_fits(S, prec, index);
        case 3: return DmFused<3>::wave_fits(S, prec, index);
        case 4: return DmFused<4>::wave_fits(S, prec, index);
        case 5: return DmFused<5>::wave_fits(S, prec, index);
        case 6: return DmFused<6>::wave_fits(S, prec, index);
        case 7: return DmFused<7>::wave_fits(S, prec, index);
        default: return false;
    }
}

DmPlan dm_plan(const Space* S, int prec, int index) {
    const int M = S->norb;
    const bool generic = M == 1 || M > 7;
    if (!generic && S->dim >= kDmWaveMinDim && wave_fits(S, prec, index)) {
        // at most about 4096 blocks: enough waves to hide the gathers'
        // latency, and a bounded set of partial sums
        const int cpt = (int) std::min<uint32_t>(16, div_up(S->dim, (uint32_t) kDmBlock * 4096u));
        return {true, cpt, (unsigned) div_up(S->dim, (uint32_t) cpt * kDmBlock)};
    }
    return {false, kDmCPT, (unsigned) div_up(S->dim, kDmTile)};
}

size_t dm_workspace_bytes(const Space* S) {
    const size_t E = (size_t) DmLayout{S->norb}.total();
    // the largest block count of any form
    size_t nb = div_up(S->dim, kDmTile);
    for (int prec = 0; prec < 2; prec++)
        for (int index = 0; index < 2; index++) nb = std::max<size_t>(nb, dm_plan(S, prec, index).nb);
    return (nb * E + E) * sizeof(double2) + sizeof(ElemDesc) * E + 256;
}

int density_matrices(const Space* S, const void* coef, double2* rho1, double2* rho2, int prec, int index,
                     void* work, hipStream_t st, const DmSync* sync) {
    if (sync && sync->rinv && S->norb > 30) {
        set_error("norb exceeds 30");
        return kInvalidArgument;
    }
    const int M = S->norb;
    const bool generic = M == 1 || M > 7;
    if (prec == kDmMixed && generic) {
        set_error("mixed precision density matrices cover 2 <= norb <= 8");
        return kInvalidArgument;
    }
    if ((generic || index == kIdxTables) && !S->tables) {
        set_error("this density-matrix form needs the jump tables; build the space with tables");
        return kNotBuilt;
    }
    const int E = DmLayout{M}.total();
    const DmPlan plan = dm_plan(S, prec, index);
    const unsigned nb = plan.nb;
    double2* part = (double2*) work;
    double2* tot = part + (size_t) nb * E;
    ElemDesc* elems = (ElemDesc*) (tot + E);
    int rc = kOk;
    if (generic) {
        std::vector<ElemDesc> el = element_list(M);
        MCTDHB_HIP_CHECK(hipMemcpyAsync(elems, el.data(), sizeof(ElemDesc) * E, hipMemcpyHostToDevice, st));
        RowCtx ctx = make_row_ctx(S->dim, S->npar, S->d_ht, S->d_map, S->d_mapot, S->d_maptt, S->d_strideot,
                                  S->d_stridett);
        k_density_generic<<<dim3(nb, E), kDmBlock, 0, st>>>(M, ctx, (const double2*) coef, elems, E, part);
        MCTDHB_HIP_CHECK(hipGetLastError());
        // the host vector must outlive the async copy
        MCTDHB_HIP_CHECK(hipStreamSynchronize(st));
    } else {
        rc = dispatch_dm(S, coef, prec, index, part, plan, st);
    }
    if (rc) return rc;
    const DmSync sy = sync ? *sync : DmSync{nullptr, nullptr, 0.0};
    if (nb <= kDmFinishBlocks) {
        k_dm_finish<<<1, 256, sy.rinv ? sizeof(double2) * 4 * M * M : 0, st>>>(M, part, (int) nb, E, tot, rho1, rho2,
                                                                               sy, 1);
        MCTDHB_HIP_CHECK(hipGetLastError());
        return kOk;
    }
    k_dm_reduce<<<E, 256, 0, st>>>(part, (int) nb, E, tot);
    k_dm_assemble<<<1, 256, 0, st>>>(M, tot, rho1, rho2);
    MCTDHB_HIP_CHECK(hipGetLastError());
    if (sy.rinv) {
        k_dm_finish<<<1, 256, sizeof(double2) * 4 * M * M, st>>>(M, nullptr, 0, 0, nullptr, rho1, nullptr, sy, 0);
        MCTDHB_HIP_CHECK(hipGetLastError());
    }
    return kOk;
}

}  // namespace mctdhb
