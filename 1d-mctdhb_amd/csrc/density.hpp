// density.hpp - rho1 / rho2 block reductions over configurations.
#pragma once

#include "space.hpp"

namespace mctdhb {

enum : int { kDmFP64 = 0, kDmMixed = 1 };

// Scratch needed by density_matrices (per-block partials and totals).
size_t dm_workspace_bytes(const Space* S);
int dm_num_elements(int M);

// rho1 (M x M row-major, rho1[k*M+l] = <a+_k a_l>) and, if rho2 is non-null,
// rho2 (M^4, index k + sM + qM^2 + lM^3 = <a+_k a+_s a_q a_l>), complex128.
// coef is complex128 (prec = kDmFP64) or complex64 (kDmMixed); index is
// kIdxTables or kIdxFly.  work must hold dm_workspace_bytes(S).
// Launch plan of the fused kernels: the one-pass wave form (tiles of
// cpt x 256 configurations) for large spaces of M >= 6, else the
// element-family passes (tiles of 16 x 256, one family per block); nb blocks
// of partial sums either way.  Measured on MI355X: the wave form takes
// 1.87 ms at N=30, M=7 against 3.7 for the passes, and the passes 0.66 ms at
// N=100, M=5 against 1.19 (at M <= 5 the wave sums cost more than they save).
struct DmPlan {
    bool wave;
    int cpt;
    unsigned nb;
};
DmPlan dm_plan(const Space* S, int prec, int index);
// Spaces from this many configurations and orbitals take the wave form.
constexpr uint32_t kDmWaveMinDim = 1u << 16;
constexpr int kDmWaveMinOrb = 6;

// Per-orbital-count fused kernels (dm_m<M>.hip): write per-block partials.
struct PkDm;
template <int M> struct DmFused {
    static bool wave_fits(const Space* S, int prec, int index);
    static int run(const Space* S, const void* c, int prec, int index, double2* part, const DmPlan& plan,
                   hipStream_t st);
    // FP64, on-the-fly indices, element-family form, count instances
    // (pack.hpp); max_grid: the largest per-instance family grid
    static int run_pk(const PkDm* tab, int count, unsigned max_grid, size_t smem, hipStream_t st);
};
// LDS of the element-family kernel (FP64 or mixed by rsz), or 0 where it
// exceeds the 64 KiB budget.
size_t dm_family_smem(const Space* S, size_t rsz);

// With sync->rinv set, also the synchronisation of rho1 (eigenvalues into
// sync->eig when set, regularisation by sync->reg, rinv = rho1^-1), as the
// standalone density_sync_small does.
struct DmSync {
    double2* rinv;
    double* eig;
    double reg;
};
// Up to this many configuration blocks, one block finishes the reduction.
constexpr unsigned kDmFinishBlocks = 64;
int density_matrices(const Space* S, const void* coef, double2* rho1, double2* rho2, int prec, int index,
                     void* work, hipStream_t st, const DmSync* sync = nullptr);

}  // namespace mctdhb
