// pack.hpp - the packed step: one launch per stage advances many instances
// by one whole MCTDHB step (reference driver.c:159-202, mctdhb_propagate_step,
// made re-entrant per instance).  Each stage kernel reads its arguments from
// a table in device memory, one entry per instance (blockIdx.y, or blockIdx.x
// for the one-workgroup stages), and runs the same block code as the
// single-instance kernel of that stage, so a packed step gives the results
// of the instance's own step bit for bit.
//
// A pack holds instances that share the stages' template parameters: the
// orbital count M, the Runge-Kutta stage count and the SIL kernel variant
// (index scheme, staged vector).  Grid lengths, particle numbers, pointers
// and step sizes differ per entry; blocks past an instance's own block count
// return at once.
#pragma once

#include "common.hpp"
#include "density.hpp"
#include "hamiltonian.hpp"
#include "rowctx.hpp"

namespace mctdhb {

// Linear half step by FFT (power-of-two length n = grid - 1): M blocks.
struct PkFft {
    double2* orb;
    const double2* tw;
    const double2* fexp;
    int grid, n, log2n;
};

// h and g from the orbitals: M^2 + M^4 blocks.
struct PkOm {
    const double2* orb;
    const double* V;
    double2* hob;
    double2* hint;
    double dx, a2, g;
    double2 a1;
    int n;
};

// Tables of the nonlinear term (orbital counts whose RK kernel does not
// form them itself): one block.
struct PkNlCoef {
    const double2* hob;
    const double2* hint;
    const double2* rho2;
    const double2* rinv;
    double2* out;
};

// Pointwise Runge-Kutta sub-step: div_up(n, 64) blocks of 64 M threads.
struct PkNl {
    double2* orb;
    const double* V;
    const double2* coef;
    const double2* hob;
    const double2* hint;
    const double2* rho2;
    const double2* rinv;
    double dt, g;
    double2 tf;
    int n;
};

// Gram-Schmidt (imaginary time): one block.
struct PkGs {
    double2* F;
    double dx;
    int n, tail, staged;
};

// FP64 coefficient block of the fused H.C row: one block.
struct PkCf {
    const double2* hob;
    const double2* hint;
    double2* cf64;
    int npar;
};

// One-workgroup SIL step.
struct PkSil {
    RowCtx ctx;
    const double2* cf;
    const uint32_t* W;
    int ldw;
    SmallSilArgs a;
};

// rho1/rho2 partial sums, one element family per block (grid: the
// instance's own padded family grid).
struct PkDm {
    RowCtx ctx;
    const double2* c;
    const uint32_t* W;
    double2* part;
    int ldw, use_sq;
    uint32_t ntiles, grid;
};

// Second-stage sums, fill and density synchronisation: one block.
struct PkDmFin {
    const double2* part;
    double2* tot;
    double2* rho1;
    double2* rho2;
    DmSync sy;
    int nblocks;
};

// Stage launchers over count table entries (device pointers).  They check
// nothing beyond what the table builder guarantees.
int pk_linear_fft(const PkFft* tab, int count, int M, size_t smem, hipStream_t st);
int pk_orbital_matrices(const PkOm* tab, int count, int M, hipStream_t st);
int pk_nl_coef(const PkNlCoef* tab, int count, int M, hipStream_t st);
// ns: 2, 4 or 6 stages (the tableau of the orbital work's RK order)
int pk_nl_rk(const PkNl* tab, int count, int M, int ns, unsigned max_blocks, hipStream_t st);
// Whether the RK kernel forms the nonlinear tables itself (no pk_nl_coef).
bool pk_nl_folds(int M, int ns);
int pk_gram_schmidt(const PkGs* tab, int count, int M, size_t smem, hipStream_t st);
int pk_hc_coefficients(const PkCf* tab, int count, int M, hipStream_t st);
int pk_sil_small(const PkSil* tab, int count, int M, int index, bool stage, size_t smem, hipStream_t st);
int pk_density(const PkDm* tab, int count, int M, unsigned max_grid, size_t smem, hipStream_t st);
// The family grid of an instance with ntiles configuration tiles (0: M not covered).
unsigned pk_density_grid(int M, unsigned ntiles);
// LDS of the Gram-Schmidt block: the staged orbitals, or 0 (not staged).
size_t pk_gs_smem(int M, int n);
int pk_dm_finish(const PkDmFin* tab, int count, int M, size_t smem, hipStream_t st);

}  // namespace mctdhb
