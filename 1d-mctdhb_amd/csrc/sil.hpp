// sil.hpp - device state of the short-iterative-Lanczos (SIL) step.
//
// Krylov vectors are kept unnormalised: v_k = s_k w_k with s_k = 1/||w_k||
// (s_0 = 1, w_0 = C), so no pass is spent rescaling a vector whose norm is
// only known after a grid-wide reduction.  Every scalar of the recurrence
// lives in this struct in device memory; kernels read it directly, so a
// whole step is enqueued without host round trips.
#pragma once

#include "common.hpp"

namespace mctdhb {

// Krylov capacity: the imaginary-time finish diagonalises with up to 50
// Lanczos vectors (reference driver.c:40-45, :300-316).
constexpr int kMaxKrylov = 64;
// reference MAX_LANCZOS_ITER (mctdhb_types.h:33): up to this many vectors the
// streaming kernels keep every coefficient in registers.
constexpr int kMaxLanczos = 15;

struct SilState {
    double s[kMaxKrylov + 1];        // scale of w_k
    double beta[kMaxKrylov + 1];     // offd (real, as carrMod returns)
    double2 alpha[kMaxKrylov + 1];   // diag (complex, as carrDot returns)
    double2 cre[kMaxKrylov + 1];     // reorthogonalisation coefficients (on w_k)
    double2 gram[kMaxKrylov + 1];    // <v_k, v_cur> for the current vector
    double2 y[kMaxKrylov + 1];       // recombination coefficients (on w_k)
    double maxbeta;
    double nrm2;                      // |C'|^2 of the recombined vector
    double eground;                   // lowest Ritz value (ground-state mode)
    unsigned int arrive;              // blocks done in the current reduce kernel
    int iter_done;
    int stop;                         // breakdown: later kernels are no-ops
    int eig_fail;
};

}  // namespace mctdhb
