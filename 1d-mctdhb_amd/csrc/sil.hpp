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

constexpr int kMaxLanczos = 15;  // reference MAX_LANCZOS_ITER (mctdhb_types.h:33)

struct SilState {
    double s[kMaxLanczos + 1];        // scale of w_k
    double beta[kMaxLanczos + 1];     // offd (real, as carrMod returns)
    double2 alpha[kMaxLanczos + 1];   // diag (complex, as carrDot returns)
    double2 cre[kMaxLanczos + 1];     // reorthogonalisation coefficients (on w_k)
    double2 gram[kMaxLanczos + 1];    // <v_k, v_cur> for the current vector
    double2 y[kMaxLanczos + 1];       // recombination coefficients (on w_k)
    double maxbeta;
    double nrm2;                      // |C'|^2 of the recombined vector
    double eground;                   // lowest Ritz value (ground-state mode)
    int iter_done;
    int stop;                         // breakdown: later kernels are no-ops
    int eig_fail;
};

}  // namespace mctdhb
