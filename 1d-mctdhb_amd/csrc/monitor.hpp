// monitor.hpp - convergence monitors on the device.
#pragma once

#include "space.hpp"

namespace mctdhb {

// result[0] (device) <- max_i |e c_i - (H c)_i| / N.  hc_work: dim complex128;
// part_work: eig_residual_part_doubles() doubles.
int eig_residual_dev(const Space* S, const double2* c, const double2* hob, const double2* hint, double e,
                     double2* hc_work, double* part_work, double* result, hipStream_t st);
size_t eig_residual_part_doubles();

}  // namespace mctdhb
