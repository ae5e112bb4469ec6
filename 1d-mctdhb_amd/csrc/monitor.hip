// monitor.hip - convergence monitors of the imaginary-time integration.
//
// eig_residual (reference src/assistant/integrator_monitor.c:9-28):
// max_i |e c_i - (H c)_i| / N, the stopping test of the imaginary-time
// driver (driver.c:47-68) next to the energy.
#include "hamiltonian.hpp"
#include "monitor.hpp"

namespace mctdhb {

constexpr int kMonBlock = 256;
constexpr int kMonGrid = 1024;

// Block maxima of |e c_i - hc_i| (all values are >= 0, so the grid maximum
// does not depend on the order of the reduction).
__global__ __launch_bounds__(kMonBlock) void k_residual_max(const double2* __restrict__ c,
                                                            const double2* __restrict__ hc, double e, uint32_t n,
                                                            double* __restrict__ part) {
    __shared__ double red[kMonBlock / 64];
    double m = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double2 r = make_double2(e * c[i].x - hc[i].x, e * c[i].y - hc[i].y);
        m = fmax(m, sqrt(r.x * r.x + r.y * r.y));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kMonBlock / 64; w++) m = fmax(m, red[w]);
        part[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(kMonBlock) void k_max_final(const double* __restrict__ part, int np, double scale,
                                                         double* __restrict__ out) {
    __shared__ double red[kMonBlock / 64];
    double m = 0;
    for (int b = threadIdx.x; b < np; b += blockDim.x) m = fmax(m, part[b]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kMonBlock / 64; w++) m = fmax(m, red[w]);
        out[0] = m * scale;
    }
}

int eig_residual_dev(const Space* S, const double2* c, const double2* hob, const double2* hint, double e,
                     double2* hc_work, double* part_work, double* result, hipStream_t st) {
    int rc = hc_apply(S, c, hob, hint, hc_work, kPrecFP64, kIndexOnTheFly, st);
    if (rc) return rc;
    k_residual_max<<<kMonGrid, kMonBlock, 0, st>>>(c, hc_work, e, S->dim, part_work);
    k_max_final<<<1, kMonBlock, 0, st>>>(part_work, kMonGrid, 1.0 / S->npar, result);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

size_t eig_residual_part_doubles() { return kMonGrid + 1; }

}  // namespace mctdhb
