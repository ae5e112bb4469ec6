// orbital.hpp - the orbital split-step on the device.
#pragma once

#include "common.hpp"

namespace mctdhb {

constexpr int kZeroBounds = 1000;     // reference ZERO_BOUNDS
constexpr int kPeriodicBounds = 1001; // reference PERIODIC_BOUNDS

// The fields of reference _OrbitalEquation the orbital step reads
// (mctdhb_types.h:136-149).
struct OrbEq {
    int norb = 0, grid_size = 0, bounds = kPeriodicBounds;
    double dx = 0, g = 0, d2coef = -0.5, tstep = 0;
    double2 d1coef{0, 0}, prop_dt{0, 0}, time_fac{1, 0};
};

// Device workspace of the orbital step for one (norb, grid, dt) setup.
struct OrbWork {
    OrbEq eq;
    int imagtime = 0;
    int rk_order = 4;  // nonlinear sub-step: 2 (Heun), 4 (classical), 5 (Butcher, 6 stages)
    int fullstep = 0;  // whole-step orbital integrator: 0 split step, 1 full-equation RK
    int robust = 0;    // full-equation RK: projector with the overlap inverse
    int dvr = 0;       // full-equation RK: linear part by the exponential DVR
    int nfft = 0, log2n = -1;  // FFT length grid_size - 1 (log2n < 0: not a power of two)
    double* d_pot = nullptr;     // grid_size
    double2* d_fexp = nullptr;   // nfft: exp(-i dt/2 (i a1 k - a2 k^2))
    double2* d_tw = nullptr;     // nfft / 2 twiddles exp(-2 pi i t / nfft)
    double2* d_cn = nullptr;     // 3 * grid_size: upper | lower | mid
    double2* d_stage = nullptr;  // 7 * norb * grid_size: up to six stages and y
    double* d_red = nullptr;     // reduction scratch
    double2* d_cnwork = nullptr; // 10 * norb * grid_size: Crank-Nicolson solve scratch
    int cn_pcr = 0;              // the parallel (PCR) CN solve applies
    double2* d_nl = nullptr;     // grid-independent factors of the nonlinear term
    double2* d_fsH = nullptr;    // norb * grid_size: H y of the full-equation step
    double2* d_fsmat = nullptr;  // 3 norb^2: projections, Gram matrix, projector coefficients
    double2* d_hsym = nullptr;   // nfft: spectral symbol of the linear part
    double2* d_dvrT = nullptr;   // nfft^2: exponential-DVR matrix, transposed (built on first use)
};

int orbwork_create(const OrbEq& eq, int imagtime, const double* pot_host, OrbWork** out);
void orbwork_destroy(OrbWork* w);
// Refresh the potential (time-dependent traps).
int orbwork_set_potential(OrbWork* w, const double* pot_host);

// Linear half steps on orb (norb x grid_size complex128, device).
int linear_fft(OrbWork* w, double2* orb, hipStream_t st);
int linear_cn(OrbWork* w, double2* orb, hipStream_t st);
// hob (M x M) and hint (M^4) from the orbitals.
int orbital_matrices(OrbWork* w, const double2* orb, double2* hob, double2* hint, hipStream_t st);
// der = -i time_fac (nonlinear(orb) + V orb) with rho2, rinv (= rho1^-1).
int nonlinear_rhs(OrbWork* w, const double2* orb, const double2* hob, const double2* hint, const double2* rho2,
                  const double2* rinv, double2* der, hipStream_t st);
// orb[k][grid - 1] = orb[k][0] (orbital_integration.c:13-20).
int periodic_bounds(OrbWork* w, double2* orb, hipStream_t st);
// Gram-Schmidt with the reference's point-by-point projection update;
// periodic_tail also copies the first grid point onto the last.
int orthonormalize(OrbWork* w, double2* orb, hipStream_t st, bool periodic_tail = false);
// One split step (linear half, h/g rebuild, RK4 nonlinear, linear half,
// Gram-Schmidt in imaginary time); hob/hint are rebuilt from the orbitals.
// periodic_tail: leave the orbitals with the periodic copy of the first
// grid point on the last (with periodic bounds), as the whole step needs.
int splitstep(OrbWork* w, int use_fft, double2* orb, double2* hob, double2* hint, const double2* rho2,
              const double2* rinv, hipStream_t st, bool periodic_tail = false);
// Full-equation Runge-Kutta orbital step (orbital_integration.c:283-348):
// linear part by FFT (der_method 0), finite differences (1) or the
// exponential DVR (2), simple or robust projector, Gram-Schmidt in
// imaginary time.  rho2 and rinv are held fixed.
int fullstep(OrbWork* w, int der_method, int robust, double2* orb, const double2* rho2, const double2* rinv,
             hipStream_t st);
// Its right-hand side at orb (orbital_integration.c:158-235).
int fullstep_rhs(OrbWork* w, int der_method, int robust, const double2* orb, const double2* rho2,
                 const double2* rinv, double2* der, hipStream_t st);
// Builds the DVR matrix on first use (allocates and synchronises).
int ensure_dvr(OrbWork* w);
// proj_k = projection of hact_k onto the orbitals (orbital_integration.c:22-129).
int projector(OrbWork* w, int robust, const double2* orb, const double2* hact, double2* proj, hipStream_t st);
// rho1 -> eigenvalues, conditional regularisation, inverse (M x M, one wave,
// LDS work space: no allocation, so the call can be captured in a graph).
int density_sync_small(int norb, double2* rho1, double2* rinv, double* eigvals, double reg_factor,
                       hipStream_t st);

}  // namespace mctdhb
