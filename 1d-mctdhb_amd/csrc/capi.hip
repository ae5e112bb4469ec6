// capi.hip - the C ABI of libmctdhb_hip.so (include/mctdhb_hip.h).
//
// Layer 1 keeps the reference's names, signatures and fail-stop error model
// (src/configurational/*.c, src/linalg/multiconfig_lanczos.c,
// src/integrator/coefficients_integration.c, src/assistant/types_definition.c)
// on host arrays, moving the work to the GPU.  Layer 2 (mctdhb_*) is the
// device-resident, status-returning API the drop-in layer is built on.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "density.hpp"
#include "hamiltonian.hpp"
#include "monitor.hpp"
#include "mctdhb_hip.h"
#include "orbital.hpp"
#include "pack.hpp"
#include "sil_api.hpp"

namespace mctdhb {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }

}  // namespace mctdhb

using namespace mctdhb;

struct mctdhb_space {
    Space* S = nullptr;
    void* dm_work = nullptr;
    // scratch of the reference-named host entry points
    double2* d_coef = nullptr;
    double2* d_out = nullptr;
    double2* d_hob = nullptr;
    double2* d_hint = nullptr;
    double2* d_rho1 = nullptr;
    double2* d_rho2 = nullptr;
    double2* d_lvec = nullptr;
    mctdhb_sil* sil = nullptr;  // cached SIL workspace of the host entry points
};

struct mctdhb_sil {
    Sil* L = nullptr;
    const mctdhb_space* sp = nullptr;
};

// Process-wide regularisation factor of the density sync of a whole step
// (reference synchronize.c:9, DEFAULT_REGULARIZATION_FACTOR of synchronize.h:23).
static double g_reg_factor = 1e-9;

// ------------------------------------------------------------------ layer 2

extern "C" const char* mctdhb_last_error(void) { return last_error(); }

extern "C" int mctdhb_device_count(int* count) {
    MCTDHB_HIP_CHECK(hipGetDeviceCount(count));
    return kOk;
}

extern "C" int mctdhb_space_create(uint16_t npar, uint16_t norb, int flags, mctdhb_space** out) {
    if (!out) {
        set_error("null output pointer");
        return kInvalidArgument;
    }
    *out = nullptr;
    Space* S = nullptr;
    int rc = space_create(npar, norb, flags, &S);
    if (rc) return rc;
    // the fixed scratch of the no-allocation H.C on this device
    if ((rc = hc_scratch_setup())) {
        space_destroy(S);
        return rc;
    }
    mctdhb_space* sp = new mctdhb_space();
    sp->S = S;
    *out = sp;
    return kOk;
}

static void free_scratch(mctdhb_space* sp) {
    hipFree(sp->dm_work);
    hipFree(sp->d_coef);
    hipFree(sp->d_out);
    hipFree(sp->d_hob);
    hipFree(sp->d_hint);
    hipFree(sp->d_rho1);
    hipFree(sp->d_rho2);
    hipFree(sp->d_lvec);
    if (sp->sil) {
        sil_destroy(sp->sil->L);
        delete sp->sil;
    }
}

extern "C" int mctdhb_space_destroy(mctdhb_space* sp) {
    if (!sp) return kOk;
    hipDeviceSynchronize();
    free_scratch(sp);
    space_destroy(sp->S);
    delete sp;
    return kOk;
}

extern "C" int mctdhb_space_info(const mctdhb_space* sp, uint32_t* dim, uint64_t* mapot_len, uint64_t* maptt_len,
                                 int* has_tables) {
    if (!sp) {
        set_error("null space");
        return kInvalidArgument;
    }
    if (dim) *dim = sp->S->dim;
    if (mapot_len) *mapot_len = sp->S->mapot_len;
    if (maptt_len) *maptt_len = sp->S->maptt_len;
    if (has_tables) *has_tables = sp->S->tables ? 1 : 0;
    return kOk;
}

extern "C" int mctdhb_space_tables(const mctdhb_space* sp, const uint16_t** ht, const uint32_t** subsdim,
                                   const uint32_t** map, const uint32_t** mapot, const uint32_t** maptt,
                                   const uint32_t** strideot, const uint32_t** stridett) {
    if (!sp) {
        set_error("null space");
        return kInvalidArgument;
    }
    const Space* S = sp->S;
    if (ht) *ht = S->d_ht;
    if (subsdim) *subsdim = S->d_subsdim;
    if (map) *map = S->d_map;
    if (mapot) *mapot = S->d_mapot;
    if (maptt) *maptt = S->d_maptt;
    if (strideot) *strideot = S->d_strideot;
    if (stridett) *stridett = S->d_stridett;
    return kOk;
}

extern "C" int mctdhb_apply_hamiltonian_dev(const mctdhb_space* sp, const void* coef, const void* hob,
                                            const void* hint, void* hc, int prec, int index, void* stream) {
    if (!sp || !coef || !hob || !hint || !hc) {
        set_error("null argument");
        return kInvalidArgument;
    }
    return hc_apply(sp->S, coef, (const double2*) hob, (const double2*) hint, hc, prec, index,
                    (hipStream_t) stream);
}

extern "C" int mctdhb_apply_hamiltonian_noalloc_dev(const mctdhb_space* sp, const void* coef, const void* hob,
                                                    const void* hint, void* hc, int prec, int index, void* stream) {
    if (!sp || !coef || !hob || !hint || !hc) {
        set_error("null argument");
        return kInvalidArgument;
    }
    return hc_apply_noalloc(sp->S, coef, (const double2*) hob, (const double2*) hint, hc, prec, index,
                            (hipStream_t) stream);
}

extern "C" size_t mctdhb_hc_coef_bytes(const mctdhb_space* sp, int prec) {
    return sp ? hc_coef_bytes(sp->S, prec) : 0;
}

extern "C" int mctdhb_hc_coefficients_dev(const mctdhb_space* sp, const void* hob, const void* hint, void* cf,
                                          int prec, void* stream) {
    if (!sp || !hob || !hint || !cf) {
        set_error("null argument");
        return kInvalidArgument;
    }
    return hc_build_coef(sp->S, (const double2*) hob, (const double2*) hint, cf, prec, (hipStream_t) stream);
}

extern "C" int mctdhb_apply_hamiltonian_cf_dev(const mctdhb_space* sp, const void* coef, const void* cf, void* hc,
                                               int prec, int index, void* stream) {
    if (!sp || !coef || !cf || !hc) {
        set_error("null argument");
        return kInvalidArgument;
    }
    return hc_apply_cf(sp->S, coef, cf, hc, prec, index, (hipStream_t) stream);
}

extern "C" int mctdhb_hc_tiling(mctdhb_space* sp, int set, int* tiling) {
    if (!sp) {
        set_error("null space");
        return kInvalidArgument;
    }
    if (set < -2 || set > 2) {
        set_error("tiling must be -2 (query), -1 (time again), 0 (one configuration per lane), 1 (pairs) or 2 "
                  "(coefficient windows)");
        return kInvalidArgument;
    }
    if (set == kTilingPair && !pair_possible(sp->S)) {
        set_error("this space has no pair work list (needs 2 <= norb <= 7 and dim >= 2)");
        return kInvalidArgument;
    }
    if (set == kTilingWindow && !hc_window_possible(sp->S)) {
        set_error("the windowed H.C needs 2 <= norb <= 5 and windows (256 + 4 N + 6 entries) within its LDS budget");
        return kInvalidArgument;
    }
    if (set >= 0) sp->S->hc_tiling = set;
    if (set == -1) {
        hc_forget_tiling(sp->S);
        const int rc = hc_tune_tiling(sp->S, nullptr);
        if (rc) return rc;
    }
    if (tiling) *tiling = sp->S->hc_tiling;
    return kOk;
}

extern "C" int mctdhb_density_dev(mctdhb_space* sp, const void* coef, void* rho1, void* rho2, int prec, int index,
                                  void* stream) {
    if (!sp || !coef || !rho1) {
        set_error("null argument");
        return kInvalidArgument;
    }
    if (prec != kDmFP64 && prec != kDmMixed) {
        set_error("unknown precision mode");
        return kInvalidArgument;
    }
    if (!sp->dm_work) MCTDHB_HIP_CHECK(hipMalloc(&sp->dm_work, dm_workspace_bytes(sp->S)));
    return density_matrices(sp->S, coef, (double2*) rho1, (double2*) rho2, prec,
                            prec == kDmMixed ? (int) kIndexOnTheFly : index, sp->dm_work, (hipStream_t) stream);
}

extern "C" int mctdhb_sil_create(const mctdhb_space* sp, int lanczos_iter, int prec, int index, mctdhb_sil** out) {
    if (!sp || !out) {
        set_error("null argument");
        return kInvalidArgument;
    }
    *out = nullptr;
    Sil* L = nullptr;
    int rc = sil_create(sp->S, lanczos_iter, prec, index, &L);
    if (rc) return rc;
    mctdhb_sil* s = new mctdhb_sil();
    s->L = L;
    s->sp = sp;
    *out = s;
    return kOk;
}

extern "C" int mctdhb_sil_destroy(mctdhb_sil* s) {
    if (!s) return kOk;
    hipDeviceSynchronize();
    sil_destroy(s->L);
    delete s;
    return kOk;
}

extern "C" int mctdhb_sil_set_small(mctdhb_sil* s, int on) {
    if (!s) {
        set_error("null argument");
        return kInvalidArgument;
    }
    s->L->small = on != 0;
    return kOk;
}

extern "C" size_t mctdhb_sil_device_bytes(const mctdhb_space* sp, int lanczos_iter, int prec) {
    return sp ? sil_device_bytes(sp->S, lanczos_iter, prec) : 0;
}

extern "C" int mctdhb_sil_step_dev(mctdhb_sil* s, void* coef, const void* hob, const void* hint, double dt_re,
                                   double dt_im, int imagtime, void* stream) {
    if (!s || !coef || !hob || !hint) {
        set_error("null argument");
        return kInvalidArgument;
    }
    return sil_step(s->L, (double2*) coef, (const double2*) hob, (const double2*) hint, dt_re, dt_im, imagtime,
                    (hipStream_t) stream);
}

extern "C" int mctdhb_lanczos_dev(mctdhb_sil* s, const void* coef, const void* hob, const void* hint, void* lvec,
                                  void* stream) {
    if (!s || !coef || !hob || !hint || !lvec) {
        set_error("null argument");
        return kInvalidArgument;
    }
    return sil_lanczos(s->L, (const double2*) coef, (const double2*) hob, (const double2*) hint, (double2*) lvec,
                       (hipStream_t) stream);
}

extern "C" int mctdhb_lowest_state_dev(mctdhb_sil* s, void* coef, const void* hob, const void* hint,
                                       double* energy, void* stream) {
    if (!s || !coef || !hob || !hint) {
        set_error("null argument");
        return kInvalidArgument;
    }
    int rc = sil_ground(s->L, (double2*) coef, (const double2*) hob, (const double2*) hint, (hipStream_t) stream);
    if (rc || !energy) return rc;
    SilState h;
    rc = sil_read_state(s->L, &h, (hipStream_t) stream);
    if (rc) return rc;
    if (h.eig_fail) {
        set_error("tridiagonal eigensolver did not converge");
        return kNumerical;
    }
    *energy = h.eground;
    return kOk;
}

extern "C" int mctdhb_eig_residual_dev(mctdhb_space* sp, const void* coef, const void* hob, const void* hint,
                                       double e, double* result, void* stream) {
    if (!sp || !coef || !hob || !hint || !result) {
        set_error("null argument");
        return kInvalidArgument;
    }
    hipStream_t st = (hipStream_t) stream;
    double2* hc = nullptr;
    double* part = nullptr;
    MCTDHB_HIP_CHECK(hipMallocAsync((void**) &hc, sizeof(double2) * sp->S->dim, st));
    MCTDHB_HIP_CHECK(hipMallocAsync((void**) &part, sizeof(double) * eig_residual_part_doubles(), st));
    int rc = eig_residual_dev(sp->S, (const double2*) coef, (const double2*) hob, (const double2*) hint, e, hc, part,
                              part + eig_residual_part_doubles() - 1, st);
    if (!rc && hipMemcpyAsync(result, part + eig_residual_part_doubles() - 1, sizeof(double), hipMemcpyDeviceToHost,
                              st) != hipSuccess) {
        set_error("device to host copy failed");
        rc = kDeviceError;
    }
    hipFreeAsync(hc, st);
    hipFreeAsync(part, st);
    if (!rc) MCTDHB_HIP_CHECK(hipStreamSynchronize(st));
    return rc;
}

extern "C" int mctdhb_sil_scalars(const mctdhb_sil* s, int* iter_done, double* diag, double* offd, void* stream) {
    if (!s) {
        set_error("null argument");
        return kInvalidArgument;
    }
    SilState h;
    int rc = sil_read_state(s->L, &h, (hipStream_t) stream);
    if (rc) return rc;
    if (h.eig_fail) {
        set_error("tridiagonal eigensolver did not converge");
        return kNumerical;
    }
    if (iter_done) *iter_done = h.iter_done;
    for (int k = 0; k < s->L->lm; k++) {
        if (diag) {
            diag[2 * k] = h.alpha[k].x;
            diag[2 * k + 1] = h.alpha[k].y;
        }
        if (offd) offd[k] = h.beta[k];
    }
    return kOk;
}

// ------------------------------------------------------------- orbitals

struct mctdhb_orbital {
    OrbWork* w = nullptr;
};

extern "C" int mctdhb_orbital_create(int norb, int grid_size, int bounds, double xi, double xf, double tstep,
                                     double g, double d2coef, double d1_re, double d1_im, int imagtime,
                                     const double* pot, mctdhb_orbital** out) {
    if (!out || !pot) {
        set_error("null argument");
        return kInvalidArgument;
    }
    *out = nullptr;
    if (bounds != kZeroBounds && bounds != kPeriodicBounds) {
        set_error("bounds must be ZERO_BOUNDS (1000) or PERIODIC_BOUNDS (1001)");
        return kInvalidArgument;
    }
    OrbEq eq;
    eq.norb = norb;
    eq.grid_size = grid_size;
    eq.bounds = bounds;
    eq.dx = (xf - xi) / (grid_size - 1);
    eq.g = g;
    eq.d2coef = d2coef;
    eq.tstep = tstep;
    eq.d1coef = make_double2(d1_re, d1_im);
    // reference types_definition.c:87-95
    eq.prop_dt = imagtime ? make_double2(0, -tstep) : make_double2(tstep, 0);
    eq.time_fac = imagtime ? make_double2(0, -1) : make_double2(1, 0);
    OrbWork* w = nullptr;
    int rc = orbwork_create(eq, imagtime, pot, &w);
    if (rc) return rc;
    mctdhb_orbital* o = new mctdhb_orbital();
    o->w = w;
    *out = o;
    return kOk;
}

extern "C" int mctdhb_orbital_destroy(mctdhb_orbital* o) {
    if (!o) return kOk;
    hipDeviceSynchronize();
    orbwork_destroy(o->w);
    delete o;
    return kOk;
}

extern "C" int mctdhb_orbital_set_potential(mctdhb_orbital* o, const double* pot) {
    if (!o || !pot) {
        set_error("null argument");
        return kInvalidArgument;
    }
    return orbwork_set_potential(o->w, pot);
}

extern "C" int mctdhb_orbital_set_interaction(mctdhb_orbital* o, double g) {
    if (!o) {
        set_error("null argument");
        return kInvalidArgument;
    }
    o->w->eq.g = g;
    return kOk;
}

extern "C" int mctdhb_orbital_set_rk_order(mctdhb_orbital* o, int order) {
    if (!o) {
        set_error("null argument");
        return kInvalidArgument;
    }
    if (order != 2 && order != 4 && order != 5) {
        set_error("Runge-Kutta order must be 2, 4 or 5");
        return kInvalidArgument;
    }
    o->w->rk_order = order;
    return kOk;
}

#define ORB_CHECK(...)                      \
    do {                                    \
        if (!o || __VA_ARGS__) {            \
            set_error("null argument");     \
            return kInvalidArgument;        \
        }                                   \
    } while (0)

extern "C" int mctdhb_linear_fft_dev(mctdhb_orbital* o, void* orb, void* stream) {
    ORB_CHECK(!orb);
    return linear_fft(o->w, (double2*) orb, (hipStream_t) stream);
}

extern "C" int mctdhb_linear_cn_dev(mctdhb_orbital* o, void* orb, void* stream) {
    ORB_CHECK(!orb);
    return linear_cn(o->w, (double2*) orb, (hipStream_t) stream);
}

extern "C" int mctdhb_orbital_matrices_dev(mctdhb_orbital* o, const void* orb, void* hob, void* hint, void* stream) {
    ORB_CHECK(!orb || !hob || !hint);
    return orbital_matrices(o->w, (const double2*) orb, (double2*) hob, (double2*) hint, (hipStream_t) stream);
}

extern "C" int mctdhb_nonlinear_rhs_dev(mctdhb_orbital* o, const void* orb, const void* hob, const void* hint,
                                        const void* rho2, const void* rinv, void* der, void* stream) {
    ORB_CHECK(!orb || !hob || !hint || !rho2 || !rinv || !der);
    return nonlinear_rhs(o->w, (const double2*) orb, (const double2*) hob, (const double2*) hint,
                         (const double2*) rho2, (const double2*) rinv, (double2*) der, (hipStream_t) stream);
}

extern "C" int mctdhb_orthonormalize_dev(mctdhb_orbital* o, void* orb, void* stream) {
    ORB_CHECK(!orb);
    return orthonormalize(o->w, (double2*) orb, (hipStream_t) stream);
}

extern "C" int mctdhb_splitstep_dev(mctdhb_orbital* o, int use_fft, void* orb, void* hob, void* hint,
                                    const void* rho2, const void* rinv, void* stream) {
    ORB_CHECK(!orb || !hob || !hint || !rho2 || !rinv);
    return splitstep(o->w, use_fft, (double2*) orb, (double2*) hob, (double2*) hint, (const double2*) rho2,
                     (const double2*) rinv, (hipStream_t) stream);
}

extern "C" int mctdhb_orbital_set_integrator(mctdhb_orbital* o, int orb_method, int der_method, int robust) {
    if (!o) {
        set_error("null argument");
        return kInvalidArgument;
    }
    if (orb_method != FULLSTEP_RUNGEKUTTA && orb_method != SPLITSTEP) {
        set_error("orbital integrator must be FULLSTEP_RUNGEKUTTA (30) or SPLITSTEP (31)");
        return kInvalidArgument;
    }
    if (der_method != 0 && der_method != DVR && der_method != SPECTRAL && der_method != FINITEDIFF) {
        set_error("derivative method must be 0, DVR (300), SPECTRAL (301) or FINITEDIFF (302)");
        return kInvalidArgument;
    }
    o->w->fullstep = orb_method == FULLSTEP_RUNGEKUTTA;
    o->w->robust = robust != 0;
    o->w->dvr = o->w->fullstep && der_method == DVR;
    return o->w->dvr ? ensure_dvr(o->w) : kOk;
}

// use_fd of the device entry points: 0 FFT, 1 finite differences, 2 DVR
static bool der_ok(int d) {
    if (d >= 0 && d <= 2) return true;
    set_error("derivative selector must be 0 (FFT), 1 (finite differences) or 2 (DVR)");
    return false;
}

extern "C" int mctdhb_fullstep_dev(mctdhb_orbital* o, int use_fd, int robust, void* orb, const void* rho2,
                                   const void* rinv, void* stream) {
    ORB_CHECK(!orb || !rho2 || !rinv);
    if (!der_ok(use_fd)) return kInvalidArgument;
    return fullstep(o->w, use_fd, robust, (double2*) orb, (const double2*) rho2, (const double2*) rinv,
                    (hipStream_t) stream);
}

extern "C" int mctdhb_fullstep_rhs_dev(mctdhb_orbital* o, int use_fd, int robust, const void* orb, const void* rho2,
                                       const void* rinv, void* der, void* stream) {
    ORB_CHECK(!orb || !rho2 || !rinv || !der);
    if (!der_ok(use_fd)) return kInvalidArgument;
    return fullstep_rhs(o->w, use_fd, robust, (const double2*) orb, (const double2*) rho2, (const double2*) rinv,
                        (double2*) der, (hipStream_t) stream);
}

extern "C" int mctdhb_projector_dev(mctdhb_orbital* o, int robust, const void* orb, const void* hact, void* proj,
                                    void* stream) {
    ORB_CHECK(!orb || !hact || !proj);
    return projector(o->w, robust, (const double2*) orb, (const double2*) hact, (double2*) proj,
                     (hipStream_t) stream);
}

// The orbital part of a whole step: the integrator set on the workspace.
// periodic_tail: end with the periodic copy of the first grid point onto
// the last (driver.c:179-182), folded into the split step's last kernel.
static int orbital_step(OrbWork* w, int use_fft, double2* orb, double2* hob, double2* hint, const double2* rho2,
                        const double2* rinv, hipStream_t st, bool periodic_tail) {
    if (w->fullstep) {
        const int rc = fullstep(w, w->dvr ? 2 : (use_fft ? 0 : 1), w->robust, orb, rho2, rinv, st);
        if (rc || !periodic_tail || w->eq.bounds != kPeriodicBounds) return rc;
        return periodic_bounds(w, orb, st);
    }
    return splitstep(w, use_fft, orb, hob, hint, rho2, rinv, st, periodic_tail);
}

extern "C" int mctdhb_density_sync_dev(mctdhb_space* sp, const void* coef, void* rho1, void* rho2, void* rinv,
                                       void* eigvals, double reg_factor, int prec, void* stream) {
    if (!sp || !coef || !rho1 || !rho2 || !rinv) {
        set_error("null argument");
        return kInvalidArgument;
    }
    if (prec != kDmFP64 && prec != kDmMixed) {
        set_error("unknown precision mode");
        return kInvalidArgument;
    }
    if (!sp->dm_work) MCTDHB_HIP_CHECK(hipMalloc(&sp->dm_work, dm_workspace_bytes(sp->S)));
    const DmSync sy{(double2*) rinv, (double*) eigvals, reg_factor};
    return density_matrices(sp->S, coef, (double2*) rho1, (double2*) rho2, prec, kIndexOnTheFly, sp->dm_work,
                            (hipStream_t) stream, &sy);
}

extern "C" int mctdhb_propagate_step_dev(mctdhb_space* sp, mctdhb_sil* s, mctdhb_orbital* o, int use_fft,
                                         void* orb, void* coef, void* hob, void* hint, void* rho1, void* rho2,
                                         void* rinv, void* stream) {
    if (!sp || !s || !o || !orb || !coef || !hob || !hint || !rho1 || !rho2 || !rinv) {
        set_error("null argument");
        return kInvalidArgument;
    }
    hipStream_t st = (hipStream_t) stream;
    OrbWork* w = o->w;
    // driver.c:179-186: orbital step and periodic bounds, then h and g
    // from the new orbitals
    int rc = orbital_step(w, use_fft, (double2*) orb, (double2*) hob, (double2*) hint, (const double2*) rho2,
                          (const double2*) rinv, st, true);
    if (rc) return rc;
    rc = orbital_matrices(w, (const double2*) orb, (double2*) hob, (double2*) hint, st);
    if (rc) return rc;
    if (sp->S->norb == 1) return kOk;
    const double2 dt = w->eq.prop_dt;
    rc = sil_step(s->L, (double2*) coef, (const double2*) hob, (const double2*) hint, dt.x, dt.y, w->imagtime, st);
    if (rc) return rc;
    return mctdhb_density_sync_dev(sp, coef, rho1, rho2, rinv, nullptr, g_reg_factor, MCTDHB_FP64, stream);
}

struct mctdhb_step_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

// Records what body enqueues on a private stream into g.  Everything body
// allocates or tunes on first use must already be done: the capture takes
// kernels and device copies only.  The capture executes nothing, so it waits
// for no work in flight (other solves keep running while one is admitted).
template <typename Body> static int capture_graph(Body body, mctdhb_step_graph* g) {
    hipStream_t cs = nullptr;
    MCTDHB_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    int rc = kOk;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        set_error("hipStreamBeginCapture failed");
        rc = kDeviceError;
    } else {
        rc = body(cs);
        const hipError_t e = hipStreamEndCapture(cs, &graph);
        if (!rc && e != hipSuccess) {
            set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            rc = kDeviceError;
        }
    }
    hipStreamDestroy(cs);
    hipGraphExec_t exec = nullptr;
    if (!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        set_error("hipGraphInstantiate failed");
        rc = kDeviceError;
    }
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    g->graph = graph;
    g->exec = exec;
    return kOk;
}

extern "C" int mctdhb_step_graph_create(mctdhb_space* sp, mctdhb_sil* s, mctdhb_orbital* o, int use_fft, void* orb,
                                        void* coef, void* hob, void* hint, void* rho1, void* rho2, void* rinv,
                                        mctdhb_step_graph** out) {
    if (!out || !sp || !s || !o || !orb || !coef || !hob || !hint || !rho1 || !rho2 || !rinv) {
        set_error("null argument");
        return kInvalidArgument;
    }
    *out = nullptr;
    if (sp->S->norb > 7) {
        set_error("step graphs cover 1 <= norb <= 7 (the fused density kernels)");
        return kInvalidArgument;
    }
    // Everything the step allocates or tunes on first use is done here, so
    // the capture below records kernels and device copies only.
    if (!sp->dm_work) MCTDHB_HIP_CHECK(hipMalloc(&sp->dm_work, dm_workspace_bytes(sp->S)));
    if (s->L->prec == kPrecMixed) {
        const int rc = hc_tune_tiling(sp->S, nullptr);
        if (rc) return rc;
    }
    mctdhb_step_graph* g = new mctdhb_step_graph();
    const int rc = capture_graph(
        [&](hipStream_t cs) {
            return mctdhb_propagate_step_dev(sp, s, o, use_fft, orb, coef, hob, hint, rho1, rho2, rinv, cs);
        },
        g);
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    return kOk;
}

extern "C" int mctdhb_step_graph_launch(mctdhb_step_graph* g, void* stream) {
    if (!g) {
        set_error("null argument");
        return kInvalidArgument;
    }
    MCTDHB_HIP_CHECK(hipGraphLaunch(g->exec, (hipStream_t) stream));
    return kOk;
}

extern "C" int mctdhb_step_graph_launch_many(mctdhb_step_graph* const* graphs, int n, void* const* streams,
                                             int nstreams, int rounds) {
    if (n < 0 || rounds < 0 || (n > 0 && (!graphs || !streams || nstreams < 1))) {
        set_error("invalid argument");
        return kInvalidArgument;
    }
    for (int i = 0; i < n; i++)
        if (!graphs[i]) {
            set_error("null graph");
            return kInvalidArgument;
        }
    for (int r = 0; r < rounds; r++)
        for (int i = 0; i < n; i++)
            MCTDHB_HIP_CHECK(hipGraphLaunch(graphs[i]->exec, (hipStream_t) streams[i % nstreams]));
    return kOk;
}

extern "C" int mctdhb_step_graph_destroy(mctdhb_step_graph* g) {
    if (!g) return kOk;
    hipDeviceSynchronize();
    hipGraphExecDestroy(g->exec);
    hipGraphDestroy(g->graph);
    delete g;
    return kOk;
}

// ------------------------------------------------------- packed steps
// Many instances' whole steps (driver.c:159-202 per instance) as one launch
// per stage (pack.hpp).  Instances are grouped by the stages' template
// parameters; each group's step is captured once as a graph of its stage
// launches over its argument tables.

struct PackGroup {
    int M = 0, ns = 0, imagtime = 0, index = 0, fold = 0;
    bool stage = false;
    int count = 0;
    void* d_tab = nullptr;  // every stage table, one allocation
    size_t fft_smem = 0, gs_smem = 0, sil_smem = 0, dm_smem = 0;
    unsigned fft_n = 0, nl_blocks = 0, dm_grid = 0;
    mctdhb_step_graph graph;
};

struct mctdhb_pack {
    std::vector<PackGroup> groups;
};

static int pack_rk_stages(int rk_order) { return rk_order == 2 ? 2 : (rk_order == 5 ? 6 : 4); }

// Why the instance's step cannot run packed, or nullptr.
static const char* pack_refusal(const mctdhb_space* sp, const mctdhb_sil* s, const mctdhb_orbital* o,
                                int use_fft) {
    const Space* S = sp->S;
    const Sil* L = s->L;
    const OrbWork* w = o->w;
    if (s->sp != sp) return "the SIL workspace belongs to another space";
    if (w->fullstep) return "packed steps take the split-step orbital integrator";
    if (!use_fft || w->log2n < 0) return "packed steps take the FFT linear step of a power-of-two length";
    if (S->norb < 2 || S->norb > 7 || w->eq.norb != S->norb) return "packed steps cover 2 <= norb <= 7";
    if (L->prec != kPrecFP64 || L->generic || !L->small || !sil_small_fits(S, L->lm))
        return "packed steps take the one-workgroup FP64 SIL step";
    const DmPlan plan = dm_plan(S, kDmFP64, kIndexOnTheFly);
    if (plan.wave || plan.nb > kDmFinishBlocks || !dm_family_smem(S, sizeof(double)))
        return "packed steps take the one-block density finish";
    return nullptr;
}

extern "C" int mctdhb_pack_eligible(const mctdhb_space* sp, const mctdhb_sil* s, const mctdhb_orbital* o,
                                    int use_fft, int* eligible) {
    if (!sp || !s || !o || !eligible) {
        set_error("null argument");
        return kInvalidArgument;
    }
    const char* why = pack_refusal(sp, s, o, use_fft);
    *eligible = why ? 0 : 1;
    if (why) set_error(why);
    return kOk;
}

extern "C" int mctdhb_pack_destroy(mctdhb_pack* p) {
    if (!p) return kOk;
    hipDeviceSynchronize();
    for (PackGroup& g : p->groups) {
        if (g.graph.exec) hipGraphExecDestroy(g.graph.exec);
        if (g.graph.graph) hipGraphDestroy(g.graph.graph);
        hipFree(g.d_tab);
    }
    delete p;
    return kOk;
}

// The stage launches of one group's step, in the order of
// mctdhb_propagate_step_dev with the FFT linear step: linear half, h and g,
// nonlinear RK sub-step (its tables first where the kernel does not form
// them), linear half, Gram-Schmidt (imaginary time), h and g, coefficient
// block, SIL step, density matrices with the synchronisation.
static int pack_enqueue(const PackGroup& g, PkFft* fft, PkOm* om, PkNlCoef* nlc, PkNl* nl, PkGs* gs, PkCf* cf,
                        PkSil* sil, PkDm* dm, PkDmFin* fin, hipStream_t st) {
    int rc = pk_linear_fft(fft, g.count, g.M, g.fft_smem, st);
    if (!rc) rc = pk_orbital_matrices(om, g.count, g.M, st);
    if (!rc && !g.fold) rc = pk_nl_coef(nlc, g.count, g.M, st);
    if (!rc) rc = pk_nl_rk(nl, g.count, g.M, g.ns, g.nl_blocks, st);
    if (!rc) rc = pk_linear_fft(fft, g.count, g.M, g.fft_smem, st);
    if (!rc && g.imagtime) rc = pk_gram_schmidt(gs, g.count, g.M, g.gs_smem, st);
    if (!rc) rc = pk_orbital_matrices(om, g.count, g.M, st);
    if (!rc) rc = pk_hc_coefficients(cf, g.count, g.M, st);
    if (!rc) rc = pk_sil_small(sil, g.count, g.M, g.index, g.stage, g.sil_smem, st);
    if (!rc) rc = pk_density(dm, g.count, g.M, g.dm_grid, g.dm_smem, st);
    if (!rc) rc = pk_dm_finish(fin, g.count, g.M, sizeof(double2) * 4 * g.M * g.M, st);
    return rc;
}

extern "C" int mctdhb_pack_create(int n, mctdhb_space* const* spaces, mctdhb_sil* const* sils,
                                  mctdhb_orbital* const* orbitals, int use_fft, void* const* orb, void* const* coef,
                                  void* const* hob, void* const* hint, void* const* rho1, void* const* rho2,
                                  void* const* rinv, mctdhb_pack** out) {
    if (!out || n < 1 || !spaces || !sils || !orbitals || !orb || !coef || !hob || !hint || !rho1 || !rho2 ||
        !rinv) {
        set_error("invalid argument");
        return kInvalidArgument;
    }
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!spaces[i] || !sils[i] || !orbitals[i] || !orb[i] || !coef[i] || !hob[i] || !hint[i] || !rho1[i] ||
            !rho2[i] || !rinv[i]) {
            set_error("null argument");
            return kInvalidArgument;
        }
        if (const char* why = pack_refusal(spaces[i], sils[i], orbitals[i], use_fft)) {
            set_error(std::string("instance ") + std::to_string(i) + ": " + why);
            return kInvalidArgument;
        }
    }
    // the solves of a group run every stage at once: each needs workspaces of its own
    // (the space holds the density partial sums, the SIL workspace the Krylov vectors,
    // the orbital workspace the Runge-Kutta tables)
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++)
            if (spaces[i] == spaces[j] || sils[i] == sils[j] || orbitals[i] == orbitals[j] ||
                coef[i] == coef[j] || orb[i] == orb[j]) {
                set_error("instances " + std::to_string(j) + " and " + std::to_string(i) +
                          " share a workspace or buffer; every solve in a pack needs its own");
                return kInvalidArgument;
            }
    // groups by the stages' template parameters, members in input order
    std::vector<std::vector<int>> members;
    std::vector<PackGroup> groups;
    for (int i = 0; i < n; i++) {
        const Space* S = spaces[i]->S;
        const Sil* L = sils[i]->L;
        const OrbWork* w = orbitals[i]->w;
        PackGroup key;
        key.M = S->norb;
        key.ns = pack_rk_stages(w->rk_order);
        key.imagtime = w->imagtime ? 1 : 0;
        key.index = L->index;
        key.stage = sil_small_stage(S);
        key.fold = pk_nl_folds(key.M, key.ns) ? 1 : 0;
        size_t gi = 0;
        for (; gi < groups.size(); gi++) {
            const PackGroup& g = groups[gi];
            if (g.M == key.M && g.ns == key.ns && g.imagtime == key.imagtime && g.index == key.index &&
                g.stage == key.stage)
                break;
        }
        if (gi == groups.size()) {
            groups.push_back(key);
            members.emplace_back();
        }
        members[gi].push_back(i);
    }
    mctdhb_pack* p = new mctdhb_pack();
    for (size_t gi = 0; gi < groups.size(); gi++) {
        PackGroup g = groups[gi];
        const std::vector<int>& mem = members[gi];
        g.count = (int) mem.size();
        std::vector<PkFft> fft(g.count);
        std::vector<PkOm> om(g.count);
        std::vector<PkNlCoef> nlc(g.count);
        std::vector<PkNl> nl(g.count);
        std::vector<PkGs> gs(g.count);
        std::vector<PkCf> cf(g.count);
        std::vector<PkSil> sil(g.count);
        std::vector<PkDm> dm(g.count);
        std::vector<PkDmFin> fin(g.count);
        for (int j = 0; j < g.count; j++) {
            const int i = mem[j];
            mctdhb_space* sp = spaces[i];
            const Space* S = sp->S;
            const Sil* L = sils[i]->L;
            const OrbWork* w = orbitals[i]->w;
            const OrbEq& q = w->eq;
            if (!sp->dm_work && hipMalloc(&sp->dm_work, dm_workspace_bytes(S)) != hipSuccess) {
                set_error("hipMalloc of the density workspace failed");
                mctdhb_pack_destroy(p);
                return kOutOfMemory;
            }
            double2* o = (double2*) orb[i];
            double2* h1 = (double2*) hob[i];
            double2* h2 = (double2*) hint[i];
            double2* c = (double2*) coef[i];
            fft[j] = PkFft{o, w->d_tw, w->d_fexp, q.grid_size, w->nfft, w->log2n};
            g.fft_smem = std::max(g.fft_smem, sizeof(double2) * w->nfft);
            om[j] = PkOm{o, w->d_pot, h1, h2, q.dx, q.d2coef, q.g, q.d1coef, q.grid_size};
            nlc[j] = PkNlCoef{h1, h2, (const double2*) rho2[i], (const double2*) rinv[i], w->d_nl};
            nl[j] = PkNl{o, w->d_pot, w->d_nl, h1, h2, (const double2*) rho2[i], (const double2*) rinv[i],
                         q.tstep, q.g, q.time_fac, q.grid_size};
            g.nl_blocks = std::max(g.nl_blocks, div_up((unsigned) q.grid_size, 64u));
            const size_t gsm = pk_gs_smem(q.norb, q.grid_size);
            gs[j] = PkGs{o, q.dx, q.grid_size, q.bounds == kPeriodicBounds ? 1 : 0, gsm ? 1 : 0};
            g.gs_smem = std::max(g.gs_smem, gsm);
            cf[j] = PkCf{h1, h2, L->cf64, S->npar};
            SmallSilArgs a{};
            a.st = L->d_st;
            a.w[0] = c;
            for (int k = 1; k < L->lm; k++) a.w[k] = L->w[k];
            a.out = L->out;
            a.lm = L->lm;
            a.dt_re = q.prop_dt.x;
            a.dt_im = q.prop_dt.y;
            a.imagtime = w->imagtime;
            const RowCtx ctx = make_row_ctx(S->dim, S->npar, S->d_ht, S->d_map, S->d_mapot, S->d_maptt,
                                            S->d_strideot, S->d_stridett, S->d_occ8);
            sil[j] = PkSil{ctx, L->cf64, S->d_W, S->ldw, a};
            g.sil_smem = std::max(g.sil_smem, sil_small_smem(S) + (g.stage ? sil_small_stage_bytes(S) : 0));
            const DmPlan plan = dm_plan(S, kDmFP64, kIndexOnTheFly);
            const unsigned grid = pk_density_grid(S->norb, plan.nb);
            double2* part = (double2*) sp->dm_work;
            dm[j] = PkDm{ctx, c, S->d_W, part, S->ldw, S->npar + 3 <= 2048 ? 1 : 0, plan.nb, grid};
            g.dm_grid = std::max(g.dm_grid, grid);
            g.dm_smem = std::max(g.dm_smem, dm_family_smem(S, sizeof(double)));
            const int E = dm_num_elements(S->norb);
            fin[j] = PkDmFin{part, part + (size_t) plan.nb * E, (double2*) rho1[i], (double2*) rho2[i],
                             DmSync{(double2*) rinv[i], nullptr, g_reg_factor}, (int) plan.nb};
        }
        // one allocation, each table at a 256-byte boundary
        size_t off = 0;
        auto place = [&](size_t bytes) {
            const size_t at = off;
            off += (bytes + 255) & ~(size_t) 255;
            return at;
        };
        const size_t o_fft = place(sizeof(PkFft) * g.count), o_om = place(sizeof(PkOm) * g.count);
        const size_t o_nlc = place(sizeof(PkNlCoef) * g.count), o_nl = place(sizeof(PkNl) * g.count);
        const size_t o_gs = place(sizeof(PkGs) * g.count), o_cf = place(sizeof(PkCf) * g.count);
        const size_t o_sil = place(sizeof(PkSil) * g.count), o_dm = place(sizeof(PkDm) * g.count);
        const size_t o_fin = place(sizeof(PkDmFin) * g.count);
        std::vector<unsigned char> host(off);
        std::memcpy(host.data() + o_fft, fft.data(), sizeof(PkFft) * g.count);
        std::memcpy(host.data() + o_om, om.data(), sizeof(PkOm) * g.count);
        std::memcpy(host.data() + o_nlc, nlc.data(), sizeof(PkNlCoef) * g.count);
        std::memcpy(host.data() + o_nl, nl.data(), sizeof(PkNl) * g.count);
        std::memcpy(host.data() + o_gs, gs.data(), sizeof(PkGs) * g.count);
        std::memcpy(host.data() + o_cf, cf.data(), sizeof(PkCf) * g.count);
        std::memcpy(host.data() + o_sil, sil.data(), sizeof(PkSil) * g.count);
        std::memcpy(host.data() + o_dm, dm.data(), sizeof(PkDm) * g.count);
        std::memcpy(host.data() + o_fin, fin.data(), sizeof(PkDmFin) * g.count);
        if (hipMalloc(&g.d_tab, off) != hipSuccess) {
            set_error("hipMalloc of the pack tables failed");
            mctdhb_pack_destroy(p);
            return kOutOfMemory;
        }
        p->groups.push_back(g);  // owned by the pack from here (freed by destroy on failure)
        PackGroup& pg = p->groups.back();
        if (hipMemcpy(pg.d_tab, host.data(), off, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("copy of the pack tables failed");
            mctdhb_pack_destroy(p);
            return kDeviceError;
        }
        unsigned char* base = (unsigned char*) pg.d_tab;
        const int rc = capture_graph(
            [&](hipStream_t cs) {
                return pack_enqueue(pg, (PkFft*) (base + o_fft), (PkOm*) (base + o_om), (PkNlCoef*) (base + o_nlc),
                                    (PkNl*) (base + o_nl), (PkGs*) (base + o_gs), (PkCf*) (base + o_cf),
                                    (PkSil*) (base + o_sil), (PkDm*) (base + o_dm), (PkDmFin*) (base + o_fin), cs);
            },
            &pg.graph);
        if (rc) {
            mctdhb_pack_destroy(p);
            return rc;
        }
    }
    *out = p;
    return kOk;
}

extern "C" int mctdhb_pack_groups(const mctdhb_pack* p, int* groups) {
    if (!p || !groups) {
        set_error("null argument");
        return kInvalidArgument;
    }
    *groups = (int) p->groups.size();
    return kOk;
}

extern "C" int mctdhb_pack_step(mctdhb_pack* p, void* const* streams, int nstreams, int rounds) {
    if (!p || rounds < 0 || !streams || nstreams < 1) {
        set_error("invalid argument");
        return kInvalidArgument;
    }
    for (int r = 0; r < rounds; r++)
        for (size_t gi = 0; gi < p->groups.size(); gi++)
            MCTDHB_HIP_CHECK(hipGraphLaunch(p->groups[gi].graph.exec, (hipStream_t) streams[gi % nstreams]));
    return kOk;
}

// ------------------------------------------------------------------ layer 1

[[noreturn]] static void fail_stop(const char* what) {
    printf("\n\nERROR: %s: %s\n\n", what, last_error());
    exit(EXIT_FAILURE);
}

static void* xmalloc(size_t bytes) {
    void* p = malloc(bytes ? bytes : 1);
    if (!p) {
        printf("\n\nMEMORY ERROR : malloc failed to allocate %zu bytes\n\n", bytes);
        exit(EXIT_FAILURE);
    }
    return p;
}

extern "C" void assert_space_parameters(uint16_t npar, uint16_t norb) {
    if (npar == 0 || norb == 0) {
        printf("\n\nInvalid number of particles %u or orbitals %u\n\n", npar, norb);
        exit(EXIT_FAILURE);
    }
    if (norb > MAX_ORBITALS) {
        printf("\nERROR: Exceeded max number of orbitals %d\n", MAX_ORBITALS);
        exit(EXIT_FAILURE);
    }
    if (npar > MAX_PARTICLES) {
        printf("\nERROR: Exceeded max number of particles %d\n", MAX_PARTICLES);
        exit(EXIT_FAILURE);
    }
}

extern "C" uint32_t space_dimension(uint16_t npar, uint16_t norb) {
    uint32_t d = 0;
    if (space_dimension_exact(npar, norb, &d) != kOk) {
        printf("\n\nINTEGER SIZE ERROR : overflow occurred representing the space dimension as 32-bit "
               "integers for %u particles and %u orbitals\n\n",
               npar, norb);
        exit(EXIT_FAILURE);
    }
    return d;
}

extern "C" uint32_t* get_subspaces_dim(uint16_t npar, uint16_t norb) {
    uint32_t* s = (uint32_t*) xmalloc(sizeof(uint32_t) * (norb + 1) * (npar + 1));
    for (uint32_t n = 0; n <= npar; n++) {
        s[n] = 0;
        s[n + (npar + 1)] = 1;
        for (uint32_t m = 2; m <= norb; m++) s[n + (npar + 1) * m] = (uint32_t) binom_u64(n + m - 1, m - 1);
    }
    return s;
}

extern "C" void set_config(uint32_t hash_i, uint16_t npar, uint16_t norb, uint16_t* conf) {
    unrank_host(hash_i, npar, norb, conf);
}

extern "C" uint32_t hash_index(uint16_t npar, uint16_t norb, uint32_t* subspaces_dim, uint16_t* conf) {
    (void) subspaces_dim;  // the closed-form rank needs no table
    return rank_host(npar, norb, conf);
}

// Device space for a MultiConfiguration (built on first use if the caller
// assembled the struct itself).
static mctdhb_space* ensure_dev(MultiConfiguration mc) {
    if (!mc->dev) {
        mctdhb_space* sp = nullptr;
        if (mctdhb_space_create(mc->npar, mc->norb, MCTDHB_SPACE_TABLES, &sp)) fail_stop("device space");
        mc->dev = sp;
    }
    return (mctdhb_space*) mc->dev;
}

template <typename T> static T* d2h_array(const T* d, size_t n) {
    T* h = (T*) xmalloc(sizeof(T) * n);
    if (n && hipMemcpy(h, d, sizeof(T) * n, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("device to host copy failed");
        fail_stop("table copy");
    }
    return h;
}

static mctdhb_space* temp_space(uint16_t npar, uint16_t norb) {
    assert_space_parameters(npar, norb);
    mctdhb_space* sp = nullptr;
    if (mctdhb_space_create(npar, norb, MCTDHB_SPACE_TABLES, &sp)) fail_stop("device space");
    return sp;
}

extern "C" uint16_t* get_hash_table(uint16_t npar, uint16_t norb) {
    mctdhb_space* sp = temp_space(npar, norb);
    uint16_t* h = d2h_array(sp->S->d_ht, (size_t) sp->S->dim * norb);
    mctdhb_space_destroy(sp);
    return h;
}

extern "C" uint32_t* get_single_jump_map(uint16_t npar, uint16_t norb, uint32_t* subsdim, uint16_t* hash_table) {
    (void) subsdim;
    (void) hash_table;
    mctdhb_space* sp = temp_space(npar, norb);
    uint32_t* h = d2h_array(sp->S->d_map, (size_t) sp->S->dim * norb * norb);
    mctdhb_space_destroy(sp);
    return h;
}

extern "C" uint32_t* get_double_diffjump_map(uint16_t npar, uint16_t norb, uint32_t* subsdim, uint16_t* hash_table,
                                             uint32_t* strides) {
    (void) subsdim;
    (void) hash_table;
    mctdhb_space* sp = temp_space(npar, norb);
    if (hipMemcpy(strides, sp->S->d_stridett, sizeof(uint32_t) * sp->S->dim, hipMemcpyDeviceToHost) != hipSuccess)
        fail_stop("table copy");
    uint32_t* h = d2h_array(sp->S->d_maptt, sp->S->maptt_len);
    mctdhb_space_destroy(sp);
    return h;
}

extern "C" uint32_t* get_double_equaljump_map(uint16_t npar, uint16_t norb, uint32_t* subsdim, uint16_t* hash_table,
                                              uint32_t* strides) {
    (void) subsdim;
    (void) hash_table;
    mctdhb_space* sp = temp_space(npar, norb);
    if (hipMemcpy(strides, sp->S->d_strideot, sizeof(uint32_t) * sp->S->dim, hipMemcpyDeviceToHost) != hipSuccess)
        fail_stop("table copy");
    uint32_t* h = d2h_array(sp->S->d_mapot, sp->S->mapot_len);
    mctdhb_space_destroy(sp);
    return h;
}

extern "C" uint32_t jump1_index(uint16_t norb, uint32_t dim, uint32_t inp_index, uint32_t* map1jump,
                                uint16_t from_orb, uint16_t to_orb) {
    return map1jump[inp_index + from_orb * dim + to_orb * dim * norb];
}

extern "C" uint32_t jump2same_index(uint16_t norb, uint16_t* conf, uint32_t conf_stride, uint32_t* mapot,
                                    uint16_t from_orb, uint16_t to_orb1, uint16_t to_orb2) {
    uint32_t c = 0;
    for (int h = 0; h < from_orb; h++) c += conf[h] > 1;
    return mapot[conf_stride + c * norb * norb + to_orb1 + to_orb2 * norb];
}

extern "C" uint32_t jump2diff_index(uint16_t norb, uint16_t* conf, uint32_t conf_stride, uint32_t* maptt,
                                    uint16_t from_orb1, uint16_t from_orb2, uint16_t to_orb1, uint16_t to_orb2) {
    if (from_orb2 < from_orb1) {
        uint16_t t = from_orb1;
        from_orb1 = from_orb2;
        from_orb2 = t;
    }
    uint32_t c = 0;
    for (int h = 0; h < from_orb1; h++)
        for (int g = h + 1; g < norb; g++) c += (conf[h] > 0 && conf[g] > 0);
    for (int g = from_orb1 + 1; g < from_orb2; g++) c += conf[g] > 0;
    return maptt[conf_stride + c * norb * norb + to_orb1 + to_orb2 * norb];
}

extern "C" MultiConfiguration get_multiconf_struct(uint16_t npar, uint16_t norb) {
    assert_space_parameters(npar, norb);
    mctdhb_space* sp = nullptr;
    if (mctdhb_space_create(npar, norb, MCTDHB_SPACE_TABLES, &sp)) fail_stop("get_multiconf_struct");
    const Space* S = sp->S;
    MultiConfiguration mc = (MultiConfiguration) xmalloc(sizeof(_MultiConfiguration));
    OperatorMappings om = (OperatorMappings) xmalloc(sizeof(_OperatorMappings));
    mc->npar = npar;
    mc->norb = norb;
    mc->dim = S->dim;
    mc->subspaces_dim = get_subspaces_dim(npar, norb);
    mc->hash_table = d2h_array(S->d_ht, (size_t) S->dim * norb);
    om->strideot = d2h_array(S->d_strideot, S->dim);
    om->stridett = d2h_array(S->d_stridett, S->dim);
    om->map = d2h_array(S->d_map, (size_t) S->dim * norb * norb);
    om->maptt = d2h_array(S->d_maptt, S->maptt_len);
    om->mapot = d2h_array(S->d_mapot, S->mapot_len);
    mc->op_maps = om;
    mc->dev = sp;
    return mc;
}

extern "C" void destroy_multiconf_struct(MultiConfiguration mc) {
    if (!mc) return;
    free(mc->op_maps->map);
    free(mc->op_maps->mapot);
    free(mc->op_maps->maptt);
    free(mc->op_maps->strideot);
    free(mc->op_maps->stridett);
    free(mc->op_maps);
    free(mc->hash_table);
    free(mc->subspaces_dim);
    mctdhb_space_destroy((mctdhb_space*) mc->dev);
    free(mc);
}

static Cmatrix alloc_cmatrix(int rows, int cols) {
    Cmatrix m = (Cmatrix) xmalloc(sizeof(Carray) * rows);
    for (int i = 0; i < rows; i++) m[i] = (Carray) xmalloc(sizeof(dcomplex) * cols);
    return m;
}

static void free_cmatrix(int rows, Cmatrix m) {
    if (!m) return;
    for (int i = 0; i < rows; i++) free(m[i]);
    free(m);
}

extern "C" ManyBodyState get_manybody_state(uint16_t npar, uint16_t norb, uint16_t grid_size) {
    assert_space_parameters(npar, norb);
    ManyBodyState s = (ManyBodyState) xmalloc(sizeof(_ManyBodyState));
    s->npar = npar;
    s->norb = norb;
    s->grid_size = grid_size;
    s->space_dim = space_dimension(npar, norb);
    const size_t m4 = (size_t) norb * norb * norb * norb;
    s->coef = (Carray) xmalloc(sizeof(dcomplex) * s->space_dim);
    s->hint = (Carray) xmalloc(sizeof(dcomplex) * m4);
    s->tb_denmat = (Carray) xmalloc(sizeof(dcomplex) * m4);
    s->ob_denmat = alloc_cmatrix(norb, norb);
    s->inv_ob_denmat = alloc_cmatrix(norb, norb);
    s->hob = alloc_cmatrix(norb, norb);
    s->orbitals = alloc_cmatrix(norb, grid_size);
    return s;
}

extern "C" void destroy_manybody_sate(ManyBodyState s) {
    if (!s) return;
    free(s->coef);
    free(s->hint);
    free(s->tb_denmat);
    free_cmatrix(s->norb, s->ob_denmat);
    free_cmatrix(s->norb, s->inv_ob_denmat);
    free_cmatrix(s->norb, s->hob);
    free_cmatrix(s->norb, s->orbitals);
    free(s);
}

extern "C" WorkspaceLanczos get_lanczos_workspace(uint16_t iter, uint32_t space_dim) {
    WorkspaceLanczos w = (WorkspaceLanczos) xmalloc(sizeof(_WorkspaceLanczos));
    w->iter = iter;
    w->space_dim = space_dim;
    w->coef_lspace = (Carray) xmalloc(sizeof(dcomplex) * iter);
    w->decomp_diag = (Carray) xmalloc(sizeof(dcomplex) * iter);
    w->decomp_offd = (Carray) xmalloc(sizeof(dcomplex) * iter);
    w->lanczos_vectors = alloc_cmatrix(iter, space_dim);
    w->lapack_diag = (Rarray) xmalloc(sizeof(double) * iter);
    w->lapack_offd = (Rarray) xmalloc(sizeof(double) * iter);
    w->lapack_eigvec = (Rarray) xmalloc(sizeof(double) * iter * iter);
    w->transform = (Carray) xmalloc(sizeof(dcomplex) * iter);
    w->hc = (Carray) xmalloc(sizeof(dcomplex) * space_dim);
    w->reortho = (Carray) xmalloc(sizeof(dcomplex) * iter);
    return w;
}

extern "C" void destroy_lanczos_workspace(WorkspaceLanczos w) {
    if (!w) return;
    free(w->coef_lspace);
    free(w->decomp_diag);
    free(w->decomp_offd);
    free_cmatrix(w->iter, w->lanczos_vectors);
    free(w->lapack_diag);
    free(w->lapack_offd);
    free(w->lapack_eigvec);
    free(w->transform);
    free(w->hc);
    free(w->reortho);
    free(w);
}

extern "C" CoefWorkspace get_coef_workspace(CoefIntegrator coef_integ_method, uint32_t space_dim, uint16_t lan_it) {
    CoefWorkspace c = (CoefWorkspace) xmalloc(sizeof(_CoefWorkspace));
    c->lan_work = nullptr;
    c->extern_work = nullptr;
    if (coef_integ_method == LANCZOS) c->lan_work = get_lanczos_workspace(lan_it, space_dim);
    return c;
}

extern "C" void destroy_coef_workspace(CoefWorkspace c) {
    if (!c) return;
    destroy_lanczos_workspace(c->lan_work);
    free(c);
}

// ---- host <-> device staging for the reference-named operators

static void stage_alloc(double2** p, size_t n) {
    if (!*p && hipMalloc((void**) p, sizeof(double2) * n) != hipSuccess) {
        set_error("device allocation failed");
        fail_stop("staging buffers");
    }
}

static void h2d(void* d, const void* h, size_t bytes) {
    if (hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("host to device copy failed");
        fail_stop("staging");
    }
}

static void d2h(void* h, const void* d, size_t bytes) {
    if (hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("device to host copy failed");
        fail_stop("staging");
    }
}

static void stage_hob_hint(mctdhb_space* sp, Cmatrix hob, Carray hint) {
    const int M = sp->S->norb;
    stage_alloc(&sp->d_hob, (size_t) M * M);
    stage_alloc(&sp->d_hint, (size_t) M * M * M * M);
    std::vector<dcomplex> packed((size_t) M * M);
    for (int k = 0; k < M; k++) memcpy(&packed[(size_t) k * M], hob[k], sizeof(dcomplex) * M);
    h2d(sp->d_hob, packed.data(), sizeof(dcomplex) * M * M);
    h2d(sp->d_hint, hint, sizeof(dcomplex) * M * M * M * M);
}

extern "C" void apply_hamiltonian(MultiConfiguration mc, Carray coef, Cmatrix hob, Carray hint, Carray hc) {
    mctdhb_space* sp = ensure_dev(mc);
    const size_t n = sp->S->dim;
    stage_alloc(&sp->d_coef, n);
    stage_alloc(&sp->d_out, n);
    stage_hob_hint(sp, hob, hint);
    h2d(sp->d_coef, coef, sizeof(dcomplex) * n);
    if (mctdhb_apply_hamiltonian_dev(sp, sp->d_coef, sp->d_hob, sp->d_hint, sp->d_out, MCTDHB_FP64,
                                     MCTDHB_INDEX_FLY, nullptr))
        fail_stop("apply_hamiltonian");
    d2h(hc, sp->d_out, sizeof(dcomplex) * n);
}

static void density_host(MultiConfiguration mc, Carray coef, bool two) {
    mctdhb_space* sp = ensure_dev(mc);
    const size_t n = sp->S->dim;
    const int M = sp->S->norb;
    stage_alloc(&sp->d_coef, n);
    stage_alloc(&sp->d_rho1, (size_t) M * M);
    if (two) stage_alloc(&sp->d_rho2, (size_t) M * M * M * M);
    h2d(sp->d_coef, coef, sizeof(dcomplex) * n);
    if (mctdhb_density_dev(sp, sp->d_coef, sp->d_rho1, two ? sp->d_rho2 : nullptr, MCTDHB_FP64,
                           sp->S->tables ? MCTDHB_INDEX_TABLES : MCTDHB_INDEX_FLY, nullptr))
        fail_stop("density matrices");
}

extern "C" void set_onebody_dm(MultiConfiguration mc, Carray coef, Cmatrix rho) {
    density_host(mc, coef, false);
    mctdhb_space* sp = (mctdhb_space*) mc->dev;
    const int M = sp->S->norb;
    std::vector<dcomplex> r((size_t) M * M);
    d2h(r.data(), sp->d_rho1, sizeof(dcomplex) * M * M);
    for (int k = 0; k < M; k++) memcpy(rho[k], &r[(size_t) k * M], sizeof(dcomplex) * M);
}

extern "C" void set_twobody_dm(MultiConfiguration mc, Carray coef, Carray rho) {
    density_host(mc, coef, true);
    mctdhb_space* sp = (mctdhb_space*) mc->dev;
    const int M = sp->S->norb;
    d2h(rho, sp->d_rho2, sizeof(dcomplex) * M * M * M * M);
}

// The cached FP64 Krylov workspace of the host entry points; max_lm is the
// caller's bound (MAX_LANCZOS_ITER for the SIL step, the device capacity for
// the ground-state diagonalisation).
static mctdhb_sil* host_sil(mctdhb_space* sp, int lm, int max_lm = MAX_LANCZOS_ITER) {
    if (sp->sil && sp->sil->L->lm != lm) {
        mctdhb_sil_destroy(sp->sil);
        sp->sil = nullptr;
    }
    if (!sp->sil) {
        if (lm < MIN_LANCZOS_ITER || lm > max_lm) {
            printf("\n\nERROR: Lanczos iterations must be between %d and %d but %d was requested\n\n",
                   MIN_LANCZOS_ITER, max_lm, lm);
            exit(EXIT_FAILURE);
        }
        if (mctdhb_sil_create(sp, lm, MCTDHB_FP64, MCTDHB_INDEX_FLY, &sp->sil)) fail_stop("Lanczos workspace");
    }
    return sp->sil;
}

extern "C" int lanczos(MultiConfiguration mc, Cmatrix hob, Carray hint, int lm, Carray diag, Carray offd,
                       Cmatrix lvec) {
    mctdhb_space* sp = ensure_dev(mc);
    const size_t n = sp->S->dim;
    mctdhb_sil* s = host_sil(sp, lm, kMaxKrylov);
    stage_alloc(&sp->d_coef, n);
    stage_hob_hint(sp, hob, hint);
    if (sp->d_lvec) {
        hipFree(sp->d_lvec);
        sp->d_lvec = nullptr;
    }
    stage_alloc(&sp->d_lvec, n * lm);
    h2d(sp->d_coef, lvec[0], sizeof(dcomplex) * n);
    if (mctdhb_lanczos_dev(s, sp->d_coef, sp->d_hob, sp->d_hint, sp->d_lvec, nullptr)) fail_stop("lanczos");
    int done = 0;
    std::vector<double> dg(2 * lm), od(lm);
    if (mctdhb_sil_scalars(s, &done, dg.data(), od.data(), nullptr)) fail_stop("lanczos");
    for (int k = 1; k < done; k++) d2h(lvec[k], sp->d_lvec + (size_t) k * n, sizeof(dcomplex) * n);
    double* dd = (double*) diag;
    double* oo = (double*) offd;
    for (int k = 0; k < done; k++) {
        dd[2 * k] = dg[2 * k];
        dd[2 * k + 1] = dg[2 * k + 1];
        if (k < lm - 1 || done < lm) {
            oo[2 * k] = od[k];
            oo[2 * k + 1] = 0;
        }
    }
    return done;
}

extern "C" double eig_residual(MultiConfiguration mc, Carray c, Cmatrix hob, Carray hint, double e) {
    mctdhb_space* sp = ensure_dev(mc);
    const size_t n = sp->S->dim;
    stage_alloc(&sp->d_coef, n);
    stage_hob_hint(sp, hob, hint);
    h2d(sp->d_coef, c, sizeof(dcomplex) * n);
    double r = 0;
    if (mctdhb_eig_residual_dev(sp, sp->d_coef, sp->d_hob, sp->d_hint, e, &r, nullptr)) fail_stop("eig_residual");
    return r;
}

extern "C" double lowest_state_lanczos(uint16_t niter, MultiConfiguration mc, Cmatrix Ho, Carray Hint, Carray C) {
    mctdhb_space* sp = ensure_dev(mc);
    const size_t n = sp->S->dim;
    mctdhb_sil* s = host_sil(sp, niter, kMaxKrylov);
    stage_alloc(&sp->d_coef, n);
    stage_hob_hint(sp, Ho, Hint);
    h2d(sp->d_coef, C, sizeof(dcomplex) * n);
    double e = 0;
    if (mctdhb_lowest_state_dev(s, sp->d_coef, sp->d_hob, sp->d_hint, &e, nullptr)) {
        printf("\n\nERROR IN DIAGONALIZATION\n\n");
        exit(EXIT_FAILURE);
    }
    int done = 0;
    if (mctdhb_sil_scalars(s, &done, nullptr, nullptr, nullptr)) fail_stop("lowest_state_lanczos");
    if (done < niter) printf("\n\nWARNING : lanczos iterations exit before expected with %d iterations\n\n", done);
    d2h(C, sp->d_coef, sizeof(dcomplex) * n);
    return e;
}

extern "C" void propagate_coef_sil(MCTDHBDataStruct mctdhb, Carray cnext) {
    MultiConfiguration mc = mctdhb->multiconfig_space;
    mctdhb_space* sp = ensure_dev(mc);
    const size_t n = sp->S->dim;
    const int lm = mctdhb->coef_workspace->lan_work->iter;
    mctdhb_sil* s = host_sil(sp, lm);
    stage_alloc(&sp->d_coef, n);
    stage_hob_hint(sp, mctdhb->state->hob, mctdhb->state->hint);
    h2d(sp->d_coef, mctdhb->state->coef, sizeof(dcomplex) * n);
    const double* dt = (const double*) &mctdhb->orb_eq->prop_dt;
    // the renormalisation in imaginary time is applied below, to state->coef only
    if (mctdhb_sil_step_dev(s, sp->d_coef, sp->d_hob, sp->d_hint, dt[0], dt[1], 0, nullptr))
        fail_stop("propagate_coef_sil");
    int done = 0;
    if (mctdhb_sil_scalars(s, &done, nullptr, nullptr, nullptr)) {
        printf("\n\nERROR IN DIAGONALIZATION\n\n");
        exit(EXIT_FAILURE);
    }
    if (done < lm)
        printf("\n\nWARNING : lanczos iterations exit before expected - %d of %d iterations\n\n", done, lm);
    d2h(cnext, sp->d_coef, sizeof(dcomplex) * n);
    memcpy(mctdhb->state->coef, cnext, sizeof(dcomplex) * n);
    if (mctdhb->integ_type == IMAGTIME) {
        double* c = (double*) mctdhb->state->coef;
        double s2 = 0;
        for (size_t i = 0; i < n; i++) s2 += c[2 * i] * c[2 * i] + c[2 * i + 1] * c[2 * i + 1];
        const double r = 1.0 / sqrt(s2);
        for (size_t i = 0; i < 2 * n; i++) c[i] *= r;
    }
}

extern "C" dcomplex total_energy(ManyBodyState psi) {
    const int M = psi->norb;
    const double* h = nullptr;
    double wr = 0, wi = 0, zr = 0, zi = 0;
    const double* r2 = (const double*) psi->tb_denmat;
    const double* g = (const double*) psi->hint;
    for (int k = 0; k < M; k++)
        for (int l = 0; l < M; l++) {
            const double* r1 = (const double*) &psi->ob_denmat[k][l];
            h = (const double*) &psi->hob[k][l];
            wr += r1[0] * h[0] - r1[1] * h[1];
            wi += r1[0] * h[1] + r1[1] * h[0];
            for (int s = 0; s < M; s++)
                for (int q = 0; q < M; q++) {
                    const size_t i = k + l * M + s * M * M + (size_t) q * M * M * M;
                    const size_t j = k + l * M + q * M * M + (size_t) s * M * M * M;
                    zr += r2[2 * i] * g[2 * j] - r2[2 * i + 1] * g[2 * j + 1];
                    zi += r2[2 * i] * g[2 * j + 1] + r2[2 * i + 1] * g[2 * j];
                }
        }
    double out[2] = {wr + zr / 2, wi + zi / 2};
    dcomplex e;
    memcpy(&e, out, sizeof(e));
    return e;
}

// ----------------------------------------------- layer 1: orbital side
//
// The reference-named orbital entry points work on the reference's host
// structs.  The device workspace of an _OrbitalEquation lives in a registry
// keyed by the struct's address (sync_orbital_matrices receives no
// workspace); it is rebuilt when a parameter it was built from changes, and
// freed by destroy_orbital_equation.

namespace {

struct DevOrbital {
    mctdhb_orbital* o = nullptr;
    // parameters the device workspace was built from
    int norb = 0, grid = 0, bounds = 0, imagtime = 0;
    double xi = 0, xf = 0, tstep = 0, d2coef = 0, d1[2] = {0, 0};
    // host <-> device staging
    double2 *d_orb = nullptr, *d_hob = nullptr, *d_hint = nullptr, *d_rho2 = nullptr, *d_rinv = nullptr;
};

std::mutex g_orb_mu;
std::unordered_map<const void*, DevOrbital*> g_orb;

void free_dev_orbital(DevOrbital* d) {
    if (!d) return;
    mctdhb_orbital_destroy(d->o);
    hipFree(d->d_orb);
    hipFree(d->d_hob);
    hipFree(d->d_hint);
    hipFree(d->d_rho2);
    hipFree(d->d_rinv);
    delete d;
}

}  // namespace

static void c2d(const dcomplex& z, double* out) { memcpy(out, &z, sizeof(double) * 2); }

// The device workspace of eq, with g and the trap refreshed from eq.
static DevOrbital* dev_orbital(OrbitalEquation eq) {
    std::lock_guard<std::mutex> lock(g_orb_mu);
    double d1[2], pdt[2];
    c2d(eq->d1coef, d1);
    c2d(eq->prop_dt, pdt);
    const int imag = pdt[1] != 0.0;
    DevOrbital*& d = g_orb[eq];
    if (d && (d->norb != eq->norb || d->grid != eq->grid_size || d->bounds != (int) eq->bounds ||
              d->imagtime != imag || d->xi != eq->xi || d->xf != eq->xf || d->tstep != eq->tstep ||
              d->d2coef != eq->d2coef || d->d1[0] != d1[0] || d->d1[1] != d1[1])) {
        hipDeviceSynchronize();
        free_dev_orbital(d);
        d = nullptr;
    }
    if (!d) {
        d = new DevOrbital();
        if (mctdhb_orbital_create(eq->norb, eq->grid_size, eq->bounds, eq->xi, eq->xf, eq->tstep, eq->g, eq->d2coef,
                                  d1[0], d1[1], imag, eq->pot_grid, &d->o)) {
            delete d;
            d = nullptr;
            fail_stop("orbital workspace");
        }
        d->norb = eq->norb;
        d->grid = eq->grid_size;
        d->bounds = eq->bounds;
        d->imagtime = imag;
        d->xi = eq->xi;
        d->xf = eq->xf;
        d->tstep = eq->tstep;
        d->d2coef = eq->d2coef;
        d->d1[0] = d1[0];
        d->d1[1] = d1[1];
        const size_t M = eq->norb;
        stage_alloc(&d->d_orb, M * eq->grid_size);
        stage_alloc(&d->d_hob, M * M);
        stage_alloc(&d->d_hint, M * M * M * M);
        stage_alloc(&d->d_rho2, M * M * M * M);
        stage_alloc(&d->d_rinv, M * M);
    } else if (mctdhb_orbital_set_potential(d->o, eq->pot_grid)) {
        fail_stop("orbital potential");
    }
    if (mctdhb_orbital_set_interaction(d->o, eq->g)) fail_stop("orbital interaction");
    return d;
}

static void orbitals_h2d(double2* d, Cmatrix orb, int norb, int grid) {
    for (int k = 0; k < norb; k++) h2d(d + (size_t) k * grid, orb[k], sizeof(dcomplex) * grid);
}

static void orbitals_d2h(Cmatrix orb, const double2* d, int norb, int grid) {
    for (int k = 0; k < norb; k++) d2h(orb[k], d + (size_t) k * grid, sizeof(dcomplex) * grid);
}

static void cmatrix_h2d(double2* d, Cmatrix m, int n) { orbitals_h2d(d, m, n, n); }
static void cmatrix_d2h(Cmatrix m, const double2* d, int n) { orbitals_d2h(m, d, n, n); }

extern "C" OrbitalEquation get_orbital_equation(char eq_name[], uint16_t norb, uint16_t grid_size,
                                                BoundaryCondition bounds, IntegratorType integ_type, double xi,
                                                double xf, double tstep, double tend, double d2coef, dcomplex d1coef,
                                                void* pot_extra_args, void* inter_extra_args,
                                                single_particle_pot pot_func, time_dependent_parameter inter_param) {
    if (grid_size > MAX_GRID_SIZE) {
        printf("\n\nERROR: Exceeded max grid size %d\n\n", MAX_GRID_SIZE);
        exit(EXIT_FAILURE);
    }
    if (grid_size < MIN_GRID_SIZE) {
        printf("\n\nERROR: Below min grid size %d\n\n", MIN_GRID_SIZE);
        exit(EXIT_FAILURE);
    }
    if (norb > MAX_ORBITALS) {
        printf("\n\nERROR: Exceeded max number of orbitals %d\n\n", MAX_ORBITALS);
        exit(EXIT_FAILURE);
    }
    if (!pot_func || !inter_param) {
        printf("\n\nTYPEERROR: Null pointer given for the trap potential or the interaction parameter\n\n");
        exit(EXIT_FAILURE);
    }
    OrbitalEquation eq = (OrbitalEquation) xmalloc(sizeof(_OrbitalEquation));
    snprintf(eq->eq_name, STR_BUFF_SIZE, "%s", eq_name ? eq_name : "");
    eq->t = 0;
    eq->norb = norb;
    eq->grid_size = grid_size;
    eq->bounds = bounds;
    eq->xi = xi;
    eq->xf = xf;
    eq->dx = (xf - xi) / (grid_size - 1);
    eq->tstep = tstep;
    eq->tend = tend;
    const double pdt[2] = {integ_type == IMAGTIME ? 0.0 : tstep, integ_type == IMAGTIME ? -tstep : 0.0};
    const double tf[2] = {integ_type == IMAGTIME ? 0.0 : 1.0, integ_type == IMAGTIME ? -1.0 : 0.0};
    memcpy(&eq->prop_dt, pdt, sizeof(pdt));
    memcpy(&eq->time_fac, tf, sizeof(tf));
    eq->grid_pts = (Rarray) xmalloc(sizeof(double) * grid_size);
    for (int i = 0; i < grid_size; i++) eq->grid_pts[i] = xi + i * eq->dx;
    eq->pot_grid = (Rarray) xmalloc(sizeof(double) * grid_size);
    pot_func(0, grid_size, eq->grid_pts, pot_extra_args, eq->pot_grid);
    eq->pot_extra_args = pot_extra_args;
    eq->inter_extra_args = inter_extra_args;
    eq->d2coef = d2coef;
    eq->d1coef = d1coef;
    eq->g = inter_param(0, inter_extra_args);
    eq->pot_func = pot_func;
    eq->inter_param = inter_param;
    return eq;
}

extern "C" void destroy_orbital_equation(OrbitalEquation eq) {
    if (!eq) return;
    {
        std::lock_guard<std::mutex> lock(g_orb_mu);
        auto it = g_orb.find(eq);
        if (it != g_orb.end()) {
            hipDeviceSynchronize();
            free_dev_orbital(it->second);
            g_orb.erase(it);
        }
    }
    free(eq->grid_pts);
    free(eq->pot_grid);
    free(eq);
}

extern "C" OrbitalWorkspace get_orbital_workspace(OrbitalEquation eq, OrbDerivative der_method) {
    DevOrbital* d = dev_orbital(eq);
    const int grid = eq->grid_size, n = grid - 1, M = eq->norb;
    OrbitalWorkspace w = (OrbitalWorkspace) xmalloc(sizeof(_OrbitalWorkspace));
    w->norb = M;
    w->grid_size = grid;
    w->orb_der_method = der_method;
    w->impr_ortho = FALSE;
    w->dvr_mat = nullptr;  // the DVR derivative is not provided
    w->cn_upper = (Carray) xmalloc(sizeof(dcomplex) * grid);
    w->cn_lower = (Carray) xmalloc(sizeof(dcomplex) * grid);
    w->cn_mid = (Carray) xmalloc(sizeof(dcomplex) * grid);
    w->cn_rhs = nullptr;
    w->orb_work1 = alloc_cmatrix(M, grid);
    w->orb_work2 = alloc_cmatrix(M, grid);
    w->extern_work = nullptr;
    w->fft_desc = d;
    w->fft_freq = (Rarray) xmalloc(sizeof(double) * n);
    w->fft_hder_exp = (Carray) xmalloc(sizeof(dcomplex) * n);
    // the constant arrays the device workspace was built with
    const OrbWork* ow = d->o->w;
    d2h(w->cn_upper, ow->d_cn, sizeof(dcomplex) * grid);
    d2h(w->cn_lower, ow->d_cn + grid, sizeof(dcomplex) * grid);
    d2h(w->cn_mid, ow->d_cn + 2 * grid, sizeof(dcomplex) * grid);
    d2h(w->fft_hder_exp, ow->d_fexp, sizeof(dcomplex) * n);
    for (int i = 0; i < n; i++)  // calculus.c:272-285
        w->fft_freq[i] = (i <= (n - 1) / 2) ? (2 * M_PI * i) / (n * eq->dx) : (2 * M_PI * (i - n)) / (n * eq->dx);
    return w;
}

extern "C" void destroy_orbital_workspace(OrbitalWorkspace w) {
    if (!w) return;
    free(w->cn_upper);
    free(w->cn_lower);
    free(w->cn_mid);
    free_cmatrix(w->norb, w->orb_work1);
    free_cmatrix(w->norb, w->orb_work2);
    free(w->fft_freq);
    free(w->fft_hder_exp);
    free(w);  // the device side belongs to the orbital equation
}

extern "C" MCTDHBDataStruct get_mctdhb_struct(IntegratorType integ_type, CoefIntegrator coef_integ_method,
                                              OrbIntegrator orb_integ_method, OrbDerivative orb_der_method,
                                              RungeKuttaOrder rk_order, uint16_t lanczos_iter, uint16_t npar,
                                              uint16_t norb, char eq_name[], BoundaryCondition bounds, double xi,
                                              double xf, uint16_t grid_size, double tstep, double tend,
                                              double d2coef, dcomplex d1coef, void* pot_extra_args,
                                              void* inter_extra_args, single_particle_pot pot_func,
                                              time_dependent_parameter inter_param) {
    const uint32_t dim = space_dimension(npar, norb);
    MCTDHBDataStruct m = (MCTDHBDataStruct) xmalloc(sizeof(_MCTDHBDataStruct));
    m->integ_type = integ_type;
    m->coef_integ_method = coef_integ_method;
    m->orb_integ_method = orb_integ_method;
    m->orb_der_method = orb_der_method;
    m->rk_order = rk_order;
    m->orb_eq = get_orbital_equation(eq_name, norb, grid_size, bounds, integ_type, xi, xf, tstep, tend, d2coef,
                                     d1coef, pot_extra_args, inter_extra_args, pot_func, inter_param);
    if (coef_integ_method == LANCZOS && (lanczos_iter > MAX_LANCZOS_ITER || lanczos_iter < 2)) {
        printf("\n\nERROR: Lanczos iterations must be between %d and %d but %u was requested\n\n", 2,
               MAX_LANCZOS_ITER, lanczos_iter);
        exit(EXIT_FAILURE);
    }
    m->coef_workspace = get_coef_workspace(coef_integ_method, dim, lanczos_iter);
    m->multiconfig_space = get_multiconf_struct(npar, norb);
    m->state = get_manybody_state(npar, norb, grid_size);
    m->orb_workspace = get_orbital_workspace(m->orb_eq, orb_der_method);
    return m;
}

extern "C" void set_mctdhb_integrator(IntegratorType integ_type, CoefIntegrator coef_integ_method,
                                      OrbIntegrator orb_integ_method, OrbDerivative orb_der_method,
                                      RungeKuttaOrder rk_order, BoundaryCondition bounds, uint16_t lanczos_iter,
                                      MCTDHBDataStruct m) {
    m->integ_type = integ_type;
    m->coef_integ_method = coef_integ_method;
    m->orb_integ_method = orb_integ_method;
    m->orb_der_method = orb_der_method;
    m->rk_order = rk_order;
    m->orb_eq->bounds = bounds;
    if (m->coef_integ_method == LANCZOS) m->coef_workspace->lan_work->iter = lanczos_iter;
}

extern "C" void destroy_mctdhb_struct(MCTDHBDataStruct m) {
    if (!m) return;
    destroy_orbital_workspace(m->orb_workspace);
    destroy_orbital_equation(m->orb_eq);
    destroy_coef_workspace(m->coef_workspace);
    destroy_multiconf_struct(m->multiconfig_space);
    destroy_manybody_sate(m->state);
    free(m);
}

extern "C" void set_regulatization_factor(double reg_fac) {
    if (reg_fac > 1e-6 || reg_fac < 1e-13) {  // synchronize.h:19-21
        printf("\n\nERROR: Regularization factor must lies between %.1E and %.1E but %.1E was given\n\n", 1e-13,
               1e-6, reg_fac);
        exit(EXIT_FAILURE);
    }
    g_reg_factor = reg_fac;
}

extern "C" double get_current_regularization_factor(void) { return g_reg_factor; }

extern "C" void set_periodic_bounds(uint16_t norb, uint16_t grid_size, Cmatrix orb) {
    for (int k = 0; k < norb; k++) orb[k][grid_size - 1] = orb[k][0];
}

extern "C" void advance_linear_crank_nicolson(OrbitalEquation eq, OrbitalWorkspace w, Cmatrix orb) {
    (void) w;
    DevOrbital* d = dev_orbital(eq);
    orbitals_h2d(d->d_orb, orb, eq->norb, eq->grid_size);
    if (mctdhb_linear_cn_dev(d->o, d->d_orb, nullptr)) fail_stop("advance_linear_crank_nicolson");
    orbitals_d2h(orb, d->d_orb, eq->norb, eq->grid_size);
}

extern "C" void advance_linear_fft(OrbitalWorkspace w, Cmatrix orb) {
    DevOrbital* d = (DevOrbital*) w->fft_desc;
    orbitals_h2d(d->d_orb, orb, w->norb, w->grid_size);
    if (mctdhb_linear_fft_dev(d->o, d->d_orb, nullptr)) fail_stop("advance_linear_fft");
    orbitals_d2h(orb, d->d_orb, w->norb, w->grid_size);
}

extern "C" void sync_orbital_matrices(OrbitalEquation eq, ManyBodyState psi) {
    DevOrbital* d = dev_orbital(eq);
    const int M = psi->norb;
    orbitals_h2d(d->d_orb, psi->orbitals, M, psi->grid_size);
    if (mctdhb_orbital_matrices_dev(d->o, d->d_orb, d->d_hob, d->d_hint, nullptr)) fail_stop("sync_orbital_matrices");
    cmatrix_d2h(psi->hob, d->d_hob, M);
    d2h(psi->hint, d->d_hint, sizeof(dcomplex) * M * M * M * M);
}

extern "C" void sync_density_matrices(MultiConfiguration mc, ManyBodyState psi) {
    mctdhb_space* sp = ensure_dev(mc);
    const size_t n = sp->S->dim;
    const int M = psi->norb;
    stage_alloc(&sp->d_coef, n);
    stage_alloc(&sp->d_rho1, (size_t) M * M);
    stage_alloc(&sp->d_rho2, (size_t) M * M * M * M);
    stage_alloc(&sp->d_out, n > (size_t) M * M ? n : (size_t) M * M);
    h2d(sp->d_coef, psi->coef, sizeof(dcomplex) * n);
    if (mctdhb_density_sync_dev(sp, sp->d_coef, sp->d_rho1, sp->d_rho2, sp->d_out, nullptr, g_reg_factor,
                                MCTDHB_FP64, nullptr))
        fail_stop("sync_density_matrices");
    cmatrix_d2h(psi->ob_denmat, sp->d_rho1, M);
    d2h(psi->tb_denmat, sp->d_rho2, sizeof(dcomplex) * M * M * M * M);
    cmatrix_d2h(psi->inv_ob_denmat, sp->d_out, M);
}

extern "C" void sync_equation_params(OrbitalEquation eq) {
    eq->g = eq->inter_param(eq->t, eq->inter_extra_args);
    eq->pot_func(eq->t, eq->grid_size, eq->grid_pts, eq->pot_extra_args, eq->pot_grid);
}

// The integrator settings of m on the device workspace of d.  Finite
// differences select Crank-Nicolson (split step) or the finite-difference
// operator (full step), the other derivative methods the FFT
// (orbital_integration.c:375-381, :142-175); DVR is not provided.
static int configure_orbital(MCTDHBDataStruct m, DevOrbital* d) {
    if (m->orb_integ_method != SPLITSTEP && m->orb_integ_method != FULLSTEP_RUNGEKUTTA) {
        printf("\n\nERROR: unknown orbital integrator %d\n\n", (int) m->orb_integ_method);
        exit(EXIT_FAILURE);
    }
    if (mctdhb_orbital_set_rk_order(d->o, m->rk_order)) fail_stop("Runge-Kutta order");
    if (mctdhb_orbital_set_integrator(d->o, m->orb_integ_method, m->orb_der_method, m->orb_workspace->impr_ortho))
        fail_stop("orbital integrator");
    return m->orb_der_method != FINITEDIFF;
}

// The orbital step of m on the device from the host state; leaves the
// orbitals (and, for the split step, h and g of the intermediate
// orbitals) in the device staging of d.
static void orbital_step_from_host(MCTDHBDataStruct m, DevOrbital* d) {
    ManyBodyState psi = m->state;
    const int M = psi->norb;
    const int use_fft = configure_orbital(m, d);
    orbitals_h2d(d->d_orb, psi->orbitals, M, psi->grid_size);
    h2d(d->d_rho2, psi->tb_denmat, sizeof(dcomplex) * M * M * M * M);
    cmatrix_h2d(d->d_rinv, psi->inv_ob_denmat, M);
    if (orbital_step(d->o->w, use_fft, d->d_orb, d->d_hob, d->d_hint, d->d_rho2, d->d_rinv, nullptr, false))
        fail_stop("orbital step");
}

// The split step alone (orbital_integration.c:351-434).
static void splitstep_dev_from_host(MCTDHBDataStruct m, DevOrbital* d) {
    if (m->orb_integ_method != SPLITSTEP) {
        printf("\n\nERROR: propagate_splitstep_orb needs the SPLITSTEP orbital integrator\n\n");
        exit(EXIT_FAILURE);
    }
    orbital_step_from_host(m, d);
}

extern "C" void propagate_splitstep_orb(MCTDHBDataStruct m, Carray orb_next) {
    DevOrbital* d = dev_orbital(m->orb_eq);
    splitstep_dev_from_host(m, d);
    ManyBodyState psi = m->state;
    const int M = psi->norb;
    orbitals_d2h(psi->orbitals, d->d_orb, M, psi->grid_size);
    cmatrix_d2h(psi->hob, d->d_hob, M);
    d2h(psi->hint, d->d_hint, sizeof(dcomplex) * M * M * M * M);
    // orb_next is the reference's row-major scratch; it receives the new orbitals
    d2h(orb_next, d->d_orb, sizeof(dcomplex) * M * psi->grid_size);
}

namespace {

// composite Simpson rule of calculus.c:8-39 with its summation order
double2 simps_host(int n, double h, const double2* f) {
    double sr = 0, si = 0;
    const int last = n % 2 == 0 ? n - 4 : n - 2;
    for (int i = 0; i < last; i += 2) {
        sr += f[i].x + 4 * f[i + 1].x + f[i + 2].x;
        si += f[i].y + 4 * f[i + 1].y + f[i + 2].y;
    }
    sr *= h / 3;
    si *= h / 3;
    if (n % 2 == 0) {
        sr += (f[n - 4].x + 3 * (f[n - 3].x + f[n - 2].x) + f[n - 1].x) * 3 * h / 8;
        si += (f[n - 4].y + 3 * (f[n - 3].y + f[n - 2].y) + f[n - 1].y) * 3 * h / 8;
    }
    return make_double2(sr, si);
}

// <f, g> = Simpson of conj(f) g (calculus.c:130-142)
double2 scalar_product_host(int n, double h, const dcomplex* f, const dcomplex* g) {
    std::vector<double2> t(n);
    const double* a = (const double*) f;
    const double* b = (const double*) g;
    for (int i = 0; i < n; i++)
        t[i] = make_double2(a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1],
                            a[2 * i] * b[2 * i + 1] - a[2 * i + 1] * b[2 * i]);
    return simps_host(n, h, t.data());
}

}  // namespace

extern "C" void propagate_fullstep_orb_rk(MCTDHBDataStruct m, Carray orb_next) {
    if (m->orb_integ_method != FULLSTEP_RUNGEKUTTA) {
        printf("\n\nERROR: propagate_fullstep_orb_rk needs the FULLSTEP_RUNGEKUTTA orbital integrator\n\n");
        exit(EXIT_FAILURE);
    }
    DevOrbital* d = dev_orbital(m->orb_eq);
    orbital_step_from_host(m, d);
    ManyBodyState psi = m->state;
    orbitals_d2h(psi->orbitals, d->d_orb, psi->norb, psi->grid_size);
    d2h(orb_next, d->d_orb, sizeof(dcomplex) * psi->norb * psi->grid_size);
}

static void projector_from_host(OrbitalEquation eq, uint16_t norb, Cmatrix orb, Cmatrix hact, Cmatrix project,
                                int robust) {
    DevOrbital* d = dev_orbital(eq);
    const int grid = eq->grid_size;
    double2* d_h = nullptr;
    stage_alloc(&d_h, (size_t) norb * grid);
    orbitals_h2d(d->d_orb, orb, norb, grid);
    orbitals_h2d(d_h, hact, norb, grid);
    if (mctdhb_projector_dev(d->o, robust, d->d_orb, d_h, d_h, nullptr)) fail_stop("projector");
    orbitals_d2h(project, d_h, norb, grid);
    hipFree(d_h);
}

extern "C" void robust_multiorb_projector(OrbitalEquation eq, uint16_t norb, Cmatrix orb, Cmatrix hact,
                                          Cmatrix project) {
    projector_from_host(eq, norb, orb, hact, project, 1);
}

extern "C" void simple_multiorb_projector(OrbitalEquation eq, uint16_t norb, Cmatrix orb, Cmatrix hact,
                                          Cmatrix project) {
    projector_from_host(eq, norb, orb, hact, project, 0);
}

// Writes exactly the entries the reference writes (orbital_matrices.c:7-20):
// the j loop starts at 2, so overlap[0][1] and overlap[1][0] keep what the
// caller's memory held.
extern "C" void set_overlap_matrix(uint16_t norb, uint16_t grid_size, double dx, Cmatrix orb, Cmatrix overlap) {
    auto put = [](dcomplex* dst, double2 v) { memcpy(dst, &v, sizeof v); };
    for (int i = 0; i < norb; i++) {
        put(&overlap[i][i], scalar_product_host(grid_size, dx, orb[i], orb[i]));
        for (int j = 2; j < norb; j++) {
            const double2 v = scalar_product_host(grid_size, dx, orb[i], orb[j]);
            put(&overlap[i][j], v);
            put(&overlap[j][i], make_double2(v.x, -v.y));
        }
    }
}

extern "C" void set_orbital_hob(OrbitalEquation eq, uint16_t norb, Cmatrix orb, Cmatrix hob) {
    DevOrbital* d = dev_orbital(eq);
    orbitals_h2d(d->d_orb, orb, norb, eq->grid_size);
    if (mctdhb_orbital_matrices_dev(d->o, d->d_orb, d->d_hob, d->d_hint, nullptr)) fail_stop("set_orbital_hob");
    cmatrix_d2h(hob, d->d_hob, norb);
}

extern "C" void set_orbital_hint(OrbitalEquation eq, uint16_t norb, Cmatrix orb, Carray hint) {
    DevOrbital* d = dev_orbital(eq);
    orbitals_h2d(d->d_orb, orb, norb, eq->grid_size);
    if (mctdhb_orbital_matrices_dev(d->o, d->d_orb, d->d_hob, d->d_hint, nullptr)) fail_stop("set_orbital_hint");
    d2h(hint, d->d_hint, sizeof(dcomplex) * norb * norb * norb * norb);
}

extern "C" void mctdhb_propagate_step(MCTDHBDataStruct m, Carray orb_works, Carray coef_works, uint32_t curr_step) {
    ManyBodyState psi = m->state;
    OrbitalEquation eq = m->orb_eq;
    const int M = psi->norb, grid = psi->grid_size;
    if (m->coef_integ_method != LANCZOS && M > 1) {
        printf("\n\nERROR: only the LANCZOS coefficient integrator is provided\n\n");
        exit(EXIT_FAILURE);
    }
    DevOrbital* d = dev_orbital(eq);
    orbital_step_from_host(m, d);
    // driver.c:176-186: new time, equation parameters, periodic bounds, h and g
    eq->t = (curr_step + 1) * eq->tstep;
    if (m->integ_type == REALTIME) {
        sync_equation_params(eq);
        if (mctdhb_orbital_set_potential(d->o, eq->pot_grid) || mctdhb_orbital_set_interaction(d->o, eq->g))
            fail_stop("equation parameters");
    }
    if (eq->bounds == PERIODIC_BOUNDS && periodic_bounds(d->o->w, d->d_orb, nullptr)) fail_stop("periodic bounds");
    if (mctdhb_orbital_matrices_dev(d->o, d->d_orb, d->d_hob, d->d_hint, nullptr)) fail_stop("orbital matrices");
    orbitals_d2h(psi->orbitals, d->d_orb, M, grid);
    cmatrix_d2h(psi->hob, d->d_hob, M);
    d2h(psi->hint, d->d_hint, sizeof(dcomplex) * M * M * M * M);
    d2h(orb_works, d->d_orb, sizeof(dcomplex) * M * grid);
    if (M == 1) return;
    propagate_coef_sil(m, coef_works);
    sync_density_matrices(m->multiconfig_space, psi);
}

// ----------------------------------------------- layer 1: job loop
//
// integration_driver (reference driver.c:203-333) with the state kept on the
// GPU for the whole run: imaginary time replays one captured step graph per
// step; real time runs the step's pieces eagerly so the trap and g of the
// new time can be evaluated on the host between the orbital step and the
// h/g rebuild, as the reference orders them (driver.c:175-186).  The host
// structs receive the state at monitor, record and end points.

namespace {

Bool g_final_diag = TRUE;          // driver.c:15
Bool g_auto_convergence = FALSE;   // driver.c:16
uint8_t g_energy_digits = 11;      // driver.c:17
double g_eig_residue_tol = 1e-6;   // driver.c:18
double g_impr_ortho_thr = 1e-8;    // driver.c:19

}  // namespace

extern "C" double overlap_residual(uint16_t norb, uint16_t grid_size, double dx, Cmatrix orb) {
    double sum = 0;
    for (int k = 0; k < norb; k++)
        for (int l = k + 1; l < norb; l++) {
            const double2 o = scalar_product_host(grid_size, dx, orb[k], orb[l]);
            sum += hypot(o.x, o.y);
        }
    return sum / norb;
}

extern "C" double avg_orbitals_norm(uint16_t norb, uint16_t grid_size, double dx, Cmatrix orb) {
    double sum = 0;
    std::vector<double2> m(grid_size);
    for (int k = 0; k < norb; k++) {
        const double* f = (const double*) orb[k];
        for (int i = 0; i < grid_size; i++) m[i] = make_double2(f[2 * i] * f[2 * i] + f[2 * i + 1] * f[2 * i + 1], 0);
        sum += sqrt(simps_host(grid_size, dx, m.data()).x);  // cplx_function_norm (calculus.c:144-155)
    }
    return sum / norb;
}

extern "C" void set_final_diagonalization(Bool shall_diag) { g_final_diag = shall_diag; }

extern "C" void set_autoconvergence_check(Bool must_check) { g_auto_convergence = must_check; }

extern "C" void set_energy_convergence_digits(uint8_t edig) {
    if (edig < 4) printf("\n\nWARNING: %u digits is too small for energy convergence criteria.\n\n", edig);
    if (edig > 13) {
        printf("\n\nERROR: %u digits for convergence is too much due to roundoff.\n\n", edig);
        exit(EXIT_FAILURE);
    }
    g_energy_digits = edig;
}

extern "C" void set_energy_convergence_eig_residual(double eig_res) {
    if (eig_res > 0.01)
        printf("\n\nWARNING: Too large eigenvalue residual to define convergence %.2lf\n\n", eig_res);
    g_eig_residue_tol = eig_res;
}

extern "C" void set_overlap_residue_threshold(double over_res) {
    if (over_res > REALTIME_OVERLAP_RESIDUE_TOL) {
        printf("\n\nERROR: Overlap residue cannot exceed %.3lf but %.3lf was given\n\n", REALTIME_OVERLAP_RESIDUE_TOL,
               over_res);
        exit(EXIT_FAILURE);
    }
    g_impr_ortho_thr = over_res;
}

namespace {

// Krylov size of the closing diagonalisation (driver.c:38-44)
uint16_t closing_lanczos_iterations(uint32_t dim) {
    if (dim / 2 > 50) return 50;
    if (dim < 4) return 0;
    return (uint16_t) (dim / 2);
}

// Device state of one run of the job loop.
struct RunState {
    MCTDHBDataStruct m = nullptr;
    mctdhb_space* sp = nullptr;
    DevOrbital* d = nullptr;
    mctdhb_sil* sil = nullptr;  // the run's own SIL workspace
    double2* d_rho1 = nullptr;
    double2* d_tmp = nullptr;   // 2 M^2 scratch of the monitor
    double* d_occ = nullptr;    // natural occupations for the monitor
    int use_fft = 1;

    // the reference's step (driver.c:158-202); with params set, the new
    // trap and g are evaluated after the orbital step
    int step(uint32_t curr_step, bool params, hipStream_t st) {
        OrbitalEquation eq = m->orb_eq;
        const int M = m->state->norb;
        // the periodic copy (driver.c:179-182) is folded into the orbital
        // step; the trap and g below do not touch the orbitals
        int rc = orbital_step(d->o->w, use_fft, d->d_orb, d->d_hob, d->d_hint, d->d_rho2, d->d_rinv, st, true);
        if (rc) return rc;
        if (params) {
            eq->t = (curr_step + 1) * eq->tstep;
            sync_equation_params(eq);
            if ((rc = mctdhb_orbital_set_potential(d->o, eq->pot_grid)) ||
                (rc = mctdhb_orbital_set_interaction(d->o, eq->g)))
                return rc;
        }
        if ((rc = mctdhb_orbital_matrices_dev(d->o, d->d_orb, d->d_hob, d->d_hint, st))) return rc;
        if (M == 1) return kOk;
        const double2 dt = d->o->w->eq.prop_dt;
        if ((rc = mctdhb_sil_step_dev(sil, sp->d_coef, d->d_hob, d->d_hint, dt.x, dt.y, d->o->w->imagtime, st)))
            return rc;
        return mctdhb_density_sync_dev(sp, sp->d_coef, d_rho1, d->d_rho2, d->d_rinv, nullptr, g_reg_factor,
                                       MCTDHB_FP64, st);
    }

    // device -> host structs; the coefficients only when asked
    void to_host(bool coef) {
        ManyBodyState psi = m->state;
        const int M = psi->norb;
        if (hipDeviceSynchronize() != hipSuccess) fail_stop("integration_driver");
        orbitals_d2h(psi->orbitals, d->d_orb, M, psi->grid_size);
        cmatrix_d2h(psi->hob, d->d_hob, M);
        d2h(psi->hint, d->d_hint, sizeof(dcomplex) * M * M * M * M);
        if (M > 1) {
            cmatrix_d2h(psi->ob_denmat, d_rho1, M);
            d2h(psi->tb_denmat, d->d_rho2, sizeof(dcomplex) * M * M * M * M);
            cmatrix_d2h(psi->inv_ob_denmat, d->d_rinv, M);
        }
        if (coef) d2h(psi->coef, sp->d_coef, sizeof(dcomplex) * psi->space_dim);
    }

    // one monitor line: [t / tend] E/N, largest and smallest natural
    // occupation over N, then the eigenvalue residue (imaginary time) or the
    // overlap residue and the norms (real time); returns E/N and the residue
    double monitor(double* eig_res_out) {
        ManyBodyState psi = m->state;
        OrbitalEquation eq = m->orb_eq;
        const int M = psi->norb, N = psi->npar;
        to_host(m->integ_type == REALTIME);
        double occ[MAX_ORBITALS] = {(double) N};
        if (M > 1) {  // eigenvalues of rho1 from a copy (no regularisation)
            if (hipMemcpy(d_tmp, d_rho1, sizeof(double2) * M * M, hipMemcpyDeviceToDevice) != hipSuccess ||
                density_sync_small(M, d_tmp, d_tmp + M * M, d_occ, 0.0, nullptr))
                fail_stop("natural occupations");
            d2h(occ, d_occ, sizeof(double) * M);
        }
        const dcomplex e = total_energy(psi);
        double ev[2];
        memcpy(ev, &e, sizeof ev);
        char fmt[32];
        snprintf(fmt, sizeof fmt, " %%%d.%dE", monitor_energy_digits + 7, monitor_energy_digits);
        printf("\n[%7.3lf / %.1lf]", eq->t, eq->tend);
        printf(fmt, ev[0] / N);
        printf("%6.2lf%10.6lf", occ[M - 1] / N, occ[0] / N);
        double res = 0;
        if (m->integ_type == IMAGTIME) {
            if (M > 1 && mctdhb_eig_residual_dev(sp, sp->d_coef, d->d_hob, d->d_hint, ev[0], &res, nullptr))
                fail_stop("eig_residual");
            printf("%11.7lf", res);
        } else {
            double c2 = 0;
            const double* c = (const double*) psi->coef;
            for (uint32_t i = 0; i < 2 * psi->space_dim; i++) c2 += c[i] * c[i];
            printf("%10.2E%11.7lf%11.7lf", overlap_residual(M, psi->grid_size, eq->dx, psi->orbitals),
                   avg_orbitals_norm(M, psi->grid_size, eq->dx, psi->orbitals), sqrt(c2));
        }
        fflush(stdout);
        if (eig_res_out) *eig_res_out = res;
        return ev[0] / N;
    }
};

}  // namespace

extern "C" void integration_driver(MCTDHBDataStruct m, uint16_t rec_nsteps, char prefix[], double tend,
                                   uint8_t monitor_rate) {
    ManyBodyState psi = m->state;
    OrbitalEquation eq = m->orb_eq;
    const int M = psi->norb, grid = psi->grid_size;
    const bool imag = m->integ_type == IMAGTIME;
    if (m->coef_integ_method != LANCZOS && M > 1) {
        printf("\n\nERROR: only the LANCZOS coefficient integrator is provided\n\n");
        exit(EXIT_FAILURE);
    }
    eq->tend = tend;
    if (rec_nsteps == 0) rec_nsteps = 1;
    if (monitor_rate == 0) monitor_rate = 1;

    RunState R;
    R.m = m;
    R.sp = ensure_dev(m->multiconfig_space);
    R.d = dev_orbital(eq);
    R.use_fft = configure_orbital(m, R.d);
    const size_t n = psi->space_dim;
    stage_alloc(&R.sp->d_coef, n);
    stage_alloc(&R.d_rho1, (size_t) M * M);
    stage_alloc(&R.d_tmp, 2 * (size_t) M * M);
    if (hipMalloc((void**) &R.d_occ, sizeof(double) * M) != hipSuccess) fail_stop("integration_driver");
    orbitals_h2d(R.d->d_orb, psi->orbitals, M, grid);
    cmatrix_h2d(R.d->d_hob, psi->hob, M);
    h2d(R.d->d_hint, psi->hint, sizeof(dcomplex) * M * M * M * M);
    h2d(R.sp->d_coef, psi->coef, sizeof(dcomplex) * n);
    if (M > 1) {
        cmatrix_h2d(R.d_rho1, psi->ob_denmat, M);
        h2d(R.d->d_rho2, psi->tb_denmat, sizeof(dcomplex) * M * M * M * M);
        cmatrix_h2d(R.d->d_rinv, psi->inv_ob_denmat, M);
        const int lm = m->coef_workspace->lan_work->iter;
        if (mctdhb_sil_create(R.sp, lm, MCTDHB_FP64, MCTDHB_INDEX_FLY, &R.sil)) fail_stop("Lanczos workspace");
        if (!R.sp->dm_work && hipMalloc(&R.sp->dm_work, dm_workspace_bytes(R.sp->S)) != hipSuccess)
            fail_stop("density workspace");
    }

    const dcomplex e0 = total_energy(psi);
    double prev_e;
    memcpy(&prev_e, &e0, sizeof prev_e);  // real part
    prev_e /= psi->npar;
    printf("\nColumns: [time] E/N, largest and smallest natural occupation / N, %s\n",
           imag ? "eigenvalue residue" : "overlap residue, orbital norm, coefficient norm");
    R.monitor(nullptr);
    // initial conditions (driver.c:236-251)
    if (imag) {
        std::string init = std::string(prefix) + "_init";
        record_raw_state((char*) init.c_str(), psi);
    } else {
        toggle_new_append_files();
        append_processed_state(prefix, psi);
        append_timestep_potential(prefix, eq);
        toggle_new_append_files();
    }

    mctdhb_step_graph graph;
    bool have_graph = false;
    if (imag && M <= 7) {  // larger orbital counts use a density path that synchronises
        const int rc = capture_graph([&](hipStream_t cs) { return R.step(0, false, cs); }, &graph);
        if (rc) fail_stop("step graph");
        have_graph = true;
    }

    const auto t0 = std::chrono::steady_clock::now();
    uint32_t steps = 0;
    double t = 0;
    while (t < tend) {
        if (have_graph) {
            if (hipGraphLaunch(graph.exec, nullptr) != hipSuccess) fail_stop("step graph launch");
            eq->t = (steps + 1) * eq->tstep;
        } else {
            if (R.step(steps, !imag, nullptr)) fail_stop("mctdhb_propagate_step");
            if (imag) eq->t = (steps + 1) * eq->tstep;
        }
        steps++;
        t = steps * eq->tstep;
        if (!imag && steps % rec_nsteps == 0) {
            R.to_host(false);
            append_processed_state(prefix, psi);
            append_timestep_potential(prefix, eq);
        }
        if (steps % monitor_rate == 0) {
            double res = 0;
            const double e = R.monitor(&res);
            if (imag && g_auto_convergence) {  // driver.c:46-67
                const double etol = pow(0.1, (double) g_energy_digits) * prev_e;
                if (fabs(e - prev_e) < etol && res < g_eig_residue_tol) break;
                prev_e = e;
            }
            if (!imag) {  // driver.c:69-94
                const double o = overlap_residual(M, grid, eq->dx, psi->orbitals);
                if (o > g_impr_ortho_thr && !m->orb_workspace->impr_ortho) {
                    printf("\n\n== Start improving orthogonality ==\n\n");
                    m->orb_workspace->impr_ortho = TRUE;  // the full-equation step's robust projector
                    if (mctdhb_orbital_set_integrator(R.d->o, m->orb_integ_method, m->orb_der_method, 1))
                        fail_stop("projector");
                }
                if (o > REALTIME_OVERLAP_RESIDUE_TOL) {
                    printf("\n\nOverlap residue fell below critical value %.3lf. Aborting\n\n",
                           REALTIME_OVERLAP_RESIDUE_TOL);
                    exit(EXIT_FAILURE);
                }
            }
        }
    }
    if (hipDeviceSynchronize() != hipSuccess) fail_stop("integration_driver");
    const double used = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (have_graph) {
        hipGraphExecDestroy(graph.exec);
        hipGraphDestroy(graph.graph);
    }
    const int secs = (int) used;
    printf("\n\n%u Total steps propagated in %.0lf(s): %d hour(s) %d minute(s)", steps, used, secs / 3600,
           (secs % 3600) / 60);

    if (imag) {  // driver.c:294-323
        if (g_auto_convergence && t < tend)
            printf("\nThe result converged before expected with t = %.1lf of %.1lf total propagation time", t,
                   tend);
        const uint16_t lan_it = closing_lanczos_iterations((uint32_t) n);
        if (g_final_diag && t >= tend && lan_it > 0 && M > 1) {
            mctdhb_sil* big = nullptr;
            if (mctdhb_sil_create(R.sp, lan_it, MCTDHB_FP64, MCTDHB_INDEX_FLY, &big)) fail_stop("Lanczos workspace");
            double e = 0;
            if (mctdhb_lowest_state_dev(big, R.sp->d_coef, R.d->d_hob, R.d->d_hint, &e, nullptr)) {
                printf("\n\nERROR IN DIAGONALIZATION\n\n");
                exit(EXIT_FAILURE);
            }
            mctdhb_sil_destroy(big);
            printf("\n\n== Additional coefficients diagonalization ==\n");
            R.monitor(nullptr);
        }
        R.to_host(true);
        char fname[STR_BUFF_SIZE];
        set_output_fname(prefix, ONE_BODY_POTENTIAL_REC, fname);
        FILE* f = fopen(fname, "w");
        if (!f) {
            printf("\n\nIOERROR: impossible to open file %s with mode w\n\n", fname);
            exit(EXIT_FAILURE);
        }
        for (int i = 0; i < grid; i++) fprintf(f, "%.14E\n", eq->pot_grid[i]);
        fclose(f);
        record_raw_state(prefix, psi);
    } else {
        R.to_host(true);
        record_time_interaction(prefix, eq, rec_nsteps);
        record_time_array(prefix, tend, eq->tstep, rec_nsteps);
    }
    if (R.sil) mctdhb_sil_destroy(R.sil);
    hipFree(R.d_rho1);
    hipFree(R.d_tmp);
    hipFree(R.d_occ);
}
