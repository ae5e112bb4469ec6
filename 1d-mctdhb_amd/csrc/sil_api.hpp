// sil_api.hpp - host driver of the device SIL step.
#pragma once

#include "sil.hpp"
#include "space.hpp"

namespace mctdhb {

struct Sil {
    const Space* S = nullptr;
    int lm = 0, prec = 0, index = 0;
    bool generic = false;
    bool small = true;  // small spaces take the one-workgroup step (mctdhb_sil_set_small)
    SilState* d_st = nullptr;
    double2* cf64 = nullptr;  // fused coefficient block, or [hob | hint] for the generic kernel
    float2* cf32 = nullptr;
    // w[0] is C (FP64) or its complex64 copy w0f.  Every vector this
    // workspace allocates has one zero element before and after it (the
    // paired H.C kernel reads w[-1] and w[dim]); wbase[] are the allocations.
    void* w[kMaxKrylov] = {};
    void* wbase[kMaxKrylov] = {};
    void* w0f = nullptr;
    void* w0fbase = nullptr;
    double2* out = nullptr;
    double* part = nullptr;
    size_t part_cap = 0;
    double* tot = nullptr;
};

int sil_create(const Space* S, int lm, int prec, int index, Sil** out);
void sil_destroy(Sil* L);
size_t sil_device_bytes(const Space* S, int lm, int prec);
// One SIL step C <- exp(-i H prop_dt) C (renormalised in imaginary time).
// C, hob (M x M), hint (M^4) are complex128 device arrays.
int sil_step(Sil* L, double2* C, const double2* hob, const double2* hint, double dt_re, double dt_im, int imagtime,
             hipStream_t st);
// C <- the Ritz vector of the lowest Ritz value (not renormalised); the
// value is left in SilState::eground.
int sil_ground(Sil* L, double2* C, const double2* hob, const double2* hint, hipStream_t st);
// Lanczos recursion only; lvec (lm x dim complex128, device) receives the
// normalised vectors.
int sil_lanczos(Sil* L, const double2* C, const double2* hob, const double2* hint, double2* lvec, hipStream_t st);
int sil_read_state(const Sil* L, SilState* host, hipStream_t st);

}  // namespace mctdhb
