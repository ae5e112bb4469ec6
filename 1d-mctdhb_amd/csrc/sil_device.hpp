// sil_device.hpp - the scalar part of the Lanczos / SIL recursion on the
// device: state initialisation, the updates after the H.C and the
// reorthogonalisation passes, the tridiagonal eigensolver and the
// exponential.  Shared by the multi-kernel step (lanczos.hip) and the
// one-workgroup step for small spaces (sil_small.hpp).
#pragma once

#include "common.hpp"
#include "sil.hpp"

namespace mctdhb {

constexpr double kBreakdownTol = 1e-14;  // reference multiconfig_lanczos.c:34

// Fresh scalars for an lm-vector recursion (one thread).
__device__ inline void sil_init_state(SilState* st, int lm) {
    for (int k = 0; k <= kMaxKrylov; k++) {
        st->s[k] = 0;
        st->beta[k] = 0;
        st->alpha[k] = make_double2(0, 0);
        st->cre[k] = make_double2(0, 0);
        st->gram[k] = make_double2(0, 0);
        st->y[k] = make_double2(0, 0);
    }
    st->s[0] = 1.0;
    st->maxbeta = 0;
    st->nrm2 = 0;
    st->eground = 0;
    st->iter_done = lm;
    st->arrive = 0;
    st->stop = 0;
    st->eig_fail = 0;
}

// After the H.C pass of iteration it: alpha_it and the reorthogonalisation
// coefficients (on w_k) for the next streaming pass.
// On the last iteration only alpha is formed.  last == 2: the H.C pass
// summed <w_it, H w_it> (the half-row form), and alpha = <v_it, H v_it -
// beta v_{it-1}> = s_it^2 <w_it, H w_it> - beta conj(<v_{it-1}, v_it>).
__device__ inline void sil_after_hc(SilState* st, const volatile double* tot, int it, int last) {
    if (it == 0) st->gram[0] = make_double2(tot[2 * (it + 1)], 0);  // <v_0, v_0> = |C|^2
    double2 alpha = make_double2(tot[2 * it], tot[2 * it + 1]);
    if (last == 2) {
        const double s = st->s[it];
        alpha = make_double2(s * s * tot[2 * it], 0);
        if (it > 0) {
            const double2 gp = st->gram[it - 1];
            alpha = alpha - make_double2(st->beta[it - 1] * gp.x, -st->beta[it - 1] * gp.y);
        }
    }
    st->alpha[it] = alpha;
    if (last) return;
    for (int k = 0; k <= it; k++) {
        const double2 d = make_double2(tot[2 * k], tot[2 * k + 1]);
        double2 ortho = d - cmul(alpha, st->gram[k]);
        if (k == it) ortho = ortho + alpha;
        st->cre[k] = cscale(ortho, st->s[k]);
    }
}

// After the streaming pass that produced w_{it+1}: beta_it, the breakdown
// test of the reference (beta / max beta < 1e-14 ends the recursion with
// it + 1 vectors) and the Gram column <v_k, v_{it+1}>.
__device__ inline void sil_after_reorth(SilState* st, const volatile double* tot, int it) {
    const double nrm2 = tot[2 * (it + 1)];
    const double beta = sqrt(nrm2);
    st->beta[it] = beta;
    if (st->maxbeta < beta) st->maxbeta = beta;
    if (beta / st->maxbeta < kBreakdownTol) {
        st->iter_done = it + 1;
        st->stop = 1;
        return;
    }
    const double sn = 1.0 / beta;
    st->s[it + 1] = sn;
    for (int k = 0; k <= it; k++)
        st->gram[k] = cscale(make_double2(tot[2 * k], tot[2 * k + 1]), st->s[k] * sn);
    st->gram[it + 1] = make_double2(nrm2 * sn * sn, 0);
}

// ------------------------------------------------ tridiagonal eigensolver
//
// Implicit-shift QL with Givens rotations (the EISPACK tql2 method) on one
// wave: lane r keeps d[r] and e[r] in registers and owns row r of the
// eigenvector matrix (in LDS), so each rotation updates all rows at once;
// the scalar recurrence runs redundantly in every lane on values broadcast
// with readlane.  n <= 64.

__device__ __forceinline__ double rlane(double x, int i) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), i);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), i);
    return __hiloint2double(hi, lo);
}

// On return lane r holds eigenvalue d (unsorted) and zr[0..n) is row r of Z
// (eigenvector j in column j).  Returns non-zero if an eigenvalue needed
// more than 60 sweeps.
__device__ inline int tridiag_eig_wave(int n, double& d, double& e, double* zr) {
    const int lane = threadIdx.x & 63;
    for (int c = 0; c < n; c++) zr[c] = (lane == c) ? 1.0 : 0.0;
    if (lane == n - 1) e = 0.0;
    for (int l = 0; l < n; l++) {
        int sweeps = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; m++) {
                const double scale = fabs(rlane(d, m)) + fabs(rlane(d, m + 1));
                if (fabs(rlane(e, m)) <= 2.220446049250313e-16 * scale) break;
            }
            if (m == l) break;
            if (++sweeps > 60) return 1;
            const double dl = rlane(d, l), el = rlane(e, l);
            // |theta| <= 1 / (2 eps) after the test above and the entries
            // of T are far from overflow: plain square roots of sums of
            // squares in place of hypot, one reciprocal per rotation
            const double theta = (rlane(d, l + 1) - dl) / (2.0 * el);
            const double rho = sqrt(theta * theta + 1.0);
            double g = rlane(d, m) - dl + el / (theta + (theta >= 0 ? rho : -rho));
            double sn = 1.0, cs = 1.0, shift = 0.0;
            bool early = false;
            for (int i = m - 1; i >= l; i--) {
                const double ei = rlane(e, i);
                const double f = sn * ei, b = cs * ei;
                double r = sqrt(f * f + g * g);
                if (lane == i + 1) e = r;
                if (r == 0.0) {
                    if (lane == i + 1) d -= shift;
                    if (lane == m) e = 0.0;
                    early = true;
                    break;
                }
                const double ir = 1.0 / r;
                sn = f * ir;
                cs = g * ir;
                g = rlane(d, i + 1) - shift;
                r = (rlane(d, i) - g) * sn + 2.0 * cs * b;
                shift = sn * r;
                if (lane == i + 1) d = g + shift;
                g = cs * r - b;
                if (lane < n) {
                    const double zi1 = zr[i + 1], zi = zr[i];
                    zr[i + 1] = sn * zi + cs * zi1;
                    zr[i] = cs * zi - sn * zi1;
                }
            }
            if (early) continue;
            if (lane == l) {
                d -= shift;
                e = g;
            }
            if (lane == m) e = 0.0;
        }
    }
    return 0;
}

// exp(-i T prop_dt) e_1 from the eigen-decomposition of the real tridiagonal
// T = (Re alpha, beta), as reference coefficients_integration.c:104-145:
// y = Z exp(-i D prop_dt) Z^T e_1.  Called by one whole wave (lanes 0..63);
// z is LDS scratch of n rows of zstride doubles.
//
// renorm: the caller renormalises the result (imaginary time).  Then the
// exponentials are taken relative to the dominant one, exp((d_k - d*) b)
// with d* the eigenvalue where d b is largest: a common factor that the
// renormalisation removes, and that keeps exp(d b) from underflowing to 0
// for every k (then C' = 0 and the renormalisation would give NaN) when
// |E| dt exceeds about 745.
__device__ inline void sil_tridiag_wave(SilState* st, double dt_re, double dt_im, double* z, int zstride,
                                        int renorm) {
    const int lane = threadIdx.x;
    const int n = st->iter_done;
    double d = lane < n ? st->alpha[lane].x : 0.0;
    double e = lane < n ? st->beta[lane] : 0.0;
    double* zr = z + lane * zstride;
    if (tridiag_eig_wave(n, d, e, zr)) {
        if (lane == 0) st->eig_fail = 1;
        return;
    }
    // lane k: aux_k = Z[0][k] exp(-i d_k prop_dt); exp(-i d dt) for complex
    // dt = a + ib is exp(d b) (cos(d a) - i sin(d a))
    double shift = -INFINITY;  // max over k of d_k b
    for (int j = 0; j < n; j++) shift = fmax(shift, rlane(d, j) * dt_im);
    if (!renorm) shift = 0.0;
    double ax = 0, ay = 0;
    if (lane < n) {
        const double z0k = z[lane];  // row 0, column lane
        const double mag = exp(d * dt_im - shift), ph = -d * dt_re;
        ax = z0k * mag * cos(ph);
        ay = z0k * mag * sin(ph);
    }
    double yx = 0, yy = 0;
    for (int j = 0; j < n; j++) {
        const double zkj = zr[j];
        yx += rlane(ax, j) * zkj;
        yy += rlane(ay, j) * zkj;
    }
    if (lane < n) st->y[lane] = make_double2(yx * st->s[lane], yy * st->s[lane]);
}

}  // namespace mctdhb
