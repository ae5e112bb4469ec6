// mxm_wave.hpp - M x M complex algebra with LDS operands for one block:
// Hermitian eigen-decomposition (cyclic Jacobi), Gauss-Jordan inverse,
// a positive-definiteness test, and the density-matrix synchronisation
// built from them (reference synchronize.c:41-56 with
// lapack_interface.c:10-163).  Lane r of the block owns row or column r
// (M <= 64); every thread of the block calls these functions.
#pragma once

#include "common.hpp"

namespace mctdhb {

__device__ __forceinline__ double2 dconj2(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    const double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// Hermitian eigen-decomposition by cyclic complex Jacobi rotations on one
// wave (n <= 64; a is overwritten, all in LDS): rotations run in the cyclic
// order, lane r applies each one to row/column r.  Eigenvalues ascending in
// lam (LDS), eigenvectors in the columns of v.  Every lane of the block
// must call it.
__device__ inline void herm_eig_wave(int n, double2* a, double* lam, double2* v, double* rowsum) {
    const int r = threadIdx.x;
    if (r < n)
        for (int c = 0; c < n; c++) v[r * n + c] = make_double2(r == c ? 1.0 : 0.0, 0.0);
    __syncthreads();
    for (int sweep = 0; sweep < 100; sweep++) {
        if (r < n) {
            double off = 0, tot = 0;
            for (int c = 0; c < n; c++) {
                const double m2 = a[r * n + c].x * a[r * n + c].x + a[r * n + c].y * a[r * n + c].y;
                tot += m2;
                if (r != c) off += m2;
            }
            rowsum[2 * r] = off;
            rowsum[2 * r + 1] = tot;
        }
        __syncthreads();
        double off = 0, tot = 0;
        for (int c = 0; c < n; c++) {
            off += rowsum[2 * c];
            tot += rowsum[2 * c + 1];
        }
        __syncthreads();
        if (off <= 1e-30 * tot || off == 0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double2 apq = a[p * n + q];
                const double mag = sqrt(apq.x * apq.x + apq.y * apq.y);
                if (mag == 0) continue;
                const double app = a[p * n + p].x, aqq = a[q * n + q].x;
                const double2 ph = cscale(apq, 1.0 / mag);
                const double tau = (aqq - app) / (2 * mag);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
                const double c = 1 / sqrt(1 + t * t), s = t * c;
                __syncthreads();
                if (r < n) {
                    const double2 arp = a[r * n + p], arq = a[r * n + q];
                    a[r * n + p] = cscale(arp, c) - cmul(cscale(dconj2(ph), s), arq);
                    a[r * n + q] = cmul(cscale(ph, s), arp) + cscale(arq, c);
                    const double2 vrp = v[r * n + p], vrq = v[r * n + q];
                    v[r * n + p] = cscale(vrp, c) - cmul(cscale(dconj2(ph), s), vrq);
                    v[r * n + q] = cmul(cscale(ph, s), vrp) + cscale(vrq, c);
                }
                __syncthreads();
                if (r < n) {
                    const double2 apr = a[p * n + r], aqr = a[q * n + r];
                    a[p * n + r] = cscale(apr, c) - cmul(cscale(ph, s), aqr);
                    a[q * n + r] = cmul(cscale(dconj2(ph), s), apr) + cscale(aqr, c);
                }
                __syncthreads();
                if (r == 0) {
                    a[p * n + q] = make_double2(0, 0);
                    a[q * n + p] = make_double2(0, 0);
                }
                __syncthreads();
            }
    }
    if (r < n) lam[r] = a[r * n + r].x;
    __syncthreads();
    for (int x = 0; x < n - 1; x++) {
        int best = x;
        for (int y = x + 1; y < n; y++)
            if (lam[y] < lam[best]) best = y;
        __syncthreads();
        if (best != x) {
            if (r == 0) {
                const double t = lam[x];
                lam[x] = lam[best];
                lam[best] = t;
            }
            if (r < n) {
                const double2 tv = v[r * n + x];
                v[r * n + x] = v[r * n + best];
                v[r * n + best] = tv;
            }
        }
        __syncthreads();
    }
}

// Gauss-Jordan inverse with partial pivoting on one wave (M <= 64): a (LDS,
// row-major, overwritten) -> inv (LDS).  Lane r owns column r; every lane of
// the block must call it.  Elimination order and arithmetic per element are
// those of a serial Gauss-Jordan sweep (lapack_interface.c:10-59 stand-in).
__device__ inline void gj_inverse_wave(int M, double2* a, double2* inv, double2* fac) {
    const int r = threadIdx.x;
    if (r < M)
        for (int c = 0; c < M; c++) inv[c * M + r] = make_double2(c == r ? 1.0 : 0.0, 0);
    __syncthreads();
    for (int col = 0; col < M; col++) {
        int piv = col;
        double best = 0;
        for (int i = col; i < M; i++) {
            const double m2 = a[i * M + col].x * a[i * M + col].x + a[i * M + col].y * a[i * M + col].y;
            if (m2 > best) {
                best = m2;
                piv = i;
            }
        }
        if (best == 0) break;
        __syncthreads();
        if (piv != col && r < M) {
            double2 t = a[col * M + r];
            a[col * M + r] = a[piv * M + r];
            a[piv * M + r] = t;
            t = inv[col * M + r];
            inv[col * M + r] = inv[piv * M + r];
            inv[piv * M + r] = t;
        }
        __syncthreads();
        const double2 d = cdiv(make_double2(1, 0), a[col * M + col]);
        __syncthreads();
        if (r < M) {
            a[col * M + r] = cmul(a[col * M + r], d);
            inv[col * M + r] = cmul(inv[col * M + r], d);
            fac[r] = a[r * M + col];
        }
        __syncthreads();
        if (r < M)
            for (int i = 0; i < M; i++) {
                if (i == col) continue;
                const double2 f = fac[i];
                if (f.x == 0 && f.y == 0) continue;
                a[i * M + r] = a[i * M + r] - cmul(f, a[col * M + r]);
                inv[i * M + r] = inv[i * M + r] - cmul(f, inv[col * M + r]);
            }
        __syncthreads();
    }
}

// True when the Hermitian matrix h (n x n, LDS, overwritten by its Cholesky
// factor) is positive definite: every pivot of the right-looking Cholesky
// elimination is positive.  One wave; lane r owns row r.
__device__ inline bool herm_posdef_wave(int n, double2* h) {
    const int r = threadIdx.x;
    for (int k = 0; k < n; k++) {
        const double piv = h[k * n + k].x;
        if (!(piv > 0)) return false;
        const double il = 1.0 / sqrt(piv);
        __syncthreads();
        if (r > k && r < n) h[r * n + k] = cscale(h[r * n + k], il);
        __syncthreads();
        if (r > k && r < n)
            for (int c = k + 1; c <= r; c++) h[r * n + c] = h[r * n + c] - cmul(h[r * n + k], dconj2(h[c * n + k]));
        __syncthreads();
    }
    return true;
}

// rho1 eigenvalues (synchronize.c:49), regularisation when the smallest
// natural occupation is below reg_factor times the largest
// (lapack_interface.c:126-163, synchronize.c:50-54), then the inverse from
// the lower triangle (lapack_interface.c:10-59).  One wave, M <= 30; lane
// r owns row r in the eigen-solve and column r in the regularisation and
// the Gauss-Jordan elimination.
//
// When the eigenvalues are not asked for, the inverse comes first: if rho1
// is positive definite and 1 / (|rho1|_F |rho1^-1|_F), a lower bound of
// lambda_min / lambda_max, clears reg_factor, no regularisation can apply
// and that inverse is the result (the same elimination on the same matrix
// as the full path).  Otherwise the full path runs.
//
// Every thread of the block calls it (threads r >= M only take part in the
// barriers); work is LDS of 4 M^2 complex values.
__device__ inline void density_sync_body(int M, double2* __restrict__ rho1, double2* __restrict__ rinv,
                                         double* __restrict__ eig_out, double reg_factor, double2* work) {
    __shared__ double lam[64], rowsum[128];
    __shared__ double2 fac[64];
    const int r = threadIdx.x;
    double2* a = work;              // M*M
    double2* v = work + M * M;      // M*M
    double2* inv = work + 2 * M * M;
    double2* rl = work + 3 * M * M;  // rho1 as regularised
    if (!eig_out) {
        // the Gauss-Jordan input of the full path without regularisation
        if (r < M)
            for (int c = 0; c < M; c++) {
                const double2 x = c == r ? make_double2(rho1[r * M + r].x, 0)
                                         : (r < c ? rho1[c * M + r] : dconj2(rho1[r * M + c]));
                a[c * M + r] = x;
                v[c * M + r] = x;
            }
        __syncthreads();
        double n2a = 0;
        if (r < M)
            for (int c = 0; c < M; c++) n2a += a[c * M + r].x * a[c * M + r].x + a[c * M + r].y * a[c * M + r].y;
        const bool pd = herm_posdef_wave(M, v);
        gj_inverse_wave(M, a, inv, fac);
        double n2i = 0;
        if (r < M)
            for (int c = 0; c < M; c++) n2i += inv[c * M + r].x * inv[c * M + r].x + inv[c * M + r].y * inv[c * M + r].y;
        if (r < M) {
            rowsum[2 * r] = n2a;
            rowsum[2 * r + 1] = n2i;
        }
        __syncthreads();
        double ta = 0, ti = 0;
        for (int c = 0; c < M; c++) {
            ta += rowsum[2 * c];
            ti += rowsum[2 * c + 1];
        }
        // the bound must clear reg_factor by more than the rounding of the
        // eigenvalue ratio the full path would compare
        if (pd && 1.0 / sqrt(ta * ti) > reg_factor + 1e-13) {
            if (r < M)
                for (int c = 0; c < M; c++) rinv[c * M + r] = inv[c * M + r];
            return;
        }
        __syncthreads();
    }
    // zheev reads the upper triangle with a real diagonal
    if (r < M)
        for (int c = 0; c < M; c++) {
            a[r * M + c] = r == c ? make_double2(rho1[r * M + r].x, 0)
                                  : (c > r ? rho1[r * M + c] : dconj2(rho1[c * M + r]));
            rl[r * M + c] = rho1[r * M + c];
        }
    __syncthreads();
    herm_eig_wave(M, a, lam, v, rowsum);
    if (eig_out && r < M) eig_out[r] = lam[r];
    if (lam[0] / lam[M - 1] < reg_factor) {
        const double x = lam[M - 1] * reg_factor;
        if (r < M)
            for (int i = 0; i < M; i++) {
                double2 z = make_double2(0, 0);
                for (int k = 0; k < M; k++)
                    z = z + cmul(cscale(v[i * M + k], x * exp(-lam[k] / x)), dconj2(v[r * M + k]));
                rl[i * M + r] = rl[i * M + r] + z;
                rho1[i * M + r] = rl[i * M + r];
            }
    }
    __syncthreads();
    // Gauss-Jordan inverse with partial pivoting of the Hermitian matrix
    // given by the lower triangle and real diagonal
    if (r < M)
        for (int c = 0; c < M; c++)
            a[c * M + r] = c == r ? make_double2(rl[r * M + r].x, 0)
                                  : (r < c ? rl[c * M + r] : dconj2(rl[r * M + c]));
    __syncthreads();
    gj_inverse_wave(M, a, inv, fac);
    if (r < M)
        for (int c = 0; c < M; c++) rinv[c * M + r] = inv[c * M + r];
}

}  // namespace mctdhb
