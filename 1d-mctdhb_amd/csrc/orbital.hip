// orbital.hip - the orbital split-step on the device.
//
// Replaces reference src/integrator/orbital_integration.c:351-434
// (propagate_splitstep_orb) and what it calls:
//   split_linear_orbitals.c:22-80, :288-402  (CN and FFT linear half steps)
//   tridiagonal_solver.c:6-212               (Thomas / Sherman-Morrison)
//   orbital_matrices.c:21-287                (h_kl, g_ksql by Simpson quadrature)
//   calculus.c:8-39, :116-191, :287-305      (Simpson, Gram-Schmidt, FD derivative)
//   split_nonlinear_orbitals.c:62-151        (nonlinear right-hand side)
//   lapack_interface.c:10-163, synchronize.c:41-56 (M x M eigen/regularise/invert)
// The odelib Runge-Kutta the reference links is absent; the nonlinear step
// uses the classical RK4 tableau.
#include <cmath>
#include <vector>

#include "mxm_wave.hpp"
#include "orbital.hpp"
#include "pack.hpp"

namespace mctdhb {

// ------------------------------------------------------------- helpers

// Weight of point p in the reference's Simpson rule on n points
// (calculus.c:8-39: composite 1/3 rule, 3/8 rule on the last 4 points when
// n is even).
__device__ __forceinline__ double simpson_w(int p, int n, double h) {
    if (n % 2 == 1) {
        if (p == 0 || p == n - 1) return h / 3;
        return (p % 2) ? 4 * h / 3 : 2 * h / 3;
    }
    const int e = n - 4;  // end of the composite part
    double w = 0;
    if (e > 0 && p <= e) w = (p == 0 || p == e) ? h / 3 : ((p % 2) ? 4 * h / 3 : 2 * h / 3);
    if (p == e) w += 3 * h / 8;
    else if (p == n - 3 || p == n - 2) w += 9 * h / 8;
    else if (p == n - 1) w += 3 * h / 8;
    return w;
}

// 4th-order FD derivative at point p with the reference's periodic wrap
// (calculus.c:287-305; the last point repeats the first).
__device__ __forceinline__ double2 dxfd_at(const double2* f, int n, int p, double r) {
    int a, b, c, d;  // (f[a] - f[b] + 8 (f[c] - f[d])) r
    if (p == 0 || p == n - 1) { a = n - 3; b = 2; c = 1; d = n - 2; }
    else if (p == 1) { a = n - 2; b = 3; c = 2; d = 0; }
    else if (p == n - 3) { a = n - 5; b = 0; c = n - 2; d = n - 4; }
    else if (p == n - 2) { a = n - 4; b = 1; c = 0; d = n - 3; }
    else { a = p - 2; b = p + 2; c = p + 1; d = p - 1; }
    const double2 fa = f[a], fb = f[b], fc = f[c], fd = f[d];
    return make_double2((fa.x - fb.x + 8 * (fc.x - fd.x)) * r, (fa.y - fb.y + 8 * (fc.y - fd.y)) * r);
}

template <int NV> __device__ __forceinline__ void block_sum_bcast(double (&v)[NV], double* lds) {
    block_sum<NV>(v, lds);
    if (threadIdx.x == 0)
        for (int t = 0; t < NV; t++) lds[t] = v[t];
    __syncthreads();
    for (int t = 0; t < NV; t++) v[t] = lds[t];
    __syncthreads();
}

// ------------------------------------------------------------- FFT step

__device__ __forceinline__ int bitrev(int i, int bits) { return __brev((unsigned) i) >> (32 - bits); }

__device__ void fft_inplace(double2* X, int n, int log2n, const double2* __restrict__ tw, bool inverse) {
    for (int s = 1; s <= log2n; s++) {
        const int half = 1 << (s - 1), step = n >> s;
        for (int j = threadIdx.x; j < n / 2; j += blockDim.x) {
            const int grp = j >> (s - 1), pos = j & (half - 1);
            const int i1 = (grp << s) + pos, i2 = i1 + half;
            double2 w = tw[pos * step];
            if (inverse) w.y = -w.y;
            const double2 b = cmul(X[i2], w), a = X[i1];
            X[i1] = a + b;
            X[i2] = a - b;
        }
        __syncthreads();
    }
}

// One block per orbital: forward DFT, multiply by exp(-i dt/2 T(k)), inverse
// DFT (both scaled 1/sqrt(n) as the reference's DFTI descriptor), last point
// = first (split_linear_orbitals.c:375-402).
__device__ __forceinline__ void fft_linear_block(double2* X, double2* __restrict__ orb, int grid, int n, int log2n,
                                                 const double2* __restrict__ tw, const double2* __restrict__ fexp,
                                                 int k) {
    double2* o = orb + (size_t) k * grid;
    for (int i = threadIdx.x; i < n; i += blockDim.x) X[bitrev(i, log2n)] = o[i];
    __syncthreads();
    fft_inplace(X, n, log2n, tw, false);
    for (int i = threadIdx.x; i < n; i += blockDim.x) X[i] = cmul(fexp[i], X[i]);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int r = bitrev(i, log2n);
        if (i < r) {
            const double2 t = X[i];
            X[i] = X[r];
            X[r] = t;
        }
    }
    __syncthreads();
    fft_inplace(X, n, log2n, tw, true);
    const double s = 1.0 / (double) n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) o[i] = cscale(X[i], s);
    __syncthreads();
    if (threadIdx.x == 0) o[n] = cscale(X[0], s);
}

__global__ __launch_bounds__(256) void k_fft_linear(double2* __restrict__ orb, int grid, int n, int log2n,
                                                    const double2* __restrict__ tw,
                                                    const double2* __restrict__ fexp) {
    extern __shared__ double2 X[];
    fft_linear_block(X, orb, grid, n, log2n, tw, fexp, blockIdx.x);
}

// Packed form: instance blockIdx.y, orbital blockIdx.x.
__global__ __launch_bounds__(256) void k_fft_linear_pk(const PkFft* __restrict__ tab) {
    extern __shared__ double2 X[];
    const PkFft a = tab[blockIdx.y];
    fft_linear_block(X, a.orb, a.grid, a.n, a.log2n, a.tw, a.fexp, blockIdx.x);
}

int pk_linear_fft(const PkFft* tab, int count, int M, size_t smem, hipStream_t st) {
    k_fft_linear_pk<<<dim3(M, count), 256, smem, st>>>(tab);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

// Lengths that are not powers of two: direct transforms, twiddles
// exp(-2 pi i t / n) for t < n (tw), one block per orbital, two passes.
__global__ __launch_bounds__(256) void k_dft_forward_mul(const double2* __restrict__ orb, int grid, int n,
                                                         const double2* __restrict__ tw,
                                                         const double2* __restrict__ fexp,
                                                         double2* __restrict__ tmp) {
    extern __shared__ double2 X[];
    const double2* o = orb + (size_t) blockIdx.x * grid;
    for (int i = threadIdx.x; i < n; i += blockDim.x) X[i] = o[i];
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        double2 acc = make_double2(0, 0);
        int t = 0;
        for (int j = 0; j < n; j++) {
            acc = cfma(X[j], tw[t], acc);
            t += k;
            if (t >= n) t -= n;
        }
        tmp[(size_t) blockIdx.x * n + k] = cmul(fexp[k], acc);
    }
}

__global__ __launch_bounds__(256) void k_dft_inverse(double2* __restrict__ orb, int grid, int n,
                                                     const double2* __restrict__ tw,
                                                     const double2* __restrict__ tmp) {
    extern __shared__ double2 X[];
    for (int i = threadIdx.x; i < n; i += blockDim.x) X[i] = tmp[(size_t) blockIdx.x * n + i];
    __syncthreads();
    double2* o = orb + (size_t) blockIdx.x * grid;
    const double s = 1.0 / (double) n;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        double2 acc = make_double2(0, 0);
        int t = 0;
        for (int j = 0; j < n; j++) {
            acc = cfma(X[j], dconj2(tw[t]), acc);
            t += k;
            if (t >= n) t -= n;
        }
        o[k] = cscale(acc, s);
        if (k == 0) o[n] = cscale(acc, s);
    }
}

int linear_fft(OrbWork* w, double2* orb, hipStream_t st) {
    const int M = w->eq.norb, grid = w->eq.grid_size, n = w->nfft;
    if (w->log2n >= 0) {
        k_fft_linear<<<M, 256, sizeof(double2) * n, st>>>(orb, grid, n, w->log2n, w->d_tw, w->d_fexp);
    } else {
        k_dft_forward_mul<<<M, 256, sizeof(double2) * n, st>>>(orb, grid, n, w->d_tw, w->d_fexp, w->d_stage);
        k_dft_inverse<<<M, 256, sizeof(double2) * n, st>>>(orb, grid, n, w->d_tw, w->d_stage);
    }
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

// ------------------------------------------------------ Crank-Nicolson

// Thomas algorithm of tridiagonal_solver.c:6-92, including the reduced
// system used when mid[0] == 0.  u, l, z: scratch of n entries.
__device__ void thomas(int n, const double2* up, const double2* lo, const double2* mid, const double2* rhs,
                       double2* ans, double2* u, double2* l, double2* z) {
    if (mid[0].x == 0 && mid[0].y == 0) {
        const double2 r0u = cdiv(rhs[0], up[0]);
        const double2 r1 = rhs[1] - cmul(mid[1], r0u);
        const double2 r2 = rhs[2] - cmul(lo[1], r0u);
        u[1] = lo[0];
        z[1] = r1;
        l[1] = make_double2(0, 0);
        u[2] = mid[2] - cmul(l[1], up[1]);
        z[2] = r2 - cmul(l[1], z[1]);
        for (int i = 2; i < n - 1; i++) {
            l[i] = cdiv(lo[i], u[i]);
            u[i + 1] = mid[i + 1] - cmul(l[i], up[i]);
            z[i + 1] = rhs[i + 1] - cmul(l[i], z[i]);
        }
        ans[n - 1] = cdiv(z[n - 1], u[n - 1]);
        for (int i = 2; i <= n - 1; i++) {
            const int k = n - i;
            ans[k] = cdiv(z[k] - cmul(up[k], ans[k + 1]), u[k]);
        }
        ans[0] = ans[1];
        ans[1] = r0u;
        return;
    }
    u[0] = mid[0];
    z[0] = rhs[0];
    for (int i = 0; i < n - 1; i++) {
        l[i] = cdiv(lo[i], u[i]);
        u[i + 1] = mid[i + 1] - cmul(l[i], up[i]);
        z[i + 1] = rhs[i + 1] - cmul(l[i], z[i]);
    }
    ans[n - 1] = cdiv(z[n - 1], u[n - 1]);
    for (int i = 2; i <= n; i++) {
        const int k = n - i;
        ans[k] = cdiv(z[k] - cmul(up[k], ans[k + 1]), u[k]);
    }
}

// One thread per orbital: CN right-hand side (split_linear_orbitals.c:288-333)
// and the (cyclic) tridiagonal solve (:335-373).  scratch: 8 * grid per orbital.
__global__ void k_cn_sequential(double2* __restrict__ orb, int M, int grid, int bounds,
                                const double2* __restrict__ cn, double2 rhs_up, double2 rhs_lo, double2 rhs_mid,
                                double2* __restrict__ scratch) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    double2* f = orb + (size_t) k * grid;
    double2* S = scratch + (size_t) k * 8 * grid;
    double2 *rhs = S, *x = S + grid, *w = S + 2 * grid, *U = S + 3 * grid;
    double2 *u = S + 4 * grid, *l = S + 5 * grid, *z = S + 6 * grid, *mid = S + 7 * grid;
    const double2 *up = cn, *lo = cn + grid, *mid0 = cn + 2 * grid;
    const int n = grid;
    // right-hand side
    if (bounds == kZeroBounds) rhs[0] = cmul(rhs_mid, f[0]) + cmul(rhs_up, f[1]);
    else rhs[0] = cmul(rhs_mid, f[0]) + cmul(rhs_up, f[1]) + cmul(rhs_lo, f[n - 2]);
    for (int i = 1; i < n - 1; i++) rhs[i] = cmul(rhs_mid, f[i]) + cmul(rhs_up, f[i + 1]) + cmul(rhs_lo, f[i - 1]);
    if (bounds == kZeroBounds) rhs[n - 1] = cmul(rhs_mid, f[n - 1]) + cmul(rhs_lo, f[n - 2]);
    else {
        rhs[n - 2] = cmul(rhs_mid, f[n - 2]) + cmul(rhs_up, f[0]) + cmul(rhs_lo, f[n - 3]);
        rhs[n - 1] = rhs[0];
    }
    if (bounds == kZeroBounds) {
        thomas(n, up, lo, mid0, rhs, f, u, l, z);
        return;
    }
    // cyclic system of size m = grid - 1 by Sherman-Morrison
    // (tridiagonal_solver.c:161-212): upper[m-1] couples row 0 to x_{m-1},
    // lower[m-1] couples row m-1 to x_0.
    const int m = n - 1;
    for (int i = 0; i < m; i++) mid[i] = mid0[i];
    double2 factor;
    if (mid[0].x == 0 && mid[0].y == 0) {
        factor = up[0];
        mid[0] = make_double2(-factor.x, -factor.y);
    } else {
        factor = mid[0];
        mid[0] = make_double2(0, 0);
    }
    for (int i = 0; i < m; i++) U[i] = make_double2(0, 0);
    U[0] = factor;
    U[m - 1] = lo[m - 1];
    const double2 V0 = make_double2(1, 0), Vm = cdiv(up[m - 1], factor);
    mid[m - 1] = mid[m - 1] - cdiv(cmul(up[m - 1], lo[m - 1]), factor);
    thomas(m, up, lo, mid, rhs, x, u, l, z);
    thomas(m, up, lo, mid, U, w, u, l, z);
    const double2 vx = cmul(V0, x[0]) + cmul(Vm, x[m - 1]);
    const double2 vw = cmul(V0, w[0]) + cmul(Vm, w[m - 1]);
    const double2 fac = cdiv(vx, make_double2(1.0 + vw.x, vw.y));
    for (int i = 0; i < m; i++) f[i] = x[i] - cmul(fac, w[i]);
    f[m] = f[0];
}

// One block per orbital: the CN right-hand side, then the tridiagonal solve
// by parallel cyclic reduction (PCR) in place of the two sequential sweeps:
// log2(m) steps, each eliminating the couplings to the rows s = 1, 2, 4, ...
// away, every row at once.  Periodic bounds (a cyclic system of size
// m = grid - 1) go through Sherman-Morrison like the reference
// (tridiagonal_solver.c:161-212), with the textbook split gamma = -b_0
// (Numerical Recipes' cyclic solver) instead of the reference's
// gamma = b_0, which leaves a zero first pivot that only a sequential sweep
// can take; both give the solution of the same system.  The matrix is
// strictly diagonally dominant (|b_i| > |a_i| + |c_i| for the CN operator),
// where PCR is stable.  Coefficients (a, b, c, two right-hand sides) live in
// the workspace scratch, double-buffered: 10 m entries per orbital.
constexpr int kCnPcrBlock = 1024;
constexpr int kCnPcrMaxRows = 2 * kCnPcrBlock;  // rows per thread <= 2
__global__ __launch_bounds__(kCnPcrBlock) void k_cn_pcr(double2* __restrict__ orb, int grid, int bounds,
                                                         const double2* __restrict__ cn, double2 rhs_up,
                                                         double2 rhs_lo, double2 rhs_mid, double2* __restrict__ scratch) {
    const int k = blockIdx.x;
    double2* f = orb + (size_t) k * grid;
    const double2 *up = cn, *lo = cn + grid, *mid0 = cn + 2 * grid;
    const int n = grid;
    const bool cyc = bounds != kZeroBounds;
    const int m = cyc ? n - 1 : n;
    double2* S = scratch + (size_t) k * 10 * grid;
    double2* X[5] = {S, S + m, S + 2 * m, S + 3 * m, S + 4 * m};              // a, b, c, d, e
    double2* Y[5] = {S + 5 * m, S + 6 * m, S + 7 * m, S + 8 * m, S + 9 * m};  // next step
    const double2 zero = make_double2(0, 0);
    // Sherman-Morrison split of the corners: beta = up[m-1] couples row 0 to
    // x_{m-1}, alpha = lo[m-1] row m-1 to x_0
    const double2 gam = make_double2(-mid0[0].x, -mid0[0].y);
    const double2 alpha = cyc ? lo[m - 1] : zero, beta = cyc ? up[m - 1] : zero;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        double2 r;  // right-hand side (split_linear_orbitals.c:288-333)
        if (i == 0) {
            r = cmul(rhs_mid, f[0]) + cmul(rhs_up, f[1]);
            if (cyc) r = r + cmul(rhs_lo, f[n - 2]);
        } else if (cyc && i == n - 2) {
            r = cmul(rhs_mid, f[n - 2]) + cmul(rhs_up, f[0]) + cmul(rhs_lo, f[n - 3]);
        } else if (!cyc && i == n - 1) {
            r = cmul(rhs_mid, f[n - 1]) + cmul(rhs_lo, f[n - 2]);
        } else {
            r = cmul(rhs_mid, f[i]) + cmul(rhs_up, f[i + 1]) + cmul(rhs_lo, f[i - 1]);
        }
        double2 b = mid0[i];
        if (cyc && i == 0) b = b - gam;
        if (cyc && i == m - 1) b = b - cdiv(cmul(alpha, beta), gam);
        X[0][i] = i >= 1 ? lo[i - 1] : zero;
        X[1][i] = b;
        X[2][i] = i <= m - 2 ? up[i] : zero;
        X[3][i] = r;
        X[4][i] = (cyc && i == 0) ? gam : ((cyc && i == m - 1) ? alpha : zero);
    }
    __syncthreads();
    const int nrhs = cyc ? 2 : 1;
    for (int s = 1; s < m; s *= 2) {
        for (int i = threadIdx.x; i < m; i += blockDim.x) {
            const int im = i - s, ip = i + s;
            const double2 ai = X[0][i], ci = X[2][i];
            double2 bn = X[1][i], an = zero, cn2 = zero, dn = X[3][i], en = nrhs > 1 ? X[4][i] : zero;
            if (im >= 0) {
                const double2 al = cdiv(make_double2(-ai.x, -ai.y), X[1][im]);
                an = cmul(al, X[0][im]);
                bn = bn + cmul(al, X[2][im]);
                dn = dn + cmul(al, X[3][im]);
                if (nrhs > 1) en = en + cmul(al, X[4][im]);
            }
            if (ip < m) {
                const double2 ga = cdiv(make_double2(-ci.x, -ci.y), X[1][ip]);
                cn2 = cmul(ga, X[2][ip]);
                bn = bn + cmul(ga, X[0][ip]);
                dn = dn + cmul(ga, X[3][ip]);
                if (nrhs > 1) en = en + cmul(ga, X[4][ip]);
            }
            Y[0][i] = an;
            Y[1][i] = bn;
            Y[2][i] = cn2;
            Y[3][i] = dn;
            Y[4][i] = en;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 5; t++) {
            double2* sw = X[t];
            X[t] = Y[t];
            Y[t] = sw;
        }
    }
    // every row decoupled: y = d / b, z = e / b
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const double2 b = X[1][i];
        Y[3][i] = cdiv(X[3][i], b);
        if (cyc) Y[4][i] = cdiv(X[4][i], b);
    }
    __syncthreads();
    double2 fac = zero;
    if (cyc) {
        // x = y - (v.y / (1 + v.z)) z with v = (1, 0, ..., beta / gamma)
        const double2 vm = cdiv(beta, gam);
        const double2 vy = Y[3][0] + cmul(vm, Y[3][m - 1]);
        const double2 vz = Y[4][0] + cmul(vm, Y[4][m - 1]);
        fac = cdiv(vy, make_double2(1.0 + vz.x, vz.y));
    }
    for (int i = threadIdx.x; i < m; i += blockDim.x) f[i] = cyc ? Y[3][i] - cmul(fac, Y[4][i]) : Y[3][i];
    if (cyc && threadIdx.x == 0) f[m] = Y[3][0] - cmul(fac, Y[4][0]);
}

int linear_cn(OrbWork* w, double2* orb, hipStream_t st) {
    const OrbEq& q = w->eq;
    // dt = prop_dt / 2, rhs coefficients of cn_rhs
    const double2 dt = cscale(q.prop_dt, 0.5);
    const double a2 = q.d2coef, dx = q.dx;
    const double2 a2dt = cscale(dt, a2 / dx / dx);
    const double2 a1dt = cscale(cmul(q.d1coef, dt), 1.0 / dx / 4);
    const double2 up = cscale(a2dt, 0.5) + a1dt, lo = cscale(a2dt, 0.5) - a1dt;
    const double2 mid = make_double2(-a2dt.x, 1.0 - a2dt.y);
    if (w->cn_pcr)
        k_cn_pcr<<<q.norb, kCnPcrBlock, 0, st>>>(orb, q.grid_size, q.bounds, w->d_cn, up, lo, mid, w->d_cnwork);
    else
        k_cn_sequential<<<div_up(q.norb, 64), 64, 0, st>>>(orb, q.norb, q.grid_size, q.bounds, w->d_cn, up, lo, mid,
                                                           w->d_cnwork);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

// --------------------------------------------------- orbital matrices

// hob[i][j] for i <= j (one block each), conj fill, real diagonal
// (orbital_matrices.c:21-70).
__device__ __forceinline__ void hob_block(int i, int j, const double2* __restrict__ orb, int M, int n, double dx,
                                          double a2, double2 a1, const double* __restrict__ V,
                                          double2* __restrict__ hob, double* red) {
    const double2 *fi = orb + (size_t) i * n, *fj = orb + (size_t) j * n;
    const double r = 1.0 / (12 * dx);
    double v[2] = {0, 0};
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        const double2 di = dxfd_at(fi, n, p, r), dj = dxfd_at(fj, n, p, r);
        double2 part = cscale(cconjmul(di, dj), -a2);
        part = part + cmul(a1, cconjmul(fi[p], dj));
        part = part + cscale(cconjmul(fi[p], fj[p]), V[p]);
        const double wp = simpson_w(p, n, dx);
        v[0] += wp * part.x;
        v[1] += wp * part.y;
    }
    block_sum<2>(v, red);
    if (threadIdx.x == 0) {
        if (i == j) hob[i * M + i] = make_double2(v[0], 0);
        else {
            hob[i * M + j] = make_double2(v[0], v[1]);
            hob[j * M + i] = make_double2(v[0], -v[1]);
        }
    }
}

// hint[a + bM + cM^2 + dM^3] = g int conj(phi_a phi_b) phi_c phi_d, one block
// per element (orbital_matrices.c:72-287 evaluates the same integrals).
__device__ __forceinline__ void hint_block(int e, const double2* __restrict__ orb, int M, int n, double dx, double g,
                                           double2* __restrict__ hint, double* red) {
    const int a = e % M, b = (e / M) % M, c = (e / (M * M)) % M, d = e / (M * M * M);
    const double2 *fa = orb + (size_t) a * n, *fb = orb + (size_t) b * n;
    const double2 *fc = orb + (size_t) c * n, *fd = orb + (size_t) d * n;
    double v[2] = {0, 0};
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        const double2 t = cmul(dconj2(cmul(fa[p], fb[p])), cmul(fc[p], fd[p]));
        const double wp = simpson_w(p, n, dx);
        v[0] += wp * t.x;
        v[1] += wp * t.y;
    }
    block_sum<2>(v, red);
    if (threadIdx.x == 0) hint[e] = make_double2(g * v[0], g * v[1]);
}

// Both in one launch: blocks [0, M^2) are the hob elements (i = b / M,
// j = b % M, only i <= j work), the next M^4 blocks the hint elements.
__device__ __forceinline__ void orbital_matrices_block(int b, const double2* __restrict__ orb, int M, int n,
                                                       double dx, double a2, double2 a1,
                                                       const double* __restrict__ V, double g,
                                                       double2* __restrict__ hob, double2* __restrict__ hint,
                                                       double* red) {
    const int mm = M * M;
    if (b < mm) {
        if (b % M >= b / M) hob_block(b / M, b % M, orb, M, n, dx, a2, a1, V, hob, red);
    } else {
        hint_block(b - mm, orb, M, n, dx, g, hint, red);
    }
}

__global__ __launch_bounds__(256) void k_orbital_matrices(const double2* __restrict__ orb, int M, int n, double dx,
                                                          double a2, double2 a1, const double* __restrict__ V,
                                                          double g, double2* __restrict__ hob,
                                                          double2* __restrict__ hint) {
    __shared__ double red[8];
    orbital_matrices_block(blockIdx.x, orb, M, n, dx, a2, a1, V, g, hob, hint, red);
}

// Packed form: instance blockIdx.y, matrix element blockIdx.x.
__global__ __launch_bounds__(256) void k_orbital_matrices_pk(const PkOm* __restrict__ tab, int M) {
    __shared__ double red[8];
    const PkOm a = tab[blockIdx.y];
    orbital_matrices_block(blockIdx.x, a.orb, M, a.n, a.dx, a.a2, a.a1, a.V, a.g, a.hob, a.hint, red);
}

int pk_orbital_matrices(const PkOm* tab, int count, int M, hipStream_t st) {
    k_orbital_matrices_pk<<<dim3(M * M + M * M * M * M, count), 256, 0, st>>>(tab, M);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

int orbital_matrices(OrbWork* w, const double2* orb, double2* hob, double2* hint, hipStream_t st) {
    const OrbEq& q = w->eq;
    const int M = q.norb;
    k_orbital_matrices<<<M * M + M * M * M * M, 256, 0, st>>>(orb, M, q.grid_size, q.dx, q.d2coef, q.d1coef,
                                                              w->d_pot, q.g, hob, hint);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

// ------------------------------------------------------ nonlinear part

// der[k][x] = -i time_fac (nonlinear(k, x) + V(x) orb[k][x]),
// nonlinear(k, x) restating split_nonlinear_orbitals.c:62-151.
__global__ __launch_bounds__(256) void k_nonlinear(const double2* __restrict__ orb, int M, int n, double g,
                                                   double2 tf, const double* __restrict__ V,
                                                   const double2* __restrict__ hob,
                                                   const double2* __restrict__ hint,
                                                   const double2* __restrict__ rho,
                                                   const double2* __restrict__ rinv, double2* __restrict__ der) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * n) return;
    const int k = idx / n, x = idx % n;
    const int mm = M * M, mmm = mm * M;
    double2 f[32];
    for (int a = 0; a < M; a++) f[a] = orb[(size_t) a * n + x];
    double2 red = make_double2(0, 0);
    for (int a = 0; a < M; a++) {
        red = red - cmul(hob[a * M + k], f[a]);
        for (int q = 0; q < M; q++) {
            double2 rm = cmul(rinv[k * M + a], rho[a + M * a + mm * q + mmm * q]);
            red = red + cscale(cmul(cmul(cmul(rm, dconj2(f[a])), f[q]), f[q]), g);
            for (int j = 0; j < M; j++) red = red - cmul(cmul(rm, f[j]), hint[j + a * M + q * mm + q * mmm]);
            for (int l = q + 1; l < M; l++) {
                rm = cscale(cmul(rinv[k * M + a], rho[a + M * a + mm * q + mmm * l]), 2.0);
                red = red + cscale(cmul(cmul(cmul(rm, dconj2(f[a])), f[l]), f[q]), g);
                for (int j = 0; j < M; j++) red = red - cmul(cmul(rm, f[j]), hint[j + a * M + l * mm + q * mmm]);
            }
        }
        for (int s = a + 1; s < M; s++) {
            for (int q = 0; q < M; q++) {
                double2 rm = cmul(rinv[k * M + a], rho[a + M * s + mm * q + mmm * q]);
                double2 sw = cmul(rinv[k * M + s], rho[a + M * s + mm * q + mmm * q]);
                double2 t = cmul(rm, dconj2(f[s])) + cmul(sw, dconj2(f[a]));
                red = red + cscale(cmul(cmul(t, f[q]), f[q]), g);
                for (int j = 0; j < M; j++) {
                    red = red - cmul(cmul(rm, f[j]), hint[j + s * M + q * mm + q * mmm]);
                    red = red - cmul(cmul(sw, f[j]), hint[j + a * M + q * mm + q * mmm]);
                }
                for (int l = q + 1; l < M; l++) {
                    rm = cscale(cmul(rinv[k * M + a], rho[a + M * s + mm * q + mmm * l]), 2.0);
                    sw = cscale(cmul(rinv[k * M + s], rho[a + M * s + mm * q + mmm * l]), 2.0);
                    t = cmul(rm, dconj2(f[s])) + cmul(sw, dconj2(f[a]));
                    red = red + cscale(cmul(cmul(t, f[l]), f[q]), g);
                    for (int j = 0; j < M; j++) {
                        red = red - cmul(cmul(rm, f[j]), hint[j + s * M + l * mm + q * mmm]);
                        red = red - cmul(cmul(sw, f[j]), hint[j + a * M + l * mm + q * mmm]);
                    }
                }
            }
        }
    }
    const double2 total = red + cscale(f[k], V[x]);
    // -i * tf * total
    const double2 mi_tf = make_double2(tf.y, -tf.x);
    der[idx] = cmul(mi_tf, total);
}

// The same term with the grid-independent factors summed once per step.
// Grouping the terms of the loops above by the orbital product they
// multiply gives
//   nonlinear(k, x) = g sum_{s, q <= l} A[k][ql][s] conj(f_s) f_q f_l
//                     - sum_j L[k][j] f_j
// with c_ql = 1 (q == l) or 2 (q < l) and, from the a == s, a < s and swap
// branches,
//   A[k][ql][s] = c_ql (sum_{a <= s} rinv[k][a] rho[a + M s + M^2 q + M^3 l]
//                       + sum_{t > s} rinv[k][t] rho[s + M t + M^2 q + M^3 l]),
//   L[k][j] = hob[j][k] + sum over the same (a, s, q, l, coefficient) of
//             coefficient * hint[...] as in the loops above.
// ql runs over q = 0..M-1, l = q..M-1.  Layout: A (M * P * M entries,
// P = M (M + 1) / 2), then L (M * M).  One block.
// Every thread of the block calls it; lpart: LDS of M^2 P entries (M <= 8),
// out: the tables (global memory or LDS).
__device__ __forceinline__ void nl_coef_block(int M, const double2* __restrict__ hob, const double2* __restrict__ hint,
                                              const double2* __restrict__ rho, const double2* __restrict__ rinv,
                                              double2* __restrict__ out, double2* __restrict__ lpart) {
    const int mm = M * M, mmm = mm * M, P = M * (M + 1) / 2, nA = M * P * M;
    const bool split = M <= 8;
    // A entries; without hob (the full-equation step) only the A table is needed
    for (int e = threadIdx.x; e < nA + (hob && !split ? mm : 0); e += blockDim.x) {
        if (e < nA) {
            const int s = e % M, ql = (e / M) % P, k = e / (M * P);
            int q = 0, rem = ql;
            while (rem >= M - q) rem -= M - q++;
            const int l = q + rem;
            double2 acc = make_double2(0, 0);
            for (int a = 0; a <= s; a++) acc = acc + cmul(rinv[k * M + a], rho[a + M * s + mm * q + mmm * l]);
            for (int t = s + 1; t < M; t++) acc = acc + cmul(rinv[k * M + t], rho[s + M * t + mm * q + mmm * l]);
            out[e] = cscale(acc, q == l ? 1.0 : 2.0);
        } else {
            const int k = (e - nA) / M, j = (e - nA) % M;
            double2 acc = hob[j * M + k];
            for (int a = 0; a < M; a++)
                for (int q = 0; q < M; q++)
                    for (int l = q; l < M; l++) {
                        const double c = q == l ? 1.0 : 2.0;
                        const double2 rka = cscale(rinv[k * M + a], c);
                        acc = acc + cmul(cmul(rka, rho[a + M * a + mm * q + mmm * l]), hint[j + a * M + l * mm + q * mmm]);
                        for (int t = a + 1; t < M; t++) {
                            const double2 r = rho[a + M * t + mm * q + mmm * l];
                            const double2 rm = cmul(rka, r), sw = cmul(cscale(rinv[k * M + t], c), r);
                            acc = acc + cmul(rm, hint[j + t * M + l * mm + q * mmm]);
                            acc = acc + cmul(sw, hint[j + a * M + l * mm + q * mmm]);
                        }
                    }
            out[e] = acc;
        }
    }
    if (!hob || !split) return;
    // L with M <= 8: thread (k, j, ql) sums the terms of one pair q <= l over
    // a (and t > a), then thread (k, j) adds the pairs in order
    for (int e = threadIdx.x; e < mm * P; e += blockDim.x) {
        const int ql = e % P, j = (e / P) % M, k = e / (P * M);
        int q = 0, rem = ql;
        while (rem >= M - q) rem -= M - q++;
        const int l = q + rem;
        const double c = q == l ? 1.0 : 2.0;
        double2 acc = make_double2(0, 0);
        for (int a = 0; a < M; a++) {
            const double2 rka = cscale(rinv[k * M + a], c);
            acc = acc + cmul(cmul(rka, rho[a + M * a + mm * q + mmm * l]), hint[j + a * M + l * mm + q * mmm]);
            for (int t = a + 1; t < M; t++) {
                const double2 r = rho[a + M * t + mm * q + mmm * l];
                acc = acc + cmul(cmul(rka, r), hint[j + t * M + l * mm + q * mmm]);
                acc = acc + cmul(cmul(cscale(rinv[k * M + t], c), r), hint[j + a * M + l * mm + q * mmm]);
            }
        }
        lpart[e] = acc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < mm; e += blockDim.x) {
        const int k = e / M, j = e % M;
        double2 acc = hob[j * M + k];
        for (int ql = 0; ql < P; ql++) acc = acc + lpart[(k * M + j) * P + ql];
        out[nA + e] = acc;
    }
}

__global__ __launch_bounds__(256) void k_nl_coef(int M, const double2* __restrict__ hob,
                                                 const double2* __restrict__ hint, const double2* __restrict__ rho,
                                                 const double2* __restrict__ rinv, double2* __restrict__ out) {
    extern __shared__ double2 lpart[];  // M^2 P partial sums of L (M <= 8)
    nl_coef_block(M, hob, hint, rho, rinv, out, lpart);
}

// Packed form: one block per instance.
__global__ __launch_bounds__(256) void k_nl_coef_pk(const PkNlCoef* __restrict__ tab, int M) {
    extern __shared__ double2 lpart[];
    const PkNlCoef a = tab[blockIdx.x];
    nl_coef_block(M, a.hob, a.hint, a.rho2, a.rinv, a.out, lpart);
}

// LDS of k_nl_coef (the split L sums for M <= 8).
static size_t nl_coef_smem(int M) { return M <= 8 ? sizeof(double2) * M * M * M * (M + 1) / 2 : 0; }

int pk_nl_coef(const PkNlCoef* tab, int count, int M, hipStream_t st) {
    k_nl_coef_pk<<<count, 256, nl_coef_smem(M), st>>>(tab, M);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

// out = y0 + dt (a_0 k_0 + a_1 k_1 + ...), the stage sums of an explicit
// tableau in ascending stage order.
struct RkSum {
    const double2* k[6];
    double a[6];
    int ns;
};

// der[k][x] = -i tf (nonlinear(k, x) + V(x) f_k(x)) with f = orb + dt * stage
// (the stage sum formed in registers), from the k_nl_coef tables.
// Grid: (grid points / 64, M); each block serves one k, so the table reads
// are uniform across the block.
template <int M>
__global__ __launch_bounds__(64) void k_nl_apply(int n, const double2* __restrict__ orb, RkSum stage, double dt,
                                                 double g, double2 tf, const double* __restrict__ V,
                                                 const double2* __restrict__ coef, double2* __restrict__ der) {
    constexpr int P = M * (M + 1) / 2;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (x >= n) return;
    double2 f[M];
#pragma unroll
    for (int a = 0; a < M; a++) {
        double2 v = orb[(size_t) a * n + x];
        if (stage.ns) {
            double2 acc = cscale(stage.k[0][(size_t) a * n + x], stage.a[0]);
            for (int j = 1; j < stage.ns; j++) acc = acc + cscale(stage.k[j][(size_t) a * n + x], stage.a[j]);
            v = v + cscale(acc, dt);
        }
        f[a] = v;
    }
    const double2* A = coef + (size_t) k * P * M;
    double2 acc[M];
#pragma unroll
    for (int s = 0; s < M; s++) acc[s] = make_double2(0, 0);
    int ql = 0;
#pragma unroll
    for (int q = 0; q < M; q++)
#pragma unroll
        for (int l = q; l < M; l++, ql++) {
            const double2 pql = cmul(f[q], f[l]);
#pragma unroll
            for (int s = 0; s < M; s++) acc[s] = acc[s] + cmul(A[ql * M + s], pql);
        }
    double2 cubic = make_double2(0, 0);
#pragma unroll
    for (int s = 0; s < M; s++) cubic = cubic + cconjmul(f[s], acc[s]);
    double2 total = cscale(cubic, g) + cscale(f[k], V[x]);
    const double2* L = coef + (size_t) M * P * M + k * M;
#pragma unroll
    for (int j = 0; j < M; j++) total = total - cmul(L[j], f[j]);
    const double2 mi_tf = make_double2(tf.y, -tf.x);
    der[(size_t) k * n + x] = cmul(mi_tf, total);
}

__global__ void k_rk_sum(const double2* y0, RkSum s, double dt, double2* out, int n);

// Grid-independent factors of the nonlinear term (orbitals up to
// kNlFusedMaxOrb; larger counts use k_nonlinear directly).
constexpr int kNlFusedMaxOrb = 8;

static int nonlinear_prepare(OrbWork* w, const double2* hob, const double2* hint, const double2* rho2,
                             const double2* rinv, hipStream_t st) {
    if (w->eq.norb > kNlFusedMaxOrb) return kOk;
    k_nl_coef<<<1, 256, nl_coef_smem(w->eq.norb), st>>>(w->eq.norb, hob, hint, rho2, rinv, w->d_nl);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

// der = rhs(orb + dt * stage) after nonlinear_prepare.
static int nonlinear_stage(OrbWork* w, const double2* orb, const RkSum& stage, const double2* hob,
                           const double2* hint, const double2* rho2, const double2* rinv, double2* der,
                           hipStream_t st) {
    const OrbEq& q = w->eq;
    const int n = q.grid_size, M = q.norb;
    if (M > kNlFusedMaxOrb) {
        const int total = M * n;
        const double2* in = orb;
        if (stage.ns) {
            double2* y = w->d_stage + 6 * (size_t) total;
            k_rk_sum<<<div_up(total, 256), 256, 0, st>>>(orb, stage, q.tstep, y, total);
            in = y;
        }
        k_nonlinear<<<div_up(total, 256), 256, 0, st>>>(in, M, n, q.g, q.time_fac, w->d_pot, hob, hint, rho2,
                                                        rinv, der);
        MCTDHB_HIP_CHECK(hipGetLastError());
        return kOk;
    }
    const dim3 grid(div_up(n, 64), M);
#define NL_CASE(m)                                                                                             \
    case m:                                                                                                   \
        k_nl_apply<m><<<grid, 64, 0, st>>>(n, orb, stage, q.tstep, q.g, q.time_fac, w->d_pot, w->d_nl, der); \
        break;
    switch (M) {
        NL_CASE(1)
        NL_CASE(2)
        NL_CASE(3)
        NL_CASE(4)
        NL_CASE(5)
        NL_CASE(6)
        NL_CASE(7)
        NL_CASE(8)
    }
#undef NL_CASE
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

int nonlinear_rhs(OrbWork* w, const double2* orb, const double2* hob, const double2* hint, const double2* rho2,
                  const double2* rinv, double2* der, hipStream_t st) {
    int rc = nonlinear_prepare(w, hob, hint, rho2, rinv, st);
    if (rc) return rc;
    return nonlinear_stage(w, orb, RkSum{}, hob, hint, rho2, rinv, der, st);
}

__global__ void k_rk4_combine(double2* __restrict__ y, const double2* __restrict__ k1, const double2* __restrict__ k2,
                              const double2* __restrict__ k3, const double2* __restrict__ k4, double dt, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double2 s = k1[i] + cscale(k2[i], 2.0) + cscale(k3[i], 2.0) + k4[i];
        y[i] = y[i] + cscale(s, dt / 6);
    }
}

// out = y0 + dt (a_0 k_0 + a_1 k_1 + ...); y0 and out may be the same array:
// each element is read before it is written.
__global__ void k_rk_sum(const double2* y0, RkSum s, double dt, double2* out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double2 acc = cscale(s.k[0][i], s.a[0]);
    for (int j = 1; j < s.ns; j++) acc = acc + cscale(s.k[j][i], s.a[j]);
    out[i] = y0[i] + cscale(acc, dt);
}

// Heun's 2nd-order and Butcher's 6-stage 5th-order tableaus (J. C. Butcher,
// J. Austral. Math. Soc. 4 (1964) 179-194), row-major a[stage][j].
constexpr double kRk2A[2 * 2] = {0, 0, 1, 0};
constexpr double kRk2B[2] = {0.5, 0.5};
constexpr double kRk5A[6 * 6] = {
    0,        0,       0,        0,         0,       0,  //
    0.25,     0,       0,        0,         0,       0,  //
    0.125,    0.125,   0,        0,         0,       0,  //
    0,        -0.5,    1,        0,         0,       0,  //
    3.0 / 16, 0,       0,        9.0 / 16,  0,       0,  //
    -3.0 / 7, 2.0 / 7, 12.0 / 7, -12.0 / 7, 8.0 / 7, 0,
};
constexpr double kRk5B[6] = {7.0 / 90, 0, 32.0 / 90, 12.0 / 90, 32.0 / 90, 7.0 / 90};

// The whole nonlinear Runge-Kutta sub-step at one grid point: the
// right-hand side at x involves only the orbitals at x (the A and L tables
// of k_nl_coef, V(x)), so the stage values stay on chip and one pass
// replaces the stage kernels.  Stage input and final sum as k_rk_sum forms
// them.
struct RkTab {
    double a[6 * 6], b[6];
};

// Block: 64 grid points x M orbitals; thread (k, x) forms orbital k's
// right-hand side at x, the stage values of all orbitals at x are shared in
// LDS between stages.
// FOLD: each block first forms the tables of k_nl_coef itself (from hob,
// hint, rho2, rinv) in LDS, in place of reading them from coef; used when
// they fit beside the stage values.
template <int M, int NS> constexpr bool nl_fold() {
    constexpr int P = M * (M + 1) / 2;  // stage values, tables (A and L), L partial sums
    return sizeof(double2) * (M * 64 * (NS + 1) + M * P * M + M * M + M * M * P) <= 60 * 1024;
}

template <int M, int NS> struct NlRkLds {
    static constexpr int P = M * (M + 1) / 2;
    static constexpr bool FOLD = nl_fold<M, NS>();
    double2 y0[M][64], ks[NS][M][64];
    double2 tables[FOLD ? M * P * M + M * M : 1], lpart[FOLD ? M * M * P : 1];
};

// Block xb of grid points (points 64 xb .. 64 xb + 63).
template <int M, int NS>
__device__ __forceinline__ void nl_rk_point_block(NlRkLds<M, NS>& L, int xb, double2* __restrict__ orb, int n,
                                                  const RkTab& tab, double dt, double g, double2 tf,
                                                  const double* __restrict__ V, const double2* __restrict__ coef_g,
                                                  const double2* __restrict__ hob, const double2* __restrict__ hint,
                                                  const double2* __restrict__ rho2,
                                                  const double2* __restrict__ rinv) {
    constexpr int P = M * (M + 1) / 2;
    constexpr bool FOLD = nl_fold<M, NS>();
    const double2* coef = coef_g;
    if constexpr (FOLD) {
        nl_coef_block(M, hob, hint, rho2, rinv, L.tables, L.lpart);
        coef = L.tables;  // the barrier below orders the table writes before use
    }
    // k is uniform across the wave: the table reads below become scalar loads
    const int xl = threadIdx.x & 63, k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int x = xb * 64 + xl;
    const bool in = x < n;
    L.y0[k][xl] = in ? orb[(size_t) k * n + x] : make_double2(0, 0);
    const double vx = in ? V[x] : 0.0;
    const double2 mi_tf = make_double2(tf.y, -tf.x);
    const double2* A = coef + (size_t) k * P * M;
    const double2* Lk = coef + (size_t) M * P * M + k * M;
    __syncthreads();
#pragma unroll 1  // stages in a loop: the table loads are not hoisted across stages
    for (int s = 0; s < NS; s++) {
        double2 f[M];
#pragma unroll
        for (int a = 0; a < M; a++) {
            if (s == 0) {
                f[a] = L.y0[a][xl];
            } else {
                double2 acc = cscale(L.ks[0][a][xl], tab.a[s * NS]);
#pragma unroll
                for (int j = 1; j < s; j++) acc = acc + cscale(L.ks[j][a][xl], tab.a[s * NS + j]);
                f[a] = L.y0[a][xl] + cscale(acc, dt);
            }
        }
        double2 acc[M];
#pragma unroll
        for (int t = 0; t < M; t++) acc[t] = make_double2(0, 0);
        int ql = 0;
#pragma unroll
        for (int q = 0; q < M; q++)
#pragma unroll
            for (int l = q; l < M; l++, ql++) {
                const double2 pql = cmul(f[q], f[l]);
#pragma unroll
                for (int t = 0; t < M; t++) acc[t] = acc[t] + cmul(A[ql * M + t], pql);
            }
        double2 cubic = make_double2(0, 0), fk = f[0];
#pragma unroll
        for (int t = 0; t < M; t++) {
            cubic = cubic + cconjmul(f[t], acc[t]);
            if (t == k) fk = f[t];  // constant indices: f stays in registers
        }
        double2 total = cscale(cubic, g) + cscale(fk, vx);
#pragma unroll
        for (int j = 0; j < M; j++) total = total - cmul(Lk[j], f[j]);
        L.ks[s][k][xl] = cmul(mi_tf, total);
        __syncthreads();
    }
    if (!in) return;
    double2 acc = cscale(L.ks[0][k][xl], tab.b[0]);
#pragma unroll
    for (int j = 1; j < NS; j++) acc = acc + cscale(L.ks[j][k][xl], tab.b[j]);
    orb[(size_t) k * n + x] = L.y0[k][xl] + cscale(acc, dt);
}

template <int M, int NS>
__global__ __launch_bounds__(64 * M) void k_nl_rk_point(double2* __restrict__ orb, int n, RkTab tab, double dt,
                                                        double g, double2 tf, const double* __restrict__ V,
                                                        const double2* __restrict__ coef_g,
                                                        const double2* __restrict__ hob,
                                                        const double2* __restrict__ hint,
                                                        const double2* __restrict__ rho2,
                                                        const double2* __restrict__ rinv) {
    __shared__ NlRkLds<M, NS> L;
    nl_rk_point_block<M, NS>(L, blockIdx.x, orb, n, tab, dt, g, tf, V, coef_g, hob, hint, rho2, rinv);
}

// Packed form: instance blockIdx.y; blocks past the instance's grid return.
template <int M, int NS>
__global__ __launch_bounds__(64 * M) void k_nl_rk_point_pk(const PkNl* __restrict__ tab_i, RkTab tab) {
    __shared__ NlRkLds<M, NS> L;
    const PkNl a = tab_i[blockIdx.y];
    if ((int) blockIdx.x * 64 >= a.n) return;  // whole block, before any barrier
    nl_rk_point_block<M, NS>(L, blockIdx.x, a.orb, a.n, tab, a.dt, a.g, a.tf, a.V, a.coef, a.hob, a.hint, a.rho2,
                             a.rinv);
}

constexpr double kRk4A_[4 * 4] = {0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 1, 0};
constexpr double kRk4B_[4] = {1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6};

// Pointwise RK sub-step for 1 <= norb <= 8 (k_nl_rk_point); returns
// kInvalidArgument (no error message) when the orbital count is not covered.
static int nl_rk_pointwise(OrbWork* w, double2* orb, const double2* hob, const double2* hint, const double2* rho2,
                           const double2* rinv, hipStream_t st) {
    const OrbEq& q = w->eq;
    const int M = q.norb, n = q.grid_size;
    if (M > 8) return kInvalidArgument;
    RkTab tab{};
    int ns = 4;
    const double *A = kRk4A_, *b = kRk4B_;
    if (w->rk_order == 2) ns = 2, A = kRk2A, b = kRk2B;
    if (w->rk_order == 5) ns = 6, A = kRk5A, b = kRk5B;
    for (int i = 0; i < ns * ns; i++) tab.a[i] = A[i];
    for (int i = 0; i < ns; i++) tab.b[i] = b[i];
    // the tables in global memory only where the kernel does not fold them
    bool fold = false;
    switch (M * 8 + ns) {
#define FOLD_CASE(m)                                                  \
    case m * 8 + 2: fold = nl_fold<m, 2>(); break;                      \
    case m * 8 + 4: fold = nl_fold<m, 4>(); break;                      \
    case m * 8 + 6: fold = nl_fold<m, 6>(); break;
        FOLD_CASE(1) FOLD_CASE(2) FOLD_CASE(3) FOLD_CASE(4) FOLD_CASE(5) FOLD_CASE(6) FOLD_CASE(7) FOLD_CASE(8)
#undef FOLD_CASE
        default: break;
    }
    if (!fold) {
        const int rc = nonlinear_prepare(w, hob, hint, rho2, rinv, st);
        if (rc) return rc;
    }
    const unsigned nb = div_up(n, 64);
#define RKP(m, s)                                                                                                 \
    k_nl_rk_point<m, s><<<nb, 64 * m, 0, st>>>(orb, n, tab, q.tstep, q.g, q.time_fac, w->d_pot, w->d_nl, hob, hint, \
                                               rho2, rinv)
#define RKP_M(m)                     \
    case m:                          \
        if (ns == 2) RKP(m, 2);      \
        else if (ns == 4) RKP(m, 4); \
        else RKP(m, 6);              \
        break;
    switch (M) {
        RKP_M(1)
        RKP_M(2)
        RKP_M(3)
        RKP_M(4)
        RKP_M(5)
        RKP_M(6)
        RKP_M(7)
        RKP_M(8)
    }
#undef RKP_M
#undef RKP
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

bool pk_nl_folds(int M, int ns) {
    switch (M * 8 + ns) {
#define FOLD_CASE(m)                                  \
    case m * 8 + 2: return nl_fold<m, 2>();           \
    case m * 8 + 4: return nl_fold<m, 4>();           \
    case m * 8 + 6: return nl_fold<m, 6>();
        FOLD_CASE(1) FOLD_CASE(2) FOLD_CASE(3) FOLD_CASE(4) FOLD_CASE(5) FOLD_CASE(6) FOLD_CASE(7) FOLD_CASE(8)
#undef FOLD_CASE
        default: return false;
    }
}

// The packed sub-step: the tableau of nl_rk_pointwise for ns stages.
int pk_nl_rk(const PkNl* tab_i, int count, int M, int ns, unsigned max_blocks, hipStream_t st) {
    RkTab tab{};
    const double *A = kRk4A_, *b = kRk4B_;
    if (ns == 2) A = kRk2A, b = kRk2B;
    else if (ns == 6) A = kRk5A, b = kRk5B;
    else if (ns != 4) {
        set_error("packed RK sub-step: 2, 4 or 6 stages");
        return kInvalidArgument;
    }
    for (int i = 0; i < ns * ns; i++) tab.a[i] = A[i];
    for (int i = 0; i < ns; i++) tab.b[i] = b[i];
    const dim3 grid(max_blocks, count);
#define RKP(m, s) k_nl_rk_point_pk<m, s><<<grid, 64 * m, 0, st>>>(tab_i, tab)
#define RKP_M(m)                     \
    case m:                          \
        if (ns == 2) RKP(m, 2);      \
        else if (ns == 4) RKP(m, 4); \
        else RKP(m, 6);              \
        break;
    switch (M) {
        RKP_M(1)
        RKP_M(2)
        RKP_M(3)
        RKP_M(4)
        RKP_M(5)
        RKP_M(6)
        RKP_M(7)
        RKP_M(8)
        default: set_error("packed RK sub-step: 1 <= norb <= 8"); return kInvalidArgument;
    }
#undef RKP_M
#undef RKP
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

static int rk_tableau(OrbWork* w, int ns, const double* A, const double* b, double2* orb, const double2* hob,
                      const double2* hint, const double2* rho2, const double2* rinv, hipStream_t st) {
    const OrbEq& q = w->eq;
    const int total = q.norb * q.grid_size;
    const unsigned nb = div_up(total, 256);
    double2* k = w->d_stage;
    int rc = nonlinear_prepare(w, hob, hint, rho2, rinv, st);
    if (rc) return rc;
    for (int s = 0; s < ns; s++) {
        RkSum r{};
        for (int j = 0; j < s; j++) {
            r.k[j] = k + (size_t) j * total;
            r.a[j] = A[s * ns + j];
        }
        r.ns = s;
        if ((rc = nonlinear_stage(w, orb, r, hob, hint, rho2, rinv, k + (size_t) s * total, st))) return rc;
    }
    RkSum r{};
    for (int j = 0; j < ns; j++) {
        r.k[j] = k + (size_t) j * total;
        r.a[j] = b[j];
    }
    r.ns = ns;
    k_rk_sum<<<nb, 256, 0, st>>>(orb, r, q.tstep, orb, total);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

// --------------------------------------------------------- Gram-Schmidt

// One block.  The reference (calculus.c:170-191) renormalises F_0, then for
// i >= 1 projects F_i point by point on each F_j, j < i, and renormalises
// it.  Its quirk: the coefficient at point p is S_p = <F_j, F_i> evaluated
// after points < p were updated, and S_{p+1} = S_p (1 - w_p |F_j[p]|^2), so
// S_p = S_0 prod_{q<p} (1 - w_q |F_j[q]|^2).  The prefix products depend on
// F_j alone, so the loop runs j-major: finish F_j (renormalise), form its
// prefix products once, then project every later F_i on it (each F_i still
// sees its projections in the order j = 0, 1, ..., i - 1).  Thread t owns
// the points [t per, (t + 1) per): it alone reads and writes them, so the
// only barriers are those of the block sums (double-buffered, one each).
__device__ __forceinline__ void gram_schmidt_block(double2* __restrict__ F, int M, int n, double dx) {
    __shared__ double red[2][16];
    __shared__ double wave_prod[4];
    const int T = blockDim.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = T >> 6;
    const int per = (n + T - 1) / T;  // contiguous points per thread
    const int p0 = min((int) threadIdx.x * per, n), p1 = min(p0 + per, n);
    int buf = 0;
    // (a, b) summed over the block, the same value in every thread
    auto sum2 = [&](double a, double b, double& sa, double& sb) {
        a = wave_sum(a);
        b = wave_sum(b);
        double* r = red[buf];
        buf ^= 1;
        if (lane == 0) {
            r[wave] = a;
            r[8 + wave] = b;
        }
        __syncthreads();
        sa = 0;
        sb = 0;
        for (int w = 0; w < nw; w++) {
            sa += r[w];
            sb += r[8 + w];
        }
    };
    for (int j = 0; j < M; j++) {
        double2* Fj = F + (size_t) j * n;
        double nr = 0, tot, unused;
        for (int p = p0; p < p1; p++) nr += simpson_w(p, n, dx) * (Fj[p].x * Fj[p].x + Fj[p].y * Fj[p].y);
        sum2(nr, 0.0, tot, unused);
        const double sc = sqrt(1.0 / tot);
        for (int p = p0; p < p1; p++) Fj[p] = cscale(Fj[p], sc);
        if (j == M - 1) break;
        // exclusive product of (1 - w_q |F_j[q]|^2) over the points before
        // this thread's first point: own chunk, wave scan, then the waves
        double prod = 1.0;
        for (int p = p0; p < p1; p++) prod *= 1.0 - simpson_w(p, n, dx) * (Fj[p].x * Fj[p].x + Fj[p].y * Fj[p].y);
        double incl = prod;
        for (int off = 1; off < 64; off <<= 1) {
            const double o = __shfl_up(incl, off, 64);
            if (lane >= off) incl *= o;
        }
        if (lane == 63) wave_prod[wave] = incl;
        double run0 = __shfl_up(incl, 1, 64);
        if (lane == 0) run0 = 1.0;
        __syncthreads();
        for (int u = 0; u < wave; u++) run0 *= wave_prod[u];
        for (int i = j + 1; i < M; i++) {
            double2* Fi = F + (size_t) i * n;
            double a = 0, b = 0;
            for (int p = p0; p < p1; p++) {
                const double2 t = cconjmul(Fj[p], Fi[p]);
                const double wp = simpson_w(p, n, dx);
                a += wp * t.x;
                b += wp * t.y;
            }
            double sa, sb;
            sum2(a, b, sa, sb);
            const double2 S0 = make_double2(sa, sb);
            double run = run0;
            for (int p = p0; p < p1; p++) {
                Fi[p] = Fi[p] - cmul(cscale(S0, run), Fj[p]);
                run *= 1.0 - simpson_w(p, n, dx) * (Fj[p].x * Fj[p].x + Fj[p].y * Fj[p].y);
            }
        }
    }
}

// The orbitals are staged in LDS when they fit (every pass above then reads
// LDS, not the cache hierarchy).  With tail set, the last grid point of
// every orbital is then set to the first (the periodic copy the whole step
// applies after the orbitals).
constexpr size_t kGsStageBytes = kLdsPerWorkgroup;
__device__ __forceinline__ void gram_schmidt_whole(double2* Fs, double2* __restrict__ Fg, int M, int n, double dx,
                                                   int tail, int staged) {
    const int tot = M * n;
    double2* F = Fg;
    if (staged) {
        for (int t = threadIdx.x; t < tot; t += blockDim.x) Fs[t] = Fg[t];
        __syncthreads();
        gram_schmidt_block(Fs, M, n, dx);
        F = Fs;
    } else {
        gram_schmidt_block(Fg, M, n, dx);
    }
    __syncthreads();  // every thread's points are final
    if (tail)
        for (int k = threadIdx.x; k < M; k += blockDim.x) F[(size_t) k * n + n - 1] = F[(size_t) k * n];
    if (staged) {
        __syncthreads();
        for (int t = threadIdx.x; t < tot; t += blockDim.x) Fg[t] = Fs[t];
    }
}

__global__ __launch_bounds__(256) void k_gram_schmidt(double2* __restrict__ Fg, int M, int n, double dx, int tail,
                                                      int staged) {
    extern __shared__ double2 Fs[];
    gram_schmidt_whole(Fs, Fg, M, n, dx, tail, staged);
}

// Packed form: one block per instance.
__global__ __launch_bounds__(256) void k_gram_schmidt_pk(const PkGs* __restrict__ tab, int M) {
    extern __shared__ double2 Fs[];
    const PkGs a = tab[blockIdx.x];
    gram_schmidt_whole(Fs, a.F, M, a.n, a.dx, a.tail, a.staged);
}

int pk_gram_schmidt(const PkGs* tab, int count, int M, size_t smem, hipStream_t st) {
    k_gram_schmidt_pk<<<count, 256, smem, st>>>(tab, M);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

size_t pk_gs_smem(int M, int n) {
    const size_t bytes = sizeof(double2) * M * n;
    return bytes <= kGsStageBytes ? bytes : 0;
}

int orthonormalize(OrbWork* w, double2* orb, hipStream_t st, bool periodic_tail) {
    const size_t bytes = sizeof(double2) * w->eq.norb * w->eq.grid_size;
    const bool staged = bytes <= kGsStageBytes;
    k_gram_schmidt<<<1, 256, staged ? bytes : 0, st>>>(orb, w->eq.norb, w->eq.grid_size, w->eq.dx,
                                                       periodic_tail ? 1 : 0, staged ? 1 : 0);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

__global__ void k_periodic(double2* orb, int M, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < M) orb[(size_t) k * n + n - 1] = orb[(size_t) k * n];
}

int periodic_bounds(OrbWork* w, double2* orb, hipStream_t st) {
    k_periodic<<<1, 64, 0, st>>>(orb, w->eq.norb, w->eq.grid_size);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

// --------------------------------------- full-equation Runge-Kutta step
//
// Reference orbital_integration.c:158-348.  Per stage, on the stage input y:
//   H_k = L y_k + g sum A[k][ql][s] conj(y_s) y_q y_l    (linear + interacting)
//   P[s][k] = <y_s, H_k>, G[s][t] = <y_s, y_t>           (Simpson)
//   C = S^-1 P (robust; S = G with the reference's overlap pattern) or C = P
//   der_k = -i tf (H_k - sum_s y_s C[s][k])
// L is the reference's finite-difference operator (split_linear_orbitals.c:
// 209-255) or its spectral one (:257-286) plus V.


// L y_k at point x by finite differences (n = grid size).
__device__ __forceinline__ double2 horb_fd_at(const double2* f, int n, int x, int bounds, double2 up, double2 lo,
                                              double c0, const double* V) {
    if (bounds == kPeriodicBounds) {
        if (x == n - 1) x = 0;  // out[n-1] = out[0]
        const int xm = x == 0 ? n - 2 : x - 1;
        const int xp = x == n - 2 ? 0 : x + 1;
        return cscale(f[x], c0 + V[x]) + cmul(up, f[xp]) + cmul(lo, f[xm]);
    }
    double2 out = cscale(f[x], c0 + V[x]);
    if (x + 1 < n) out = out + cmul(up, f[x + 1]);
    if (x > 0) out = out + cmul(lo, f[x - 1]);
    return out;
}

struct FsLin {
    int fd, bounds;
    double2 up, lo;
    double c0;
};

// H[k][x] = (finite-difference L y_k, or the spectral L y_k already in H)
// + g * interacting part from the A table.  Grid (n / 64, M), one k per block.
template <int M>
__global__ __launch_bounds__(64) void k_fs_haction(int n, const double2* __restrict__ y, FsLin lin,
                                                   const double* __restrict__ V, const double2* __restrict__ coef,
                                                   double g, double2* __restrict__ H) {
    constexpr int P = M * (M + 1) / 2;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (x >= n) return;
    double2 f[M];
#pragma unroll
    for (int a = 0; a < M; a++) f[a] = y[(size_t) a * n + x];
    const double2* A = coef + (size_t) k * P * M;
    double2 acc[M];
#pragma unroll
    for (int s = 0; s < M; s++) acc[s] = make_double2(0, 0);
    int ql = 0;
#pragma unroll
    for (int q = 0; q < M; q++)
#pragma unroll
        for (int l = q; l < M; l++, ql++) {
            const double2 pql = cmul(f[q], f[l]);
#pragma unroll
            for (int s = 0; s < M; s++) acc[s] = acc[s] + cmul(A[ql * M + s], pql);
        }
    double2 cubic = make_double2(0, 0);
#pragma unroll
    for (int s = 0; s < M; s++) cubic = cubic + cconjmul(f[s], acc[s]);
    const size_t o = (size_t) k * n + x;
    const double2 l = lin.fd ? horb_fd_at(y + (size_t) k * n, n, x, lin.bounds, lin.up, lin.lo, lin.c0, V) : H[o];
    H[o] = l + cscale(cubic, g);
}

// The same for any orbital count, orbitals read from memory.
__global__ __launch_bounds__(64) void k_fs_haction_any(int M, int n, const double2* __restrict__ y, FsLin lin,
                                                       const double* __restrict__ V,
                                                       const double2* __restrict__ coef, double g,
                                                       double2* __restrict__ H) {
    const int P = M * (M + 1) / 2;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (x >= n) return;
    const double2* A = coef + (size_t) k * P * M;
    double2 cubic = make_double2(0, 0);
    for (int s = 0; s < M; s++) {
        double2 acc = make_double2(0, 0);
        int ql = 0;
        for (int q = 0; q < M; q++)
            for (int l = q; l < M; l++, ql++)
                acc = acc + cmul(A[ql * M + s], cmul(y[(size_t) q * n + x], y[(size_t) l * n + x]));
        cubic = cubic + cconjmul(y[(size_t) s * n + x], acc);
    }
    const size_t o = (size_t) k * n + x;
    const double2 l = lin.fd ? horb_fd_at(y + (size_t) k * n, n, x, lin.bounds, lin.up, lin.lo, lin.c0, V) : H[o];
    H[o] = l + cscale(cubic, g);
}

// dst_k = IDFT(sym * DFT(src_k)) + V src_k, dst[n] = dst[0]: the spectral
// linear part (split_linear_orbitals.c:257-286); one block per orbital.
__global__ __launch_bounds__(256) void k_fft_apply(const double2* __restrict__ src, double2* __restrict__ dst,
                                                   int grid, int n, int log2n, const double2* __restrict__ tw,
                                                   const double2* __restrict__ sym, const double* __restrict__ V) {
    extern __shared__ double2 X[];
    const double2* f = src + (size_t) blockIdx.x * grid;
    double2* o = dst + (size_t) blockIdx.x * grid;
    for (int i = threadIdx.x; i < n; i += blockDim.x) X[bitrev(i, log2n)] = f[i];
    __syncthreads();
    fft_inplace(X, n, log2n, tw, false);
    for (int i = threadIdx.x; i < n; i += blockDim.x) X[i] = cmul(sym[i], X[i]);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int r = bitrev(i, log2n);
        if (i < r) {
            const double2 t = X[i];
            X[i] = X[r];
            X[r] = t;
        }
    }
    __syncthreads();
    fft_inplace(X, n, log2n, tw, true);
    const double sc = 1.0 / (double) n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) o[i] = cscale(X[i], sc) + cscale(f[i], V[i]);
    __syncthreads();
    if (threadIdx.x == 0) o[n] = cscale(X[0], sc) + cscale(f[0], V[0]);
}

// Inverse direct transform for lengths that are not powers of two, written
// to dst with V src added (pairs with k_dft_forward_mul).
__global__ __launch_bounds__(256) void k_dft_inverse_add(double2* __restrict__ dst, const double2* __restrict__ src,
                                                         int grid, int n, const double2* __restrict__ tw,
                                                         const double2* __restrict__ tmp,
                                                         const double* __restrict__ V) {
    extern __shared__ double2 X[];
    for (int i = threadIdx.x; i < n; i += blockDim.x) X[i] = tmp[(size_t) blockIdx.x * n + i];
    __syncthreads();
    double2* o = dst + (size_t) blockIdx.x * grid;
    const double2* f = src + (size_t) blockIdx.x * grid;
    const double sc = 1.0 / (double) n;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        double2 acc = make_double2(0, 0);
        int t = 0;
        for (int j = 0; j < n; j++) {
            acc = cfma(X[j], dconj2(tw[t]), acc);
            t += k;
            if (t >= n) t -= n;
        }
        o[k] = cscale(acc, sc) + cscale(f[k], V[k]);
        if (k == 0) o[n] = o[0];
    }
}

// The exponential-DVR matrix of the derivative terms (split_linear_orbitals.c:
// 82-165), n = grid_size - 1, stored transposed (DT[l * n + j] = D[j][l]) so
// the product below reads it coalesced.  One thread per entry j >= i;
// the entry j < i is its (anti-)Hermitian partner, as in the reference.
__global__ __launch_bounds__(256) void k_dvr_build(int n, double length, double d2coef, double2 d1coef,
                                                   double2* __restrict__ DT) {
    const long long e = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long) n * n) return;
    const int i = (int) (e / n), j = (int) (e % n);
    if (j < i) return;
    double2 s1 = make_double2(0, 0), s2 = make_double2(0, 0);
    for (int k = 0; k < n; k++) {
        const int kmom = k - n / 2;
        const double w1 = -(2 * M_PI * kmom / length);           // d1 diagonal: i * w1
        const double w2 = -(2 * M_PI * kmom / length) * (2 * M_PI * kmom / length);
        double sn = 0, cs = 1;
        if (j != i) sincos(2 * kmom * (j - i) * M_PI / n, &sn, &cs);
        // s1 += i w1 e / n, s2 += w2 e / n with e = cos + i sin
        s1 = s1 + make_double2(-w1 * sn / n, w1 * cs / n);
        s2 = s2 + make_double2(w2 * cs / n, w2 * sn / n);
    }
    const double2 dij = cscale(s2, d2coef) + cmul(d1coef, s1);
    DT[(size_t) j * n + i] = dij;  // D[i][j]
    if (j != i) {                  // D[j][i] = d2 conj(s2) - d1 conj(s1)
        const double2 dji = cscale(make_double2(s2.x, -s2.y), d2coef) - cmul(d1coef, make_double2(s1.x, -s1.y));
        DT[(size_t) i * n + j] = dji;
    }
}

// H_k[j] = sum_l D[j][l] y_k[l] + V[j] y_k[j] (j < n), H_k[n] = H_k[0]
// (orbital_integration.c:159-169).  One thread per row j, all orbitals.
template <int M>
__global__ __launch_bounds__(256) void k_dvr_apply(int n, int grid, const double2* __restrict__ DT,
                                                   const double2* __restrict__ y, const double* __restrict__ V,
                                                   double2* __restrict__ H) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double2 acc[M];
#pragma unroll
    for (int k = 0; k < M; k++) acc[k] = make_double2(0, 0);
    for (int l = 0; l < n; l++) {
        const double2 d = DT[(size_t) l * n + j];
#pragma unroll
        for (int k = 0; k < M; k++) acc[k] = acc[k] + cmul(d, y[(size_t) k * grid + l]);
    }
#pragma unroll
    for (int k = 0; k < M; k++) {
        const double2 h = acc[k] + cscale(y[(size_t) k * grid + j], V[j]);
        H[(size_t) k * grid + j] = h;
        if (j == 0) H[(size_t) k * grid + n] = h;
    }
}

__global__ __launch_bounds__(256) void k_dvr_apply_any(int M, int n, int grid, const double2* __restrict__ DT,
                                                       const double2* __restrict__ y, const double* __restrict__ V,
                                                       double2* __restrict__ H) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    for (int k = 0; k < M; k++) {
        double2 acc = make_double2(0, 0);
        for (int l = 0; l < n; l++) acc = acc + cmul(DT[(size_t) l * n + j], y[(size_t) k * grid + l]);
        const double2 h = acc + cscale(y[(size_t) k * grid + j], V[j]);
        H[(size_t) k * grid + j] = h;
        if (j == 0) H[(size_t) k * grid + n] = h;
    }
}

int ensure_dvr(OrbWork* w) {
    if (w->d_dvrT) return kOk;
    const int n = w->nfft;
    if (hipMalloc((void**) &w->d_dvrT, sizeof(double2) * (size_t) n * n) != hipSuccess) {
        w->d_dvrT = nullptr;
        set_error("device allocation failed for the DVR matrix");
        return kOutOfMemory;
    }
    const long long total = (long long) n * n;
    k_dvr_build<<<(unsigned) ((total + 255) / 256), 256>>>(n, n * w->eq.dx, w->eq.d2coef, w->eq.d1coef, w->d_dvrT);
    MCTDHB_HIP_CHECK(hipGetLastError());
    MCTDHB_HIP_CHECK(hipDeviceSynchronize());
    return kOk;
}

// mats[0..M^2): P[a][b] = <y_a, H_b>; with the Gram part, mats[M^2..2M^2):
// G[a][b] = <y_a, y_b>.  One block per entry, Simpson weights.
__global__ __launch_bounds__(256) void k_fs_dots(int M, int n, double dx, const double2* __restrict__ y,
                                                 const double2* __restrict__ H, double2* __restrict__ mats) {
    __shared__ double red[16];
    const int e = blockIdx.x, mm = M * M;
    const int a = (e % mm) / M, b = e % M;
    const double2* fa = y + (size_t) a * n;
    const double2* fb = (e < mm ? H : y) + (size_t) b * n;
    double v[2] = {0, 0};
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        const double2 t = cconjmul(fa[p], fb[p]);
        const double w = simpson_w(p, n, dx);
        v[0] += w * t.x;
        v[1] += w * t.y;
    }
    block_sum<2>(v, red);
    if (threadIdx.x == 0) mats[e] = make_double2(v[0], v[1]);
}

// C = P (simple projector) or S^-1 P (robust).  S is the overlap matrix as
// the reference fills it (orbital_matrices.c:7-20): its j loop starts at 2,
// so S[0][1] and S[1][0] are never set (quirk Q2; zero here, uninitialised
// memory in the reference); the inverse reads the lower triangle and the
// real diagonal (lapack_interface.c:10-59).  One wave.
__global__ __launch_bounds__(64) void k_fs_coef(int M, int robust, double2* __restrict__ mats) {
    __shared__ double2 a[30 * 30], inv[30 * 30], fac[64];
    const int r = threadIdx.x, mm = M * M;
    const double2 *P = mats, *G = mats + mm;
    double2* C = mats + 2 * mm;
    if (!robust) {
        for (int e = r; e < mm; e += blockDim.x) C[e] = P[e];
        return;
    }
    auto lower = [&](int i, int j) {  // S[i][j], i > j
        if (j >= 2) return G[i * M + j];
        if (i >= 2) return dconj2(G[j * M + i]);
        return make_double2(0, 0);
    };
    if (r < M)
        for (int c = 0; c < M; c++)
            a[c * M + r] = c == r ? make_double2(G[r * M + r].x, 0) : (r < c ? lower(c, r) : dconj2(lower(r, c)));
    __syncthreads();
    gj_inverse_wave(M, a, inv, fac);
    if (r < M)  // lane r: column r of C = S^-1 P
        for (int s = 0; s < M; s++) {
            double2 acc = make_double2(0, 0);
            for (int l = 0; l < M; l++) acc = acc + cmul(inv[s * M + l], P[l * M + r]);
            C[s * M + r] = acc;
        }
}

// out_k = sum_s y_s C[s][k] (project) or -i tf (H_k - that) (der).
__global__ __launch_bounds__(256) void k_fs_der(int M, int n, const double2* __restrict__ y,
                                                const double2* __restrict__ H, const double2* __restrict__ mats,
                                                double2 tf, int project, double2* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * n) return;
    const int k = i / n, x = i % n;
    const double2* C = mats + 2 * M * M;
    double2 p = make_double2(0, 0);
    for (int s = 0; s < M; s++) p = p + cmul(y[(size_t) s * n + x], C[s * M + k]);
    if (project) {
        out[i] = p;
        return;
    }
    out[i] = cmul(make_double2(tf.y, -tf.x), H[i] - p);
}

// P and C for the orbitals y and the functions H (k_fs_dots + k_fs_coef).
static int fs_projection(OrbWork* w, int robust, const double2* y, const double2* H, hipStream_t st) {
    const int M = w->eq.norb, n = w->eq.grid_size;
    k_fs_dots<<<(robust ? 2 : 1) * M * M, 256, 0, st>>>(M, n, w->eq.dx, y, H, w->d_fsmat);
    k_fs_coef<<<1, 64, 0, st>>>(M, robust, w->d_fsmat);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

// der = right-hand side of the full equation at y; the A table in d_nl
// must be current.  der_method: 0 FFT, 1 finite differences, 2 DVR.
static int fs_rhs(OrbWork* w, int der_method, int robust, const double2* y, double2* der, hipStream_t st) {
    const OrbEq& q = w->eq;
    const int M = q.norb, n = q.grid_size, nf = w->nfft;
    const int use_fd = der_method == 1;
    double2* H = w->d_fsH;
    if (der_method == 2) {
        if (!w->d_dvrT) {
            set_error("the DVR matrix is not built (ensure_dvr)");
            return kInvalidArgument;
        }
        if (M <= 8) {
#define DVR_CASE(m) \
    case m: k_dvr_apply<m><<<div_up(nf, 256), 256, 0, st>>>(nf, n, w->d_dvrT, y, w->d_pot, H); break;
            switch (M) {
                DVR_CASE(1)
                DVR_CASE(2)
                DVR_CASE(3)
                DVR_CASE(4)
                DVR_CASE(5)
                DVR_CASE(6)
                DVR_CASE(7)
                DVR_CASE(8)
            }
#undef DVR_CASE
        } else {
            k_dvr_apply_any<<<div_up(nf, 256), 256, 0, st>>>(M, nf, n, w->d_dvrT, y, w->d_pot, H);
        }
    } else if (!use_fd) {
        if (w->log2n >= 0) {
            k_fft_apply<<<M, 256, sizeof(double2) * nf, st>>>(y, H, n, nf, w->log2n, w->d_tw, w->d_hsym, w->d_pot);
        } else {
            k_dft_forward_mul<<<M, 256, sizeof(double2) * nf, st>>>(y, n, nf, w->d_tw, w->d_hsym, w->d_cnwork);
            k_dft_inverse_add<<<M, 256, sizeof(double2) * nf, st>>>(H, y, n, nf, w->d_tw, w->d_cnwork, w->d_pot);
        }
    }
    FsLin lin{};
    lin.fd = use_fd;
    lin.bounds = q.bounds;
    const double2 a1 = q.d1coef;
    const double a2 = q.d2coef, dx = q.dx;
    lin.up = make_double2(a2 / dx / dx + a1.x / dx / 2, a1.y / dx / 2);
    lin.lo = make_double2(a2 / dx / dx - a1.x / dx / 2, -a1.y / dx / 2);
    lin.c0 = -2 * a2 / dx / dx;
    const dim3 grid(div_up(n, 64), M);
#define FS_CASE(m)                                                                                  \
    case m:                                                                                        \
        k_fs_haction<m><<<grid, 64, 0, st>>>(n, y, lin, w->d_pot, w->d_nl, q.g, H);              \
        break;
    switch (M) {
        FS_CASE(1)
        FS_CASE(2)
        FS_CASE(3)
        FS_CASE(4)
        FS_CASE(5)
        FS_CASE(6)
        FS_CASE(7)
        FS_CASE(8)
        default: k_fs_haction_any<<<grid, 64, 0, st>>>(M, n, y, lin, w->d_pot, w->d_nl, q.g, H);
    }
#undef FS_CASE
    MCTDHB_HIP_CHECK(hipGetLastError());
    int rc = fs_projection(w, robust, y, H, st);
    if (rc) return rc;
    k_fs_der<<<div_up(M * n, 256), 256, 0, st>>>(M, n, y, H, w->d_fsmat, q.time_fac, 0, der);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

int fullstep_rhs(OrbWork* w, int der_method, int robust, const double2* orb, const double2* rho2,
                 const double2* rinv, double2* der, hipStream_t st) {
    int rc = der_method == 2 ? ensure_dvr(w) : kOk;
    if (rc) return rc;
    k_nl_coef<<<1, 256, 0, st>>>(w->eq.norb, nullptr, nullptr, rho2, rinv, w->d_nl);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return fs_rhs(w, der_method, robust, orb, der, st);
}

int projector(OrbWork* w, int robust, const double2* orb, const double2* hact, double2* proj, hipStream_t st) {
    const int M = w->eq.norb, n = w->eq.grid_size;
    int rc = fs_projection(w, robust, orb, hact, st);
    if (rc) return rc;
    k_fs_der<<<div_up(M * n, 256), 256, 0, st>>>(M, n, orb, hact, w->d_fsmat, w->eq.time_fac, 1, proj);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

constexpr double kRk4A[4 * 4] = {0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 1, 0};
constexpr double kRk4B[4] = {1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6};

int fullstep(OrbWork* w, int der_method, int robust, double2* orb, const double2* rho2, const double2* rinv,
             hipStream_t st) {
    const OrbEq& q = w->eq;
    const int M = q.norb, total = M * q.grid_size;
    const unsigned nb = div_up(total, 256);
    int rc = der_method == 2 ? ensure_dvr(w) : kOk;
    if (rc) return rc;
    if (q.bounds == kPeriodicBounds && (rc = periodic_bounds(w, orb, st))) return rc;
    k_nl_coef<<<1, 256, 0, st>>>(M, nullptr, nullptr, rho2, rinv, w->d_nl);
    int ns = 4;
    const double *A = kRk4A, *b = kRk4B;
    if (w->rk_order == 2) ns = 2, A = kRk2A, b = kRk2B;
    if (w->rk_order == 5) ns = 6, A = kRk5A, b = kRk5B;
    double2* k = w->d_stage;
    double2* y = w->d_stage + 6 * (size_t) total;
    for (int s = 0; s < ns; s++) {
        const double2* in = orb;
        if (s > 0) {
            RkSum r{};
            for (int j = 0; j < s; j++) {
                r.k[j] = k + (size_t) j * total;
                r.a[j] = A[s * ns + j];
            }
            r.ns = s;
            k_rk_sum<<<nb, 256, 0, st>>>(orb, r, q.tstep, y, total);
            in = y;
        }
        if ((rc = fs_rhs(w, der_method, robust, in, k + (size_t) s * total, st))) return rc;
    }
    RkSum r{};
    for (int j = 0; j < ns; j++) {
        r.k[j] = k + (size_t) j * total;
        r.a[j] = b[j];
    }
    r.ns = ns;
    k_rk_sum<<<nb, 256, 0, st>>>(orb, r, q.tstep, orb, total);
    MCTDHB_HIP_CHECK(hipGetLastError());
    if (w->imagtime) rc = orthonormalize(w, orb, st);
    return rc;
}

// The reference first copies the first grid point onto the last
// (orbital_integration.c:370-373); with periodic bounds both linear half
// steps neither read the last point nor keep it (they write it from the
// first, split_linear_orbitals.c:364-365 and :397), so that copy changes
// nothing and is not made here.
int splitstep(OrbWork* w, int use_fft, double2* orb, double2* hob, double2* hint, const double2* rho2,
              const double2* rinv, hipStream_t st, bool periodic_tail) {
    const OrbEq& q = w->eq;
    const int total = q.norb * q.grid_size;
    int rc = use_fft ? linear_fft(w, orb, st) : linear_cn(w, orb, st);
    if (rc) return rc;
    rc = orbital_matrices(w, orb, hob, hint, st);
    if (rc) return rc;
    if (q.norb <= 8) {
        rc = nl_rk_pointwise(w, orb, hob, hint, rho2, rinv, st);
    } else if (w->rk_order == 2) {
        rc = rk_tableau(w, 2, kRk2A, kRk2B, orb, hob, hint, rho2, rinv, st);
    } else if (w->rk_order == 5) {
        rc = rk_tableau(w, 6, kRk5A, kRk5B, orb, hob, hint, rho2, rinv, st);
    } else {
        // classical RK4: stages at orb, orb + dt/2 k1, orb + dt/2 k2, orb + dt k3
        double2 *k1 = w->d_stage, *k2 = k1 + total, *k3 = k2 + total, *k4 = k3 + total;
        const double dt = q.tstep;
        const unsigned nb = div_up(total, 256);
        RkSum r{};
        r.a[0] = 0.5;
        if ((rc = nonlinear_prepare(w, hob, hint, rho2, rinv, st))) return rc;
        if ((rc = nonlinear_stage(w, orb, r, hob, hint, rho2, rinv, k1, st))) return rc;
        r.ns = 1;
        r.k[0] = k1;
        if ((rc = nonlinear_stage(w, orb, r, hob, hint, rho2, rinv, k2, st))) return rc;
        r.k[0] = k2;
        if ((rc = nonlinear_stage(w, orb, r, hob, hint, rho2, rinv, k3, st))) return rc;
        r.k[0] = k3;
        r.a[0] = 1.0;
        if ((rc = nonlinear_stage(w, orb, r, hob, hint, rho2, rinv, k4, st))) return rc;
        k_rk4_combine<<<nb, 256, 0, st>>>(orb, k1, k2, k3, k4, dt, total);
    }
    if (rc) return rc;
    rc = use_fft ? linear_fft(w, orb, st) : linear_cn(w, orb, st);
    if (rc) return rc;
    // in real time the linear step already made the periodic copy
    if (w->imagtime) rc = orthonormalize(w, orb, st, periodic_tail && q.bounds == kPeriodicBounds);
    return rc;
}

// ---------------------------------------------------- M x M algebra

__global__ __launch_bounds__(64) void k_density_sync_small(int M, double2* __restrict__ rho1,
                                                           double2* __restrict__ rinv, double* __restrict__ eig_out,
                                                           double reg_factor) {
    extern __shared__ double2 work[];  // 4 * M * M
    density_sync_body(M, rho1, rinv, eig_out, reg_factor, work);
}

int density_sync_small(int norb, double2* rho1, double2* rinv, double* eigvals, double reg_factor, hipStream_t st) {
    if (norb > 30) {
        set_error("norb exceeds 30");
        return kInvalidArgument;
    }
    // at most 4 * 30 * 30 * 16 = 57600 bytes of LDS
    k_density_sync_small<<<1, 64, sizeof(double2) * 4 * norb * norb, st>>>(norb, rho1, rinv, eigvals, reg_factor);
    MCTDHB_HIP_CHECK(hipGetLastError());
    return kOk;
}

// ------------------------------------------------------------- set-up

int orbwork_create(const OrbEq& eq, int imagtime, const double* pot_host, OrbWork** out) {
    *out = nullptr;
    if (eq.norb < 1 || eq.norb > 30 || eq.grid_size < 6) {
        set_error("invalid orbital equation (norb in 1..30, grid_size >= 6)");
        return kInvalidArgument;
    }
    OrbWork* w = new OrbWork();
    w->eq = eq;
    w->imagtime = imagtime;
    const int grid = eq.grid_size, n = grid - 1, M = eq.norb;
    w->nfft = n;
    w->log2n = (n & (n - 1)) == 0 ? __builtin_ctz((unsigned) n) : -1;
    if (w->log2n >= 0 && n > 4096) w->log2n = -1;
    if (n > 4096) {
        set_error("grid_size - 1 above 4096 is not supported by the FFT kernels");
        delete w;
        return kInvalidArgument;
    }
    // FFT exponential (split_linear_orbitals.c:56-80) and twiddles on the host
    std::vector<double2> fexp(n), tw(n);
    const double2 dt = make_double2(0.5 * eq.prop_dt.x, 0.5 * eq.prop_dt.y);
    for (int i = 0; i < n; i++) {
        const double freq = (i <= (n - 1) / 2) ? (2 * M_PI * i) / (n * eq.dx) : (2 * M_PI * (i - n)) / (n * eq.dx);
        // -i dt (i a1 k - a2 k^2) = dt (a1 k) + i dt a2 k^2
        const double2 ia1k = make_double2(-eq.d1coef.y * freq, eq.d1coef.x * freq);
        const double2 arg0 = make_double2(ia1k.x - eq.d2coef * freq * freq, ia1k.y);
        // z = -i * dt * arg0
        const double2 dta = make_double2(dt.x * arg0.x - dt.y * arg0.y, dt.x * arg0.y + dt.y * arg0.x);
        const double2 z = make_double2(dta.y, -dta.x);
        const double m = exp(z.x);
        fexp[i] = make_double2(m * cos(z.y), m * sin(z.y));
        const long long r = i % n;
        const double ang = 2 * M_PI * (double) r / (double) n;
        tw[i] = make_double2(cos(ang), -sin(ang));
    }
    // CN tridiagonal (split_linear_orbitals.c:22-54)
    std::vector<double2> cn(3 * (size_t) grid, make_double2(0, 0));
    {
        const double a2 = eq.d2coef, dx = eq.dx;
        const double2 a2dt = make_double2(dt.x * a2 / dx / dx, dt.y * a2 / dx / dx);
        const double2 a1dt = make_double2((eq.d1coef.x * dt.x - eq.d1coef.y * dt.y) / dx / 4,
                                          (eq.d1coef.x * dt.y + eq.d1coef.y * dt.x) / dx / 4);
        const double2 minus_half = make_double2(-0.5 * a2dt.x, -0.5 * a2dt.y);
        for (int i = 0; i < grid; i++) cn[2 * grid + i] = make_double2(a2dt.x, 1.0 + a2dt.y);
        for (int i = 0; i < grid - 2; i++) {
            cn[i] = make_double2(minus_half.x - a1dt.x, minus_half.y - a1dt.y);
            cn[grid + i] = make_double2(minus_half.x + a1dt.x, minus_half.y + a1dt.y);
        }
        if (eq.bounds == kZeroBounds) {
            cn[grid - 2] = make_double2(minus_half.x - a1dt.x, minus_half.y - a1dt.y);
            cn[grid + grid - 2] = make_double2(minus_half.x + a1dt.x, minus_half.y + a1dt.y);
        } else {
            cn[grid - 2] = make_double2(minus_half.x + a1dt.x, minus_half.y + a1dt.y);
            cn[grid + grid - 2] = make_double2(minus_half.x - a1dt.x, minus_half.y - a1dt.y);
        }
    }
    bool ok = hipMalloc((void**) &w->d_pot, sizeof(double) * grid) == hipSuccess;
    ok = ok && hipMalloc((void**) &w->d_fexp, sizeof(double2) * n) == hipSuccess;
    ok = ok && hipMalloc((void**) &w->d_tw, sizeof(double2) * n) == hipSuccess;
    ok = ok && hipMalloc((void**) &w->d_cn, sizeof(double2) * 3 * grid) == hipSuccess;
    ok = ok && hipMalloc((void**) &w->d_stage, sizeof(double2) * 7 * (size_t) M * grid) == hipSuccess;
    ok = ok && hipMalloc((void**) &w->d_red, sizeof(double) * 1024) == hipSuccess;
    ok = ok && hipMalloc((void**) &w->d_cnwork, sizeof(double2) * 10 * (size_t) M * grid) == hipSuccess;
    ok = ok && hipMalloc((void**) &w->d_nl, sizeof(double2) * ((size_t) M * M * M * (M + 1) / 2 + M * M)) ==
                   hipSuccess;
    ok = ok && hipMalloc((void**) &w->d_fsH, sizeof(double2) * (size_t) M * grid) == hipSuccess;
    ok = ok && hipMalloc((void**) &w->d_fsmat, sizeof(double2) * 3 * M * M) == hipSuccess;
    ok = ok && hipMalloc((void**) &w->d_hsym, sizeof(double2) * n) == hipSuccess;
    if (!ok) {
        orbwork_destroy(w);
        set_error("device allocation failed for the orbital workspace");
        return kOutOfMemory;
    }
    MCTDHB_HIP_CHECK(hipMemcpy(w->d_fexp, fexp.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    MCTDHB_HIP_CHECK(hipMemcpy(w->d_tw, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    MCTDHB_HIP_CHECK(hipMemcpy(w->d_cn, cn.data(), sizeof(double2) * 3 * grid, hipMemcpyHostToDevice));
    // the parallel CN solve needs a nonzero first diagonal (its cyclic split
    // divides by it; the reference's reduced system for mid[0] == 0 stays on
    // the sequential sweep) and at most two rows per thread
    {
        const double2 m0 = cn[2 * (size_t) grid];
        w->cn_pcr = (m0.x != 0 || m0.y != 0) && grid >= 4 && grid <= kCnPcrMaxRows + 1 ? 1 : 0;
    }
    // spectral symbol of the linear part: i freq d1 - freq^2 d2 (split_linear_orbitals.c:275-278)
    std::vector<double2> sym(n);
    for (int i = 0; i < n; i++) {
        const double freq = (i <= (n - 1) / 2) ? (2 * M_PI * i) / (n * eq.dx) : (2 * M_PI * (i - n)) / (n * eq.dx);
        sym[i] = make_double2(-freq * eq.d1coef.y - freq * freq * eq.d2coef, freq * eq.d1coef.x);
    }
    MCTDHB_HIP_CHECK(hipMemcpy(w->d_hsym, sym.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    int rc = orbwork_set_potential(w, pot_host);
    if (rc) {
        orbwork_destroy(w);
        return rc;
    }
    *out = w;
    return kOk;
}

int orbwork_set_potential(OrbWork* w, const double* pot_host) {
    MCTDHB_HIP_CHECK(hipMemcpy(w->d_pot, pot_host, sizeof(double) * w->eq.grid_size, hipMemcpyHostToDevice));
    return kOk;
}

void orbwork_destroy(OrbWork* w) {
    if (!w) return;
    hipFree(w->d_pot);
    hipFree(w->d_fexp);
    hipFree(w->d_tw);
    hipFree(w->d_cn);
    hipFree(w->d_stage);
    hipFree(w->d_red);
    hipFree(w->d_cnwork);
    hipFree(w->d_nl);
    hipFree(w->d_fsH);
    hipFree(w->d_fsmat);
    hipFree(w->d_hsym);
    hipFree(w->d_dvrT);
    delete w;
}

}  // namespace mctdhb
