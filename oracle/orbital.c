/*
 * orbital.c (oracle) - the orbital split-step and everything it calls:
 * Simpson quadrature, finite differences, Gram-Schmidt, h_kl and g_ksql
 * integrals, Crank-Nicolson and FFT linear half-steps, the nonlinear
 * right-hand side, the RK2/RK4/RK5 nonlinear step, the M x M Hermitian
 * eigen/inverse pieces of the density sync, and one full MCTDHB step.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Orbitals are stored row-major: orb[k * grid_size + x].
 *
 * Substitutions for code absent from this machine (both parity unpinned by
 * any reference test):
 *  - MKL DFTI (reference types_definition.c:224-230) -> orc_dft, a direct
 *    transform for odd lengths and radix-2 decimation for powers of two,
 *    with DFTI's 1/sqrt(n) forward and backward scale;
 *  - odelib cplx_rungekutta{2,4,5} (orbital_integration.c:384-416) ->
 *    Heun's 2nd-order, the classical 4th-order and Butcher's 6-stage
 *    5th-order tableaus (only their convergence orders are checked);
 *  - LAPACKE zheev / zhesv (lapack_interface.c:10-124) -> cyclic complex
 *    Jacobi rotations and Gauss-Jordan elimination with partial pivoting.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.141592653589793

/* Reference calculus.c:8-39 (3/8 rule on the last 4 points for even n). */
ocplx
orc_csimps(uint32_t n, double h, const ocplx* f)
{
    ocplx acc = 0;
    if (n < 3) return NAN;
    if (n % 2 == 0)
    {
        for (uint32_t i = 0; i < n - 4; i += 2) acc += f[i] + 4 * f[i + 1] + f[i + 2];
        acc *= h / 3;
        acc += (f[n - 4] + 3 * (f[n - 3] + f[n - 2]) + f[n - 1]) * 3 * h / 8;
        return acc;
    }
    for (uint32_t i = 0; i < n - 2; i += 2) acc += f[i] + 4 * f[i + 1] + f[i + 2];
    acc *= h / 3;
    return acc;
}

/* Reference calculus.c:41-72 */
double
orc_rsimps(uint32_t n, double h, const double* f)
{
    double acc = 0;
    if (n < 3) return NAN;
    if (n % 2 == 0)
    {
        for (uint32_t i = 0; i < n - 4; i += 2) acc += f[i] + 4 * f[i + 1] + f[i + 2];
        acc *= h / 3;
        acc += (f[n - 4] + 3 * (f[n - 3] + f[n - 2]) + f[n - 1]) * 3 * h / 8;
        return acc;
    }
    for (uint32_t i = 0; i < n - 2; i += 2) acc += f[i] + 4 * f[i + 1] + f[i + 2];
    acc *= h / 3;
    return acc;
}

/* Reference calculus.c:287-305: 4th-order central difference with periodic
 * wrap; the last point repeats the first. */
void
orc_dxfd(uint32_t n, double dx, const ocplx* f, ocplx* dfdx)
{
    double r = 1.0 / (12 * dx);
    dfdx[0] = (f[n - 3] - f[2] + 8 * (f[1] - f[n - 2])) * r;
    dfdx[1] = (f[n - 2] - f[3] + 8 * (f[2] - f[0])) * r;
    dfdx[n - 3] = (f[n - 5] - f[0] + 8 * (f[n - 2] - f[n - 4])) * r;
    dfdx[n - 2] = (f[n - 4] - f[1] + 8 * (f[0] - f[n - 3])) * r;
    dfdx[n - 1] = dfdx[0];
    for (uint32_t i = 2; i < n - 3; i++) dfdx[i] = (f[i - 2] - f[i + 2] + 8 * (f[i + 1] - f[i - 1])) * r;
}

/* Reference calculus.c:130-142 */
ocplx
orc_scalar_product(uint32_t n, double h, const ocplx* fstar, const ocplx* f)
{
    ocplx* integ = malloc(sizeof(ocplx) * n);
    for (uint32_t i = 0; i < n; i++) integ[i] = conj(fstar[i]) * f[i];
    ocplx r = orc_csimps(n, h, integ);
    free(integ);
    return r;
}

/* Reference calculus.c:116-128 */
void
orc_renormalize(uint32_t n, double dx, double new_norm, ocplx* f)
{
    double* m2 = malloc(sizeof(double) * n);
    for (uint32_t i = 0; i < n; i++) m2[i] = creal(f[i]) * creal(f[i]) + cimag(f[i]) * cimag(f[i]);
    double renorm = new_norm * sqrt(1.0 / orc_rsimps(n, dx, m2));
    for (uint32_t i = 0; i < n; i++) f[i] = f[i] * renorm;
    free(m2);
}

/* Reference calculus.c:170-191, including its point-by-point update: the
 * projection <F_j, F_i> is re-evaluated before every grid point while F_i
 * is being modified. */
void
orc_orthonormalize(uint32_t npts, double dx, uint32_t nfun, ocplx* F)
{
    orc_renormalize(npts, dx, 1.0, &F[0]);
    for (uint32_t i = 1; i < nfun; i++)
    {
        ocplx* Fi = &F[(size_t) i * npts];
        for (uint32_t j = 0; j < i; j++)
        {
            const ocplx* Fj = &F[(size_t) j * npts];
            for (uint32_t k = 0; k < npts; k++) Fi[k] -= orc_scalar_product(npts, dx, Fj, Fi) * Fj[k];
        }
        orc_renormalize(npts, dx, 1.0, Fi);
    }
}

/* Reference orbital_matrices.c:21-70 */
void
orc_orbital_hob(const orc_orbeq* eq, const ocplx* orb, ocplx* hob)
{
    const int    M = eq->norb, n = eq->grid_size;
    const double dx = eq->dx, a2 = eq->d2coef;
    const ocplx  a1 = eq->d1coef_re + I * eq->d1coef_im;
    const double* V = eq->pot;
    ocplx *ddxi = malloc(sizeof(ocplx) * n), *ddxj = malloc(sizeof(ocplx) * n);
    ocplx* integ = malloc(sizeof(ocplx) * n);
    for (int i = 0; i < M; i++)
    {
        const ocplx* oi = &orb[(size_t) i * n];
        orc_dxfd(n, dx, oi, ddxi);
        for (int j = i + 1; j < M; j++)
        {
            const ocplx* oj = &orb[(size_t) j * n];
            orc_dxfd(n, dx, oj, ddxj);
            for (int k = 0; k < n; k++)
            {
                ocplx part = -a2 * conj(ddxi[k]) * ddxj[k];
                part = part + a1 * conj(oi[k]) * ddxj[k];
                part = part + V[k] * conj(oi[k]) * oj[k];
                integ[k] = part;
            }
            ocplx part = orc_csimps(n, dx, integ);
            hob[i * M + j] = part;
            hob[j * M + i] = conj(part);
        }
        for (int k = 0; k < n; k++)
        {
            ocplx part = -a2 * conj(ddxi[k]) * ddxi[k];
            part = part + a1 * conj(oi[k]) * ddxi[k];
            part = part + V[k] * conj(oi[k]) * oi[k];
            integ[k] = part;
        }
        hob[i * M + i] = creal(orc_csimps(n, dx, integ));
    }
    free(ddxi);
    free(ddxj);
    free(integ);
}

/* Quadruple (a, b, c, d) -> a + bM + cM^2 + dM^3 */
#define QI(a, b, c, d) ((a) + (b) * M + (c) * M2 + (d) * M3)

/* g * int conj(phi_a phi_b) phi_c phi_d, with the factors multiplied in the
 * order the reference writes them. */
static ocplx
four_integral(const orc_orbeq* eq, const ocplx* orb, int a, int b, int c, int d, ocplx* integ)
{
    const int    n = eq->grid_size;
    const ocplx *pa = &orb[(size_t) a * n], *pb = &orb[(size_t) b * n];
    const ocplx *pc = &orb[(size_t) c * n], *pd = &orb[(size_t) d * n];
    for (int i = 0; i < n; i++) integ[i] = conj(pa[i] * pb[i]) * pc[i] * pd[i];
    return eq->g * orc_csimps(n, eq->dx, integ);
}

/* Reference orbital_matrices.c:72-287: each distinct integral is evaluated
 * once (same integrand factor order) and written to every index its
 * permutation/conjugation symmetry reaches. */
void
orc_orbital_hint(const orc_orbeq* eq, const ocplx* orb, ocplx* hint)
{
    const int M = eq->norb, M2 = M * M, M3 = M2 * M;
    ocplx*    integ = malloc(sizeof(ocplx) * eq->grid_size);
    ocplx     v;
    for (int k = 0; k < M; k++)
    {
        hint[QI(k, k, k, k)] = four_integral(eq, orb, k, k, k, k, integ);
        for (int s = k + 1; s < M; s++)
        {
            v = four_integral(eq, orb, k, s, k, k, integ);
            hint[QI(k, s, k, k)] = v;
            hint[QI(s, k, k, k)] = v;
            hint[QI(k, k, k, s)] = conj(v);
            hint[QI(k, k, s, k)] = conj(v);

            v = four_integral(eq, orb, s, k, s, s, integ);
            hint[QI(s, k, s, s)] = v;
            hint[QI(k, s, s, s)] = v;
            hint[QI(s, s, s, k)] = conj(v);
            hint[QI(s, s, k, s)] = conj(v);

            v = four_integral(eq, orb, k, s, s, k, integ);
            hint[QI(k, s, s, k)] = v;
            hint[QI(s, k, s, k)] = v;
            hint[QI(s, k, k, s)] = v;
            hint[QI(k, s, k, s)] = v;

            v = four_integral(eq, orb, k, k, s, s, integ);
            hint[QI(k, k, s, s)] = v;
            hint[QI(s, s, k, k)] = conj(v);

            for (int q = s + 1; q < M; q++)
            {
                v = four_integral(eq, orb, k, s, q, k, integ);
                hint[QI(k, s, q, k)] = v;
                hint[QI(k, s, k, q)] = v;
                hint[QI(s, k, k, q)] = v;
                hint[QI(s, k, q, k)] = v;
                hint[QI(k, q, s, k)] = conj(v);
                hint[QI(k, q, k, s)] = conj(v);
                hint[QI(q, k, k, s)] = conj(v);
                hint[QI(q, k, s, k)] = conj(v);

                v = four_integral(eq, orb, s, k, q, s, integ);
                hint[QI(s, k, q, s)] = v;
                hint[QI(k, s, q, s)] = v;
                hint[QI(k, s, s, q)] = v;
                hint[QI(s, k, s, q)] = v;
                hint[QI(s, q, k, s)] = conj(v);
                hint[QI(s, q, s, k)] = conj(v);
                hint[QI(q, s, s, k)] = conj(v);
                hint[QI(q, s, k, s)] = conj(v);

                v = four_integral(eq, orb, q, s, k, q, integ);
                hint[QI(q, s, k, q)] = v;
                hint[QI(q, s, q, k)] = v;
                hint[QI(s, q, q, k)] = v;
                hint[QI(s, q, k, q)] = v;
                hint[QI(k, q, s, q)] = conj(v);
                hint[QI(k, q, q, s)] = conj(v);
                hint[QI(q, k, s, q)] = conj(v);
                hint[QI(q, k, q, s)] = conj(v);

                v = four_integral(eq, orb, k, k, q, s, integ);
                hint[QI(k, k, q, s)] = v;
                hint[QI(k, k, s, q)] = v;
                hint[QI(q, s, k, k)] = conj(v);
                hint[QI(s, q, k, k)] = conj(v);

                v = four_integral(eq, orb, s, s, k, q, integ);
                hint[QI(s, s, k, q)] = v;
                hint[QI(s, s, q, k)] = v;
                hint[QI(k, q, s, s)] = conj(v);
                hint[QI(q, k, s, s)] = conj(v);

                v = four_integral(eq, orb, q, q, k, s, integ);
                hint[QI(q, q, k, s)] = v;
                hint[QI(q, q, s, k)] = v;
                hint[QI(k, s, q, q)] = conj(v);
                hint[QI(s, k, q, q)] = conj(v);

                for (int l = q + 1; l < M; l++)
                {
                    v = four_integral(eq, orb, k, s, q, l, integ);
                    hint[QI(k, s, q, l)] = v;
                    hint[QI(k, s, l, q)] = v;
                    hint[QI(s, k, q, l)] = v;
                    hint[QI(s, k, l, q)] = v;
                    hint[QI(q, l, k, s)] = conj(v);
                    hint[QI(l, q, k, s)] = conj(v);
                    hint[QI(l, q, s, k)] = conj(v);
                    hint[QI(q, l, s, k)] = conj(v);

                    v = four_integral(eq, orb, k, q, s, l, integ);
                    hint[QI(k, q, s, l)] = v;
                    hint[QI(k, q, l, s)] = v;
                    hint[QI(q, k, s, l)] = v;
                    hint[QI(q, k, l, s)] = v;
                    hint[QI(s, l, k, q)] = conj(v);
                    hint[QI(s, l, q, k)] = conj(v);
                    hint[QI(l, s, q, k)] = conj(v);
                    hint[QI(l, s, k, q)] = conj(v);

                    v = four_integral(eq, orb, k, l, s, q, integ);
                    hint[QI(k, l, s, q)] = v;
                    hint[QI(k, l, q, s)] = v;
                    hint[QI(l, k, s, q)] = v;
                    hint[QI(l, k, q, s)] = v;
                    hint[QI(s, q, k, l)] = conj(v);
                    hint[QI(s, q, l, k)] = conj(v);
                    hint[QI(q, s, l, k)] = conj(v);
                    hint[QI(q, s, k, l)] = conj(v);
                }
            }
        }
    }
    free(integ);
}
#undef QI

/* Reference split_linear_orbitals.c:22-54 (dt = prop_dt / 2). */
void
orc_cn_tridiagonal(const orc_orbeq* eq, ocplx* upper, ocplx* lower, ocplx* mid)
{
    const ocplx  dt = 0.5 * (eq->prop_dt_re + I * eq->prop_dt_im);
    const int    n = eq->grid_size;
    const double a2 = eq->d2coef, dx = eq->dx;
    const ocplx  a1 = eq->d1coef_re + I * eq->d1coef_im;
    for (int i = 0; i < n; i++) mid[i] = I + a2 * dt / dx / dx;
    for (int i = 0; i < n - 2; i++)
    {
        upper[i] = -a2 * dt / dx / dx / 2 - a1 * dt / dx / 4;
        lower[i] = -a2 * dt / dx / dx / 2 + a1 * dt / dx / 4;
    }
    if (eq->bounds == 1000)
    {
        upper[n - 2] = -a2 * dt / dx / dx / 2 - a1 * dt / dx / 4;
        lower[n - 2] = -a2 * dt / dx / dx / 2 + a1 * dt / dx / 4;
    }
    else
    {
        upper[n - 2] = -a2 * dt / dx / dx / 2 + a1 * dt / dx / 4;
        lower[n - 2] = -a2 * dt / dx / dx / 2 - a1 * dt / dx / 4;
    }
}

/* Reference split_linear_orbitals.c:288-333 */
void
orc_cn_rhs(const orc_orbeq* eq, const ocplx* f, ocplx* out)
{
    const ocplx  dt = 0.5 * (eq->prop_dt_re + I * eq->prop_dt_im);
    const int    n = eq->grid_size;
    const double a2 = eq->d2coef, dx = eq->dx;
    const ocplx  a1 = eq->d1coef_re + I * eq->d1coef_im;
    const ocplx  up = a2 * dt / dx / dx / 2 + a1 * dt / dx / 4;
    const ocplx  lo = a2 * dt / dx / dx / 2 - a1 * dt / dx / 4;
    const ocplx  mid = I - a2 * dt / dx / dx;
    if (eq->bounds == 1000)
        out[0] = mid * f[0] + up * f[1];
    else
        out[0] = mid * f[0] + up * f[1] + lo * f[n - 2];
    for (int i = 1; i < n - 1; i++) out[i] = mid * f[i] + up * f[i + 1] + lo * f[i - 1];
    if (eq->bounds == 1000)
        out[n - 1] = mid * f[n - 1] + lo * f[n - 2];
    else
    {
        out[n - 2] = mid * f[n - 2] + up * f[0] + lo * f[n - 3];
        out[n - 1] = out[0];
    }
}

/* Reference tridiagonal_solver.c:6-92 (Thomas algorithm, with the reduced
 * system used when mid[0] == 0). */
void
orc_solve_tridiag(uint32_t n, const ocplx* upper, const ocplx* lower, const ocplx* mid,
                  const ocplx* rhs, ocplx* ans)
{
    ocplx *u = malloc(sizeof(ocplx) * n), *l = malloc(sizeof(ocplx) * n);
    ocplx* z = malloc(sizeof(ocplx) * n);
    if (cabs(mid[0]) == 0)
    {
        ocplx r1 = rhs[1] - mid[1] * rhs[0] / upper[0];
        ocplx r2 = rhs[2] - lower[1] * rhs[0] / upper[0];
        u[1] = lower[0];
        z[1] = r1;
        l[1] = 0;
        u[2] = mid[2] - l[1] * upper[1];
        z[2] = r2 - l[1] * z[1];
        for (uint32_t i = 2; i < n - 1; i++)
        {
            l[i] = lower[i] / u[i];
            u[i + 1] = mid[i + 1] - l[i] * upper[i];
            z[i + 1] = rhs[i + 1] - l[i] * z[i];
        }
        ans[n - 1] = z[n - 1] / u[n - 1];
        for (uint32_t i = 2; i <= n - 1; i++)
        {
            uint32_t k = n - i;
            ans[k] = (z[k] - upper[k] * ans[k + 1]) / u[k];
        }
        ans[0] = ans[1];
        ans[1] = rhs[0] / upper[0];
    }
    else
    {
        u[0] = mid[0];
        z[0] = rhs[0];
        for (uint32_t i = 0; i < n - 1; i++)
        {
            l[i] = lower[i] / u[i];
            u[i + 1] = mid[i + 1] - l[i] * upper[i];
            z[i + 1] = rhs[i + 1] - l[i] * z[i];
        }
        ans[n - 1] = z[n - 1] / u[n - 1];
        for (uint32_t i = 2; i <= n; i++)
        {
            uint32_t k = n - i;
            ans[k] = (z[k] - upper[k] * ans[k + 1]) / u[k];
        }
    }
    free(u);
    free(l);
    free(z);
}

/* Reference tridiagonal_solver.c:161-212: cyclic system through
 * Sherman-Morrison with gamma = mid[0] (or -upper[0] if mid[0] == 0).
 * upper[n-1] couples row 0 to x_{n-1} and lower[n-1] row n-1 to x_0. */
void
orc_solve_cyclic_tridiag_sm(uint32_t n, const ocplx* upper, const ocplx* lower, ocplx* mid,
                            const ocplx* rhs, ocplx* ans)
{
    ocplx *x = malloc(sizeof(ocplx) * n), *w = malloc(sizeof(ocplx) * n);
    ocplx *U = calloc(n, sizeof(ocplx)), *V = calloc(n, sizeof(ocplx));
    ocplx  keep0 = mid[0], keepn = mid[n - 1], factor;
    if (cabs(mid[0]) == 0)
    {
        factor = upper[0];
        mid[0] = -factor;
    }
    else
    {
        factor = mid[0];
        mid[0] = 0;
    }
    U[0] = factor;
    V[0] = 1;
    U[n - 1] = lower[n - 1];
    V[n - 1] = upper[n - 1] / factor;
    mid[n - 1] = mid[n - 1] - upper[n - 1] * lower[n - 1] / factor;
    orc_solve_tridiag(n, upper, lower, mid, rhs, x);
    orc_solve_tridiag(n, upper, lower, mid, U, w);
    ocplx vx = 0, vw = 0;
    for (uint32_t i = 0; i < n; i++) vx = vx + V[i] * x[i];
    for (uint32_t i = 0; i < n; i++) vw = vw + V[i] * w[i];
    factor = vx / (1.0 + vw);
    for (uint32_t i = 0; i < n; i++) ans[i] = x[i] + (-1) * factor * w[i];
    mid[0] = keep0;
    mid[n - 1] = keepn;
    free(x);
    free(w);
    free(U);
    free(V);
}

/* Reference split_linear_orbitals.c:335-373 */
void
orc_advance_linear_cn(const orc_orbeq* eq, ocplx* orb)
{
    const int n = eq->grid_size;
    ocplx *up = malloc(sizeof(ocplx) * n), *lo = malloc(sizeof(ocplx) * n);
    ocplx *mid = malloc(sizeof(ocplx) * n), *rhs = malloc(sizeof(ocplx) * n);
    orc_cn_tridiagonal(eq, up, lo, mid);
    int npts = (eq->bounds == 1000) ? n : n - 1;
    for (int k = 0; k < eq->norb; k++)
    {
        ocplx* o = &orb[(size_t) k * n];
        orc_cn_rhs(eq, o, rhs);
        if (eq->bounds == 1000)
            orc_solve_tridiag(npts, up, lo, mid, rhs, o);
        else
        {
            orc_solve_cyclic_tridiag_sm(npts, up, lo, mid, rhs, o);
            o[npts] = o[0];
        }
    }
    free(up);
    free(lo);
    free(mid);
    free(rhs);
}

/* Reference split_linear_orbitals.c:56-80 */
void
orc_hder_fftexp(const orc_orbeq* eq, ocplx* hder_exp)
{
    const ocplx  dt = 0.5 * (eq->prop_dt_re + I * eq->prop_dt_im);
    const int    m = eq->grid_size - 1;
    const double dx = eq->dx, d2 = eq->d2coef;
    const ocplx  d1 = eq->d1coef_re + I * eq->d1coef_im;
    for (int i = 0; i < m; i++)
    {
        double freq;
        if (i <= (m - 1) / 2)
            freq = (2 * ORC_PI * i) / (m * dx);
        else
            freq = (2 * ORC_PI * (i - m)) / (m * dx);
        hder_exp[i] = cexp(-I * dt * (I * d1 * freq - d2 * freq * freq));
    }
}

static ocplx
twiddle(int64_t num, int64_t n, int sign)
{
    int64_t r = num % n;
    if (r < 0) r += n;
    double ang = 2.0 * ORC_PI * (double) r / (double) n;
    return cos(ang) + I * sign * sin(ang);
}

static void
fft_rec(int n, int stride, int sign, const ocplx* in, ocplx* out, ocplx* scratch)
{
    if (n == 1)
    {
        out[0] = in[0];
        return;
    }
    int h = n / 2;
    fft_rec(h, 2 * stride, sign, in, scratch, out);
    fft_rec(h, 2 * stride, sign, in + stride, scratch + h, out + h);
    for (int k = 0; k < h; k++)
    {
        ocplx t = twiddle(k, n, sign) * scratch[h + k];
        out[k] = scratch[k] + t;
        out[k + h] = scratch[k] - t;
    }
}

/* out_k = scale * sum_j in_j exp(sign * 2 pi i j k / n) */
void
orc_dft(int n, int sign, double scale, const ocplx* in, ocplx* out)
{
    ocplx* tmp = malloc(sizeof(ocplx) * n);
    if ((n & (n - 1)) == 0)
    {
        ocplx* scratch = malloc(sizeof(ocplx) * n);
        fft_rec(n, 1, sign, in, tmp, scratch);
        free(scratch);
    }
    else
    {
        for (int k = 0; k < n; k++)
        {
            ocplx acc = 0;
            for (int j = 0; j < n; j++) acc += in[j] * twiddle((int64_t) j * k, n, sign);
            tmp[k] = acc;
        }
    }
    for (int k = 0; k < n; k++) out[k] = tmp[k] * scale;
    free(tmp);
}

/* Reference split_linear_orbitals.c:375-402 */
void
orc_advance_linear_fft(const orc_orbeq* eq, ocplx* orb)
{
    const int m = eq->grid_size - 1;
    double    scale = 1.0 / sqrt((double) m);
    ocplx *   fw = malloc(sizeof(ocplx) * m), *ex = malloc(sizeof(ocplx) * m);
    orc_hder_fftexp(eq, ex);
    for (int k = 0; k < eq->norb; k++)
    {
        ocplx* o = &orb[(size_t) k * eq->grid_size];
        orc_dft(m, -1, scale, o, fw);
        for (int i = 0; i < m; i++) fw[i] = ex[i] * fw[i];
        orc_dft(m, +1, scale, fw, o);
        o[m] = o[0];
    }
    free(fw);
    free(ex);
}

/* Reference split_nonlinear_orbitals.c:62-151 */
ocplx
orc_orb_full_nonlinear(int k, int n, double g, int m, int grid_size, const ocplx* orb,
                       const ocplx* hob, const ocplx* hint, const ocplx* rho,
                       const ocplx* rinv)
{
    const int mm = m * m, mmm = mm * m;
#define ORB(a) orb[(size_t) (a) * grid_size + n]
    ocplx red = 0, rm, sw;
    for (int a = 0; a < m; a++)
    {
        red -= hob[a * m + k] * ORB(a);
        for (int q = 0; q < m; q++)
        {
            rm = rinv[k * m + a] * rho[a + m * a + mm * q + mmm * q];
            red += g * rm * conj(ORB(a)) * ORB(q) * ORB(q);
            for (int j = 0; j < m; j++) red -= rm * ORB(j) * hint[j + a * m + q * mm + q * mmm];
            for (int l = q + 1; l < m; l++)
            {
                rm = 2 * rinv[k * m + a] * rho[a + m * a + mm * q + mmm * l];
                red += g * rm * conj(ORB(a)) * ORB(l) * ORB(q);
                for (int j = 0; j < m; j++)
                    red -= rm * ORB(j) * hint[j + a * m + l * mm + q * mmm];
            }
        }
        for (int s = a + 1; s < m; s++)
        {
            for (int q = 0; q < m; q++)
            {
                rm = rinv[k * m + a] * rho[a + m * s + mm * q + mmm * q];
                sw = rinv[k * m + s] * rho[a + m * s + mm * q + mmm * q];
                red += g * (rm * conj(ORB(s)) + sw * conj(ORB(a))) * ORB(q) * ORB(q);
                for (int j = 0; j < m; j++)
                {
                    red -= rm * ORB(j) * hint[j + s * m + q * mm + q * mmm];
                    red -= sw * ORB(j) * hint[j + a * m + q * mm + q * mmm];
                }
                for (int l = q + 1; l < m; l++)
                {
                    rm = 2 * rinv[k * m + a] * rho[a + m * s + mm * q + mmm * l];
                    sw = 2 * rinv[k * m + s] * rho[a + m * s + mm * q + mmm * l];
                    red += g * (rm * conj(ORB(s)) + sw * conj(ORB(a))) * ORB(l) * ORB(q);
                    for (int j = 0; j < m; j++)
                    {
                        red -= rm * ORB(j) * hint[j + s * m + l * mm + q * mmm];
                        red -= sw * ORB(j) * hint[j + a * m + l * mm + q * mmm];
                    }
                }
            }
        }
    }
#undef ORB
    return red;
}

/* Reference orbital_integration.c:242-281:
 * der[k][x] = -i time_fac (nonlinear(k, x) + V(x) orb[k][x]). */
void
orc_nonlinear_rhs(const orc_orbeq* eq, const ocplx* orb, const ocplx* hob, const ocplx* hint,
                  const ocplx* rho2, const ocplx* rinv, ocplx* der)
{
    const int   M = eq->norb, n = eq->grid_size;
    const ocplx tf = eq->time_fac_re + I * eq->time_fac_im;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < M; k++)
        for (int x = 0; x < n; x++)
            der[(size_t) k * n + x] =
                -I * tf *
                (orc_orb_full_nonlinear(k, x, eq->g, M, n, orb, hob, hint, rho2, rinv) +
                 eq->pot[x] * orb[(size_t) k * n + x]);
}

/* ------------------------------------------------------ M x M algebra */

/* Hermitian eigen-decomposition by cyclic complex Jacobi rotations.
 * eigvec (n x n row-major) holds eigenvectors in its columns, eigenvalues
 * ascending (the zheev convention).  Returns 0, or 1 without convergence. */
int
orc_herm_eig(int n, const ocplx* A, double* eigvals, ocplx* eigvec)
{
    ocplx* a = malloc(sizeof(ocplx) * n * n);
    memcpy(a, A, sizeof(ocplx) * n * n);
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) eigvec[r * n + c] = (r == c) ? 1.0 : 0.0;
    int ok = 0;
    for (int sweep = 0; sweep < 100; sweep++)
    {
        double off = 0, tot = 0;
        for (int r = 0; r < n; r++)
            for (int c = 0; c < n; c++)
            {
                double v = cabs(a[r * n + c]);
                tot += v * v;
                if (r != c) off += v * v;
            }
        if (off <= 1e-30 * tot || off == 0)
        {
            ok = 1;
            break;
        }
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++)
            {
                ocplx  apq = a[p * n + q];
                double mag = cabs(apq);
                if (mag == 0) continue;
                double app = creal(a[p * n + p]), aqq = creal(a[q * n + q]);
                /* rotation zeroing a[p][q]: with phase e = apq/|apq| the 2x2
                 * block becomes real symmetric [app mag; mag aqq]. */
                ocplx  ph = apq / mag;
                double tau = (aqq - app) / (2 * mag);
                double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
                double c = 1 / sqrt(1 + t * t), s = t * c;
                /* columns: J = [c, s*ph; -s*conj(ph), c] applied as A <- J^H A J */
                for (int r = 0; r < n; r++)
                {
                    ocplx arp = a[r * n + p], arq = a[r * n + q];
                    a[r * n + p] = c * arp - s * conj(ph) * arq;
                    a[r * n + q] = s * ph * arp + c * arq;
                }
                for (int r = 0; r < n; r++)
                {
                    ocplx apr = a[p * n + r], aqr = a[q * n + r];
                    a[p * n + r] = c * apr - s * ph * aqr;
                    a[q * n + r] = s * conj(ph) * apr + c * aqr;
                }
                a[p * n + q] = 0;
                a[q * n + p] = 0;
                for (int r = 0; r < n; r++)
                {
                    ocplx vrp = eigvec[r * n + p], vrq = eigvec[r * n + q];
                    eigvec[r * n + p] = c * vrp - s * conj(ph) * vrq;
                    eigvec[r * n + q] = s * ph * vrp + c * vrq;
                }
            }
    }
    for (int r = 0; r < n; r++) eigvals[r] = creal(a[r * n + r]);
    for (int x = 0; x < n - 1; x++)
    {
        int best = x;
        for (int y = x + 1; y < n; y++)
            if (eigvals[y] < eigvals[best]) best = y;
        if (best != x)
        {
            double t = eigvals[x];
            eigvals[x] = eigvals[best];
            eigvals[best] = t;
            for (int r = 0; r < n; r++)
            {
                ocplx tv = eigvec[r * n + x];
                eigvec[r * n + x] = eigvec[r * n + best];
                eigvec[r * n + best] = tv;
            }
        }
    }
    free(a);
    return ok ? 0 : 1;
}

/* Inverse of a Hermitian matrix given by its lower triangle and real
 * diagonal (the 'L' storage zhesv reads in lapack_interface.c:10-59), by
 * Gauss-Jordan elimination with partial pivoting.  Returns 0, or k > 0 if
 * pivot k is exactly zero (singular). */
int
orc_herm_inverse(int n, const ocplx* A, ocplx* Ainv)
{
    ocplx* a = malloc(sizeof(ocplx) * n * n);
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++)
        {
            if (r == c) a[r * n + c] = creal(A[r * n + r]);
            else if (c < r) a[r * n + c] = A[r * n + c];
            else a[r * n + c] = conj(A[c * n + r]);
            Ainv[r * n + c] = (r == c) ? 1.0 : 0.0;
        }
    int info = 0;
    for (int col = 0; col < n; col++)
    {
        int piv = col;
        for (int r = col + 1; r < n; r++)
            if (cabs(a[r * n + col]) > cabs(a[piv * n + col])) piv = r;
        if (cabs(a[piv * n + col]) == 0)
        {
            info = col + 1;
            break;
        }
        if (piv != col)
            for (int c = 0; c < n; c++)
            {
                ocplx t = a[col * n + c];
                a[col * n + c] = a[piv * n + c];
                a[piv * n + c] = t;
                t = Ainv[col * n + c];
                Ainv[col * n + c] = Ainv[piv * n + c];
                Ainv[piv * n + c] = t;
            }
        ocplx inv = 1.0 / a[col * n + col];
        for (int c = 0; c < n; c++)
        {
            a[col * n + c] *= inv;
            Ainv[col * n + c] *= inv;
        }
        for (int r = 0; r < n; r++)
        {
            if (r == col) continue;
            ocplx f = a[r * n + col];
            if (f == 0) continue;
            for (int c = 0; c < n; c++)
            {
                a[r * n + c] -= f * a[col * n + c];
                Ainv[r * n + c] -= f * Ainv[col * n + c];
            }
        }
    }
    free(a);
    return info;
}

/* Reference lapack_interface.c:126-163: A <- A + V diag(x exp(-lambda/x)) V^H. */
void
orc_regularization(int n, double x, ocplx* A)
{
    ocplx*  V = malloc(sizeof(ocplx) * n * n);
    double* lam = malloc(sizeof(double) * n);
    /* zheev reads the upper triangle with a real diagonal */
    ocplx* Au = malloc(sizeof(ocplx) * n * n);
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++)
            Au[r * n + c] = (r == c) ? creal(A[r * n + r]) : (c > r ? A[r * n + c] : conj(A[c * n + r]));
    orc_herm_eig(n, Au, lam, V);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
        {
            ocplx z = 0;
            for (int k = 0; k < n; k++) z += V[i * n + k] * (x * exp(-lam[k] / x)) * conj(V[j * n + k]);
            A[i * n + j] = A[i * n + j] + z;
        }
    free(V);
    free(lam);
    free(Au);
}

/* Reference synchronize.c:41-56 */
void
orc_sync_density(const orc_space* S, const ocplx* coef, ocplx* rho1, ocplx* rho2, ocplx* rinv,
                 double reg_factor)
{
    const int M = S->norb;
    double    occ[64];
    ocplx     V[64 * 64];
    orc_onebody_dm(S, coef, rho1);
    orc_twobody_dm(S, coef, rho2);
    ocplx Au[64 * 64];
    for (int r = 0; r < M; r++)
        for (int c = 0; c < M; c++)
            Au[r * M + c] = (r == c) ? creal(rho1[r * M + r]) : (c > r ? rho1[r * M + c] : conj(rho1[c * M + r]));
    orc_herm_eig(M, Au, occ, V);
    if (occ[0] / occ[M - 1] < reg_factor) orc_regularization(M, occ[M - 1] * reg_factor, rho1);
    orc_herm_inverse(M, rho1, rinv);
}

/* Classical RK4 step of the nonlinear orbital equation with h, g, rho2
 * and rho1^-1 held fixed (the role odelib plays at orbital_integration.c:384-416). */
static void
rk4_nonlinear(const orc_orbeq* eq, ocplx* orb, const ocplx* hob, const ocplx* hint,
              const ocplx* rho2, const ocplx* rinv)
{
    const size_t n = (size_t) eq->norb * eq->grid_size;
    const double dt = eq->tstep;
    ocplx *k1 = malloc(sizeof(ocplx) * n), *k2 = malloc(sizeof(ocplx) * n);
    ocplx *k3 = malloc(sizeof(ocplx) * n), *k4 = malloc(sizeof(ocplx) * n);
    ocplx* y = malloc(sizeof(ocplx) * n);
    orc_nonlinear_rhs(eq, orb, hob, hint, rho2, rinv, k1);
    for (size_t i = 0; i < n; i++) y[i] = orb[i] + 0.5 * dt * k1[i];
    orc_nonlinear_rhs(eq, y, hob, hint, rho2, rinv, k2);
    for (size_t i = 0; i < n; i++) y[i] = orb[i] + 0.5 * dt * k2[i];
    orc_nonlinear_rhs(eq, y, hob, hint, rho2, rinv, k3);
    for (size_t i = 0; i < n; i++) y[i] = orb[i] + dt * k3[i];
    orc_nonlinear_rhs(eq, y, hob, hint, rho2, rinv, k4);
    for (size_t i = 0; i < n; i++) orb[i] = orb[i] + dt / 6 * (k1[i] + 2 * k2[i] + 2 * k3[i] + k4[i]);
    free(k1);
    free(k2);
    free(k3);
    free(k4);
    free(y);
}

/* The 2nd- and 5th-order explicit tableaus: Heun's trapezoidal method and
 * Butcher's six-stage fifth-order method (J. C. Butcher, J. Austral. Math.
 * Soc. 4 (1964) 179-194).  odelib's own tableaus are not available. */
static const double RK2_A[2 * 2] = {0, 0, 1, 0};
static const double RK2_B[2] = {0.5, 0.5};
static const double RK5_A[6 * 6] = {
    0,          0,         0,          0,           0,         0,
    0.25,       0,         0,          0,           0,         0,
    0.125,      0.125,     0,          0,           0,         0,
    0,          -0.5,      1,          0,           0,         0,
    3.0 / 16,   0,         0,          9.0 / 16,    0,         0,
    -3.0 / 7,   2.0 / 7,   12.0 / 7,   -12.0 / 7,   8.0 / 7,   0,
};
static const double RK5_B[6] = {7.0 / 90, 0, 32.0 / 90, 12.0 / 90, 32.0 / 90, 7.0 / 90};

/* One explicit RK step with an s-stage tableau; stage sums run over j in
 * ascending order: y = y0 + dt * (a_0 k_0 + a_1 k_1 + ...). */
static void
tableau_nonlinear(const orc_orbeq* eq, int s, const double* A, const double* b, ocplx* orb,
                  const ocplx* hob, const ocplx* hint, const ocplx* rho2, const ocplx* rinv)
{
    const size_t n = (size_t) eq->norb * eq->grid_size;
    const double dt = eq->tstep;
    ocplx* k = malloc(sizeof(ocplx) * n * s);
    ocplx* y = malloc(sizeof(ocplx) * n);
    for (int st = 0; st < s; st++)
    {
        if (st == 0)
        {
            orc_nonlinear_rhs(eq, orb, hob, hint, rho2, rinv, k);
            continue;
        }
        for (size_t i = 0; i < n; i++)
        {
            ocplx acc = A[st * s] * k[i];
            for (int j = 1; j < st; j++) acc += A[st * s + j] * k[j * n + i];
            y[i] = orb[i] + dt * acc;
        }
        orc_nonlinear_rhs(eq, y, hob, hint, rho2, rinv, k + st * n);
    }
    for (size_t i = 0; i < n; i++)
    {
        ocplx acc = b[0] * k[i];
        for (int j = 1; j < s; j++) acc += b[j] * k[j * n + i];
        orb[i] = orb[i] + dt * acc;
    }
    free(k);
    free(y);
}

/* The nonlinear sub-step of order eq->rk_order (reference
 * orbital_integration.c:384-416 switches on RK2 / RK4 / RK5). */
void
orc_rk_nonlinear(const orc_orbeq* eq, ocplx* orb, const ocplx* hob, const ocplx* hint,
                 const ocplx* rho2, const ocplx* rinv)
{
    if (eq->rk_order == 2)
        tableau_nonlinear(eq, 2, RK2_A, RK2_B, orb, hob, hint, rho2, rinv);
    else if (eq->rk_order == 5)
        tableau_nonlinear(eq, 6, RK5_A, RK5_B, orb, hob, hint, rho2, rinv);
    else
        rk4_nonlinear(eq, orb, hob, hint, rho2, rinv);
}

static void
periodic_bounds(const orc_orbeq* eq, ocplx* orb)
{
    for (int k = 0; k < eq->norb; k++)
        orb[(size_t) k * eq->grid_size + eq->grid_size - 1] = orb[(size_t) k * eq->grid_size];
}

/* Reference orbital_integration.c:351-434 (split-step: linear half step,
 * h/g rebuild, nonlinear RK step, linear half step, Gram-Schmidt in
 * imaginary time).  hob/hint are overwritten by the rebuild. */
void
orc_splitstep_orb(const orc_orbeq* eq, int imagtime, int use_fft, ocplx* orb, ocplx* hob,
                  ocplx* hint, const ocplx* rho2, const ocplx* rinv)
{
    if (eq->bounds == 1001) periodic_bounds(eq, orb);
    if (use_fft)
        orc_advance_linear_fft(eq, orb);
    else
        orc_advance_linear_cn(eq, orb);
    orc_orbital_hob(eq, orb, hob);
    orc_orbital_hint(eq, orb, hint);
    orc_rk_nonlinear(eq, orb, hob, hint, rho2, rinv);
    if (use_fft)
        orc_advance_linear_fft(eq, orb);
    else
        orc_advance_linear_cn(eq, orb);
    if (imagtime) orc_orthonormalize(eq->grid_size, eq->dx, eq->norb, orb);
}

/* Reference driver.c:159-202 for a time-independent trap and interaction. */
void
orc_propagate_step(const orc_space* S, orc_orbeq* eq, int imagtime, int use_fft,
                   int lanczos_iter, ocplx* orb, ocplx* coef, ocplx* hob, ocplx* hint,
                   ocplx* rho1, ocplx* rho2, ocplx* rinv, ocplx* lvec_work)
{
    orc_splitstep_orb(eq, imagtime, use_fft, orb, hob, hint, rho2, rinv);
    if (eq->bounds == 1001) periodic_bounds(eq, orb);
    orc_orbital_hob(eq, orb, hob);
    orc_orbital_hint(eq, orb, hint);
    if (eq->norb == 1) return;
    orc_sil_step(S, hob, hint, lanczos_iter, eq->prop_dt_re, eq->prop_dt_im, imagtime, coef,
                 lvec_work);
    orc_sync_density(S, coef, rho1, rho2, rinv, 1e-9);
}
