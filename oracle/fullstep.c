/*
 * fullstep.c (oracle) - the full-equation Runge-Kutta orbital step and the
 * pieces it calls: the linear part by finite differences or FFT, the
 * interacting part, the overlap matrix (with quirk Q2), the simple and the
 * robust projector onto the orbital space, and the time derivative.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Orbitals are stored row-major: orb[k * grid_size + x].  The Runge-Kutta
 * tableaus are the ones orbital.c uses for the split step (odelib is not
 * available; parity unpinned, convergence order checked).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.141592653589793

/* Reference split_linear_orbitals.c:209-255 (finite differences). */
void
orc_linear_horb_fd(const orc_orbeq* eq, const ocplx* f, ocplx* out)
{
    const int    npts = eq->grid_size;
    const double a2 = eq->d2coef, dx = eq->dx;
    const ocplx  a1 = eq->d1coef_re + I * eq->d1coef_im;
    const ocplx  upper = a2 / dx / dx + a1 / dx / 2;
    const ocplx  lower = a2 / dx / dx - a1 / dx / 2;
    const double* pot = eq->pot;
    ocplx mid = -2 * a2 / dx / dx + pot[0];
    if (eq->bounds == 1000)
        out[0] = mid * f[0] + upper * f[1];
    else
        out[0] = mid * f[0] + upper * f[1] + lower * f[npts - 2];
    for (int i = 1; i < npts - 1; i++)
    {
        mid = -2 * a2 / dx / dx + pot[i];
        out[i] = mid * f[i] + upper * f[i + 1] + lower * f[i - 1];
    }
    if (eq->bounds == 1000)
    {
        mid = -2 * a2 / dx / dx + pot[npts - 1];
        out[npts - 1] = mid * f[npts - 1] + lower * f[npts - 2];
    } else
    {
        mid = -2 * a2 / dx / dx + pot[npts - 2];
        out[npts - 2] = mid * f[npts - 2] + upper * f[0] + lower * f[npts - 3];
        out[npts - 1] = out[0];
    }
}

/* Reference split_linear_orbitals.c:257-286 (spectral derivative). */
void
orc_linear_horb_fft(const orc_orbeq* eq, const ocplx* f, ocplx* out)
{
    const int    n = eq->grid_size - 1;
    const double d2 = eq->d2coef, dx = eq->dx, scale = 1.0 / sqrt((double) n);
    const ocplx  d1 = eq->d1coef_re + I * eq->d1coef_im;
    ocplx *      fw = malloc(sizeof(ocplx) * n), *bw = malloc(sizeof(ocplx) * n);
    orc_dft(n, -1, scale, f, fw);
    for (int i = 0; i < n; i++)
    {
        const double freq = (i <= (n - 1) / 2) ? (2 * ORC_PI * i) / (n * dx) : (2 * ORC_PI * (i - n)) / (n * dx);
        fw[i] *= (freq * I * d1 - freq * freq * d2);
    }
    orc_dft(n, +1, scale, fw, bw);
    for (int i = 0; i < n; i++) out[i] = bw[i] + eq->pot[i] * f[i];
    out[n] = out[0];
    free(fw);
    free(bw);
}

/* Reference split_linear_orbitals.c:82-165: the exponential-DVR matrix of
 * the derivative terms, n = grid_size - 1 rows, row-major. */
void
orc_expdvr_mat(const orc_orbeq* eq, ocplx* mat)
{
    const int    n = eq->grid_size - 1;
    const double length = n * eq->dx, d2coef = eq->d2coef;
    const ocplx  d1coef = eq->d1coef_re + I * eq->d1coef_im;
    ocplx *      d1 = malloc(sizeof(ocplx) * n * n), *d2 = malloc(sizeof(ocplx) * n * n);
    for (int i = 0; i < n; i++)
    {
        for (int j = i + 1; j < n; j++)
        {
            ocplx s1 = 0, s2 = 0;
            for (int k = 0; k < n; k++)
            {
                const int   kmom = k - n / 2;
                const ocplx e = cexp(2 * I * kmom * (j - i) * ORC_PI / n);
                s1 += -(2 * I * ORC_PI * kmom / length) * e / n;
                s2 += -(2 * ORC_PI * kmom / length) * (2 * ORC_PI * kmom / length) * e / n;
            }
            d1[i * n + j] = s1;
            d1[j * n + i] = -conj(s1);
            d2[i * n + j] = s2;
            d2[j * n + i] = conj(s2);
        }
        ocplx s1 = 0, s2 = 0;
        for (int k = 0; k < n; k++)
        {
            const int kmom = k - n / 2;
            s1 = s1 + -(2 * I * ORC_PI * kmom / length) / n;
            s2 = s2 + -(2 * ORC_PI * kmom / length) * (2 * ORC_PI * kmom / length) / n;
        }
        d1[i * n + i] = s1;
        d2[i * n + i] = s2;
    }
    for (int i = 0; i < n * n; i++) mat[i] = d2coef * d2[i] + d1coef * d1[i];
    free(d1);
    free(d2);
}

/* Reference orbital_integration.c:159-169 (DVR branch): the matrix times
 * the first n points, plus V, last point = first. */
void
orc_linear_horb_dvr(const orc_orbeq* eq, const ocplx* mat, const ocplx* f, ocplx* out)
{
    const int n = eq->grid_size - 1;
    for (int j = 0; j < n; j++)
    {
        ocplx acc = 0;
        for (int l = 0; l < n; l++) acc += mat[j * n + l] * f[l];
        out[j] = acc + eq->pot[j] * f[j];
    }
    out[n] = out[0];
}

/* Reference split_nonlinear_orbitals.c:4-60 */
ocplx
orc_orb_interacting_part(int k, int n, double g, int m, int grid_size, const ocplx* orb, const ocplx* rho,
                         const ocplx* rinv)
{
    const int mm = m * m, mmm = mm * m;
#define ORB(a) orb[(size_t) (a) * grid_size + n]
    ocplx red = 0, grho, swap;
    for (int a = 0; a < m; a++)
    {
        for (int q = 0; q < m; q++)
        {
            grho = g * rinv[k * m + a] * rho[a + m * a + mm * q + mmm * q];
            red = red + grho * conj(ORB(a)) * ORB(q) * ORB(q);
            for (int l = q + 1; l < m; l++)
            {
                grho = 2 * g * rinv[k * m + a] * rho[a + m * a + mm * q + mmm * l];
                red = red + grho * conj(ORB(a)) * ORB(l) * ORB(q);
            }
        }
        for (int s = a + 1; s < m; s++)
        {
            for (int q = 0; q < m; q++)
            {
                grho = g * rinv[k * m + a] * rho[a + m * s + mm * q + mmm * q];
                swap = g * rinv[k * m + s] * rho[a + m * s + mm * q + mmm * q];
                red = red + (grho * conj(ORB(s)) + swap * conj(ORB(a))) * ORB(q) * ORB(q);
                for (int l = q + 1; l < m; l++)
                {
                    grho = 2 * g * rinv[k * m + a] * rho[a + m * s + mm * q + mmm * l];
                    swap = 2 * g * rinv[k * m + s] * rho[a + m * s + mm * q + mmm * l];
                    red = red + (grho * conj(ORB(s)) + swap * conj(ORB(a))) * ORB(l) * ORB(q);
                }
            }
        }
    }
#undef ORB
    return red;
}

/* Reference orbital_matrices.c:7-20.  Quirk Q2: the inner loop starts at
 * j = 2, so overlap[0][1] and overlap[1][0] are never set; the reference
 * leaves them as whatever the allocation held, here they are zero. */
void
orc_overlap_matrix(int norb, int grid, double dx, const ocplx* orb, ocplx* overlap)
{
    memset(overlap, 0, sizeof(ocplx) * norb * norb);
    for (int i = 0; i < norb; i++)
    {
        overlap[i * norb + i] = orc_scalar_product(grid, dx, &orb[(size_t) i * grid], &orb[(size_t) i * grid]);
        for (int j = 2; j < norb; j++)
        {
            overlap[i * norb + j] =
                orc_scalar_product(grid, dx, &orb[(size_t) i * grid], &orb[(size_t) j * grid]);
            overlap[j * norb + i] = conj(overlap[i * norb + j]);
        }
    }
}

/* Reference orbital_integration.c:22-93 (robust, with the inverse of the
 * overlap matrix) and :94-129 (simple). */
void
orc_multiorb_projector(const orc_orbeq* eq, int robust, const ocplx* orb, const ocplx* hact, ocplx* project)
{
    const int    norb = eq->norb, npts = eq->grid_size;
    const double dx = eq->dx;
    ocplx*       po = malloc(sizeof(ocplx) * norb * norb);
    ocplx*       ov = malloc(sizeof(ocplx) * norb * norb);
    ocplx*       inv = malloc(sizeof(ocplx) * norb * norb);
    if (robust)
    {
        orc_overlap_matrix(norb, npts, dx, orb, ov);
        orc_herm_inverse(norb, ov, inv);
    }
    for (int k = 0; k < norb; k++)
        for (int s = 0; s < norb; s++)
            po[s * norb + k] = orc_scalar_product(npts, dx, &orb[(size_t) s * npts], &hact[(size_t) k * npts]);
    for (int k = 0; k < norb; k++)
        for (int i = 0; i < npts; i++)
        {
            ocplx proj = 0;
            for (int s = 0; s < norb; s++)
            {
                if (robust)
                    for (int l = 0; l < norb; l++)
                        proj += orb[(size_t) s * npts + i] * inv[s * norb + l] * po[l * norb + k];
                else
                    proj += orb[(size_t) s * npts + i] * po[s * norb + k];
            }
            project[(size_t) k * npts + i] = proj;
        }
    free(po);
    free(ov);
    free(inv);
}

/* Reference orbital_integration.c:158-235: der = -i time_fac (H orb - P H orb)
 * with H = linear part (der: 0 FFT, 1 finite differences, 2 exponential
 * DVR) plus the interacting part. */
void
orc_fullstep_rhs(const orc_orbeq* eq, int der_method, int robust, const ocplx* orb, const ocplx* rho2,
                 const ocplx* rinv, ocplx* der)
{
    const int    norb = eq->norb, npts = eq->grid_size;
    const size_t n = (size_t) norb * npts;
    const ocplx  tf = eq->time_fac_re + I * eq->time_fac_im;
    ocplx *      hact = malloc(sizeof(ocplx) * n), *lin = malloc(sizeof(ocplx) * n), *proj = malloc(sizeof(ocplx) * n);
    ocplx*       dvr = NULL;
    if (der_method == 2)
    {
        dvr = malloc(sizeof(ocplx) * (npts - 1) * (npts - 1));
        orc_expdvr_mat(eq, dvr);
    }
    for (int k = 0; k < norb; k++)
    {
        if (der_method == 1)
            orc_linear_horb_fd(eq, &orb[(size_t) k * npts], &lin[(size_t) k * npts]);
        else if (der_method == 2)
            orc_linear_horb_dvr(eq, dvr, &orb[(size_t) k * npts], &lin[(size_t) k * npts]);
        else
            orc_linear_horb_fft(eq, &orb[(size_t) k * npts], &lin[(size_t) k * npts]);
    }
    free(dvr);
    for (int k = 0; k < norb; k++)
        for (int j = 0; j < npts; j++)
            hact[(size_t) k * npts + j] =
                orc_orb_interacting_part(k, j, eq->g, norb, npts, orb, rho2, rinv) + lin[(size_t) k * npts + j];
    orc_multiorb_projector(eq, robust, orb, hact, proj);
    for (size_t i = 0; i < n; i++) der[i] = -I * tf * (hact[i] - proj[i]);
    free(hact);
    free(lin);
    free(proj);
}

/* The tableaus of orbital.c: Heun (RK2), classical RK4, Butcher's RK5. */
static const double FS_RK2_A[4] = {0, 0, 1, 0};
static const double FS_RK2_B[2] = {0.5, 0.5};
static const double FS_RK4_A[16] = {0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 1, 0};
static const double FS_RK4_B[4] = {1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6};
static const double FS_RK5_A[36] = {
    0,        0,       0,        0,         0,       0, 0.25,     0,       0,        0,         0,       0,
    0.125,    0.125,   0,        0,         0,       0, 0,        -0.5,    1,        0,         0,       0,
    3.0 / 16, 0,       0,        9.0 / 16,  0,       0, -3.0 / 7, 2.0 / 7, 12.0 / 7, -12.0 / 7, 8.0 / 7, 0,
};
static const double FS_RK5_B[6] = {7.0 / 90, 0, 32.0 / 90, 12.0 / 90, 32.0 / 90, 7.0 / 90};

/* Reference orbital_integration.c:283-348: periodic copy, one explicit RK
 * step of the full equation with rho2 and rho1^-1 fixed, Gram-Schmidt in
 * imaginary time. */
void
orc_fullstep_orb(const orc_orbeq* eq, int imagtime, int der_method, int robust, ocplx* orb, const ocplx* rho2,
                 const ocplx* rinv)
{
    const int     norb = eq->norb, npts = eq->grid_size;
    const size_t  n = (size_t) norb * npts;
    const double  dt = eq->tstep;
    int           s;
    const double *A, *b;
    if (eq->rk_order == 2) s = 2, A = FS_RK2_A, b = FS_RK2_B;
    else if (eq->rk_order == 5) s = 6, A = FS_RK5_A, b = FS_RK5_B;
    else s = 4, A = FS_RK4_A, b = FS_RK4_B;
    if (eq->bounds == 1001)
        for (int k = 0; k < norb; k++) orb[(size_t) k * npts + npts - 1] = orb[(size_t) k * npts];
    ocplx* k = malloc(sizeof(ocplx) * n * s);
    ocplx* y = malloc(sizeof(ocplx) * n);
    for (int st = 0; st < s; st++)
    {
        if (st == 0)
        {
            orc_fullstep_rhs(eq, der_method, robust, orb, rho2, rinv, k);
            continue;
        }
        for (size_t i = 0; i < n; i++)
        {
            ocplx acc = A[st * s] * k[i];
            for (int j = 1; j < st; j++) acc += A[st * s + j] * k[j * n + i];
            y[i] = orb[i] + dt * acc;
        }
        orc_fullstep_rhs(eq, der_method, robust, y, rho2, rinv, k + st * n);
    }
    for (size_t i = 0; i < n; i++)
    {
        ocplx acc = b[0] * k[i];
        for (int j = 1; j < s; j++) acc += b[j] * k[j * n + i];
        orb[i] = orb[i] + dt * acc;
    }
    free(k);
    free(y);
    if (imagtime) orc_orthonormalize(npts, eq->dx, norb, orb);
}
