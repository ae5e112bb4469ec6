/*
 * lanczos.c (oracle) - vector kernels, Lanczos tridiagonalisation, the
 * symmetric tridiagonal eigensolver and the short-iterative-Lanczos (SIL)
 * coefficient step.  TEST INFRASTRUCTURE ONLY.
 *
 * The reference calls MKL LAPACKE_dstev (coefficients_integration.c:111-118,
 * multiconfig_lanczos.c:143-149); MKL is absent here, so orc_tridiag_eig is
 * an implicit-shift QL iteration (the EISPACK tql2 / Golub-Van Loan 8.3
 * method) returning the same eigen-decomposition in dstev's row-major layout
 * (eigenvector j in column j).  The SIL propagator exp(-i T dt) e1 does not
 * depend on eigenvector signs or eigenvalue order.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Reference basic_linalg.c:193-199 (conjugate on the first argument). */
ocplx
orc_carr_dot(uint32_t n, const ocplx* a, const ocplx* b)
{
    ocplx z = 0;
    for (uint32_t i = 0; i < n; i++) z = z + conj(a[i]) * b[i];
    return z;
}

/* Reference basic_linalg.c:217-226 */
double
orc_carr_mod(uint32_t n, const ocplx* a)
{
    double s = 0;
    for (uint32_t i = 0; i < n; i++) s = s + creal(a[i]) * creal(a[i]) + cimag(a[i]) * cimag(a[i]);
    return sqrt(s);
}

/* Reference multiconfig_lanczos.c:11-90.  lvec holds lm rows of dim
 * entries; row 0 is the input vector.  Returns the iterations done. */
int
orc_lanczos(const orc_space* S, const ocplx* hob, const ocplx* hint, int lm, ocplx* diag,
            ocplx* offd, ocplx* lvec)
{
    const uint32_t nc = S->dim;
    ocplx*         out = malloc(sizeof(ocplx) * nc);
    ocplx*         ortho = malloc(sizeof(ocplx) * lm);
    double         max_beta = 0;
    const double   tol = 1e-14;

    orc_apply_hamiltonian(S, &lvec[0], hob, hint, out);
    diag[0] = orc_carr_dot(nc, &lvec[0], out);
    for (int it = 0; it < lm - 1; it++)
    {
        const ocplx* vi = &lvec[(size_t) it * nc];
        for (uint32_t j = 0; j < nc; j++) out[j] = out[j] - diag[it] * vi[j];
        /* one classical Gram-Schmidt pass against every previous vector */
        for (int k = 0; k <= it; k++) ortho[k] = orc_carr_dot(nc, &lvec[(size_t) k * nc], out);
        for (uint32_t j = 0; j < nc; j++)
            for (int k = 0; k <= it; k++) out[j] -= lvec[(size_t) k * nc + j] * ortho[k];
        offd[it] = orc_carr_mod(nc, out);
        if (max_beta < creal(offd[it])) max_beta = creal(offd[it]);
        if (creal(offd[it]) / max_beta < tol)
        {
            free(ortho);
            free(out);
            return it + 1;
        }
        ocplx  scale = 1.0 / offd[it];
        ocplx* vn = &lvec[(size_t) (it + 1) * nc];
        for (uint32_t j = 0; j < nc; j++) vn[j] = out[j] * scale;
        orc_apply_hamiltonian(S, vn, hob, hint, out);
        for (uint32_t j = 0; j < nc; j++) out[j] = out[j] - offd[it] * vi[j];
        diag[it + 1] = orc_carr_dot(nc, vn, out);
    }
    free(ortho);
    free(out);
    return lm;
}

/* Symmetric tridiagonal eigen-decomposition by implicit-shift QL with
 * Givens rotations.  d[0..n) diagonal (overwritten by eigenvalues, sorted
 * ascending), e[0..n-1) sub-diagonal (destroyed), z (n x n row-major) gets
 * the orthonormal eigenvectors in its columns.  Returns 0, or 1 if an
 * eigenvalue failed to converge in 60 sweeps. */
int
orc_tridiag_eig(int n, double* d, double* e, double* z)
{
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) z[r * n + c] = (r == c) ? 1.0 : 0.0;
    if (n < 1) return 0;
    e[n - 1] = 0.0;
    for (int l = 0; l < n; l++)
    {
        int sweeps = 0, m;
        for (;;)
        {
            /* find the first negligible sub-diagonal at or after l */
            for (m = l; m < n - 1; m++)
            {
                double scale = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= 2.220446049250313e-16 * scale) break;
            }
            if (m == l) break;
            if (++sweeps > 60) return 1;
            /* Wilkinson-type shift from the leading 2x2 block */
            double theta = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double rho = hypot(theta, 1.0);
            double g = d[m] - d[l] + e[l] / (theta + (theta >= 0 ? rho : -rho));
            double sn = 1.0, cs = 1.0, shift = 0.0;
            int    i;
            int    early = 0;
            for (i = m - 1; i >= l; i--)
            {
                double f = sn * e[i];
                double b = cs * e[i];
                double r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0)
                {
                    d[i + 1] -= shift;
                    e[m] = 0.0;
                    early = 1;
                    break;
                }
                sn = f / r;
                cs = g / r;
                g = d[i + 1] - shift;
                r = (d[i] - g) * sn + 2.0 * cs * b;
                shift = sn * r;
                d[i + 1] = g + shift;
                g = cs * r - b;
                for (int row = 0; row < n; row++)
                {
                    double zi1 = z[row * n + i + 1];
                    double zi = z[row * n + i];
                    z[row * n + i + 1] = sn * zi + cs * zi1;
                    z[row * n + i] = cs * zi - sn * zi1;
                }
            }
            if (early) continue;
            d[l] -= shift;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    /* ascending order, as dstev returns */
    for (int a = 0; a < n - 1; a++)
    {
        int best = a;
        for (int b = a + 1; b < n; b++)
            if (d[b] < d[best]) best = b;
        if (best != a)
        {
            double t = d[a];
            d[a] = d[best];
            d[best] = t;
            for (int row = 0; row < n; row++)
            {
                t = z[row * n + a];
                z[row * n + a] = z[row * n + best];
                z[row * n + best] = t;
            }
        }
    }
    return 0;
}

/* Reference coefficients_integration.c:59-160.  coef (dim) is advanced in
 * place by one step exp(-i H prop_dt); lvec_work holds lm*dim entries.
 * Returns the Lanczos iterations done, or -1 on eigensolver failure. */
int
orc_sil_step(const orc_space* S, const ocplx* hob, const ocplx* hint, int lm, double prop_dt_re,
             double prop_dt_im, int imagtime, ocplx* coef, ocplx* lvec)
{
    const uint32_t dim = S->dim;
    const ocplx    prop_dt = prop_dt_re + I * prop_dt_im;
    ocplx          diag[32], offd[32], cl[32], aux[32];
    double         d[32], e[32], z[32 * 32];

    offd[lm - 1] = 0;
    memcpy(lvec, coef, sizeof(ocplx) * dim);
    int done = orc_lanczos(S, hob, hint, lm, diag, offd, lvec);
    for (int k = 0; k < done; k++)
    {
        d[k] = creal(diag[k]);
        e[k] = creal(offd[k]);
    }
    if (orc_tridiag_eig(done, d, e, z) != 0) return -1;
    for (int k = 0; k < done; k++) cl[k] = (k == 0) ? 1.0 : 0.0;
    for (int k = 0; k < done; k++)
    {
        aux[k] = 0;
        for (int j = 0; j < done; j++) aux[k] += z[j * done + k] * cl[j];
        aux[k] = aux[k] * cexp(-I * d[k] * prop_dt);
    }
    for (int k = 0; k < done; k++)
    {
        cl[k] = 0;
        for (int j = 0; j < done; j++) cl[k] += z[k * done + j] * aux[j];
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < dim; i++)
    {
        ocplx acc = 0;
        for (int j = 0; j < done; j++) acc += lvec[(size_t) j * dim + i] * cl[j];
        coef[i] = acc;
    }
    if (imagtime)
    {
        double renorm = 1.0 / orc_carr_mod(dim, coef);
        for (uint32_t i = 0; i < dim; i++) coef[i] = coef[i] * renorm;
    }
    return done;
}

/* Reference multiconfig_lanczos.c:92-172: niter-step Lanczos, then the
 * lowest eigenvector of T mapped back.  Returns the lowest eigenvalue. */
double
orc_lowest_state_lanczos(const orc_space* S, const ocplx* hob, const ocplx* hint, int niter,
                         ocplx* coef)
{
    const uint32_t dim = S->dim;
    ocplx*         lvec = malloc(sizeof(ocplx) * (size_t) niter * dim);
    ocplx*         diag = malloc(sizeof(ocplx) * niter);
    ocplx*         offd = malloc(sizeof(ocplx) * niter);
    double*        d = malloc(sizeof(double) * niter);
    double*        e = malloc(sizeof(double) * niter);
    double*        z = malloc(sizeof(double) * (size_t) niter * niter);
    offd[niter - 1] = 0;
    memcpy(lvec, coef, sizeof(ocplx) * dim);
    int done = orc_lanczos(S, hob, hint, niter, diag, offd, lvec);
    for (int k = 0; k < done; k++)
    {
        d[k] = creal(diag[k]);
        e[k] = creal(offd[k]);
    }
    double lowest = NAN;
    if (orc_tridiag_eig(done, d, e, z) == 0)
    {
        int    arg = 0;
        double sentinel = 1e15;
        for (int k = 0; k < done; k++)
            if (sentinel > d[k])
            {
                sentinel = d[k];
                arg = k;
            }
        for (uint32_t i = 0; i < dim; i++)
        {
            ocplx acc = 0;
            for (int k = 0; k < done; k++) acc += lvec[(size_t) k * dim + i] * z[k * done + arg];
            coef[i] = acc;
        }
        lowest = sentinel;
    }
    free(lvec);
    free(diag);
    free(offd);
    free(d);
    free(e);
    free(z);
    return lowest;
}
