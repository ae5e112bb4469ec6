/*
 * configurational.c (oracle) - H.C, one- and two-body density matrices and
 * the total energy, restated from the reference.  TEST INFRASTRUCTURE ONLY.
 *
 *   H = sum_kl h_kl a+_k a_l + 1/2 sum_ksql g[k + sM + qM^2 + lM^3] a+_k a+_s a_q a_l
 *
 * The rule blocks and their order follow src/configurational/hamiltonian.c
 * (one-body :44-55, Rules 1-9 :57-401, hc = w + z/2 at :402), so the FP64
 * sum for each configuration is accumulated in the reference's order.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

void
orc_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
}

int
orc_get_threads(void)
{
    return omp_get_max_threads();
}

/* Chunk counters that locate a double-jump block inside a configuration's
 * mapot / maptt segment (reference hamiltonian.c:92-97 and :236-248). */
static inline uint32_t
ot_chunk(const uint16_t* occ, int k)
{
    uint32_t c = 0;
    for (int h = 0; h < k; h++) c += occ[h] > 1;
    return c;
}

static inline uint32_t
tt_chunk(const uint16_t* occ, int norb, int k, int s)
{
    uint32_t c = 0;
    for (int h = 0; h < k; h++)
        for (int g = h + 1; g < norb; g++) c += (occ[h] > 0 && occ[g] > 0);
    for (int g = k + 1; g < s; g++) c += occ[g] > 0;
    return c;
}

/* Reference hamiltonian.c:6-406 */
void
orc_apply_hamiltonian(const orc_space* S, const ocplx* coef, const ocplx* hob,
                      const ocplx* hint, ocplx* hc)
{
    const uint32_t  dim = S->dim;
    const int       M = S->norb;
    const uint32_t  M2 = (uint32_t) M * M, M3 = M2 * M;
    const uint16_t* ht = S->hash_table;
    const uint32_t *map = S->map, *mapot = S->mapot, *maptt = S->maptt;
    const uint32_t *sot = S->strideot, *stt = S->stridett;

#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < dim; ii++)
    {
        const uint32_t  i = (uint32_t) ii;
        const uint16_t* occ = &ht[(size_t) i * M];
        ocplx           w = 0, z = 0;
        double          bf;
        uint32_t        j;

        /* one-body part */
        for (int k = 0; k < M; k++)
        {
            if (occ[k] < 1) continue;
            w = w + hob[k * M + k] * occ[k] * coef[i];
            for (int l = 0; l < M; l++)
            {
                if (l == k) continue;
                bf = sqrt((double) occ[k] * (occ[l] + 1));
                j = map[i + k * dim + l * M * dim];
                w = w + hob[k * M + l] * bf * coef[j];
            }
        }
        /* Rule 1: a+_k a+_k a_k a_k */
        for (int k = 0; k < M; k++)
        {
            bf = occ[k] * (occ[k] - 1);
            z += hint[k + M * k + M2 * k + M3 * k] * coef[i] * bf;
        }
        /* Rule 2: a+_k a+_s a_k a_s, k < s (4 index permutations) */
        for (int k = 0; k < M; k++)
        {
            if (occ[k] < 1) continue;
            for (int s = k + 1; s < M; s++)
            {
                bf = occ[k] * occ[s];
                z += 4 * hint[k + s * M + k * M2 + s * M3] * bf * coef[i];
            }
        }
        /* Rule 3: a+_k a+_k a_q a_q */
        for (int k = 0; k < M; k++)
        {
            if (occ[k] < 2) continue;
            for (int q = 0; q < M; q++)
            {
                if (q == k) continue;
                bf = sqrt((double) (occ[k] - 1) * occ[k] * (occ[q] + 1) * (occ[q] + 2));
                j = mapot[sot[i] + ot_chunk(occ, k) * M2 + q + q * M];
                z += hint[k + k * M + q * M2 + q * M3] * coef[j] * bf;
            }
        }
        /* Rule 4: a+_k a+_k a_k a_l */
        for (int k = 0; k < M; k++)
        {
            if (occ[k] < 2) continue;
            for (int l = 0; l < M; l++)
            {
                if (l == k) continue;
                bf = (occ[k] - 1) * sqrt((double) occ[k] * (occ[l] + 1));
                j = map[i + k * dim + l * M * dim];
                z += 2 * hint[k + k * M + k * M2 + l * M3] * coef[j] * bf;
            }
        }
        /* Rule 5: a+_k a+_s a_s a_s */
        for (int k = 0; k < M; k++)
        {
            if (occ[k] < 1) continue;
            for (int s = 0; s < M; s++)
            {
                if (s == k || occ[s] < 1) continue;
                bf = occ[s] * sqrt((double) occ[k] * (occ[s] + 1));
                j = map[i + k * dim + s * M * dim];
                z += 2 * hint[k + s * M + s * M2 + s * M3] * coef[j] * bf;
            }
        }
        /* Rule 6.0: a+_k a+_k a_q a_l, k < q < l */
        for (int k = 0; k < M; k++)
        {
            if (occ[k] < 2) continue;
            for (int q = k + 1; q < M; q++)
                for (int l = q + 1; l < M; l++)
                {
                    bf = sqrt((double) occ[k] * (occ[k] - 1) * (occ[q] + 1) * (occ[l] + 1));
                    j = mapot[sot[i] + ot_chunk(occ, k) * M2 + q + l * M];
                    z += 2 * hint[k + k * M + q * M2 + l * M3] * coef[j] * bf;
                }
        }
        /* Rule 6.1: q < k < l */
        for (int q = 0; q < M; q++)
            for (int k = q + 1; k < M; k++)
            {
                if (occ[k] < 2) continue;
                for (int l = k + 1; l < M; l++)
                {
                    bf = sqrt((double) occ[k] * (occ[k] - 1) * (occ[q] + 1) * (occ[l] + 1));
                    j = mapot[sot[i] + ot_chunk(occ, k) * M2 + q + l * M];
                    z += 2 * hint[k + k * M + q * M2 + l * M3] * coef[j] * bf;
                }
            }
        /* Rule 6.2: q < l < k */
        for (int q = 0; q < M; q++)
            for (int l = q + 1; l < M; l++)
                for (int k = l + 1; k < M; k++)
                {
                    if (occ[k] < 2) continue;
                    bf = sqrt((double) occ[k] * (occ[k] - 1) * (occ[q] + 1) * (occ[l] + 1));
                    j = mapot[sot[i] + ot_chunk(occ, k) * M2 + q + l * M];
                    z += 2 * hint[k + k * M + q * M2 + l * M3] * coef[j] * bf;
                }
        /* Rule 7.0: a+_k a+_s a_q a_q, q < k < s */
        for (int q = 0; q < M; q++)
            for (int k = q + 1; k < M; k++)
            {
                if (occ[k] < 1) continue;
                for (int s = k + 1; s < M; s++)
                {
                    if (occ[s] < 1) continue;
                    bf = sqrt((double) occ[k] * occ[s] * (occ[q] + 1) * (occ[q] + 2));
                    j = maptt[stt[i] + tt_chunk(occ, M, k, s) * M2 + q + q * M];
                    z += 2 * hint[k + s * M + q * M2 + q * M3] * coef[j] * bf;
                }
            }
        /* Rule 7.1: k < q < s */
        for (int k = 0; k < M; k++)
        {
            if (occ[k] < 1) continue;
            for (int q = k + 1; q < M; q++)
                for (int s = q + 1; s < M; s++)
                {
                    if (occ[s] < 1) continue;
                    bf = sqrt((double) occ[k] * occ[s] * (occ[q] + 1) * (occ[q] + 2));
                    j = maptt[stt[i] + tt_chunk(occ, M, k, s) * M2 + q + q * M];
                    z += 2 * hint[k + s * M + q * M2 + q * M3] * coef[j] * bf;
                }
        }
        /* Rule 7.2: k < s < q */
        for (int k = 0; k < M; k++)
        {
            if (occ[k] < 1) continue;
            for (int s = k + 1; s < M; s++)
            {
                if (occ[s] < 1) continue;
                for (int q = s + 1; q < M; q++)
                {
                    bf = sqrt((double) occ[k] * occ[s] * (occ[q] + 1) * (occ[q] + 2));
                    j = maptt[stt[i] + tt_chunk(occ, M, k, s) * M2 + q + q * M];
                    z += 2 * hint[k + s * M + q * M2 + q * M3] * coef[j] * bf;
                }
            }
        }
        /* Rule 8: a+_k a+_s a_s a_l, k, s, l distinct */
        for (int s = 0; s < M; s++)
        {
            if (occ[s] < 1) continue;
            for (int k = 0; k < M; k++)
            {
                if (occ[k] < 1 || k == s) continue;
                for (int l = 0; l < M; l++)
                {
                    if (l == k || l == s) continue;
                    bf = occ[s] * sqrt((double) occ[k] * (occ[l] + 1));
                    j = map[i + k * dim + l * M * dim];
                    z += 4 * hint[k + s * M + s * M2 + l * M3] * coef[j] * bf;
                }
            }
        }
        /* Rule 9: a+_k a+_s a_q a_l, all distinct, k < s, q < l */
        for (int k = 0; k < M; k++)
        {
            if (occ[k] < 1) continue;
            for (int s = k + 1; s < M; s++)
            {
                if (occ[s] < 1) continue;
                for (int q = 0; q < M; q++)
                {
                    if (q == s || q == k) continue;
                    for (int l = q + 1; l < M; l++)
                    {
                        if (l == k || l == s) continue;
                        bf = sqrt((double) occ[k] * occ[s] * (occ[q] + 1) * (occ[l] + 1));
                        j = maptt[stt[i] + tt_chunk(occ, M, k, s) * M2 + q + l * M];
                        z += 4 * hint[k + s * M + q * M2 + l * M3] * coef[j] * bf;
                    }
                }
            }
        }
        hc[i] = w + 0.5 * z;
    }
}

/* Reference density_matrices.c:8-52 (rho[k][l] = <a+_k a_l>, Hermitian fill). */
void
orc_onebody_dm(const orc_space* S, const ocplx* coef, ocplx* rho)
{
    const uint32_t  dim = S->dim;
    const int       M = S->norb;
    const uint16_t* ht = S->hash_table;
    const uint32_t* map = S->map;
    for (int k = 0; k < M; k++)
    {
        ocplx acc = 0;
        for (uint32_t i = 0; i < dim; i++)
        {
            double m2 = creal(coef[i]) * creal(coef[i]) + cimag(coef[i]) * cimag(coef[i]);
            acc = acc + m2 * ht[k + i * M];
        }
        rho[k * M + k] = acc;
        for (int l = k + 1; l < M; l++)
        {
            acc = 0;
            for (uint32_t i = 0; i < dim; i++)
            {
                uint32_t nk = ht[k + M * i];
                if (nk < 1) continue;
                uint32_t nl = ht[l + M * i];
                uint32_t j = map[i + k * dim + l * M * dim];
                acc += conj(coef[i]) * coef[j] * sqrt((double) (nl + 1) * nk);
            }
            rho[k * M + l] = acc;
            rho[l * M + k] = conj(acc);
        }
    }
}

/* One reduction over configurations.  Each thread sums a static block of
 * configurations; the per-thread partials are then added in thread order,
 * so the result is reproducible for a fixed thread count. */
#define DM_REDUCE(RESULT, ...)                                                         \
    do                                                                                 \
    {                                                                                  \
        ocplx _parts[256];                                                             \
        int   _nt = 1;                                                                 \
        _Pragma("omp parallel")                                                        \
        {                                                                              \
            ocplx _part = 0;                                                           \
            _Pragma("omp for schedule(static) nowait")                                 \
            for (int64_t _ii = 0; _ii < dim; _ii++)                                     \
            {                                                                          \
                const uint32_t  i = (uint32_t) _ii;                                    \
                const uint16_t* occ = &ht[(size_t) i * M];                             \
                __VA_ARGS__                                                            \
            }                                                                          \
            _parts[omp_get_thread_num()] = _part;                                      \
            _Pragma("omp single") _nt = omp_get_num_threads();                         \
        }                                                                              \
        ocplx _tot = 0;                                                                \
        for (int _t = 0; _t < _nt; _t++) _tot += _parts[_t];                           \
        RESULT = _tot;                                                                 \
    } while (0)

/* Reference density_matrices.c:54-443; element index k + sM + qM^2 + lM^3
 * holds <a+_k a+_s a_q a_l>. */
void
orc_twobody_dm(const orc_space* S, const ocplx* coef, ocplx* rho)
{
    const uint32_t  dim = S->dim;
    const int       M = S->norb;
    const uint32_t  M2 = (uint32_t) M * M, M3 = M2 * M;
    const uint16_t* ht = S->hash_table;
    const uint32_t *map = S->map, *mapot = S->mapot, *maptt = S->maptt;
    const uint32_t *sot = S->strideot, *stt = S->stridett;
    ocplx           r;

    /* Rule 1 */
    for (int k = 0; k < M; k++)
    {
        DM_REDUCE(r, {
            if (occ[k] < 2) continue;
            double m2 = creal(coef[i]) * creal(coef[i]) + cimag(coef[i]) * cimag(coef[i]);
            double bf = occ[k] * (occ[k] - 1);
            _part += m2 * bf;
        });
        rho[k + M * k + M2 * k + M3 * k] = r;
    }
    /* Rule 2 */
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++)
        {
            DM_REDUCE(r, {
                double m2 = creal(coef[i]) * creal(coef[i]) + cimag(coef[i]) * cimag(coef[i]);
                double bf = occ[k] * occ[s];
                _part += m2 * bf;
            });
            rho[k + s * M + k * M2 + s * M3] = r;
            rho[s + k * M + k * M2 + s * M3] = r;
            rho[s + k * M + s * M2 + k * M3] = r;
            rho[k + s * M + s * M2 + k * M3] = r;
        }
    /* Rule 3 */
    for (int k = 0; k < M; k++)
        for (int q = k + 1; q < M; q++)
        {
            DM_REDUCE(r, {
                if (occ[k] < 2) continue;
                uint32_t j = mapot[sot[i] + ot_chunk(occ, k) * M2 + q + q * M];
                double   bf =
                    sqrt((double) (occ[k] - 1) * occ[k] * (occ[q] + 1) * (occ[q] + 2));
                _part += conj(coef[i]) * coef[j] * bf;
            });
            rho[k + k * M + q * M2 + q * M3] = r;
            rho[q + q * M + k * M2 + k * M3] = conj(r);
        }
    /* Rule 4 */
    for (int k = 0; k < M; k++)
        for (int l = k + 1; l < M; l++)
        {
            DM_REDUCE(r, {
                if (occ[k] < 2) continue;
                uint32_t j = map[i + k * dim + l * M * dim];
                double   bf = (occ[k] - 1) * sqrt((double) occ[k] * (occ[l] + 1));
                _part += conj(coef[i]) * coef[j] * bf;
            });
            rho[k + k * M + k * M2 + l * M3] = r;
            rho[k + k * M + l * M2 + k * M3] = r;
            rho[l + k * M + k * M2 + k * M3] = conj(r);
            rho[k + l * M + k * M2 + k * M3] = conj(r);
        }
    /* Rule 5 */
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++)
        {
            DM_REDUCE(r, {
                if (occ[k] < 1 || occ[s] < 1) continue;
                uint32_t j = map[i + k * dim + s * M * dim];
                double   bf = occ[s] * sqrt((double) occ[k] * (occ[s] + 1));
                _part += conj(coef[i]) * coef[j] * bf;
            });
            rho[k + s * M + s * M2 + s * M3] = r;
            rho[s + k * M + s * M2 + s * M3] = r;
            rho[s + s * M + s * M2 + k * M3] = conj(r);
            rho[s + s * M + k * M2 + s * M3] = conj(r);
        }
    /* Rule 6 (6.0 k<q<l, 6.1 q<k<l, 6.2 q<l<k share one body) */
    for (int pass = 0; pass < 3; pass++)
        for (int a = 0; a < M; a++)
            for (int b = a + 1; b < M; b++)
                for (int c = b + 1; c < M; c++)
                {
                    int k, q, l;
                    if (pass == 0) { k = a; q = b; l = c; }
                    else if (pass == 1) { q = a; k = b; l = c; }
                    else { q = a; l = b; k = c; }
                    DM_REDUCE(r, {
                        if (occ[k] < 2) continue;
                        uint32_t j = mapot[sot[i] + ot_chunk(occ, k) * M2 + l + q * M];
                        double   bf = sqrt((double) occ[k] * (occ[k] - 1) * (occ[q] + 1) *
                                           (occ[l] + 1));
                        _part += conj(coef[i]) * coef[j] * bf;
                    });
                    rho[k + k * M + q * M2 + l * M3] = r;
                    rho[k + k * M + l * M2 + q * M3] = r;
                    rho[l + q * M + k * M2 + k * M3] = conj(r);
                    rho[q + l * M + k * M2 + k * M3] = conj(r);
                }
    /* Rule 7 (7.0 s<k<l, 7.1 k<s<l, 7.2 k<l<s share one body) */
    for (int pass = 0; pass < 3; pass++)
        for (int a = 0; a < M; a++)
            for (int b = a + 1; b < M; b++)
                for (int c = b + 1; c < M; c++)
                {
                    int k, s, l;
                    if (pass == 0) { s = a; k = b; l = c; }
                    else if (pass == 1) { k = a; s = b; l = c; }
                    else { k = a; l = b; s = c; }
                    DM_REDUCE(r, {
                        if (occ[k] < 1 || occ[s] < 1) continue;
                        uint32_t j = map[i + k * dim + l * M * dim];
                        double   bf = occ[s] * sqrt((double) occ[k] * (occ[l] + 1));
                        _part += conj(coef[i]) * coef[j] * bf;
                    });
                    rho[k + s * M + s * M2 + l * M3] = r;
                    rho[s + k * M + s * M2 + l * M3] = r;
                    rho[s + k * M + l * M2 + s * M3] = r;
                    rho[k + s * M + l * M2 + s * M3] = r;
                    rho[l + s * M + s * M2 + k * M3] = conj(r);
                    rho[s + l * M + s * M2 + k * M3] = conj(r);
                    rho[s + l * M + k * M2 + s * M3] = conj(r);
                    rho[l + s * M + k * M2 + s * M3] = conj(r);
                }
    /* Rule 8: k < s, q < l, {q, l} disjoint from {k, s} */
    for (int k = 0; k < M; k++)
        for (int s = k + 1; s < M; s++)
            for (int q = 0; q < M; q++)
            {
                if (q == s || q == k) continue;
                for (int l = q + 1; l < M; l++)
                {
                    if (l == k || l == s) continue;
                    DM_REDUCE(r, {
                        if (occ[k] < 1 || occ[s] < 1) continue;
                        uint32_t j = maptt[stt[i] + tt_chunk(occ, M, k, s) * M2 + q + l * M];
                        double   bf =
                            sqrt((double) occ[k] * occ[s] * (occ[q] + 1) * (occ[l] + 1));
                        _part += conj(coef[i]) * coef[j] * bf;
                    });
                    rho[k + s * M + q * M2 + l * M3] = r;
                    rho[s + k * M + q * M2 + l * M3] = r;
                    rho[k + s * M + l * M2 + q * M3] = r;
                    rho[s + k * M + l * M2 + q * M3] = r;
                }
            }
}

/* Reference integrator_monitor.c:55-88:
 * E = sum_kl rho1[k][l] h[k][l] + 1/2 sum rho2[k,l,s,q] g[k,l,q,s]. */
void
orc_total_energy(int M, const ocplx* rho1, const ocplx* rho2, const ocplx* hob,
                 const ocplx* hint, ocplx* energy)
{
    ocplx z = 0, w = 0;
    int   M2 = M * M, M3 = M2 * M;
    for (int k = 0; k < M; k++)
        for (int l = 0; l < M; l++)
        {
            w = w + rho1[k * M + l] * hob[k * M + l];
            for (int s = 0; s < M; s++)
                for (int q = 0; q < M; q++)
                    z = z + rho2[k + l * M + s * M2 + q * M3] * hint[k + l * M + q * M2 + s * M3];
        }
    *energy = w + z / 2;
}
