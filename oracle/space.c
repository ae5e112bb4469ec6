/*
 * space.c (oracle) - Fock-space hashing and jump tables, restated from the
 * reference src/configurational/space.c.  TEST INFRASTRUCTURE ONLY: used by
 * tests and the CPU baseline, never by the product path.
 *
 * Ordering convention (reference space.c:138-179): configurations are ranked
 * with the LAST orbital most significant; index 0 holds every particle in
 * orbital 0 and index dim-1 every particle in orbital norb-1.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

/* Reference space.c:84-117: the multiply/divide running product with the
 * 32-bit overflow guard.  Returns 0 on success, -1 where the reference
 * would abort with "INTEGER SIZE ERROR". */
int
orc_space_dimension(uint16_t npar, uint16_t norb, uint32_t* dim)
{
    int64_t prod = 1, div = 2, f;
    if (norb > npar)
    {
        for (f = (int64_t) npar + norb - 1; f > norb - 1; f--)
        {
            if (prod > (int64_t) UINT32_MAX / f) return -1;
            prod *= f;
            if (prod % div == 0 && div <= npar)
            {
                prod /= div;
                div++;
            }
        }
        for (f = div; f <= npar; f++) prod /= f;
        *dim = (uint32_t) prod;
        return 0;
    }
    for (f = (int64_t) npar + norb - 1; f > npar; f--)
    {
        if (prod > (int64_t) UINT32_MAX / f) return -1;
        prod *= f;
        if (prod % div == 0 && div <= norb - 1)
        {
            prod /= div;
            div++;
        }
    }
    for (f = div; f < norb; f++) prod /= f;
    *dim = (uint32_t) prod;
    return 0;
}

static uint32_t
dim_or_die(uint16_t npar, uint16_t norb)
{
    uint32_t d = 0;
    if (orc_space_dimension(npar, norb, &d) != 0) abort();
    return d;
}

/* Reference space.c:119-135: subsdim[n + (N+1)*m] = dim(n particles, m
 * orbitals), with m = 0 -> 0 and m = 1 -> 1. */
void
orc_subspaces_dim(uint16_t npar, uint16_t norb, uint32_t* subsdim)
{
    for (uint32_t n = 0; n <= npar; n++)
    {
        subsdim[n] = 0;
        subsdim[n + (npar + 1)] = 1;
        for (uint32_t m = 2; m <= norb; m++)
            subsdim[n + (npar + 1) * m] = dim_or_die((uint16_t) n, (uint16_t) m);
    }
}

/* Reference space.c:137-156: greedy unranking, highest orbital first. */
void
orc_set_config(uint32_t hash_i, uint16_t npar, uint16_t norb, uint16_t* conf)
{
    int m = norb - 1;
    memset(conf, 0, sizeof(uint16_t) * norb);
    while (hash_i > 0)
    {
        while (hash_i < dim_or_die(npar, (uint16_t) m)) m--;
        hash_i -= dim_or_die(npar, (uint16_t) m);
        conf[m]++;
        npar--;
    }
    if (npar > 0) conf[0] += npar;
}

/* Reference space.c:158-178: rank = sum over particles in orbital m >= 1
 * (taken from the top) of dim(remaining particles, m orbitals). */
uint32_t
orc_hash_index(uint16_t npar, uint16_t norb, const uint32_t* subsdim, const uint16_t* conf)
{
    uint32_t rank = 0, left = npar;
    for (int m = norb - 1; m > 0; m--)
    {
        for (int c = conf[m]; c > 0; c--)
        {
            rank += subsdim[left + (uint32_t) (npar + 1) * m];
            left--;
        }
    }
    return rank;
}

/* The greedy unranking of orc_set_config with dim(n, m) read from the
 * subspace-dimension table instead of recomputed (same values for m >= 1,
 * which is all the greedy loop visits). */
static void
set_config_lut(uint32_t hash_i, uint16_t npar, uint16_t norb, const uint32_t* subsdim,
               uint16_t* conf)
{
    uint32_t stride = (uint32_t) npar + 1, left = npar;
    int      m = norb - 1;
    memset(conf, 0, sizeof(uint16_t) * norb);
    while (hash_i > 0)
    {
        while (hash_i < subsdim[left + stride * m]) m--;
        hash_i -= subsdim[left + stride * m];
        conf[m]++;
        left--;
    }
    if (left > 0) conf[0] += left;
}

/* Reference space.c:180-199 */
void
orc_hash_table(uint16_t npar, uint16_t norb, uint16_t* ht)
{
    uint32_t  dim = dim_or_die(npar, norb);
    uint32_t* subsdim = malloc(sizeof(uint32_t) * (npar + 1) * (norb + 1));
    orc_subspaces_dim(npar, norb, subsdim);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < dim; i++)
        set_config_lut((uint32_t) i, npar, norb, subsdim, &ht[(size_t) norb * i]);
    free(subsdim);
}

/* Reference space.c:201-251: map[i + k*dim + l*norb*dim] = rank(i - e_k + e_l)
 * for n_k >= 1 (k == l included), 0 otherwise. */
void
orc_single_jump_map(uint16_t npar, uint16_t norb, const uint32_t* subsdim, const uint16_t* ht,
                    uint32_t* map)
{
    uint32_t dim = dim_or_die(npar, norb);
    memset(map, 0, sizeof(uint32_t) * (size_t) dim * norb * norb);
#pragma omp parallel
    {
        uint16_t occ[64];
#pragma omp for schedule(static)
        for (int64_t i = 0; i < dim; i++)
        {
            memcpy(occ, &ht[norb * i], sizeof(uint16_t) * norb);
            for (int k = 0; k < norb; k++)
            {
                if (occ[k] < 1) continue;
                for (int l = 0; l < norb; l++)
                {
                    occ[k]--;
                    occ[l]++;
                    map[i + (size_t) k * dim + (size_t) l * norb * dim] =
                        orc_hash_index(npar, norb, subsdim, occ);
                    occ[k]++;
                    occ[l]--;
                }
            }
        }
    }
}

/* Number of (config, k<s both occupied) chunks; reference space.c:254-293. */
uint64_t
orc_count_diffjump(uint32_t dim, uint16_t norb, const uint16_t* ht)
{
    uint64_t n = 0;
    for (uint32_t i = 0; i < dim; i++)
    {
        int occupied = 0;
        for (int k = 0; k < norb; k++) occupied += ht[k + (size_t) i * norb] > 0;
        n += (uint64_t) occupied * (occupied - 1) / 2;
    }
    return n;
}

/* Number of (config, k with n_k >= 2) chunks; reference space.c:360-395. */
uint64_t
orc_count_equaljump(uint32_t dim, uint16_t norb, const uint16_t* ht)
{
    uint64_t n = 0;
    for (uint32_t i = 0; i < dim; i++)
        for (int k = 0; k < norb; k++) n += ht[k + (size_t) i * norb] > 1;
    return n;
}

/* Reference space.c:296-358: chunk order (k, s) lexicographic over occupied
 * pairs k < s; inside a chunk entry (l + q*norb) = rank(i - e_k - e_s + e_l + e_q).
 * strides[i] = chunks before config i, times norb^2. */
void
orc_double_diffjump_map(uint16_t npar, uint16_t norb, const uint32_t* subsdim,
                        const uint16_t* ht, uint32_t* strides, uint32_t* map)
{
    uint32_t dim = dim_or_die(npar, norb);
    uint32_t m2 = (uint32_t) norb * norb;
    uint64_t acc = 0;
    for (uint32_t i = 0; i < dim; i++)
    {
        strides[i] = (uint32_t) (acc * m2);
        int occupied = 0;
        for (int k = 0; k < norb; k++) occupied += ht[k + (size_t) i * norb] > 0;
        acc += (uint64_t) occupied * (occupied - 1) / 2;
    }
#pragma omp parallel
    {
        uint16_t occ[64];
#pragma omp for schedule(static)
        for (int64_t i = 0; i < dim; i++)
        {
            memcpy(occ, &ht[norb * i], sizeof(uint16_t) * norb);
            uint32_t chunk = 0;
            for (int k = 0; k < norb; k++)
            {
                if (occ[k] < 1) continue;
                for (int s = k + 1; s < norb; s++)
                {
                    if (occ[s] < 1) continue;
                    uint32_t base = strides[i] + chunk * m2;
                    for (int l = 0; l < norb; l++)
                    {
                        for (int q = 0; q < norb; q++)
                        {
                            occ[k]--;
                            occ[s]--;
                            occ[l]++;
                            occ[q]++;
                            map[base + l + q * norb] = orc_hash_index(npar, norb, subsdim, occ);
                            occ[k]++;
                            occ[s]++;
                            occ[l]--;
                            occ[q]--;
                        }
                    }
                    chunk++;
                }
            }
        }
    }
}

/* Reference space.c:398-451: one chunk per orbital k with n_k >= 2, in k
 * order; entry (l + q*norb) = rank(i - 2 e_k + e_l + e_q). */
void
orc_double_equaljump_map(uint16_t npar, uint16_t norb, const uint32_t* subsdim,
                         const uint16_t* ht, uint32_t* strides, uint32_t* map)
{
    uint32_t dim = dim_or_die(npar, norb);
    uint32_t m2 = (uint32_t) norb * norb;
    uint64_t acc = 0;
    for (uint32_t i = 0; i < dim; i++)
    {
        strides[i] = (uint32_t) (acc * m2);
        for (int k = 0; k < norb; k++) acc += ht[k + (size_t) i * norb] > 1;
    }
#pragma omp parallel
    {
        uint16_t occ[64];
#pragma omp for schedule(static)
        for (int64_t i = 0; i < dim; i++)
        {
            memcpy(occ, &ht[norb * i], sizeof(uint16_t) * norb);
            uint32_t chunk = 0;
            for (int k = 0; k < norb; k++)
            {
                if (occ[k] < 2) continue;
                uint32_t base = strides[i] + chunk * m2;
                for (int l = 0; l < norb; l++)
                {
                    for (int q = 0; q < norb; q++)
                    {
                        occ[k] -= 2;
                        occ[l]++;
                        occ[q]++;
                        map[base + l + q * norb] = orc_hash_index(npar, norb, subsdim, occ);
                        occ[k] += 2;
                        occ[l]--;
                        occ[q]--;
                    }
                }
                chunk++;
            }
        }
    }
}

/* Reference types_definition.c:110-141 (build order: hash table, map,
 * maptt, mapot). */
orc_space*
orc_space_create(uint16_t npar, uint16_t norb)
{
    uint32_t dim;
    if (npar == 0 || norb == 0 || norb > 30 || npar > 10000) return NULL;
    if (orc_space_dimension(npar, norb, &dim) != 0) return NULL;
    orc_space* S = calloc(1, sizeof(orc_space));
    S->npar = npar;
    S->norb = norb;
    S->dim = dim;
    uint64_t m2 = (uint64_t) norb * norb;
    S->subsdim = malloc(sizeof(uint32_t) * (npar + 1) * (norb + 1));
    orc_subspaces_dim(npar, norb, S->subsdim);
    S->hash_table = malloc(sizeof(uint16_t) * (size_t) dim * norb);
    orc_hash_table(npar, norb, S->hash_table);
    S->strideot = malloc(sizeof(uint32_t) * (size_t) dim);
    S->stridett = malloc(sizeof(uint32_t) * (size_t) dim);
    S->map = malloc(sizeof(uint32_t) * (size_t) dim * m2);
    orc_single_jump_map(npar, norb, S->subsdim, S->hash_table, S->map);
    S->maptt_len = orc_count_diffjump(dim, norb, S->hash_table) * m2;
    S->mapot_len = orc_count_equaljump(dim, norb, S->hash_table) * m2;
    S->maptt = calloc(S->maptt_len + 1, sizeof(uint32_t));
    S->mapot = calloc(S->mapot_len + 1, sizeof(uint32_t));
    orc_double_diffjump_map(npar, norb, S->subsdim, S->hash_table, S->stridett, S->maptt);
    orc_double_equaljump_map(npar, norb, S->subsdim, S->hash_table, S->strideot, S->mapot);
    return S;
}

void
orc_space_destroy(orc_space* S)
{
    if (!S) return;
    free(S->subsdim);
    free(S->hash_table);
    free(S->strideot);
    free(S->stridett);
    free(S->map);
    free(S->maptt);
    free(S->mapot);
    free(S);
}
