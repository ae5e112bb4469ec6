/*
 * mctdhb_types.h - data structures of the MCTDHB configurational core.
 *
 * Same struct names, field names, field order and enum values as the
 * reference include/mctdhb_types.h:12-245, so code written against the
 * reference compiles against this header.  Two substitutions, both for
 * types the reference took from Intel MKL (mctdhb_types.h:17-18):
 *   - DFTI_DESCRIPTOR_HANDLE fft_desc  -> void* fft_desc (opaque plan handle)
 *   - MKL_Complex16                    -> mctdhb_complex16 {double real, imag}
 * Structs that own device mirrors carry one extra trailing field `dev`
 * (opaque), appended so every reference field keeps its offset.
 */
#ifndef MCTDHB_TYPES_H
#define MCTDHB_TYPES_H

#include <inttypes.h>
#include <stdint.h>

#ifdef __cplusplus
typedef double _Complex dcomplex;
extern "C" {
#else
#include <complex.h>
typedef double complex dcomplex;
#endif

#define PI 3.141592653589793
#define STR_BUFF_SIZE 128
#define MAX_GRID_SIZE 2049
#define MIN_GRID_SIZE 50
#define MAX_ORBITALS 30
#define MAX_PARTICLES 10000
#define MAX_LANCZOS_ITER 15
#define MIN_LANCZOS_ITER 2

typedef struct
{
    double real, imag;
} mctdhb_complex16;

typedef int*              Iarray;
typedef int**             Imatrix;
typedef double*           Rarray;
typedef double**          Rmatrix;
typedef dcomplex*         Carray;
typedef dcomplex**        Cmatrix;
typedef mctdhb_complex16* MKLCarray;

typedef enum
{
    FALSE = 0,
    TRUE = 1
} Bool;

typedef enum
{
    IMAGTIME = 10,
    REALTIME = 11
} IntegratorType;

typedef enum
{
    LANCZOS = 20,
    RUNGEKUTTA = 21
} CoefIntegrator;

typedef enum
{
    FULLSTEP_RUNGEKUTTA = 30,
    SPLITSTEP = 31,
} OrbIntegrator;

typedef enum
{
    DVR = 300,
    SPECTRAL = 301,
    FINITEDIFF = 302
} OrbDerivative;

typedef enum
{
    RK2 = 2,
    RK4 = 4,
    RK5 = 5
} RungeKuttaOrder;

typedef enum
{
    ZERO_BOUNDS = 1000,
    PERIODIC_BOUNDS = 1001
} BoundaryCondition;

#define DEFAULT_INTEGRATION_TYPE IMAGTIME
#define DEFAULT_COEF_INTEGRATOR LANCZOS
#define DEFAULT_ORB_INTEGRATOR FULLSTEP_RUNGEKUTTA
#define DEFAULT_ORB_DERIVATIVE DVR
#define DEFAULT_RUNGEKUTTA_ORDER RK5
#define DEFAULT_BOUNDARY_CONDITION PERIODIC_BOUNDS
#define DEFAULT_LANCZOS_ITER 5

typedef void (*single_particle_pot)(
    double t, uint16_t npts, double* pts, void* params, double* grid_pot);

typedef double (*time_dependent_parameter)(double t, void* params);

typedef struct
{
    char                     eq_name[STR_BUFF_SIZE];
    uint16_t                 norb, grid_size;
    BoundaryCondition        bounds;
    double                   t, xi, xf, dx, tstep, tend, g, d2coef;
    dcomplex                 d1coef, prop_dt, time_fac;
    Rarray                   grid_pts;
    Rarray                   pot_grid;
    void*                    pot_extra_args;
    void*                    inter_extra_args;
    single_particle_pot      pot_func;
    time_dependent_parameter inter_param;
} _OrbitalEquation;

typedef _OrbitalEquation* OrbitalEquation;

typedef struct
{
    uint32_t *map, *mapot, *maptt, *strideot, *stridett;
} _OperatorMappings;

typedef _OperatorMappings* OperatorMappings;

typedef struct
{
    uint16_t         npar, norb;
    uint32_t         dim;
    uint16_t*        hash_table;
    uint32_t*        subspaces_dim;
    OperatorMappings op_maps;
    void*            dev; /* appended: device-resident space (mctdhb_space*) */
} _MultiConfiguration;

typedef _MultiConfiguration* MultiConfiguration;

typedef struct
{
    uint16_t npar, norb, grid_size;
    uint32_t space_dim;
    Carray   coef, hint, tb_denmat;
    Cmatrix  hob, orbitals, ob_denmat, inv_ob_denmat;
} _ManyBodyState;

typedef _ManyBodyState* ManyBodyState;

typedef struct
{
    uint16_t iter;
    uint32_t space_dim;
    Cmatrix  lanczos_vectors;
    Carray   decomp_diag, decomp_offd, coef_lspace, transform, hc, reortho;
    Rarray   lapack_diag, lapack_offd, lapack_eigvec;
} _WorkspaceLanczos;

typedef _WorkspaceLanczos* WorkspaceLanczos;

typedef struct
{
    uint16_t      norb, grid_size;
    OrbDerivative orb_der_method;
    Bool          impr_ortho;
    Carray        dvr_mat;
    Carray        cn_upper, cn_lower, cn_mid, cn_rhs;
    Cmatrix       orb_work1, orb_work2;
    void*         extern_work;
    void*         fft_desc; /* opaque plan handle (MKL DFTI descriptor in the reference) */
    Rarray        fft_freq;
    Carray        fft_hder_exp;
} _OrbitalWorkspace;

typedef _OrbitalWorkspace* OrbitalWorkspace;

typedef struct
{
    WorkspaceLanczos lan_work;
    void*            extern_work;
} _CoefWorkspace;

typedef _CoefWorkspace* CoefWorkspace;

typedef struct
{
    IntegratorType     integ_type;
    CoefIntegrator     coef_integ_method;
    OrbIntegrator      orb_integ_method;
    OrbDerivative      orb_der_method;
    RungeKuttaOrder    rk_order;
    MultiConfiguration multiconfig_space;
    OrbitalEquation    orb_eq;
    CoefWorkspace      coef_workspace;
    OrbitalWorkspace   orb_workspace;
    ManyBodyState      state;
} _MCTDHBDataStruct;

typedef _MCTDHBDataStruct* MCTDHBDataStruct;

#ifdef __cplusplus
}
#endif

#endif
