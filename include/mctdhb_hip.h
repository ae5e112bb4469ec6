/*
 * mctdhb_hip.h - C ABI of libmctdhb_hip.so, the MI355X (gfx950) MCTDHB
 * configurational core.
 *
 * Two layers, both plain C (no device or torch types in any signature):
 *
 * 1. Reference-named drop-in entry points.  Same names, argument order and
 *    meaning as the reference C library; host arrays in, host arrays out,
 *    the work done on the GPU.  They keep the reference's fail-stop error
 *    behaviour (message on stdout, exit(EXIT_FAILURE)).  Each declaration
 *    cites the reference interface it replaces (file:line).
 *
 * 2. Device-resident entry points (mctdhb_*).  Inputs and outputs are device
 *    pointers (complex128 = interleaved doubles, complex64 = interleaved
 *    floats), `stream` is a hipStream_t passed as void* (NULL = default
 *    stream), calls are asynchronous on that stream, and every function
 *    returns 0 or an error code with the message in mctdhb_last_error().
 */
#ifndef MCTDHB_HIP_H
#define MCTDHB_HIP_H

#include <stddef.h>

#include "mctdhb_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ================================================ 1. reference drop-in */

/* include/configurational/space.h:14-221, src/configurational/space.c */
void      assert_space_parameters(uint16_t npar, uint16_t norb);                           /* space.c:61-82 */
uint32_t  space_dimension(uint16_t npar, uint16_t norb);                                   /* space.c:84-117 */
uint32_t* get_subspaces_dim(uint16_t npar, uint16_t norb);                                 /* space.c:119-135 */
void      set_config(uint32_t hash_i, uint16_t npar, uint16_t norb, uint16_t* conf);       /* space.c:137-156 */
uint32_t  hash_index(uint16_t npar, uint16_t norb, uint32_t* subspaces_dim, uint16_t* conf); /* space.c:158-178 */
uint16_t* get_hash_table(uint16_t npar, uint16_t norb);                                    /* space.c:180-199 */
uint32_t* get_single_jump_map(uint16_t npar, uint16_t norb, uint32_t* subsdim,
                              uint16_t* hash_table);                                       /* space.c:201-251 */
uint32_t* get_double_diffjump_map(uint16_t npar, uint16_t norb, uint32_t* subsdim,
                                  uint16_t* hash_table, uint32_t* strides);                /* space.c:296-358 */
uint32_t* get_double_equaljump_map(uint16_t npar, uint16_t norb, uint32_t* subsdim,
                                   uint16_t* hash_table, uint32_t* strides);               /* space.c:398-451 */
uint32_t  jump1_index(uint16_t norb, uint32_t dim, uint32_t inp_index, uint32_t* map1jump,
                      uint16_t from_orb, uint16_t to_orb);                                 /* space.c:453-463 */
uint32_t  jump2same_index(uint16_t norb, uint16_t* conf, uint32_t conf_stride, uint32_t* mapot,
                          uint16_t from_orb, uint16_t to_orb1, uint16_t to_orb2);          /* space.c:465-486 */
uint32_t  jump2diff_index(uint16_t norb, uint16_t* conf, uint32_t conf_stride, uint32_t* maptt,
                          uint16_t from_orb1, uint16_t from_orb2, uint16_t to_orb1,
                          uint16_t to_orb2);                                               /* space.c:488-524 */

/* include/assistant/types_definition.h:22-113, src/assistant/types_definition.c */
MultiConfiguration get_multiconf_struct(uint16_t npar, uint16_t norb);          /* types_definition.c:110-141 */
void               destroy_multiconf_struct(MultiConfiguration multiconf);       /* types_definition.c:347-360 */
ManyBodyState      get_manybody_state(uint16_t npar, uint16_t norb, uint16_t grid_size); /* :143-160 */
void               destroy_manybody_sate(ManyBodyState state);                   /* :361-373 (reference spelling) */
WorkspaceLanczos   get_lanczos_workspace(uint16_t iter, uint32_t space_dim);     /* :162-179 */
void               destroy_lanczos_workspace(WorkspaceLanczos lan_work);         /* :374-389 */
CoefWorkspace      get_coef_workspace(CoefIntegrator coef_integ_method, uint32_t space_dim,
                                      uint16_t lan_it);                          /* :181-198 */
void               destroy_coef_workspace(CoefWorkspace coef_work);              /* :408-421 */
OrbitalEquation    get_orbital_equation(char eq_name[], uint16_t norb, uint16_t grid_size,
                                        BoundaryCondition bounds, IntegratorType integ_type, double xi,
                                        double xf, double tstep, double tend, double d2coef, dcomplex d1coef,
                                        void* pot_extra_args, void* inter_extra_args,
                                        single_particle_pot pot_func,
                                        time_dependent_parameter inter_param); /* :30-108 */
void               destroy_orbital_equation(OrbitalEquation orbeq);             /* :339-345 */
/* fft_desc holds the device workspace; dvr_mat is NULL (DVR not provided) */
OrbitalWorkspace   get_orbital_workspace(OrbitalEquation eq_desc, OrbDerivative der_method); /* :200-244 */
void               destroy_orbital_workspace(OrbitalWorkspace orb_work);        /* :391-406 */
MCTDHBDataStruct   get_mctdhb_struct(IntegratorType integ_type, CoefIntegrator coef_integ_method,
                                     OrbIntegrator orb_integ_method, OrbDerivative orb_der_method,
                                     RungeKuttaOrder rk_order, uint16_t lanczos_iter, uint16_t npar,
                                     uint16_t norb, char eq_name[], BoundaryCondition bounds, double xi,
                                     double xf, uint16_t grid_size, double tstep, double tend, double d2coef,
                                     dcomplex d1coef, void* pot_extra_args, void* inter_extra_args,
                                     single_particle_pot pot_func,
                                     time_dependent_parameter inter_param);  /* :246-314 */
void               set_mctdhb_integrator(IntegratorType integ_type, CoefIntegrator coef_integ_method,
                                         OrbIntegrator orb_integ_method, OrbDerivative orb_der_method,
                                         RungeKuttaOrder rk_order, BoundaryCondition bounds,
                                         uint16_t lanczos_iter, MCTDHBDataStruct mctdhb); /* :316-331 */
void               destroy_mctdhb_struct(MCTDHBDataStruct mctdhb);              /* :422-432 */

/* include/configurational/hamiltonian.h:7-13 */
void apply_hamiltonian(MultiConfiguration multiconf, Carray coef, Cmatrix hob, Carray hint,
                       Carray hc);                                               /* hamiltonian.c:6-406 */

/* include/configurational/density_matrices.h:7-12 */
void set_onebody_dm(MultiConfiguration multiconf, Carray coef, Cmatrix rho);    /* density_matrices.c:8-52 */
void set_twobody_dm(MultiConfiguration multiconf, Carray coef, Carray rho);     /* density_matrices.c:54-443 */

/* include/linalg/multiconfig_lanczos.h:31-39 */
int lanczos(MultiConfiguration multiconf, Cmatrix hob_mat, Carray hint_mat, int iter, Carray diag,
            Carray offd, Cmatrix lvec);                                          /* multiconfig_lanczos.c:11-90 */

/* include/linalg/multiconfig_lanczos.h:57-63 */
double lowest_state_lanczos(uint16_t niter, MultiConfiguration multiconf, Cmatrix Ho, Carray Hint,
                            Carray C);                                           /* multiconfig_lanczos.c:92-172 */

/* include/integrator/split_linear_orbitals.h:59-62 */
void advance_linear_crank_nicolson(OrbitalEquation eq_desc, OrbitalWorkspace orb_work,
                                   Cmatrix orb);                                 /* split_linear_orbitals.c:335-373 */
void advance_linear_fft(OrbitalWorkspace orb_work, Cmatrix orb);                 /* split_linear_orbitals.c:375-402 */

/* include/integrator/synchronize.h:36-57 */
void   set_regulatization_factor(double reg_fac);                                /* synchronize.c:11-26 (reference spelling) */
double get_current_regularization_factor(void);                                  /* synchronize.c:28-32 */
void   sync_orbital_matrices(OrbitalEquation eq_desc, ManyBodyState psi);        /* synchronize.c:34-39 */
void   sync_density_matrices(MultiConfiguration mc_space, ManyBodyState psi);    /* synchronize.c:41-56 */
void   sync_equation_params(OrbitalEquation eq_desc);                            /* synchronize.c:58-67 */

/* include/integrator/orbital_integration.h:7-37.  rk_order RK2 / RK4 / RK5;
 * derivatives FINITEDIFF, SPECTRAL or DVR (full step).  orb_next is
 * scratch.  The robust projector's overlap matrix has the reference's
 * unset [0][1] / [1][0] entries (quirk Q2) as zero. */
void set_periodic_bounds(uint16_t norb, uint16_t grid_size, Cmatrix orb);       /* orbital_integration.c:13-20 */
void robust_multiorb_projector(OrbitalEquation eq_desc, uint16_t norb, Cmatrix orb, Cmatrix hact_orb,
                               Cmatrix project);                                 /* orbital_integration.c:22-93 */
void simple_multiorb_projector(OrbitalEquation eq_desc, uint16_t norb, Cmatrix orb, Cmatrix hact_orb,
                               Cmatrix project);                                 /* orbital_integration.c:94-129 */
void propagate_fullstep_orb_rk(MCTDHBDataStruct mctdhb, Carray orb_next);        /* orbital_integration.c:283-348 */
void propagate_splitstep_orb(MCTDHBDataStruct mctdhb, Carray orb_next);          /* orbital_integration.c:351-434 */

/* include/function_tools/orbital_matrices.h:19-32.  set_overlap_matrix
 * writes exactly the entries the reference writes (its j loop starts at 2). */
void set_overlap_matrix(uint16_t norb, uint16_t grid_size, double dx, Cmatrix orb,
                        Cmatrix overlap);                                        /* orbital_matrices.c:7-20 */
void set_orbital_hob(OrbitalEquation eq_desc, uint16_t norb, Cmatrix orb, Cmatrix hob); /* orbital_matrices.c:22-70 */
void set_orbital_hint(OrbitalEquation eq_desc, uint16_t norb, Cmatrix orb, Carray hint); /* orbital_matrices.c:72-287 */

/* include/integrator/driver.h:66-71.  Coefficient integrator LANCZOS only;
 * orbital integrator SPLITSTEP or FULLSTEP_RUNGEKUTTA; orb_works and
 * coef_works are scratch. */
void mctdhb_propagate_step(MCTDHBDataStruct mctdhb, Carray orb_works, Carray coef_works,
                           uint32_t curr_step);                                  /* driver.c:158-202 */

/* include/integrator/coefficients_integration.h:25-26 */
void propagate_coef_sil(MCTDHBDataStruct mctdhb, Carray cnext);                  /* coefficients_integration.c:59-160 */

/* include/assistant/integrator_monitor.h */
double   eig_residual(MultiConfiguration multiconf, Carray c, Cmatrix hob, Carray hint,
                      double e);                                                 /* integrator_monitor.c:9-28 */
dcomplex total_energy(ManyBodyState psi);                                        /* integrator_monitor.c:55-88 */
double   overlap_residual(uint16_t norb, uint16_t grid_size, double dx, Cmatrix orb); /* integrator_monitor.c:30-42 */
double   avg_orbitals_norm(uint16_t norb, uint16_t grid_size, double dx, Cmatrix orb); /* integrator_monitor.c:44-53 */

/* include/integrator/driver.h: the job loop and its settings.  The loop
 * keeps the state on the GPU (imaginary time replays one captured step
 * graph per step) and copies it to the host structs at monitor, record and
 * end points.  Printing is a compact monitor line, not the reference's
 * column layout. */
#define REALTIME_OVERLAP_RESIDUE_TOL 0.001                                       /* driver.h:7 */
void set_final_diagonalization(Bool shall_diag);                                 /* driver.c:97-101 */
void set_autoconvergence_check(Bool must_check);                                 /* driver.c:103-107 */
void set_energy_convergence_digits(uint8_t edig);                                /* driver.c:109-128 */
void set_energy_convergence_eig_residual(double eig_res);                        /* driver.c:130-141 */
void set_overlap_residue_threshold(double over_res);                             /* driver.c:143-157 */
void integration_driver(MCTDHBDataStruct mctdhb, uint16_t rec_nsteps, char prefix[], double tend,
                        uint8_t monitor_rate);                                   /* driver.c:203-333 */

/* include/function_tools/builtin_potential.h, builtin_time_parameter.h */
single_particle_pot get_builtin_pot(char pot_name[]);                            /* builtin_potential.c:215-260 */
void potfunc_harmonic(double t, uint16_t npts, Rarray x, void* params, Rarray V); /* builtin_potential.c:36-50 */
void potfunc_doublewell(double t, uint16_t npts, Rarray x, void* params, Rarray V); /* :51-69 */
void potfunc_harmonicgauss(double t, uint16_t npts, Rarray x, void* params, Rarray V); /* :70-94 */
void potfunc_barrier(double t, uint16_t npts, Rarray x, void* params, Rarray V);  /* :95-121 */
void potfunc_time_trapezoid_barrier(double t, uint16_t npts, Rarray x, void* params, Rarray V); /* :123-143 */
void potfunc_opticallattice(double t, uint16_t npts, Rarray x, void* params, Rarray V); /* :145-161 */
void potfunc_time_trapezoid_opticallattice(double t, uint16_t npts, Rarray x, void* params,
                                           Rarray V);                            /* :163-180 */
void potfunc_step(double t, uint16_t npts, Rarray x, void* params, Rarray V);    /* :182-196 */
void potfunc_square_well(double t, uint16_t npts, Rarray x, void* params, Rarray V); /* :198-213 */
double timeparam_linear_ramp(double t, void* params);                           /* builtin_time_parameter.c:6-15 */
time_dependent_parameter get_builtin_param_func(char funcname[]);               /* builtin_time_parameter.c:17-30 */

/* include/assistant/dataio.h: text formats.  Complex numbers are written
 * as " (%.16E%+.16Ej)" tokens (numpy.loadtxt reads them as complex128) and
 * read with the scanf formats below; '#' starts a comment line. */
#define MIN_PARAMS_INLINE 21                                                     /* dataio.h:32 */
#define PARAMS_FNAME_SUFFIX "_mctdhb_parameters.dat"                             /* dataio.h:34 */
typedef enum { COMMON_INP, MULTIPLE_INP, LAST_JOB_OUT } JobsInputHandle;         /* dataio.h:70-75 */
typedef enum {
    ORBITALS_REC,
    COEFFICIENTS_REC,
    ONE_BODY_MATRIX_REC,
    TWO_BODY_MATRIX_REC,
    PARAMETERS_REC,
    ONE_BODY_POTENTIAL_REC
} RecordDataType;                                                                /* dataio.h:78-86 */
extern char    orb_cplx_read_fmt[STR_BUFF_SIZE];                                 /* dataio.c:14 */
extern char    coef_cplx_read_fmt[STR_BUFF_SIZE];                                /* dataio.c:15 */
extern char    inp_dirname[STR_BUFF_SIZE];                                       /* dataio.c:16 */
extern char    out_dirname[STR_BUFF_SIZE];                                       /* dataio.c:17 */
extern char    integrator_desc_fname[STR_BUFF_SIZE];                             /* dataio.c:18 */
extern uint8_t monitor_energy_digits;                                            /* dataio.c:20 */
extern Bool    new_empty_append_files;                                           /* dataio.c:29 */
void toggle_new_append_files(void);                                              /* dataio.c:41-50 */
void set_orbitals_from_file(char fname[], ManyBodyState psi);                    /* dataio.c:52-66 */
void set_coef_from_file(char fname[], ManyBodyState psi);                        /* dataio.c:68-72 */
MCTDHBDataStruct get_mctdhb_from_files(char par_fname[], char integ_fname[], uint32_t line,
                                       single_particle_pot custom_pot_fun,
                                       time_dependent_parameter custom_inter_fun, void* custom_pot_params,
                                       void* custom_inter_params);               /* dataio.c:74-293 */
MCTDHBDataStruct full_setup_mctdhb_current_dir(char fprefix[], uint32_t job_num, JobsInputHandle which_inp,
                                               single_particle_pot custom_pot_fun,
                                               time_dependent_parameter custom_inter_fun,
                                               void* custom_pot_params,
                                               void* custom_inter_params);       /* dataio.c:295-378 */
uint32_t auto_number_of_jobs(char prefix[]);                                     /* dataio.c:519-527 */
void set_output_fname(char prefix[], RecordDataType id, char* fname);            /* dataio.c:529-558 */
void append_processed_state(char prefix[], ManyBodyState psi);                   /* dataio.c:705-746 */
void record_raw_state(char prefix[], ManyBodyState psi);                         /* dataio.c:747-764 */
void append_timestep_potential(char prefix[], OrbitalEquation eq_desc);          /* dataio.c:765-781 */
void record_time_interaction(char prefix[], OrbitalEquation eq_desc, uint16_t recn); /* dataio.c:782-813 */
void record_time_array(char prefix[], double tend, double tstep, uint16_t recn); /* dataio.c:814-835 */
void append_mctdhb_parameters(char prefix[], MCTDHBDataStruct mctdhb);           /* dataio.c:836-866 */

/* ================================================= 2. device-resident */

typedef struct mctdhb_space mctdhb_space; /* configurational space on one GPU */
typedef struct mctdhb_sil   mctdhb_sil;   /* Krylov workspace of the SIL step */

/* status codes */
#define MCTDHB_OK 0
#define MCTDHB_EINVAL 1
#define MCTDHB_EOVERFLOW 2
#define MCTDHB_EDEVICE 3
#define MCTDHB_ENOMEM 4
#define MCTDHB_ENOTBUILT 5
#define MCTDHB_ENUMERIC 6

/* space flags */
#define MCTDHB_SPACE_TABLES 1 /* also build the reference-layout jump tables */

/* arithmetic forms */
#define MCTDHB_FP64 0  /* complex128 throughout (the parity form) */
#define MCTDHB_MIXED 1 /* complex64 H.C and density sums (Kahan), FP64 Lanczos scalars */

/* index schemes */
#define MCTDHB_INDEX_TABLES 0 /* read the reference jump tables */
#define MCTDHB_INDEX_FLY 1    /* closed-form rank deltas, no table traffic */

const char* mctdhb_last_error(void);
int         mctdhb_device_count(int* count);

/* Build the Fock-space hash table (and with MCTDHB_SPACE_TABLES the
 * map/mapot/maptt/strides of get_multiconf_struct) on the current device. */
int mctdhb_space_create(uint16_t npar, uint16_t norb, int flags, mctdhb_space** out);
int mctdhb_space_destroy(mctdhb_space* space);
int mctdhb_space_info(const mctdhb_space* space, uint32_t* dim, uint64_t* mapot_len,
                      uint64_t* maptt_len, int* has_tables);
/* Device pointers of the tables (NULL where not built). */
int mctdhb_space_tables(const mctdhb_space* space, const uint16_t** hash_table,
                        const uint32_t** subspaces_dim, const uint32_t** map, const uint32_t** mapot,
                        const uint32_t** maptt, const uint32_t** strideot, const uint32_t** stridett);

/* hc = H coef.  hob: M x M row-major complex128; hint: M^4 complex128 with
 * hint[k + sM + qM^2 + lM^3].  coef/hc complex128 (FP64) or complex64 (MIXED). */
int mctdhb_apply_hamiltonian_dev(const mctdhb_space* space, const void* coef, const void* hob,
                                 const void* hint, void* hc, int prec, int index, void* stream);

/* mctdhb_apply_hamiltonian_dev that allocates no device memory.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The (hob, hint)-derived coefficient block goes to a fixed
 * 128 KiB device array of the library, one per GPU (enough for every
 * norb <= 7; norb 1 and > 7 read hob and hint directly); the space's device
 * must be current.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (different streams, or threads) race; serialise them.
 * The mixed form runs one configuration per lane unless the paired tiling
 * was already chosen on this space by an allocating call.  Same arguments,
 * results and errors as mctdhb_apply_hamiltonian_dev. */
int mctdhb_apply_hamiltonian_noalloc_dev(const mctdhb_space* space, const void* coef, const void* hob,
                                         const void* hint, void* hc, int prec, int index, void* stream);

/* The (hob, hint)-derived coefficient block of the H.C kernels, built once
 * and reused across applications with the same one- and two-body matrices.
 * cf must hold mctdhb_hc_coef_bytes(space, prec) bytes of device memory. */
size_t mctdhb_hc_coef_bytes(const mctdhb_space* space, int prec);
int    mctdhb_hc_coefficients_dev(const mctdhb_space* space, const void* hob, const void* hint, void* cf,
                                  int prec, void* stream);
int    mctdhb_apply_hamiltonian_cf_dev(const mctdhb_space* space, const void* coef, const void* cf, void* hc,
                                       int prec, int index, void* stream);

/* Tiling of the mixed H.C kernel on this space: 0 one configuration per
 * lane, 1 two configurations of one run per lane (one 16-byte gather serves
 * both), 2 one configuration per lane with the coefficients each block
 * reaches staged in LDS, one window per far group of jump types
 * (2 <= norb <= 5, N <= 185; the plain product only, the SIL step's
 * Lanczos forms keep tiling 0 under it).  By default the possible ones are
 * timed at the first mixed call on a space of a new (N, M) and the fastest
 * kept for every space of that shape in the process (so a solve runs the
 * same kernels alone and among others).
 * set: -2 query only, -1 time again, 0, 1 or 2 force; *tiling
 * (may be NULL) receives the current choice (-1: not yet timed).  Callers
 * that change it must not have a mixed SIL workspace open on the space. */
int mctdhb_hc_tiling(mctdhb_space* space, int set, int* tiling);

/* rho1 (M x M, <a+_k a_l>) and rho2 (M^4, <a+_k a+_s a_q a_l>), complex128;
 * rho2 may be NULL.  coef complex128 (FP64) or complex64 (MIXED). */
int mctdhb_density_dev(mctdhb_space* space, const void* coef, void* rho1, void* rho2, int prec,
                       int index, void* stream);

/* SIL workspace for lanczos_iter Krylov vectors (2..64; up to 15 the
 * streaming kernels keep all coefficients in registers). */
int mctdhb_sil_create(const mctdhb_space* space, int lanczos_iter, int prec, int index,
                      mctdhb_sil** out);
int mctdhb_sil_destroy(mctdhb_sil* sil);
/* FP64 steps of spaces up to 8192 configurations, 2 <= norb <= 7 and at most
 * 15 Krylov vectors run as one workgroup (the whole recursion in one
 * kernel); on = 0 keeps them on the multi-kernel path.  Default on. */
int mctdhb_sil_set_small(mctdhb_sil* sil, int on);
size_t mctdhb_sil_device_bytes(const mctdhb_space* space, int lanczos_iter, int prec);
/* coef (complex128, device) <- exp(-i H prop_dt) coef; renormalised when
 * imagtime != 0 (prop_dt = -i dt in imaginary time, dt in real time). */
int mctdhb_sil_step_dev(mctdhb_sil* sil, void* coef, const void* hob, const void* hint,
                        double prop_dt_re, double prop_dt_im, int imagtime, void* stream);
/* The Lanczos recursion alone; lvec (lanczos_iter x dim complex128) gets the
 * normalised Krylov vectors, lvec row 0 = coef. */
int mctdhb_lanczos_dev(mctdhb_sil* sil, const void* coef, const void* hob, const void* hint,
                       void* lvec, void* stream);
/* coef <- Ritz vector of the lowest Ritz value of the lanczos_iter-step
 * recursion started from coef (not renormalised); *energy (host, may be
 * NULL; synchronises) <- that Ritz value.  multiconfig_lanczos.c:93-172 */
int mctdhb_lowest_state_dev(mctdhb_sil* sil, void* coef, const void* hob, const void* hint, double* energy,
                            void* stream);
/* *result (host; synchronises) <- max_i |e c_i - (H c)_i| / N for coef
 * (complex128, device) with the FP64 H.C.  integrator_monitor.c:9-28 */
int mctdhb_eig_residual_dev(mctdhb_space* space, const void* coef, const void* hob, const void* hint, double e,
                            double* result, void* stream);
/* Host copy of the recursion scalars of the last call (synchronises):
 * diag (2*lanczos_iter doubles, complex), offd (lanczos_iter doubles). */
int mctdhb_sil_scalars(const mctdhb_sil* sil, int* iter_done, double* diag, double* offd,
                       void* stream);

/* ---- orbital step (device-resident) ---------------------------------- */

typedef struct mctdhb_orbital mctdhb_orbital; /* orbital-equation workspace */

/* Grid of grid_size points on [xi, xf] (dx = (xf - xi)/(grid_size - 1), the
 * last point repeating the first for PERIODIC_BOUNDS), kinetic coefficients
 * d2coef (real) and d1coef (complex), contact interaction g, step tstep in
 * imaginary (prop_dt = -i tstep) or real time, as reference
 * get_orbital_equation / get_orbital_workspace (types_definition.c:30-108,
 * :200-244).  pot: grid_size doubles on the host. */
int mctdhb_orbital_create(int norb, int grid_size, int bounds, double xi, double xf, double tstep, double g,
                          double d2coef, double d1coef_re, double d1coef_im, int imagtime, const double* pot,
                          mctdhb_orbital** out);
int mctdhb_orbital_destroy(mctdhb_orbital* w);
int mctdhb_orbital_set_potential(mctdhb_orbital* w, const double* pot);
/* Contact interaction g for later launches (synchronize.c:59-68 refreshes it
 * from the time-dependent parameter in real time). */
int mctdhb_orbital_set_interaction(mctdhb_orbital* w, double g);
/* Order of the nonlinear Runge-Kutta sub-step: 2, 4 (default) or 5
 * (RungeKuttaOrder, mctdhb_types.h:80-85; orbital_integration.c:384-416). */
int mctdhb_orbital_set_rk_order(mctdhb_orbital* w, int order);
/* Device arrays: orb norb x grid_size, hob M x M, hint M^4, rho2 M^4,
 * rinv M x M (inverse one-body density matrix), all complex128. */
int mctdhb_linear_fft_dev(mctdhb_orbital* w, void* orb, void* stream);       /* split_linear_orbitals.c:375-402 */
int mctdhb_linear_cn_dev(mctdhb_orbital* w, void* orb, void* stream);        /* split_linear_orbitals.c:335-373 */
int mctdhb_orbital_matrices_dev(mctdhb_orbital* w, const void* orb, void* hob, void* hint,
                                void* stream);                                /* synchronize.c:34-39 */
int mctdhb_nonlinear_rhs_dev(mctdhb_orbital* w, const void* orb, const void* hob, const void* hint,
                             const void* rho2, const void* rinv, void* der,
                             void* stream);                                   /* orbital_integration.c:242-281 */
int mctdhb_orthonormalize_dev(mctdhb_orbital* w, void* orb, void* stream);   /* calculus.c:170-191 */
int mctdhb_splitstep_dev(mctdhb_orbital* w, int use_fft, void* orb, void* hob, void* hint, const void* rho2,
                         const void* rinv, void* stream);                     /* orbital_integration.c:351-434 */
/* Full-equation Runge-Kutta orbital step: linear part by FFT (use_fd 0),
 * finite differences (1) or the exponential DVR (2; its n^2 matrix is built
 * on the GPU on first use), simple or robust projector, Gram-Schmidt in
 * imaginary time; rho2 and rinv held fixed. */
int mctdhb_fullstep_dev(mctdhb_orbital* w, int use_fd, int robust, void* orb, const void* rho2, const void* rinv,
                        void* stream);                                        /* orbital_integration.c:283-348 */
int mctdhb_fullstep_rhs_dev(mctdhb_orbital* w, int use_fd, int robust, const void* orb, const void* rho2,
                            const void* rinv, void* der, void* stream);       /* orbital_integration.c:158-235 */
/* proj = projection of hact (norb x grid) onto the orbitals */
int mctdhb_projector_dev(mctdhb_orbital* w, int robust, const void* orb, const void* hact, void* proj,
                         void* stream);                                       /* orbital_integration.c:22-129 */
/* The orbital integrator of mctdhb_propagate_step_dev and the step graph:
 * orb_method SPLITSTEP (31, default) or FULLSTEP_RUNGEKUTTA (30).  With the
 * full step, der_method DVR (300) selects the DVR linear part (the matrix is
 * built here); otherwise use_fft of the step selects FFT or finite
 * differences.  der_method 0 means "as use_fft says". */
int mctdhb_orbital_set_integrator(mctdhb_orbital* w, int orb_method, int der_method, int robust);
/* rho1, rho2 of coef, then natural occupations (eigvals, M doubles, device,
 * may be NULL), regularisation and rinv = rho1^-1 (synchronize.c:41-56). */
int mctdhb_density_sync_dev(mctdhb_space* space, const void* coef, void* rho1, void* rho2, void* rinv,
                            void* eigvals, double reg_factor, int prec, void* stream);
/* One MCTDHB time step for a time-independent trap and interaction
 * (driver.c:159-202): orbital split step, h/g rebuild, SIL coefficient
 * step, density sync with the regularisation factor of
 * set_regulatization_factor (default 1e-9).  All arrays device-resident. */
int mctdhb_propagate_step_dev(mctdhb_space* space, mctdhb_sil* sil, mctdhb_orbital* w, int use_fft, void* orb,
                              void* coef, void* hob, void* hint, void* rho1, void* rho2, void* rinv,
                              void* stream);

/* The same step recorded once as a HIP graph and replayed per step: the
 * body of the reference's time loop (driver.c:159-202 inside
 * driver.c:283-320) as one launch.  The graph holds the buffer addresses
 * given here and the scalars of the moment of capture (dt, g, Runge-Kutta
 * order, Lanczos size); the trap is read from the orbital workspace at each
 * launch, so mctdhb_orbital_set_potential between launches takes effect.
 * Covers 1 <= norb <= 7 (the fused density kernels).  Replays run in stream
 * order on the stream given to launch. */
typedef struct mctdhb_step_graph mctdhb_step_graph;
int mctdhb_step_graph_create(mctdhb_space* space, mctdhb_sil* sil, mctdhb_orbital* w, int use_fft, void* orb,
                             void* coef, void* hob, void* hint, void* rho1, void* rho2, void* rinv,
                             mctdhb_step_graph** out);
int mctdhb_step_graph_launch(mctdhb_step_graph* g, void* stream);
/* `rounds` replays of n step graphs of independent solves: in each round
 * graph i is launched on streams[i % nstreams], so every solve's steps stay
 * in order on its own stream.  One call in place of rounds * n launches from
 * the caller (many solves at once, reference driver.c:159-202 per solve). */
int mctdhb_step_graph_launch_many(mctdhb_step_graph* const* graphs, int n, void* const* streams, int nstreams,
                                  int rounds);
int mctdhb_step_graph_destroy(mctdhb_step_graph* g);

/* Packed steps: the step of mctdhb_propagate_step_dev (reference
 * driver.c:159-202, mctdhb_propagate_step, once per solve) for many small
 * solves at once, one kernel launch per stage over a table of the solves'
 * arguments in device memory instead of one launch per stage and solve.
 * A solve qualifies when its step takes the FFT linear part (grid_size - 1 a
 * power of two), the split-step orbital integrator, 2 <= norb <= 7 and the
 * one-workgroup FP64 SIL step (complex128 workspace, NC <= 8192); eligible
 * receives 1 or 0 (then mctdhb_last_error names the reason).  create groups
 * the n solves by the stages' kernel variants (norb, Runge-Kutta order,
 * imaginary or real time, SIL index scheme and LDS staging) and records
 * each group's step once as a graph; the buffers and the scalars of the
 * moment of creation are kept as with mctdhb_step_graph_create.  step
 * launches `rounds` steps of every group, group g on streams[g % nstreams];
 * a solve's results are those of its own mctdhb_propagate_step_dev, bit for
 * bit.  The solves of a pack share no space, SIL or orbital workspace and no
 * coefficient or orbital buffer (create refuses that), and a solve must not
 * be stepped by a pack and a step graph at the same time. */
typedef struct mctdhb_pack mctdhb_pack;
int mctdhb_pack_eligible(const mctdhb_space* space, const mctdhb_sil* sil, const mctdhb_orbital* w, int use_fft,
                         int* eligible);
int mctdhb_pack_create(int n, mctdhb_space* const* spaces, mctdhb_sil* const* sils, mctdhb_orbital* const* w,
                       int use_fft, void* const* orb, void* const* coef, void* const* hob, void* const* hint,
                       void* const* rho1, void* const* rho2, void* const* rinv, mctdhb_pack** out);
int mctdhb_pack_groups(const mctdhb_pack* p, int* groups);
int mctdhb_pack_step(mctdhb_pack* p, void* const* streams, int nstreams, int rounds);
int mctdhb_pack_destroy(mctdhb_pack* p);

#ifdef __cplusplus
}
#endif

#endif
