// dataio.hip - the reference's text formats: the parameters line and the
// integrator description read at set-up, the orbital and coefficient
// files, and the records written during and after a run (reference
// src/assistant/dataio.c).  Host code only.
//
// Numbers are complex text tokens "(re+imj)", the form numpy writes and
// reads (numpy.loadtxt(..., dtype=complex128)); reading uses the
// reference's scanf formats, kept in the same global variables.  The
// helpers the reference takes from its text-array library (line counting,
// comment skipping, array streaming) are written here: a comment line
// starts with '#'.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "mctdhb_hip.h"

// reference dataio.c:14-29
char orb_cplx_read_fmt[STR_BUFF_SIZE] = " (%lf%lfj) ";
char coef_cplx_read_fmt[STR_BUFF_SIZE] = " (%lf%lfj)";
char inp_dirname[STR_BUFF_SIZE] = "input/";
char out_dirname[STR_BUFF_SIZE] = "output/";
char integrator_desc_fname[STR_BUFF_SIZE] = "mctdhb_integrator.conf";
uint8_t monitor_energy_digits = 10;
Bool new_empty_append_files = FALSE;

namespace {

// Written complex and real numbers: 17 significant digits, enough to read
// every double back exactly.
constexpr const char* kCplxBefore = " (%.16E%+.16Ej)";
constexpr const char* kCplxBoth = " (%.16E%+.16Ej) ";
constexpr const char* kRealBefore = " %.16E";

FILE* open_or_die(const char* fname, const char* mode) {
    FILE* f = fopen(fname, mode);
    if (!f) {
        printf("\n\nIOERROR: impossible to open file %s with mode %s\n\n", fname, mode);
        exit(EXIT_FAILURE);
    }
    return f;
}

// Move past blank lines and lines whose first non-blank character is '#'.
void skip_comments(FILE* f) {
    for (;;) {
        int c;
        do c = fgetc(f);
        while (c == ' ' || c == '\t' || c == '\r' || c == '\n');
        if (c != '#') {
            if (c != EOF) ungetc(c, f);
            return;
        }
        while (c != '\n' && c != EOF) c = fgetc(f);
    }
}

void skip_line(FILE* f) {
    int c;
    do c = fgetc(f);
    while (c != '\n' && c != EOF);
}

// Non-blank, non-comment lines of a file.
uint32_t data_lines(const char* fname) {
    FILE* f = open_or_die(fname, "r");
    uint32_t n = 0;
    int c, first = 0;  // first non-blank character of the current line (0: none yet)
    while ((c = fgetc(f)) != EOF) {
        if (c == '\n') {
            n += first && first != '#';
            first = 0;
        } else if (!first && c != ' ' && c != '\t' && c != '\r') {
            first = c;
        }
    }
    n += first && first != '#';
    fclose(f);
    return n;
}

void read_cplx(FILE* f, const char* fmt, size_t n, dcomplex* out, const char* fname) {
    for (size_t i = 0; i < n; i++) {
        double re, im;
        if (fscanf(f, fmt, &re, &im) != 2) {
            printf("\n\nIOERROR: reading complex number %zu of %zu from file %s\n\n", i + 1, n, fname);
            exit(EXIT_FAILURE);
        }
        double v[2] = {re, im};
        memcpy(&out[i], v, sizeof v);
    }
}

void put_cplx(FILE* f, const char* fmt, dcomplex z) {
    double v[2];
    memcpy(v, &z, sizeof v);
    fprintf(f, fmt, v[0], v[1]);
}

// One line of n complex values, appended.
void append_line(const char* fname, size_t n, const dcomplex* a) {
    if (new_empty_append_files) remove(fname);
    FILE* f = open_or_die(fname, "a");
    for (size_t i = 0; i < n; i++) put_cplx(f, kCplxBefore, a[i]);
    fputc('\n', f);
    fclose(f);
}

// One line holding the rows of a Cmatrix back to back, appended.
void append_rows(const char* fname, int rows, int cols, Cmatrix m) {
    if (new_empty_append_files) remove(fname);
    FILE* f = open_or_die(fname, "a");
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < cols; c++) put_cplx(f, kCplxBefore, m[r][c]);
    fputc('\n', f);
    fclose(f);
}

void report_integrator_warning(const char* fname, int which, const char* what) {
    printf("\n\nWARNING: Reading %d value : %s. Source file name %s\n\n", which, what, fname);
}

}  // namespace

extern "C" void toggle_new_append_files(void) { new_empty_append_files = new_empty_append_files ? FALSE : TRUE; }

// grid rows x orbital columns in the file (dataio.c:52-66)
extern "C" void set_orbitals_from_file(char fname[], ManyBodyState psi) {
    FILE* f = open_or_die(fname, "r");
    skip_comments(f);
    const int M = psi->norb, n = psi->grid_size;
    dcomplex* row = (dcomplex*) malloc(sizeof(dcomplex) * M);
    for (int j = 0; j < n; j++) {
        read_cplx(f, orb_cplx_read_fmt, M, row, fname);
        for (int i = 0; i < M; i++) psi->orbitals[i][j] = row[i];
    }
    free(row);
    fclose(f);
}

// one coefficient per line (dataio.c:68-72)
extern "C" void set_coef_from_file(char fname[], ManyBodyState psi) {
    FILE* f = open_or_die(fname, "r");
    skip_comments(f);
    read_cplx(f, coef_cplx_read_fmt, psi->space_dim, psi->coef, fname);
    fclose(f);
}

// The 21 fields of one parameters line and the integrator description
// (dataio.c:74-293).
extern "C" MCTDHBDataStruct get_mctdhb_from_files(char par_fname[], char integ_fname[], uint32_t line,
                                                 single_particle_pot custom_pot_fun,
                                                 time_dependent_parameter custom_inter_fun, void* custom_pot_params,
                                                 void* custom_inter_params) {
    const uint32_t nlines = data_lines(par_fname);
    if (line > nlines) {
        printf("\n\nIOERROR : Requested to read parameters from line %u but file %s has %u lines\n\n", line,
               par_fname, nlines);
        exit(EXIT_FAILURE);
    }
    double* gpar = (double*) calloc(5, sizeof(double));
    double* potpar = (double*) calloc(5, sizeof(double));
    FILE* f = open_or_die(par_fname, "r");
    skip_comments(f);
    for (uint32_t i = line; i > 1; i--) {
        skip_line(f);
        skip_comments(f);
    }
    unsigned npar, norb, grid_size;
    double xi, xf, tstep, tend, d2coef, im_d1coef;
    char g_name[STR_BUFF_SIZE], pot_name[STR_BUFF_SIZE];
    const int got = fscanf(f, "%u %u %u %lf %lf %lf %lf %lf %lf %127s %lf %lf %lf %lf %lf %127s %lf %lf %lf %lf %lf",
                           &npar, &norb, &grid_size, &xi, &xf, &tstep, &tend, &d2coef, &im_d1coef, g_name, &gpar[0],
                           &gpar[1], &gpar[2], &gpar[3], &gpar[4], pot_name, &potpar[0], &potpar[1], &potpar[2],
                           &potpar[3], &potpar[4]);
    fclose(f);
    if (got != 21) {
        printf("\n\nIOERROR: Expected to read %d in line %u of file %s. However fscanf returned %d\n\n", 21, line,
               par_fname, got);
        exit(EXIT_FAILURE);
    }
    single_particle_pot pot_func = get_builtin_pot(pot_name);
    time_dependent_parameter g_func = get_builtin_param_func(g_name);
    void* pot_params = potpar;
    void* inter_params = gpar;
    if (!pot_func) pot_func = custom_pot_fun;
    if (custom_pot_params) {
        pot_params = custom_pot_params;
        free(potpar);
    }
    if (!g_func) g_func = custom_inter_fun;
    if (custom_inter_params) {
        inter_params = custom_inter_params;
        free(gpar);
    }
    // integrator description: seven integers, each after optional comments
    unsigned v[7] = {DEFAULT_INTEGRATION_TYPE, DEFAULT_COEF_INTEGRATOR, DEFAULT_ORB_INTEGRATOR,
                     DEFAULT_ORB_DERIVATIVE,   DEFAULT_RUNGEKUTTA_ORDER, DEFAULT_BOUNDARY_CONDITION,
                     DEFAULT_LANCZOS_ITER};
    static const char* what[7] = {"time integration type", "coef integration method", "orbital integration method",
                                  "Derivatives method", "global runge-kutta order", "Boundary conditions",
                                  "Lanczos iterations"};
    if ((f = fopen(integ_fname, "r")) != nullptr) {
        bool warned = false;
        for (int k = 0; k < 7; k++) {
            skip_comments(f);
            unsigned x;
            if (fscanf(f, "%u", &x) == 1) {
                v[k] = x;
            } else {
                report_integrator_warning(integ_fname, k + 1, what[k]);
                warned = true;
            }
        }
        fclose(f);
        if (warned)
            printf("\n\nWARNING: some of the integration config could not be set from file %s, which were "
                   "listed above\n\n",
                   integ_fname);
    } else {
        printf("\n\nWARNING: could not set integrator descriptor, file %s not found. Using default values.\n\n",
               integ_fname);
    }
    const double d1[2] = {0.0, im_d1coef};
    dcomplex d1coef;
    memcpy(&d1coef, d1, sizeof d1);
    return get_mctdhb_struct((IntegratorType) v[0], (CoefIntegrator) v[1], (OrbIntegrator) v[2],
                             (OrbDerivative) v[3], (RungeKuttaOrder) v[4], (uint16_t) (uint8_t) v[6],
                             (uint16_t) npar, (uint16_t) norb, pot_name, (BoundaryCondition) v[5], xi, xf,
                             (uint16_t) grid_size, tstep, tend, d2coef, d1coef, pot_params, inter_params, pot_func,
                             g_func);
}

// <inp_dirname><prefix>_mctdhb_parameters.dat line job_num, orbitals and
// coefficients by the job-input rule, then densities and h/g
// (dataio.c:295-378).
extern "C" MCTDHBDataStruct full_setup_mctdhb_current_dir(char fprefix[], uint32_t job_num,
                                                          JobsInputHandle which_inp,
                                                          single_particle_pot custom_pot_fun,
                                                          time_dependent_parameter custom_inter_fun,
                                                          void* custom_pot_params, void* custom_inter_params) {
    std::string par = std::string(inp_dirname) + fprefix + PARAMS_FNAME_SUFFIX;
    MCTDHBDataStruct m = get_mctdhb_from_files((char*) par.c_str(), integrator_desc_fname, job_num, custom_pot_fun,
                                               custom_inter_fun, custom_pot_params, custom_inter_params);
    auto input_name = [&](const char* what) {
        std::string dir = inp_dirname;
        uint32_t job = 1;
        if (which_inp == MULTIPLE_INP) job = job_num;
        if (which_inp == LAST_JOB_OUT && job_num > 1) {
            dir = out_dirname;
            job = job_num - 1;
        }
        return dir + fprefix + "_job" + std::to_string(job) + what;
    };
    std::string orb = input_name("_orb.dat"), coef = input_name("_coef.dat");
    set_orbitals_from_file((char*) orb.c_str(), m->state);
    set_coef_from_file((char*) coef.c_str(), m->state);
    sync_density_matrices(m->multiconfig_space, m->state);
    sync_orbital_matrices(m->orb_eq, m->state);
    return m;
}

// dataio.c:519-527
extern "C" uint32_t auto_number_of_jobs(char prefix[]) {
    std::string par = std::string(inp_dirname) + prefix + PARAMS_FNAME_SUFFIX;
    return data_lines(par.c_str());
}

// dataio.c:529-558
extern "C" void set_output_fname(char prefix[], RecordDataType id, char* fname) {
    const char* suffix = "";
    switch (id) {
        case ORBITALS_REC: suffix = "_orb.dat"; break;
        case COEFFICIENTS_REC: suffix = "_coef.dat"; break;
        case PARAMETERS_REC: suffix = PARAMS_FNAME_SUFFIX; break;
        case ONE_BODY_MATRIX_REC: suffix = "_obmat.dat"; break;
        case TWO_BODY_MATRIX_REC: suffix = "_tbmat.dat"; break;
        case ONE_BODY_POTENTIAL_REC: suffix = "_obpotential.dat"; break;
    }
    snprintf(fname, STR_BUFF_SIZE, "%s%s%s", out_dirname, prefix, suffix);
}

// rho2 (M^4), rho1 (M x M) and the orbitals (M x grid) as one appended
// line each (dataio.c:705-746)
extern "C" void append_processed_state(char prefix[], ManyBodyState psi) {
    char fname[STR_BUFF_SIZE];
    const int M = psi->norb;
    set_output_fname(prefix, TWO_BODY_MATRIX_REC, fname);
    append_line(fname, (size_t) M * M * M * M, psi->tb_denmat);
    set_output_fname(prefix, ONE_BODY_MATRIX_REC, fname);
    append_rows(fname, M, M, psi->ob_denmat);
    set_output_fname(prefix, ORBITALS_REC, fname);
    append_rows(fname, M, psi->grid_size, psi->orbitals);
}

// coefficients as a column, orbitals as grid rows x orbital columns
// (dataio.c:747-764)
extern "C" void record_raw_state(char prefix[], ManyBodyState psi) {
    char fname[STR_BUFF_SIZE];
    set_output_fname(prefix, COEFFICIENTS_REC, fname);
    FILE* f = open_or_die(fname, "w");
    for (uint32_t i = 0; i < psi->space_dim; i++) {
        put_cplx(f, kCplxBefore, psi->coef[i]);
        fputc('\n', f);
    }
    fclose(f);
    set_output_fname(prefix, ORBITALS_REC, fname);
    f = open_or_die(fname, "w");
    for (int j = 0; j < psi->grid_size; j++) {
        for (int k = 0; k < psi->norb; k++) put_cplx(f, kCplxBoth, psi->orbitals[k][j]);
        fputc('\n', f);
    }
    fclose(f);
}

// dataio.c:766-781
extern "C" void append_timestep_potential(char prefix[], OrbitalEquation eq) {
    char fname[STR_BUFF_SIZE];
    set_output_fname(prefix, ONE_BODY_POTENTIAL_REC, fname);
    if (new_empty_append_files) remove(fname);
    FILE* f = open_or_die(fname, "a");
    for (int i = 0; i < eq->grid_size; i++) fprintf(f, kRealBefore, eq->pot_grid[i]);
    fputc('\n', f);
    fclose(f);
}

// g(t) for t = 0 and for every step whose index is NOT a multiple of recn:
// the reference's test reads `if (prop_steps % recn)` (dataio.c:783-813),
// kept as written.
extern "C" void record_time_interaction(char prefix[], OrbitalEquation eq, uint16_t recn) {
    std::string fname = std::string(out_dirname) + prefix + "_interaction.dat";
    FILE* f = open_or_die(fname.c_str(), "w");
    fprintf(f, "%.10E\n", eq->inter_param(0, eq->inter_extra_args));
    uint32_t steps = 0;
    double t = 0;
    while (t < eq->tend) {
        steps++;
        if (steps % recn) {
            t = steps * eq->tstep;
            fprintf(f, "%.10E\n", eq->inter_param(t, eq->inter_extra_args));
        }
    }
    fclose(f);
}

// 0 and every recn-th step time (dataio.c:815-835)
extern "C" void record_time_array(char prefix[], double tend, double tstep, uint16_t recn) {
    std::string fname = std::string(out_dirname) + prefix + "_timesteps.dat";
    FILE* f = open_or_die(fname.c_str(), "w");
    fprintf(f, "%.6lf\n", 0.0);
    uint32_t steps = 0;
    while (steps * tstep < tend) {
        steps++;
        if (steps % recn == 0) fprintf(f, "%.6lf\n", steps * tstep);
    }
    fclose(f);
}

// One line of 11 fields: npar norb grid xi xf tstep tend d2coef Im(d1coef)
// g E/N (dataio.c:837-866)
extern "C" void append_mctdhb_parameters(char prefix[], MCTDHBDataStruct m) {
    char fname[STR_BUFF_SIZE];
    set_output_fname(prefix, PARAMETERS_REC, fname);
    if (new_empty_append_files) remove(fname);
    const dcomplex e = total_energy(m->state);
    double ev[2], d1[2];
    memcpy(ev, &e, sizeof ev);
    memcpy(d1, &m->orb_eq->d1coef, sizeof d1);
    FILE* f = open_or_die(fname, "a");
    fprintf(f, "%u %u %u %.10lf %.10lf %.10lf %.2lf %.1lf %.10lf %.10E %.10E\n", (unsigned) m->state->npar,
            (unsigned) m->state->norb, (unsigned) m->state->grid_size, m->orb_eq->xi, m->orb_eq->xf,
            m->orb_eq->tstep, m->orb_eq->tend, m->orb_eq->d2coef, d1[1], m->orb_eq->g,
            ev[0] / m->state->npar);
    fclose(f);
}
