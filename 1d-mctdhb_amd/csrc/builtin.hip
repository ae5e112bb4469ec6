// builtin.hip - the reference's builtin trap potentials and time-dependent
// interaction parameter (host functions, evaluated on the grid once per
// real-time step).  Names, parameter order and formulas are those of
// reference src/function_tools/builtin_potential.c and
// builtin_time_parameter.c, so parameter files written for the reference
// select the same functions here.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "mctdhb_hip.h"

namespace {

// value going linearly from v0 at t = 0 to vf at t = period, vf when
// period <= 0 (builtin_potential.c:8-13); not clamped after the period
double linear_sweep(double t, double v0, double vf, double period) {
    return period <= 0 ? vf : v0 + (vf - v0) * t / period;
}

// 0 -> top in ramp, top for hold, top -> 0 in ramp, then 0
// (builtin_potential.c:15-35)
double trapezoid_sweep(double t, double top, double ramp, double hold) {
    if (t <= ramp) return linear_sweep(t, 0, top, ramp);
    if (t < ramp + hold) return top;
    if (t < 2 * ramp + hold) return top - (t - ramp - hold) / ramp * top;
    return 0;
}

const double* P(void* params) { return (const double*) params; }

void require_width(double width, const double* x) {
    if (width < x[1] - x[0]) {
        printf("\n\nERROR: linear potential barrier requires a width greater than spatial grid step size\n\n");
        exit(EXIT_FAILURE);
    }
}

}  // namespace

// params: omega0, omegaf, period (builtin_potential.c:37-50)
extern "C" void potfunc_harmonic(double t, uint16_t npts, Rarray x, void* params, Rarray V) {
    const double w = linear_sweep(t, P(params)[0], P(params)[1], P(params)[2]);
    for (int i = 0; i < npts; i++) V[i] = 0.5 * w * w * x[i] * x[i];
}

// params: b, a0, af, period; V = b^2 (x^2 - a^2 / b^2)^2 (builtin_potential.c:52-69)
extern "C" void potfunc_doublewell(double t, uint16_t npts, Rarray x, void* params, Rarray V) {
    const double b = P(params)[0];
    const double a = linear_sweep(t, P(params)[1], P(params)[2], P(params)[3]);
    const double b2 = b * b, shift = a * a / b2;
    for (int i = 0; i < npts; i++) {
        const double u = x[i] * x[i] - shift;
        V[i] = b2 * u * u;
    }
}

// params: omega, width, h0, hf, period (builtin_potential.c:71-93)
extern "C" void potfunc_harmonicgauss(double t, uint16_t npts, Rarray x, void* params, Rarray V) {
    const double w = P(params)[0], width = P(params)[1];
    const double h = linear_sweep(t, P(params)[2], P(params)[3], P(params)[4]);
    require_width(width, x);
    for (int i = 0; i < npts; i++) {
        const double x2 = x[i] * x[i];
        V[i] = w * w * x2 / 2 + h * exp(-x2 / (width * width));
    }
}

// params: width, h0, hf, period, background (builtin_potential.c:95-119)
extern "C" void potfunc_barrier(double t, uint16_t npts, Rarray x, void* params, Rarray V) {
    const double width = P(params)[0], bg = P(params)[4];
    const double h = linear_sweep(t, P(params)[1], P(params)[2], P(params)[3]);
    require_width(width, x);
    for (int i = 0; i < npts; i++) {
        V[i] = bg;
        if (fabs(x[i]) <= width / 2) {
            const double c = cos(x[i] * PI / width);
            V[i] += h * c * c;
        }
    }
}

// params: width, top, ramp, hold (builtin_potential.c:121-140)
extern "C" void potfunc_time_trapezoid_barrier(double t, uint16_t npts, Rarray x, void* params, Rarray V) {
    const double width = P(params)[0];
    const double h = trapezoid_sweep(t, P(params)[1], P(params)[2], P(params)[3]);
    for (int i = 0; i < npts; i++) {
        V[i] = 0;
        if (fabs(x[i]) <= width / 2) {
            const double c = cos(x[i] * PI / width);
            V[i] = h * c * c;
        }
    }
}

// params: k, h0, hf, period; V = h sin^2(pi k x / L), L = x[n-1] - x[0]
// (builtin_potential.c:142-158)
extern "C" void potfunc_opticallattice(double t, uint16_t npts, Rarray x, void* params, Rarray V) {
    const double len = x[npts - 1] - x[0], k = P(params)[0];
    const double h = linear_sweep(t, P(params)[1], P(params)[2], P(params)[3]);
    for (int i = 0; i < npts; i++) {
        const double s = sin(PI * k * x[i] / len);
        V[i] = h * s * s;
    }
}

// params: k, top, ramp, hold (builtin_potential.c:160-177)
extern "C" void potfunc_time_trapezoid_opticallattice(double t, uint16_t npts, Rarray x, void* params, Rarray V) {
    const double len = x[npts - 1] - x[0], k = P(params)[0];
    const double h = trapezoid_sweep(t, P(params)[1], P(params)[2], P(params)[3]);
    for (int i = 0; i < npts; i++) {
        const double s = sin(PI * k * x[i] / len);
        V[i] = h * s * s;
    }
}

// params: xs, v0, vf, period; V = v(t) for x >= xs (builtin_potential.c:179-193)
extern "C" void potfunc_step(double t, uint16_t npts, Rarray x, void* params, Rarray V) {
    const double xs = P(params)[0];
    const double v = linear_sweep(t, P(params)[1], P(params)[2], P(params)[3]);
    for (int i = 0; i < npts; i++) V[i] = x[i] >= xs ? v : 0;
}

// params: xi, xf, v_out0, v_outf, period.  The reference's mask keeps the
// outside value only where x <= xi and x >= xf, i.e. zero wherever
// x > xi or x < xf (builtin_potential.c:195-212, kept as written).
extern "C" void potfunc_square_well(double t, uint16_t npts, Rarray x, void* params, Rarray V) {
    const double xi = P(params)[0], xf = P(params)[1];
    const double v = linear_sweep(t, P(params)[2], P(params)[3], P(params)[4]);
    for (int i = 0; i < npts; i++) V[i] = (x[i] > xi || x[i] < xf) ? 0 : v;
}

// builtin_potential.c:214-260; "custom" gives NULL (the caller's function)
extern "C" single_particle_pot get_builtin_pot(char pot_name[]) {
    static const struct {
        const char* name;
        single_particle_pot f;
    } table[] = {
        {"harmonic", potfunc_harmonic},
        {"doublewell", potfunc_doublewell},
        {"harmonicgauss", potfunc_harmonicgauss},
        {"barrier", potfunc_barrier},
        {"trapezoid_barrier", potfunc_time_trapezoid_barrier},
        {"opticallattice", potfunc_opticallattice},
        {"trapezoid_opticallattice", potfunc_time_trapezoid_opticallattice},
        {"potfunc_step", potfunc_step},
        {"potfunc_square_well", potfunc_square_well},
    };
    for (const auto& e : table)
        if (strcmp(pot_name, e.name) == 0) return e.f;
    if (strcmp(pot_name, "custom") == 0) return nullptr;
    printf("\n\nERROR: Potential '%s' not valid\n\n", pot_name);
    exit(EXIT_FAILURE);
}

// params: g0, gf, period; gf from t >= period on (builtin_time_parameter.c:6-15)
extern "C" double timeparam_linear_ramp(double t, void* params) {
    const double g0 = P(params)[0], gf = P(params)[1], period = P(params)[2];
    if (t >= period) return gf;
    return g0 + (gf - g0) * t / period;
}

// builtin_time_parameter.c:17-30
extern "C" time_dependent_parameter get_builtin_param_func(char funcname[]) {
    if (strcmp(funcname, "linear_ramp") == 0) return timeparam_linear_ramp;
    if (strcmp(funcname, "custom") == 0) return nullptr;
    printf("\n\nIOERROR : Builtin time function '%s' not found\n\n", funcname);
    exit(EXIT_FAILURE);
}
