/*
 * hhg_oracle.h — CPU restatement of the HigherHarmonics per-k-point Bloch
 * solve and current reduction.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle the HIP product path
 * is compared with, and the CPU baseline that bench.py times.  Only tests/, the smoke
 * entry point and bench.py's CPU-baseline leg load it.  Product code never
 * links, loads or calls it.
 *
 * Everything is FP64 (reference: sources/HHG/GlobalDefinitions.hpp:14).
 */
#ifndef HHG_ORACLE_H
#define HHG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_MODEL_DIRAC2D = 0, OR_MODEL_PIFLUX2D = 1, OR_MODEL_HONEYCOMB = 2, OR_MODEL_DIRAC3D = 3, OR_MODEL_PIFLUX3D = 4 };
enum {
    OR_PULSE_COS2 = 0,
    OR_PULSE_GAUSS = 1,
    OR_PULSE_DOUBLE_GAUSS = 2,
    OR_PULSE_DOUBLE_COS2 = 3,
    OR_PULSE_CONTINUOUS = 4,
    OR_PULSE_SPLINE = 5
};
enum { OR_RELAX_NONE = 0, OR_RELAX_TO_INITIAL = 1, OR_RELAX_T2 = 2, OR_RELAX_EQUILIBRIUM = 3 };
enum { OR_METHOD_RKF78 = 0, OR_METHOD_MAGNUS4 = 1, OR_METHOD_MAGNUS_MOMENTS = 2 };
enum { OR_PROPOSAL_RESTART = 0, OR_PROPOSAL_CARRY = 1 };

typedef struct {
    int32_t kind;
    int32_t relax;
    double v;          /* Dirac-2D velocity */
    double mass;       /* Dirac-2D gap parameter Delta */
    double hopping;    /* PiFlux / honeycomb t */
    double theta;      /* polarisation angle of A(t) in the kx-ky plane */
    double fermi_energy;
    double beta;       /* inverse temperature; <= 0 means T = 0 */
    double relax_time;        /* equilibrium relaxation: the diagonal time */
    double relax_time_offdiag; /* equilibrium relaxation: the off-diagonal time (<= 0: none) */
    int32_t ic_rescue;        /* 3D pi-flux rescue of the reference's drivers: 1 continuum grid, 2 uniform grid */
    int32_t eq_form;          /* equilibrium relaxation targets: 0 the reference's formulas; 1 the HIP
                                 kernels' one-exponential form, restated for its CPU check */
    double split_dispersion;  /* 3D pi-flux, > 0: j split at |m(k - A)| into two outputs of N + 1 samples */
    int32_t perp_current;     /* honeycomb, 1: a second output of N + 1 samples, the reference's j_perp
                                 (Honeycomb.cpp:328-334) */
    int32_t pad2;
} or_model;

typedef struct {
    int32_t kind;
    int32_t pad;
    double amp_a;      /* peak vector-potential amplitude of pulse A */
    double amp_b;      /* of pulse B (double pulses) */
    double n_cycles;   /* cos^2: envelope lasts 2*pi*n_cycles */
    double sigma;      /* Gaussian: envelope width */
    double t0;         /* cos^2: window begin; Gaussian: centre of pulse A */
    double delay;      /* B pulse delay relative to A */
    double cep;        /* carrier-envelope phase added to the carrier */
    /* OR_PULSE_SPLINE: per interval 4 coefficients (or_spline_coefficients) of
     * the cubic in x = (t - t_i)/dt; A = 0 outside [s_t0, s_t_last] */
    const double* coef;
    int64_t n_intervals;
    double s_t0, s_dt, s_t_last;
} or_pulse;

typedef struct {
    double t_begin;
    double t_end;
    int32_t n_measurements;
    int32_t n_subdivisions;
} or_time;

typedef struct {
    int32_t method;
    int32_t proposal;
    double rtol;
    double atol;
    int32_t max_fail;        /* consecutive rejections before a trajectory is marked failed */
    int32_t magnus_substeps; /* Magnus steps per sampling interval */
    int32_t max_steps;       /* accepted steps per sampling interval before a trajectory is marked failed
                                (<= 0: 100000) */
    int32_t pad;
} or_solver;

typedef struct {
    int64_t attempts;
    int64_t rejects;
    int64_t failed_trajectories;
    int64_t rhs_evals;
} or_stats;

/* The equilibrium relaxations' Fermi pair (P - M, P M) with P = f(c + x),
 * M = f(c - x), f(e) = 1 / (1 + exp(beta e)): form 0 the reference's two Fermi
 * functions, 1 the HIP kernels' one-exponential form. */
void or_fermi_pair(int form, double c, double x, double beta, double* pm, double* pp);

/* Pulse A(t) (vector potential including its amplitude). */
double or_pulse_A(const or_pulse* p, double t);

/* Clamped cubic spline through n >= 2 samples spaced dt with end slopes d0, d1
 * (the interpolant of Boost's cardinal_cubic_b_spline): coef[4 (n - 1)], per
 * interval c with y(t_i + x dt) = c0 + x (c1 + x (c2 + x c3)).  Returns 0, or -1
 * for bad arguments. */
int or_spline_coefficients(const double* y, int64_t n, double dt, double d0, double d1, double* coef);

/* Gauss-Legendre nodes/weights on [-1,1], ascending nodes. */
void or_gauss_legendre(int n, double* x, double* w);

/* Fehlberg 7(8) tableau as used by the oracle: c[13], a[13*13] row-major,
 * b8[13] (propagated), b7[13] (embedded). */
void or_rkf78_tableau(double* c, double* a, double* b8, double* b7);

/* Points are (kx, ky, kz); kz is used by the 3D pi-flux lattice only. */

/* Initial Bloch vector of one k-point. */
void or_initial_state(const or_model* m, double kx, double ky, double kz, double s[3]);

/* Bloch right-hand side ds/dt at time t (used by tests against scipy).
 * s_ref: the initial state (relaxation toward it) or the state at the start
 * of the step (equilibrium relaxation). */
void or_rhs(const or_model* m, const or_pulse* p, double kx, double ky, double kz, const double s_ref[3],
            const double s[3], double t, double out[3]);

/* The four pulse moments of one sampling interval [t0, t0 + h] that the
 * reference's Dirac Magnus step uses (Laser::magnus_coefficients). */
void or_magnus_moments(const or_pulse* p, double t0, double h, double mom[4]);

/* One moment-based Magnus step of the 3D Dirac cone at (k_z, kappa) over an
 * interval of length h with the interval's moments (DiracSystem::time_evolution_magnus). */
void or_magnus_moments_step(double kz, double kappa, const double mom[4], double h, double s[3]);

/* Current vertex g(s, A) of one k-point (unweighted). */
double or_vertex(const or_model* m, double kx, double ky, double kz, double A, const double s[3]);

/* One k-point: Bloch vector at every sample t_i, s_out[(N+1)*3]. Returns 0 or
 * 1 if the controller failed. */
int or_trajectory(const or_model* m, const or_pulse* p, const or_time* tc, const or_solver* sv,
                  double kx, double ky, double kz, double* s_out, or_stats* st);

/* One k-point's attempted steps in every sampling interval, att_out[N]
 * (interval i = t_{i-1}..t_i at index i - 1).  Returns as or_trajectory. */
int or_trajectory_attempts(const or_model* m, const or_pulse* p, const or_time* tc, const or_solver* sv,
                           double kx, double ky, double kz, int32_t* att_out, or_stats* st);

/* Band occupations of one point of the 3D pi-flux lattice with equilibrium
 * relaxation (PiFlux::evolve_occupation_numbers, PiFlux.cpp:222-318):
 * occ_out[(N+1)*2] = (lower, upper) at every sample.  The shifted momentum of
 * the equilibrium targets persists across right-hand-side calls (its rescued
 * cos x stays once set).  Returns 0, 1 if the controller failed, -1 if the
 * model, relaxation or method is not that. */
int or_occupations(const or_model* m, const or_pulse* p, const or_time* tc, const or_solver* sv,
                   double kx, double ky, double kz, double* occ_out, or_stats* st);

/* j(t_i) = sum_k w_k g_k(t_i), i = 0..N, deterministic for any thread count
 * (kz may be null: all zero); with split_dispersion > 0, j_out holds 2 (N + 1)
 * samples (below the cut, then the rest), with perp_current 1 (honeycomb) j
 * then j_perp.  Returns the number of failed trajectories. */
int64_t or_current_density(const or_model* m, const or_pulse* p, const or_time* tc, const or_solver* sv,
                           const double* kx, const double* ky, const double* kz, const double* w, int64_t n_k,
                           double* j_out, or_stats* st, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
