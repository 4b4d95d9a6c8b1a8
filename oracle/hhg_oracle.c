/*
 * hhg_oracle.c — CPU restatement of the HigherHarmonics per-k-point Bloch solve.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle the HIP path is compared with, and the
 * CPU baseline that bench.py times.  Nothing in the product links or loads this file.
 *
 * What it restates (reference paths relative to the reference repository),
 * and what pins each part (tests/test_oracle.py, tests/test_host_abi.py):
 *  - pulses: Laser::laser_function = A0 * envelope(t) * sin(t)
 *      (sources/HHG/Laser/Laser.hpp:32-54), cos^2 envelope
 *      (sources/HHG/Laser/CosineLaser.cpp:22-26), Gaussian envelope
 *      (sources/HHG/Laser/GaussLaser.cpp:26-28), A/B pair where B is the same
 *      base pulse evaluated at t - delay (sources/HHG/Laser/DoubleLaser.hpp:63-67),
 *      continuous A0 sin t; pinned by closed forms and E = -dA/dt against
 *      finite differences;
 *  - sampled pulses through the clamped cubic spline (the function Boost's
 *      cardinal_cubic_b_spline with end slopes represents, Laser.hpp:8-11),
 *      here in slope (Hermite) form; pinned against scipy's clamped CubicSpline
 *      and by convergence; the last bits of Boost's B-spline evaluation are
 *      parity unpinned (Boost is absent);
 *  - Bloch equation ds/dt = 2 d(k - A(t)) x s in the lab frame
 *      (PiFlux: sources/HHG/Systems/PiFlux.cpp:89-94, prefactor 4t = 2 * 2t;
 *       honeycomb: sources/HHG/Systems/Honeycomb.cpp:130-138, omega = 2t(Re g, -Im g, 0));
 *      the 3D Dirac cone in its band frame (DiracSystem.cpp:143-153); pinned
 *      against scipy DOP853 on the von Neumann equation of H = d.sigma built
 *      independently (the 3D Dirac cone against the lab-frame Weyl cone);
 *  - relaxation toward the initial state (sources/HHG/Systems/DiracSystem.cpp:231-233)
 *      and, for the 2D configurations, transverse T2 dephasing about d(k - A(t));
 *      pinned by the closed-form decay without field;
 *  - the honeycomb relaxation toward the instantaneous equilibrium at k - A(t)
 *      (sources/HHG/Systems/Honeycomb.cpp:160-199); pinned by the thermal state
 *      being its own target and the post-pulse closed form;
 *  - the 3D pi-flux lattice ds/dt = 4t m x s, m = (cos kx, cos ky, cos(kz - A))
 *      (sources/HHG/Systems/PiFlux.cpp:89-94), its thermal initial state
 *      (PiFlux.cpp:519-549), its relaxation toward the instantaneous
 *      equilibrium (PiFlux.cpp:112-182, diagonal_sigma PiFlux.cpp:1037-1049)
 *      and the drivers' rescues of vanishing cosines (PiFlux.cpp:825-829,
 *      978-993); pinned by |s| of the pure state, the equal-rate and T = 0
 *      fixed points, the post-pulse closed form and the rescued point's
 *      finiteness;
 *  - the split current (PiFlux::current_per_energy, PiFlux.cpp:954-1035):
 *      pinned by the halves summing to the unsplit current and by the cuts
 *      that put everything on one side;
 *  - the occupation run (PiFlux::evolve_occupation_numbers, PiFlux.cpp:222-318)
 *      with its persistent shifted momentum: pinned by the closed form
 *      sqrt((|s| -+ s.m^)/2) of the plain trajectory and by the set of points
 *      on which the persistent rescue changes the run; the rescued cos x value
 *      itself is pinned only through that set;
 *  - Boost.Odeint runge_kutta_fehlberg78 + make_controlled + integrate_adaptive,
 *      restarted at dt = Delta / n_subdivisions every sampling interval
 *      (sources/HHG/Systems/DiracSystem.cpp:159-173, Honeycomb.cpp:145-157,
 *       PiFlux.cpp:96-110; tolerances passed by the caller); the tableau by
 *      all order conditions in exact rationals, the controller by its step
 *      sequence on problems with known error behaviour;
 *  - the quadrature-weighted current sum j(t_i) = sum_k w_k g_k(t_i)
 *      (sources/hhg_gauss.cpp:71-133, sources/HHG/Systems/Honeycomb.cpp:225-336,
 *       sources/HHG/Systems/PiFlux.cpp:794-863); pinned by exact quadrature of
 *      polynomials and the reference's own Gauss-Legendre tables;
 *  - a 4th-order (two-node Gauss) Magnus step with the closed-form SO(3)
 *      exponential, the structural analogue of the reference's Magnus path
 *      (sources/HHG/Systems/DiracSystem.cpp:176-211,
 *       sources/HHG/Systems/DiracDetail/MagnusMatrix.cpp:64-74), and the
 *      reference's moment Magnus step (MagnusMatrix.cpp:25-74, Laser.cpp:40-55);
 *      pinned by their observed orders (4, and 3 with the reference's moments).
 *
 * Boost.Odeint is a third-party dependency absent from the reference tree
 * (cmake/user_settings.cmake:3 points at Boost 1.78.0).  Its published
 * controlled_runge_kutta / default_error_checker / default_step_adjuster /
 * integrate_adaptive algorithm is restated in rkf78_* below.
 */
#include "hhg_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846264338327950288
#define OR_SQRT3 1.73205080756887729352744634150587237

/* ------------------------------------------------------------------------- */
/* Pulses                                                                     */
/* ------------------------------------------------------------------------- */

/* cos^2 envelope on (t0, t0 + 2 pi n): 0.5 (1 - cos((t - t0) / n)).
 * sources/HHG/Laser/CosineLaser.cpp:22-26, duration CosineLaser.hpp:13-15. */
static double env_cos2(double t, double t0, double n) {
    const double te = t0 + 2.0 * OR_PI * n;
    if (t <= t0 || t >= te) return 0.0;
    return 0.5 * (1.0 - cos((t - t0) / n));
}

/* Gaussian envelope exp(-(t - c)^2 / (2 sigma^2)).
 * sources/HHG/Laser/GaussLaser.cpp:26-28. */
static double env_gauss(double t, double c, double sigma) {
    const double x = t - c;
    return exp(-(x * x) / (2.0 * sigma * sigma));
}

double or_pulse_A(const or_pulse* p, double t) {
    switch (p->kind) {
    case OR_PULSE_COS2:
        return p->amp_a * env_cos2(t, p->t0, p->n_cycles) * sin(t + p->cep);
    case OR_PULSE_GAUSS:
        return p->amp_a * env_gauss(t, p->t0, p->sigma) * sin(t + p->cep);
    case OR_PULSE_DOUBLE_GAUSS:
        /* DoubleLaser.hpp:63-67: B = base laser at (t - delay). */
        return p->amp_a * env_gauss(t, p->t0, p->sigma) * sin(t + p->cep) +
               p->amp_b * env_gauss(t - p->delay, p->t0, p->sigma) * sin(t - p->delay + p->cep);
    case OR_PULSE_DOUBLE_COS2:
        return p->amp_a * env_cos2(t, p->t0, p->n_cycles) * sin(t + p->cep) +
               p->amp_b * env_cos2(t - p->delay, p->t0, p->n_cycles) * sin(t - p->delay + p->cep);
    case OR_PULSE_CONTINUOUS:
        /* ContinuousLaser: envelope 1 (sources/HHG/Laser/ContinuousLaser.cpp:7-9) */
        return p->amp_a * sin(t + p->cep);
    case OR_PULSE_SPLINE: {
        /* Laser::laser_function of a splined pulse (Laser.hpp:32-37): 0 beyond t_end */
        if (!(t >= p->s_t0) || t > p->s_t_last) return 0.0;
        const double u = (t - p->s_t0) * (1.0 / p->s_dt);
        int64_t i = (int64_t)u;
        if (i > p->n_intervals - 1) i = p->n_intervals - 1;
        const double x = u - (double)i;
        const double* c = p->coef + 4 * i;
        return c[0] + x * (c[1] + x * (c[2] + x * c[3]));
    }
    default:
        return NAN;
    }
}

/* The clamped cubic spline in its slope (Hermite) form: the slopes m_i solve
 * m_{i-1} + 4 m_i + m_{i+1} = 3 (y_{i+1} - y_{i-1}) / h with m_0 = d0,
 * m_{n-1} = d1 (C2 continuity at the inner nodes), and each interval is the
 * cubic Hermite interpolant of (y_i, m_i), (y_{i+1}, m_{i+1}).  Boost's
 * cardinal_cubic_b_spline (Laser.hpp:8-11, a third-party header absent here)
 * represents the same unique C2 interpolant with these end slopes in a
 * B-spline basis. */
int or_spline_coefficients(const double* y, int64_t n, double dt, double d0, double d1, double* coef) {
    if (n < 2 || !(dt > 0.0) || !y || !coef) return -1;
    double* m = (double*)malloc(sizeof(double) * (size_t)n);
    double* cp = (double*)malloc(sizeof(double) * (size_t)n);
    if (!m || !cp) {
        free(m);
        free(cp);
        return -1;
    }
    m[0] = d0;
    m[n - 1] = d1;
    if (n > 2) {
        /* unknowns m_1..m_{n-2}: tridiagonal (1, 4, 1), Thomas with the known end slopes moved right */
        const int64_t k = n - 2;
        for (int64_t j = 0; j < k; ++j) {
            const int64_t i = j + 1;
            double r = 3.0 * (y[i + 1] - y[i - 1]) / dt;
            if (j == 0) r -= d0;
            if (j == k - 1) r -= d1;
            if (j == 0) {
                cp[j] = 1.0 / 4.0;
                m[i] = r / 4.0;
            } else {
                const double den = 4.0 - cp[j - 1];
                cp[j] = 1.0 / den;
                m[i] = (r - m[i - 1]) / den;
            }
        }
        for (int64_t j = k - 2; j >= 0; --j) m[j + 1] -= cp[j] * m[j + 2];
    }
    for (int64_t i = 0; i + 1 < n; ++i) {
        double* c = coef + 4 * i;
        const double dy = y[i + 1] - y[i];
        c[0] = y[i];
        c[1] = dt * m[i];
        c[2] = 3.0 * dy - dt * (2.0 * m[i] + m[i + 1]);
        c[3] = -2.0 * dy + dt * (m[i] + m[i + 1]);
    }
    free(m);
    free(cp);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Gauss-Legendre nodes (Newton on P_n in long double)                       */
/* ------------------------------------------------------------------------- */

void or_gauss_legendre(int n, double* x, double* w) {
    /* Nodes are symmetric; compute the positive half by Newton's method on
     * the three-term recurrence (n+1) P_{n+1} = (2n+1) x P_n - n P_{n-1}. */
    const int m = (n + 1) / 2;
    for (int i = 0; i < m; ++i) {
        long double z = cosl((long double)OR_PI * ((long double)i + 0.75L) / ((long double)n + 0.5L));
        long double pp = 0.0L;
        for (int it = 0; it < 100; ++it) {
            long double p1 = 1.0L, p2 = 0.0L;
            for (int j = 1; j <= n; ++j) {
                long double p3 = p2;
                p2 = p1;
                p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / (long double)j;
            }
            pp = (long double)n * (z * p1 - p2) / (z * z - 1.0L);
            long double z1 = z;
            z = z1 - p1 / pp;
            if (fabsl(z - z1) <= 4.0L * LDBL_EPSILON) {
                /* one more evaluation at the converged node for the weight */
                p1 = 1.0L; p2 = 0.0L;
                for (int j = 1; j <= n; ++j) {
                    long double p3 = p2;
                    p2 = p1;
                    p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / (long double)j;
                }
                pp = (long double)n * (z * p1 - p2) / (z * z - 1.0L);
                break;
            }
        }
        const long double wi = 2.0L / ((1.0L - z * z) * pp * pp);
        /* ascending order: negative node first */
        x[i] = (double)(-z);
        w[i] = (double)wi;
        x[n - 1 - i] = (double)z;
        w[n - 1 - i] = (double)wi;
    }
    if (n & 1) x[n / 2] = 0.0;
}

/* ------------------------------------------------------------------------- */
/* Fehlberg 7(8) tableau (E. Fehlberg, NASA TR R-287, 1968, Table X)          */
/* ------------------------------------------------------------------------- */

static const double FC[13] = {0.0,       2.0 / 27.0, 1.0 / 9.0, 1.0 / 6.0, 5.0 / 12.0, 1.0 / 2.0, 5.0 / 6.0,
                              1.0 / 6.0, 2.0 / 3.0,  1.0 / 3.0, 1.0,       0.0,        1.0};

static const double FA[13][12] = {
    {0},
    {2.0 / 27.0},
    {1.0 / 36.0, 1.0 / 12.0},
    {1.0 / 24.0, 0.0, 1.0 / 8.0},
    {5.0 / 12.0, 0.0, -25.0 / 16.0, 25.0 / 16.0},
    {1.0 / 20.0, 0.0, 0.0, 1.0 / 4.0, 1.0 / 5.0},
    {-25.0 / 108.0, 0.0, 0.0, 125.0 / 108.0, -65.0 / 27.0, 125.0 / 54.0},
    {31.0 / 300.0, 0.0, 0.0, 0.0, 61.0 / 225.0, -2.0 / 9.0, 13.0 / 900.0},
    {2.0, 0.0, 0.0, -53.0 / 6.0, 704.0 / 45.0, -107.0 / 9.0, 67.0 / 90.0, 3.0},
    {-91.0 / 108.0, 0.0, 0.0, 23.0 / 108.0, -976.0 / 135.0, 311.0 / 54.0, -19.0 / 60.0, 17.0 / 6.0, -1.0 / 12.0},
    {2383.0 / 4100.0, 0.0, 0.0, -341.0 / 164.0, 4496.0 / 1025.0, -301.0 / 82.0, 2133.0 / 4100.0, 45.0 / 82.0,
     45.0 / 164.0, 18.0 / 41.0},
    {3.0 / 205.0, 0.0, 0.0, 0.0, 0.0, -6.0 / 41.0, -3.0 / 205.0, -3.0 / 41.0, 3.0 / 41.0, 6.0 / 41.0, 0.0},
    {-1777.0 / 4100.0, 0.0, 0.0, -341.0 / 164.0, 4496.0 / 1025.0, -289.0 / 82.0, 2193.0 / 4100.0, 51.0 / 82.0,
     33.0 / 164.0, 12.0 / 41.0, 0.0, 1.0},
};

/* 8th-order weights (propagated: odeint's runge_kutta_fehlberg78 has order 8,
 * error order 7) and the embedded 7th-order weights. */
static const double FB8[13] = {0.0, 0.0, 0.0, 0.0, 0.0, 34.0 / 105.0, 9.0 / 35.0, 9.0 / 35.0, 9.0 / 280.0,
                               9.0 / 280.0, 0.0, 41.0 / 840.0, 41.0 / 840.0};
static const double FB7[13] = {41.0 / 840.0, 0.0, 0.0, 0.0, 0.0, 34.0 / 105.0, 9.0 / 35.0, 9.0 / 35.0,
                               9.0 / 280.0, 9.0 / 280.0, 41.0 / 840.0, 0.0, 0.0};

void or_rkf78_tableau(double* c, double* a, double* b8, double* b7) {
    for (int i = 0; i < 13; ++i) {
        c[i] = FC[i];
        b8[i] = FB8[i];
        b7[i] = FB7[i];
        for (int j = 0; j < 13; ++j) a[i * 13 + j] = (j < 12 && j < i) ? FA[i][j] : 0.0;
    }
}

/* ------------------------------------------------------------------------- */
/* Models                                                                     */
/* ------------------------------------------------------------------------- */

/* Honeycomb nearest-neighbour vectors, sources/HHG/Systems/Honeycomb.hpp:23-27. */
static const double NN[3][2] = {{1.0, 0.0}, {-0.5, 0.5 * OR_SQRT3}, {-0.5, -0.5 * OR_SQRT3}};

/* d(q) with q = k - A e_theta.  2D Dirac: (v qx, v qy, Delta);
 * PiFlux-2D: 2t (cos qx, cos qy, 0); honeycomb: t (Re g, -Im g, 0),
 * g(q) = sum_delta exp(i q.delta) (Honeycomb.cpp:44-50). */
static void model_d(const or_model* m, double qx, double qy, double d[3]) {
    switch (m->kind) {
    case OR_MODEL_DIRAC2D:
        d[0] = m->v * qx;
        d[1] = m->v * qy;
        d[2] = m->mass;
        break;
    case OR_MODEL_PIFLUX2D:
        d[0] = 2.0 * m->hopping * cos(qx);
        d[1] = 2.0 * m->hopping * cos(qy);
        d[2] = 0.0;
        break;
    case OR_MODEL_HONEYCOMB: {
        double gr = 0.0, gi = 0.0;
        for (int n = 0; n < 3; ++n) {
            const double ph = NN[n][0] * qx + NN[n][1] * qy;
            gr += cos(ph);
            gi += sin(ph);
        }
        d[0] = m->hopping * gr;
        d[1] = -m->hopping * gi;
        d[2] = 0.0;
        break;
    }
    default:
        d[0] = d[1] = d[2] = NAN;
    }
}

/* is_zero of sources/HHG/GlobalDefinitions.hpp:61-67: biased exponent below
 * 0b01111010011, i.e. |x| < 2^-44 */
static int or_is_zero(double x) { return fabs(x) < 0x1p-44; }

/* RESCUE_TRAFO of sources/HHG/Systems/PiFlux.cpp:50 */
#define OR_RESCUE 1e-4

/* cos k_x of a 3D pi-flux point: on the uniform grid of the split current the
 * reference replaces it by RESCUE |cos k_z| where cos k_x and cos k_y vanish
 * and cos k_z < 0, for the whole run (PiFlux.cpp:978-986). */
static double piflux_cos_x(const or_model* m, double kx, double ky, double kz, int* rescued) {
    const double cx = cos(kx), cz = cos(kz);
    const int r = m->ic_rescue == 2 && or_is_zero(cx) && or_is_zero(cos(ky)) && cz < 0.0;
    if (rescued) *rescued = r;
    return r ? fabs(cz) * OR_RESCUE : cx;
}

/* The field vector d of the Bloch equation ds/dt = 2 d x s at vector
 * potential A.  2D models: d(k - A e_theta).  3D Dirac cone in the band frame
 * of the reference (sources/HHG/Systems/DiracSystem.cpp:143-153), the point
 * given as (k_z, kappa) in (kx, ky), field along z:
 * d = -(A kappa / |k|, 0, |k| - A k_z / |k|). */
static void model_field(const or_model* m, double kx, double ky, double kz, double A, double d[3]) {
    if (m->kind == OR_MODEL_PIFLUX3D) {
        /* d = 2t m, m = (cos kx, cos ky, cos(kz - A)) (PiFlux.cpp:89-94, prefactor 4t = 2 * 2t) */
        d[0] = 2.0 * m->hopping * piflux_cos_x(m, kx, ky, kz, NULL);
        d[1] = 2.0 * m->hopping * cos(ky);
        d[2] = 2.0 * m->hopping * cos(kz - A);
        return;
    }
    if (m->kind == OR_MODEL_DIRAC3D) {
        const double k = sqrt(kx * kx + ky * ky);
        d[0] = -A * ky / k;
        d[1] = 0.0;
        d[2] = -(k - A * kx / k);
        return;
    }
    model_d(m, kx - A * cos(m->theta), ky - A * sin(m->theta), d);
}

/* Fermi function, sources/HHG/GlobalDefinitions.hpp:36-41 (beta <= 0: T = 0). */
static double fermi(double e, double beta) {
    if (!(beta > 0.0)) return e != 0.0 ? (e < 0.0 ? 1.0 : 0.0) : 0.5;
    return 1.0 / (1.0 + exp(beta * e));
}

/* Thermal state of the 3D pi-flux lattice at m = (cx, cy, cz): PiFlux::ic_sigma
 * (PiFlux.cpp:519-549) with alpha^2 = f(-E_F + 2t|m|), beta^2 = f(-E_F - 2t|m|)
 * (PiFlux.cpp:420-426) and z_eps = cz + |m|. */
static void piflux3d_ic(const or_model* m, double cx, double cy, double cz, double s[3]) {
    const double disp = sqrt(cx * cx + cy * cy + cz * cz);
    const double a2 = fermi(-m->fermi_energy + 2.0 * m->hopping * disp, m->beta);
    const double b2 = fermi(-m->fermi_energy - 2.0 * m->hopping * disp, m->beta);
    const double diff = a2 - b2, prod = 2.0 * sqrt(a2 * b2);
    const double ze = cz + disp, nrm = disp * ze;
    s[0] = (diff * cx * ze + prod * (cy * cy + cz * ze)) / nrm;
    s[1] = (diff * cy * ze - prod * cx * cy) / nrm;
    s[2] = (diff * cz * ze - prod * cx * ze) / nrm;
}

/* The HIP kernels' one-exponential Fermi pair (eq_form 1), restated for its
 * CPU check: P = f(c + x), M = f(c - x) with q = exp(-beta x), w = exp(beta c),
 * P - M = w (q^2 - 1) / ((q + w)(1 + w q)), P M = q / ((q + w)(1 + w q)),
 * q^2 - 1 = (q - 1)(q + 1) from q - 1 = expm1(-beta x); outside
 * beta > 0, |beta c| <= 345 the two Fermi functions as they stand. */
static void fermi_pair_product(double c, double x, double beta, double* pm, double* pp) {
    const double y = beta * c;
    if (beta > 0.0 && fabs(y) <= 345.0) {
        const double w = exp(y);
        const double q = exp(-beta * x);
        const double r = 1.0 / ((q + w) * fma(w, q, 1.0));
        *pm = w * ((q - 1.0) * (q + 1.0)) * r;
        *pp = q * r;
    } else {
        const double fp = fermi(c + x, beta), fm = fermi(c - x, beta);
        *pm = fp - fm;
        *pp = fp * fm;
    }
}

/* The Fermi pair (P - M, P M) of P = f(c + x), M = f(c - x): form 0 as the
 * reference evaluates it (two Fermi functions), 1 in the HIP kernels' form. */
void or_fermi_pair(int form, double c, double x, double beta, double* pm, double* pp) {
    if (form == 1) {
        fermi_pair_product(c, x, beta, pm, pp);
        return;
    }
    const double fp = fermi(c + x, beta), fm = fermi(c - x, beta);
    *pm = fp - fm;
    *pp = fp * fm;
}

/* eq_form 1 of piflux3d_equilibrium: the HIP kernels' arrangement (band-frame
 * phase from G = nz g without dividing by nz, is_zero(|g|) as |G| < 2^-44 nz,
 * three divisions). */
static void piflux3d_equilibrium_product(const or_model* m, double cx, double cy, double cz, const double s_ref[3],
                                         double D[3], double O[3]) {
    if (or_is_zero(cx) && or_is_zero(cy) && cz < 0.0) {
        cx = fabs(cz) * OR_RESCUE;
        cz *= (1.0 - 0.5 * OR_RESCUE * OR_RESCUE);
    }
    const double disp = sqrt(cx * cx + cy * cy + cz * cz);
    double diff, pp;
    fermi_pair_product(-m->fermi_energy, 2.0 * m->hopping * disp, m->beta, &diff, &pp);
    const double p2 = 2.0 * sqrt(pp);
    const double z = cz + disp, nz = 2.0 * disp * z;
    const double xz = 2.0 * cx * z, yz = 2.0 * cy * z, xy = 2.0 * cx * cy;
    const double G0 = s_ref[0] * (z * z - cx * cx + cy * cy) - xy * s_ref[1] - xz * s_ref[2];
    const double G1 = s_ref[1] * (z * z + cx * cx - cy * cy) - xy * s_ref[0] - yz * s_ref[2];
    const double L = sqrt(G0 * G0 + G1 * G1);
    const int flat = L < 0x1p-44 * nz;
    const double u = flat ? 0.0 : p2 / L;
    const double prod = flat ? p2 : u * G0, imag = u * G1;
    const double inv = 1.0 / (disp * z);
    const double fd = diff * z * inv;
    D[0] = cx * fd;
    D[1] = cy * fd;
    D[2] = cz * fd;
    O[0] = (prod * (cy * cy + cz * z) + imag * (-cx * cy)) * inv;
    O[1] = (prod * (-cx * cy) + imag * (cx * cx + cz * z)) * inv;
    O[2] = (prod * (-cx * z) + imag * (-cy * z)) * inv;
}

/* Relaxation targets of the 3D pi-flux lattice at shifted m = (cx, cy, cz),
 * update_equilibrium_state of PiFlux.cpp:134-166: the diagonal part is
 * (alpha^2 - beta^2) m/|m|; the off-diagonal part keeps the band-frame phase
 * of s_ref (diagonal_sigma, PiFlux.cpp:1037-1049) with modulus 2 alpha beta. */
static void piflux3d_equilibrium(const or_model* m, double cx, double cy, double cz, const double s_ref[3],
                                 double D[3], double O[3]) {
    if (m->eq_form == 1) {
        piflux3d_equilibrium_product(m, cx, cy, cz, s_ref, D, O);
        return;
    }
    if (or_is_zero(cx) && or_is_zero(cy) && cz < 0.0) {
        cx = fabs(cz) * OR_RESCUE;
        cz *= (1.0 - 0.5 * OR_RESCUE * OR_RESCUE);
    }
    const double disp = sqrt(cx * cx + cy * cy + cz * cz);
    const double a2 = fermi(-m->fermi_energy + 2.0 * m->hopping * disp, m->beta);
    const double b2 = fermi(-m->fermi_energy - 2.0 * m->hopping * disp, m->beta);
    const double diff = a2 - b2;
    /* diagonal_sigma(s_ref, m), first two components */
    const double z = cz + disp, nz = 2.0 * disp * z;
    const double xz = 2.0 * cx * z, yz = 2.0 * cy * z, xy = 2.0 * cx * cy;
    const double g0 = (s_ref[0] * (z * z - cx * cx + cy * cy) - xy * s_ref[1] - xz * s_ref[2]) / nz;
    const double g1 = (s_ref[1] * (z * z + cx * cx - cy * cy) - xy * s_ref[0] - yz * s_ref[2]) / nz;
    const double xyl = sqrt(g0 * g0 + g1 * g1);
    const double p2 = 2.0 * sqrt(a2 * b2);
    const double prod = p2 * (or_is_zero(xyl) ? 1.0 : g0 / xyl);
    const double imag = p2 * (or_is_zero(xyl) ? 0.0 : g1 / xyl);
    const double ze = cz + disp, nrm = disp * ze;
    const double fd = diff * ze / nrm;
    D[0] = cx * fd;
    D[1] = cy * fd;
    D[2] = cz * fd;
    O[0] = (prod * (cy * cy + cz * ze) + imag * (-cx * cy)) / nrm;
    O[1] = (prod * (-cx * cy) + imag * (cx * cx + cz * ze)) / nrm;
    O[2] = (prod * (-cx * ze) + imag * (-cy * ze)) / nrm;
}

/* Honeycomb::ic_sigma (Honeycomb.cpp:355-361) at the field d = t(Re g, -Im g, 0)
 * with the occupations of Honeycomb.cpp:347-353: alpha2 = f(E_F - t|g|),
 * beta2 = f(E_F + t|g|).  Also the instantaneous equilibrium of the honeycomb
 * relaxation, evaluated at the shifted momentum (Honeycomb.cpp:169-177). */
static void honeycomb_state(const or_model* m, const double d[3], double s[3]) {
    const double e = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double g = e / m->hopping; /* |g| */
    const double a2 = fermi(m->fermi_energy - m->hopping * g, m->beta);
    const double b2 = fermi(m->fermi_energy + m->hopping * g, m->beta);
    const double diff = a2 - b2, prod = 2.0 * sqrt(a2 * b2);
    if (e == 0.0) { s[0] = s[1] = 0.0; s[2] = -prod; return; }
    /* Re g = d0 / t, Im g = -d1 / t */
    s[0] = -diff * (d[0] / m->hopping) / g;
    s[1] = diff * (-d[1] / m->hopping) / g;
    s[2] = -prod;
}

void or_initial_state(const or_model* m, double kx, double ky, double kz, double s[3]) {
    if (m->kind == OR_MODEL_PIFLUX3D) {
        /* the continuum-limit driver rescues cos k_z of the fresh momentum of
         * every positive z node (PiFlux.cpp:825-829: cos kx = cos ky = 0 there);
         * the x/y loops overwrite cos kx, so only the (1 - RESCUE^2/2) factor of
         * cos kz survives into ic_sigma.  The -z node reuses the last x/y values
         * (PiFlux.cpp:845-849) and is not rescued. */
        int rescued = 0;
        const double cx = piflux_cos_x(m, kx, ky, kz, &rescued);
        double cz = cos(kz);
        if ((m->ic_rescue == 1 && kz > 0.0 && cz < 0.0) || rescued) cz *= (1.0 - 0.5 * OR_RESCUE * OR_RESCUE);
        piflux3d_ic(m, cx, cos(ky), cz, s);
        return;
    }
    if (m->kind == OR_MODEL_DIRAC3D) {
        /* DiracSystem.cpp:155-158: alpha0 = f(E_F + |k|), beta0 = f(E_F - |k|),
         * s = (2 alpha0 beta0, 0, alpha0^2 - beta0^2) */
        const double k = sqrt(kx * kx + ky * ky);
        const double a0 = fermi(m->fermi_energy + k, m->beta);
        const double b0 = fermi(m->fermi_energy - k, m->beta);
        s[0] = 2.0 * a0 * b0;
        s[1] = 0.0;
        s[2] = a0 * a0 - b0 * b0;
        return;
    }
    double d[3];
    model_d(m, kx, ky, d);
    const double e = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (m->kind == OR_MODEL_HONEYCOMB) {
        honeycomb_state(m, d, s);
        return;
    }
    /* 2D Dirac / PiFlux: band-diagonal thermal state (n_upper - n_lower) d_hat,
     * equal to -d_hat at T = 0, E_F = 0 as in DiracSystem.cpp:155-158. */
    if (e == 0.0) { s[0] = s[1] = s[2] = 0.0; return; }
    const double nu = fermi(e - m->fermi_energy, m->beta);
    const double nl = fermi(-e - m->fermi_energy, m->beta);
    const double f = (nu - nl) / e;
    s[0] = f * d[0];
    s[1] = f * d[1];
    s[2] = f * d[2];
}

void or_rhs(const or_model* m, const or_pulse* p, double kx, double ky, double kz, const double s_ref[3],
            const double s[3], double t, double out[3]) {
    const double A = or_pulse_A(p, t);
    if (m->kind == OR_MODEL_PIFLUX3D) {
        /* PiFlux.cpp:89-94 / 169-180: 4t (m x s), then the relaxation */
        const double cx = piflux_cos_x(m, kx, ky, kz, NULL), cy = cos(ky), cz = cos(kz - A), pf = 4.0 * m->hopping;
        out[0] = (cy * s[2] - cz * s[1]) * pf;
        out[1] = (cz * s[0] - cx * s[2]) * pf;
        out[2] = (cx * s[1] - cy * s[0]) * pf;
        if (m->relax == OR_RELAX_EQUILIBRIUM) {
            double D[3], O[3];
            piflux3d_equilibrium(m, cx, cy, cz, s_ref, D, O);
            const double rd = 1.0 / m->relax_time;
            const double ro = m->relax_time_offdiag > 0.0 ? 1.0 / m->relax_time_offdiag : 0.0;
            for (int c = 0; c < 3; ++c) out[c] -= rd * (s[c] / 3.0 - D[c]) + ro * (s[c] * (2.0 / 3.0) - O[c]);
        }
        return;
    }
    const double* s_init = s_ref;
    double d[3];
    model_field(m, kx, ky, kz, A, d);
    const double wx = 2.0 * d[0], wy = 2.0 * d[1], wz = 2.0 * d[2];
    out[0] = wy * s[2] - wz * s[1];
    out[1] = wz * s[0] - wx * s[2];
    out[2] = wx * s[1] - wy * s[0];
    if (m->relax == OR_RELAX_TO_INITIAL) {
        const double r = 1.0 / m->relax_time;
        for (int c = 0; c < 3; ++c) out[c] -= (s[c] - s_init[c]) * r;
    } else if (m->relax == OR_RELAX_EQUILIBRIUM) {
        /* honeycomb: -(s - s_eq(k - A))/tau (Honeycomb.cpp:180-191) */
        const double r = 1.0 / m->relax_time;
        double se[3];
        if (m->eq_form == 1) {
            /* the HIP kernels' form: (P - M) (d0, d1) / |d|, -2 sqrt(P M), c = E_F, x = |d| */
            const double e = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            double pm, pp;
            fermi_pair_product(m->fermi_energy, e, m->beta, &pm, &pp);
            const double f = e == 0.0 ? 0.0 : pm / e;
            se[0] = f * d[0];
            se[1] = f * d[1];
            se[2] = -2.0 * sqrt(pp);
        } else {
            honeycomb_state(m, d, se);
        }
        for (int c = 0; c < 3; ++c) out[c] -= r * (s[c] - se[c]);
    } else if (m->relax == OR_RELAX_T2) {
        const double r = 1.0 / m->relax_time;
        const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double proj = dd > 0.0 ? (d[0] * s[0] + d[1] * s[1] + d[2] * s[2]) / dd : 0.0;
        for (int c = 0; c < 3; ++c) out[c] -= (s[c] - proj * d[c]) * r;
    }
}

double or_vertex(const or_model* m, double kx, double ky, double kz, double A, const double s[3]) {
    /* 3D pi-flux lattice: sin(k_z - A) s_z (PiFlux.cpp:831-838) */
    if (m->kind == OR_MODEL_PIFLUX3D) return sin(kz - A) * s[2];
    /* 3D Dirac cone: the band inversion rho = s_z (DiracSystem.cpp:167-169);
     * the weight carries k_z kappa / |k| (DiracSystem.cpp:279) */
    if (m->kind == OR_MODEL_DIRAC3D) return s[2];
    const double ct = cos(m->theta), st = sin(m->theta);
    const double qx = kx - A * ct, qy = ky - A * st;
    switch (m->kind) {
    case OR_MODEL_DIRAC2D:
        /* -(d d / d k_par) . s */
        return -m->v * (ct * s[0] + st * s[1]);
    case OR_MODEL_PIFLUX2D:
        /* reference vertex sin(k - A) s (PiFlux.cpp:831-838) along the field */
        return ct * sin(qx) * s[0] + st * sin(qy) * s[1];
    case OR_MODEL_HONEYCOMB: {
        /* shifted_grad_gamma (Honeycomb.cpp:52-61), contraction Honeycomb.cpp:273-279,
         * projection on the polarisation Honeycomb.cpp:328-334. */
        double gxr = 0, gxi = 0, gyr = 0, gyi = 0;
        for (int n = 0; n < 3; ++n) {
            const double ph = NN[n][0] * qx + NN[n][1] * qy;
            const double c = cos(ph), sn = sin(ph);
            gxr += NN[n][0] * c;
            gxi += NN[n][0] * sn;
            gyr += NN[n][1] * c;
            gyi += NN[n][1] * sn;
        }
        const double jx = gxr * s[0] + gxi * s[1];
        const double jy = gyr * s[0] + gyi * s[1];
        return ct * jx + st * jy;
    }
    default:
        return NAN;
    }
}

/* The honeycomb's j_perp vertex: the reference's check buffer_x sin(theta) +
 * buffer_y cos(theta) (Honeycomb.cpp:328-334; theta = 0 there, so it is the
 * y component of the current), with buffer_x, buffer_y the components of
 * shifted_grad_gamma contracted with (s_x, s_y) (Honeycomb.cpp:273-279). */
static double or_vertex_perp(const or_model* m, double kx, double ky, double A, const double s[3]) {
    const double ct = cos(m->theta), st = sin(m->theta);
    const double qx = kx - A * ct, qy = ky - A * st;
    double gxr = 0, gxi = 0, gyr = 0, gyi = 0;
    for (int n = 0; n < 3; ++n) {
        const double ph = NN[n][0] * qx + NN[n][1] * qy;
        const double c = cos(ph), sn = sin(ph);
        gxr += NN[n][0] * c;
        gxi += NN[n][0] * sn;
        gyr += NN[n][1] * c;
        gyi += NN[n][1] * sn;
    }
    return st * (gxr * s[0] + gxi * s[1]) + ct * (gyr * s[0] + gyi * s[1]);
}

/* ------------------------------------------------------------------------- */
/* RKF78 with the odeint controller                                           */
/* ------------------------------------------------------------------------- */

/* The shifted momentum of PiFlux::evolve_occupation_numbers (PiFlux.cpp:238,
 * 252-257): one object for the whole run, so that the rescue of its cos x,
 * once made, stays; cz is its cos z as the last update left it. */
typedef struct {
    double cx, cz;
} or_frame;

/* update_z to z - A and the rescue (PiFlux.cpp:253-257) */
static void frame_update(or_frame* f, double cy, double cz) {
    if (or_is_zero(f->cx) && or_is_zero(cy) && cz < 0.0) {
        f->cx = fabs(cz) * OR_RESCUE;
        cz *= (1.0 - 0.5 * OR_RESCUE * OR_RESCUE);
    }
    f->cz = cz;
}

typedef struct {
    const or_model* m;
    const or_pulse* p;
    double kx, ky, kz;
    double s_init[3];
    double s_step[3]; /* state at the start of the current attempt */
    double rtol, atol;
    int max_fail;
    long max_steps;
    or_stats* st;
    or_frame* frame;  /* non-null: the occupation run's persistent shifted momentum */
} or_ctx;

static void ctx_rhs(or_ctx* c, const double s[3], double t, double out[3]) {
    c->st->rhs_evals++;
    if (c->frame) {
        /* right_side of PiFlux.cpp:287-299: 4t (m x s) at the point's own m,
         * the equilibrium targets at the persistent shifted momentum */
        const or_model* m = c->m;
        const double A = or_pulse_A(c->p, t);
        const double cx = piflux_cos_x(m, c->kx, c->ky, c->kz, NULL), cy = cos(c->ky), cz = cos(c->kz - A);
        const double pf = 4.0 * m->hopping;
        out[0] = (cy * s[2] - cz * s[1]) * pf;
        out[1] = (cz * s[0] - cx * s[2]) * pf;
        out[2] = (cx * s[1] - cy * s[0]) * pf;
        frame_update(c->frame, cy, cz);
        double D[3], O[3];
        piflux3d_equilibrium(m, c->frame->cx, cy, c->frame->cz, c->s_step, D, O);
        const double rd = 1.0 / m->relax_time;
        const double ro = m->relax_time_offdiag > 0.0 ? 1.0 / m->relax_time_offdiag : 0.0;
        for (int q = 0; q < 3; ++q) out[q] -= rd * (s[q] / 3.0 - D[q]) + ro * (s[q] * (2.0 / 3.0) - O[q]);
        return;
    }
    or_rhs(c->m, c->p, c->kx, c->ky, c->kz, c->m->relax == OR_RELAX_EQUILIBRIUM ? c->s_step : c->s_init, s, t, out);
}

/* controlled_runge_kutta::try_step for an explicit error stepper: derivative at
 * the start, one embedded step, error = max_i |err_i| / (atol + rtol (|x_i| +
 * dt |dxdt_i|)), reject with dt *= max(0.9 err^(-1/(q_err - 1)), 0.2),
 * q_err = 7; accept, and for err < 0.5 grow dt *= 0.9 max(err, 5^-q)^(-1/q),
 * q = 8 (at most 4.5x).  Returns 1 on acceptance. */
static int rkf78_try_step(or_ctx* c, double s[3], double* t, double* dt) {
    double k[13][3];
    const double h = *dt;
    /* equilibrium relaxation reads the state at the start of the step inside
     * every stage (the integrator's own state variable, PiFlux.cpp:144) */
    c->s_step[0] = s[0];
    c->s_step[1] = s[1];
    c->s_step[2] = s[2];
    ctx_rhs(c, s, *t, k[0]);
    for (int i = 1; i < 13; ++i) {
        double x[3];
        for (int q = 0; q < 3; ++q) {
            double acc = 0.0;
            for (int j = 0; j < i; ++j) acc += FA[i][j] * k[j][q];
            x[q] = s[q] + h * acc;
        }
        ctx_rhs(c, x, *t + FC[i] * h, k[i]);
    }
    double snew[3], err = 0.0;
    int has_nan = 0;
    for (int q = 0; q < 3; ++q) {
        double acc = 0.0;
        for (int j = 0; j < 13; ++j) acc += FB8[j] * k[j][q];
        snew[q] = s[q] + h * acc;
        const double e = h * (41.0 / 840.0) * (k[0][q] + k[10][q] - k[11][q] - k[12][q]);
        const double sc = c->atol + c->rtol * (fabs(s[q]) + fabs(h) * fabs(k[0][q]));
        const double r = fabs(e) / sc;
        if (isnan(r))
            has_nan = 1;
        else if (r > err)
            err = r;
    }
    /* a NaN error is not > 1 and is accepted, as in odeint; the caller's
     * finiteness check then marks the trajectory failed */
    if (has_nan) err = NAN;
    c->st->attempts++;
    if (err > 1.0) {
        double f = 0.9 * pow(err, -1.0 / 6.0);
        if (f < 0.2) f = 0.2;
        *dt = h * f;
        c->st->rejects++;
        return 0;
    }
    *t += h;
    s[0] = snew[0];
    s[1] = snew[1];
    s[2] = snew[2];
    if (err < 0.5) {
        const double e8 = pow(5.0, -8.0);
        if (err < e8) err = e8;
        *dt = h * 0.9 * pow(err, -1.0 / 8.0);
    }
    return 1;
}

/* integrate_adaptive over [t0, t1] starting from step dt (by reference: on
 * return dt holds the controller's proposal; last_unclipped the proposal that
 * was in force before the final step was clipped to t1).  Loop and clip
 * tests use the absolute DBL_EPSILON exactly like odeint's less_with_sign.
 * Returns 0, or 1 when a step needed more than max_fail attempts, when an
 * accepted step left t unchanged (a step below half an ulp of t: odeint would
 * repeat it forever), or when the interval took more than max_steps accepted
 * steps. */
static int rkf78_integrate(or_ctx* c, double s[3], double t0, double t1, double* dt, double* last_unclipped) {
    double t = t0;
    double h = *dt;
    double unclipped = h;
    long steps = 0;
    while (t1 - t > DBL_EPSILON) {
        if (steps >= c->max_steps) return 1;
        unclipped = h;
        if ((t + h) - t1 > DBL_EPSILON) h = t1 - t;
        int tries = 0;
        const double t_prev = t;
        const double s_prev[3] = {s[0], s[1], s[2]};
        for (;;) {
            if (!(h > 0.0) || !isfinite(h)) return 1;
            const int ok = rkf78_try_step(c, s, &t, &h);
            ++tries;
            if (ok) break;
            if (tries > c->max_fail) return 1;
        }
        if (!(t > t_prev)) {  /* the step is discarded: the state stays where it was */
            s[0] = s_prev[0];
            s[1] = s_prev[1];
            s[2] = s_prev[2];
            return 1;
        }
        ++steps;
        if (!isfinite(s[0]) || !isfinite(s[1]) || !isfinite(s[2])) return 1;
    }
    *dt = h;
    if (last_unclipped) *last_unclipped = unclipped;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Magnus 4 (two-node Gauss) with the closed-form rotation                   */
/* ------------------------------------------------------------------------- */

static void magnus4_step(const or_model* m, const or_pulse* p, double kx, double ky, double kz, double s[3],
                         double t, double h) {
    const double c1 = 0.5 - OR_SQRT3 / 6.0, c2 = 0.5 + OR_SQRT3 / 6.0;
    double w1[3], w2[3];
    const double A1 = or_pulse_A(p, t + c1 * h), A2 = or_pulse_A(p, t + c2 * h);
    model_field(m, kx, ky, kz, A1, w1);
    model_field(m, kx, ky, kz, A2, w2);
    for (int q = 0; q < 3; ++q) { w1[q] *= 2.0; w2[q] *= 2.0; }
    /* Omega = h/2 (w1 + w2) + sqrt(3)/12 h^2 (w2 x w1) */
    const double f = OR_SQRT3 / 12.0 * h * h;
    double O[3];
    O[0] = 0.5 * h * (w1[0] + w2[0]) + f * (w2[1] * w1[2] - w2[2] * w1[1]);
    O[1] = 0.5 * h * (w1[1] + w2[1]) + f * (w2[2] * w1[0] - w2[0] * w1[2]);
    O[2] = 0.5 * h * (w1[2] + w2[2]) + f * (w2[0] * w1[1] - w2[1] * w1[0]);
    const double th2 = O[0] * O[0] + O[1] * O[1] + O[2] * O[2];
    const double th = sqrt(th2);
    double sa, sb; /* sin(th)/th, (1 - cos th)/th^2 */
    if (th < 1e-4) {
        sa = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
        sb = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
    } else {
        sa = sin(th) / th;
        const double hs = sin(0.5 * th);
        sb = 2.0 * hs * hs / th2;
    }
    double oxs[3] = {O[1] * s[2] - O[2] * s[1], O[2] * s[0] - O[0] * s[2], O[0] * s[1] - O[1] * s[0]};
    double oxoxs[3] = {O[1] * oxs[2] - O[2] * oxs[1], O[2] * oxs[0] - O[0] * oxs[2], O[0] * oxs[1] - O[1] * oxs[0]};
    for (int q = 0; q < 3; ++q) s[q] += sa * oxs[q] + sb * oxoxs[q];
}

/* ------------------------------------------------------------------------- */
/* The reference's moment-based Magnus step of the 3D Dirac cone              */
/* ------------------------------------------------------------------------- */

/* Laser::magnus_coefficients (sources/HHG/Laser/Laser.cpp:40-55): 30-point
 * Gauss-Legendre on [0, 1] (sources/HHG/Laser/gauss.hpp:14-80) of A(t0 + h x)
 * against 1 and the table's polynomials 2x - 1, 6x(x - 1) + 1 and
 * 20x^3 - 30x^2 + 12x + 1 (gauss.hpp:82-102), scaled by 1, 3, 5, 7.  The last
 * polynomial is the shifted Legendre P3 with the sign of its constant term
 * flipped; it is evaluated as the reference evaluates it. */
void or_magnus_moments(const or_pulse* p, double t0, double h, double mom[4]) {
    double xi[30], wi[30];
    or_gauss_legendre(30, xi, wi);
    mom[0] = mom[1] = mom[2] = mom[3] = 0.0;
    for (int i = 0; i < 30; ++i) {
        const double x = 0.5 * (xi[i] + 1.0), w = 0.5 * wi[i];
        const double a = w * or_pulse_A(p, t0 + h * x);
        mom[0] += a;
        mom[1] += a * (2.0 * x - 1.0);
        mom[2] += a * (6.0 * x * (x - 1.0) + 1.0);
        mom[3] += a * (20.0 * x * x * x - 30.0 * x * x + 12.0 * x + 1.0);
    }
    mom[1] *= 3.0;
    mom[2] *= 5.0;
    mom[3] *= 7.0;
}

/* DiracSystem::time_evolution_magnus (sources/HHG/Systems/DiracSystem.cpp:176-211):
 * with K = 2|k| h, Kp = 2 kappa h, Kz = 2 k_z h (Magnus(2|k|, 2 kappa, 2 k_z, h),
 * MagnusMatrix.cpp:12-23) and the moments divided by |k| (al, be, ga, de), the
 * generator Omega = [[0, a, b], [-a, 0, c], [-b, -c, 0]] has the entries of
 * MagnusMatrix.cpp:35-62 (grouped here by q = Kz^2 + Kp^2), and s <- exp(Omega) s.
 * The exponential is Rodrigues' formula about w = (-c, b, -a) with the half-angle
 * form of 1 - cos (the reference divides cos - 1 by theta^2 directly,
 * MagnusMatrix.cpp:64-74; the two agree to rounding). */
void or_magnus_moments_step(double kz, double kappa, const double mom[4], double h, double s[3]) {
    const double kk = sqrt(kz * kz + kappa * kappa);
    const double K = 2.0 * kk * h, Kp = 2.0 * kappa * h, Kz = 2.0 * kz * h;
    const double al = mom[0] / kk, be = mom[1] / kk, ga = mom[2] / kk, de = mom[3] / kk;
    const double K2 = K * K, K3 = K2 * K, K4 = K2 * K2;
    const double Kp2 = Kp * Kp, Kz2 = Kz * Kz, q = Kz2 + Kp2;
    const double al2 = al * al, al3 = al2 * al, al4 = al2 * al2, be2 = be * be, be3 = be2 * be, ga2 = ga * ga;
    const double a = al * Kz - K +
                     Kp2 * ((al2 * be2 - al3 * ga) * K * q / 2520.0 + al2 * ga * K2 * Kz / 1260.0 -
                            al * be2 * K2 * Kz / 1080.0 - al * ga * K3 / 2520.0 + be2 * K3 / 1890.0 + be2 * K / 60.0 -
                            be * de * K / 420.0 + ga2 * K / 420.0);
    const double b = K * Kp *
                     (-al4 * be * q * q + 4.0 * al3 * be * K * Kz * q - 6.0 * al2 * be * K2 * Kz2 -
                      2.0 * al2 * be * K2 * Kp2 + (-42.0 * al2 * be + 18.0 * al2 * de - 36.0 * al * be * ga + 18.0 * be3) * q +
                      (4.0 * al * be * K3 + 84.0 * al * be * K - 36.0 * al * de * K + 36.0 * be * ga * K) * Kz -
                      be * K4 - 42.0 * be * K2 - 2520.0 * be + 18.0 * de * K2) /
                     15120.0;
    const double c = Kp *
                     (3.0 * (al2 * be2 - al3 * ga) * K * Kz * q + 9.0 * al2 * ga * K2 * Kz2 + 3.0 * al2 * ga * K2 * Kp2 -
                      6.0 * al * be2 * K2 * Kz2 + al * be2 * K2 * Kp2 - 9.0 * al * ga * K3 * Kz - 7560.0 * al +
                      (3.0 * be2 * K3 + 126.0 * be2 * K - 18.0 * be * de * K + 18.0 * ga2 * K) * Kz + 3.0 * ga * K4) /
                     7560.0;
    const double O[3] = {-c, b, -a};
    const double th2 = a * a + b * b + c * c;
    const double th = sqrt(th2);
    double sa, sb;
    if (th < 1e-4) {
        sa = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
        sb = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
    } else {
        sa = sin(th) / th;
        const double hs = sin(0.5 * th);
        sb = 2.0 * hs * hs / th2;
    }
    double oxs[3] = {O[1] * s[2] - O[2] * s[1], O[2] * s[0] - O[0] * s[2], O[0] * s[1] - O[1] * s[0]};
    double oxoxs[3] = {O[1] * oxs[2] - O[2] * oxs[1], O[2] * oxs[0] - O[0] * oxs[2], O[0] * oxs[1] - O[1] * oxs[0]};
    for (int r = 0; r < 3; ++r) s[r] += sa * oxs[r] + sb * oxoxs[r];
}

/* ------------------------------------------------------------------------- */
/* Trajectories and the current sum                                           */
/* ------------------------------------------------------------------------- */

static double sample_time(const or_time* tc, int i) {
    const double delta = (tc->t_end - tc->t_begin) / tc->n_measurements;
    return tc->t_begin + (double)i * delta;
}

/* Corners of the cos^2 envelopes (A'' jumps at the support edges) that lie
 * strictly inside a sampling interval, ascending, with the interval index i
 * (t_{i-1} < t < t_i).  The carried controller integrates up to each corner
 * and continues from there; the reference's restart controller does not
 * (it has no such corners to meet with its short restart steps), so this is
 * used only with OR_PROPOSAL_CARRY.  Edges within 1e-9 Delta of a sample
 * time and edges of a zero-amplitude component are skipped.  The product
 * computes the same list with the same double operations. */
static int envelope_corners(const or_pulse* p, const or_time* tc, double* t_k, int* i_k) {
    double e[4];
    int ne = 0;
    if (p->kind != OR_PULSE_COS2 && p->kind != OR_PULSE_DOUBLE_COS2) return 0;
    if (p->amp_a != 0.0) {
        e[ne++] = p->t0;
        e[ne++] = p->t0 + 2.0 * OR_PI * p->n_cycles;
    }
    if (p->kind == OR_PULSE_DOUBLE_COS2 && p->amp_b != 0.0) {
        const double b = p->t0 + p->delay;
        e[ne++] = b;
        e[ne++] = b + 2.0 * OR_PI * p->n_cycles;
    }
    for (int a = 1; a < ne; ++a) /* insertion sort */
        for (int b = a; b > 0 && e[b] < e[b - 1]; --b) {
            const double tmp = e[b];
            e[b] = e[b - 1];
            e[b - 1] = tmp;
        }
    const double delta = (tc->t_end - tc->t_begin) / tc->n_measurements;
    int n = 0;
    for (int q = 0; q < ne; ++q) {
        const double x = (e[q] - tc->t_begin) / delta;
        if (!(x > 0.0 && x < (double)tc->n_measurements)) continue;
        const double fl = floor(x);
        if (x - fl < 1e-9 || fl + 1.0 - x < 1e-9) continue;
        if (n > 0 && t_k[n - 1] == e[q]) continue;
        t_k[n] = e[q];
        i_k[n] = (int)fl + 1;
        ++n;
    }
    return n;
}

/* Evolves one k-point and calls emit(i, s) at every sample.  Returns 1 on a
 * controller failure (the remaining samples are emitted from the frozen
 * state), 0 otherwise, or -1 for an unsupported combination. */
typedef void (*emit_fn)(void* user, int i, const double s[3]);

static int evolve(const or_model* m, const or_pulse* p, const or_time* tc, const or_solver* sv, double kx, double ky,
                  double kz, or_stats* st, emit_fn emit, void* user, or_frame* frame) {
    or_ctx c;
    c.frame = frame;
    c.m = m;
    c.p = p;
    c.kx = kx;
    c.ky = ky;
    c.kz = kz;
    c.rtol = sv->rtol;
    c.atol = sv->atol;
    c.max_fail = sv->max_fail > 0 ? sv->max_fail : 500;
    c.max_steps = sv->max_steps > 0 ? sv->max_steps : 100000;
    c.st = st;
    double s[3];
    if (m->kind == OR_MODEL_HONEYCOMB && m->relax == OR_RELAX_EQUILIBRIUM) {
        /* the relaxing honeycomb starts in the equilibrium at A(t_begin)
         * (Honeycomb.cpp:193-194) */
        const double A0 = or_pulse_A(p, tc->t_begin);
        or_initial_state(m, kx - A0 * cos(m->theta), ky - A0 * sin(m->theta), kz, s);
    } else {
        or_initial_state(m, kx, ky, kz, s);
    }
    c.s_init[0] = s[0];
    c.s_init[1] = s[1];
    c.s_init[2] = s[2];
    if (frame) {
        /* the occupation run: 3D pi-flux lattice, equilibrium relaxation, RKF78;
         * the shifted momentum starts as the point's own and is updated to
         * A(t_begin) before the integration (PiFlux.cpp:238, 285) */
        if (m->kind != OR_MODEL_PIFLUX3D || m->relax != OR_RELAX_EQUILIBRIUM || sv->method != OR_METHOD_RKF78)
            return -1;
        frame->cx = piflux_cos_x(m, kx, ky, kz, NULL);
        frame_update(frame, cos(ky), cos(kz - or_pulse_A(p, tc->t_begin)));
    }
    emit(user, 0, s);
    /* equilibrium relaxation exists for the 3D pi-flux lattice, which has no
     * other relaxation, and the honeycomb lattice */
    if (m->relax == OR_RELAX_EQUILIBRIUM && m->kind != OR_MODEL_PIFLUX3D && m->kind != OR_MODEL_HONEYCOMB) return -1;
    if (m->kind == OR_MODEL_PIFLUX3D && m->relax != OR_RELAX_NONE && m->relax != OR_RELAX_EQUILIBRIUM) return -1;
    const int N = tc->n_measurements;
    const double delta = (tc->t_end - tc->t_begin) / N;
    if (sv->method == OR_METHOD_MAGNUS_MOMENTS) {
        /* one step per sampling interval, 3D Dirac cone only, no relaxation */
        if (m->kind != OR_MODEL_DIRAC3D || m->relax != OR_RELAX_NONE) return -1;
        for (int i = 1; i <= N; ++i) {
            double mom[4];
            or_magnus_moments(p, sample_time(tc, i - 1), delta, mom);
            or_magnus_moments_step(kx, ky, mom, delta, s);
            st->attempts += 1;
            emit(user, i, s);
        }
        return 0;
    }
    if (sv->method == OR_METHOD_MAGNUS4) {
        if (m->relax != OR_RELAX_NONE) return -1;
        const int ms = sv->magnus_substeps > 0 ? sv->magnus_substeps : 1;
        const double h = delta / ms;
        for (int i = 1; i <= N; ++i) {
            const double t0 = sample_time(tc, i - 1);
            for (int q = 0; q < ms; ++q) magnus4_step(m, p, kx, ky, kz, s, t0 + q * h, h);
            st->attempts += ms;
            emit(user, i, s);
        }
        return 0;
    }
    int failed = 0;
    const double dt0 = delta / tc->n_subdivisions;
    double dt = dt0;
    const int carry = sv->proposal == OR_PROPOSAL_CARRY;
    double t_k[4];
    int i_k[4];
    const int n_k = carry ? envelope_corners(p, tc, t_k, i_k) : 0;
    int q = 0;
    for (int i = 1; i <= N; ++i) {
        if (!failed) {
            double h = carry ? dt : dt0;
            double unclipped = h;
            double ta = sample_time(tc, i - 1);
            /* carry mode continues from the proposal of the last unclipped
             * step when the final step was clipped, also across corners */
            for (; q < n_k && i_k[q] == i && !failed; ++q) {
                if (rkf78_integrate(&c, s, ta, t_k[q], &h, &unclipped)) failed = 1;
                h = h > unclipped ? h : unclipped;
                ta = t_k[q];
            }
            if (!failed && rkf78_integrate(&c, s, ta, sample_time(tc, i), &h, &unclipped)) failed = 1;
            if (!failed) dt = h > unclipped ? h : unclipped;
        }
        emit(user, i, s);
    }
    return failed;
}

typedef struct {
    double* out;
} traj_user;

static void emit_traj(void* user, int i, const double s[3]) {
    traj_user* u = (traj_user*)user;
    u->out[3 * i + 0] = s[0];
    u->out[3 * i + 1] = s[1];
    u->out[3 * i + 2] = s[2];
}

int or_trajectory(const or_model* m, const or_pulse* p, const or_time* tc, const or_solver* sv, double kx, double ky,
                  double kz, double* s_out, or_stats* st) {
    traj_user u = {s_out};
    or_stats local = {0, 0, 0, 0};
    const int r = evolve(m, p, tc, sv, kx, ky, kz, st ? st : &local, emit_traj, &u, NULL);
    if (r > 0 && st) st->failed_trajectories++;
    return r;
}

typedef struct {
    int32_t* att;
    const or_stats* st;
    int64_t seen;
} att_user;

static void emit_att(void* user, int i, const double s[3]) {
    (void)s;
    att_user* u = (att_user*)user;
    if (i > 0) u->att[i - 1] = (int32_t)(u->st->attempts - u->seen);
    u->seen = u->st->attempts;
}

int or_trajectory_attempts(const or_model* m, const or_pulse* p, const or_time* tc, const or_solver* sv, double kx,
                           double ky, double kz, int32_t* att_out, or_stats* st) {
    or_stats local = {0, 0, 0, 0};
    or_stats* use = st ? st : &local;
    att_user u = {att_out, use, use->attempts};
    const int r = evolve(m, p, tc, sv, kx, ky, kz, use, emit_att, &u, NULL);
    if (r > 0) use->failed_trajectories++;
    return r;
}

typedef struct {
    double* out;
    const or_frame* frame;
    double cy;
} occ_user;

/* compute_occupations of PiFlux.cpp:228-236: the state in the band frame of
 * the shifted momentum (diagonal_sigma, PiFlux.cpp:1037-1049), then
 * (lower, upper) = sqrt((|d| -+ d_z) / 2) */
static void emit_occ(void* user, int i, const double s[3]) {
    occ_user* u = (occ_user*)user;
    const double mx = u->frame->cx, my = u->cy, mz = u->frame->cz;
    const double disp = sqrt(mx * mx + my * my + mz * mz);
    const double z = mz + disp, nrm = 2.0 * disp * z;
    const double xz = 2.0 * mx * z, yz = 2.0 * my * z, xy = 2.0 * mx * my;
    const double d0 = (s[0] * (z * z - mx * mx + my * my) - xy * s[1] - xz * s[2]) / nrm;
    const double d1 = (s[1] * (z * z + mx * mx - my * my) - xy * s[0] - yz * s[2]) / nrm;
    const double d2 = (s[2] * (z * z - mx * mx - my * my) + xz * s[0] + yz * s[1]) / nrm;
    const double n = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    u->out[2 * i + 0] = sqrt(0.5 * (-d2 + n));
    u->out[2 * i + 1] = sqrt(0.5 * (d2 + n));
}

int or_occupations(const or_model* m, const or_pulse* p, const or_time* tc, const or_solver* sv, double kx, double ky,
                   double kz, double* occ_out, or_stats* st) {
    or_frame frame = {0.0, 0.0};
    occ_user u = {occ_out, &frame, cos(ky)};
    or_stats local = {0, 0, 0, 0};
    const int r = evolve(m, p, tc, sv, kx, ky, kz, st ? st : &local, emit_occ, &u, &frame);
    if (r > 0 && st) st->failed_trajectories++;
    return r;
}

typedef struct {
    const or_model* m;
    const double* A;  /* A(t_i) */
    double kx, ky, kz, w;
    double* row;
    int split_offset; /* 0: no split; else N + 1, the row offset of the part above the cut */
    int perp_offset;  /* 0: no j_perp; else N + 1, the row offset of the honeycomb's j_perp */
} acc_user;

static void emit_acc(void* user, int i, const double s[3]) {
    acc_user* u = (acc_user*)user;
    const double g = u->w * or_vertex(u->m, u->kx, u->ky, u->kz, u->A[i], s);
    if (u->split_offset) {
        /* PiFlux::current_per_energy (PiFlux.cpp:998-1006): the instantaneous
         * dispersion |m(k - A)| against the cut */
        const double cx = piflux_cos_x(u->m, u->kx, u->ky, u->kz, NULL), cy = cos(u->ky);
        const double cz = cos(u->kz - u->A[i]);
        const double disp = sqrt(cx * cx + cy * cy + cz * cz);
        u->row[disp < u->m->split_dispersion ? i : u->split_offset + i] += g;
        return;
    }
    if (u->perp_offset) u->row[u->perp_offset + i] += u->w * or_vertex_perp(u->m, u->kx, u->ky, u->A[i], s);
    u->row[i] += g;
}

#define OR_CHUNK 64

int64_t or_current_density(const or_model* m, const or_pulse* p, const or_time* tc, const or_solver* sv,
                           const double* kx, const double* ky, const double* kz, const double* w, int64_t n_k,
                           double* j_out, or_stats* st, int n_threads) {
    const int N = tc->n_measurements;
    const int split = m->kind == OR_MODEL_PIFLUX3D && m->split_dispersion > 0.0;
    const int perp = m->kind == OR_MODEL_HONEYCOMB && m->perp_current == 1;
    const int W = split || perp ? 2 * (N + 1) : N + 1; /* output samples */
    const int64_t n_chunks = (n_k + OR_CHUNK - 1) / OR_CHUNK;
    double* A = (double*)malloc(sizeof(double) * (size_t)(N + 1));
    double* partial = (double*)calloc((size_t)n_chunks * (size_t)W, sizeof(double));
    if (!A || !partial) {
        free(A);
        free(partial);
        return -1;
    }
    for (int i = 0; i <= N; ++i) A[i] = or_pulse_A(p, sample_time(tc, i));
    int64_t failed = 0, attempts = 0, rejects = 0, rhs = 0;
    int unsupported = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : failed, attempts, rejects, rhs, unsupported)
    for (int64_t c = 0; c < n_chunks; ++c) {
        or_stats ls = {0, 0, 0, 0};
        double* row = partial + (size_t)c * (size_t)W;
        const int64_t k_end = (c + 1) * OR_CHUNK < n_k ? (c + 1) * OR_CHUNK : n_k;
        for (int64_t k = c * OR_CHUNK; k < k_end; ++k) {
            const double z = kz ? kz[k] : 0.0;
            acc_user u = {m, A, kx[k], ky[k], z, w[k], row, split ? N + 1 : 0, perp ? N + 1 : 0};
            const int r = evolve(m, p, tc, sv, kx[k], ky[k], z, &ls, emit_acc, &u, NULL);
            if (r > 0) failed++;
            if (r < 0) unsupported++;
        }
        attempts += ls.attempts;
        rejects += ls.rejects;
        rhs += ls.rhs_evals;
    }
    /* deterministic compensated (Neumaier) sum over chunks in index order */
    for (int i = 0; i < W; ++i) {
        double sum = 0.0, comp = 0.0;
        for (int64_t c = 0; c < n_chunks; ++c) {
            const double x = partial[(size_t)c * (size_t)W + i];
            const double t = sum + x;
            if (fabs(sum) >= fabs(x))
                comp += (sum - t) + x;
            else
                comp += (x - t) + sum;
            sum = t;
        }
        j_out[i] = sum + comp;
    }
    free(A);
    free(partial);
    if (st) {
        st->attempts += attempts;
        st->rejects += rejects;
        st->rhs_evals += rhs;
        st->failed_trajectories += failed;
    }
    if (unsupported) return -1;
    return failed;
}
