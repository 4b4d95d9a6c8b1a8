/* The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only;
 * built by `make sanitize`, run by tests/test_sanitizers.py): every model,
 * pulse kind, relaxation and method on a few k-points, both proposals, the
 * occupations, the multi-threaded current sum with its split, and the
 * controller's failure paths (unreachable tolerance, no forward progress, the
 * step cap).  TEST INFRASTRUCTURE ONLY, like the oracle itself.
 * Exit status 0 when every run returns what it should. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "hhg_oracle.h"

static int failures = 0;
#define EXPECT(c)                                                                       \
    do {                                                                                \
        if (!(c)) {                                                                     \
            fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                                 \
        }                                                                               \
    } while (0)

int main(void) {
    enum { N = 48, NS = 120 };
    double y[NS], coef[4 * (NS - 1)];
    for (int i = 0; i < NS; ++i) y[i] = 0.4 * sin(0.09 * i) * exp(-0.0004 * (i - 60) * (i - 60));
    EXPECT(or_spline_coefficients(y, NS, 0.1, 0.0, 0.0, coef) == 0);
    for (int n = 1; n <= 64; ++n) {
        double x[64], w[64];
        or_gauss_legendre(n, x, w);
    }
    static double s_out[(N + 1) * 3], occ[(N + 1) * 2], j[2 * (N + 1)];
    const or_time tc = {-6.0, 6.0, N, 50};
    for (int model = 0; model <= 4; ++model)
        for (int kind = 0; kind <= 5; ++kind)
            for (int relax = 0; relax <= 3; ++relax)
                for (int method = 0; method <= 2; ++method)
                    for (int proposal = 0; proposal <= 1; ++proposal) {
                        or_model m = {0};
                        m.kind = model, m.relax = relax, m.v = 1.0, m.mass = 0.1, m.hopping = 1.0;
                        m.fermi_energy = 0.05, m.beta = 10.0, m.relax_time = 5.0, m.relax_time_offdiag = 7.0;
                        or_pulse p = {0};
                        p.kind = kind, p.amp_a = 0.6, p.amp_b = 0.3, p.n_cycles = 1.5, p.sigma = 2.0, p.t0 = -5.0;
                        p.delay = 1.0, p.cep = 0.2;
                        if (kind == 5) p.coef = coef, p.n_intervals = NS - 1, p.s_t0 = -6.0, p.s_dt = 0.1, p.s_t_last = 5.9;
                        const or_solver sv = {method, proposal, 1e-9, 1e-9, 500, 2, 0, 0};
                        or_stats st = {0, 0, 0, 0};
                        const int r = or_trajectory(&m, &p, &tc, &sv, 0.4, 0.3, 0.5, s_out, &st);
                        EXPECT(r == 0 || r == -1); /* -1: a combination the oracle does not provide */
                        if (r == 0)
                            for (int i = 0; i < 3 * (N + 1); ++i) EXPECT(isfinite(s_out[i]));
                    }
    {
        or_model m = {0};
        m.kind = 4, m.relax = 3, m.hopping = 1.0, m.relax_time = 4.0, m.relax_time_offdiag = 6.0, m.beta = 5.0;
        m.ic_rescue = 2;
        or_pulse p = {0};
        p.kind = 0, p.amp_a = 0.8, p.n_cycles = 1.5, p.t0 = -5.0;
        const or_solver sv = {0, 0, 1e-9, 1e-9, 500, 1, 0, 0};
        EXPECT(or_occupations(&m, &p, &tc, &sv, 0.0, 1.5707963267948966, 3.0, occ, NULL) == 0);
        m.split_dispersion = 0.866;
        double kx[70], ky[70], kz[70], w[70];
        for (int i = 0; i < 70; ++i) kx[i] = 0.05 * i, ky[i] = 0.3, kz[i] = -1.0 + 0.03 * i, w[i] = 1.0 / 70;
        or_stats st = {0, 0, 0, 0};
        EXPECT(or_current_density(&m, &p, &tc, &sv, kx, ky, kz, w, 70, j, &st, 4) == 0);
    }
    {
        /* failure paths: unreachable tolerance; a step below half an ulp of t; the step cap */
        or_model m = {0};
        m.kind = 0, m.v = 1.0, m.mass = 0.1;
        or_pulse p = {0};
        p.kind = 0, p.amp_a = 0.5, p.n_cycles = 2.0, p.t0 = -6.28;
        const or_time t1 = {0.3, 4.3, 64, 50};
        static double s64[65 * 3];
        const or_solver tight = {0, 0, 1e-300, 1e-300, 1, 1, 0, 0};
        EXPECT(or_trajectory(&m, &p, &t1, &tight, 0.4, 0.1, 0.0, s64, NULL) == 1);
        or_model fast = m;
        fast.v = 1.0e6;
        or_pulse cont = {0};
        cont.kind = 4, cont.amp_a = 0.5;
        const or_time far = {1.0e10, 1.0e10 + 1.0, 8, 50};
        const or_solver sv = {0, 0, 1e-10, 1e-10, 500, 1, 0, 0};
        static double s_far[9 * 3];
        EXPECT(or_trajectory(&fast, &cont, &far, &sv, 0.4, 0.2, 0.0, s_far, NULL) == 1);
        const or_solver capped = {0, 0, 1e-6, 1e-6, 500, 1, 3, 0};
        EXPECT(or_trajectory(&m, &p, &t1, &capped, 0.3, 0.2, 0.0, s64, NULL) == 1);
    }
    printf("oracle_check: %d failed expectations\n", failures);
    return failures == 0 ? 0 : 1;
}
