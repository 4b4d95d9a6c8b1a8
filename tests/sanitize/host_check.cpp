// Host-only checks of libhhg_amd's C++ host code (hhg_host.cpp) built with
// AddressSanitizer and UndefinedBehaviorSanitizer: Gauss-Legendre nodes,
// every k-grid kind, splines, every pulse kind, spectra, the reference's
// tabulated lasers, the exact accumulators and job validation on valid and
// invalid descriptions.  No GPU: none of these functions touches one.
// Exit status 0 when every expectation holds.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../higherharmonics_amd/csrc/hhg_host.hpp"

static int failures = 0;
#define EXPECT(c)                                                          \
    do {                                                                   \
        if (!(c)) {                                                        \
            std::fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                    \
        }                                                                  \
    } while (0)

static hhg_pulse_desc pulse(int kind) {
    hhg_pulse_desc p{};
    p.kind = kind;
    p.amp_a = 0.7;
    p.amp_b = 0.4;
    p.n_cycles = 3.0;
    p.sigma = 2.0;
    p.t0 = -1.0;
    p.delay = 2.5;
    p.cep = 0.3;
    return p;
}

int main() {
    for (int n : {1, 2, 3, 7, 30, 64, 255, 1024}) {
        std::vector<double> x(n), w(n);
        EXPECT(hhg_gauss_legendre(n, x.data(), w.data()) == HHG_SUCCESS);
        double sum = 0.0;
        for (double v : w) sum += v;
        EXPECT(std::fabs(sum - 2.0) < 1e-12);
    }
    std::vector<double> ex_x{0.1, -0.2, 0.3}, ex_y{0.5, 0.1, -0.4}, ex_w{1.0, 2.0, 3.0}, ex_z{0.2, 0.1, 0.0};
    hhg_kgrid_desc grids[8]{};
    grids[0].kind = HHG_KGRID_GL_BOX, grids[0].nx = 8, grids[0].ny = 6, grids[0].x0 = -3, grids[0].x1 = 3,
    grids[0].y0 = -2, grids[0].y1 = 2;
    grids[1].kind = HHG_KGRID_GL_HEX, grids[1].nx = 6, grids[1].ny = 6;
    grids[2].kind = HHG_KGRID_EXPLICIT, grids[2].kx = ex_x.data(), grids[2].ky = ex_y.data(), grids[2].w = ex_w.data(),
    grids[2].kz = ex_z.data(), grids[2].n_points = 3;
    grids[3].kind = HHG_KGRID_DIRAC_CYLINDER, grids[3].nx = 8, grids[3].ny = 5, grids[3].x1 = 3.0;
    grids[4].kind = HHG_KGRID_PIFLUX_CONTINUUM, grids[4].nx = 2, grids[4].ny = 1;
    grids[5].kind = HHG_KGRID_PIFLUX_UNIFORM, grids[5].nx = 3;
    grids[6].kind = HHG_KGRID_PIFLUX_OCCUPATION, grids[6].nx = 3;
    grids[7].kind = HHG_KGRID_DIRAC_GAUSS, grids[7].nx = 4, grids[7].ny = 3, grids[7].x1 = 2.0;
    for (const auto& g : grids) {
        const int64_t n = hhg_kgrid_size(&g);
        EXPECT(n > 0);
        std::vector<double> kx(n), ky(n), kz(n), w(n);
        EXPECT(hhg_kgrid_build(&g, kx.data(), ky.data(), kz.data(), w.data()) == HHG_SUCCESS);
        for (int64_t i = 0; i < n; ++i) EXPECT(std::isfinite(kx[i]) && std::isfinite(w[i]));
    }
    for (int n = 2; n <= 20; ++n) {
        std::vector<double> y(n), c(4 * (n - 1));
        for (int i = 0; i < n; ++i) y[i] = std::sin(0.7 * i);
        EXPECT(hhg_spline_coefficients(y.data(), n, 0.1, 0.0, 0.5, c.data()) == HHG_SUCCESS);
    }
    std::vector<double> samples(200);
    for (int i = 0; i < 200; ++i) samples[i] = 0.3 * std::sin(0.05 * i);
    std::vector<double> t(100), A(100);
    for (int i = 0; i < 100; ++i) t[i] = -5.0 + 0.2 * i;
    for (int kind = 0; kind <= 5; ++kind) {
        hhg_pulse_desc p = pulse(kind);
        if (kind == HHG_PULSE_SPLINE) p.samples = samples.data(), p.n_samples = 200, p.sample_t0 = -2.0, p.sample_dt = 0.05;
        EXPECT(hhg_pulse_evaluate(&p, t.data(), A.data(), 100) == HHG_SUCCESS);
        for (double a : A) EXPECT(std::isfinite(a));
    }
    {
        const int N = 63, pad = 8, m = pad * (N + 1) / 2 + 1;
        std::vector<double> j(N + 1), f(m), re(m), im(m);
        for (int i = 0; i <= N; ++i) j[i] = std::cos(0.3 * i);
        hhg_time_desc tc{-3.0, 5.0, N, 50};
        EXPECT(hhg_trapezoidal_spectrum(j.data(), &tc, pad, f.data(), re.data(), im.data()) == HHG_SUCCESS);
        hhg_time_desc tw{-3.0, 5.0, 64, 50};
        std::vector<double> jw(65), fw(33), rw(33), iw(33);
        EXPECT(hhg_welch_spectrum(jw.data(), &tw, fw.data(), rw.data(), iw.data()) == HHG_SUCCESS);
    }
    {
        std::vector<double> tab(1000);
        hhg_pulse_desc out{};
        double amp = 0.0;
        for (int gauss = 0; gauss <= 1; ++gauss)
            EXPECT(hhg_reference_double_laser(gauss, 1.0, 1.0, 0.3, 4.0, 0.5, HHG_ACTIVE_BOTH, tab.data(), &out, &amp) ==
                   HHG_SUCCESS);
        std::vector<double> EA(201), EB(201);
        for (int i = 0; i < 201; ++i) EA[i] = std::sin(0.1 * i) * std::exp(-0.001 * (i - 100) * (i - 100)), EB[i] = 0.5 * EA[i];
        const int64_t n = hhg_reference_experimental_laser_size(1.0);
        EXPECT(n > 0);
        std::vector<double> es(n);
        EXPECT(hhg_reference_experimental_laser(1.0, 1.0, 0.3, 0.2, HHG_ACTIVE_BOTH, EA.data(), EB.data(), 201, es.data(),
                                                n, &out, &amp) == HHG_SUCCESS);
    }
    {
        std::vector<uint64_t> a{5, 0, ~0ull, 1}, b{~0ull, 0, 7, ~0ull};
        hhg_accumulators_add(a.data(), b.data(), 2);
        std::vector<double> j(2);
        hhg_accumulators_to_double(a.data(), 2, j.data());
        EXPECT(std::isfinite(j[0]) && std::isfinite(j[1]));
    }
    {
        hhg_job_desc jd{};
        jd.model.kind = HHG_MODEL_HONEYCOMB, jd.model.hopping = 1.0;
        jd.pulse = pulse(HHG_PULSE_COS2);
        jd.pulse.cep = 0.0;
        jd.time = hhg_time_desc{0.0, 10.0, 100, 50};
        hhg_solver_desc sv{};
        sv.rtol = sv.atol = 1e-10;
        EXPECT(hhg::validate_job(jd, sv, 1, 0) == HHG_SUCCESS);
        hhg_job_desc bad = jd;
        bad.pulse.amp_a = std::nan("");
        EXPECT(hhg::validate_job(bad, sv, 1, 0) == HHG_ERROR_INVALID_ARGUMENT);
        bad = jd;
        bad.model.kind = 9;
        EXPECT(hhg::validate_job(bad, sv, 1, 0) == HHG_ERROR_INVALID_ARGUMENT);
        bad = jd;
        bad.model.theta = 0.2;
        EXPECT(hhg::validate_job(bad, sv, 1, 0) == HHG_ERROR_UNSUPPORTED);
        bad = jd;
        samples[17] = std::nan("");
        bad.pulse.kind = HHG_PULSE_SPLINE, bad.pulse.samples = samples.data(), bad.pulse.n_samples = 200,
        bad.pulse.sample_dt = 0.05;
        EXPECT(hhg::validate_job(bad, sv, 1, 0) == HHG_ERROR_INVALID_ARGUMENT);
        EXPECT(std::string(hhg_last_error_string()).find("17") != std::string::npos);
    }
    std::printf("host_check: %d failed expectations\n", failures);
    return failures == 0 ? 0 : 1;
}
