// hhg_host.cpp — host utilities of the C ABI (no GPU needed): Gauss-Legendre
// nodes, Brillouin-zone grids, host pulse evaluation and the end-corrected
// spectrum.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hhg_amd.h"
#include "hhg_host.hpp"

namespace hhg {

namespace {
thread_local std::string g_last_error;
}

hhg_status fail(hhg_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

void clear_error() { g_last_error.clear(); }

const char* last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// Gauss-Legendre nodes: Newton iteration on P_n in long double from the
// Tricomi initial guess, weights 2 / ((1 - x^2) P_n'(x)^2).  Replaces the
// Boost-generated tables of sources/HHG/thread_gauss.hpp:21-311, which stop
// at N = 960.
// ---------------------------------------------------------------------------
void gauss_legendre(int n, double* x, double* w) {
    const long double pi = 3.141592653589793238462643383279502884L;
    const int half = n / 2;
    for (int i = 1; i <= half; ++i) {
        // Tricomi: x_i ~ (1 - (n-1)/(8 n^3)) cos(pi (4i - 1) / (4n + 2)), the i-th largest root
        const long double th = pi * (4.0L * i - 1.0L) / (4.0L * n + 2.0L);
        long double z = (1.0L - (n - 1.0L) / (8.0L * n * n * n)) * cosl(th);
        long double dp = 0.0L;
        for (int it = 0; it < 64; ++it) {
            long double p0 = 1.0L, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) p0 = 1.0L, p1 = z;
            dp = n * (p0 - z * p1) / (1.0L - z * z);
            const long double dz = p1 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-19L * fabsl(z) + 1e-30L) {
                p0 = 1.0L;
                p1 = z;
                for (int k = 2; k <= n; ++k) {
                    const long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                dp = n * (p0 - z * p1) / (1.0L - z * z);
                break;
            }
        }
        const long double wi = 2.0L / ((1.0L - z * z) * dp * dp);
        x[n - i] = (double)z;
        w[n - i] = (double)wi;
        x[i - 1] = -(double)z;
        w[i - 1] = (double)wi;
    }
    if (n % 2 == 1) {
        // centre node: P_n'(0) from the recurrence of P_{n-1}(0)
        long double p0 = 1.0L, p1 = 0.0L;
        for (int k = 2; k <= n; ++k) {
            const long double p2 = (-(k - 1.0L) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        const long double dp = n * p0;  // P_n'(0) = n P_{n-1}(0)
        x[half] = 0.0;
        w[half] = (double)(2.0L / (dp * dp));
    }
}

int64_t kgrid_size(const hhg_kgrid_desc& g) {
    if (g.kind == HHG_KGRID_EXPLICIT) return g.n_points;
    if (g.kind == HHG_KGRID_DIRAC_CYLINDER) {
        if (g.nx < 2 || g.ny < 1) return 0;
        return (int64_t)(g.nx - 2) * (int64_t)g.ny;  // z = 1..nx-1 without nx/2
    }
    if (g.kind == HHG_KGRID_PIFLUX_CONTINUUM) {
        if (g.nx < 1 || g.ny < 1) return 0;
        const int64_t C = g.ny, F = 8 * C;
        return 2 * (int64_t)g.nx * (2 * F * F + 20 * C * C);
    }
    if (g.kind == HHG_KGRID_PIFLUX_UNIFORM) {
        if (g.nx < 1 || g.nx > 2048) return 0;
        return 2 * (int64_t)g.nx * g.nx * g.nx;
    }
    if (g.kind == HHG_KGRID_PIFLUX_OCCUPATION) {
        if (g.nx < 1 || g.nx > 65536) return 0;
        return (int64_t)g.nx * g.nx;
    }
    return (int64_t)g.nx * (int64_t)g.ny;
}

hhg_status kgrid_build(const hhg_kgrid_desc& g, double* kx, double* ky, double* kz, double* w) {
    const int64_t n_all = kgrid_size(g);
    if (kz) std::fill(kz, kz + std::max<int64_t>(0, n_all), 0.0);
    if (g.kind == HHG_KGRID_EXPLICIT) {
        if (g.n_points < 0 || (g.n_points > 0 && (!g.kx || !g.ky || !g.w)))
            return fail(HHG_ERROR_INVALID_ARGUMENT, "explicit k-grid needs kx, ky, w");
        std::memcpy(kx, g.kx, sizeof(double) * g.n_points);
        std::memcpy(ky, g.ky, sizeof(double) * g.n_points);
        std::memcpy(w, g.w, sizeof(double) * g.n_points);
        if (kz && g.kz) std::memcpy(kz, g.kz, sizeof(double) * g.n_points);
        return HHG_SUCCESS;
    }
    if (g.kind == HHG_KGRID_PIFLUX_CONTINUUM) {
        // PiFlux::current_density_continuum_limit (PiFlux.cpp:794-863): z = +-pi x at the nx positive nodes of
        // the 2 nx-point rule, weight w_z; improved_xy_integral (PiFlux.cpp:665-792): four Gauss patches on
        // [0, pi] x [0, pi/2] around edge = 0.35 pi, each dimension a 2N-point rule for the patch's N
        // (N_fine = 8 N_coarse, N_coarse = ny), weight w_x w_y (x_high - x_low)/2 (y_high - y_low)/2
        if (g.nx < 1 || g.ny < 1) return fail(HHG_ERROR_INVALID_ARGUMENT, "pi-flux grid needs nx, ny >= 1");
        if (!kz) return fail(HHG_ERROR_INVALID_ARGUMENT, "pi-flux grid needs the third coordinate");
        const double pi = 3.14159265358979323846, edge = 0.35 * pi;
        const int C = g.ny, F = 8 * C;
        struct Patch {
            double xl, xh;
            int nx;
            double yl, yh;
            int ny;
        };
        const Patch patches[4] = {{edge, pi - edge, F, edge, 0.5 * pi, F / 2},
                                  {0.0, pi, 3 * C, 0.0, edge, C},
                                  {0.0, edge, C, edge, 0.5 * pi, C},
                                  {pi - edge, pi, C, edge, 0.5 * pi, C}};
        std::vector<double> gz(2 * g.nx), wz(2 * g.nx);
        gauss_legendre(2 * g.nx, gz.data(), wz.data());
        int64_t k = 0;
        for (int z = g.nx; z < 2 * g.nx; ++z) {  // positive nodes
            for (int sgn = 1; sgn >= -1; sgn -= 2) {
                const double zz = sgn * pi * gz[z];
                for (const Patch& P : patches) {
                    std::vector<double> gx(2 * P.nx), wx(2 * P.nx), gy(2 * P.ny), wy(2 * P.ny);
                    gauss_legendre(2 * P.nx, gx.data(), wx.data());
                    gauss_legendre(2 * P.ny, gy.data(), wy.data());
                    const double tw = 0.5 * (P.xh - P.xl) * 0.5 * (P.yh - P.yl);
                    for (int i = 0; i < 2 * P.nx; ++i)
                        for (int j = 0; j < 2 * P.ny; ++j, ++k) {
                            kx[k] = 0.5 * (P.xh - P.xl) * gx[i] + 0.5 * (P.xh + P.xl);
                            ky[k] = 0.5 * (P.yh - P.yl) * gy[j] + 0.5 * (P.yh + P.yl);
                            kz[k] = zz;
                            w[k] = wz[z] * (wx[i] * tw) * wy[j];
                        }
                }
            }
        }
        return HHG_SUCCESS;
    }
    if (g.kind == HHG_KGRID_PIFLUX_UNIFORM) {
        // PiFlux::current_per_energy (PiFlux.cpp:954-1035): k = (x, y, z) pi/N for i = z N^2 + y N + x,
        // each followed by its -z copy (k.invert()), weight 1/N^3 (the final division)
        if (g.nx < 1 || g.nx > 2048) return fail(HHG_ERROR_INVALID_ARGUMENT, "uniform pi-flux grid needs 1 <= nx <= 2048");
        if (!kz) return fail(HHG_ERROR_INVALID_ARGUMENT, "pi-flux grid needs the third coordinate");
        const int64_t N = g.nx;
        const double pi = 3.14159265358979323846, d = pi / (double)N, wt = 1.0 / (double)(N * N * N);
        int64_t k = 0;
        for (int64_t i = 0; i < N * N * N; ++i) {
            const int64_t z = i / (N * N), y = (i / N) % N, x = i % N;
            for (int sgn = 1; sgn >= -1; sgn -= 2, ++k) {
                kx[k] = x * d;
                ky[k] = y * d;
                kz[k] = sgn * (d * z);
                w[k] = wt;
            }
        }
        return HHG_SUCCESS;
    }
    if (g.kind == HHG_KGRID_PIFLUX_OCCUPATION) {
        // PiFlux::compute_occupation_numbers (PiFlux.cpp:909-952): k = (x pi/N, pi/2, z pi/N)
        // for i = x N + z; the weight 1/N^2 is not used by the reference (it stores per-k maps)
        if (g.nx < 1 || g.nx > 65536)
            return fail(HHG_ERROR_INVALID_ARGUMENT, "pi-flux occupation grid needs 1 <= nx <= 65536");
        if (!kz) return fail(HHG_ERROR_INVALID_ARGUMENT, "pi-flux grid needs the third coordinate");
        const int64_t N = g.nx;
        const double pi = 3.14159265358979323846, d = pi / (double)N, wt = 1.0 / (double)(N * N);
        for (int64_t i = 0; i < N * N; ++i) {
            kx[i] = (double)(i / N) * d;
            ky[i] = 0.5 * pi;
            kz[i] = d * (double)(i % N);
            w[i] = wt;
        }
        return HHG_SUCCESS;
    }
    if (g.kind == HHG_KGRID_DIRAC_GAUSS) {
        // sources/hhg_gauss.cpp:71-133: Gauss-Legendre in k_z (x1 x_z, both signs) and in kappa
        // (kappa_max (x_r + 1) / 2, DiracSystem.cpp:257-267), integrand weight k_z kappa / |k|
        // (hhg_gauss.cpp:24-26), with the substitutions' Jacobians x1 and kappa_max / 2
        if (g.nx < 1 || g.ny < 1 || !(g.x1 > 0.0))
            return fail(HHG_ERROR_INVALID_ARGUMENT, "Dirac Gauss grid needs nx, ny >= 1, x1 (band width) > 0");
        std::vector<double> gz(g.nx), wz(g.nx), gk(g.ny), wk(g.ny);
        gauss_legendre(g.nx, gz.data(), wz.data());
        gauss_legendre(g.ny, gk.data(), wk.data());
        const double max_k = g.x1;
        int64_t k = 0;
        for (int z = 0; z < g.nx; ++z) {
            const double kz = max_k * gz[z];
            const double kmax = std::sqrt(max_k * max_k - kz * kz);
            for (int r = 0; r < g.ny; ++r, ++k) {
                const double ka = 0.5 * kmax * (gk[r] + 1.0);
                kx[k] = kz;
                ky[k] = ka;
                w[k] = max_k * wz[z] * 0.5 * kmax * wk[r] * kz * ka / std::sqrt(kz * kz + ka * ka);
            }
        }
        return HHG_SUCCESS;
    }
    if (g.kind == HHG_KGRID_DIRAC_CYLINDER) {
        // DiracSystem::compute_current_density (DiracSystem.cpp:280-306): k_z = (z - nx/2) dk,
        // z = 1..nx-1 except nx/2, dk = 2 max_k / nx; kappa on (0, sqrt(max_k^2 - k_z^2))
        // (DiracSystem.cpp:259-263) by ny-point Gauss-Legendre (the reference refines an adaptive
        // trapezoid there); integrand weight k_z kappa / |k| (DiracSystem.cpp:279)
        if (g.nx < 4 || g.ny < 1 || !(g.x1 > 0.0))
            return fail(HHG_ERROR_INVALID_ARGUMENT, "Dirac cylinder grid needs nx >= 4, ny >= 1, x1 (band width) > 0");
        std::vector<double> gk(g.ny), wk(g.ny);
        gauss_legendre(g.ny, gk.data(), wk.data());
        const double max_k = g.x1, dz = 2.0 * max_k / g.nx;
        int64_t k = 0;
        for (int z = 1; z < g.nx; ++z) {
            if (z == g.nx / 2) continue;
            const double kz = (z - g.nx / 2) * dz;
            const double kmax = std::sqrt(max_k * max_k - kz * kz);
            for (int r = 0; r < g.ny; ++r, ++k) {
                const double ka = 0.5 * kmax * (gk[r] + 1.0);
                kx[k] = kz;
                ky[k] = ka;
                w[k] = dz * 0.5 * kmax * wk[r] * kz * ka / std::sqrt(kz * kz + ka * ka);
            }
        }
        return HHG_SUCCESS;
    }
    if (g.nx <= 0 || g.ny <= 0) return fail(HHG_ERROR_INVALID_ARGUMENT, "k-grid needs nx, ny > 0");
    std::vector<double> gx(g.nx), wx(g.nx), gy(g.ny), wy(g.ny);
    gauss_legendre(g.nx, gx.data(), wx.data());
    gauss_legendre(g.ny, gy.data(), wy.data());
    if (g.kind == HHG_KGRID_GL_BOX) {
        if (!(g.x1 > g.x0) || !(g.y1 > g.y0)) return fail(HHG_ERROR_INVALID_ARGUMENT, "k-grid box must have x1 > x0, y1 > y0");
        const double hx = 0.5 * (g.x1 - g.x0), cx = 0.5 * (g.x1 + g.x0);
        const double hy = 0.5 * (g.y1 - g.y0), cy = 0.5 * (g.y1 + g.y0);
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j) {
                const int64_t k = (int64_t)i * g.ny + j;
                kx[k] = hx * gx[i] + cx;
                ky[k] = hy * gy[j] + cy;
                w[k] = (wx[i] * hx) * (wy[j] * hy);
            }
        return HHG_SUCCESS;
    }
    if (g.kind == HHG_KGRID_GL_HEX) {
        // Honeycomb.cpp:229-268: x on [-2pi/3, 2pi/3], y on +-(4pi/(3 sqrt3) - |x|/sqrt3)
        const double pi = 3.14159265358979323846;
        const double sqrt3 = 1.73205080756887729353;
        const double hx = 2.0 * pi / 3.0;
        for (int i = 0; i < g.nx; ++i) {
            const double x = hx * gx[i];
            const double ymax = 4.0 * pi / (3.0 * sqrt3) - std::fabs(x) / sqrt3;
            for (int j = 0; j < g.ny; ++j) {
                const int64_t k = (int64_t)i * g.ny + j;
                kx[k] = x;
                ky[k] = ymax * gy[j];
                w[k] = (hx * wx[i]) * (ymax * wy[j]);
            }
        }
        return HHG_SUCCESS;
    }
    return fail(HHG_ERROR_INVALID_ARGUMENT, "unknown k-grid kind");
}

// ---------------------------------------------------------------------------
// Host pulse evaluation (Laser::laser_function, sources/HHG/Laser/Laser.hpp:32).
// ---------------------------------------------------------------------------
static double env_cos2_h(const hhg_pulse_desc& p, double t) {
    const double te = p.t0 + 2.0 * 3.14159265358979323846 * p.n_cycles;
    if (t <= p.t0 || t >= te) return 0.0;
    return 0.5 * (1.0 - std::cos((t - p.t0) / p.n_cycles));
}

static double env_gauss_h(const hhg_pulse_desc& p, double t) {
    const double x = t - p.t0;
    return std::exp(-(x * x) / (2.0 * p.sigma * p.sigma));
}

// Clamped cubic spline through n >= 2 samples spaced h with end slopes d0, d1:
// the C2 piecewise cubic that the reference's Boost cardinal_cubic_b_spline
// represents in a B-spline basis (Laser.hpp:8-11).  Solved here for the second
// derivatives M_i (tridiagonal, diagonally dominant, Thomas elimination):
//   2 M_0 + M_1 = 6/h ((y_1 - y_0)/h - d0)
//   M_{i-1} + 4 M_i + M_{i+1} = 6/h^2 (y_{i+1} - 2 y_i + y_{i-1})
//   M_{n-2} + 2 M_{n-1} = 6/h (d1 - (y_{n-1} - y_{n-2})/h)
// then stored per interval as a cubic in x = (t - t_i)/h.
hhg_status spline_coefficients(const double* y, int64_t n, double h, double d0, double d1, double* coef) {
    if (n < 2 || !(h > 0.0)) return fail(HHG_ERROR_INVALID_ARGUMENT, "spline needs n >= 2 samples and dt > 0");
    std::vector<double> diag(n), rhs(n), M(n);
    const double s6 = 6.0 / h;
    for (int64_t i = 0; i < n; ++i) {
        diag[i] = (i == 0 || i == n - 1) ? 2.0 : 4.0;
        if (i == 0)
            rhs[i] = s6 * ((y[1] - y[0]) / h - d0);
        else if (i == n - 1)
            rhs[i] = s6 * (d1 - (y[n - 1] - y[n - 2]) / h);
        else
            rhs[i] = s6 * ((y[i + 1] - 2.0 * y[i] + y[i - 1]) / h);
    }
    // off-diagonals are 1: forward elimination, back substitution
    for (int64_t i = 1; i < n; ++i) {
        const double f = 1.0 / diag[i - 1];
        diag[i] -= f;
        rhs[i] -= f * rhs[i - 1];
    }
    M[n - 1] = rhs[n - 1] / diag[n - 1];
    for (int64_t i = n - 2; i >= 0; --i) M[i] = (rhs[i] - M[i + 1]) / diag[i];
    const double h2 = h * h;
    for (int64_t i = 0; i + 1 < n; ++i) {
        double* c = coef + 4 * i;
        c[0] = y[i];
        c[1] = (y[i + 1] - y[i]) - h2 * (2.0 * M[i] + M[i + 1]) / 6.0;
        c[2] = 0.5 * h2 * M[i];
        c[3] = h2 * (M[i + 1] - M[i]) / 6.0;
    }
    return HHG_SUCCESS;
}

SplinePulse::SplinePulse(const hhg_pulse_desc& p) {
    n = p.n_samples - 1;
    coef.assign(4 * std::max<int64_t>(n, 0), 0.0);
    status = spline_coefficients(p.samples, p.n_samples, p.sample_dt, p.slope_begin, p.slope_end, coef.data());
    t0 = p.sample_t0;
    inv_h = 1.0 / p.sample_dt;
    const double last = p.sample_t0 + (double)n * p.sample_dt;
    t_last = p.t_cut > p.sample_t0 ? std::min(p.t_cut, last) : last;
}

// the device's spline_A with the same operations
double SplinePulse::operator()(double t) const {
    if (!(t >= t0) || t > t_last) return 0.0;
    const double u = (t - t0) * inv_h;
    int64_t i = (int64_t)u;
    i = i < n - 1 ? i : n - 1;
    const double x = u - (double)i;
    const double* c = coef.data() + 4 * i;
    return c[0] + x * (c[1] + x * (c[2] + x * c[3]));
}

// ---------------------------------------------------------------------------
// The reference's tabulated lasers in its own units.  The experiment's layout
// and units are data of the reference (ExperimentParams.hpp:5-13) and are kept
// as it defines them.
// ---------------------------------------------------------------------------
namespace ref_exp {
constexpr double exp_dt = 0.03335555;  // ps between measurements
constexpr int N_extra = 16, N_experiment = 201, begin_zero = 8, zero_buff = 8;
constexpr double laser_end = 6.67111 + (N_extra + begin_zero + zero_buff) * exp_dt;  // ps
constexpr double exp_photon_energy = 6.091420442472572;                                // meV
constexpr double unified_t_max = 1.5 * laser_end;                                      // ps
constexpr double hbar = 6.582119569509065698e-13;                                      // meV s
}  // namespace ref_exp

static double add_active(int active, double a, double b) {
    return active == HHG_ACTIVE_BOTH ? a + b : (active == HHG_ACTIVE_A ? a : b);
}

static hhg_pulse_desc spline_desc(const double* samples, int64_t n, double dt, double t_cut) {
    hhg_pulse_desc d{};
    d.kind = HHG_PULSE_SPLINE;
    d.samples = samples;
    d.n_samples = n;
    d.sample_t0 = 0.0;
    d.sample_dt = dt;
    d.t_cut = t_cut;
    return d;
}

// DoubleLaser<CosineLaser | GaussLaser> (DoubleLaser.hpp:8-74): A = base pulse of
// field 1.6 E0, B = 0.7 E0 at t - shift, 1000 samples over [0, 1.5 duration]
hhg_status reference_double_laser(int gauss, double photon_energy, double E0, double ratio, double n_cycles,
                                  double delay_ps, int active, double* samples, hhg_pulse_desc* out, double* amp) {
    using namespace ref_exp;
    if (!(photon_energy > 0.0) || !(n_cycles > 0.0) || active < 0 || active > 2 || !samples || !out)
        return fail(HHG_ERROR_INVALID_ARGUMENT, "double laser needs photon_energy > 0, n_cycles > 0, a known "
                                                "active pulse and output buffers");
    constexpr double pi = 3.14159265358979323846;
    const double e_ph = exp_photon_energy * photon_energy;
    const double duration = gauss ? 2.0 * pi * 3.0 * n_cycles : 2.0 * pi * n_cycles;  // base_duration
    const double t_max = 1.5 * duration;
    constexpr int N = 1000;
    const double dt = t_max / (N - 1);
    const double shift = photon_energy * (delay_ps * exp_photon_energy / (1e12 * hbar));
    hhg_pulse_desc a{};
    a.kind = gauss ? HHG_PULSE_GAUSS : HHG_PULSE_COS2;
    a.amp_a = ratio * 1e11 * (E0 * 1.6) / e_ph;  // Laser.cpp:24-25
    if (gauss) {
        a.sigma = pi * n_cycles / std::sqrt(2.0 * std::log(1.0 / 0.05));  // GaussLaser.cpp:5-7
        a.t0 = 0.5 * duration;
    } else {
        a.n_cycles = n_cycles;
    }
    hhg_pulse_desc b = a;
    b.amp_a = ratio * 1e11 * (E0 * 0.7) / e_ph;
    for (int i = 0; i < N; ++i) {
        const double t = i * dt;
        samples[i] = add_active(active, pulse_eval(a, t), pulse_eval(b, t - shift));
    }
    *out = spline_desc(samples, N, dt, t_max);
    if (amp) *amp = ratio * 1e11 * E0 / e_ph;
    return HHG_SUCCESS;
}

// ExperimentalLaser::vector_potential (ExperimentalLaser.cpp:28-58): A = -sum E dt
// after begin_zero zeros, a quartic tail that returns A smoothly to 0 over
// N_extra - 1 steps, zero_buff + 1 zeros
static std::vector<double> experiment_vector_potential(const double* E, double dt) {
    using namespace ref_exp;
    std::vector<double> ret(N_experiment + N_extra + zero_buff + begin_zero, 0.0);
    for (int i = 0; i < N_experiment; ++i) ret[i + begin_zero] = ret[i + begin_zero - 1] - E[i] * dt;
    const double last = ret[N_experiment + begin_zero - 1];
    const double prime = -E[N_experiment - 1];
    const double primeprime = 0.5 * (E[N_experiment - 2] - E[N_experiment - 1]) / dt;
    const double end = (N_extra - 1) * dt;
    const double third =
        -4.0 * (primeprime / (2.0 * end) + (3.0 * prime) / (4.0 * end * end) + last / (end * end * end));
    const double fourth = -0.25 * (3.0 * third / end + 2.0 * primeprime / (end * end) + prime / (end * end * end));
    for (int i = 1; i < N_extra; ++i) {
        const double t = i * dt;
        ret[N_experiment + begin_zero - 1 + i] =
            last + prime * t + primeprime * t * t + third * t * t * t + fourth * t * t * t * t;
    }
    for (size_t i = ret.size() - begin_zero - 1; i < ret.size(); ++i) ret[i] = 0.0;
    return ret;
}

int64_t reference_experimental_size(double photon_energy) {
    using namespace ref_exp;
    if (!(photon_energy > 0.0)) return -1;
    const double e_ph = photon_energy * exp_photon_energy;
    const double t_end = unified_t_max * (e_ph / (1e12 * hbar));
    const double dt = e_ph * (exp_dt / (1e12 * hbar));
    return (int64_t)(t_end / dt) + 2;  // N_temp + 1 with N_temp = t_end / dt + 1
}

hhg_status reference_experimental_laser(double photon_energy, double E0, double ratio, double shift_ps, int active,
                                        const double* E_A, const double* E_B, int n_measured, double* samples,
                                        int64_t n_samples, hhg_pulse_desc* out, double* amp) {
    using namespace ref_exp;
    if (!(photon_energy > 0.0) || active < 0 || active > 2 || !E_A || !E_B || !samples || !out)
        return fail(HHG_ERROR_INVALID_ARGUMENT, "experimental laser needs photon_energy > 0, a known active "
                                                "pulse, both fields and output buffers");
    if (n_measured != N_experiment)
        return fail(HHG_ERROR_INVALID_ARGUMENT, "experimental laser needs the experiment's 201 field samples");
    if (n_samples != reference_experimental_size(photon_energy))
        return fail(HHG_ERROR_INVALID_ARGUMENT, "experimental laser: n_samples must be "
                                                "hhg_reference_experimental_laser_size(photon_energy)");
    const double e_ph = photon_energy * exp_photon_energy;
    const double to_unitless = e_ph / (1e12 * hbar);
    const double t_end = unified_t_max * to_unitless;
    const double dt = e_ph * (exp_dt / (1e12 * hbar));
    const double unitless_laser_end = e_ph * (laser_end / (1e12 * hbar));
    const double shift = shift_ps * to_unitless;
    const std::vector<double> va = experiment_vector_potential(E_A, exp_dt);
    const std::vector<double> vb = experiment_vector_potential(E_B, exp_dt);
    const SplinePulse sa(spline_desc(va.data(), (int64_t)va.size(), dt, 0.0));
    const SplinePulse sb(spline_desc(vb.data(), (int64_t)vb.size(), dt, 0.0));
    if (sa.status != HHG_SUCCESS) return sa.status;
    if (sb.status != HHG_SUCCESS) return sb.status;
    const double scale = (0.1 / hbar) * E0 * ratio;  // (MV ps / cm) -> Peierls phase, ExperimentalLaser.cpp:220-225
    for (int64_t i = 0; i < n_samples; ++i) {
        const double t = 0.0 + dt * (double)i;
        double a, b;
        if (shift >= 0.0) {
            a = t <= unitless_laser_end ? sa(t) : 0.0;
            b = (t - shift >= 0.0 && t - shift <= unitless_laser_end) ? sb(t - shift) : 0.0;
        } else {
            a = (t + shift >= 0.0 && t + shift <= unitless_laser_end) ? sa(t + shift) : 0.0;
            b = t <= unitless_laser_end ? sb(t) : 0.0;
        }
        samples[i] = scale * add_active(active, a, b);
    }
    *out = spline_desc(samples, n_samples, dt, t_end);
    if (amp) *amp = ratio * 1e11 * E0 / e_ph;
    return HHG_SUCCESS;
}

double pulse_eval(const hhg_pulse_desc& p, double t) {
    switch (p.kind) {
    case HHG_PULSE_COS2: return p.amp_a * env_cos2_h(p, t) * std::sin(t + p.cep);
    case HHG_PULSE_GAUSS: return p.amp_a * env_gauss_h(p, t) * std::sin(t + p.cep);
    case HHG_PULSE_DOUBLE_GAUSS: {
        const double tb = t - p.delay;
        return p.amp_a * env_gauss_h(p, t) * std::sin(t + p.cep) + p.amp_b * env_gauss_h(p, tb) * std::sin(tb + p.cep);
    }
    case HHG_PULSE_DOUBLE_COS2: {
        const double tb = t - p.delay;
        return p.amp_a * env_cos2_h(p, t) * std::sin(t + p.cep) + p.amp_b * env_cos2_h(p, tb) * std::sin(tb + p.cep);
    }
    case HHG_PULSE_CONTINUOUS: return p.amp_a * std::sin(t + p.cep);  // ContinuousLaser.cpp:7-9
    }
    return NAN;
}

// ---------------------------------------------------------------------------
// FFT: iterative radix-2 for powers of two, Bluestein's chirp-z for any n.
// ---------------------------------------------------------------------------
using cplx = std::complex<double>;

static void fft_pow2(std::vector<cplx>& a, bool inverse) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    const double sgn = inverse ? 1.0 : -1.0;
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t half = len / 2;
        // twiddles computed directly per index to avoid recurrence drift
        std::vector<cplx> tw(half);
        for (size_t k = 0; k < half; ++k) {
            const double ang = sgn * 2.0 * 3.14159265358979323846 * (double)k / (double)len;
            tw[k] = cplx(std::cos(ang), std::sin(ang));
        }
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < half; ++k) {
                const cplx u = a[i + k], v = a[i + k + half] * tw[k];
                a[i + k] = u + v;
                a[i + k + half] = u - v;
            }
    }
}

// X_k = sum_n x_n exp(-2 pi i k n / n_total), k = 0..n_total-1
static std::vector<cplx> dft_any(const std::vector<double>& x) {
    const size_t n = x.size();
    size_t m = 1;
    while (m < 2 * n - 1) m <<= 1;
    // chirp c_k = exp(-i pi k^2 / n), with k^2 reduced mod 2n exactly
    std::vector<cplx> chirp(n);
    for (size_t k = 0; k < n; ++k) {
        const unsigned long long kk = ((unsigned long long)k * (unsigned long long)k) % (2ull * n);
        const double ang = -3.14159265358979323846 * (double)kk / (double)n;
        chirp[k] = cplx(std::cos(ang), std::sin(ang));
    }
    std::vector<cplx> a(m, 0.0), b(m, 0.0);
    for (size_t k = 0; k < n; ++k) a[k] = x[k] * chirp[k];
    b[0] = std::conj(chirp[0]);
    for (size_t k = 1; k < n; ++k) b[k] = b[m - k] = std::conj(chirp[k]);
    fft_pow2(a, false);
    fft_pow2(b, false);
    for (size_t k = 0; k < m; ++k) a[k] *= b[k];
    fft_pow2(a, true);
    std::vector<cplx> X(n);
    for (size_t k = 0; k < n; ++k) X[k] = a[k] / (double)m * chirp[k];
    return X;
}

// TrapezoidalFFT::compute (sources/HHG/Fourier/TrapezoidalFFT.cpp:11-49) on j
// zero-padded to zero_padding (N+1) samples (sources/hhg.cpp:277-282).  The
// reference's conventions are kept: the end correction reads the last element
// of the padded input, the transform term carries delta_t twice and the phase
// factor is exp(+i w t_begin) next to a forward exp(-i ...) transform.
hhg_status trapezoidal_spectrum(const double* j, const hhg_time_desc& tc, int zero_padding, double* freq, double* re,
                                double* im) {
    if (tc.n_measurements <= 0 || zero_padding <= 0 || !(tc.t_end > tc.t_begin))
        return fail(HHG_ERROR_INVALID_ARGUMENT, "spectrum needs n_measurements > 0, zero_padding > 0, t_end > t_begin");
    const int64_t n_in = (int64_t)zero_padding * (tc.n_measurements + 1);
    std::vector<double> x(n_in, 0.0);
    for (int i = 0; i <= tc.n_measurements; ++i) x[i] = j[i];
    const std::vector<cplx> X = dft_any(x);
    const int64_t n_fft = n_in / 2 + 1;
    const double delta = (tc.t_end - tc.t_begin) / tc.n_measurements;
    const double tiny = 5.684341886080802e-14;  // is_zero threshold, GlobalDefinitions.hpp:57-71
    for (int64_t i = 0; i < n_fft; ++i) {
        const double f = 2.0 * (double)i * 3.14159265358979323846 / ((double)n_in * delta);
        const double th = f * delta;
        double W;
        cplx a0;
        if (std::fabs(th) < tiny) {
            W = 1.0;
            a0 = cplx(-0.5, 0.0);
        } else {
            W = 2.0 * (1.0 - std::cos(th)) / (th * th);
            a0 = cplx(-(1.0 - std::cos(th)), th - std::sin(th)) / (th * th);
        }
        const cplx corr = a0 * x[0] + std::conj(a0) * std::polar(1.0, f * (tc.t_end - tc.t_begin)) * x[n_in - 1];
        const cplx out = std::polar(delta * W, f * tc.t_begin) * X[i] + corr;
        freq[i] = f;
        re[i] = delta * out.real();
        im[i] = delta * out.imag();
    }
    return HHG_SUCCESS;
}

// The continuous-laser spectrum of sources/hhg.cpp:264-275: the first N samples
// of j times the Welch window 1 - (i - N/2)^2 / (N/2)^2 (WelchWindow.hpp:11-13),
// real-to-complex DFT X_k = sum_i x_i exp(-2 pi i k i / N), k = 0..N/2 (FFTW r2c,
// FFT.cpp:8-42), and the reference's frequency axis k / Delta (hhg.cpp:272-274).
hhg_status welch_spectrum(const double* j, const hhg_time_desc& tc, double* freq, double* re, double* im) {
    const int N = tc.n_measurements;
    if (N < 2 || !(tc.t_end > tc.t_begin))
        return fail(HHG_ERROR_INVALID_ARGUMENT, "spectrum needs n_measurements >= 2, t_end > t_begin");
    std::vector<double> x(N);
    for (int i = 0; i < N; ++i) {
        const double u = (double)i - 0.5 * N;
        x[i] = j[i] * (1.0 - u * u / (0.25 * (double)N * (double)N));
    }
    const std::vector<cplx> X = dft_any(x);
    const double delta = (tc.t_end - tc.t_begin) / N;
    for (int k = 0; k <= N / 2; ++k) {
        freq[k] = (double)k / delta;
        re[k] = X[k].real();
        im[k] = X[k].imag();
    }
    return HHG_SUCCESS;
}

}  // namespace hhg

// ---------------------------------------------------------------------------
// C ABI: host utilities
// ---------------------------------------------------------------------------
namespace hhg {

// Checks one job description against the solver before any device work: every
// refusal is INVALID_ARGUMENT (bad or non-finite values) or UNSUPPORTED (a
// combination not provided), with a message naming the job.
hhg_status validate_job(const hhg_job_desc& jd, const hhg_solver_desc& sv, int n_kgrids, int idx) {
    const std::string at = "job " + std::to_string(idx) + ": ";
    const hhg_model_desc& m = jd.model;
    const hhg_pulse_desc& p = jd.pulse;
    const hhg_time_desc& t = jd.time;
    if (jd.kgrid < 0 || jd.kgrid >= n_kgrids) return fail(HHG_ERROR_INVALID_ARGUMENT, at + "k-grid index out of range");
    // every number that enters a right-hand side or a pulse must be finite (a
    // NaN or infinity would only surface as non-finite trajectories)
    const double finite_needed[] = {m.v, m.mass, m.hopping, m.theta, m.fermi_energy, m.beta, m.relax_time,
                                    m.relax_time_offdiag, m.split_dispersion, p.amp_a, p.amp_b, p.n_cycles, p.sigma,
                                    p.t0, p.delay, p.cep, t.t_begin, t.t_end};
    for (double x : finite_needed)
        if (!std::isfinite(x))
            return fail(HHG_ERROR_INVALID_ARGUMENT, at + "model, pulse and time parameters must be finite");
    if (p.kind == HHG_PULSE_SPLINE && p.samples && p.n_samples > 0 && p.n_samples <= (int64_t)1 << 30) {
        for (int64_t i = 0; i < p.n_samples; ++i)
            if (!std::isfinite(p.samples[i]))
                return fail(HHG_ERROR_INVALID_ARGUMENT, at + "spline pulse sample " + std::to_string(i) + " is not finite");
        if (!std::isfinite(p.sample_dt) || !std::isfinite(p.t_cut))
            return fail(HHG_ERROR_INVALID_ARGUMENT, at + "spline pulse spacing and cut must be finite");
    }
    if (m.kind < 0 || m.kind > 4) return fail(HHG_ERROR_INVALID_ARGUMENT, at + "unknown model kind");
    if (m.relax < 0 || m.relax > 3) return fail(HHG_ERROR_INVALID_ARGUMENT, at + "unknown relaxation kind");
    const bool piflux3d = m.kind == HHG_MODEL_PIFLUX3D;
    if (piflux3d && m.relax != HHG_RELAX_NONE && m.relax != HHG_RELAX_EQUILIBRIUM)
        return fail(HHG_ERROR_UNSUPPORTED, at + "the 3D pi-flux lattice relaxes toward its equilibrium only");
    if (!piflux3d && m.kind != HHG_MODEL_HONEYCOMB && m.relax == HHG_RELAX_EQUILIBRIUM)
        return fail(HHG_ERROR_UNSUPPORTED,
                    at + "equilibrium relaxation is the 3D pi-flux and honeycomb lattices'");
    // the pi-flux equilibrium relaxation reads the integrator's state at the
    // start of every step (PiFlux.cpp:144), so j(t) depends on the step
    // sequence: only the reference's restart controller reproduces it
    if (piflux3d && m.relax == HHG_RELAX_EQUILIBRIUM && sv.proposal != HHG_PROPOSAL_RESTART)
        return fail(HHG_ERROR_UNSUPPORTED,
                    at + "equilibrium relaxation depends on the step sequence: it needs HHG_PROPOSAL_RESTART");
    if (m.relax != HHG_RELAX_NONE && !(m.relax_time > 0.0))
        return fail(HHG_ERROR_INVALID_ARGUMENT, at + "relaxation needs relax_time > 0");
    const bool lattice = m.kind == HHG_MODEL_PIFLUX2D || m.kind == HHG_MODEL_HONEYCOMB || piflux3d;
    if (lattice && m.theta != 0.0)
        return fail(HHG_ERROR_UNSUPPORTED, at + "lattice models take the field along one axis (theta = 0)");
    if (m.kind == HHG_MODEL_DIRAC3D && m.theta != 0.0)
        return fail(HHG_ERROR_UNSUPPORTED, at + "the 3D Dirac cone takes the field along z (theta = 0)");
    if (lattice && !(m.hopping > 0.0)) return fail(HHG_ERROR_INVALID_ARGUMENT, at + "lattice models need hopping > 0");
    if (p.kind < 0 || p.kind > 5) return fail(HHG_ERROR_INVALID_ARGUMENT, at + "unknown pulse kind");
    if (p.kind == HHG_PULSE_SPLINE &&
        (!p.samples || p.n_samples < 2 || p.n_samples > (int64_t)1 << 30 || !(p.sample_dt > 0.0) ||
         !std::isfinite(p.sample_t0) || !std::isfinite(p.slope_begin) || !std::isfinite(p.slope_end)))
        return fail(HHG_ERROR_INVALID_ARGUMENT,
                    at + "spline pulse needs 2 <= n_samples <= 2^30 samples, sample_dt > 0, finite t0 and slopes");
    const bool gauss = p.kind == HHG_PULSE_GAUSS || p.kind == HHG_PULSE_DOUBLE_GAUSS;
    const bool cos2 = p.kind == HHG_PULSE_COS2 || p.kind == HHG_PULSE_DOUBLE_COS2;
    if (gauss && !(p.sigma > 0.0)) return fail(HHG_ERROR_INVALID_ARGUMENT, at + "Gaussian pulse needs sigma > 0");
    if (cos2 && !(p.n_cycles > 0.0)) return fail(HHG_ERROR_INVALID_ARGUMENT, at + "cos^2 pulse needs n_cycles > 0");
    if (t.n_measurements < 1 || t.n_subdivisions < 1 || !(t.t_end > t.t_begin))
        return fail(HHG_ERROR_INVALID_ARGUMENT, at + "time grid needs N >= 1, n_subdivisions >= 1, t_end > t_begin");
    const double delta = (t.t_end - t.t_begin) / t.n_measurements;
    if (delta > 0.25 || (cos2 && delta / p.n_cycles > 0.25))
        return fail(HHG_ERROR_UNSUPPORTED,
                    at + "sampling interval must not exceed 0.25 (>= 26 samples per optical cycle)");
    if (sv.method != HHG_METHOD_RKF78 && m.relax != HHG_RELAX_NONE)
        return fail(HHG_ERROR_UNSUPPORTED, at + "the Magnus step is unitary: relaxation needs RKF78");
    if (m.split_dispersion != 0.0 && (!piflux3d || sv.method != HHG_METHOD_RKF78 || !(m.split_dispersion > 0.0)))
        return fail(HHG_ERROR_UNSUPPORTED, at + "the split current is the 3D pi-flux lattice's with RKF78 and a cut > 0");
    if (m.perp_current != 0 && (m.perp_current != 1 || m.kind != HHG_MODEL_HONEYCOMB || sv.method != HHG_METHOD_RKF78))
        return fail(HHG_ERROR_UNSUPPORTED, at + "j_perp (perp_current 1) is the honeycomb's with RKF78");
    if (sv.method == HHG_METHOD_MAGNUS_MOMENTS && m.kind != HHG_MODEL_DIRAC3D)
        return fail(HHG_ERROR_UNSUPPORTED, at + "the moment Magnus step is the 3D Dirac cone's (HHG_MODEL_DIRAC3D)");
    return HHG_SUCCESS;
}

}  // namespace hhg

extern "C" {

int32_t hhg_abi_version(void) { return HHG_AMD_ABI_VERSION; }

const char* hhg_last_error_string(void) { return hhg::last_error(); }

hhg_status hhg_gauss_legendre(int32_t n, double* x, double* w) {
    hhg::clear_error();
    if (n <= 0 || !x || !w) return hhg::fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_gauss_legendre: n > 0 and outputs required");
    hhg::gauss_legendre(n, x, w);
    return HHG_SUCCESS;
}

int64_t hhg_kgrid_size(const hhg_kgrid_desc* g) { return g ? hhg::kgrid_size(*g) : -1; }

hhg_status hhg_kgrid_build(const hhg_kgrid_desc* g, double* kx, double* ky, double* kz, double* w) {
    hhg::clear_error();
    if (!g || !kx || !ky || !w) return hhg::fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_kgrid_build: null argument");
    return hhg::kgrid_build(*g, kx, ky, kz, w);
}

hhg_status hhg_pulse_evaluate(const hhg_pulse_desc* p, const double* t, double* A, int64_t n) {
    hhg::clear_error();
    if (!p || (n > 0 && (!t || !A))) return hhg::fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_pulse_evaluate: null argument");
    if (p->kind < 0 || p->kind > 5) return hhg::fail(HHG_ERROR_INVALID_ARGUMENT, "unknown pulse kind");
    if (p->kind == HHG_PULSE_SPLINE) {
        if (!p->samples) return hhg::fail(HHG_ERROR_INVALID_ARGUMENT, "spline pulse needs samples");
        const hhg::SplinePulse sp(*p);
        if (sp.status != HHG_SUCCESS) return sp.status;
        for (int64_t i = 0; i < n; ++i) A[i] = sp(t[i]);
        return HHG_SUCCESS;
    }
    for (int64_t i = 0; i < n; ++i) A[i] = hhg::pulse_eval(*p, t[i]);
    return HHG_SUCCESS;
}

hhg_status hhg_reference_double_laser(int32_t gauss, double photon_energy, double E0, double model_ratio,
                                      double n_cycles, double delay_ps, int32_t active, double* samples,
                                      hhg_pulse_desc* out, double* momentum_amplitude) {
    hhg::clear_error();
    return hhg::reference_double_laser(gauss, photon_energy, E0, model_ratio, n_cycles, delay_ps, active, samples, out,
                                       momentum_amplitude);
}

int64_t hhg_reference_experimental_laser_size(double photon_energy) {
    return hhg::reference_experimental_size(photon_energy);
}

hhg_status hhg_reference_experimental_laser(double photon_energy, double E0, double model_ratio, double shift_ps,
                                            int32_t active, const double* E_A, const double* E_B, int32_t n_measured,
                                            double* samples, int64_t n_samples, hhg_pulse_desc* out,
                                            double* momentum_amplitude) {
    hhg::clear_error();
    return hhg::reference_experimental_laser(photon_energy, E0, model_ratio, shift_ps, active, E_A, E_B, n_measured,
                                             samples, n_samples, out, momentum_amplitude);
}

hhg_status hhg_spline_coefficients(const double* samples, int64_t n, double dt, double slope_begin, double slope_end,
                                   double* coef) {
    hhg::clear_error();
    if (!samples || !coef) return hhg::fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_spline_coefficients: null argument");
    return hhg::spline_coefficients(samples, n, dt, slope_begin, slope_end, coef);
}

hhg_status hhg_trapezoidal_spectrum(const double* j, const hhg_time_desc* time, int32_t zero_padding, double* freq,
                                    double* re, double* im) {
    hhg::clear_error();
    if (!j || !time || !freq || !re || !im)
        return hhg::fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_trapezoidal_spectrum: null argument");
    return hhg::trapezoidal_spectrum(j, *time, zero_padding, freq, re, im);
}

hhg_status hhg_welch_spectrum(const double* j, const hhg_time_desc* time, double* freq, double* re, double* im) {
    hhg::clear_error();
    if (!j || !time || !freq || !re || !im) return hhg::fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_welch_spectrum: null argument");
    return hhg::welch_spectrum(j, *time, freq, re, im);
}

void hhg_accumulators_to_double(const uint64_t* acc, int64_t n, double* j) {
    for (int64_t i = 0; i < n; ++i) {
        const long long hi = (long long)acc[2 * i + 1];
        j[i] = (double)hi * (1.0 / 65536.0) + (double)acc[2 * i] * (1.0 / (65536.0 * 18446744073709551616.0));
    }
}

void hhg_accumulators_add(uint64_t* total, const uint64_t* part, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t lo = total[2 * i] + part[2 * i];
        const uint64_t carry = lo < total[2 * i] ? 1u : 0u;
        total[2 * i] = lo;
        total[2 * i + 1] += part[2 * i + 1] + carry;
    }
}

}  // extern "C"
