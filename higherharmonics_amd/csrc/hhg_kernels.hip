// hhg_kernels.hip — table, clear and finalize kernels, and the run-time
// dispatch onto the per-model solve kernels (hhg_solve_<model>.hip).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "hhg_device.hpp"
#include "hhg_kernels.hpp"

namespace hhg {

// A(t_i) for the vertex; with magnus = 1 A at the two Gauss nodes of every
// Magnus substep; with magnus = 2 the four pulse moments of every sampling
// interval for the reference's Dirac Magnus step (Laser::magnus_coefficients,
// sources/HHG/Laser/Laser.cpp:40-55): 30-point Gauss-Legendre on [0, 1] of
// A(t_{i-1} + Delta x) against 1, 2x - 1, 6x(x - 1) + 1 and
// 20x^3 - 30x^2 + 12x + 1 (sources/HHG/Laser/gauss.hpp:82-102; the last is the
// shifted Legendre P3 with the sign of its constant term flipped, evaluated as
// the reference evaluates it), scaled by 1, 3, 5, 7.
template <int P>
__global__ void pulse_table_kernel(const JobDev* jobs, int job, double* At, double* Pt, double* Am, int with_magnus,
                                   Gauss30 g) {
    JobRef J = job_ref(jobs + job);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= J.n_meas) {
        const double A = pulse_A<P>(J, J.t_begin + (double)i * J.delta);
        At[J.sample_base + i] = A;
        double sa = 0.0, ca = 1.0;
        if (model_has_phase(J.model)) sincos(model_phase_scale(J.model) * A, &sa, &ca);
        Pt[2 * (J.sample_base + i)] = sa;
        Pt[2 * (J.sample_base + i) + 1] = ca;
    }
    if (with_magnus == 2) {
        if (i < J.n_meas) {
            const double t0 = J.t_begin + (double)i * J.delta;
            double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
            for (int q = 0; q < 30; ++q) {
                const double x = g.x[q];
                const double a = g.w[q] * pulse_A<P>(J, t0 + J.delta * x);
                m0 += a;
                m1 += a * (2.0 * x - 1.0);
                m2 += a * (6.0 * x * (x - 1.0) + 1.0);
                m3 += a * (20.0 * x * x * x - 30.0 * x * x + 12.0 * x + 1.0);
            }
            double* out = Am + J.magnus_base + 4 * i;
            out[0] = m0;
            out[1] = 3.0 * m1;
            out[2] = 5.0 * m2;
            out[3] = 7.0 * m3;
        }
    } else if (with_magnus) {
        const int ms = J.magnus_sub;
        if (i < (int64_t)J.n_meas * ms) {
            const int64_t iv = i / ms, q = i % ms;
            const double h = J.delta / ms;
            const double t = (J.t_begin + (double)iv * J.delta) + (double)q * h;
            const double c1 = 0.5 - kSqrt3 / 6.0, c2 = 0.5 + kSqrt3 / 6.0;
            Am[J.magnus_base + 2 * i] = pulse_A<P>(J, t + c1 * h);
            Am[J.magnus_base + 2 * i + 1] = pulse_A<P>(J, t + c2 * h);
        }
    }
}

// Nominal-step table of one job: one thread per sampling interval replays the
// controller's maximal-growth sequence from the restart step (the same
// floating-point operations as rkf78_interval) and the whole-interval step,
// and stores A at the 13 stage times of each candidate, or for the phase
// models (model_has_phase) sin and cos of f A at the 10 distinct stage times.
// Unused slots hold t = NaN so they never match.
template <int P>
__global__ void nominal_table_kernel(const JobDev* jobs, int job, double* tab) {
    JobRef J = job_ref(jobs + job);
    const int64_t ii = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= J.n_meas) return;
    const int model = J.model;
    const bool phase = phase_table(model, J.relax);
    const int W = nom_stride(model, J.relax);
    double* row = tab + J.nominal_base + ii * nom_row(model, J.relax);
    const double t0 = J.t_begin + (double)ii * J.delta;
    const double t1 = J.t_begin + (double)(ii + 1) * J.delta;
    auto put = [&](int slot, double t, double h) {
        double* c = row + slot * W;
        c[0] = t;
        c[1] = h;
        if (phase) {
            const double fs = model_phase_scale(model);
            for (int d = 0; d < kPhaseStages; ++d)
                sincos(fs * pulse_A<P>(J, t + F78::c[phase_stage_of(d)] * h), &c[2 + 2 * d], &c[3 + 2 * d]);
        } else {
            for (int j = 0; j < 13; ++j) c[2 + j] = pulse_A<P>(J, t + F78::c[j] * h);
        }
    };
    auto put_empty = [&](int slot) {
        double* c = row + slot * W;
        for (int j = 0; j < W; ++j) c[j] = __builtin_nan("");
    };
    double t = t0, h = J.delta / J.n_sub;
    for (int slot = 0; slot < kNomRestartSlots; ++slot) {
        if (!(t1 - t > kDblEps)) {
            put_empty(slot);
            continue;
        }
        if ((t + h) - t1 > kDblEps) h = t1 - t;
        put(slot, t, h);
        t += h;
        h *= max_growth();
    }
    put(kNomRestartSlots, t0, t1 - t0);
}

// Zeroes the counters and accumulators (a fill kernel rather than a memset so
// the sequence can be captured in a graph and replayed).
__global__ void clear_kernel(unsigned long long* p, int64_t n_words) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_words) p[i] = 0ull;
}

__global__ void finalize_kernel(const unsigned long long* acc, int64_t n, double* j) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) j[i] = fix_to_double(acc + 2 * i);
}

// ---------------------------------------------------------------------------
// Run-time model dispatch
// ---------------------------------------------------------------------------
template <class F>
hipError_t by_model(int model, F&& f) {
    switch (model) {
    case MODEL_DIRAC2D: return f(std::integral_constant<int, MODEL_DIRAC2D>{});
    case MODEL_PIFLUX2D: return f(std::integral_constant<int, MODEL_PIFLUX2D>{});
    case MODEL_HONEYCOMB: return f(std::integral_constant<int, MODEL_HONEYCOMB>{});
    case MODEL_DIRAC3D: return f(std::integral_constant<int, MODEL_DIRAC3D>{});
    case MODEL_PIFLUX3D: return f(std::integral_constant<int, MODEL_PIFLUX3D>{});
    }
    return hipErrorInvalidValue;
}

template <class F>
hipError_t by_pulse(int pulse, F&& f) {
    switch (pulse) {
    case PULSE_COS2: return f(std::integral_constant<int, PULSE_COS2>{});
    case PULSE_GAUSS: return f(std::integral_constant<int, PULSE_GAUSS>{});
    case PULSE_DOUBLE_GAUSS: return f(std::integral_constant<int, PULSE_DOUBLE_GAUSS>{});
    case PULSE_DOUBLE_COS2: return f(std::integral_constant<int, PULSE_DOUBLE_COS2>{});
    case PULSE_CONTINUOUS: return f(std::integral_constant<int, PULSE_CONTINUOUS>{});
    case PULSE_SPLINE: return f(std::integral_constant<int, PULSE_SPLINE>{});
    }
    return hipErrorInvalidValue;
}

hipError_t launch_solve(int method, int model, int pulse, int relax, int learned, const KernelArgs& a, int blocks,
                        int threads, size_t smem, hipStream_t st) {
    return by_model(model, [&](auto M) {
        return solve_model<decltype(M)::value>(method, pulse, relax, learned, a, blocks, threads, smem, st, nullptr);
    });
}

const void* solve_kernel_fn(int method, int model, int pulse, int relax, int learned) {
    const void* fn = nullptr;
    KernelArgs dummy{};
    const hipError_t e = by_model(model, [&](auto M) {
        return solve_model<decltype(M)::value>(method, pulse, relax, learned, dummy, 0, 0, 0, nullptr, &fn);
    });
    return e == hipSuccess ? fn : nullptr;
}

hipError_t launch_trajectories(int method, int model, int pulse, int relax, const TrajArgs& x) {
    return by_model(model, [&](auto M) { return traj_model<decltype(M)::value>(method, pulse, relax, x); });
}

hipError_t launch_cost_probe(int model, int pulse, const KernelArgs& a, int n_probe, float* cost, size_t smem,
                             hipStream_t st) {
    return by_model(model, [&](auto M) { return cost_model<decltype(M)::value>(pulse, a, n_probe, cost, smem, st); });
}

hipError_t launch_first_step(int model, int pulse, const KernelArgs& a, double* h_first, size_t smem,
                             hipStream_t st) {
    return by_model(model, [&](auto M) { return first_step_model<decltype(M)::value>(pulse, a, h_first, smem, st); });
}

hipError_t launch_pulse_table(int pulse, const JobDev* jobs, int job, int64_t n_threads, double* At, double* Pt,
                              double* Am, int with_magnus, const Gauss30& g, hipStream_t st) {
    const int blocks = (int)((n_threads + 255) / 256);
    return by_pulse(pulse, [&](auto P) {
        hipLaunchKernelGGL((pulse_table_kernel<decltype(P)::value>), dim3(blocks), dim3(256), 0, st, jobs, job, At, Pt,
                           Am, with_magnus, g);
        return hipGetLastError();
    });
}

hipError_t launch_nominal_table(int pulse, const JobDev* jobs, int job, int64_t n_meas, double* tab,
                                hipStream_t st) {
    const int blocks = (int)((n_meas + 127) / 128);
    return by_pulse(pulse, [&](auto P) {
        hipLaunchKernelGGL((nominal_table_kernel<decltype(P)::value>), dim3(blocks), dim3(128), 0, st, jobs, job, tab);
        return hipGetLastError();
    });
}

hipError_t launch_clear(unsigned long long* p, int64_t n_words, hipStream_t st) {
    if (n_words <= 0) return hipSuccess;
    hipLaunchKernelGGL(clear_kernel, dim3((unsigned)((n_words + 255) / 256)), dim3(256), 0, st, p, n_words);
    return hipGetLastError();
}

hipError_t launch_finalize(const unsigned long long* acc, int64_t n, double* j, hipStream_t st) {
    hipLaunchKernelGGL(finalize_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, acc, n, j);
    return hipGetLastError();
}

}  // namespace hhg
