// hhg_kernels.hip — table, clear and finalize kernels, and the run-time
// dispatch onto the per-model solve kernels (hhg_solve_<model>.hip).
#include <hip/hip_runtime.h>

#include "hhg_device.hpp"
#include "hhg_kernels.hpp"

namespace hhg {

// A(t_i) for the vertex and A at the two Gauss nodes of every Magnus substep.
template <int P>
__global__ void pulse_table_kernel(const JobDev* jobs, int job, double* At, double* Am, int with_magnus) {
    JobRef J = job_ref(jobs + job);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= J.n_meas) At[J.sample_base + i] = pulse_A<P>(J, J.t_begin + (double)i * J.delta);
    if (with_magnus) {
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
// and stores A at the 13 stage times of each candidate.  Unused slots hold
// t = NaN so they never match.
template <int P>
__global__ void nominal_table_kernel(const JobDev* jobs, int job, double* tab) {
    JobRef J = job_ref(jobs + job);
    const int64_t ii = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= J.n_meas) return;
    double* row = tab + J.nominal_base + ii * kNomRow;
    const double t0 = J.t_begin + (double)ii * J.delta;
    const double t1 = J.t_begin + (double)(ii + 1) * J.delta;
    auto put = [&](int slot, double t, double h) {
        double* c = row + slot * kNomStride;
        c[0] = t;
        c[1] = h;
        for (int j = 0; j < 13; ++j) c[2 + j] = pulse_A<P>(J, t + F78::c[j] * h);
    };
    auto put_empty = [&](int slot) {
        double* c = row + slot * kNomStride;
        for (int j = 0; j < kNomStride; ++j) c[j] = __builtin_nan("");
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
hipError_t launch_solve(int method, int model, int pulse, int relax, bool learned, const KernelArgs& a, int blocks,
                        int threads, size_t smem, hipStream_t st) {
    switch (model) {
    case MODEL_DIRAC2D: return solve_model<MODEL_DIRAC2D>(method, pulse, relax, learned, a, blocks, threads, smem, st, nullptr);
    case MODEL_PIFLUX2D: return solve_model<MODEL_PIFLUX2D>(method, pulse, relax, learned, a, blocks, threads, smem, st, nullptr);
    case MODEL_HONEYCOMB: return solve_model<MODEL_HONEYCOMB>(method, pulse, relax, learned, a, blocks, threads, smem, st, nullptr);
    }
    return hipErrorInvalidValue;
}

const void* solve_kernel_fn(int method, int model, int pulse, int relax, bool learned) {
    const void* fn = nullptr;
    KernelArgs dummy{};
    hipError_t e = hipErrorInvalidValue;
    switch (model) {
    case MODEL_DIRAC2D: e = solve_model<MODEL_DIRAC2D>(method, pulse, relax, learned, dummy, 0, 0, 0, nullptr, &fn); break;
    case MODEL_PIFLUX2D: e = solve_model<MODEL_PIFLUX2D>(method, pulse, relax, learned, dummy, 0, 0, 0, nullptr, &fn); break;
    case MODEL_HONEYCOMB: e = solve_model<MODEL_HONEYCOMB>(method, pulse, relax, learned, dummy, 0, 0, 0, nullptr, &fn); break;
    }
    return e == hipSuccess ? fn : nullptr;
}

hipError_t launch_trajectories(int method, int model, int pulse, int relax, const TrajArgs& x) {
    switch (model) {
    case MODEL_DIRAC2D: return traj_model<MODEL_DIRAC2D>(method, pulse, relax, x);
    case MODEL_PIFLUX2D: return traj_model<MODEL_PIFLUX2D>(method, pulse, relax, x);
    case MODEL_HONEYCOMB: return traj_model<MODEL_HONEYCOMB>(method, pulse, relax, x);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_cost_probe(int model, int pulse, const KernelArgs& a, int n_probe, float* cost, size_t smem,
                             hipStream_t st) {
    switch (model) {
    case MODEL_DIRAC2D: return cost_model<MODEL_DIRAC2D>(pulse, a, n_probe, cost, smem, st);
    case MODEL_PIFLUX2D: return cost_model<MODEL_PIFLUX2D>(pulse, a, n_probe, cost, smem, st);
    case MODEL_HONEYCOMB: return cost_model<MODEL_HONEYCOMB>(pulse, a, n_probe, cost, smem, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_pulse_table(int pulse, const JobDev* jobs, int job, int64_t n_threads, double* At, double* Am,
                              int with_magnus, hipStream_t st) {
    const int blocks = (int)((n_threads + 255) / 256);
    switch (pulse) {
    case PULSE_COS2:
        hipLaunchKernelGGL((pulse_table_kernel<PULSE_COS2>), dim3(blocks), dim3(256), 0, st, jobs, job, At, Am, with_magnus);
        break;
    case PULSE_GAUSS:
        hipLaunchKernelGGL((pulse_table_kernel<PULSE_GAUSS>), dim3(blocks), dim3(256), 0, st, jobs, job, At, Am, with_magnus);
        break;
    case PULSE_DOUBLE_GAUSS:
        hipLaunchKernelGGL((pulse_table_kernel<PULSE_DOUBLE_GAUSS>), dim3(blocks), dim3(256), 0, st, jobs, job, At, Am,
                           with_magnus);
        break;
    case PULSE_DOUBLE_COS2:
        hipLaunchKernelGGL((pulse_table_kernel<PULSE_DOUBLE_COS2>), dim3(blocks), dim3(256), 0, st, jobs, job, At, Am,
                           with_magnus);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_nominal_table(int pulse, const JobDev* jobs, int job, int64_t n_meas, double* tab,
                                hipStream_t st) {
    const int blocks = (int)((n_meas + 127) / 128);
    switch (pulse) {
    case PULSE_COS2:
        hipLaunchKernelGGL((nominal_table_kernel<PULSE_COS2>), dim3(blocks), dim3(128), 0, st, jobs, job, tab);
        break;
    case PULSE_GAUSS:
        hipLaunchKernelGGL((nominal_table_kernel<PULSE_GAUSS>), dim3(blocks), dim3(128), 0, st, jobs, job, tab);
        break;
    case PULSE_DOUBLE_GAUSS:
        hipLaunchKernelGGL((nominal_table_kernel<PULSE_DOUBLE_GAUSS>), dim3(blocks), dim3(128), 0, st, jobs, job,
                           tab);
        break;
    case PULSE_DOUBLE_COS2:
        hipLaunchKernelGGL((nominal_table_kernel<PULSE_DOUBLE_COS2>), dim3(blocks), dim3(128), 0, st, jobs, job, tab);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
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
