// hhg_kernels.hpp — launch interface between the plan code and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "hhg_internal.hpp"

namespace hhg {

struct KernelArgs {
    const JobDev* jobs;
    const int2* items;  // {job, chunk}
    int n_items;
    int queue_chunk;
    int pad0, pad1;
    const double* kx;
    const double* ky;
    const double* w;
    const int64_t* perm;  // nullptr: identity lane -> k map
    const double* At;
    const double* Am;
    const double* nom;  // nominal-step pulse table (nullptr: evaluate every stage)
    unsigned long long* acc;  // 2 words per output sample
    Counters* ctr;
    SolverDev sp;
};

hipError_t launch_solve(int method, int model, int pulse, int relax, const KernelArgs& a, int blocks, int threads,
                        hipStream_t st);
const void* solve_kernel_fn(int method, int model, int pulse, int relax);
hipError_t launch_trajectories(int method, int model, int pulse, int relax, const JobDev* jobs, const double* kx,
                               const double* ky, const int64_t* sel, int64_t n_sel, const double* Am, SolverDev sp,
                               double* s_out, Counters* ctr, hipStream_t st);
hipError_t launch_pulse_table(int pulse, const JobDev* jobs, int job, int64_t n_threads, double* At, double* Am,
                              int with_magnus, hipStream_t st);
hipError_t launch_nominal_table(int pulse, const JobDev* jobs, int job, int64_t n_meas, double* tab,
                                hipStream_t st);
hipError_t launch_clear(unsigned long long* p, int64_t n_words, hipStream_t st);
hipError_t launch_finalize(const unsigned long long* acc, int64_t n, double* j, hipStream_t st);

}  // namespace hhg
