// hhg_kernels.hpp — launch interface between the plan code and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "hhg_internal.hpp"

namespace hhg {

struct MlpDev;

// Learned-proposal kernels: LEARN_NONE (plain and carried controllers; the
// learned first-interval policy is first_step_kernel followed by this kernel
// reading KernelArgs::h_first) and LEARN_SCHEDULED (the network proposes the
// first step of every infer_every-th interval inside the loop), each with or
// without LOG_STEPS: only kernels built with it write step logs.  The logging
// code holds the network's features per lane through the interval; without it
// the headline's honeycomb kernel needs 215 instead of 247 VGPRs and spills
// fewer scalar registers into vector lanes.
enum : int { LEARN_NONE = 0, LEARN_SCHEDULED = 2, LOG_STEPS = 4 };
__host__ __device__ constexpr int learn_mode(int learn) { return learn & 3; }

struct KernelArgs {
    const JobDev* jobs;
    const int2* items;  // {job, chunk}
    int n_items;
    int queue_chunk;
    int infer_every;     // learned proposal, scheduled policy: intervals between network evaluations
    int fallback_after;  // learned proposal: rejections before the plain controller takes over
    int infer_policy;    // learned proposal: HHG_INFER_FIRST or HHG_INFER_SCHEDULED
    int count_waves;     // nonzero: count the attempts every wave issues (Counters::wave_attempts)
    int block_acc;       // nonzero: block accumulation in LDS (the block pops waves x queue_chunk items)
    int acc_lds_off;     // byte offset of the LDS accumulators in the dynamic LDS
    int acc_lds_samples; // samples the LDS accumulators hold (the group's largest job output)
    int loop_only;       // nonzero: no whole-interval attempt outside the controller's loop
    const double* kx;
    const double* ky;
    const double* kz;  // 3D pi-flux lattice only (nullptr otherwise)
    const double* w;
    const int64_t* perm;  // nullptr: identity lane -> k map; else per job relative k index
    const double* At;
    const double* Pt;  // phase models: (sin f A_i, cos f A_i) per sample (model_has_phase)
    const double* Am;
    const double* nom;  // nominal-step pulse table (nullptr: evaluate every stage)
    const double* h_first;  // per trajectory: the first interval's first step (nullptr: Delta/n_sub)
    unsigned long long* acc;  // 2 words per output sample
    Counters* ctr;
    const MlpDev* mlp;        // learned proposal / cost probe network
    float* log;               // step log records (nullptr: no logging)
    unsigned long long* log_count;
    long long log_cap;
    int log_stride;
    int pad;
    SolverDev sp;
    unsigned long long* job_fail;  // per job: failed, non-finite trajectories (nullptr: not counted)
};

struct TrajArgs {
    const JobDev* jobs;
    const double* kx;
    const double* ky;
    const double* kz;
    const int64_t* sel;
    int64_t n_sel;
    const double* Am;
    SolverDev sp;
    double* s_out;
    Counters* ctr;
    hipStream_t st;
};

// per-model entry points (explicitly instantiated in hhg_solve_<model>.hip)
template <int M>
hipError_t solve_model(int method, int pulse, int relax, int learned, const KernelArgs& a, int blocks, int threads,
                       size_t smem, hipStream_t st, const void** fn);
template <int M>
hipError_t traj_model(int method, int pulse, int relax, const TrajArgs& x);
template <int M>
hipError_t cost_model(int pulse, const KernelArgs& a, int n_probe, float* cost, size_t smem, hipStream_t st);

hipError_t launch_solve(int method, int model, int pulse, int relax, int learned, const KernelArgs& a, int blocks,
                        int threads, size_t smem, hipStream_t st);
const void* solve_kernel_fn(int method, int model, int pulse, int relax, int learned);
// One RKF78 launch over the items of two groups, keys {model, pulse, relax} in
// either order: the fused kernel's address, or nullptr where none is built;
// launch_solve_fused returns hipErrorInvalidValue for such a pair.
const void* solve_fused_fn(const int key1[3], const int key2[3], int learned);
hipError_t launch_solve_fused(const int key1[3], const int key2[3], int learned, const KernelArgs& a, int blocks,
                              int threads, size_t smem, hipStream_t st);
hipError_t launch_trajectories(int method, int model, int pulse, int relax, const TrajArgs& x);
// 3D pi-flux lattice with equilibrium relaxation, RKF78: x.s_out = occupations [n_sel][N+1][2]
hipError_t launch_occupations(int pulse, const TrajArgs& x);
hipError_t launch_cost_probe(int model, int pulse, const KernelArgs& a, int n_probe, float* cost, size_t smem,
                             hipStream_t st);
template <int M>
hipError_t first_step_model(int pulse, const KernelArgs& a, double* h_first, size_t smem, hipStream_t st);
// the network's first-interval step of every trajectory of a group's items (h_first[traj_base + k])
hipError_t launch_first_step(int model, int pulse, const KernelArgs& a, double* h_first, size_t smem,
                             hipStream_t st);
// 30-point Gauss-Legendre nodes and weights on [0, 1] (moment Magnus tables)
struct Gauss30 {
    double x[30], w[30];
};

hipError_t launch_pulse_table(int pulse, const JobDev* jobs, int job, int64_t n_threads, double* At, double* Pt,
                              double* Am, int with_magnus, const Gauss30& g, hipStream_t st);
hipError_t launch_nominal_table(int pulse, const JobDev* jobs, int job, int64_t n_meas, double* tab,
                                hipStream_t st);
hipError_t launch_clear(unsigned long long* p, int64_t n_words, hipStream_t st);
hipError_t launch_finalize(const unsigned long long* acc, int64_t n, double* j, hipStream_t st);

}  // namespace hhg
