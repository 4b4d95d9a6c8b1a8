// hhg_kernels.hip — gfx950 kernels of the Bloch hot path.
//
// Work decomposition: a work item is 64 consecutive k-points of one job, one
// per lane of a wavefront.  Lanes advance sampling interval by sampling
// interval (reference loop: sources/HHG/Systems/DiracSystem.cpp:166-173); at
// every sample t_i the wave sums w_k g_k(t_i) with a shuffle butterfly, lane
// (i mod 64) keeps the sum, and every 64 samples the wave adds its 64 sums to
// the job's exact 128-bit fixed-point accumulators (one vector atomic pair).
// In persistent mode a grid sized to the resident wave count pops items from
// a device counter until the list is exhausted.
#include <hip/hip_runtime.h>

#include "hhg_device.hpp"
#include "hhg_kernels.hpp"

namespace hhg {

#ifndef HHG_WAVES_PER_EU
#define HHG_WAVES_PER_EU 2
#endif

namespace {

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Copies a wave-uniform record into scalar registers.  The tables are only
// read by the kernels, but the compiler cannot prove that the atomics to the
// accumulators leave them untouched, so without this the 200-byte job record
// would be held in vector registers of every lane.
template <class T>
__device__ __forceinline__ T load_uniform(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "record must be dword sized");
    T out;
    const int* src = reinterpret_cast<const int*>(p);
    int* dst = reinterpret_cast<int*>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return out;
}

// Per-lane buffering of one sample sum and the 64-sample flush.
struct SampleSink {
    unsigned long long* acc;  // job accumulators, 2 words per sample
    int N;
    double buf;
    bool overflow;
    __device__ __forceinline__ void put(int i, double total, int lane) {
        const int slot = i & (WAVE - 1);
        if (lane == slot) buf = total;
        if (slot == WAVE - 1 || i == N) {
            const int base = i - slot;
            if (lane <= slot) overflow |= !fix_accumulate(acc + 2 * (int64_t)(base + lane), buf);
        }
    }
};

struct LaneCounters {
    uint32_t att = 0, rej = 0, fail = 0, nonfin = 0, fallback = 0;
};

__device__ __forceinline__ void flush_counters(Counters* ctr, LaneCounters& lc) {
    const uint32_t a = wave_sum_u32(lc.att), r = wave_sum_u32(lc.rej), f = wave_sum_u32(lc.fail),
                   n = wave_sum_u32(lc.nonfin), b = wave_sum_u32(lc.fallback);
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        if (a) atomicAdd(&ctr->attempts, (unsigned long long)a);
        if (r) atomicAdd(&ctr->rejects, (unsigned long long)r);
        if (f) atomicAdd(&ctr->failed, (unsigned long long)f);
        if (n) atomicAdd(&ctr->nonfinite, (unsigned long long)n);
        if (b) atomicAdd(&ctr->fallbacks, (unsigned long long)b);
    }
    lc = LaneCounters{};
}

template <int M, int P, int R>
__device__ void solve_item_rkf78(JobRef J, int chunk, const KernelArgs& a, LaneCounters& lc) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t kk = (int64_t)chunk * WAVE + lane;
    const bool active = kk < J.n_k;
    double kx = 0.0, ky = 0.0, w = 0.0;
    if (active) {
        const int64_t idx = J.k_base + (a.perm ? a.perm[J.k_base + kk] : kk);
        kx = a.kx[idx];
        ky = a.ky[idx];
        w = a.w[idx];
    }
    Bloch<M, P, R> f(J, kx, ky);
    double s[3];
    f.initial(s);
    const int N = J.n_meas;
    SampleSink sink{a.acc + 2 * J.sample_base, N, 0.0, false};
    const double* At = a.At + J.sample_base;
    sink.put(0, wave_sum(active ? w * f.vertex(At[0], s) : 0.0), lane);

    int status = active ? 0 : 3;
    const bool restart = a.sp.proposal == PROPOSAL_RESTART;
    const double h_restart = J.delta / J.n_sub;
    double h_carry = h_restart;
    for (int i = 1; i <= N; ++i) {
        if (status == 0) {
            const double t0 = J.t_begin + (double)(i - 1) * J.delta;
            const double t1 = J.t_begin + (double)i * J.delta;
            double h = restart ? h_restart : h_carry;
            double unclipped = h;
            const double* nom = a.nom ? a.nom + J.nominal_base + (int64_t)(i - 1) * kNomRow : nullptr;
            rkf78_interval(f, s, t0, t1, h, unclipped, a.sp, lc.att, lc.rej, status, nom);
            h_carry = h > unclipped ? h : unclipped;
        }
        const double g = (status != 3) ? w * f.vertex(At[i], s) : 0.0;
        sink.put(i, wave_sum(g), lane);
    }
    if (status == 1) lc.fail++;
    if (status == 2) lc.nonfin++;
    if (sink.overflow) lc.nonfin++;
}

template <int M>
__device__ void solve_item_magnus(JobRef J, int chunk, const KernelArgs& a, LaneCounters& lc) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t kk = (int64_t)chunk * WAVE + lane;
    const bool active = kk < J.n_k;
    double kx = 0.0, ky = 0.0, w = 0.0;
    if (active) {
        const int64_t idx = J.k_base + (a.perm ? a.perm[J.k_base + kk] : kk);
        kx = a.kx[idx];
        ky = a.ky[idx];
        w = a.w[idx];
    }
    Bloch<M, PULSE_COS2, RELAX_NONE> f(J, kx, ky);  // the pulse enters only through the Am table
    double s[3];
    f.initial(s);
    const int N = J.n_meas;
    const int ms = J.magnus_sub;
    const double h = J.delta / ms;
    SampleSink sink{a.acc + 2 * J.sample_base, N, 0.0, false};
    const double* At = a.At + J.sample_base;
    const double* Am = a.Am + J.magnus_base;
    sink.put(0, wave_sum(active ? w * f.vertex(At[0], s) : 0.0), lane);
    for (int i = 1; i <= N; ++i) {
        for (int q = 0; q < ms; ++q) {
            const int64_t m = 2 * ((int64_t)(i - 1) * ms + q);
            magnus4_step(f, s, Am[m], Am[m + 1], h);
        }
        const double g = active ? w * f.vertex(At[i], s) : 0.0;
        sink.put(i, wave_sum(g), lane);
    }
    if (active) lc.att += (uint32_t)N * (uint32_t)ms;
    if (active && !(isfinite(s[0]) && isfinite(s[1]) && isfinite(s[2]))) lc.nonfin++;
    if (sink.overflow) lc.nonfin++;
}

template <int METHOD, int M, int P, int R>
__device__ __forceinline__ void solve_item(JobRef J, int chunk, const KernelArgs& a, LaneCounters& lc) {
    if constexpr (METHOD == 0)
        solve_item_rkf78<M, P, R>(J, chunk, a, lc);
    else
        solve_item_magnus<M>(J, chunk, a, lc);
}

}  // namespace

// One kernel per (method, model, pulse, relaxation); a launch only carries
// items of its own group.  Every wave pops items from the group's queue
// counter until the list is exhausted; a grid of one wave per item makes each
// wave solve one item, a grid of the resident wave count is the persistent
// schedule.
template <int METHOD, int M, int P, int R>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(HHG_WAVES_PER_EU, HHG_WAVES_PER_EU))) void
solve_kernel(KernelArgs a) {
    const int lane = threadIdx.x & (WAVE - 1);
    LaneCounters lc;
    for (;;) {
        int base = 0;
        if (lane == 0) base = (int)atomicAdd(&a.ctr->queue, (unsigned long long)a.queue_chunk);
        base = uniform(base);
        if (base >= a.n_items) break;
        const int end = min(base + a.queue_chunk, a.n_items);
        for (int it = base; it < end; ++it) {
            const int2 item = load_uniform(a.items + it);
            solve_item<METHOD, M, P, R>(job_ref(a.jobs + item.x), item.y, a, lc);
        }
    }
    flush_counters(a.ctr, lc);
}

// Per-k trajectories of one job (tests / inspection): one lane per selected
// k-point, Bloch vector at every sample.
template <int METHOD, int M, int P, int R>
__global__ __launch_bounds__(64) void trajectory_kernel(const JobDev* jobs, const double* kx, const double* ky,
                                                        const int64_t* sel, int64_t n_sel, const double* Am,
                                                        SolverDev sp, double* s_out, Counters* ctr) {
    const int64_t t_idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t_idx >= n_sel) return;
    JobRef J = job_ref(jobs);
    const int64_t idx = J.k_base + sel[t_idx];
    Bloch<M, P, R> f(J, kx[idx], ky[idx]);
    double s[3];
    f.initial(s);
    const int N = J.n_meas;
    double* out = s_out + t_idx * (int64_t)(N + 1) * 3;
    out[0] = s[0];
    out[1] = s[1];
    out[2] = s[2];
    uint32_t att = 0, rej = 0;
    int status = 0;
    const double h_restart = J.delta / J.n_sub;
    double h_carry = h_restart;
    for (int i = 1; i <= N; ++i) {
        if constexpr (METHOD == 0) {
            if (status == 0) {
                const double t0 = J.t_begin + (double)(i - 1) * J.delta;
                const double t1 = J.t_begin + (double)i * J.delta;
                double h = (sp.proposal == PROPOSAL_RESTART) ? h_restart : h_carry;
                double unclipped = h;
                rkf78_interval(f, s, t0, t1, h, unclipped, sp, att, rej, status, nullptr);
                h_carry = h > unclipped ? h : unclipped;
            }
        } else {
            const int ms = J.magnus_sub;
            const double h = J.delta / ms;
            for (int q = 0; q < ms; ++q) {
                const int64_t m = J.magnus_base + 2 * ((int64_t)(i - 1) * ms + q);
                magnus4_step(f, s, Am[m], Am[m + 1], h);
            }
            att += ms;
        }
        out[3 * i + 0] = s[0];
        out[3 * i + 1] = s[1];
        out[3 * i + 2] = s[2];
    }
    atomicAdd(&ctr->attempts, (unsigned long long)att);
    atomicAdd(&ctr->rejects, (unsigned long long)rej);
    if (status == 1) atomicAdd(&ctr->failed, 1ull);
    if (status == 2) atomicAdd(&ctr->nonfinite, 1ull);
}

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
// Host-side launchers: run-time (method, model, pulse, relax) -> template.
// Magnus takes no relaxation and reads the pulse from its table, so it has one
// instantiation per model.
// ---------------------------------------------------------------------------
template <template <int, int, int, int> class Op, int ME, int MO, int P>
static hipError_t dispatch_r(int relax, typename Op<0, 0, 0, 0>::Args& args) {
    switch (relax) {
    case RELAX_NONE: return Op<ME, MO, P, RELAX_NONE>::run(args);
    case RELAX_TO_INITIAL: return Op<ME, MO, P, RELAX_TO_INITIAL>::run(args);
    case RELAX_T2: return Op<ME, MO, P, RELAX_T2>::run(args);
    }
    return hipErrorInvalidValue;
}

template <template <int, int, int, int> class Op, int ME, int MO>
static hipError_t dispatch_p(int pulse, int relax, typename Op<0, 0, 0, 0>::Args& args) {
    if constexpr (ME == 1) {
        return Op<ME, MO, PULSE_COS2, RELAX_NONE>::run(args);
    } else {
        switch (pulse) {
        case PULSE_COS2: return dispatch_r<Op, ME, MO, PULSE_COS2>(relax, args);
        case PULSE_GAUSS: return dispatch_r<Op, ME, MO, PULSE_GAUSS>(relax, args);
        case PULSE_DOUBLE_GAUSS: return dispatch_r<Op, ME, MO, PULSE_DOUBLE_GAUSS>(relax, args);
        case PULSE_DOUBLE_COS2: return dispatch_r<Op, ME, MO, PULSE_DOUBLE_COS2>(relax, args);
        }
        return hipErrorInvalidValue;
    }
}

template <template <int, int, int, int> class Op>
static hipError_t dispatch(int method, int model, int pulse, int relax, typename Op<0, 0, 0, 0>::Args& args) {
#define HHG_D(ME, MO) \
    if (method == ME && model == MO) return dispatch_p<Op, ME, MO>(pulse, relax, args);
    HHG_D(0, MODEL_DIRAC2D)
    HHG_D(0, MODEL_PIFLUX2D)
    HHG_D(0, MODEL_HONEYCOMB)
    HHG_D(1, MODEL_DIRAC2D)
    HHG_D(1, MODEL_PIFLUX2D)
    HHG_D(1, MODEL_HONEYCOMB)
#undef HHG_D
    return hipErrorInvalidValue;
}

struct SolveLaunch {
    const KernelArgs* a;
    int blocks, threads;
    hipStream_t st;
    const void* fn;
};

template <int ME, int MO, int P, int R>
struct SolveOp {
    using Args = SolveLaunch;
    static hipError_t run(Args& x) {
        if (x.a == nullptr) {
            x.fn = reinterpret_cast<const void*>(&solve_kernel<ME, MO, P, R>);
            return hipSuccess;
        }
        hipLaunchKernelGGL((solve_kernel<ME, MO, P, R>), dim3(x.blocks), dim3(x.threads), 0, x.st, *x.a);
        return hipGetLastError();
    }
};

hipError_t launch_solve(int method, int model, int pulse, int relax, const KernelArgs& a, int blocks, int threads,
                        hipStream_t st) {
    SolveLaunch x{&a, blocks, threads, st, nullptr};
    return dispatch<SolveOp>(method, model, pulse, relax, x);
}

const void* solve_kernel_fn(int method, int model, int pulse, int relax) {
    SolveLaunch x{nullptr, 0, 0, nullptr, nullptr};
    if (dispatch<SolveOp>(method, model, pulse, relax, x) != hipSuccess) return nullptr;
    return x.fn;
}

struct TrajLaunch {
    const JobDev* jobs;
    const double *kx, *ky;
    const int64_t* sel;
    int64_t n_sel;
    const double* Am;
    SolverDev sp;
    double* s_out;
    Counters* ctr;
    hipStream_t st;
};

template <int ME, int MO, int P, int R>
struct TrajOp {
    using Args = TrajLaunch;
    static hipError_t run(Args& x) {
        const int blocks = (int)((x.n_sel + 63) / 64);
        hipLaunchKernelGGL((trajectory_kernel<ME, MO, P, R>), dim3(blocks), dim3(64), 0, x.st, x.jobs, x.kx, x.ky,
                           x.sel, x.n_sel, x.Am, x.sp, x.s_out, x.ctr);
        return hipGetLastError();
    }
};

hipError_t launch_trajectories(int method, int model, int pulse, int relax, const JobDev* jobs, const double* kx,
                               const double* ky, const int64_t* sel, int64_t n_sel, const double* Am, SolverDev sp,
                               double* s_out, Counters* ctr, hipStream_t st) {
    TrajLaunch x{jobs, kx, ky, sel, n_sel, Am, sp, s_out, ctr, st};
    return dispatch<TrajOp>(method, model, pulse, relax, x);
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
