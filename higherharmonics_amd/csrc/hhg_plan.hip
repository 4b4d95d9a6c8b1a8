// hhg_plan.hip — plans: validation, k-grid/job tables in HBM, launches,
// statistics, autotuning and graph capture behind the C ABI.
//
// A plan replaces the reference's Dispatcher (sources/HHG/Dispatch/Dispatcher.hpp:13-33):
// it is built once from a description (the Dispatcher constructors,
// DiracDispatcher.cpp:25-62, PiFluxDispatcher.cpp:46-105,
// HoneycombDispatcher.cpp:28-56) and then computes j(t) (Dispatcher::compute).
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <limits>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/hhg_amd.h"
#include "hhg_host.hpp"
#include "hhg_internal.hpp"
#include "hhg_kernels.hpp"
#include "hhg_mlp.hpp"

using hhg::fail;

namespace {

constexpr int kCounterWords = (int)(sizeof(hhg::Counters) / sizeof(unsigned long long));

struct Group {
    int model, pulse, relax;
    int item_begin, n_items;
    int waves_per_block, queue_chunk, blocks;
    int acc_samples;  // the largest output of the group's jobs (samples)
    int block_acc;    // block accumulation in LDS (size_group decides)
    int fused;        // the fused launch over two groups' items: the second group's key follows
    int model2, pulse2, relax2;
};

std::string hip_msg(const char* what, hipError_t e) { return std::string(what) + ": " + hipGetErrorString(e); }

// A roctx range around a host call (rocprofv3 --marker-trace shows it beside
// the kernels the call enqueued).
struct TraceRange {
    explicit TraceRange(const char* what) { roctxRangePushA(what); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

}  // namespace

// Static workspace of hhg_trajectories_noalloc: the selection indices, the
// trajectory output and the counters of one call, for callers that cannot
// allocate device memory.  Every GPU the library runs on holds its own copy
// for the life of the process; calls on one GPU share it, so calls that
// overlap on one GPU race.
constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;  // per device: the address of its copy of scratch

// The device address of `scratch` on `device`, looked up once per device
// with that device current (the caller has made it current).
hhg_status scratch_on_device(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(HHG_ERROR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return HHG_SUCCESS;
}

constexpr size_t round_up_256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

struct hhg_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    hhg_solver_desc solver{};
    hhg_launch_desc launch{};
    int n_jobs = 0;
    std::vector<hhg::JobDev> jobs;
    std::vector<hhg_job_desc> job_desc;
    std::vector<int64_t> offsets;  // n_jobs + 1
    int64_t n_out = 0;
    int64_t n_kpts = 0;
    std::vector<Group> groups;     // per (model, pulse, relax): first-step and cost-probe launches
    std::vector<Group> launches;   // the solve launches: the groups, or one fused launch over two groups
    int num_cus = 0;
    size_t lds_per_block = 0;  // the device's dynamic LDS limit per block

    double *d_kx = nullptr, *d_ky = nullptr, *d_kz = nullptr, *d_w = nullptr;
    double* d_spl = nullptr;  // spline pulse coefficients of every spline job, concatenated
    hhg::JobDev* d_jobs = nullptr;
    int2* d_items = nullptr;
    double *d_At = nullptr, *d_Pt = nullptr, *d_Am = nullptr, *d_nom = nullptr;
    double* d_h_first = nullptr;  // learned first-interval steps, one per trajectory
    unsigned long long* d_words = nullptr;  // [counters per group][accumulators]
    int64_t n_words = 0;
    double* d_j = nullptr;
    hhg::Counters* d_aux = nullptr;  // counters of the trajectory and occupation launches
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_done = nullptr;  // recorded after every execution, on the stream it used
    bool executed = false;
    bool last_host = false;        // the last execution was hhg_plan_execute[_accumulators]
    double last_wall_ms = 0.0;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    const hhg::MlpDev* d_mlp = nullptr;  // owned by the hhg_mlp handle
    int64_t* d_perm = nullptr;
    float* d_log = nullptr;
    int64_t log_cap = 0;
    int log_stride = 16;

    bool learned() const { return solver.proposal == HHG_PROPOSAL_LEARNED; }
    // learned first-interval proposals: first_step_kernel, then the carried controller's kernel
    bool first_step() const { return learned() && solver.infer_policy != HHG_INFER_SCHEDULED; }
    // the solve kernels' learned-proposal instantiation (LEARN_* of hhg_solve.hpp): the network runs
    // inside the solve kernel only for the scheduled policy
    // the solve kernels' LEARN argument: the scheduled policy's in-loop network,
    // and the step-logging variant while a step log is set
    int solve_learn() const {
        return (learned() && solver.infer_policy == HHG_INFER_SCHEDULED ? hhg::LEARN_SCHEDULED : hhg::LEARN_NONE) |
               (d_log ? hhg::LOG_STEPS : 0);
    }
    // dynamic LDS of a group's solve launch: the network (scheduled learned
    // policy), then the block accumulators (2 words per sample)
    size_t mlp_lds() const {
        return hhg::learn_mode(solve_learn()) != hhg::LEARN_NONE ? (size_t)hhg::kMlpLdsBytes : 0;
    }
    size_t acc_lds_off() const { return (mlp_lds() + 15) & ~(size_t)15; }
    size_t smem(const Group& g) const {
        return g.block_acc ? acc_lds_off() + 16 * (size_t)g.acc_samples : mlp_lds();
    }

    // at least one counter block, so trajectory launches have one even when
    // this shard owns no work items
    int64_t n_counter_blocks() const { return std::max<int64_t>(1, (int64_t)groups.size()); }
    hhg::Counters* counters(int g) const { return reinterpret_cast<hhg::Counters*>(d_words) + g; }
    unsigned long long* acc() const { return d_words + n_counter_blocks() * kCounterWords; }
    // per job: failed and non-finite trajectories of the last execution, after the accumulators
    unsigned long long* job_fail() const { return acc() + 2 * n_out; }
};

namespace {

hhg_status check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return HHG_SUCCESS;
    if (e == hipErrorOutOfMemory) return fail(HHG_ERROR_OUT_OF_MEMORY, hip_msg(what, e));
    return fail(HHG_ERROR_HIP, hip_msg(what, e));
}

#define HHG_TRY(expr, what)                                 \
    do {                                                    \
        const hhg_status _st = check_hip((expr), (what));   \
        if (_st != HHG_SUCCESS) return _st;                 \
    } while (0)

hhg::SolverDev solver_dev(const hhg_solver_desc& sv) {
    hhg::SolverDev sp{};
    sp.rtol = sv.rtol;
    sp.atol = sv.atol;
    sp.max_fail = sv.max_fail > 0 ? sv.max_fail : 500;
    sp.proposal = sv.proposal;
    sp.max_steps = sv.max_steps > 0 ? sv.max_steps : hhg::kDefaultMaxSteps;
    return sp;
}

// Corners of the cos^2 envelopes (A'' jumps at the support edges) strictly
// inside a sampling interval, ascending.  An RKF78 step across a corner has an
// error far above its embedded estimate; the carried and learned proposals
// take whole-interval steps, so they integrate up to each corner and continue
// from there.  The plain controller keeps the reference's per-interval restart
// sequence unchanged.  Edges within 1e-9 Delta of a sample time and edges of a
// zero-amplitude component are skipped (the oracle's envelope_corners makes
// the same list with the same operations).
void set_envelope_corners(hhg::JobDev& J, const hhg_pulse_desc& p, const hhg_solver_desc& sv) {
    J.n_kinks = 0;
    if (sv.method != HHG_METHOD_RKF78 || sv.proposal == HHG_PROPOSAL_RESTART) return;
    if (p.kind != HHG_PULSE_COS2 && p.kind != HHG_PULSE_DOUBLE_COS2) return;
    std::vector<double> e;
    if (p.amp_a != 0.0) {
        e.push_back(p.t0);
        e.push_back(p.t0 + 2.0 * 3.14159265358979323846 * p.n_cycles);
    }
    if (p.kind == HHG_PULSE_DOUBLE_COS2 && p.amp_b != 0.0) {
        const double b = p.t0 + p.delay;
        e.push_back(b);
        e.push_back(b + 2.0 * 3.14159265358979323846 * p.n_cycles);
    }
    std::sort(e.begin(), e.end());
    for (double t : e) {
        const double x = (t - J.t_begin) / J.delta;
        if (!(x > 0.0 && x < (double)J.n_meas)) continue;
        const double fl = std::floor(x);
        if (x - fl < 1e-9 || fl + 1.0 - x < 1e-9) continue;
        if (J.n_kinks > 0 && J.kink_t[J.n_kinks - 1] == t) continue;
        J.kink_t[J.n_kinks] = t;
        J.kink_i[J.n_kinks] = (int32_t)fl + 1;
        ++J.n_kinks;
    }
}

hhg::JobDev make_job(const hhg_job_desc& jd, const hhg_solver_desc& sv) {
    hhg::JobDev J{};
    const hhg_model_desc& m = jd.model;
    const hhg_pulse_desc& p = jd.pulse;
    J.model = m.kind;
    J.relax = m.relax;
    J.v = m.v;
    J.mass = m.mass;
    J.hop = m.hopping;
    J.ct = std::cos(m.theta);
    J.st = std::sin(m.theta);
    if (m.theta == 0.0) J.ct = 1.0, J.st = 0.0;
    J.fermi = m.fermi_energy;
    J.beta = m.beta;
    J.relax_rate = m.relax != HHG_RELAX_NONE ? 1.0 / m.relax_time : 0.0;
    J.relax_rate2 = m.relax == HHG_RELAX_EQUILIBRIUM && m.relax_time_offdiag > 0.0 ? 1.0 / m.relax_time_offdiag : 0.0;
    // one-exponential Fermi pair of the equilibrium relaxations: w = exp(beta c)
    // within [1e-150, 1e150], so that (q + w)(1 + w q) stays finite and nonzero
    {
        const double c = m.kind == HHG_MODEL_PIFLUX3D ? -m.fermi_energy : m.fermi_energy;
        const double y = m.beta * c;
        J.fermi_w = m.relax == HHG_RELAX_EQUILIBRIUM && m.beta > 0.0 && std::fabs(y) <= 345.0 ? std::exp(y) : 0.0;
    }
    J.split_cut = m.split_dispersion;
    J.perp_current = m.perp_current;
    J.n_outputs = J.split_cut > 0.0 || J.perp_current ? 2 : 1;
    J.pulse = p.kind;
    J.theta_zero = m.theta == 0.0;
    J.amp_a = p.amp_a;
    J.amp_b = p.amp_b;
    J.inv_n = p.n_cycles > 0.0 ? 1.0 / p.n_cycles : 0.0;
    J.t_env_end = p.t0 + 2.0 * 3.14159265358979323846 * p.n_cycles;
    J.inv_2s2 = p.sigma > 0.0 ? 1.0 / (2.0 * p.sigma * p.sigma) : 0.0;
    J.t0 = p.t0;
    J.delay = p.delay;
    J.cep = p.cep;
    J.t_begin = jd.time.t_begin;
    J.delta = (jd.time.t_end - jd.time.t_begin) / jd.time.n_measurements;
    J.n_meas = jd.time.n_measurements;
    J.n_sub = jd.time.n_subdivisions;
    J.magnus_sub = sv.method == HHG_METHOD_MAGNUS4 ? std::max(1, sv.magnus_substeps) : 0;
    set_envelope_corners(J, p, sv);
    return J;
}

void destroy_device(hhg_plan* p) {
    if (p->graph_exec) hipGraphExecDestroy(p->graph_exec);
    if (p->graph) hipGraphDestroy(p->graph);
    for (auto& e : p->ev)
        if (e) hipEventDestroy(e);
    if (p->ev_done) hipEventDestroy(p->ev_done);
    hipFree(p->d_aux);
    hipFree(p->d_kx);
    hipFree(p->d_ky);
    hipFree(p->d_kz);
    hipFree(p->d_spl);
    hipFree(p->d_w);
    hipFree(p->d_jobs);
    hipFree(p->d_items);
    hipFree(p->d_At);
    hipFree(p->d_Pt);
    hipFree(p->d_h_first);
    hipFree(p->d_Am);
    hipFree(p->d_nom);
    hipFree(p->d_perm);
    hipFree(p->d_log);
    hipFree(p->d_words);
    hipFree(p->d_j);
    if (p->stream) hipStreamDestroy(p->stream);
}

// the solve kernel a group launches (a fused group: the two groups' fused kernel)
const void* group_fn(const hhg_plan* p, const Group& g) {
    if (g.fused) {
        const int k1[3] = {g.model, g.pulse, g.relax}, k2[3] = {g.model2, g.pulse2, g.relax2};
        return hhg::solve_fused_fn(k1, k2, p->solve_learn());
    }
    return hhg::solve_kernel_fn(p->solver.method, g.model, g.pulse, g.relax, p->solve_learn());
}

int max_block_threads(const hhg_plan* p, const Group& g) {
    const void* fn = group_fn(p, g);
    hipFuncAttributes attr{};
    if (!fn || hipFuncGetAttributes(&attr, fn) != hipSuccess) return 256;
    return attr.maxThreadsPerBlock;
}

int blocks_per_cu(const hhg_plan* p, const Group& g, int threads, size_t smem) {
    const void* fn = group_fn(p, g);
    int per_cu = 0;
    if (smem > p->lds_per_block) return 0;
    if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, smem) != hipSuccess) return 0;
    return per_cu;
}

int resident_blocks(const hhg_plan* p, const Group& g, int threads) {
    return std::max(1, blocks_per_cu(p, g, threads, p->smem(g))) * p->num_cus;
}

// Launch shape of a group.  Persistent launches with block_reduce 2
// accumulate in LDS per block where the block's LDS copy of the largest job
// output fits; waves_per_block 0 takes 8 there (one block of the kernel's 8
// waves per CU), else 4.
void size_group(const hhg_plan* p, Group& g, int waves_per_block, int queue_chunk, int persistent) {
    const bool want_block = persistent && p->launch.block_reduce == 2;
    auto fits = [&](int wpb) {
        Group t = g;
        t.block_acc = 1;
        return blocks_per_cu(p, t, 64 * wpb, p->smem(t)) > 0;
    };
    const int wpb_max = std::max(1, max_block_threads(p, g) / 64);
    int wpb = waves_per_block;
    if (wpb <= 0) wpb = (want_block && wpb_max >= 8 && fits(8)) ? 8 : std::min(4, wpb_max);
    g.waves_per_block = wpb;
    g.queue_chunk = queue_chunk;
    g.block_acc = want_block && fits(wpb) ? 1 : 0;
    const int items_blocks = (g.n_items + wpb - 1) / wpb;
    if (persistent) {
        const int res = resident_blocks(p, g, 64 * wpb);
        g.blocks = std::max(1, std::min(items_blocks, res));
    } else {
        g.queue_chunk = 1;
        g.blocks = std::max(1, items_blocks);
    }
}

hhg::KernelArgs launch_args(const hhg_plan* p, const Group& g, int gi, int n_items_limit) {
    hhg::KernelArgs a{};
    a.jobs = p->d_jobs;
    a.items = p->d_items + g.item_begin;
    a.n_items = n_items_limit > 0 ? std::min(g.n_items, n_items_limit) : g.n_items;
    a.queue_chunk = g.queue_chunk;
    a.kx = p->d_kx;
    a.ky = p->d_ky;
    a.kz = p->d_kz;
    a.w = p->d_w;
    a.At = p->d_At;
    a.Pt = p->d_Pt;
    a.Am = p->d_Am;
    a.nom = p->d_nom;
    a.acc = p->acc();
    a.ctr = p->counters(gi);
    a.job_fail = p->job_fail();
    a.perm = p->d_perm;
    a.mlp = p->d_mlp;
    a.infer_every = p->solver.mlp_infer_every > 0 ? p->solver.mlp_infer_every : 1;
    a.fallback_after = p->solver.fallback_after > 0 ? p->solver.fallback_after : 2;
    a.infer_policy = p->solver.infer_policy;
    a.count_waves = p->launch.count_wave_attempts;
    a.loop_only = p->launch.loop_only;
    a.block_acc = g.block_acc;
    a.acc_lds_off = (int)p->acc_lds_off();
    a.acc_lds_samples = g.acc_samples;
    a.log = p->d_log;
    a.log_count = &p->counters(0)->log_count;
    a.log_cap = p->log_cap;
    a.log_stride = p->log_stride;
    a.sp = solver_dev(p->solver);
    return a;
}

// The execution's kernels on `st`.  events: the timing events around the
// solve launches and the reduction (a graph capture leaves them out: run()
// times a graph replay from outside the graph).
hhg_status enqueue(hhg_plan* p, hipStream_t st, int n_items_limit, bool events = true) {
    HHG_TRY(hhg::launch_clear(p->d_words, p->n_words, st), "clear");
    if (events) HHG_TRY(hipEventRecord(p->ev[0], st), "event");
    if (p->first_step())
        for (size_t gi = 0; gi < p->groups.size(); ++gi) {
            const hhg::KernelArgs a = launch_args(p, p->groups[gi], (int)gi, 0);
            HHG_TRY(hhg::launch_first_step(p->groups[gi].model, p->groups[gi].pulse, a, p->d_h_first,
                                           hhg::kMlpLdsBytes, st),
                    "first-step launch");
        }
    for (size_t li = 0; li < p->launches.size(); ++li) {
        const Group& g = p->launches[li];
        hhg::KernelArgs a = launch_args(p, g, (int)li, n_items_limit);
        if (p->first_step()) a.h_first = p->d_h_first;
        int blocks = g.blocks;
        if (n_items_limit > 0) blocks = std::min(blocks, (a.n_items + g.waves_per_block - 1) / g.waves_per_block);
        blocks = std::max(1, blocks);
        if (g.fused) {
            const int k1[3] = {g.model, g.pulse, g.relax}, k2[3] = {g.model2, g.pulse2, g.relax2};
            HHG_TRY(hhg::launch_solve_fused(k1, k2, p->solve_learn(), a, blocks, 64 * g.waves_per_block, p->smem(g),
                                            st),
                    "fused solve launch");
        } else {
            HHG_TRY(hhg::launch_solve(p->solver.method, g.model, g.pulse, g.relax, p->solve_learn(), a, blocks,
                                      64 * g.waves_per_block, p->smem(g), st),
                    "solve launch");
        }
    }
    if (events) HHG_TRY(hipEventRecord(p->ev[1], st), "event");
    HHG_TRY(hhg::launch_finalize(p->acc(), p->n_out, p->d_j, st), "finalize");
    if (events) HHG_TRY(hipEventRecord(p->ev[2], st), "event");
    return HHG_SUCCESS;
}

hhg_status collect_stats(hhg_plan* p, hhg_stats* stats, double wall_ms) {
    std::vector<hhg::Counters> c(p->groups.size());
    HHG_TRY(hipMemcpy(c.data(), p->d_words, sizeof(hhg::Counters) * c.size(), hipMemcpyDeviceToHost), "counters");
    hhg_stats s{};
    for (const auto& j : p->jobs) s.trajectories += j.n_k;
    for (const auto& x : c) {
        s.attempts += (int64_t)x.attempts;
        s.rejects += (int64_t)x.rejects;
        s.failed_trajectories += (int64_t)x.failed;
        s.nonfinite_trajectories += (int64_t)x.nonfinite;
        s.proposal_fallbacks += (int64_t)x.fallbacks;
        s.wave_attempts += (int64_t)x.wave_attempts;
        s.accumulator_overflows += (int64_t)x.overflows;
    }
    float a = 0.f, b = 0.f;
    if (hipEventElapsedTime(&a, p->ev[0], p->ev[1]) == hipSuccess) s.solve_ms = a;
    if (hipEventElapsedTime(&b, p->ev[1], p->ev[2]) == hipSuccess) s.reduce_ms = b;
    s.total_ms = wall_ms;
    if (stats) *stats = s;
    if (s.accumulator_overflows > 0)
        return fail(HHG_ERROR_NONFINITE, std::to_string(s.accumulator_overflows) +
                                             " lanes' sample sums left the accumulators' range (|sum| >= 2^46)");
    if (s.nonfinite_trajectories > 0)
        return fail(HHG_ERROR_NONFINITE, std::to_string(s.nonfinite_trajectories) + " trajectories became non-finite");
    if (s.failed_trajectories > 0)
        return fail(HHG_ERROR_TOLERANCE,
                    std::to_string(s.failed_trajectories) + " trajectories exceeded the step controller's limits");
    return HHG_SUCCESS;
}

hhg_status run(hhg_plan* p, hipStream_t st, int n_items_limit = 0) {
    if (p->graph_exec && n_items_limit == 0 && st == p->stream) {
        // solve_ms of a replay covers the whole graph (the clear and finalize
        // kernels add microseconds); reduce_ms is 0
        TraceRange range("hhg graph launch");
        HHG_TRY(hipEventRecord(p->ev[0], st), "event");
        HHG_TRY(hipGraphLaunch(p->graph_exec, st), "graph launch");
        HHG_TRY(hipEventRecord(p->ev[1], st), "event");
        HHG_TRY(hipEventRecord(p->ev[2], st), "event");
    } else {
        TraceRange range("hhg enqueue");
        const hhg_status r = enqueue(p, st, n_items_limit);
        if (r != HHG_SUCCESS) return r;
    }
    HHG_TRY(hipEventRecord(p->ev_done, st), "event");
    p->executed = true;
    return HHG_SUCCESS;
}

// waits for the last execution, on whichever stream it ran
hhg_status wait_last(hhg_plan* p) {
    HHG_TRY(hipSetDevice(p->device), "hipSetDevice");
    if (p->executed) HHG_TRY(hipEventSynchronize(p->ev_done), "synchronize");
    return HHG_SUCCESS;
}

}  // namespace

extern "C" {

hhg_status hhg_plan_create(const hhg_plan_desc* d, hhg_plan** out) {
    hhg::clear_error();
    if (!d || !out) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_plan_create: null argument");
    *out = nullptr;
    if (d->n_jobs < 1 || !d->jobs || d->n_kgrids < 1 || !d->kgrids)
        return fail(HHG_ERROR_INVALID_ARGUMENT, "plan needs >= 1 job and >= 1 k-grid");
    const int n_ranks = d->n_ranks > 0 ? d->n_ranks : 1;
    if (d->rank < 0 || d->rank >= n_ranks) return fail(HHG_ERROR_INVALID_ARGUMENT, "rank outside [0, n_ranks)");
    const hhg_solver_desc& sv = d->solver;
    if (sv.method != HHG_METHOD_RKF78 && sv.method != HHG_METHOD_MAGNUS4 && sv.method != HHG_METHOD_MAGNUS_MOMENTS)
        return fail(HHG_ERROR_INVALID_ARGUMENT, "unknown method");
    if (sv.method == HHG_METHOD_RKF78 && (!(sv.rtol >= 0.0) || !(sv.atol >= 0.0) || !(sv.rtol + sv.atol > 0.0) ||
                                          !std::isfinite(sv.rtol) || !std::isfinite(sv.atol)))
        return fail(HHG_ERROR_INVALID_ARGUMENT, "RKF78 needs finite rtol, atol >= 0, not both 0");
    if (sv.proposal < 0 || sv.proposal > 2) return fail(HHG_ERROR_INVALID_ARGUMENT, "unknown proposal");
    if (sv.proposal == HHG_PROPOSAL_LEARNED && (!sv.mlp || sv.method != HHG_METHOD_RKF78))
        return fail(HHG_ERROR_INVALID_ARGUMENT, "the learned proposal needs RKF78 and a network (solver.mlp)");
    if (d->launch.group_by_cost < 0)
        return fail(HHG_ERROR_INVALID_ARGUMENT, "group_by_cost must be >= 0 (the number of cost classes; 1 = 64)");
    if (d->launch.group_by_cost && !sv.mlp)
        return fail(HHG_ERROR_INVALID_ARGUMENT, "grouping by predicted cost needs a network (solver.mlp)");
    if (sv.mlp && hhg::mlp_device(sv.mlp) != d->device)
        return fail(HHG_ERROR_INVALID_ARGUMENT, "the network lives on another device than the plan");
    if (sv.mlp_infer_every < 0 || sv.fallback_after < 0 || sv.max_steps < 0)
        return fail(HHG_ERROR_INVALID_ARGUMENT, "mlp_infer_every, fallback_after and max_steps must be >= 0");
    if (sv.infer_policy != HHG_INFER_FIRST && sv.infer_policy != HHG_INFER_SCHEDULED)
        return fail(HHG_ERROR_INVALID_ARGUMENT, "unknown infer_policy");
    for (int i = 0; i < d->n_jobs; ++i) {
        const hhg_status st = hhg::validate_job(d->jobs[i], sv, d->n_kgrids, i);
        if (st != HHG_SUCCESS) return st;
    }
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev < 1) return fail(HHG_ERROR_NO_DEVICE, "no HIP device");
    if (d->device < 0 || d->device >= n_dev) return fail(HHG_ERROR_INVALID_ARGUMENT, "device ordinal out of range");

    hhg_plan* p = new hhg_plan();
    p->device = d->device;
    p->solver = sv;
    p->launch = d->launch;
    p->n_jobs = d->n_jobs;
    p->job_desc.assign(d->jobs, d->jobs + d->n_jobs);
    auto bail = [&](hhg_status st) {
        destroy_device(p);
        delete p;
        return st;
    };
    if (hipSetDevice(p->device) != hipSuccess) return bail(fail(HHG_ERROR_HIP, "hipSetDevice failed"));
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, p->device) != hipSuccess)
            return bail(fail(HHG_ERROR_HIP, "hipGetDeviceProperties failed"));
        p->num_cus = prop.multiProcessorCount;
        p->lds_per_block = prop.sharedMemPerBlock;
    }

    // k-grids, concatenated
    std::vector<int64_t> grid_base(d->n_kgrids + 1, 0), grid_size(d->n_kgrids);
    for (int g = 0; g < d->n_kgrids; ++g) {
        grid_size[g] = hhg::kgrid_size(d->kgrids[g]);
        if (grid_size[g] < 0) return bail(fail(HHG_ERROR_INVALID_ARGUMENT, "bad k-grid size"));
        grid_base[g + 1] = grid_base[g] + grid_size[g];
    }
    p->n_kpts = grid_base[d->n_kgrids];
    std::vector<double> kx(p->n_kpts), ky(p->n_kpts), kz(p->n_kpts), w(p->n_kpts);
    for (int g = 0; g < d->n_kgrids; ++g) {
        const hhg_status st = hhg::kgrid_build(d->kgrids[g], kx.data() + grid_base[g], ky.data() + grid_base[g],
                                               kz.data() + grid_base[g], w.data() + grid_base[g]);
        if (st != HHG_SUCCESS) return bail(st);
    }
    bool has_kz = false;  // only the 3D pi-flux lattice reads the third coordinate
    for (int i = 0; i < d->n_jobs; ++i) has_kz |= d->jobs[i].model.kind == HHG_MODEL_PIFLUX3D;

    // jobs: shard = contiguous block of 64-point chunks of each job's grid
    p->offsets.assign(d->n_jobs + 1, 0);
    int64_t magnus_words = 0, nominal_words = 0, traj_total = 0;
    std::vector<double> spl;             // spline pulse coefficients, concatenated
    std::vector<int64_t> spl_off(d->n_jobs, -1);
    for (int i = 0; i < d->n_jobs; ++i) {
        hhg::JobDev J = make_job(d->jobs[i], sv);
        const int g = d->jobs[i].kgrid;
        if (J.model == hhg::MODEL_PIFLUX3D)
            J.ic_rescue = d->kgrids[g].kind == HHG_KGRID_PIFLUX_CONTINUUM ? 1
                          : (d->kgrids[g].kind == HHG_KGRID_PIFLUX_UNIFORM ||
                             d->kgrids[g].kind == HHG_KGRID_PIFLUX_OCCUPATION) ? 2
                                                                               : 0;
        const int64_t n_chunks = (grid_size[g] + hhg::WAVE - 1) / hhg::WAVE;
        const int64_t per_rank = (n_chunks + n_ranks - 1) / n_ranks;
        const int64_t c0 = std::min(n_chunks, per_rank * d->rank), c1 = std::min(n_chunks, c0 + per_rank);
        J.k_base = grid_base[g] + c0 * hhg::WAVE;
        J.n_k = std::max<int64_t>(0, std::min(grid_size[g], c1 * hhg::WAVE) - c0 * hhg::WAVE);
        J.n_items = (int32_t)(c1 - c0);
        J.sample_base = p->offsets[i];
        J.traj_base = traj_total;
        traj_total += J.n_k;
        p->offsets[i + 1] = p->offsets[i] + J.n_outputs * (int64_t)(J.n_meas + 1);
        J.magnus_base = magnus_words;
        if (sv.method == HHG_METHOD_MAGNUS4) magnus_words += 2 * (int64_t)J.n_meas * J.magnus_sub;
        if (sv.method == HHG_METHOD_MAGNUS_MOMENTS) magnus_words += 4 * (int64_t)J.n_meas;
        J.nominal_base = nominal_words;
        if (sv.method == HHG_METHOD_RKF78) nominal_words += (int64_t)J.n_meas * hhg::nom_row(J.model, J.relax);
        J.row_base = 0;
        const hhg_pulse_desc& pd = d->jobs[i].pulse;
        if (pd.kind == HHG_PULSE_SPLINE) {
            const hhg::SplinePulse sp(pd);
            if (sp.status != HHG_SUCCESS) return bail(sp.status);
            spl_off[i] = (int64_t)spl.size();
            spl.insert(spl.end(), sp.coef.begin(), sp.coef.end());
            J.spl_t0 = sp.t0;
            J.spl_inv_h = sp.inv_h;
            J.spl_t_last = sp.t_last;
            J.spl_n = (int32_t)sp.n;
        }
        p->jobs.push_back(J);
    }
    p->n_out = p->offsets[d->n_jobs];

    // groups by (model, pulse, relax); inside a group, longest jobs first
    std::map<std::tuple<int, int, int>, std::vector<int>> by_key;
    for (int i = 0; i < d->n_jobs; ++i) {
        const auto& J = p->jobs[i];
        const int pulse = sv.method != HHG_METHOD_RKF78 ? 0 : J.pulse;  // Magnus kernels read only tables
        by_key[{J.model, pulse, J.relax}].push_back(i);
    }
    std::vector<int2> items;
    for (auto& kv : by_key) {
        auto& js = kv.second;
        std::stable_sort(js.begin(), js.end(), [&](int a, int b) { return p->jobs[a].n_meas > p->jobs[b].n_meas; });
        Group g{};
        std::tie(g.model, g.pulse, g.relax) = kv.first;
        g.item_begin = (int)items.size();
        for (int j : js)
            for (int c = 0; c < p->jobs[j].n_items; ++c) items.push_back(make_int2(j, c));
        g.n_items = (int)items.size() - g.item_begin;
        g.acc_samples = 0;
        for (int j : js)
            g.acc_samples = std::max(g.acc_samples, (int)(p->offsets[j + 1] - p->offsets[j]));
        if (g.n_items > 0) p->groups.push_back(g);
    }
    if (items.size() > (size_t)0x7fffffff) return bail(fail(HHG_ERROR_UNSUPPORTED, "more than 2^31 work items"));

    // device memory
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(HHG_ERROR_HIP, "stream creation failed"));
    for (auto& e : p->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(HHG_ERROR_HIP, "event creation failed"));
    if (hipEventCreateWithFlags(&p->ev_done, hipEventDisableTiming) != hipSuccess)
        return bail(fail(HHG_ERROR_HIP, "event creation failed"));
    p->n_words = p->n_counter_blocks() * kCounterWords + 2 * p->n_out + 2 * (int64_t)d->n_jobs;
    const size_t nk = sizeof(double) * std::max<int64_t>(1, p->n_kpts);
    hhg_status st = HHG_SUCCESS;
    auto A = [&](void** ptr, size_t bytes, const char* what) {
        if (st == HHG_SUCCESS) st = check_hip(hipMalloc(ptr, std::max<size_t>(bytes, 8)), what);
    };
    A((void**)&p->d_kx, nk, "kx");
    A((void**)&p->d_ky, nk, "ky");
    if (has_kz) A((void**)&p->d_kz, nk, "kz");
    A((void**)&p->d_w, nk, "w");
    A((void**)&p->d_jobs, sizeof(hhg::JobDev) * p->jobs.size(), "jobs");
    A((void**)&p->d_items, sizeof(int2) * std::max<size_t>(1, items.size()), "items");
    A((void**)&p->d_At, sizeof(double) * p->n_out, "pulse samples");
    A((void**)&p->d_Pt, 2 * sizeof(double) * p->n_out, "pulse sample phases");
    A((void**)&p->d_Am, sizeof(double) * std::max<int64_t>(1, magnus_words), "magnus table");
    if (nominal_words > 0) A((void**)&p->d_nom, sizeof(double) * nominal_words, "nominal-step table");
    A((void**)&p->d_words, sizeof(unsigned long long) * p->n_words, "accumulators");
    A((void**)&p->d_j, sizeof(double) * p->n_out, "j");
    A((void**)&p->d_aux, sizeof(hhg::Counters), "trajectory counters");
    if (p->first_step()) A((void**)&p->d_h_first, sizeof(double) * std::max<int64_t>(1, traj_total), "first steps");
    if (st != HHG_SUCCESS) return bail(st);
    auto C = [&](void* dst, const void* src, size_t bytes, const char* what) {
        if (st == HHG_SUCCESS && bytes) st = check_hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), what);
    };
    C(p->d_kx, kx.data(), sizeof(double) * p->n_kpts, "upload kx");
    C(p->d_ky, ky.data(), sizeof(double) * p->n_kpts, "upload ky");
    if (has_kz) C(p->d_kz, kz.data(), sizeof(double) * p->n_kpts, "upload kz");
    C(p->d_w, w.data(), sizeof(double) * p->n_kpts, "upload w");
    if (!spl.empty()) {
        A((void**)&p->d_spl, sizeof(double) * spl.size(), "spline coefficients");
        C(p->d_spl, spl.data(), sizeof(double) * spl.size(), "upload spline coefficients");
        for (int i = 0; i < d->n_jobs; ++i)
            if (spl_off[i] >= 0) p->jobs[i].spl = p->d_spl + spl_off[i];
    }
    C(p->d_jobs, p->jobs.data(), sizeof(hhg::JobDev) * p->jobs.size(), "upload jobs");
    C(p->d_items, items.data(), sizeof(int2) * items.size(), "upload items");
    if (st != HHG_SUCCESS) return bail(st);
    hhg::Gauss30 g30{};
    {
        double xi[30], wi[30];
        hhg::gauss_legendre(30, xi, wi);
        for (int q = 0; q < 30; ++q) g30.x[q] = 0.5 * (xi[q] + 1.0), g30.w[q] = 0.5 * wi[q];
    }
    const int with_magnus = sv.method == HHG_METHOD_MAGNUS4 ? 1 : sv.method == HHG_METHOD_MAGNUS_MOMENTS ? 2 : 0;
    for (int i = 0; i < d->n_jobs; ++i) {
        const auto& J = p->jobs[i];
        const int64_t n = std::max<int64_t>(J.n_meas + 1, (int64_t)J.n_meas * J.magnus_sub);
        st = check_hip(
            hhg::launch_pulse_table(J.pulse, p->d_jobs, i, n, p->d_At, p->d_Pt, p->d_Am, with_magnus, g30, p->stream),
            "pulse table");
        if (st != HHG_SUCCESS) return bail(st);
        if (p->d_nom) {
            st = check_hip(hhg::launch_nominal_table(J.pulse, p->d_jobs, i, J.n_meas, p->d_nom, p->stream),
                           "nominal-step table");
            if (st != HHG_SUCCESS) return bail(st);
        }
    }
    if ((st = check_hip(hipStreamSynchronize(p->stream), "pulse table sync")) != HHG_SUCCESS) return bail(st);

    if (sv.mlp) p->d_mlp = hhg::mlp_device_record(sv.mlp);

    // intervals the cost probe sums the network's attempts output over
#ifndef HHG_COST_PROBES
#define HHG_COST_PROBES 16
#endif
    constexpr int kCostProbes = HHG_COST_PROBES;
    // lockstep grouping: order each job's k-points by the network's predicted
    // attempts so that the lanes of a wave need similar work
    if (p->launch.group_by_cost) {
        float* d_cost = nullptr;
        if ((st = check_hip(hipMalloc(&d_cost, sizeof(float) * std::max<int64_t>(1, p->n_kpts)), "cost")) != HHG_SUCCESS)
            return bail(st);
        if ((st = check_hip(hipMalloc(&p->d_perm, sizeof(int64_t) * std::max<int64_t>(1, p->n_kpts)), "perm")) !=
            HHG_SUCCESS) {
            hipFree(d_cost);
            return bail(st);
        }
        for (const auto& g : p->groups) {
            hhg::KernelArgs a{};
            a.jobs = p->d_jobs;
            a.items = p->d_items + g.item_begin;
            a.n_items = g.n_items;
            a.kx = p->d_kx;
            a.ky = p->d_ky;
            a.kz = p->d_kz;
            a.At = p->d_At;
            a.mlp = p->d_mlp;
            if (st == HHG_SUCCESS)
                st = check_hip(hhg::launch_cost_probe(g.model, g.pulse, a, kCostProbes, d_cost,
                                                      hhg::kMlpLdsBytes, p->stream),
                               "cost probe");
        }
        std::vector<float> cost(p->n_kpts);
        if (st == HHG_SUCCESS) st = check_hip(hipStreamSynchronize(p->stream), "cost probe sync");
        if (st == HHG_SUCCESS)
            st = check_hip(hipMemcpy(cost.data(), d_cost, sizeof(float) * p->n_kpts, hipMemcpyDeviceToHost), "cost");
        hipFree(d_cost);
        if (st != HHG_SUCCESS) return bail(st);
        // cost classes: the quantiles of the job's predictions; costliest class
        // first, grid order inside a class
        const int64_t n_cls = p->launch.group_by_cost == 1 ? 64 : p->launch.group_by_cost;
        for (auto& x : cost)  // a NaN prediction counts as the costliest (keeps the sorts well-ordered)
            if (x != x) x = std::numeric_limits<float>::infinity();
        std::vector<int64_t> perm(p->n_kpts, 0);
        for (const auto& J : p->jobs) {
            if (J.n_k == 0) continue;
            const float* c = cost.data() + J.k_base;
            std::vector<float> sorted(c, c + J.n_k);
            std::sort(sorted.begin(), sorted.end());
            const int64_t nc = std::min<int64_t>(n_cls, J.n_k);
            std::vector<float> edge;  // class q holds costs in [edge[q - 1], edge[q])
            for (int64_t q = 1; q < nc; ++q) edge.push_back(sorted[(size_t)((q * J.n_k) / nc)]);
            auto cls = [&](int64_t r) {
                return (int64_t)(std::upper_bound(edge.begin(), edge.end(), c[r]) - edge.begin());
            };
            std::vector<int64_t> order(J.n_k), key(J.n_k);
            for (int64_t r = 0; r < J.n_k; ++r) order[r] = r, key[r] = cls(r);
            std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return key[x] > key[y]; });
            for (int64_t r = 0; r < J.n_k; ++r) perm[J.k_base + r] = order[r];
        }
        if ((st = check_hip(hipMemcpy(p->d_perm, perm.data(), sizeof(int64_t) * p->n_kpts, hipMemcpyHostToDevice),
                            "upload perm")) != HHG_SUCCESS)
            return bail(st);
    }

    // launch shapes
    const int wpb = p->launch.waves_per_block;  // 0: chosen per group
    const int qc = p->launch.queue_chunk > 0 ? p->launch.queue_chunk : 1;
    if (wpb < 0 || wpb > 8 || (wpb & (wpb - 1)))
        return bail(fail(HHG_ERROR_INVALID_ARGUMENT, "waves_per_block must be 0, 1, 2, 4 or 8"));
    if (p->launch.block_reduce < 0 || p->launch.block_reduce > 2)
        return bail(fail(HHG_ERROR_INVALID_ARGUMENT, "block_reduce must be 0, 1 or 2"));
    // one solve launch per group, or one fused launch where the plan has two
    // RKF78 groups whose fused kernel is built (launch.separate_launches 0)
    p->launches = p->groups;
    if (sv.method == HHG_METHOD_RKF78 && p->groups.size() == 2 && !p->launch.separate_launches) {
        Group f = p->groups[0];
        const Group& g2 = p->groups[1];
        f.fused = 1;
        f.model2 = g2.model, f.pulse2 = g2.pulse, f.relax2 = g2.relax;
        f.item_begin = 0;
        f.n_items = g2.item_begin + g2.n_items;
        f.acc_samples = std::max(f.acc_samples, g2.acc_samples);
        if (group_fn(p, f)) p->launches.assign(1, f);
    }
    for (auto& g : p->launches)
        if (64 * wpb > max_block_threads(p, g))
            return bail(fail(HHG_ERROR_INVALID_ARGUMENT, "waves_per_block exceeds the kernel's block size limit"));
    for (auto& g : p->launches) {
        size_group(p, g, wpb, qc, p->launch.persistent);
        if (p->launch.block_reduce == 2 && p->launch.persistent && !g.block_acc)
            return bail(fail(HHG_ERROR_UNSUPPORTED, "block_reduce 2: a job's output does not fit the block's LDS"));
    }
    *out = p;
    return HHG_SUCCESS;
}

int64_t hhg_plan_output_size(const hhg_plan* p) { return p ? p->n_out : -1; }

int32_t hhg_plan_solve_launches(const hhg_plan* p) { return p ? (int32_t)p->launches.size() : -1; }

hhg_status hhg_plan_job_offsets(const hhg_plan* p, int64_t* offsets) {
    hhg::clear_error();
    if (!p || !offsets) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_plan_job_offsets: null argument");
    std::copy(p->offsets.begin(), p->offsets.end(), offsets);
    return HHG_SUCCESS;
}

hhg_status hhg_plan_execute(hhg_plan* p, double* j_out, hhg_stats* stats) {
    hhg::clear_error();
    if (!p || !j_out) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_plan_execute: null argument");
    TraceRange range("hhg_plan_execute");
    const auto t0 = std::chrono::steady_clock::now();
    HHG_TRY(hipSetDevice(p->device), "hipSetDevice");
    hhg_status st = run(p, p->stream);
    if (st != HHG_SUCCESS) return st;
    HHG_TRY(hipMemcpyAsync(j_out, p->d_j, sizeof(double) * p->n_out, hipMemcpyDeviceToHost, p->stream), "copy j");
    HHG_TRY(hipStreamSynchronize(p->stream), "synchronize");
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    p->last_wall_ms = ms;
    return collect_stats(p, stats, ms);
}

hhg_status hhg_plan_execute_accumulators(hhg_plan* p, uint64_t* acc_out, hhg_stats* stats) {
    hhg::clear_error();
    if (!p || !acc_out) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_plan_execute_accumulators: null argument");
    TraceRange range("hhg_plan_execute_accumulators");
    const auto t0 = std::chrono::steady_clock::now();
    HHG_TRY(hipSetDevice(p->device), "hipSetDevice");
    hhg_status st = run(p, p->stream);
    if (st != HHG_SUCCESS) return st;
    HHG_TRY(hipMemcpyAsync(acc_out, p->acc(), sizeof(uint64_t) * 2 * p->n_out, hipMemcpyDeviceToHost, p->stream),
            "copy accumulators");
    HHG_TRY(hipStreamSynchronize(p->stream), "synchronize");
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    p->last_wall_ms = ms;
    return collect_stats(p, stats, ms);
}

hhg_status hhg_plan_execute_device(hhg_plan* p, double* j_dev, void* stream) {
    hhg::clear_error();
    if (!p || !j_dev) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_plan_execute_device: null argument");
    TraceRange range("hhg_plan_execute_device");
    HHG_TRY(hipSetDevice(p->device), "hipSetDevice");
    hipStream_t st = stream ? (hipStream_t)stream : p->stream;
    hhg_status r = run(p, st);
    if (r != HHG_SUCCESS) return r;
    p->last_wall_ms = 0.0;
    HHG_TRY(hipMemcpyAsync(j_dev, p->d_j, sizeof(double) * p->n_out, hipMemcpyDeviceToDevice, st), "copy j");
    return HHG_SUCCESS;
}

hhg_status hhg_plan_last_stats(hhg_plan* p, hhg_stats* stats) {
    hhg::clear_error();
    if (!p || !stats) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_plan_last_stats: null argument");
    if (!p->executed) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_plan_last_stats: the plan has not been executed");
    const hhg_status st = wait_last(p);
    if (st != HHG_SUCCESS) return st;
    return collect_stats(p, stats, p->last_wall_ms);
}

hhg_status hhg_plan_autotune(hhg_plan* p, hhg_launch_desc* chosen) {
    hhg::clear_error();
    if (!p) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_plan_autotune: null argument");
    TraceRange range("hhg_plan_autotune");
    HHG_TRY(hipSetDevice(p->device), "hipSetDevice");
    if (p->graph_exec) return fail(HHG_ERROR_UNSUPPORTED, "autotune before capturing the graph");
    // Tune on a prefix of every group's items: enough waves to fill the chip a
    // few times, so the measured rate is the steady-state one.
    const int wpbs[] = {1, 2, 4, 8};
    const int chunks[] = {1, 2, 4};
    double best = 1e300;
    int best_w = 4, best_c = 1;
    const std::vector<Group> saved = p->launches;
    int wpb_max = 8;
    for (const auto& g : p->launches) wpb_max = std::min(wpb_max, max_block_threads(p, g) / 64);
    for (int wpb : wpbs)
        for (int qc : chunks) {
            if (wpb > wpb_max) continue;  // beyond a kernel's block size limit
            int limit = 0;
            for (auto& g : p->launches) {
                size_group(p, g, wpb, qc, 1);
                limit = std::max(limit, 3 * g.blocks * wpb);
            }
            hhg_status st = enqueue(p, p->stream, limit);
            if (st != HHG_SUCCESS) return st;
            HHG_TRY(hipStreamSynchronize(p->stream), "autotune sync");
            float ms = 0.f;
            HHG_TRY(hipEventElapsedTime(&ms, p->ev[0], p->ev[1]), "autotune timing");
            // normalise by the items actually run
            int64_t ran = 0;
            for (auto& g : p->launches) ran += std::min(g.n_items, limit);
            const double per_item = ms / std::max<int64_t>(1, ran);
            if (per_item < best) best = per_item, best_w = wpb, best_c = qc;
        }
    p->launches = saved;
    p->launch.waves_per_block = best_w;
    p->launch.queue_chunk = best_c;
    p->launch.persistent = 1;
    for (auto& g : p->launches) size_group(p, g, best_w, best_c, 1);
    // the tuning runs were partial solves: leave no counts of them behind
    HHG_TRY(hhg::launch_clear(p->d_words, p->n_words, p->stream), "clear");
    HHG_TRY(hipStreamSynchronize(p->stream), "autotune sync");
    p->executed = false;
    if (chosen) *chosen = p->launch;
    return HHG_SUCCESS;
}

hhg_status hhg_plan_capture_graph(hhg_plan* p) {
    hhg::clear_error();
    if (!p) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_plan_capture_graph: null argument");
    TraceRange range("hhg_plan_capture_graph");
    HHG_TRY(hipSetDevice(p->device), "hipSetDevice");
    if (p->graph_exec) return HHG_SUCCESS;
    HHG_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    const hhg_status st = enqueue(p, p->stream, 0, false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(p->stream, &g);
    if (st != HHG_SUCCESS) {
        if (g) hipGraphDestroy(g);
        return st;
    }
    HHG_TRY(e, "end capture");
    p->graph = g;
    HHG_TRY(hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0), "graph instantiate");
    return HHG_SUCCESS;
}

hhg_status hhg_plan_set_step_log(hhg_plan* p, int64_t capacity, int32_t stride) {
    hhg::clear_error();
    if (!p || capacity < 0 || stride < 1) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_plan_set_step_log: bad argument");
    if (p->solver.method != HHG_METHOD_RKF78) return fail(HHG_ERROR_UNSUPPORTED, "step logs need RKF78");
    HHG_TRY(hipSetDevice(p->device), "hipSetDevice");
    hipFree(p->d_log);
    p->d_log = nullptr;
    p->log_cap = 0;
    if (capacity > 0) {
        HHG_TRY(hipMalloc(&p->d_log, sizeof(float) * (hhg::kMlpIn + hhg::kMlpOut) * capacity), "step log");
        p->log_cap = capacity;
    }
    p->log_stride = stride;
    // the solves now run the logging (or the plain) kernel variant: launch
    // shapes for it, and a captured graph holds the old kernels and arguments
    for (auto& g : p->launches) size_group(p, g, g.waves_per_block, g.queue_chunk, p->launch.persistent);
    if (p->graph_exec) hipGraphExecDestroy(p->graph_exec), p->graph_exec = nullptr;
    if (p->graph) hipGraphDestroy(p->graph), p->graph = nullptr;
    return HHG_SUCCESS;
}

hhg_status hhg_plan_get_step_log(hhg_plan* p, float* records, int64_t capacity, int64_t* n_records) {
    hhg::clear_error();
    if (!p || !n_records || (capacity > 0 && !records))
        return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_plan_get_step_log: null argument");
    const hhg_status ws = wait_last(p);
    if (ws != HHG_SUCCESS) return ws;
    hhg::Counters c{};
    HHG_TRY(hipMemcpy(&c, p->d_words, sizeof(c), hipMemcpyDeviceToHost), "log count");
    const int64_t n = std::min<int64_t>((int64_t)c.log_count, p->log_cap);
    *n_records = n;
    const int64_t m = std::min(n, capacity);
    if (m > 0)
        HHG_TRY(hipMemcpy(records, p->d_log, sizeof(float) * (hhg::kMlpIn + hhg::kMlpOut) * m, hipMemcpyDeviceToHost),
                "step log");
    return HHG_SUCCESS;
}

hhg_status hhg_plan_destroy(hhg_plan* p) {
    hhg::clear_error();
    if (!p) return HHG_SUCCESS;
    hipSetDevice(p->device);
    destroy_device(p);
    delete p;
    return HHG_SUCCESS;
}

hhg_status hhg_trajectories(hhg_plan* p, const int64_t* k_index, int64_t n_sel, double* s_out) {
    hhg::clear_error();
    if (!p || (n_sel > 0 && (!k_index || !s_out))) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_trajectories: null argument");
    if (n_sel <= 0) return HHG_SUCCESS;
    const hhg::JobDev& J = p->jobs[0];
    for (int64_t i = 0; i < n_sel; ++i)
        if (k_index[i] < 0 || k_index[i] >= J.n_k)
            return fail(HHG_ERROR_INVALID_ARGUMENT, "k index outside job 0's k-points");
    HHG_TRY(hipSetDevice(p->device), "hipSetDevice");
    int64_t* d_sel = nullptr;
    double* d_s = nullptr;
    const size_t out_bytes = sizeof(double) * 3 * (size_t)(J.n_meas + 1) * (size_t)n_sel;
    hhg_status st = check_hip(hipMalloc(&d_sel, sizeof(int64_t) * n_sel), "selection");
    if (st == HHG_SUCCESS) st = check_hip(hipMalloc(&d_s, out_bytes), "trajectory output");
    if (st == HHG_SUCCESS) st = check_hip(hipMemcpy(d_sel, k_index, sizeof(int64_t) * n_sel, hipMemcpyHostToDevice), "upload");
    const int64_t aux_words = (int64_t)(sizeof(hhg::Counters) / sizeof(unsigned long long));
    if (st == HHG_SUCCESS)
        st = check_hip(hhg::launch_clear(reinterpret_cast<unsigned long long*>(p->d_aux), aux_words, p->stream), "clear");
    if (st == HHG_SUCCESS) {
        const hhg::SolverDev sp = solver_dev(p->solver);
        const int pulse = p->solver.method == HHG_METHOD_MAGNUS4 ? 0 : J.pulse;
        hhg::TrajArgs x{p->d_jobs, p->d_kx, p->d_ky, p->d_kz, d_sel, n_sel, p->d_Am, sp, d_s, p->d_aux, p->stream};
        st = check_hip(hhg::launch_trajectories(p->solver.method, J.model, pulse, J.relax, x), "trajectory launch");
    }
    if (st == HHG_SUCCESS) st = check_hip(hipStreamSynchronize(p->stream), "synchronize");
    if (st == HHG_SUCCESS) st = check_hip(hipMemcpy(s_out, d_s, out_bytes, hipMemcpyDeviceToHost), "download");
    hipFree(d_sel);
    hipFree(d_s);
    return st;
}

hhg_status hhg_trajectories_noalloc(hhg_plan* p, const int64_t* k_index, int64_t n_sel, double* s_out) {
    hhg::clear_error();
    if (!p || (n_sel > 0 && (!k_index || !s_out)))
        return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_trajectories_noalloc: null argument");
    if (n_sel <= 0) return HHG_SUCCESS;
    const hhg::JobDev& J = p->jobs[0];
    for (int64_t i = 0; i < n_sel; ++i)
        if (k_index[i] < 0 || k_index[i] >= J.n_k)
            return fail(HHG_ERROR_INVALID_ARGUMENT, "k index outside job 0's k-points");
    // workspace layout: counters, selection, output (each 256-byte aligned)
    const size_t off_sel = round_up_256(sizeof(hhg::Counters));
    const size_t out_bytes = sizeof(double) * 3 * (size_t)(J.n_meas + 1) * (size_t)n_sel;
    const size_t off_out = off_sel + round_up_256(sizeof(int64_t) * (size_t)n_sel);
    if (n_sel > (int64_t)(kScratchBytes / sizeof(int64_t)) || off_out + out_bytes > kScratchBytes)
        return fail(HHG_ERROR_INVALID_ARGUMENT,
                    "hhg_trajectories_noalloc: " + std::to_string(n_sel) + " trajectories of " +
                        std::to_string(J.n_meas + 1) + " samples need more than the " +
                        std::to_string(kScratchBytes >> 20) + " MiB static workspace");
    HHG_TRY(hipSetDevice(p->device), "hipSetDevice");
    unsigned char* base = nullptr;
    const hhg_status sst = scratch_on_device(p->device, &base);
    if (sst != HHG_SUCCESS) return sst;
    auto* d_ctr = reinterpret_cast<hhg::Counters*>(base);
    auto* d_sel = reinterpret_cast<int64_t*>(base + off_sel);
    auto* d_s = reinterpret_cast<double*>(base + off_out);
    HHG_TRY(hipMemcpyAsync(d_sel, k_index, sizeof(int64_t) * n_sel, hipMemcpyHostToDevice, p->stream), "upload");
    // the counters start at zero (fill kernel: the sequence may be captured)
    HHG_TRY(hhg::launch_clear(reinterpret_cast<unsigned long long*>(d_ctr),
                              (int64_t)(sizeof(hhg::Counters) / sizeof(unsigned long long)), p->stream),
            "clear");
    const hhg::SolverDev sp = solver_dev(p->solver);
    const int pulse = p->solver.method == HHG_METHOD_MAGNUS4 ? 0 : J.pulse;
    hhg::TrajArgs x{p->d_jobs, p->d_kx, p->d_ky, p->d_kz, d_sel, n_sel, p->d_Am, sp, d_s, d_ctr, p->stream};
    HHG_TRY(hhg::launch_trajectories(p->solver.method, J.model, pulse, J.relax, x), "trajectory launch");
    HHG_TRY(hipMemcpyAsync(s_out, d_s, out_bytes, hipMemcpyDeviceToHost, p->stream), "download");
    HHG_TRY(hipStreamSynchronize(p->stream), "synchronize");
    return HHG_SUCCESS;
}

hhg_status hhg_plan_job_failures(hhg_plan* p, int64_t* failed, int64_t* nonfinite) {
    hhg::clear_error();
    if (!p || !failed || !nonfinite) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_plan_job_failures: null argument");
    const int n_jobs = (int)p->jobs.size();
    std::vector<unsigned long long> w(2 * (size_t)n_jobs);
    const hhg_status ws = wait_last(p);
    if (ws != HHG_SUCCESS) return ws;
    HHG_TRY(hipStreamSynchronize(p->stream), "synchronize");
    HHG_TRY(hipMemcpy(w.data(), p->job_fail(), sizeof(unsigned long long) * w.size(), hipMemcpyDeviceToHost),
            "job failure counts");
    for (int j = 0; j < n_jobs; ++j) {
        failed[j] = (int64_t)w[2 * j];
        nonfinite[j] = (int64_t)w[2 * j + 1];
    }
    return HHG_SUCCESS;
}

hhg_status hhg_occupations(hhg_plan* p, const int64_t* k_index, int64_t n_sel, double* occ_out) {
    hhg::clear_error();
    if (!p || (n_sel > 0 && (!k_index || !occ_out)))
        return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_occupations: null argument");
    const hhg::JobDev& J = p->jobs[0];
    if (J.model != hhg::MODEL_PIFLUX3D || J.relax != hhg::RELAX_EQUILIBRIUM || p->solver.method != HHG_METHOD_RKF78)
        return fail(HHG_ERROR_INVALID_ARGUMENT,
                    "hhg_occupations: job 0 must be the 3D pi-flux lattice with equilibrium relaxation, solver RKF78");
    if (n_sel <= 0) return HHG_SUCCESS;
    for (int64_t i = 0; i < n_sel; ++i)
        if (k_index[i] < 0 || k_index[i] >= J.n_k)
            return fail(HHG_ERROR_INVALID_ARGUMENT, "k index outside job 0's k-points");
    HHG_TRY(hipSetDevice(p->device), "hipSetDevice");
    int64_t* d_sel = nullptr;
    double* d_o = nullptr;
    const size_t out_bytes = sizeof(double) * 2 * (size_t)(J.n_meas + 1) * (size_t)n_sel;
    hhg_status st = check_hip(hipMalloc(&d_sel, sizeof(int64_t) * n_sel), "selection");
    if (st == HHG_SUCCESS) st = check_hip(hipMalloc(&d_o, out_bytes), "occupation output");
    if (st == HHG_SUCCESS) st = check_hip(hipMemcpy(d_sel, k_index, sizeof(int64_t) * n_sel, hipMemcpyHostToDevice), "upload");
    const int64_t aux_words = (int64_t)(sizeof(hhg::Counters) / sizeof(unsigned long long));
    if (st == HHG_SUCCESS)
        st = check_hip(hhg::launch_clear(reinterpret_cast<unsigned long long*>(p->d_aux), aux_words, p->stream), "clear");
    if (st == HHG_SUCCESS) {
        const hhg::SolverDev sp = solver_dev(p->solver);
        hhg::TrajArgs x{p->d_jobs, p->d_kx, p->d_ky, p->d_kz, d_sel, n_sel, p->d_Am, sp, d_o, p->d_aux, p->stream};
        st = check_hip(hhg::launch_occupations(J.pulse, x), "occupation launch");
    }
    if (st == HHG_SUCCESS) st = check_hip(hipStreamSynchronize(p->stream), "synchronize");
    if (st == HHG_SUCCESS) st = check_hip(hipMemcpy(occ_out, d_o, out_bytes, hipMemcpyDeviceToHost), "download");
    hipFree(d_sel);
    hipFree(d_o);
    return st;
}

}  // extern "C"
