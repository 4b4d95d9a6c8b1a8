// hhg_solve.hpp — the solve kernels (RKF78 plain/carried/learned, Magnus),
// the per-k trajectory kernel and the cost probe, as templates instantiated
// once per model in hhg_solve_<model>.hip.
//
// Work decomposition: a work item is 64 consecutive k-points of one job, one
// per lane of a wavefront.  Lanes advance sampling interval by sampling
// interval (reference loop: sources/HHG/Systems/DiracSystem.cpp:166-173); at
// every sample t_i the wave sums w_k g_k(t_i) with a shuffle butterfly, lane
// (i mod 64) keeps the sum, and every 64 samples the wave adds its 64 sums to
// the job's exact 128-bit fixed-point accumulators (one vector atomic pair).
// Waves pop items from a device counter until the list is exhausted.
#pragma once
#include <hip/hip_runtime.h>

#include "hhg_device.hpp"
#include "hhg_kernels.hpp"
#include "hhg_mlp.hpp"

namespace hhg {

// Waves per SIMD each solve kernel is compiled for (measured on MI355X): the
// Magnus kernels fit 4 without spills (config 2's Magnus-4 leg 16.8 -> 22.6 M
// trajectories/s going from 2 to 4); the RKF78 kernels need 200-256 VGPRs and
// run at 2.  At 3 the learned 2D Dirac kernel ran 4-7 % faster but spilled 58
// VGPRs to scratch in its interval loop (43 GB of scratch writes per solve);
// at 4 the plain kernel gained 3-6 % and the carried controller lost 11 %.
// The equilibrium-relaxation kernels spill at 2 (3D pi-flux 468, honeycomb
// 280 bytes of scratch per lane); at 1 (blocks of at most 4 waves) they spill
// into AGPRs instead and use no scratch, but ran slower on the reference
// systems' configurations (3D pi-flux 874 vs 725 ms, honeycomb 751 vs 679 ms,
// spline double pulse 406 vs 328 ms), so they stay at 2 (HHG_EQ_WAVES).
// HHG_WAVES_PER_EU overrides all of them (occupancy experiments).

#ifndef HHG_EQ_WAVES
#define HHG_EQ_WAVES 2
#endif
template <int METHOD, int M, int R, int LEARN>
struct SolveOccupancy {
#ifdef HHG_WAVES_PER_EU
    static constexpr int waves = HHG_WAVES_PER_EU;
#else
    static constexpr int waves = METHOD != METHOD_RKF78 ? 4 : R == RELAX_EQUILIBRIUM ? HHG_EQ_WAVES : 2;
#endif
    // one wave per SIMD needs blocks of at most 4 waves
    static constexpr int max_threads = waves == 1 ? 256 : 512;
};

namespace solve_detail {

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Copies a wave-uniform record into scalar registers.
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

extern __shared__ __align__(16) unsigned char solve_dyn_lds[];

// Per-lane buffering of one sample sum and the 64-sample flush, into the
// job's accumulators in HBM or, in block-accumulation launches, into the
// block's LDS copy of them (lds >= 0: its first word's index in the LDS
// accumulator region at KernelArgs::acc_lds_off).
struct SampleSink {
    unsigned long long* acc;  // job accumulators, 2 words per sample
    int N;
    double buf;
    bool overflow;
    int lds;                  // -1: HBM; else the sink's first word in the LDS accumulators
    int lds_off;              // byte offset of the LDS accumulators in the dynamic LDS
    __device__ __forceinline__ void put(int i, double total, int lane) {
        const int slot = i & (WAVE - 1);
        if (lane == slot) buf = total;
        if (slot == WAVE - 1 || i == N) {
            const int base = i - slot;
            if (lane <= slot) {
                unsigned long long lo, hi;
                if (!fix_split(buf, lo, hi)) {
                    overflow = true;
                } else if (lds >= 0) {
                    unsigned long long* l = reinterpret_cast<unsigned long long*>(solve_dyn_lds + lds_off);
                    fix_add_lds(l + lds + 2 * (base + lane), lo, hi);
                } else {
                    fix_add(acc + 2 * (int64_t)(base + lane), lo, hi);
                }
            }
        }
    }
};

// Per-lane counts of one work item; flush_counters adds their 64-bit wave
// sums to the launch's counters and the item's job after every item, so no
// 32-bit lane or wave total can wrap however long the launch runs.
struct LaneCounters {
    uint32_t att = 0, rej = 0, fail = 0, nonfin = 0, fallback = 0, overflow = 0;
    unsigned long long wave_att = 0;  // lane 0: attempts the wave issued (count_waves launches)
};

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t o = __shfl_xor(v, off, WAVE);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ void flush_counters(Counters* ctr, unsigned long long* job_fail, int job, LaneCounters& lc) {
    const unsigned long long a = wave_sum_u64(lc.att), r = wave_sum_u64(lc.rej), f = wave_sum_u64(lc.fail),
                             n = wave_sum_u64(lc.nonfin), b = wave_sum_u64(lc.fallback),
                             o = wave_sum_u64(lc.overflow);
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        if (a) atomicAdd(&ctr->attempts, a);
        if (r) atomicAdd(&ctr->rejects, r);
        if (f) atomicAdd(&ctr->failed, f);
        if (n) atomicAdd(&ctr->nonfinite, n);
        if (b) atomicAdd(&ctr->fallbacks, b);
        if (o) atomicAdd(&ctr->overflows, o);
        if (lc.wave_att) atomicAdd(&ctr->wave_attempts, lc.wave_att);
        if (job_fail && f) atomicAdd(job_fail + 2 * job, f);
        if (job_fail && n) atomicAdd(job_fail + 2 * job + 1, n);
    }
    lc = LaneCounters{};
}

// Network features of one k-point at the start of interval i (times
// t_{i-1}..t_i): k, |d(k - A)|, gap at k, A, E = -dA/dt (central difference
// of the sample table), log(h_prev / Delta), log10 of the last error ratio
// clamped below at 5^-8 (the controller grows the step by its largest factor
// for every ratio below 5^-8, and the tiny-error path reports those as 0).
template <class F>
__device__ __forceinline__ void interval_features(const F& f, JobRef J, const double* At, int i, double gap_k,
                                                  double h_prev, double last_err, float feat[kMlpIn]) {
    // FP32 is enough for network inputs; only the differences stay FP64
    const float inv_delta = __fdividef(1.0f, (float)J.delta);
    const double A0 = At[i - 1];
    const float dA = i >= 2 ? (float)(At[i] - At[i - 2]) * (0.5f * inv_delta) : (float)(At[i] - At[i - 1]) * inv_delta;
    double d[3];
    f.d_of(A0, d);
    feat[0] = (float)f.kx;
    feat[1] = (float)f.ky;
    feat[2] = __fsqrt_rn((float)(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
    feat[3] = (float)gap_k;
    feat[4] = (float)A0;
    feat[5] = -dA;
    feat[6] = __logf((float)h_prev * inv_delta);
    feat[7] = __log10f((float)(last_err > (1.0 / 390625.0) ? last_err : (1.0 / 390625.0)));
}

// third coordinate of a k-point (3D pi-flux lattice; the other models have none)
template <int M>
__device__ __forceinline__ double load_kz(const double* kz, int64_t idx) {
    if constexpr (M == MODEL_PIFLUX3D) return kz[idx];
    else return 0.0;
}

// The launch arguments through an opaque copy of the kernel-argument segment
// address (the kernels that solve RKF78 items take KernelArgs as their only
// parameter, so the segment starts with it).  As with fresh_job, a copy made
// in every pass of a loop keeps the loop from holding the fields it reads in
// scalar registers from kernel entry on.
using ConstArgs = const __attribute__((address_space(4))) KernelArgs;
__device__ __forceinline__ ConstArgs& fresh_args() {
    uint64_t p = reinterpret_cast<uint64_t>(__builtin_amdgcn_kernarg_segment_ptr());
    asm volatile("" : "+s"(p));
    return *reinterpret_cast<ConstArgs*>(p);
}

template <int M, int P, int R, int LEARN>
__device__ void solve_item_rkf78(JobRef J, int chunk, const KernelArgs& a, LaneCounters& lc, bool to_lds) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t kk = (int64_t)chunk * WAVE + lane;
    const bool active = kk < J.n_k;
    double kx = 0.0, ky = 0.0, kz = 0.0, w = 0.0;
    if (active) {
        const int64_t idx = J.k_base + (a.perm ? a.perm[J.k_base + kk] : kk);
        kx = a.kx[idx];
        ky = a.ky[idx];
        kz = load_kz<M>(a.kz, idx);
        w = a.w[idx];
    }
    Bloch<M, P, R> f(J, kx, ky, kz);
    double s[3];
    f.initial(s);
    double gap_k;
    {
        double d[3];
        f.d_of(0.0, d);
        gap_k = 2.0 * sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    }
    const int N = J.n_meas;
    SampleSink sink{a.acc + 2 * J.sample_base, N, 0.0, false, to_lds ? 0 : -1, a.acc_lds_off};
    const double* At = a.At + J.sample_base;
    const double* Pt = a.Pt + 2 * J.sample_base;  // phase models: (sin f A_i, cos f A_i)
    // 3D pi-flux split current (PiFlux.cpp:998-1006): the samples of points
    // with |m(k - A)| >= split_cut go to the job's second output
    // honeycomb j_perp (Honeycomb.cpp:328-334): the second output
    SampleSink sink_hi{a.acc + 2 * (J.sample_base + N + 1), N, 0.0, false, to_lds ? 2 * (N + 1) : -1, a.acc_lds_off};
    const bool split = M == MODEL_PIFLUX3D && J.split_cut > 0.0;
    const bool perp = M == MODEL_HONEYCOMB && J.perp_current;
    const double w_perp = perp && active ? w * f.perp_factor() : 0.0;
    // what the current vertex reads of the field at sample i: A_i, or
    // (sin f A_i, cos f A_i) where the table holds phases (kPhase)
    struct Field { double x, y; };
    auto field = [&](int i) -> Field {
        if constexpr (Bloch<M, P, R>::kPhase) return {Pt[2 * i], Pt[2 * i + 1]};
        else return {At[i], 0.0};
    };
    auto emit = [&](int i, double g) {
        if constexpr (M == MODEL_PIFLUX3D) {
            if (split) {
                const bool lo = f.dispersion_phase(Pt[2 * i], Pt[2 * i + 1]) < J.split_cut;
                sink.put(i, wave_sum(lo ? g : 0.0), lane);
                sink_hi.put(i, wave_sum(lo ? 0.0 : g), lane);
                return;
            }
        }
        if constexpr (M == MODEL_HONEYCOMB) {
            if (perp) sink_hi.put(i, wave_sum(w_perp * f.vertex_perp_phase(Pt[2 * i], Pt[2 * i + 1], s)), lane);
        }
        sink.put(i, wave_sum(g), lane);
    };
    auto vertex = [&](const Field& fa) {
        if constexpr (Bloch<M, P, R>::kPhase) return f.vertex_phase(fa.x, fa.y, s);
        else return f.vertex(fa.x, s);
    };
    {
        const Field f0 = field(0);
        emit(0, active ? w * vertex(f0) : 0.0);
    }

    int status = active ? 0 : 3;
    const bool restart = a.sp.proposal == PROPOSAL_RESTART;
    const double h_restart = J.delta / J.n_sub;
    // the learned first-interval policy's proposal (first_step_kernel)
    double h_carry = (a.h_first && active) ? a.h_first[J.traj_base + kk] : h_restart;
    // an empty statement that reads h_carry: its load completes here, before
    // the interval loop, so no wait inside the loop has to count it
    asm volatile("" ::"v"(h_carry));
    const bool logging = (LEARN & LOG_STEPS) != 0 && a.log != nullptr;
    // step logs record exact error ratios; the network's feature clamps the
    // last ratio at 5^-8, which the tiny-error path's 0 satisfies.  Scheduled
    // learned solves hand an interval whose first step was rejected
    // fallback_after times back to the restart step Delta/n_sub.
    const IntervalCtl ctl{learn_mode(LEARN) == LEARN_SCHEDULED ? a.fallback_after : 0, h_restart,
                          logging ? TINY_EXACT : TINY_ZERO};
    IntervalOut io{};
    io.last_err = 1e-3;
    // learned: the network proposes an interval's first step in place of the
    // controller's carried step (a non-finite proposal restarts the interval at
    // Delta/n_sub); the interval itself runs the carried controller's code
    auto propose = [&](int i, bool want, float feat[kMlpIn]) {
        // every lane of the wave takes part in the inference
        interval_features(f, J, At, i, gap_k, h_carry, io.last_err, feat);
        float out[kMlpOut];
        mlp_infer_wave(a.mlp, solve_dyn_lds, feat, out);
        const float r = __expf(out[0]);
        if (want && status == 0) {
            if (r == r && r > 0.0f) h_carry = (double)fminf(fmaxf(r, 1e-4f), 4.5f) * J.delta;
            else h_carry = h_restart, lc.fallback++;
        }
    };
    int next_infer = 1;  // scheduled: interval of the next network proposal
    int kq = 0;  // next envelope corner
    // The table slots the next interval's steps read first (the restart
    // sequence's, or the whole-interval step's) are touched an interval ahead,
    // one load per 64-byte line and one of the last word (every line of the
    // range holds a touched word), so that the steps find them in the caches.
    // The loaded value is only read by an empty statement at the top of the
    // next interval, which keeps the load from being removed.  The sample's
    // field is loaded before the steps, so that it has arrived when they end.
    // Both hold registers across the steps: the equilibrium kernels, whose
    // registers are all in use, do neither.
    constexpr bool kEarlyLoads = R != RELAX_EQUILIBRIUM;
    constexpr int W = nom_stride(M, R);
    const int touch_first = restart ? 0 : kNomRestartSlots * W;
    const int touch_words = restart ? kNomRestartSlots * W : W;
    const int touch_lines = (touch_words + 7) / 8 + 1;
    uint32_t touched = 0;
    for (int i = 1; i <= N; ++i) {
        ConstArgs& A = fresh_args();  // the arguments the interval reads
        asm volatile("" ::"v"(touched));
        if (kEarlyLoads && A.nom && i < N && lane < touch_lines) {
            const double* nrow = A.nom + J.nominal_base + (int64_t)i * nom_row(M, R) + touch_first;
            const int wd = 8 * lane < touch_words ? 8 * lane : touch_words - 1;
            touched = reinterpret_cast<const uint32_t*>(nrow + wd)[0];
        }
        Field fi{};
        if constexpr (kEarlyLoads) fi = field(i);
        float feat[kMlpIn];
        const bool log_this = logging && status == 0 && ((i + 7 * lane) % a.log_stride) == 0;
        bool featured = false;
        if constexpr (learn_mode(LEARN) == LEARN_SCHEDULED) {
            if (i == next_infer) {
                next_infer += A.infer_every;
                propose(i, true, feat);
                featured = true;
            }
        }
        if (log_this && !featured) interval_features(f, J, At, i, gap_k, h_carry, io.last_err, feat);
        const uint32_t att0 = lc.att;
        bool corner = false;
        if (status == 0) {
            const double t0 = J.t_begin + (double)(i - 1) * J.delta;
            const double t1 = J.t_begin + (double)i * J.delta;
            double h = restart ? h_restart : h_carry;
            double unclipped = h;
            // an envelope corner inside the interval (listed only for carried and
            // learned proposals): integrate up to it, continue from there with the
            // controller's carried step; the table's whole-interval step does not apply
            corner = kq < J.n_kinks && J.kink_i[kq] == i;
            const double* nom =
                (A.nom && !corner) ? A.nom + J.nominal_base + (int64_t)(i - 1) * nom_row(M, R) : nullptr;
            const SolverDev sp{A.sp.rtol, A.sp.atol, A.sp.max_fail, A.sp.proposal, A.sp.max_steps, 0};
            IntervalCtl c = ctl;
            // Whole-interval steps: when a carried or proposed step covers the
            // interval, the loop clips it to the interval and its first attempt
            // is the table's whole-interval slot.  When that holds for every
            // active lane, that attempt runs outside the controller's loop, with
            // one program path for the wave; it stands when every lane accepts
            // it and it ends the interval, and it then leaves exactly the state
            // the loop would.  Otherwise the interval runs through the loop from
            // its start.  (A second copy of the attempt: the equilibrium kernels,
            // whose registers are all in use, keep the loop alone.)
            constexpr bool kWholeIntervalAttempt = R != RELAX_EQUILIBRIUM;
            bool done = false;
            const double hw = t1 - t0;  // the loop's clip of a step that covers the interval
            const bool whole = nom != nullptr && !restart && sp.max_steps >= 1 && t1 - t0 > kDblEps &&
                               (t0 + h) - t1 > kDblEps;
            if (kWholeIntervalAttempt && !A.loop_only && __all(whole)) {
                const double* cv = nom + kNomRestartSlots * W;
                double row[W];
#pragma unroll
                for (int j = 0; j < W; ++j) row[j] = cv[j];
                const bool hit = row[0] == t0 && row[1] == hw;
                StageVals<Bloch<M, P, R>::kPhase> sv;
                stage_values_from_row<Bloch<M, P, R>>(row, sv);
                double snew[3];
                const double err = rkf78_attempt(f, s, hw, sv, snew, sp.rtol, sp.atol, c.tiny_mode);
                const double t_next = t0 + hw;
                const bool ok = hit && !(err > 1.0) && t_next > t0 && !(t1 - t_next > kDblEps);
                if (__all(ok)) {
                    s[0] = snew[0];
                    s[1] = snew[1];
                    s[2] = snew[2];
                    ++lc.att;
                    io.attempts = 1;
                    io.fell_back = false;
                    io.last_err = err;
                    io.h_acc = hw;
                    io.err_acc = err;
                    h = hw;
                    grow_step(h, err);
                    if (!(isfinite(s[0]) && isfinite(s[1]) && isfinite(s[2]))) status = 2;
                    done = true;
                }
            }
            double ta = t0;
            for (int q = kq; !done; ++q) {
                const bool stop = q < J.n_kinks && J.kink_i[q] == i;
                const double tb = stop ? J.kink_t[q] : t1;
                if (status == 0) {
                    rkf78_interval(f, s, ta, tb, h, unclipped, sp, lc.att, lc.rej, status, nom, c, io);
                    if (io.fell_back) lc.fallback++;
                }
                if (!stop) break;
                h = h > unclipped ? h : unclipped;
                ta = tb;
                c.fallback_after = 0;  // the fallback concerns the interval's first step only
            }
            h_carry = h > unclipped ? h : unclipped;
        }
        if (log_this && status == 0 && !corner) {
            const unsigned long long slot = atomicAdd(a.log_count, 1ull);
            if (slot < (unsigned long long)a.log_cap) {
                float* rec = a.log + slot * (kMlpIn + kMlpOut);
#pragma unroll
                for (int q = 0; q < kMlpIn; ++q) rec[q] = feat[q];
                // target: the step the error model of the first accepted step
                // allows, h 0.9 err^(-1/8) without the growth cap (near-independent
                // of the step that was tried: err scales as h^8)
                const double e = io.err_acc > 1e-16 ? io.err_acc : 1e-16;
                rec[kMlpIn] = (float)(log(io.h_acc / J.delta) + log(0.9) - 0.125 * log(e));
                rec[kMlpIn + 1] = (float)log((double)io.attempts);
            }
        }
        // the wave issues as many attempts in this interval as its busiest lane
        if (A.count_waves) lc.wave_att += wave_max_u32(lc.att - att0);
        while (kq < J.n_kinks && J.kink_i[kq] <= i) ++kq;
        if constexpr (!kEarlyLoads) fi = field(i);
        emit(i, (status != 3) ? w * vertex(fi) : 0.0);
    }
    if (status == 1) lc.fail++;
    if (status == 2) lc.nonfin++;
    if (sink.overflow || sink_hi.overflow) lc.overflow++;
}

template <int M>
__device__ void solve_item_magnus(JobRef J, int chunk, const KernelArgs& a, LaneCounters& lc, bool to_lds) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t kk = (int64_t)chunk * WAVE + lane;
    const bool active = kk < J.n_k;
    double kx = 0.0, ky = 0.0, kz = 0.0, w = 0.0;
    if (active) {
        const int64_t idx = J.k_base + (a.perm ? a.perm[J.k_base + kk] : kk);
        kx = a.kx[idx];
        ky = a.ky[idx];
        kz = load_kz<M>(a.kz, idx);
        w = a.w[idx];
    }
    Bloch<M, PULSE_COS2, RELAX_NONE> f(J, kx, ky, kz);  // the pulse enters only through the Am table
    double s[3];
    f.initial(s);
    const int N = J.n_meas;
    const int ms = J.magnus_sub;
    const double h = J.delta / ms;
    SampleSink sink{a.acc + 2 * J.sample_base, N, 0.0, false, to_lds ? 0 : -1, a.acc_lds_off};
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
    if (a.count_waves) lc.wave_att += (unsigned long long)N * (unsigned long long)ms;
    if (active && !(isfinite(s[0]) && isfinite(s[1]) && isfinite(s[2]))) lc.nonfin++;
    if (sink.overflow) lc.overflow++;
}

// The reference's moment-based Magnus step of the 3D Dirac cone, one per
// sampling interval; Am holds the four pulse moments of every interval.
template <int M>
__device__ void solve_item_magnus_moments(JobRef J, int chunk, const KernelArgs& a, LaneCounters& lc, bool to_lds) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t kk = (int64_t)chunk * WAVE + lane;
    const bool active = kk < J.n_k;
    double kx = 1.0, ky = 1.0, w = 0.0;  // idle lanes evolve a harmless point with weight 0
    if (active) {
        const int64_t idx = J.k_base + (a.perm ? a.perm[J.k_base + kk] : kk);
        kx = a.kx[idx];
        ky = a.ky[idx];
        w = a.w[idx];
    }
    Bloch<M, PULSE_CONTINUOUS, RELAX_NONE> f(J, kx, ky);  // the pulse enters only through the moments
    double s[3];
    f.initial(s);
    const MomentLane L = moment_lane(kx, ky, J.delta);
    const int N = J.n_meas;
    SampleSink sink{a.acc + 2 * J.sample_base, N, 0.0, false, to_lds ? 0 : -1, a.acc_lds_off};
    const double* Am = a.Am + J.magnus_base;
    sink.put(0, wave_sum(active ? w * s[2] : 0.0), lane);
    for (int i = 1; i <= N; ++i) {
        const double* m4 = Am + 4 * (int64_t)(i - 1);
        const double mom[4] = {m4[0], m4[1], m4[2], m4[3]};
        magnus_moments_step(L, mom, s);
        sink.put(i, wave_sum(active ? w * s[2] : 0.0), lane);
    }
    if (active) lc.att += (uint32_t)N;
    if (a.count_waves) lc.wave_att += (unsigned long long)N;
    if (active && !(isfinite(s[0]) && isfinite(s[1]) && isfinite(s[2]))) lc.nonfin++;
    if (sink.overflow) lc.overflow++;
}

}  // namespace solve_detail

namespace solve_detail {

// The persistent work loop shared by the solve kernels.  Work is popped from
// the launch's queue counter.  Wave accumulation: each wave pops queue_chunk
// items and adds its sample sums to HBM.  Block accumulation: the block pops
// waves x queue_chunk items; its waves add their sample sums to the block's
// LDS copy of the current job's accumulators, which goes to HBM (one 128-bit
// addition per sample) when the block moves to another job and at the end;
// items of another job than the round's first (at a job boundary) add to HBM
// directly.  item(J, chunk, lc, to_lds) solves one work item.
template <int LEARN, class Item>
__device__ __forceinline__ void solve_loop(const KernelArgs& a, Item&& item_fn) {
    const int lane = threadIdx.x & (WAVE - 1);
    if constexpr (learn_mode(LEARN) != LEARN_NONE) {
        mlp_stage_lds(a.mlp, solve_dyn_lds);
        __syncthreads();
    }
    __shared__ int s_base;
    unsigned long long* lacc = reinterpret_cast<unsigned long long*>(solve_dyn_lds + a.acc_lds_off);
    const bool block = a.block_acc != 0;
    const int waves = blockDim.x / WAVE, wave = threadIdx.x / WAVE;
    const int per_pop = block ? waves * a.queue_chunk : a.queue_chunk;
    int cur_job = -1;
    auto flush_lds = [&](int job) {  // every thread of the block: LDS words read, zeroed, added to HBM
        JobRef J = job_ref(a.jobs + job);
        const int n_s = J.n_outputs * (J.n_meas + 1);
        for (int q = threadIdx.x; q < n_s; q += blockDim.x) {
            const unsigned long long lo = lacc[2 * q], hi = lacc[2 * q + 1];
            lacc[2 * q] = 0ull;
            lacc[2 * q + 1] = 0ull;
            if (lo | hi) fix_add(a.acc + 2 * (J.sample_base + q), lo, hi);
        }
    };
    if (block)
        for (int q = threadIdx.x; q < 2 * a.acc_lds_samples; q += blockDim.x) lacc[q] = 0ull;
    for (;;) {
        int base;
        if (block) {
            __syncthreads();  // the previous round's waves are done with s_base and the LDS sums
            if (threadIdx.x == 0) s_base = (int)atomicAdd(&a.ctr->queue, (unsigned long long)per_pop);
            __syncthreads();
            base = s_base;
        } else {
            base = 0;
            if (lane == 0) base = (int)atomicAdd(&a.ctr->queue, (unsigned long long)per_pop);
            base = uniform(base);
        }
        if (base >= a.n_items) break;
        if (block) {
            const int lead = a.items[base].x;
            if (lead != cur_job) {
                if (cur_job >= 0) flush_lds(cur_job);
                cur_job = lead;
                __syncthreads();  // zeroed before any wave adds
            }
        }
        for (int c = 0; c < a.queue_chunk; ++c) {
            const int it = block ? base + c * waves + wave : base + c;
            if (it >= a.n_items) break;
            const int2 item = load_uniform(a.items + it);
            LaneCounters lc;
            item_fn(job_ref(a.jobs + item.x), item.y, lc, block && item.x == cur_job);
            // the item's counts, its failures attributed to its job
            flush_counters(a.ctr, a.job_fail, item.x, lc);
        }
    }
    if (block && cur_job >= 0) flush_lds(cur_job);
}

}  // namespace solve_detail

// One kernel per (method, model, pulse, relaxation, learned); a launch only
// carries items of its own group.  Every wave pops items from the group's
// queue counter until the list is exhausted: a grid of one wave per item
// makes each wave solve one item, a grid of the resident wave count is the
// persistent schedule.
template <int METHOD, int M, int P, int R, int LEARN>
__global__ __attribute__((amdgpu_flat_work_group_size(1, SolveOccupancy<METHOD, M, R, LEARN>::max_threads),
                          amdgpu_waves_per_eu(SolveOccupancy<METHOD, M, R, LEARN>::waves,
                                              SolveOccupancy<METHOD, M, R, LEARN>::waves))) void
solve_kernel(KernelArgs a) {
    using namespace solve_detail;
    solve_loop<LEARN>(a, [&](JobRef J, int chunk, LaneCounters& lc, bool to_lds) {
        if constexpr (METHOD == METHOD_RKF78)
            solve_item_rkf78<M, P, R, LEARN>(J, chunk, a, lc, to_lds);
        else if constexpr (METHOD == METHOD_MAGNUS4)
            solve_item_magnus<M>(J, chunk, a, lc, to_lds);
        else
            solve_item_magnus_moments<M>(J, chunk, a, lc, to_lds);
    });
}

// A (model, pulse, relaxation) group of RKF78 jobs.
template <int M_, int P_, int R_>
struct GroupKey {
    static constexpr int M = M_, P = P_, R = R_;
};

// One RKF78 launch over the items of two groups (mixed sweeps): the waves
// share one work queue, so the launch has one tail instead of one per group.
// The item's job record (constant address space, a scalar load) selects the
// group's solver, a branch that is uniform over the wave.
template <class G1, class G2, int LEARN>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void solve_kernel_fused(KernelArgs a) {
    using namespace solve_detail;
    solve_loop<LEARN>(a, [&](JobRef J, int chunk, LaneCounters& lc, bool to_lds) {
        if (J.model == G1::M && J.pulse == G1::P && J.relax == G1::R)
            solve_item_rkf78<G1::M, G1::P, G1::R, LEARN>(J, chunk, a, lc, to_lds);
        else
            solve_item_rkf78<G2::M, G2::P, G2::R, LEARN>(J, chunk, a, lc, to_lds);
    });
}

// Per-k trajectories of one job (tests / inspection): one lane per selected
// k-point, Bloch vector at every sample.
template <int METHOD, int M, int P, int R>
__global__ __launch_bounds__(64) void trajectory_kernel(const JobDev* jobs, const double* kx, const double* ky,
                                                        const double* kz, const int64_t* sel, int64_t n_sel,
                                                        const double* Am, SolverDev sp, double* s_out,
                                                        Counters* ctr) {
    const int64_t t_idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t_idx >= n_sel) return;
    JobRef J = job_ref(jobs);
    const int64_t idx = J.k_base + sel[t_idx];
    Bloch<M, P, R> f(J, kx[idx], ky[idx], solve_detail::load_kz<M>(kz, idx));
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
    IntervalCtl ctl{0, h_restart, TINY_ZERO};  // nothing here reads the error ratio
    IntervalOut io{};
    int kq = 0;
    for (int i = 1; i <= N; ++i) {
        if constexpr (METHOD == 0) {
            if (status == 0) {
                const double t0 = J.t_begin + (double)(i - 1) * J.delta;
                const double t1 = J.t_begin + (double)i * J.delta;
                double h = (sp.proposal == PROPOSAL_RESTART) ? h_restart : h_carry;
                double unclipped = h;
                double ta = t0;
                for (int q = kq;; ++q) {  // envelope corners, as in the solve kernel
                    const bool stop = q < J.n_kinks && J.kink_i[q] == i;
                    const double tb = stop ? J.kink_t[q] : t1;
                    if (status == 0) rkf78_interval(f, s, ta, tb, h, unclipped, sp, att, rej, status, nullptr, ctl, io);
                    if (!stop) break;
                    h = h > unclipped ? h : unclipped;
                    ta = tb;
                }
                h_carry = h > unclipped ? h : unclipped;
            }
            while (kq < J.n_kinks && J.kink_i[kq] <= i) ++kq;
        } else if constexpr (METHOD == 1) {
            const int ms = J.magnus_sub;
            const double h = J.delta / ms;
            for (int q = 0; q < ms; ++q) {
                const int64_t m = J.magnus_base + 2 * ((int64_t)(i - 1) * ms + q);
                magnus4_step(f, s, Am[m], Am[m + 1], h);
            }
            att += ms;
        } else {
            const double* m4 = Am + J.magnus_base + 4 * (int64_t)(i - 1);
            const double mom[4] = {m4[0], m4[1], m4[2], m4[3]};
            magnus_moments_step(moment_lane(kx[idx], ky[idx], J.delta), mom, s);
            att += 1;
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

// Band occupations of the 3D pi-flux lattice with equilibrium relaxation
// (PiFlux::evolve_occupation_numbers, PiFlux.cpp:222-318): one lane per
// selected k-point, RKF78 with the persistent shifted momentum, (lower, upper)
// at every sample.
template <int P>
__global__ __launch_bounds__(64) void occupation_kernel(const JobDev* jobs, const double* kx, const double* ky,
                                                        const double* kz, const int64_t* sel, int64_t n_sel,
                                                        SolverDev sp, double* occ_out, Counters* ctr) {
    const int64_t t_idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t_idx >= n_sel) return;
    JobRef J = job_ref(jobs);
    const int64_t idx = J.k_base + sel[t_idx];
    Bloch<MODEL_PIFLUX3D, P, RELAX_EQUILIBRIUM, true> f(J, kx[idx], ky[idx], kz[idx]);
    double s[3], o[2];
    f.initial(s);
    f.begin_run(pulse_A<P>(J, J.t_begin));
    const int N = J.n_meas;
    double* out = occ_out + t_idx * (int64_t)(N + 1) * 2;
    f.occupations(s, o);
    out[0] = o[0];
    out[1] = o[1];
    uint32_t att = 0, rej = 0;
    int status = 0;
    const double h_restart = J.delta / J.n_sub;
    double h_carry = h_restart;
    IntervalCtl ctl{0, h_restart, TINY_ZERO};  // nothing here reads the error ratio
    IntervalOut io{};
    int kq = 0;
    for (int i = 1; i <= N; ++i) {
        if (status == 0) {
            const double t0 = J.t_begin + (double)(i - 1) * J.delta;
            const double t1 = J.t_begin + (double)i * J.delta;
            double h = (sp.proposal == PROPOSAL_RESTART) ? h_restart : h_carry;
            double unclipped = h;
            double ta = t0;
            for (int q = kq;; ++q) {  // envelope corners, as in the solve kernel
                const bool stop = q < J.n_kinks && J.kink_i[q] == i;
                const double tb = stop ? J.kink_t[q] : t1;
                if (status == 0) rkf78_interval(f, s, ta, tb, h, unclipped, sp, att, rej, status, nullptr, ctl, io);
                if (!stop) break;
                h = h > unclipped ? h : unclipped;
                ta = tb;
            }
            h_carry = h > unclipped ? h : unclipped;
        }
        while (kq < J.n_kinks && J.kink_i[kq] <= i) ++kq;
        f.occupations(s, o);
        out[2 * i + 0] = o[0];
        out[2 * i + 1] = o[1];
    }
    atomicAdd(&ctr->attempts, (unsigned long long)att);
    atomicAdd(&ctr->rejects, (unsigned long long)rej);
    if (status == 1) atomicAdd(&ctr->failed, 1ull);
    if (status == 2) atomicAdd(&ctr->nonfinite, 1ull);
}

// Predicted cost of every k-point of a group's items for lockstep grouping:
// the network's attempts-per-interval output summed over n_probe evenly
// spaced intervals (features with h_prev = Delta and a 1e-3 error ratio).
template <int M, int P>
__global__ __launch_bounds__(256) void cost_probe_kernel(KernelArgs a, int n_probe, float* cost) {
    using namespace solve_detail;
    const int lane = threadIdx.x & (WAVE - 1);
    mlp_stage_lds(a.mlp, solve_dyn_lds);
    __syncthreads();
    const int it = uniform((int)(blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE));
    if (it >= a.n_items) return;
    const int2 item = load_uniform(a.items + it);
    JobRef J = job_ref(a.jobs + item.x);
    const int64_t kk = (int64_t)item.y * WAVE + lane;
    const bool active = kk < J.n_k;
    const int64_t idx = J.k_base + (active ? kk : 0);
    Bloch<M, P, RELAX_NONE> f(J, a.kx[idx], a.ky[idx], load_kz<M>(a.kz, idx));
    double d[3];
    f.d_of(0.0, d);
    const double gap_k = 2.0 * sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double* At = a.At + J.sample_base;
    float total = 0.0f;
    for (int q = 0; q < n_probe; ++q) {
        const int i = 1 + (int)(((int64_t)q * J.n_meas) / n_probe);
        float feat[kMlpIn], out[kMlpOut];
        interval_features(f, J, At, i, gap_k, J.delta, 1e-3, feat);
        mlp_infer_wave(a.mlp, solve_dyn_lds, feat, out);
        total += __expf(out[1]);
    }
    if (active) cost[J.k_base + kk] = total;
}

// First-interval proposals of the learned solve (HHG_INFER_FIRST): for every
// trajectory of a group's items the network's first step at the start of the
// first interval, with the features the solve's own first interval has
// (carried step Delta/n_sub, last error ratio 1e-3); a non-finite proposal
// leaves Delta/n_sub and counts a fallback.  The solve then runs the carried
// controller's kernel from these steps.
template <int M, int P>
__global__ __launch_bounds__(256) void first_step_kernel(KernelArgs a, double* h_first) {
    using namespace solve_detail;
    const int lane = threadIdx.x & (WAVE - 1);
    mlp_stage_lds(a.mlp, solve_dyn_lds);
    __syncthreads();
    const int it = uniform((int)(blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE));
    if (it >= a.n_items) return;
    const int2 item = load_uniform(a.items + it);
    JobRef J = job_ref(a.jobs + item.x);
    const int64_t kk = (int64_t)item.y * WAVE + lane;
    const bool active = kk < J.n_k;
    const int64_t idx = J.k_base + (active ? (a.perm ? a.perm[J.k_base + kk] : kk) : 0);
    Bloch<M, P, RELAX_NONE> f(J, a.kx[idx], a.ky[idx], load_kz<M>(a.kz, idx));
    double d[3];
    f.d_of(0.0, d);
    const double gap_k = 2.0 * sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double h_restart = J.delta / J.n_sub;
    float feat[kMlpIn], out[kMlpOut];
    interval_features(f, J, a.At + J.sample_base, 1, gap_k, h_restart, 1e-3, feat);
    mlp_infer_wave(a.mlp, solve_dyn_lds, feat, out);
    const float r = __expf(out[0]);
    const bool ok = r == r && r > 0.0f;
    if (active) h_first[J.traj_base + kk] = ok ? (double)fminf(fmaxf(r, 1e-4f), 4.5f) * J.delta : h_restart;
    const unsigned long long nf = wave_sum_u64((active && !ok) ? 1ull : 0ull);
    if (lane == 0 && nf) atomicAdd(&a.ctr->fallbacks, nf);
}

// Run-time (method, pulse, relax, learned) -> template, for one model.  The
// 3D pi-flux lattice has no relaxation or its equilibrium relaxation; the
// honeycomb lattice has all four; the other models have all but the
// equilibrium relaxation.
template <int M>
struct ModelLaunch {
    template <int R>
    static constexpr bool has_relax = M == MODEL_PIFLUX3D    ? (R == RELAX_NONE || R == RELAX_EQUILIBRIUM)
                                      : M == MODEL_HONEYCOMB ? true
                                                             : (R != RELAX_EQUILIBRIUM);

    template <int P, int R, int L>
    static hipError_t solve_t(const KernelArgs& a, int blocks, int threads, size_t smem, hipStream_t st,
                              const void** fn) {
        if (fn) {
            *fn = reinterpret_cast<const void*>(&solve_kernel<0, M, P, R, L>);
            return hipSuccess;
        }
        hipLaunchKernelGGL((solve_kernel<0, M, P, R, L>), dim3(blocks), dim3(threads), smem, st, a);
        return hipGetLastError();
    }
    template <int P, int R>
    static hipError_t solve_l(int learn, const KernelArgs& a, int blocks, int threads, size_t smem, hipStream_t st,
                              const void** fn) {
        if constexpr (has_relax<R>) {
            switch (learn) {
            case LEARN_NONE: return solve_t<P, R, LEARN_NONE>(a, blocks, threads, smem, st, fn);
            case LEARN_SCHEDULED: return solve_t<P, R, LEARN_SCHEDULED>(a, blocks, threads, smem, st, fn);
            case LEARN_NONE | LOG_STEPS: return solve_t<P, R, LEARN_NONE | LOG_STEPS>(a, blocks, threads, smem, st, fn);
            case LEARN_SCHEDULED | LOG_STEPS:
                return solve_t<P, R, LEARN_SCHEDULED | LOG_STEPS>(a, blocks, threads, smem, st, fn);
            }
        }
        return hipErrorInvalidValue;
    }
    template <int P>
    static hipError_t solve_r(int relax, int learned, const KernelArgs& a, int blocks, int threads, size_t smem,
                              hipStream_t st, const void** fn) {
        switch (relax) {
        case RELAX_NONE: return solve_l<P, RELAX_NONE>(learned, a, blocks, threads, smem, st, fn);
        case RELAX_TO_INITIAL: return solve_l<P, RELAX_TO_INITIAL>(learned, a, blocks, threads, smem, st, fn);
        case RELAX_T2: return solve_l<P, RELAX_T2>(learned, a, blocks, threads, smem, st, fn);
        case RELAX_EQUILIBRIUM: return solve_l<P, RELAX_EQUILIBRIUM>(learned, a, blocks, threads, smem, st, fn);
        }
        return hipErrorInvalidValue;
    }
    static hipError_t solve(int method, int pulse, int relax, int learned, const KernelArgs& a, int blocks,
                            int threads, size_t smem, hipStream_t st, const void** fn) {
        if (method == METHOD_MAGNUS4) {
            if (fn) {
                *fn = reinterpret_cast<const void*>(&solve_kernel<1, M, PULSE_COS2, RELAX_NONE, LEARN_NONE>);
                return hipSuccess;
            }
            hipLaunchKernelGGL((solve_kernel<1, M, PULSE_COS2, RELAX_NONE, LEARN_NONE>), dim3(blocks), dim3(threads), smem,
                               st, a);
            return hipGetLastError();
        }
        if (method == METHOD_MAGNUS_MOMENTS) {
            if constexpr (M == MODEL_DIRAC3D) {
                if (fn) {
                    *fn = reinterpret_cast<const void*>(&solve_kernel<2, M, PULSE_CONTINUOUS, RELAX_NONE, LEARN_NONE>);
                    return hipSuccess;
                }
                hipLaunchKernelGGL((solve_kernel<2, M, PULSE_CONTINUOUS, RELAX_NONE, LEARN_NONE>), dim3(blocks),
                                   dim3(threads), smem, st, a);
                return hipGetLastError();
            }
            return hipErrorInvalidValue;
        }
        switch (pulse) {
        case PULSE_COS2: return solve_r<PULSE_COS2>(relax, learned, a, blocks, threads, smem, st, fn);
        case PULSE_GAUSS: return solve_r<PULSE_GAUSS>(relax, learned, a, blocks, threads, smem, st, fn);
        case PULSE_DOUBLE_GAUSS: return solve_r<PULSE_DOUBLE_GAUSS>(relax, learned, a, blocks, threads, smem, st, fn);
        case PULSE_DOUBLE_COS2: return solve_r<PULSE_DOUBLE_COS2>(relax, learned, a, blocks, threads, smem, st, fn);
        case PULSE_CONTINUOUS: return solve_r<PULSE_CONTINUOUS>(relax, learned, a, blocks, threads, smem, st, fn);
        case PULSE_SPLINE: return solve_r<PULSE_SPLINE>(relax, learned, a, blocks, threads, smem, st, fn);
        }
        return hipErrorInvalidValue;
    }

    template <int P, int R>
    static hipError_t traj_t(int method, const TrajArgs& x) {
        if constexpr (has_relax<R>) return traj_i<P, R>(method, x);
        else return hipErrorInvalidValue;
    }
    template <int P, int R>
    static hipError_t traj_i(int method, const TrajArgs& x) {
        const int blocks = (int)((x.n_sel + 63) / 64);
        if (method == METHOD_MAGNUS_MOMENTS) {
            if constexpr (M == MODEL_DIRAC3D)
                hipLaunchKernelGGL((trajectory_kernel<2, M, PULSE_CONTINUOUS, RELAX_NONE>), dim3(blocks), dim3(64), 0,
                                   x.st, x.jobs, x.kx, x.ky, x.kz, x.sel, x.n_sel, x.Am, x.sp, x.s_out, x.ctr);
            else
                return hipErrorInvalidValue;
        } else if (method == METHOD_MAGNUS4)
            hipLaunchKernelGGL((trajectory_kernel<1, M, PULSE_COS2, RELAX_NONE>), dim3(blocks), dim3(64), 0, x.st,
                               x.jobs, x.kx, x.ky, x.kz, x.sel, x.n_sel, x.Am, x.sp, x.s_out, x.ctr);
        else
            hipLaunchKernelGGL((trajectory_kernel<0, M, P, R>), dim3(blocks), dim3(64), 0, x.st, x.jobs, x.kx, x.ky,
                               x.kz, x.sel, x.n_sel, x.Am, x.sp, x.s_out, x.ctr);
        return hipGetLastError();
    }
    template <int P>
    static hipError_t traj_r(int method, int relax, const TrajArgs& x) {
        switch (relax) {
        case RELAX_NONE: return traj_t<P, RELAX_NONE>(method, x);
        case RELAX_TO_INITIAL: return traj_t<P, RELAX_TO_INITIAL>(method, x);
        case RELAX_T2: return traj_t<P, RELAX_T2>(method, x);
        case RELAX_EQUILIBRIUM: return traj_t<P, RELAX_EQUILIBRIUM>(method, x);
        }
        return hipErrorInvalidValue;
    }
    static hipError_t traj(int method, int pulse, int relax, const TrajArgs& x) {
        switch (pulse) {
        case PULSE_COS2: return traj_r<PULSE_COS2>(method, relax, x);
        case PULSE_GAUSS: return traj_r<PULSE_GAUSS>(method, relax, x);
        case PULSE_DOUBLE_GAUSS: return traj_r<PULSE_DOUBLE_GAUSS>(method, relax, x);
        case PULSE_DOUBLE_COS2: return traj_r<PULSE_DOUBLE_COS2>(method, relax, x);
        case PULSE_CONTINUOUS: return traj_r<PULSE_CONTINUOUS>(method, relax, x);
        case PULSE_SPLINE: return traj_r<PULSE_SPLINE>(method, relax, x);
        }
        return hipErrorInvalidValue;
    }

    static hipError_t first_step(int pulse, const KernelArgs& a, double* h_first, size_t smem, hipStream_t st) {
        const int blocks = (a.n_items + 3) / 4;
        switch (pulse) {
        case PULSE_COS2:
            hipLaunchKernelGGL((first_step_kernel<M, PULSE_COS2>), dim3(blocks), dim3(256), smem, st, a, h_first);
            break;
        case PULSE_GAUSS:
            hipLaunchKernelGGL((first_step_kernel<M, PULSE_GAUSS>), dim3(blocks), dim3(256), smem, st, a, h_first);
            break;
        case PULSE_DOUBLE_GAUSS:
            hipLaunchKernelGGL((first_step_kernel<M, PULSE_DOUBLE_GAUSS>), dim3(blocks), dim3(256), smem, st, a,
                               h_first);
            break;
        case PULSE_DOUBLE_COS2:
            hipLaunchKernelGGL((first_step_kernel<M, PULSE_DOUBLE_COS2>), dim3(blocks), dim3(256), smem, st, a,
                               h_first);
            break;
        case PULSE_CONTINUOUS:
            hipLaunchKernelGGL((first_step_kernel<M, PULSE_CONTINUOUS>), dim3(blocks), dim3(256), smem, st, a,
                               h_first);
            break;
        case PULSE_SPLINE:
            hipLaunchKernelGGL((first_step_kernel<M, PULSE_SPLINE>), dim3(blocks), dim3(256), smem, st, a, h_first);
            break;
        default:
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }

    static hipError_t cost(int pulse, const KernelArgs& a, int n_probe, float* cost_out, size_t smem, hipStream_t st) {
        const int blocks = (a.n_items + 3) / 4;
        switch (pulse) {
        case PULSE_COS2:
            hipLaunchKernelGGL((cost_probe_kernel<M, PULSE_COS2>), dim3(blocks), dim3(256), smem, st, a, n_probe,
                               cost_out);
            break;
        case PULSE_GAUSS:
            hipLaunchKernelGGL((cost_probe_kernel<M, PULSE_GAUSS>), dim3(blocks), dim3(256), smem, st, a, n_probe,
                               cost_out);
            break;
        case PULSE_DOUBLE_GAUSS:
            hipLaunchKernelGGL((cost_probe_kernel<M, PULSE_DOUBLE_GAUSS>), dim3(blocks), dim3(256), smem, st, a,
                               n_probe, cost_out);
            break;
        case PULSE_DOUBLE_COS2:
            hipLaunchKernelGGL((cost_probe_kernel<M, PULSE_DOUBLE_COS2>), dim3(blocks), dim3(256), smem, st, a,
                               n_probe, cost_out);
            break;
        case PULSE_CONTINUOUS:
            hipLaunchKernelGGL((cost_probe_kernel<M, PULSE_CONTINUOUS>), dim3(blocks), dim3(256), smem, st, a,
                               n_probe, cost_out);
            break;
        case PULSE_SPLINE:
            hipLaunchKernelGGL((cost_probe_kernel<M, PULSE_SPLINE>), dim3(blocks), dim3(256), smem, st, a, n_probe,
                               cost_out);
            break;
        default:
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
};

}  // namespace hhg

namespace hhg {

template <int M>
hipError_t solve_model(int method, int pulse, int relax, int learned, const KernelArgs& a, int blocks, int threads,
                       size_t smem, hipStream_t st, const void** fn) {
    return ModelLaunch<M>::solve(method, pulse, relax, learned, a, blocks, threads, smem, st, fn);
}

template <int M>
hipError_t traj_model(int method, int pulse, int relax, const TrajArgs& x) {
    return ModelLaunch<M>::traj(method, pulse, relax, x);
}

template <int M>
hipError_t cost_model(int pulse, const KernelArgs& a, int n_probe, float* cost, size_t smem, hipStream_t st) {
    return ModelLaunch<M>::cost(pulse, a, n_probe, cost, smem, st);
}

template <int M>
hipError_t first_step_model(int pulse, const KernelArgs& a, double* h_first, size_t smem, hipStream_t st) {
    return ModelLaunch<M>::first_step(pulse, a, h_first, smem, st);
}

}  // namespace hhg

#define HHG_INSTANTIATE_MODEL(M)                                                                                 \
    namespace hhg {                                                                                              \
    template hipError_t solve_model<M>(int, int, int, int, const KernelArgs&, int, int, size_t, hipStream_t,     \
                                       const void**);                                                            \
    template hipError_t traj_model<M>(int, int, int, const TrajArgs&);                                           \
    template hipError_t cost_model<M>(int, const KernelArgs&, int, float*, size_t, hipStream_t);                 \
    template hipError_t first_step_model<M>(int, const KernelArgs&, double*, size_t, hipStream_t);              \
    }
