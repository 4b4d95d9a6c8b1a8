// hhg_internal.hpp — layouts shared by the HIP kernels and the plan code.
//
// Device memory of a plan (all FP64 unless stated):
//   kx, ky, w        : SoA k-points of every k-grid, concatenated
//   kz               : third coordinate (only when a job uses the 3D pi-flux lattice)
//   perm (int64)     : optional lane -> k-point map per job (cost grouping)
//   jobs (JobDev)    : one record per job
//   items (int2)     : work items {job, chunk}; one item = 64 consecutive
//                      k-points of one job = one wavefront
//   At               : A(t_i), i = 0..N, per job (vertex evaluation)
//   Am               : A at the two Gauss nodes of every Magnus substep, or the
//                      four pulse moments of every interval (moment Magnus)
//   partial          : one row of N+1 wave sums per item, per job contiguous
//   seg_sum/seg_comp : level-1 compensated sums over item segments
//   j                : the result, sum over jobs of N+1 samples
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hhg {

enum : int {
    MODEL_DIRAC2D = 0,
    MODEL_PIFLUX2D = 1,
    MODEL_HONEYCOMB = 2,
    MODEL_DIRAC3D = 3,
    MODEL_PIFLUX3D = 4,
    N_MODELS = 5
};
enum : int {
    PULSE_COS2 = 0,
    PULSE_GAUSS = 1,
    PULSE_DOUBLE_GAUSS = 2,
    PULSE_DOUBLE_COS2 = 3,
    PULSE_CONTINUOUS = 4,
    PULSE_SPLINE = 5,
    N_PULSES = 6
};
enum : int { METHOD_RKF78 = 0, METHOD_MAGNUS4 = 1, METHOD_MAGNUS_MOMENTS = 2 };
enum : int { RELAX_NONE = 0, RELAX_TO_INITIAL = 1, RELAX_T2 = 2, RELAX_EQUILIBRIUM = 3 };
enum : int { PROPOSAL_RESTART = 0, PROPOSAL_CARRY = 1, PROPOSAL_LEARNED = 2 };

constexpr int WAVE = 64;

// Nominal-step pulse table: per sampling interval, kNomSlots candidate steps
// (t, h, A(t + c_j h) for the 13 Fehlberg stages).  Slots 0..5 replay the
// reference controller from its restart step Delta/n_sub growing at the
// maximal factor until the clipped last step; slot 6 is the single step
// covering the whole interval.
constexpr int kNomSlots = 7;
constexpr int kNomRestartSlots = 6;
constexpr int kNomStride = 15;  // t, h, 13 stage values
constexpr int kNomRow = kNomSlots * kNomStride;

// Lattice models read the field only through the phase f A of the shifted
// momentum (honeycomb: sin and cos of (kx - A)/2; pi-flux: cos(kx - A) or
// cos(kz - A)), and sin(f (k - A)) = sin(f k) cos(f A) - cos(f k) sin(f A).
// Their nominal-step table stores (sin f A, cos f A) at the 10 distinct
// Fehlberg stage times instead of A (c_7 = c_3, c_11 = c_0, c_12 = c_10), so a
// hit replaces the per-lane sincos of every stage by an angle subtraction; the
// sample table holds (sin f A_i, cos f A_i) of every phase model.  (The
// equilibrium-relaxation kernels, at their register limit, spill a little more
// with the phases, and still run 17-25 % faster for the sincos they drop.)
__host__ __device__ constexpr bool model_has_phase(int m) {
    return m == MODEL_PIFLUX2D || m == MODEL_HONEYCOMB || m == MODEL_PIFLUX3D;
}
__host__ __device__ constexpr bool phase_table(int m, int r) { return (void)r, model_has_phase(m); }
__host__ __device__ constexpr double model_phase_scale(int m) { return m == MODEL_HONEYCOMB ? 0.5 : 1.0; }
constexpr int kPhaseStages = 10;
// Fehlberg stage -> distinct stage time
__host__ __device__ constexpr int phase_stage(int i) {
    return i == 7 ? 3 : i == 11 ? 0 : i == 12 ? 9 : (i >= 8 ? i - 1 : i);
}
// the distinct stage times' Fehlberg stage
__host__ __device__ constexpr int phase_stage_of(int d) { return d >= 7 ? d + 1 : d; }
__host__ __device__ constexpr int nom_stride(int m, int r) { return phase_table(m, r) ? 2 + 2 * kPhaseStages : kNomStride; }
__host__ __device__ constexpr int nom_row(int m, int r) { return kNomSlots * nom_stride(m, r); }

struct JobDev {
    // model
    int32_t model, relax;
    double v, mass, hop;
    double ct, st;        // cos/sin of the polarisation angle
    double fermi, beta;   // beta <= 0: T = 0
    double relax_rate;    // 1 / relax_time (equilibrium relaxation: the diagonal rate)
    double relax_rate2;   // equilibrium relaxation: the off-diagonal rate
    double fermi_w;       // equilibrium relaxation: exp(beta c), c = -E_F (3D pi-flux) or E_F (honeycomb);
                          // 0 where the one-exponential Fermi pair does not apply (T = 0, |beta E_F| > 345)
    // pulse
    int32_t pulse, theta_zero;
    double amp_a, amp_b;
    double inv_n;         // 1 / n_cycles (cos^2)
    double t_env_end;     // t0 + 2 pi n_cycles (cos^2)
    double inv_2s2;       // 1 / (2 sigma^2) (Gaussian)
    double t0, delay, cep;
    // spline pulse: 4 coefficients per interval, intervals of length 1/spl_inv_h
    // from spl_t0; A = 0 outside [spl_t0, spl_t_last]
    const double* spl;
    double spl_t0, spl_inv_h, spl_t_last;
    int32_t spl_n, pad3;
    // time grid
    double t_begin, delta;
    int32_t n_meas, n_sub;
    int32_t magnus_sub;
    int32_t ic_rescue;    // 3D pi-flux rescue of the reference's drivers: 1 continuum grid, 2 uniform grid
    double split_cut;     // 3D pi-flux: > 0 splits j at |m(k - A)| = split_cut into two outputs
    // k-grid slice and output placement
    int64_t k_base, n_k;
    int64_t row_base;     // first partial row double of this job
    int64_t sample_base;  // offset of this job in At and j
    int64_t magnus_base;  // offset of this job in Am
    int64_t nominal_base; // offset of this job in the nominal-step table (N rows of kNomRow)
    int64_t item_base;    // first item of this job (items are job-major before reordering)
    int64_t traj_base;    // first trajectory of this job in per-trajectory arrays (first-step proposals)
    int32_t n_items;
    int32_t n_outputs;    // currents per sample: 2 with split_cut > 0 or perp_current, else 1
    // corners of the cos^2 envelopes inside sampling intervals (carried and
    // learned proposals stop there): times, interval index i (t_{i-1} < t < t_i)
    double kink_t[4];
    int32_t kink_i[4];
    int32_t n_kinks;
    int32_t perp_current; // honeycomb: the second output is the reference's j_perp (Honeycomb.cpp:328-334)
};

struct Counters {
    unsigned long long attempts;
    unsigned long long rejects;
    unsigned long long failed;
    unsigned long long nonfinite;
    unsigned long long fallbacks;
    unsigned long long queue;  // work-queue head
    unsigned long long log_count;
    unsigned long long wave_attempts;  // attempts issued per wave (the largest lane count of every interval)
    unsigned long long overflows;      // lanes whose sample sums left the accumulators' range
    unsigned long long pad[7];
};

struct SolverDev {
    double rtol, atol;
    int32_t max_fail;   // rejections of one step before the trajectory fails
    int32_t proposal;
    int32_t max_steps;  // accepted steps per sampling interval before the trajectory fails
    int32_t pad;
};

// Accepted steps per sampling interval before a trajectory is marked failed
// when the caller gives none: far above any converging problem (the restart
// controller takes 4 per interval), low enough that a step size pinned at the
// rounding floor of t ends the trajectory in a fraction of a second.
constexpr int kDefaultMaxSteps = 100000;

}  // namespace hhg
