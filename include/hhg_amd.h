/*
 * hhg_amd.h — C ABI of the MI355X HigherHarmonics Bloch-equation hot path.
 *
 * One library (libhhg_amd.so) exposes the per-k-point Bloch solve and the
 * quadrature-weighted current reduction j(t) as plain C: descriptors in,
 * caller-owned double buffers out, status codes instead of exceptions.
 *
 * Reference interfaces this ABI replaces (paths relative to the reference
 * repository majesticrock/HigherHarmonics):
 *   - Dispatcher::compute(rank, n_ranks, n_z) filling current_density_time
 *       (sources/HHG/Dispatch/Dispatcher.hpp:17,29) -> hhg_plan_execute
 *   - DiracSystem / PiFlux / Honeycomb ::compute_current_density(laser,
 *       time_config, rank, n_ranks, ...) (sources/HHG/Systems/DiracSystem.hpp:58-59,
 *       PiFlux.hpp:92-93, Honeycomb.hpp:70-71) -> hhg_plan_create + hhg_plan_execute
 *   - per-k time_evolution_sigma / time_evolution_magnus /
 *       time_evolution_diagonal_relaxation (DiracSystem.hpp:40-47,
 *       Honeycomb.hpp:64-68, PiFlux.hpp:80-81) -> hhg_trajectories
 *   - PiFlux::current_per_energy(laser, time_config, N) (PiFlux.hpp:98-99,
 *       PiFlux.cpp:954-1035) -> a job with hhg_model_desc.split_dispersion on
 *       HHG_KGRID_PIFLUX_UNIFORM
 *   - PiFlux::compute_occupation_numbers(laser, time_config, N) (PiFlux.hpp:95-96,
 *       PiFlux.cpp:909-952) -> hhg_occupations over HHG_KGRID_PIFLUX_OCCUPATION
 *   - Laser::laser_function(t) (sources/HHG/Laser/Laser.hpp:32) -> hhg_pulse_evaluate
 *   - TimeIntegrationConfig {t_begin, t_end, n_measurements, n_subdivisions}
 *       (sources/HHG/TimeIntegrationConfig.hpp:7-15) -> hhg_time_desc
 *   - gauss::container<N> tables (sources/HHG/thread_gauss.hpp:21-311) -> hhg_gauss_legendre
 *   - TrapezoidalFFT::compute (sources/HHG/Fourier/TrapezoidalFFT.cpp:11-49) with
 *       the x8 zero padding of sources/hhg.cpp:277-282 -> hhg_trapezoidal_spectrum
 *
 * Units are the reference's dimensionless ones: time in 1/omega_L, energy in
 * hbar omega_L, A(t) as a momentum shift k -> k - A(t) e_theta.
 * Everything is FP64.
 *
 * Threading: a plan is not thread-safe; distinct plans may be used
 * concurrently from different threads.  All device memory of a plan is owned
 * by the plan and released by hhg_plan_destroy.
 */
#ifndef HHG_AMD_H
#define HHG_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HHG_AMD_ABI_VERSION 5

typedef enum hhg_status {
    HHG_SUCCESS = 0,
    HHG_ERROR_INVALID_ARGUMENT = 1, /* bad descriptor (reference: std::invalid_argument) */
    HHG_ERROR_HIP = 2,              /* a HIP runtime call failed */
    HHG_ERROR_TOLERANCE = 3,        /* >= 1 trajectory failed its step controller (reference: odeint
                                       step_adjustment_error); j is still written, see hhg_stats */
    HHG_ERROR_OUT_OF_MEMORY = 4,
    HHG_ERROR_UNSUPPORTED = 5,      /* combination not provided (reference: std::runtime_error) */
    HHG_ERROR_NO_DEVICE = 6,
    HHG_ERROR_NONFINITE = 7         /* a pulse or state value became NaN/Inf */
} hhg_status;

/* ---- models: d(k - A(t) e_theta); Bloch equation ds/dt = 2 d x s [+ relaxation] ---- */
typedef enum hhg_model_kind {
    HHG_MODEL_DIRAC2D = 0,   /* d = (v qx, v qy, mass) */
    HHG_MODEL_PIFLUX2D = 1,  /* d = 2t (cos qx, cos qy, 0); cf. PiFlux.cpp:89-94 */
    HHG_MODEL_HONEYCOMB = 2, /* d = t (Re g, -Im g, 0), g = sum_delta e^{i q.delta}; Honeycomb.cpp:44-50,130-138 */
    HHG_MODEL_DIRAC3D = 3,   /* the reference's 3D Dirac cone in its band frame, field along z, point (k_z, kappa)
                                in (kx, ky): d = -(A kappa/|k|, 0, |k| - A k_z/|k|), DiracSystem.cpp:138-174;
                                current vertex s_z; lengths in hbar omega_L / v_F */
    HHG_MODEL_PIFLUX3D = 4   /* the reference's 3D pi-flux lattice, field along z, point (x, y, z) in (kx, ky, kz):
                                d = 2t (cos x, cos y, cos(z - A)), PiFlux.cpp:73-111; initial state PiFlux.cpp:519-549
                                with the occupations of PiFlux.cpp:420-426; current vertex sin(z - A) s_z
                                (PiFlux.cpp:831-838) */
} hhg_model_kind;

typedef enum hhg_relax_kind {
    HHG_RELAX_NONE = 0,
    HHG_RELAX_TO_INITIAL = 1, /* -(s - s0)/tau, DiracSystem.cpp:231-233 */
    HHG_RELAX_T2 = 2,         /* -(s - (s.d^)d^)/T2 about the instantaneous d(k - A(t)) */
    HHG_RELAX_EQUILIBRIUM = 3 /* toward the instantaneous equilibrium of d(k - A(t)).  HHG_MODEL_PIFLUX3D: diagonal
                                 time relax_time, off-diagonal time relax_time_offdiag, coherence phase of the
                                 step-start state (PiFlux.cpp:113-220; restart proposal only).  HHG_MODEL_HONEYCOMB:
                                 -(s - s_eq(k - A))/relax_time, starting in s_eq(k - A(t_begin))
                                 (Honeycomb.cpp:160-199) */
} hhg_relax_kind;

typedef struct hhg_model_desc {
    int32_t kind;        /* hhg_model_kind */
    int32_t relax;       /* hhg_relax_kind */
    double v;            /* Dirac-2D velocity */
    double mass;         /* Dirac-2D gap parameter */
    double hopping;      /* PiFlux-2D / honeycomb hopping t */
    double theta;        /* polarisation angle (Honeycomb.hpp:88-90 fixes 0) */
    double fermi_energy; /* E_F in hbar omega_L */
    double beta;         /* 1/T in 1/(hbar omega_L); <= 0 means T = 0 */
    double relax_time;   /* tau or T2 in 1/omega_L */
    double relax_time_offdiag; /* HHG_RELAX_EQUILIBRIUM: off-diagonal relaxation time in 1/omega_L (<= 0: none) */
    double split_dispersion;   /* HHG_MODEL_PIFLUX3D with RKF78, > 0: the job's output is two currents of N + 1
                                  samples, first the points with |m(k - A(t_i))| below this value, then the rest
                                  (PiFlux::current_per_energy, PiFlux.cpp:954-1035, cut sqrt(3)/2) */
    int32_t perp_current;      /* HHG_MODEL_HONEYCOMB with RKF78, 1: the job's output is two currents of N + 1
                                  samples, j along the polarisation, then the reference's consistency check
                                  j_perp = j_x sin(theta) + j_y cos(theta) (Honeycomb.cpp:328-334; theta = 0,
                                  so j_y), which vanishes on the mirror-symmetric hexagonal grid */
    int32_t pad;
} hhg_model_desc;

/* ---- pulses: A(t) = amplitude * envelope * sin(t + cep) ---- */
typedef enum hhg_pulse_kind {
    HHG_PULSE_COS2 = 0,         /* CosineLaser.cpp:22-26, window (t0, t0 + 2 pi n) */
    HHG_PULSE_GAUSS = 1,        /* GaussLaser.cpp:26-28, centre t0, width sigma */
    HHG_PULSE_DOUBLE_GAUSS = 2, /* DoubleLaser<GaussLaser>: A(t) + B(t - delay), DoubleLaser.hpp:63-67 */
    HHG_PULSE_DOUBLE_COS2 = 3,  /* DoubleLaser<CosineLaser> */
    HHG_PULSE_CONTINUOUS = 4,   /* ContinuousLaser.cpp:7-9, envelope 1 */
    HHG_PULSE_SPLINE = 5        /* a sampled pulse through the cubic spline the reference builds with Boost's
                                   cardinal_cubic_b_spline (Laser.hpp:8-11, 32-37; DoubleLaser.hpp:48-73,
                                   ExperimentalLaser.cpp:229-257): the C2 interpolant of the samples with the given
                                   end slopes; A = 0 outside [sample_t0, t_cut] */
} hhg_pulse_kind;

typedef struct hhg_pulse_desc {
    int32_t kind; /* hhg_pulse_kind */
    int32_t pad;
    double amp_a;    /* peak of pulse A */
    double amp_b;    /* peak of pulse B (double pulses) */
    double n_cycles; /* cos^2 envelope length 2 pi n_cycles */
    double sigma;    /* Gaussian width */
    double t0;       /* cos^2 window begin / Gaussian centre of pulse A */
    double delay;    /* pulse B delay */
    double cep;      /* carrier-envelope phase */
    /* HHG_PULSE_SPLINE: samples[i] = A(sample_t0 + i sample_dt), i < n_samples (>= 2); the end slopes
     * (the reference passes 0 and 0); t_cut: A = 0 beyond it (the reference's Laser::t_end, Laser.hpp:35;
     * t_cut <= sample_t0 means the last sample time).  The samples are read at plan creation. */
    const double* samples;
    int64_t n_samples;
    double sample_t0, sample_dt;
    double slope_begin, slope_end;
    double t_cut;
} hhg_pulse_desc;

/* TimeIntegrationConfig: samples t_i = t_begin + i (t_end - t_begin)/N, i = 0..N;
 * the reference controller restarts each interval at dt = Delta/n_subdivisions. */
typedef struct hhg_time_desc {
    double t_begin;
    double t_end;
    int32_t n_measurements;
    int32_t n_subdivisions;
} hhg_time_desc;

/* ---- Brillouin-zone quadrature ---- */
typedef enum hhg_kgrid_kind {
    HHG_KGRID_GL_BOX = 0,  /* nx x ny Gauss-Legendre on [x0,x1] x [y0,y1], x-major */
    HHG_KGRID_GL_HEX = 1,  /* hexagonal BZ of Honeycomb.cpp:229-268 (nx x ny nodes) */
    HHG_KGRID_EXPLICIT = 2, /* caller's points and weights */
    HHG_KGRID_DIRAC_CYLINDER = 3, /* (k_z, kappa) of the 3D Dirac cone: k_z = (z - nx/2) dk, z = 1..nx-1 except nx/2,
                                     dk = 2 x1 / nx (DiracSystem.cpp:280-306); kappa: ny-point Gauss-Legendre on
                                     (0, sqrt(x1^2 - k_z^2)) where the reference refines an adaptive trapezoid;
                                     weight dk w_kappa k_z kappa / |k| (DiracSystem.cpp:279); x1 = band width */
    HHG_KGRID_PIFLUX_CONTINUUM = 4, /* (x, y, z) of the 3D pi-flux lattice, PiFlux::current_density_continuum_limit
                                      (PiFlux.cpp:794-863) with improved_xy_integral (PiFlux.cpp:665-792): z = +-pi x_z
                                      at the nx positive nodes of the 2 nx-point Gauss-Legendre rule, four Gauss
                                      patches in (x, y) on [0, pi] x [0, pi/2] with ny = N_coarse (PiFlux.hpp:12-29,
                                      the reference's nx = 32, ny = 4); weights as the reference's, without Jacobians
                                      of the z map.  Jobs on this grid start, like the reference's continuum
                                      driver, from cos z scaled by 1 - 5e-9 at the +z nodes with cos z < 0
                                      (PiFlux.cpp:825-829) */
    HHG_KGRID_PIFLUX_UNIFORM = 5,  /* (x pi/N, y pi/N, +-z pi/N), x, y, z < N = nx, weight 1/N^3 each: the uniform
                                      grid of PiFlux::current_per_energy (PiFlux.cpp:954-1035).  Where cos x and
                                      cos y vanish (is_zero) and cos z < 0 the reference's rescue replaces cos x by
                                      1e-4 |cos z| for the whole run and scales the initial cos z by 1 - 5e-9
                                      (PiFlux.cpp:978-986); jobs on this grid do the same */
    HHG_KGRID_PIFLUX_OCCUPATION = 6, /* (x pi/N, pi/2, z pi/N) for i = x N + z, x, z < N = nx, weight 1/N^2: the
                                      slice of PiFlux::compute_occupation_numbers (PiFlux.cpp:909-952), with the
                                      rescue of HHG_KGRID_PIFLUX_UNIFORM (PiFlux.cpp:935-938) */
    HHG_KGRID_DIRAC_GAUSS = 7 /* (k_z, kappa) of the 3D Dirac cone by Gauss-Legendre in both, the quadrature of
                                 sources/hhg_gauss.cpp:71-133: k_z = x1 x_z over the nx-point rule (both signs),
                                 kappa = kappa_max(k_z) (x_r + 1) / 2 over the ny-point rule,
                                 kappa_max = sqrt(x1^2 - k_z^2) (DiracSystem.cpp:257-267); weight
                                 x1 w_z kappa_max / 2 w_r k_z kappa / |k| (hhg_gauss.cpp:24-26, 77; the Jacobians
                                 x1 and kappa_max / 2 of its substitutions included, so that it integrates what
                                 HHG_KGRID_DIRAC_CYLINDER does); x1 = band width */
} hhg_kgrid_kind;

typedef struct hhg_kgrid_desc {
    int32_t kind; /* hhg_kgrid_kind */
    int32_t nx;
    int32_t ny;
    int32_t pad;
    double x0, x1, y0, y1;
    const double* kx; /* explicit grids only: n_points each */
    const double* ky;
    const double* w;
    int64_t n_points;
    const double* kz; /* explicit grids only, optional: third coordinate (3D pi-flux points) */
} hhg_kgrid_desc;

/* One job = one (model, pulse, time grid, k-grid) current j(t). */
typedef struct hhg_job_desc {
    hhg_model_desc model;
    hhg_pulse_desc pulse;
    hhg_time_desc time;
    int32_t kgrid; /* index into hhg_plan_desc.kgrids */
    int32_t pad;
} hhg_job_desc;

typedef enum hhg_method {
    HHG_METHOD_RKF78 = 0,
    HHG_METHOD_MAGNUS4 = 1,        /* two-node Gauss Magnus, magnus_substeps per sampling interval */
    HHG_METHOD_MAGNUS_MOMENTS = 2  /* the reference's Dirac Magnus step: one per sampling interval from four
                                      30-point moments of A (Laser.cpp:40-55, MagnusMatrix.cpp:12-76,
                                      DiracSystem.cpp:176-211); HHG_MODEL_DIRAC3D without relaxation only */
} hhg_method;

typedef enum hhg_proposal {
    HHG_PROPOSAL_RESTART = 0, /* reference: every interval starts at dt = Delta/n_sub */
    HHG_PROPOSAL_CARRY = 1,   /* carry the controller's step across intervals (ablation) */
    HHG_PROPOSAL_LEARNED = 2  /* MLP proposes each interval's first step; controller still judges */
} hhg_proposal;

/* When the network proposes an interval's first step in place of the
 * controller's carried step (HHG_PROPOSAL_LEARNED).  The other intervals start
 * from the carried step; the interval runs the carried controller either way. */
typedef enum hhg_infer_policy {
    HHG_INFER_FIRST = 0,    /* the first interval only: where the sampling interval is shorter than the
                               error-limited step the carried step is then one attempt per interval, and
                               the interval loop is exactly the carried controller's */
    HHG_INFER_SCHEDULED = 1 /* every trajectory, every mlp_infer_every-th interval (fallback_after applies) */
} hhg_infer_policy;

typedef struct hhg_mlp hhg_mlp;

typedef struct hhg_solver_desc {
    int32_t method;   /* hhg_method */
    int32_t proposal; /* hhg_proposal */
    double rtol;
    double atol;
    int32_t max_fail;        /* attempts per step before a trajectory fails (odeint: 500) */
    int32_t magnus_substeps; /* Magnus steps per sampling interval */
    const hhg_mlp* mlp;      /* HHG_PROPOSAL_LEARNED and cost grouping: the step-proposal network */
    int32_t mlp_infer_every; /* HHG_INFER_SCHEDULED: intervals between network evaluations (0: every interval) */
    int32_t fallback_after;  /* HHG_INFER_SCHEDULED: rejections of an interval's first step (proposed or
                                carried) after which the interval restarts at Delta/n_sub (0: 2); a
                                non-finite proposal restarts its interval there at once (both policies) */
    int32_t max_steps;       /* accepted steps per sampling interval before a trajectory fails (0: 100000); a
                                trajectory also fails when an accepted step leaves t unchanged.  odeint has
                                neither bound and would loop forever in both cases */
    int32_t infer_policy;    /* hhg_infer_policy, for HHG_PROPOSAL_LEARNED */
} hhg_solver_desc;

typedef struct hhg_launch_desc {
    int32_t waves_per_block; /* 0 = 4, or 8 with block accumulation; or 1, 2, 4, 8 */
    int32_t queue_chunk;     /* 0 = default/autotuned: work items grabbed per queue pop */
    int32_t persistent;      /* 1 = persistent grid + device work queue, 0 = one wave per item */
    int32_t group_by_cost;   /* > 0: order k-points inside a job by the network's predicted step count
                                (needs solver.mlp) so that lanes of a wave need similar work: 1 = 64 cost
                                classes (quantiles of the prediction), n >= 2 = n classes; inside a class
                                the k-grid order is kept, so neighbouring points (whose step counts rise
                                and fall together in time) stay in one wave.  A full sort by the prediction
                                (n = the job's point count) scatters them and costs more than it saves */
    int32_t count_wave_attempts; /* 1 = count hhg_stats.wave_attempts (a few instructions per interval) */
    int32_t block_reduce;    /* persistent launches: 0 or 1 = every wave adds its exact sample sums to the job's
                                accumulators in HBM; 2 = the waves of a block add them to an LDS copy of the
                                job's accumulators, which goes to HBM once per job and block (far less HBM
                                traffic, about 10 % more solve time on MI355X; HHG_ERROR_UNSUPPORTED where a
                                job's output does not fit the block's LDS).  j is bitwise the same either way */
    int32_t separate_launches; /* 0 = a plan whose RKF78 jobs form two (model, pulse, relaxation) groups with a
                                  fused kernel (honeycomb double cos^2 with PiFlux-2D double Gaussian, no
                                  relaxation) solves both in one launch over one work queue; 1 = one launch per
                                  group.  j is bitwise the same either way */
    int32_t loop_only;         /* RKF78: 0 = where every active lane of a wave takes the nominal-step table's
                                  whole-interval step first (carried and proposed steps that cover the
                                  interval), that attempt runs outside the controller's loop and stands when
                                  every lane accepts it; 1 = every interval through the loop.  j and the
                                  counters are bitwise the same either way (a check of the first form) */
} hhg_launch_desc;

typedef struct hhg_plan_desc {
    int32_t n_kgrids;
    int32_t n_jobs;
    const hhg_kgrid_desc* kgrids;
    const hhg_job_desc* jobs;
    hhg_solver_desc solver;
    hhg_launch_desc launch;
    int32_t rank;    /* shard of the k-points (reference: contiguous block per MPI rank) */
    int32_t n_ranks; /* 1 on a single GPU */
    int32_t device;  /* HIP device ordinal */
    int32_t pad;
} hhg_plan_desc;

typedef struct hhg_stats {
    int64_t trajectories;
    int64_t attempts;            /* attempted RKF78 steps (or Magnus steps) */
    int64_t rejects;
    int64_t failed_trajectories; /* controller gave up (HHG_ERROR_TOLERANCE) */
    int64_t nonfinite_trajectories;
    int64_t proposal_fallbacks;  /* learned proposals replaced by the plain controller */
    int64_t wave_attempts;       /* launch.count_wave_attempts: attempts issued per wavefront of 64
                                    trajectories, i.e. the sum over waves and intervals of the largest
                                    per-trajectory attempt count (0 when not counted) */
    int64_t accumulator_overflows; /* wavefront lanes whose sample sums left the exact accumulators'
                                      range (|sum| >= 2^46): j is wrong where this is nonzero */
    double solve_ms;             /* device time of the solve kernel(s); after a graph replay
                                    (hhg_plan_capture_graph) the time of the whole graph */
    double reduce_ms;            /* device time of the j reduction (0 after a graph replay) */
    double total_ms;             /* host wall time of hhg_plan_execute */
} hhg_stats;

typedef struct hhg_plan hhg_plan;

/* Version of this ABI (HHG_AMD_ABI_VERSION). */
int32_t hhg_abi_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* hhg_last_error_string(void);

/* Builds the k-grids, uploads them and the job tables to the device and
 * allocates all scratch.  Replaces constructing a Dispatcher
 * (sources/HHG/Dispatch/DiracDispatcher.cpp:25-62 and siblings). */
hhg_status hhg_plan_create(const hhg_plan_desc* desc, hhg_plan** plan);

/* Number of output doubles: sum over jobs of (n_measurements + 1). */
int64_t hhg_plan_output_size(const hhg_plan* plan);

/* offsets[n_jobs + 1]: job j's samples are j_out[offsets[j] .. offsets[j+1]). */
hhg_status hhg_plan_job_offsets(const hhg_plan* plan, int64_t* offsets);

/* Solves every trajectory and writes j(t) of every job into the caller-owned
 * host buffer j_out (hhg_plan_output_size doubles).  stats may be NULL.
 * Replaces Dispatcher::compute (sources/HHG/Dispatch/Dispatcher.hpp:29). */
hhg_status hhg_plan_execute(hhg_plan* plan, double* j_out, hhg_stats* stats);

/* As hhg_plan_execute but leaves j in device memory j_dev (hhg_plan_output_size
 * doubles, caller-allocated on the plan's device) and does not synchronise
 * the host beyond what the plan's stream requires; stream is a hipStream_t or
 * NULL for the plan's own stream.  It does not wait for the solve, so it
 * reports no controller failures: hhg_plan_last_stats does. */
hhg_status hhg_plan_execute_device(hhg_plan* plan, double* j_dev, void* stream);

/* Waits for the plan's last execution (on whichever stream it was enqueued)
 * and returns its statistics and status as hhg_plan_execute would have
 * (HHG_ERROR_TOLERANCE / HHG_ERROR_NONFINITE with j still written).
 * total_ms is 0 after hhg_plan_execute_device. */
hhg_status hhg_plan_last_stats(hhg_plan* plan, hhg_stats* stats);

/* Tries launch shapes on the plan's own work and keeps the fastest
 * (waves_per_block x queue_chunk); writes the choice into chosen (may be NULL). */
hhg_status hhg_plan_autotune(hhg_plan* plan, hhg_launch_desc* chosen);

/* Number of solve-kernel launches one execution of the plan makes (one per
 * (model, pulse, relaxation) group, or 1 for a fused pair); -1 for NULL.
 * No counterpart in the reference, whose dispatchers run one system per call. */
int32_t hhg_plan_solve_launches(const hhg_plan* plan);

/* Records the plan's kernels into a HIP graph; later executes replay it. */
hhg_status hhg_plan_capture_graph(hhg_plan* plan);

hhg_status hhg_plan_destroy(hhg_plan* plan);

/* Per-job counts of the last execution (waits for it on whichever stream it
 * was enqueued): failed[j] trajectories whose controller gave up
 * (HHG_ERROR_TOLERANCE) and nonfinite[j] whose state became non-finite, for
 * j < n_jobs (the per-pulse counts behind the plan-wide totals of hhg_stats;
 * the reference throws from odeint on the first such trajectory,
 * DiracSystem.cpp:167-173).  hhg_trajectories and hhg_occupations keep their
 * own counters and leave these alone; hhg_plan_autotune runs partial solves
 * and leaves zeros. */
hhg_status hhg_plan_job_failures(hhg_plan* plan, int64_t* failed, int64_t* nonfinite);

/* Per-k Bloch vectors at every sample of job 0 for the listed k indices of
 * its k-grid: s_out[n_sel][N+1][3].  Replaces time_evolution_sigma /
 * time_evolution_magnus for inspection and tests. */
hhg_status hhg_trajectories(hhg_plan* plan, const int64_t* k_index, int64_t n_sel, double* s_out);

/* hhg_trajectories without device allocation: the same kernel, inputs and
 * outputs (s_out[n_sel][N+1][3] in host memory, bitwise equal), with the
 * selection, the output staging and the counters in a static device
 * workspace of 16 MiB that every GPU the library runs on holds for the life of
 * the process.  No counterpart in the reference: for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc).
 * Calls needing more than 16 MiB (8 n_sel + 24 n_sel (N + 1) bytes plus 512)
 * fail with HHG_ERROR_INVALID_ARGUMENT; calls that overlap on one GPU share
 * the workspace and race. */
hhg_status hhg_trajectories_noalloc(hhg_plan* plan, const int64_t* k_index, int64_t n_sel, double* s_out);

/* Band occupations of job 0 for the listed k indices of its k-grid:
 * occ_out[n_sel][N+1][2] = (lower, upper) at every sample.  Replaces
 * PiFlux::evolve_occupation_numbers (PiFlux.cpp:222-318, called from
 * compute_occupation_numbers, PiFlux.cpp:909-952): job 0 is the 3D pi-flux
 * lattice with HHG_RELAX_EQUILIBRIUM and the plan's solver RKF78.  As in the
 * reference the shifted momentum of the equilibrium targets persists across
 * right-hand-side calls, so once its rescue (cos x, cos y zero, cos(z - A) < 0)
 * has set cos x = 1e-4 |cos(z - A)| that value stays for the rest of the run;
 * the occupations are sqrt((|d| -+ d_z)/2) of the state in the band frame of
 * that shifted momentum as left by the last right-hand-side call of the
 * interval (diagonal_sigma, PiFlux.cpp:1037-1049). */
hhg_status hhg_occupations(hhg_plan* plan, const int64_t* k_index, int64_t n_sel, double* occ_out);

/* ---- host utilities (no GPU needed) ---- */

/* Gauss-Legendre nodes/weights on [-1, 1], nodes ascending. */
hhg_status hhg_gauss_legendre(int32_t n, double* x, double* w);

/* Number of points of a k-grid, and the points/weights themselves (kz may be
 * null except for HHG_KGRID_PIFLUX_CONTINUUM; it is 0 for 2D grids). */
int64_t hhg_kgrid_size(const hhg_kgrid_desc* g);
hhg_status hhg_kgrid_build(const hhg_kgrid_desc* g, double* kx, double* ky, double* kz, double* w);

/* Coefficients of the clamped cubic spline through n >= 2 samples spaced dt (the
 * function Boost's cardinal_cubic_b_spline builds, Laser.hpp:8-11): for interval i,
 * 4 doubles c with A(t_i + x dt) = c0 + x (c1 + x (c2 + x c3)), x in [0, 1];
 * coef holds 4 (n - 1) doubles. */
hhg_status hhg_spline_coefficients(const double* samples, int64_t n, double dt, double slope_begin, double slope_end,
                                   double* coef);

/* ---- the reference's tabulated lasers, in its units ---- */
typedef enum hhg_active_pulse {
    HHG_ACTIVE_BOTH = 0, /* Active::Both */
    HHG_ACTIVE_A = 1,    /* Active::A */
    HHG_ACTIVE_B = 2     /* Active::B */
} hhg_active_pulse;

/* DoubleLaser<CosineLaser> (gauss = 0) or DoubleLaser<GaussLaser> (gauss = 1), DoubleLaser.hpp:8-74:
 * photon_energy in units of the experiment's 6.0914 meV (ExperimentParams.hpp:12), E0 in MV/cm,
 * model_ratio the system's laser_model_ratio, delay of pulse B in ps; pulses A and B of field 1.6 E0 and
 * 0.7 E0; samples[1000] over [0, 1.5 x the base duration] become a HHG_PULSE_SPLINE pulse in *out
 * (out->samples points to samples).  *momentum_amplitude (may be null) receives the laser's own
 * amplitude (Laser.cpp:33-38). */
hhg_status hhg_reference_double_laser(int32_t gauss, double photon_energy, double E0, double model_ratio,
                                      double n_cycles, double delay_ps, int32_t active, double* samples,
                                      hhg_pulse_desc* out, double* momentum_amplitude);

/* ExperimentalLaser (ExperimentalLaser.cpp:8-257) from measured fields E_A, E_B [MV/cm] of the
 * reference's experiment layout (201 samples 0.03335555 ps apart, ExperimentParams.hpp:5-11): the
 * vector potential of each field with its polynomial tail, splined, shifted by shift_ps (B later for
 * shift >= 0), scaled by 0.1 E0 model_ratio / hbar and tabulated on the spline's own step.  Returns the
 * number of samples for photon_energy (units of 6.0914 meV), or fills samples[n] and *out. */
int64_t hhg_reference_experimental_laser_size(double photon_energy);
hhg_status hhg_reference_experimental_laser(double photon_energy, double E0, double model_ratio, double shift_ps,
                                            int32_t active, const double* E_A, const double* E_B, int32_t n_measured,
                                            double* samples, int64_t n_samples, hhg_pulse_desc* out,
                                            double* momentum_amplitude);

/* A(t) at n times (host evaluation, Laser::laser_function). */
hhg_status hhg_pulse_evaluate(const hhg_pulse_desc* p, const double* t, double* A, int64_t n);

/* End-corrected Fourier integral of j (n_measurements + 1 samples), zero
 * padded to zero_padding * (N + 1); outputs have zero_padding*(N+1)/2 + 1
 * entries.  Keeps the reference's conventions (see hhg_amd spectrum notes). */
hhg_status hhg_trapezoidal_spectrum(const double* j, const hhg_time_desc* time, int32_t zero_padding,
                                    double* freq, double* re, double* im);

/* Spectrum of a continuous-laser run (sources/hhg.cpp:264-275): the first N
 * samples of j times the Welch window (WelchWindow.hpp:11-13), real-to-complex
 * DFT (FFTW r2c, FFT.cpp:8-42), frequencies k / Delta; outputs N/2 + 1 entries. */
hhg_status hhg_welch_spectrum(const double* j, const hhg_time_desc* time, double* freq, double* re, double* im);

/* ---- step logs (training data for the proposal network) ---- */

/* Enables logging in later executes: for every stride-th interval of every
 * trajectory (staggered by lane) one record of 10 floats: the 8 network
 * features at the interval start, log(h_opt / Delta) with h_opt = h 0.9
 * max(err, 1e-16)^(-1/8) of the interval's first accepted step h (error ratio
 * err; the controller's growth rule without its 4.5 cap), and log(attempts in
 * the interval).  Intervals split at an envelope corner are not logged.
 * At most capacity records are kept.  capacity = 0 disables logging.  No
 * counterpart in the reference (its solver keeps no step history). */
hhg_status hhg_plan_set_step_log(hhg_plan* plan, int64_t capacity, int32_t stride);

/* Copies the records of the last execute (at most capacity) and their count. */
hhg_status hhg_plan_get_step_log(hhg_plan* plan, float* records, int64_t capacity, int64_t* n_records);

/* ---- step-proposal network: 8 -> 64 -> 64 -> 64 -> 2, ReLU, BF16 weights and
 * activations with FP32 accumulation; FP32 master weights (no counterpart in
 * the reference) ---- */

int64_t hhg_mlp_param_count(void);
hhg_status hhg_mlp_create(int32_t device, uint64_t seed, hhg_mlp** mlp);
hhg_status hhg_mlp_destroy(hhg_mlp* mlp);
/* affine normalisation: x_n = (x - in_mean) / in_std, y = y_n * out_std + out_mean */
hhg_status hhg_mlp_set_normalization(hhg_mlp* mlp, const float* in_mean, const float* in_std, const float* out_mean,
                                     const float* out_std);
hhg_status hhg_mlp_get_params(const hhg_mlp* mlp, float* params);
hhg_status hhg_mlp_set_params(hhg_mlp* mlp, const float* params);
/* forward pass of n samples (x: n x 8, y: n x 2), one thread per sample */
hhg_status hhg_mlp_forward(hhg_mlp* mlp, const float* x, int64_t n, float* y);
/* the same network evaluated by the solve kernel's wave-level MFMA path */
hhg_status hhg_mlp_forward_mfma(hhg_mlp* mlp, const float* x, int64_t n, float* y);
/* gradient of the mean loss 0.5 |y_n - t_n|^2 over n samples w.r.t. the FP32 parameters */
hhg_status hhg_mlp_gradients(hhg_mlp* mlp, const float* x, const float* t, int64_t n, float* grad, float* loss);
/* Adam on shuffled batches; epoch_loss[epochs] receives the mean batch loss per epoch (may be NULL) */
hhg_status hhg_mlp_train(hhg_mlp* mlp, const float* x, const float* t, int64_t n, int32_t epochs, int32_t batch,
                         float lr, uint64_t seed, float* epoch_loss);

/* ---- exact accumulators (shard reduction) ---- */

/* As hhg_plan_execute, but returns j as the exact signed 128-bit fixed-point
 * sums the device accumulates (2 words per sample: low, high; value =
 * (high * 2^64 + low) * 2^-80).  Shards of the same plan description added
 * with hhg_accumulators_add equal the unsharded accumulators bit for bit,
 * whatever the order: this is the exact replacement of the reference's
 * MPI_Reduce(SUM) of j (sources/hhg.cpp:241-243). */
hhg_status hhg_plan_execute_accumulators(hhg_plan* plan, uint64_t* acc_out, hhg_stats* stats);

/* total += part (n samples, 128-bit two's complement with carry). */
void hhg_accumulators_add(uint64_t* total, const uint64_t* part, int64_t n);

/* acc (n samples) -> j as doubles. */
void hhg_accumulators_to_double(const uint64_t* acc, int64_t n, double* j);

#ifdef __cplusplus
}
#endif
#endif
