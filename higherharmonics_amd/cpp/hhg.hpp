// hhg.hpp — C++ host interface over the C ABI (include/hhg_amd.h), shaped
// like the reference's host classes so a caller of the reference finds the
// same names and argument meanings:
//
//   HHG::TimeIntegrationConfig      sources/HHG/TimeIntegrationConfig.hpp:7-15
//   HHG::Laser::ContinuousLaser     sources/HHG/Laser/ContinuousLaser.hpp:6-11
//   HHG::Laser::CosineLaser         sources/HHG/Laser/CosineLaser.hpp:6-19
//   HHG::Laser::GaussLaser          sources/HHG/Laser/GaussLaser.hpp:6-21
//   HHG::Laser::DoubleLaser         sources/HHG/Laser/DoubleLaser.hpp:8-74
//   Laser::laser_function(t)        sources/HHG/Laser/Laser.hpp:32
//   HHG::Systems::DiracSystem       sources/HHG/Systems/DiracSystem.hpp:10-91 (3D cone)
//   HHG::Systems::*::compute_current_density(laser, time_config, rank, n_ranks)
//                                   sources/HHG/Systems/DiracSystem.hpp:58-59,
//                                   PiFlux.hpp:92-93, Honeycomb.hpp:70-71
//   HHG::Dispatch::Plan             the role of Dispatcher::compute
//                                   (sources/HHG/Dispatch/Dispatcher.hpp:29)
//
// Errors: the C ABI returns status codes; this layer turns them into the
// exceptions the reference throws (std::invalid_argument for bad input,
// std::runtime_error otherwise).  Every solve runs on the GPU.
#pragma once

#include <algorithm>
#include <array>
#include <cstddef>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hhg_amd.h"

namespace HHG {

using h_float = double;
constexpr h_float pi = 3.14159265358979323846;

inline void check(hhg_status st) {
    if (st == HHG_SUCCESS) return;
    const std::string msg = hhg_last_error_string();
    if (st == HHG_ERROR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error("hhg status " + std::to_string((int)st) + ": " + msg);
}

struct TimeIntegrationConfig {
    h_float t_begin{};
    h_float t_end{};
    int n_measurements{};
    int n_subdivisions{};

    h_float measure_every() const { return (t_end - t_begin) / n_measurements; }
    h_float dt() const { return measure_every() / n_subdivisions; }
    hhg_time_desc desc() const { return hhg_time_desc{t_begin, t_end, n_measurements, n_subdivisions}; }
};

namespace Laser {

// A(t) = amplitude * envelope(t) * sin(t + cep); amplitude is the reference's
// momentum_amplitude (Laser.hpp:17) already in dimensionless units.
struct Laser {
    h_float momentum_amplitude{};
    h_float t_begin{};
    h_float t_end{};
    virtual ~Laser() = default;
    virtual hhg_pulse_desc desc() const = 0;

    h_float laser_function(h_float t) const {
        const hhg_pulse_desc d = desc();
        h_float A = 0.0;
        check(hhg_pulse_evaluate(&d, &t, &A, 1));
        return A;
    }
};

// The reference's amplitude conversion (Laser.cpp:24-25): momentum amplitude
// = model_ratio * 1e11 * E_0 / photon_energy, E_0 in MV/cm, photon energy in meV.
inline h_float momentum_amplitude(h_float photon_energy, h_float E_0, h_float model_ratio) {
    return model_ratio * 1e11 * E_0 / photon_energy;
}

// envelope 1 (ContinuousLaser.cpp:7-9)
struct ContinuousLaser : Laser {
    h_float cep{};
    explicit ContinuousLaser(h_float amplitude, h_float cep_ = 0.0) : cep(cep_) { momentum_amplitude = amplitude; }
    hhg_pulse_desc desc() const override {
        hhg_pulse_desc d{};
        d.kind = HHG_PULSE_CONTINUOUS;
        d.amp_a = momentum_amplitude;
        d.cep = cep;
        return d;
    }
};

// cos^2 envelope on (t_begin, t_begin + 2 pi n_cycles); begin_shift = pi * t0_offset
struct CosineLaser : Laser {
    h_float n_cycles{}, cep{};
    CosineLaser(h_float amplitude, h_float n_cycles_, h_float begin_shift = 0.0, h_float cep_ = 0.0)
        : n_cycles(n_cycles_), cep(cep_) {
        momentum_amplitude = amplitude;
        t_begin = begin_shift;
        t_end = begin_shift + 2 * pi * n_cycles_;
    }
    hhg_pulse_desc desc() const override {
        hhg_pulse_desc d{};
        d.kind = HHG_PULSE_COS2;
        d.amp_a = momentum_amplitude;
        d.n_cycles = n_cycles;
        d.t0 = t_begin;
        d.cep = cep;
        return d;
    }
};

// Gaussian envelope exp(-(t - center)^2 / 2 sigma^2), window center +- window * sigma
struct GaussLaser : Laser {
    h_float sigma{}, center{}, cep{};
    GaussLaser(h_float amplitude, h_float sigma_, h_float center_ = 0.0, h_float cep_ = 0.0, h_float window = 5.0)
        : sigma(sigma_), center(center_), cep(cep_) {
        momentum_amplitude = amplitude;
        t_begin = center_ - window * sigma_;
        t_end = center_ + window * sigma_;
    }
    // sigma of the reference's GaussLaser(n_cycles) (GaussLaser.cpp:7)
    static h_float reference_sigma(h_float n_cycles) { return pi * n_cycles / std::sqrt(2.0 * std::log(1.0 / 0.05)); }
    hhg_pulse_desc desc() const override {
        hhg_pulse_desc d{};
        d.kind = HHG_PULSE_GAUSS;
        d.amp_a = momentum_amplitude;
        d.sigma = sigma;
        d.t0 = center;
        d.cep = cep;
        return d;
    }
};

// A(t) + B(t - delay), B the same base pulse with amplitude amplitude_b
template <class BaseLaser>
struct DoubleLaser : Laser {
    BaseLaser base;
    h_float amplitude_b{}, delay{};
    DoubleLaser(BaseLaser base_, h_float amplitude_b_, h_float delay_)
        : base(std::move(base_)), amplitude_b(amplitude_b_), delay(delay_) {
        momentum_amplitude = base.momentum_amplitude;
        t_begin = std::min(base.t_begin, base.t_begin + delay_);
        t_end = std::max(base.t_end, base.t_end + delay_);
    }
    hhg_pulse_desc desc() const override {
        hhg_pulse_desc d = base.desc();
        d.kind = d.kind == HHG_PULSE_GAUSS ? HHG_PULSE_DOUBLE_GAUSS : HHG_PULSE_DOUBLE_COS2;
        d.amp_b = amplitude_b;
        d.delay = delay;
        return d;
    }
};

// A sampled pulse through its clamped cubic spline (HHG_PULSE_SPLINE): the
// tables the reference builds with Boost's cardinal_cubic_b_spline
// (Laser.hpp:8-11, 32-37), A = 0 beyond t_end.
struct SplineLaser : Laser {
    std::vector<h_float> samples;
    h_float sample_t0{}, sample_dt{}, slope_begin{}, slope_end{};
    SplineLaser() = default;
    SplineLaser(std::vector<h_float> samples_, h_float t0, h_float dt, h_float t_cut, h_float amplitude = 0.0)
        : samples(std::move(samples_)), sample_t0(t0), sample_dt(dt) {
        momentum_amplitude = amplitude;
        t_begin = t0;
        t_end = t_cut;
    }
    hhg_pulse_desc desc() const override {
        hhg_pulse_desc d{};
        d.kind = HHG_PULSE_SPLINE;
        d.samples = samples.data();
        d.n_samples = (int64_t)samples.size();
        d.sample_t0 = sample_t0;
        d.sample_dt = sample_dt;
        d.slope_begin = slope_begin;
        d.slope_end = slope_end;
        d.t_cut = t_end;
        return d;
    }
};

enum class Active { Both = HHG_ACTIVE_BOTH, A = HHG_ACTIVE_A, B = HHG_ACTIVE_B };

// The experiment's photon energy in meV (ExperimentParams.hpp:12): the tabulated
// lasers take their photon energy in units of it.
constexpr h_float exp_photon_energy = 6.091420442472572;

// DoubleLaser<CosineLaser | GaussLaser> in the reference's units (DoubleLaser.hpp:26-35)
inline SplineLaser reference_double_laser(bool gauss, h_float photon_energy, h_float E_0, h_float model_ratio,
                                          h_float n_cycles, h_float delay_ps, Active active = Active::Both) {
    std::vector<h_float> samples(1000);
    hhg_pulse_desc d{};
    h_float amp = 0.0;
    check(hhg_reference_double_laser(gauss ? 1 : 0, photon_energy, E_0, model_ratio, n_cycles, delay_ps,
                                     (int32_t)active, samples.data(), &d, &amp));
    return SplineLaser(std::move(samples), d.sample_t0, d.sample_dt, d.t_cut, amp);
}

// ExperimentalLaser in the reference's units (ExperimentalLaser.cpp:8-26) from the
// measured fields E_A, E_B [MV/cm] (201 samples each)
inline SplineLaser reference_experimental_laser(h_float photon_energy, h_float E_0, h_float model_ratio,
                                                h_float shift_ps, const std::vector<h_float>& E_A,
                                                const std::vector<h_float>& E_B, Active active = Active::Both) {
    if (E_A.size() != E_B.size()) throw std::invalid_argument("E_A and E_B differ in length");
    const int64_t n = hhg_reference_experimental_laser_size(photon_energy);
    if (n < 2) throw std::invalid_argument("photon_energy must be > 0");
    std::vector<h_float> samples((size_t)n);
    hhg_pulse_desc d{};
    h_float amp = 0.0;
    check(hhg_reference_experimental_laser(photon_energy, E_0, model_ratio, shift_ps, (int32_t)active, E_A.data(),
                                           E_B.data(), (int32_t)E_A.size(), samples.data(), n, &d, &amp));
    return SplineLaser(std::move(samples), d.sample_t0, d.sample_dt, d.t_cut, amp);
}

}  // namespace Laser

namespace Dispatch {

struct Solver {
    int method = HHG_METHOD_RKF78;
    int proposal = HHG_PROPOSAL_RESTART;
    h_float rtol = 1e-10, atol = 1e-10;
    int max_fail = 500;
    int magnus_substeps = 1;
    const hhg_mlp* mlp = nullptr;
    int mlp_infer_every = 1;
    int fallback_after = 0;
    int max_steps = 0;  // accepted steps per sampling interval before a trajectory fails (0: the library's)
    int infer_policy = HHG_INFER_FIRST;
    hhg_solver_desc desc() const {
        hhg_solver_desc d{};
        d.method = method;
        d.proposal = proposal;
        d.rtol = rtol;
        d.atol = atol;
        d.max_fail = max_fail;
        d.magnus_substeps = magnus_substeps;
        d.mlp = mlp;
        d.mlp_infer_every = mlp_infer_every;
        d.fallback_after = fallback_after;
        d.max_steps = max_steps;
        d.infer_policy = infer_policy;
        return d;
    }
};

// A compiled set of jobs on one device; compute() returns one j(t) per job.
class Plan {
   public:
    Plan(const std::vector<hhg_kgrid_desc>& kgrids, const std::vector<hhg_job_desc>& jobs, const Solver& solver,
         int rank = 0, int n_ranks = 1, int device = 0) {
        hhg_plan_desc d{};
        d.n_kgrids = (int32_t)kgrids.size();
        d.kgrids = kgrids.data();
        d.n_jobs = (int32_t)jobs.size();
        d.jobs = jobs.data();
        d.solver = solver.desc();
        d.launch.persistent = 1;
        d.rank = rank;
        d.n_ranks = n_ranks;
        d.device = device;
        hhg_plan* p = nullptr;
        check(hhg_plan_create(&d, &p));
        plan_.reset(p);
        offsets_.resize(jobs.size() + 1);
        check(hhg_plan_job_offsets(p, offsets_.data()));
    }

    std::vector<std::vector<h_float>> compute(hhg_stats* stats = nullptr) {
        std::vector<h_float> flat((size_t)hhg_plan_output_size(plan_.get()));
        check(hhg_plan_execute(plan_.get(), flat.data(), stats));
        std::vector<std::vector<h_float>> out;
        for (size_t j = 0; j + 1 < offsets_.size(); ++j)
            out.emplace_back(flat.begin() + offsets_[j], flat.begin() + offsets_[j + 1]);
        return out;
    }

    // Bloch vectors of job 0 at every sample for the listed k indices: [n_sel][N+1][3]
    std::vector<h_float> trajectories(const std::vector<int64_t>& k_index, int n_samples) {
        std::vector<h_float> out(k_index.size() * (size_t)n_samples * 3);
        check(hhg_trajectories(plan_.get(), k_index.data(), (int64_t)k_index.size(), out.data()));
        return out;
    }

    // (lower, upper) band occupations of job 0 for the listed k indices: [n_sel][N+1][2]
    std::vector<h_float> occupations(const std::vector<int64_t>& k_index, int n_samples) {
        std::vector<h_float> out(k_index.size() * (size_t)n_samples * 2);
        check(hhg_occupations(plan_.get(), k_index.data(), (int64_t)k_index.size(), out.data()));
        return out;
    }

    // failed and non-finite trajectories per job of the last compute()
    std::pair<std::vector<int64_t>, std::vector<int64_t>> job_failures() const {
        const size_t n = offsets_.size() - 1;
        std::vector<int64_t> failed(n), nonfinite(n);
        check(hhg_plan_job_failures(plan_.get(), failed.data(), nonfinite.data()));
        return {failed, nonfinite};
    }

    void autotune() { check(hhg_plan_autotune(plan_.get(), nullptr)); }
    void capture_graph() { check(hhg_plan_capture_graph(plan_.get())); }

   private:
    struct Del {
        void operator()(hhg_plan* p) const { hhg_plan_destroy(p); }
    };
    std::unique_ptr<hhg_plan, Del> plan_;
    std::vector<int64_t> offsets_;
};

}  // namespace Dispatch

namespace Systems {

// Common part of the three systems: a model descriptor plus its k-grid.
class System {
   public:
    hhg_model_desc model{};
    hhg_kgrid_desc kgrid{};
    Dispatch::Solver solver{};

    // j(t_i), i = 0..N (the reference's compute_current_density)
    std::vector<h_float> compute_current_density(const Laser::Laser* laser, const TimeIntegrationConfig& time_config,
                                                 int rank = 0, int n_ranks = 1, hhg_stats* stats = nullptr) const {
        hhg_job_desc job{};
        job.model = model;
        job.pulse = laser->desc();
        job.time = time_config.desc();
        job.kgrid = 0;
        Dispatch::Plan plan({kgrid}, {job}, solver, rank, n_ranks);
        return take_perp(plan.compute(stats)[0]);
    }

    // the same, executing one plan `repeats` times (timing): the solve time of
    // every execution in *solve_ms, the statistics of the last in *stats
    std::vector<h_float> compute_current_density_repeated(const Laser::Laser* laser,
                                                          const TimeIntegrationConfig& time_config, int repeats,
                                                          std::vector<double>* solve_ms,
                                                          hhg_stats* stats = nullptr) const {
        hhg_job_desc job{};
        job.model = model;
        job.pulse = laser->desc();
        job.time = time_config.desc();
        job.kgrid = 0;
        Dispatch::Plan plan({kgrid}, {job}, solver);
        std::vector<h_float> j;
        for (int r = 0; r < std::max(1, repeats); ++r) {
            hhg_stats st{};
            j = plan.compute(&st)[0];
            if (solve_ms) solve_ms->push_back(st.solve_ms);
            if (stats) *stats = st;
        }
        return take_perp(std::move(j));
    }

    // model.perp_current (honeycomb): j_perp of the last computation, N + 1 samples
    mutable std::vector<h_float> last_j_perp;

    // The Bloch vectors of single points at every sample, [point][N+1][3]: the
    // time evolutions the reference's debug outputs record
    // (compute_current_density_debug, DiracSystem.cpp:333-363, PiFlux.cpp:320-383);
    // kz only for the 3D pi-flux lattice (may be empty)
    std::vector<h_float> trajectories_at(const Laser::Laser* laser, const TimeIntegrationConfig& time_config,
                                         const std::vector<h_float>& kx, const std::vector<h_float>& ky,
                                         const std::vector<h_float>& kz) const {
        std::vector<h_float> w(kx.size(), 1.0);
        hhg_kgrid_desc g{};
        g.kind = HHG_KGRID_EXPLICIT;
        g.kx = kx.data();
        g.ky = ky.data();
        g.w = w.data();
        g.kz = kz.empty() ? nullptr : kz.data();
        g.n_points = (int64_t)kx.size();
        hhg_job_desc job{};
        job.model = model;
        job.model.perp_current = 0;
        job.pulse = laser->desc();
        job.time = time_config.desc();
        job.kgrid = 0;
        Dispatch::Plan plan({g}, {job}, solver);
        std::vector<int64_t> sel(kx.size());
        for (size_t i = 0; i < sel.size(); ++i) sel[i] = (int64_t)i;
        return plan.trajectories(sel, time_config.n_measurements + 1);
    }

    virtual ~System() = default;

   protected:
    System() = default;

    // With model.perp_current the job's output is j, then j_perp: keep j_perp
    // and warn where it does not vanish, as Honeycomb::compute_current_density
    // does (Honeycomb.cpp:328-334)
    std::vector<h_float> take_perp(std::vector<h_float> j) const {
        if (!model.perp_current) return j;
        const size_t n = j.size() / 2;
        last_j_perp.assign(j.begin() + (std::ptrdiff_t)n, j.end());
        j.resize(n);
        for (size_t i = 0; i < n; ++i) {
            const h_float jp = last_j_perp[i];
            if (std::abs(jp) > 1e-6 * std::abs(j[i]) && std::abs(jp) > 1e-6)
                std::cerr << "Warning: j_perp = " << jp << "\t\tj_par = " << j[i] << "!\n";
        }
        return j;
    }
};

// 2D massive Dirac cone d = (v qx, v qy, Delta) on a GL box |kx|, |ky| <= k_max
class DiracSystem2D : public System {
   public:
    DiracSystem2D(h_float v, h_float mass, h_float k_max, int nx, int ny) {
        model.kind = HHG_MODEL_DIRAC2D;
        model.v = v;
        model.mass = mass;
        kgrid.kind = HHG_KGRID_GL_BOX;
        kgrid.nx = nx;
        kgrid.ny = ny;
        kgrid.x0 = kgrid.y0 = -k_max;
        kgrid.x1 = kgrid.y1 = k_max;
    }
    DiracSystem2D& dephasing(h_float T2) {
        model.relax = HHG_RELAX_T2;
        model.relax_time = T2;
        return *this;
    }
    DiracSystem2D& diagonal_relaxation(h_float tau) {
        model.relax = HHG_RELAX_TO_INITIAL;
        model.relax_time = tau;
        return *this;
    }
};

// The reference's 3D Dirac cone (DiracSystem.hpp:10-91) in its own units:
// temperature in K, E_F in meV, v_F in m/s, band width in photon energies,
// photon energy in meV, diagonal relaxation time in fs (<= 0: none).  Field
// along z; points (k_z, kappa) on the grid of compute_current_density with
// n_kappa Gauss-Legendre nodes in kappa.  Without relaxation the reference's
// moment Magnus step is used (DiracSystem.cpp:285), with relaxation RKF78 at
// its tolerances abs 1e-12, rel 1e-8 (DiracSystem.cpp:14-15, 361-406).
class DiracSystem : public System {
   public:
    static constexpr h_float hbar = 6.582119569509065698e-13;  // meV s (GlobalDefinitions.hpp:29)
    static constexpr h_float k_B = 0.08617333262145177434;     // meV / K (GlobalDefinitions.hpp:30)
    h_float v_F, photon_energy;

    DiracSystem(h_float temperature, h_float E_F, h_float v_F_, h_float band_width, h_float photon_energy_,
                h_float diagonal_relaxation_time, int n_z, int n_kappa)
        : v_F(v_F_), photon_energy(photon_energy_) {
        model.kind = HHG_MODEL_DIRAC3D;
        model.fermi_energy = E_F / photon_energy;
        model.beta = temperature > 0.0 ? photon_energy / (k_B * temperature) : 0.0;  // 0: T = 0
        if (diagonal_relaxation_time > 0.0) {
            model.relax = HHG_RELAX_TO_INITIAL;
            // inverse time (1e15 hbar) / (tau photon_energy) in omega_L (DiracSystem.cpp:64)
            model.relax_time = (diagonal_relaxation_time * photon_energy) / (1e15 * hbar);
            solver.method = HHG_METHOD_RKF78;
            solver.atol = 1e-12;
            solver.rtol = 1e-8;
        } else {
            solver.method = HHG_METHOD_MAGNUS_MOMENTS;
        }
        kgrid.kind = HHG_KGRID_DIRAC_CYLINDER;
        kgrid.nx = n_z;
        kgrid.ny = n_kappa;
        kgrid.x1 = band_width;
    }
    // hbar v_F / photon energy (DiracSystem.hpp:30-32)
    h_float laser_model_ratio(h_float photon_energy_) const { return hbar * v_F / photon_energy_; }
    // Gauss-Legendre in k_z as well, the quadrature of sources/hhg_gauss.cpp:71-133
    DiracSystem& gauss_in_z() {
        kgrid.kind = HHG_KGRID_DIRAC_GAUSS;
        return *this;
    }
};

// The reference's 3D pi-flux lattice (PiFlux.hpp:25-147) in its own units:
// temperature in K, E_F in meV, v_F in m/s, band width in photon energies,
// photon energy in meV, relaxation times in fs (diagonal <= 0: none;
// off-diagonal <= 0: the diagonal one, PiFluxDispatcher.cpp:18-34).  Field
// along z on the continuum-limit grid (z_range, N_coarse; the reference's
// 32 and 4, PiFlux.hpp:12-29); RKF78 at abs 1e-12, rel 1e-8 with the restart
// controller (PiFlux.cpp:20-21, 96-110); with relaxation toward the
// instantaneous equilibrium (PiFlux.cpp:113-220).
class PiFlux : public System {
   public:
    static constexpr h_float sqrt_3 = 1.732050807568877293527446341;   // GlobalDefinitions.hpp:32
    static constexpr h_float sqrt_12 = 3.464101615137754587054892683;  // GlobalDefinitions.hpp:33
    h_float v_F, band_width, photon_energy;

    PiFlux(h_float temperature, h_float E_F, h_float v_F_, h_float band_width_, h_float photon_energy_,
           h_float diagonal_relaxation_time, h_float offdiagonal_relaxation_time, int z_range = 32, int n_coarse = 4)
        : v_F(v_F_), band_width(band_width_), photon_energy(photon_energy_) {
        model.kind = HHG_MODEL_PIFLUX3D;
        model.fermi_energy = E_F / photon_energy;
        model.beta = temperature > 0.0 ? photon_energy / (DiracSystem::k_B * temperature) : 0.0;  // 0: T = 0
        model.hopping = band_width / sqrt_12;                                                     // PiFlux.cpp:66
        if (diagonal_relaxation_time > 0.0) {
            // inverse times (1e15 hbar) / (tau photon_energy) in omega_L (PiFlux.cpp:68-69)
            const h_float tau_o = offdiagonal_relaxation_time > 0.0 ? offdiagonal_relaxation_time
                                                                     : diagonal_relaxation_time;
            model.relax = HHG_RELAX_EQUILIBRIUM;
            model.relax_time = (diagonal_relaxation_time * photon_energy) / (1e15 * DiracSystem::hbar);
            model.relax_time_offdiag = (tau_o * photon_energy) / (1e15 * DiracSystem::hbar);
        }
        solver.method = HHG_METHOD_RKF78;
        solver.proposal = HHG_PROPOSAL_RESTART;
        solver.atol = 1e-12;
        solver.rtol = 1e-8;
        kgrid.kind = HHG_KGRID_PIFLUX_CONTINUUM;
        kgrid.nx = z_range;
        kgrid.ny = n_coarse;
    }
    // the lattice constant sqrt3 hbar v_F / (photon_energy band_width) (PiFlux.cpp:67, PiFlux.hpp:78-80)
    h_float laser_model_ratio() const { return sqrt_3 * DiracSystem::hbar * v_F / (photon_energy * band_width); }

    // PiFlux::compute_occupation_numbers(laser, time_config, N) (PiFlux.cpp:909-952): the
    // (lower, upper) band occupations on the slice (x pi/N, pi/2, z pi/N) at every sample,
    // stored as the reference's OccupationContainer maps [sample][x][column] with column
    // (z - round(A(t_i) / (pi/N))) mod N (coordinate_shift, PiFlux.cpp:918-921); needs the
    // relaxation (a diagonal relaxation time > 0)
    struct Occupations {
        int N = 0, n_samples = 0;
        std::vector<h_float> lower, upper;  // [n_samples][N][N]
    };
    Occupations compute_occupation_numbers(const Laser::Laser* laser, const TimeIntegrationConfig& time_config,
                                           int N) const {
        if (model.relax != HHG_RELAX_EQUILIBRIUM)
            throw std::invalid_argument("occupation numbers need the relaxation (diagonal_relaxation_time > 0)");
        hhg_job_desc job{};
        job.model = model;
        job.pulse = laser->desc();
        job.time = time_config.desc();
        job.kgrid = 0;
        hhg_kgrid_desc g{};
        g.kind = HHG_KGRID_PIFLUX_OCCUPATION;
        g.nx = N;
        Dispatch::Plan plan({g}, {job}, solver);
        Occupations o;
        o.N = N;
        o.n_samples = time_config.n_measurements + 1;
        std::vector<int64_t> sel((size_t)N * N);
        for (size_t i = 0; i < sel.size(); ++i) sel[i] = (int64_t)i;
        const std::vector<h_float> occ = plan.occupations(sel, o.n_samples);
        o.lower.assign((size_t)o.n_samples * N * N, 0.0);
        o.upper.assign((size_t)o.n_samples * N * N, 0.0);
        const h_float dz = pi / N, me = time_config.measure_every();
        h_float t_begin = time_config.t_begin, t_end = t_begin + me;  // accumulated as PiFlux.cpp:941-949
        for (int i = 0; i < o.n_samples; ++i) {
            const int shift = (int)std::round(laser->laser_function(t_begin) / dz);
            for (int x = 0; x < N; ++x)
                for (int z = 0; z < N; ++z) {
                    const int col = (((z - shift) % N) + N) % N;
                    const size_t src = (((size_t)x * N + z) * o.n_samples + i) * 2;
                    const size_t dst = ((size_t)i * N + x) * N + col;
                    o.lower[dst] = occ[src];
                    o.upper[dst] = occ[src + 1];
                }
            t_begin = t_end;
            t_end += me;
        }
        return o;
    }

    // PiFlux::current_per_energy(laser, time_config, N) (PiFlux.cpp:954-1035): the current
    // on the uniform grid (x, y, +-z) pi/N split by the instantaneous dispersion
    // |m(k - A(t_i))| at each sample: {below sqrt3/2 (Dirac-like states), the rest},
    // n_measurements + 1 samples each
    std::array<std::vector<h_float>, 2> current_per_energy(const Laser::Laser* laser,
                                                           const TimeIntegrationConfig& time_config, int N,
                                                           hhg_stats* stats = nullptr, int repeats = 1,
                                                           std::vector<double>* solve_ms = nullptr) const {
        hhg_job_desc job{};
        job.model = model;
        job.model.split_dispersion = 0.5 * sqrt_3;
        job.pulse = laser->desc();
        job.time = time_config.desc();
        job.kgrid = 0;
        hhg_kgrid_desc g{};
        g.kind = HHG_KGRID_PIFLUX_UNIFORM;
        g.nx = N;
        Dispatch::Plan plan({g}, {job}, solver);
        std::vector<h_float> j;
        for (int r = 0; r < std::max(1, repeats); ++r) {  // > 1: the same plan again (timing)
            hhg_stats st{};
            j = plan.compute(&st)[0];
            if (solve_ms) solve_ms->push_back(st.solve_ms);
            if (stats) *stats = st;
        }
        const size_t n = j.size() / 2;
        return {std::vector<h_float>(j.begin(), j.begin() + (std::ptrdiff_t)n),
                std::vector<h_float>(j.begin() + (std::ptrdiff_t)n, j.end())};
    }
};

// PiFlux square lattice d = 2t (cos(kx - A), cos ky, 0) over [-pi, pi]^2
class PiFlux2D : public System {
   public:
    PiFlux2D(h_float hopping, int nx, int ny) {
        model.kind = HHG_MODEL_PIFLUX2D;
        model.hopping = hopping;
        kgrid.kind = HHG_KGRID_GL_BOX;
        kgrid.nx = nx;
        kgrid.ny = ny;
        kgrid.x0 = kgrid.y0 = -pi;
        kgrid.x1 = kgrid.y1 = pi;
    }
};

// Honeycomb nearest-neighbour model over the hexagonal BZ (Honeycomb.cpp:225-336).
// Dimensionless: hopping t and the GL sizes.  Reference units (Honeycomb.cpp:83-92):
// temperature in K, E_F in meV, v_F in m/s, band width in photon energies, photon
// energy in meV, diagonal relaxation time in fs (<= 0: none; else relaxation toward
// the instantaneous equilibrium, Honeycomb.cpp:160-199), N_k positive GL nodes per
// axis (Honeycomb.hpp:12-21); RKF78 at abs 1e-12, rel 1e-8 (Honeycomb.cpp:14-15).
class Honeycomb : public System {
   public:
    h_float v_F = 0.0, band_width = 0.0, photon_energy = 0.0;

    Honeycomb(h_float hopping, int nx, int ny) {
        model.kind = HHG_MODEL_HONEYCOMB;
        model.hopping = hopping;
        kgrid.kind = HHG_KGRID_GL_HEX;
        kgrid.nx = nx;
        kgrid.ny = ny;
    }
    Honeycomb(h_float temperature, h_float E_F, h_float v_F_, h_float band_width_, h_float photon_energy_,
              h_float diagonal_relaxation_time, int N_k = 64)
        : Honeycomb(band_width_ / 6.0, 2 * N_k, 2 * N_k) {  // hopping W/6 (Honeycomb.cpp:90)
        model.perp_current = 1;  // the reference's j_perp check on every run (Honeycomb.cpp:328-334)
        v_F = v_F_;
        band_width = band_width_;
        photon_energy = photon_energy_;
        model.fermi_energy = E_F / photon_energy;
        model.beta = temperature > 0.0 ? photon_energy / (DiracSystem::k_B * temperature) : 0.0;  // 0: T = 0
        if (diagonal_relaxation_time > 0.0) {
            // inverse time (1e15 hbar) / (tau photon_energy) in omega_L (Honeycomb.cpp:92)
            model.relax = HHG_RELAX_EQUILIBRIUM;
            model.relax_time = (diagonal_relaxation_time * photon_energy) / (1e15 * DiracSystem::hbar);
        }
        solver.method = HHG_METHOD_RKF78;
        solver.proposal = HHG_PROPOSAL_RESTART;
        solver.atol = 1e-12;
        solver.rtol = 1e-8;
    }
    // the lattice constant 4 hbar v_F / (photon_energy band_width) (Honeycomb.cpp:91, Honeycomb.hpp:60-62)
    h_float laser_model_ratio() const { return 4.0 * DiracSystem::hbar * v_F / (photon_energy * band_width); }
};

}  // namespace Systems

namespace Fourier {

// TrapezoidalFFT::compute with the x8 zero padding of sources/hhg.cpp:277-282
struct Spectrum {
    std::vector<h_float> frequencies, real, imag;
};

// continuous lasers: Welch window + real-to-complex DFT (sources/hhg.cpp:264-275)
inline Spectrum welch_spectrum(const std::vector<h_float>& j, const TimeIntegrationConfig& tc) {
    const size_t n = (size_t)tc.n_measurements / 2 + 1;
    Spectrum s{std::vector<h_float>(n), std::vector<h_float>(n), std::vector<h_float>(n)};
    const hhg_time_desc td = tc.desc();
    check(hhg_welch_spectrum(j.data(), &td, s.frequencies.data(), s.real.data(), s.imag.data()));
    return s;
}

inline Spectrum trapezoidal_spectrum(const std::vector<h_float>& j, const TimeIntegrationConfig& tc,
                                     int zero_padding = 8) {
    const size_t n = (size_t)(zero_padding * (tc.n_measurements + 1)) / 2 + 1;
    Spectrum s{std::vector<h_float>(n), std::vector<h_float>(n), std::vector<h_float>(n)};
    const hhg_time_desc td = tc.desc();
    check(hhg_trapezoidal_spectrum(j.data(), &td, zero_padding, s.frequencies.data(), s.real.data(), s.imag.data()));
    return s;
}

}  // namespace Fourier
}  // namespace HHG
