// hhg_run — command-line driver over the C++ host interface, the role of the
// reference's sources/hhg.cpp main(): read a key-value configuration, build
// the system, laser and time grid (the Dispatcher constructors,
// sources/HHG/Dispatch/*Dispatcher.cpp), compute j(t) on the GPU, take the
// spectrum and write one JSON document (gzip-compressed when the output name
// ends in .gz, as the reference's current_density.json.gz).
//
//   hhg_run <config> [output.json[.gz]]
//
// Configuration: whitespace-separated "key value" lines, '#' starts a comment
// (the reference's InputFileReader format, sources/hhg.cpp:58-80).
//
// system_type Dirac | PiFlux | Honeycomb: the reference's 3D Dirac cone, 3D
// pi-flux lattice and honeycomb lattice in the reference's own units and keys
// (sources/hhg.cpp:63-80, DiracDispatcher.cpp:15-58, PiFluxDispatcher.cpp:14-110,
// HoneycombDispatcher.cpp:12-60):
//   T [K], E_F [meV], v_F [m/s], band_width [photon energies],
//   field_amplitude [MV/cm], photon_energy [meV], diagonal_relaxation_time and
//   offdiagonal_relaxation_time [fs] (<= 0: none / the diagonal one), laser_type
//   continuous | cosine, n_laser_cycles, t0_offset (cos^2 begin in units of
//   pi); N = n_laser_cycles * 1024 samples (hhg.cpp:87-88).
//   Dirac: without relaxation the moment Magnus step, else RKF78; n_z, and
//   n_kappa (Gauss-Legendre nodes in kappa, default 32) in place of the
//   reference's adaptive kappa trapezoid; z_rule gauss puts n_z Gauss-Legendre
//   nodes in k_z too, the quadrature of the reference's hhg_gauss.cpp.
//   PiFlux: RKF78, the continuum-limit grid z_range x n_coarse (32, 4), a cos^2
//   run continues 2 pi after the pulse (PiFluxDispatcher.cpp:51-53).
//   Honeycomb: RKF78, the hexagonal-BZ grid with N_k positive nodes per axis
//   (64), relaxation toward the instantaneous equilibrium, cos^2 + 2 pi
//   (HoneycombDispatcher.cpp:31-35).
//   gauss (PiFlux): the reference's GaussLaser over [t0_offset, t0_offset + 6 pi n].
//   Tabulated lasers (cubic-spline pulses): dcos, dcosA, dcosB, dgauss, dgaussA,
//   dgaussB (PiFlux; B delayed by t0_offset ps) and exp, expA, expB (all three;
//   experiment_data names a file of the measured fields, two columns E_A E_B in
//   MV/cm, 201 rows; B shifted by t0_offset ps); their photon_energy is in
//   units of the experiment's 6.0914 meV (Dispatcher.cpp:10-19).
//   The output carries the reference's fields (hhg.cpp:301-326); continuous
//   lasers use the Welch-window spectrum.
//   PiFlux only: occupations yes | no | only writes occupations.json.gz (band
//   occupation maps lower_band / upper_band [500][n_z][n_z] of the relaxing
//   lattice, hhg.cpp:120-170) and split_current yes | no | only writes
//   split_current.json.gz (dirac_current / non_dirac_current on the n_z^3
//   uniform grid, hhg.cpp:174-219), both next to the output file; "only" skips
//   what follows.  With either, tabulated pulses stop at 0.75 t_end.
//   Every file output gets a <name>metadata.json.gz beside it (build commit,
//   date, compiler, ABI version), the role of the reference's metadata files.
//
// The 2D models in dimensionless units (time 1/omega_L, energy hbar omega_L):
//   system_type       Dirac2D | PiFlux2D | Honeycomb2D
//   laser_type        cosine | gauss | double_gauss | double_cosine | continuous
//   field_amplitude   A0 (momentum amplitude)
//   n_laser_cycles    cos^2 length in cycles (default 8)
//   t0_offset         window begin of the cos^2 pulse in units of pi (default -n_laser_cycles)
//   sigma, center     Gaussian width and centre (default 4 pi, 0)
//   amplitude_b, delay   second pulse of a double pulse
//   cep               carrier-envelope phase
//   v, mass, k_max    Dirac-2D parameters (1, 0.1, 3)
//   hopping           lattice hopping t (1)
//   nx, ny            k-grid size
//   T2, tau           transverse dephasing / relaxation to the initial state (0 = off)
//   n_measurements    sampling intervals (default n_laser_cycles * 1024, sources/hhg.cpp:87-88)
//   n_subdivisions    controller restart divisor (50)
//   extra_time        time appended after a cos^2 pulse (2 pi for the lattices,
//                     HoneycombDispatcher.cpp:35, PiFluxDispatcher.cpp:53)
//   method            rkf78 | magnus4;  rtol, atol, magnus_substeps
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>

#include <zlib.h>

#include "hhg.hpp"

namespace {

std::map<std::string, std::string> read_config(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::invalid_argument("cannot open " + path);
    std::map<std::string, std::string> kv;
    std::string line;
    while (std::getline(in, line)) {
        const auto hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ls(line);
        std::string k, v;
        if (ls >> k >> v) kv[k] = v;
    }
    return kv;
}

double num(const std::map<std::string, std::string>& kv, const std::string& k, double def) {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::stod(it->second);
}

std::string str(const std::map<std::string, std::string>& kv, const std::string& k, const std::string& def) {
    auto it = kv.find(k);
    return it == kv.end() ? def : it->second;
}

void write_array(std::ostream& o, const char* name, const std::vector<double>& v, bool last = false) {
    o << "  \"" << name << "\": [";
    char buf[40];
    for (size_t i = 0; i < v.size(); ++i) {
        std::snprintf(buf, sizeof buf, "%.17g", v[i]);
        o << (i ? ", " : "") << buf;
    }
    o << "]" << (last ? "\n" : ",\n");
}

std::string num_str(double v) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

// the reference's SI strings: std::to_string of a double ("%f")
std::string fixed_str(double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%f", v);
    return buf;
}

std::string time_stamp() {
    const std::time_t now = std::time(nullptr);
    char buf[64];
    std::strftime(buf, sizeof buf, "%Y-%m-%d %H:%M:%S", std::localtime(&now));
    return buf;
}

// [n_s][N][N] maps as nested arrays
void write_maps(std::ostream& o, const char* name, const std::vector<double>& v, int n_s, int N, bool last = false) {
    o << "  \"" << name << "\": [";
    char buf[40];
    for (int i = 0; i < n_s; ++i) {
        o << (i ? ", [" : "[");
        for (int x = 0; x < N; ++x) {
            o << (x ? ", [" : "[");
            for (int z = 0; z < N; ++z) {
                std::snprintf(buf, sizeof buf, "%.17g", v[((size_t)i * N + x) * N + z]);
                o << (z ? ", " : "") << buf;
            }
            o << "]";
        }
        o << "]";
    }
    o << "]" << (last ? "\n" : ",\n");
}

#ifndef HHG_GIT_COMMIT
#define HHG_GIT_COMMIT "unknown"
#endif

// the metadata file written beside every output (the role of the reference's
// metadata.json.gz, hhg.cpp:165-169, 212-216, 337-341: information on the build
// that produced the data)
std::string metadata_json() {
    std::ostringstream o;
    o << "{\n  \"hhg_git_commit\": \"" << HHG_GIT_COMMIT << "\",\n  \"hhg_build_date\": \"" << __DATE__ << " "
      << __TIME__ << "\",\n  \"hhg_compiler\": \"" << __VERSION__ << "\",\n  \"hhg_abi_version\": "
      << hhg_abi_version() << ",\n  \"hhg_target\": \"gfx950\",\n  \"hhg_run_time\": \"" << time_stamp()
      << "\"\n}\n";
    return o.str();
}

// a file next to the main output (the working directory without one)
std::string sibling(int argc, char** argv, const std::string& name) {
    if (argc <= 2) return name;
    const std::string path = argv[2];
    const auto slash = path.find_last_of('/');
    return slash == std::string::npos ? name : path.substr(0, slash + 1) + name;
}

// a plain file, or gzip when the name ends in .gz
void write_file(const std::string& path, const std::string& text) {
    if (path.size() > 3 && path.compare(path.size() - 3, 3, ".gz") == 0) {
        gzFile f = gzopen(path.c_str(), "wb");
        if (!f || gzwrite(f, text.data(), (unsigned)text.size()) != (int)text.size() || gzclose(f) != Z_OK)
            throw std::runtime_error("cannot write " + path);
        return;
    }
    std::ofstream out(path);
    if (!out || !(out << text)) throw std::runtime_error("cannot write " + path);
}

// stdout, or a plain file (gzip when the name ends in .gz) with metadata.json.gz beside it
void emit(const std::string& text, int argc, char** argv) {
    if (argc <= 2) {
        std::cout << text;
        return;
    }
    write_file(argv[2], text);
    write_file(sibling(argc, argv, "metadata.json.gz"), metadata_json());
}

// The reference's 3D runs in its own units (Dirac/PiFlux dispatchers, hhg.cpp main).
int run_reference_units(const std::string& system, const std::map<std::string, std::string>& kv, int argc,
                        char** argv) {
    using namespace HHG;
    const double T = num(kv, "T", 0.0), E_F = num(kv, "E_F", 0.0), v_F = num(kv, "v_F", 1.5e5);
    const double band_width = num(kv, "band_width", 10.0), E0 = num(kv, "field_amplitude", 1.0);
    const double photon_energy = num(kv, "photon_energy", 5.25), tau = num(kv, "diagonal_relaxation_time", -1.0);
    const double tau_o = num(kv, "offdiagonal_relaxation_time", -1.0);
    const std::string laser_type = str(kv, "laser_type", "continuous");
    const int n_cycles = (int)num(kv, "n_laser_cycles", 8);
    const double t0_offset = num(kv, "t0_offset", 0.0);
    const int n_z = (int)num(kv, "n_z", 64), n_kappa = (int)num(kv, "n_kappa", 32);
    const int z_range = (int)num(kv, "z_range", 32), n_coarse = (int)num(kv, "n_coarse", 4);
    const int N = (int)num(kv, "n_measurements", n_cycles * 1024);

    const int N_k = (int)num(kv, "N_k", 64);
    const bool dirac = system == "Dirac", honeycomb = system == "Honeycomb";
    // occupation maps and split current: yes | no | only (hhg.cpp:79-80, 120-219); with
    // either the PiFlux dispatcher stops tabulated pulses at 0.75 t_end (PiFluxDispatcher.cpp:44)
    const std::string occupations = str(kv, "occupations", "no"), split_current = str(kv, "split_current", "no");
    const bool extra_outputs = occupations != "no" || split_current != "no";
    if (extra_outputs && system != "PiFlux")
        throw std::invalid_argument("occupations and split_current are outputs of system_type PiFlux");
    const double t_end_scale = extra_outputs ? 0.75 : 1.0;
    // the tabulated lasers take the photon energy in units of the experiment's
    // (Dispatcher.cpp:10-19); the lattices' systems see the scaled energy, the
    // Dirac system the raw one with the scaled energy only in its laser ratio
    // (DiracDispatcher.cpp:38-40)
    const bool is_exp = laser_type == "exp" || laser_type == "expA" || laser_type == "expB";
    const bool is_double = laser_type.rfind("dcos", 0) == 0 || laser_type.rfind("dgauss", 0) == 0;
    const bool scaled = is_exp || is_double;
    const double e_sys = (scaled && !dirac) ? photon_energy * Laser::exp_photon_energy : photon_energy;
    Systems::DiracSystem dsys(T, E_F, v_F, band_width, e_sys, tau, n_z, n_kappa);
    const std::string z_rule = str(kv, "z_rule", "uniform");
    if (z_rule == "gauss")
        dsys.gauss_in_z();  // hhg_gauss.cpp's Gauss-Legendre (k_z, kappa) quadrature
    else if (z_rule != "uniform")
        throw std::invalid_argument("z_rule is uniform (DiracSystem.cpp:280-306) or gauss (hhg_gauss.cpp:71-133)");
    Systems::PiFlux psys(T, E_F, v_F, band_width, e_sys, tau, tau_o, z_range, n_coarse);
    Systems::Honeycomb hsys(T, E_F, v_F, band_width, e_sys, tau, N_k);
    const Systems::System& sys = dirac ? static_cast<const Systems::System&>(dsys)
                                 : honeycomb ? static_cast<const Systems::System&>(hsys)
                                             : psys;
    const double ratio = dirac ? dsys.laser_model_ratio(scaled ? photon_energy * Laser::exp_photon_energy : photon_energy)
                         : honeycomb ? hsys.laser_model_ratio()
                                     : psys.laser_model_ratio();
    const double A0 = Laser::momentum_amplitude(photon_energy, E0, ratio);
    double laser_photon_energy = photon_energy;
    std::unique_ptr<Laser::Laser> laser;
    TimeIntegrationConfig tc;
    auto active_of = [&](const std::string& base) {
        if (laser_type == base + "A") return Laser::Active::A;
        if (laser_type == base + "B") return Laser::Active::B;
        return Laser::Active::Both;
    };
    if (laser_type == "continuous") {
        laser = std::make_unique<Laser::ContinuousLaser>(A0);
        tc = {-n_cycles * pi, n_cycles * pi, N, 50};
    } else if (laser_type == "cosine") {
        laser = std::make_unique<Laser::CosineLaser>(A0, (double)n_cycles, pi * t0_offset);
        // the lattice continues one cycle after the pulse so the relaxation sets in
        tc = {laser->t_begin, laser->t_end + (dirac ? 0.0 : 2.0 * pi), N, 50};
    } else if (laser_type == "gauss" && !dirac && !honeycomb) {
        // GaussLaser(n_cycles, begin shift t0_offset): window [t0, t0 + 6 pi n], centre in the
        // middle, sigma = pi n / sqrt(2 ln 20) (GaussLaser.cpp:9-12, GaussLaser.hpp:13-16,
        // PiFluxDispatcher.cpp:86-88)
        const double base = 2.0 * pi * 3.0 * n_cycles;
        auto g = std::make_unique<Laser::GaussLaser>(A0, Laser::GaussLaser::reference_sigma(n_cycles),
                                                     t0_offset + 0.5 * base);
        g->t_begin = t0_offset;
        g->t_end = t0_offset + base;
        laser = std::move(g);
        tc = {laser->t_begin, t_end_scale * laser->t_end, N, 50};
    } else if (is_double && !dirac && !honeycomb) {
        // DoubleLaser<CosineLaser | GaussLaser> with the B delay t0_offset in ps (PiFluxDispatcher.cpp:74-104)
        const bool gauss = laser_type.rfind("dgauss", 0) == 0;
        laser = std::make_unique<Laser::SplineLaser>(Laser::reference_double_laser(
            gauss, photon_energy, E0, ratio, (double)n_cycles, t0_offset, active_of(gauss ? "dgauss" : "dcos")));
        laser_photon_energy = photon_energy * Laser::exp_photon_energy;
        tc = {laser->t_begin, t_end_scale * laser->t_end, N, 50};
    } else if (is_exp) {
        // ExperimentalLaser from the measured fields in experiment_data (two columns E_A E_B
        // in MV/cm, 201 rows), B shifted by t0_offset ps (DiracDispatcher.cpp:38-51)
        const std::string path = str(kv, "experiment_data", "");
        std::ifstream in(path);
        if (!in) throw std::invalid_argument("laser_type " + laser_type + " needs experiment_data <file>");
        std::vector<double> ea, eb;
        std::string line;
        while (std::getline(in, line)) {
            const auto hash = line.find('#');
            if (hash != std::string::npos) line.resize(hash);
            std::istringstream ls(line);
            double a, b;
            if (ls >> a >> b) ea.push_back(a), eb.push_back(b);
        }
        laser = std::make_unique<Laser::SplineLaser>(
            Laser::reference_experimental_laser(photon_energy, E0, ratio, t0_offset, ea, eb, active_of("exp")));
        laser_photon_energy = photon_energy * Laser::exp_photon_energy;
        tc = {laser->t_begin, t_end_scale * laser->t_end, N, 50};
    } else {
        throw std::invalid_argument("Laser type '" + laser_type + "' is not recognized!");
    }
    const bool continuous = laser_type == "continuous";

    // the fields every output of hhg.cpp carries (hhg.cpp:141-159) and the PiFlux
    // dispatcher's special_information (PiFluxDispatcher.cpp:136-141)
    auto head = [&](std::ostringstream& o, const TimeIntegrationConfig& t) {
        const int n_t = t.n_measurements;
        std::vector<double> lf(n_t + (continuous ? 0 : 1));
        for (size_t i = 0; i < lf.size(); ++i) lf[i] = laser->laser_function(t.t_begin + (double)i * t.measure_every());
        o << "{\n  \"time\": \"" << time_stamp() << "\", \"N\": " << n_t << ",\n";
        write_array(o, "laser_function", lf);
        o << "  \"t_begin\": " << num_str(t.t_begin) << ", \"t_end\": " << num_str(t.t_end)
          << ", \"n_laser_cycles\": " << n_cycles << ", \"n_measurements\": " << n_t + (continuous ? 0 : 1)
          << ",\n  \"T\": " << num_str(T) << ", \"E_F\": " << num_str(E_F) << ", \"v_F\": " << num_str(v_F)
          << ", \"band_width\": " << num_str(band_width) << ", \"field_amplitude\": " << num_str(E0)
          << ", \"photon_energy\": " << num_str(laser_photon_energy) << ", \"laser_type\": \"" << laser_type
          << "\", \"diagonal_relaxation_time\": " << num_str(tau)
          << ", \"offdiagonal_relaxation_time\": " << num_str(tau_o) << ",\n  \"system_type\": \"" << system
          << "\", \"n_z\": " << n_z << ", \"t0_offset\": " << num_str(t0_offset) << ", \"lattice_constant\": \""
          << fixed_str(1e12 * psys.laser_model_ratio()) << " pm\", \"hopping_element\": \""
          << fixed_str(psys.model.hopping * e_sys) << " meV\",\n";
    };
    if (occupations != "no") {
        // PiFlux::compute_occupation_numbers on n_z x n_z points, 499 intervals (hhg.cpp:120-170)
        TimeIntegrationConfig tco = tc;
        tco.n_measurements = 499;
        const auto occ = psys.compute_occupation_numbers(laser.get(), tco, n_z);
        const int n_out = tco.n_measurements + (continuous ? 0 : 1);
        std::ostringstream o;
        head(o, tco);
        write_maps(o, "lower_band", occ.lower, n_out, n_z);
        write_maps(o, "upper_band", occ.upper, n_out, n_z, true);
        o << "}\n";
        write_file(sibling(argc, argv, "occupations.json.gz"), o.str());
        write_file(sibling(argc, argv, "occupations_metadata.json.gz"), metadata_json());
        if (occupations == "only") return 0;
    }
    if (split_current != "no") {
        // PiFlux::current_per_energy on the n_z^3 x 2 uniform grid (hhg.cpp:174-219)
        hhg_stats sst{};
        std::vector<double> sruns;
        const auto sc = psys.current_per_energy(laser.get(), tc, n_z, &sst, (int)num(kv, "timing_repeats", 1), &sruns);
        std::ostringstream o;
        head(o, tc);
        write_array(o, "dirac_current", sc[0]);
        write_array(o, "non_dirac_current", sc[1]);
        o << "  \"trajectories\": " << sst.trajectories << ", \"attempted_steps\": " << sst.attempts
          << ", \"rejects\": " << sst.rejects << ", \"solve_ms\": " << num_str(sst.solve_ms) << ",\n";
        write_array(o, "solve_ms_runs", sruns, true);
        o << "}\n";
        write_file(sibling(argc, argv, "split_current.json.gz"), o.str());
        write_file(sibling(argc, argv, "split_current_metadata.json.gz"), metadata_json());
        if (split_current == "only") return 0;
    }

    // debug_data yes | no | only (hhg.cpp:78, 235-238): the time evolutions of 25 single
    // points beside the current (compute_current_density_debug; only: no current)
    const std::string debug_data = str(kv, "debug_data", "no");
    if (debug_data != "no" && debug_data != "yes" && debug_data != "only")
        throw std::invalid_argument("debug_data is yes, no or only");
    if (debug_data != "no" && honeycomb)  // Honeycomb.cpp:111-115
        throw std::runtime_error("Honeycomb::compute_current_density_debug is not implemented yet!");

    hhg_stats stats{};
    std::vector<double> runs_ms;
    const int repeats = (int)num(kv, "timing_repeats", 1);  // > 1: execute the plan that often (timing)
    const std::vector<double> j = debug_data == "only"
                                      ? std::vector<double>((size_t)N + 1, 0.0)
                                      : sys.compute_current_density_repeated(laser.get(), tc, repeats, &runs_ms, &stats);
    const Fourier::Spectrum sp = continuous ? Fourier::welch_spectrum(j, tc) : Fourier::trapezoidal_spectrum(j, tc);
    std::vector<double> A(N + (continuous ? 0 : 1));
    for (size_t i = 0; i < A.size(); ++i) A[i] = laser->laser_function(tc.t_begin + (double)i * tc.measure_every());

    std::ostringstream o;
    o << "{\n";
    write_array(o, "laser_function", A);
    o << "  \"N\": " << N << ", \"t_begin\": " << num_str(tc.t_begin) << ", \"t_end\": " << num_str(tc.t_end)
      << ", \"n_laser_cycles\": " << n_cycles << ", \"n_measurements\": " << N + (continuous ? 0 : 1) << ",\n";
    write_array(o, "current_density_time", j);
    write_array(o, "current_density_frequency_real", sp.real);
    write_array(o, "current_density_frequency_imag", sp.imag);
    o << "  \"T\": " << num_str(T) << ", \"E_F\": " << num_str(E_F) << ", \"v_F\": " << num_str(v_F)
      << ", \"band_width\": " << num_str(band_width) << ", \"field_amplitude\": " << num_str(E0)
      << ", \"photon_energy\": " << num_str(laser_photon_energy) << ", \"laser_type\": \"" << laser_type
      << "\", \"diagonal_relaxation_time\": " << num_str(tau)
      << ", \"offdiagonal_relaxation_time\": " << num_str(tau_o) << ",\n";
    write_array(o, "frequencies", sp.frequencies);
    o << "  \"zero_padding\": 8, \"system_type\": \"" << system << "\", \"n_z\": " << n_z;
    if (dirac)
        o << ", \"n_kappa\": " << n_kappa;
    else if (honeycomb)
        o << ", \"N_k\": " << N_k;
    else
        o << ", \"z_range\": " << z_range << ", \"n_coarse\": " << n_coarse;
    o << ", \"t0_offset\": " << num_str(t0_offset) << ",\n";
    o << "  \"momentum_amplitude\": " << num_str(laser->momentum_amplitude) << ", \"trajectories\": " << stats.trajectories
      << ", \"attempted_steps\": " << stats.attempts << ", \"rejects\": " << stats.rejects
      << ", \"solve_ms\": " << num_str(stats.solve_ms) << ",\n";
    // the honeycomb's consistency check j_perp (Honeycomb.cpp:328-334; the
    // reference only prints its warnings, which the host library repeats)
    if (honeycomb) write_array(o, "current_density_perp", sys.last_j_perp);
    write_array(o, "solve_ms_runs", runs_ms, true);
    o << "}\n";
    emit(o.str(), argc, argv);
    if (debug_data == "no") return 0;

    // the 25 debug points (n_debug_points, GlobalDefinitions.hpp:34) and their
    // recorded quantity s_z: the 3D Dirac cone at k_z = (n_z/10) dk, kappa_i =
    // (i + 1) kappa_max(k_z) / 26, weighted by k_z kappa / |k|
    // (DiracSystem.cpp:333-363, 427-458); the 3D pi-flux lattice at
    // (pi/2, (0.5 - 0.02 i) pi, 0.49 pi) (PiFlux.cpp:320-383)
    constexpr int n_debug = 25;
    std::vector<double> px(n_debug), py(n_debug), pz, weight(n_debug, 1.0);
    if (dirac) {
        const double dk = 2.0 * band_width / n_z, k_z = (double)(n_z / 10) * dk;
        const double kappa_max = std::sqrt(band_width * band_width - k_z * k_z);
        for (int i = 0; i < n_debug; ++i) {
            px[i] = k_z;
            py[i] = (i + 1) * kappa_max / (n_debug + 1);
            weight[i] = k_z * py[i] / std::hypot(k_z, py[i]);
        }
    } else {
        pz.assign(n_debug, 0.49 * HHG::pi);
        for (int i = 0; i < n_debug; ++i) px[i] = 0.5 * HHG::pi, py[i] = 0.5 * HHG::pi - i * 0.02 * HHG::pi;
    }
    const std::vector<double> traj = sys.trajectories_at(laser.get(), tc, px, py, pz);
    const size_t n_s = (size_t)N + 1;
    std::ostringstream d;
    d << "{\n  \"time\": \"" << time_stamp() << "\",\n";
    write_array(d, "frequencies", sp.frequencies);
    std::vector<std::vector<double>> rho(n_debug, std::vector<double>(n_s)), fre(n_debug), fim(n_debug);
    for (int i = 0; i < n_debug; ++i) {
        for (size_t t = 0; t < n_s; ++t) rho[i][t] = weight[i] * traj[((size_t)i * n_s + t) * 3 + 2];
        if (continuous) {  // the reference transforms the debug sets of pulsed runs only (hhg.cpp:287-291)
            fre[i].assign(n_s, 0.0);
            fim[i].assign(n_s, 0.0);
        } else {
            const Fourier::Spectrum s = Fourier::trapezoidal_spectrum(rho[i], tc, 1);
            fre[i] = s.real;
            fim[i] = s.imag;
        }
    }
    auto write_rows = [&](const char* name, const std::vector<std::vector<double>>& rows, bool last) {
        d << "  \"" << name << "\": [\n";
        for (int i = 0; i < n_debug; ++i) {
            std::ostringstream r;
            write_array(r, "x", rows[i], true);
            std::string s = r.str();  // "  \"x\": [...]\n"
            s = s.substr(s.find('['));
            s.pop_back();
            d << "    " << s << (i + 1 < n_debug ? ",\n" : "\n");
        }
        d << "  ]" << (last ? "\n" : ",\n");
    };
    write_rows("time_evolutions", rho, false);
    write_rows("debug_fft_real", fre, false);
    write_rows("debug_fft_imag", fim, true);
    d << "}\n";
    write_file(sibling(argc, argv, "time_evolution.json.gz"), d.str());
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    using namespace HHG;
    if (argc < 2) {
        std::cerr << "usage: hhg_run <config> [output.json]\n";
        return 2;
    }
    try {
        const auto kv = read_config(argv[1]);
        const std::string system = str(kv, "system_type", "Dirac2D");
        if (system == "Dirac" || system == "PiFlux" || system == "Honeycomb")
            return run_reference_units(system, kv, argc, argv);
        const std::string laser_type = str(kv, "laser_type", "cosine");
        const double A0 = num(kv, "field_amplitude", 0.5);
        const double n_cycles = num(kv, "n_laser_cycles", 8);
        const bool lattice = system != "Dirac2D";

        std::unique_ptr<Laser::Laser> laser;
        if (laser_type == "cosine" || laser_type == "double_cosine") {
            Laser::CosineLaser base(A0, n_cycles, pi * num(kv, "t0_offset", -n_cycles), num(kv, "cep", 0.0));
            if (laser_type == "cosine")
                laser = std::make_unique<Laser::CosineLaser>(base);
            else
                laser = std::make_unique<Laser::DoubleLaser<Laser::CosineLaser>>(base, num(kv, "amplitude_b", A0),
                                                                                  num(kv, "delay", 0.0));
        } else if (laser_type == "gauss" || laser_type == "double_gauss") {
            Laser::GaussLaser base(A0, num(kv, "sigma", 4 * pi), num(kv, "center", 0.0), num(kv, "cep", 0.0));
            if (laser_type == "gauss")
                laser = std::make_unique<Laser::GaussLaser>(base);
            else
                laser = std::make_unique<Laser::DoubleLaser<Laser::GaussLaser>>(base, num(kv, "amplitude_b", A0),
                                                                                 num(kv, "delay", 0.0));
        } else if (laser_type == "continuous") {
            laser = std::make_unique<Laser::ContinuousLaser>(A0, num(kv, "cep", 0.0));
        } else {
            throw std::invalid_argument("Laser type '" + laser_type + "' is not recognized!");
        }

        const bool cos2 = laser_type == "cosine" || laser_type == "double_cosine";
        const double extra = num(kv, "extra_time", (cos2 && lattice) ? 2 * pi : 0.0);
        const bool continuous = laser_type == "continuous";
        const double t_b = continuous ? -n_cycles * pi : laser->t_begin;   // DiracDispatcher.cpp:34-35
        const double t_e = continuous ? n_cycles * pi : laser->t_end + extra;
        TimeIntegrationConfig tc{t_b, t_e, (int)num(kv, "n_measurements", n_cycles * 1024),
                                 (int)num(kv, "n_subdivisions", 50)};

        std::unique_ptr<Systems::System> sys;
        const int nx = (int)num(kv, "nx", 128), ny = (int)num(kv, "ny", 64);
        if (system == "Dirac2D") {
            auto d = std::make_unique<Systems::DiracSystem2D>(num(kv, "v", 1.0), num(kv, "mass", 0.1),
                                                              num(kv, "k_max", 3.0), nx, ny);
            if (num(kv, "T2", 0.0) > 0) d->dephasing(num(kv, "T2", 0.0));
            if (num(kv, "tau", 0.0) > 0) d->diagonal_relaxation(num(kv, "tau", 0.0));
            sys = std::move(d);
        } else if (system == "PiFlux2D") {
            sys = std::make_unique<Systems::PiFlux2D>(num(kv, "hopping", 1.0), nx, ny);
        } else if (system == "Honeycomb2D") {
            sys = std::make_unique<Systems::Honeycomb>(num(kv, "hopping", 1.0), nx, ny);
        } else {
            throw std::invalid_argument("System type '" + system + "' is not recognized!");
        }
        if (lattice && num(kv, "tau", 0.0) > 0) {
            sys->model.relax = HHG_RELAX_TO_INITIAL;
            sys->model.relax_time = num(kv, "tau", 0.0);
        }
        sys->solver.method = str(kv, "method", "rkf78") == "magnus4" ? HHG_METHOD_MAGNUS4 : HHG_METHOD_RKF78;
        sys->solver.rtol = num(kv, "rtol", 1e-10);
        sys->solver.atol = num(kv, "atol", 1e-10);
        sys->solver.magnus_substeps = (int)num(kv, "magnus_substeps", 1);

        hhg_stats stats{};
        const std::vector<double> j = sys->compute_current_density(laser.get(), tc, 0, 1, &stats);
        const Fourier::Spectrum sp = continuous ? Fourier::welch_spectrum(j, tc) : Fourier::trapezoidal_spectrum(j, tc);
        std::vector<double> times(j.size()), A(j.size());
        for (size_t i = 0; i < j.size(); ++i) {
            times[i] = tc.t_begin + (double)i * tc.measure_every();
            A[i] = laser->laser_function(times[i]);
        }

        std::ostringstream o;
        o << "{\n";
        o << "  \"system_type\": \"" << system << "\", \"laser_type\": \"" << laser_type << "\",\n";
        o << "  \"t_begin\": " << tc.t_begin << ", \"t_end\": " << tc.t_end << ", \"n_measurements\": "
          << tc.n_measurements << ", \"zero_padding\": 8,\n";
        o << "  \"trajectories\": " << stats.trajectories << ", \"attempted_steps\": " << stats.attempts
          << ", \"rejects\": " << stats.rejects << ", \"solve_ms\": " << stats.solve_ms << ",\n";
        write_array(o, "time", times);
        write_array(o, "laser_function", A);
        write_array(o, "current_density_time", j);
        write_array(o, "frequencies", sp.frequencies);
        write_array(o, "current_density_frequency_real", sp.real);
        write_array(o, "current_density_frequency_imag", sp.imag, true);
        o << "}\n";
        emit(o.str(), argc, argv);
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "hhg_run: " << e.what() << "\n";
        return 1;
    }
}
