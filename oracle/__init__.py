"""CPU oracle for the HigherHarmonics Bloch hot path.

TEST INFRASTRUCTURE ONLY.  This package is the oracle against which the HIP
product path is compared, and the CPU baseline that bench.py times.  Only
``tests/``, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import it; product code never does.

* ``hhg_oracle.c`` (built by ``oracle/Makefile`` into ``libhhg_oracle.so`` and
  the optimised ``libhhg_oracle_fast.so``) restates the reference's pulses,
  Bloch right-hand sides, Boost.Odeint RKF78 controller with per-interval
  restart, Magnus step and quadrature-weighted current sum.
* :func:`trapezoidal_spectrum` restates the reference's end-corrected Fourier
  integral (sources/HHG/Fourier/TrapezoidalFFT.cpp:11-49) with the x8 zero
  padding of sources/hhg.cpp:36,277-282, on numpy's FFT instead of FFTW.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

MODEL_DIRAC2D, MODEL_PIFLUX2D, MODEL_HONEYCOMB, MODEL_DIRAC3D, MODEL_PIFLUX3D = 0, 1, 2, 3, 4
PULSE_COS2, PULSE_GAUSS, PULSE_DOUBLE_GAUSS, PULSE_DOUBLE_COS2, PULSE_CONTINUOUS, PULSE_SPLINE = 0, 1, 2, 3, 4, 5
RELAX_NONE, RELAX_TO_INITIAL, RELAX_T2, RELAX_EQUILIBRIUM = 0, 1, 2, 3
METHOD_RKF78, METHOD_MAGNUS4, METHOD_MAGNUS_MOMENTS = 0, 1, 2
PROPOSAL_RESTART, PROPOSAL_CARRY = 0, 1


class Model(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("relax", ctypes.c_int32), ("v", ctypes.c_double),
                ("mass", ctypes.c_double), ("hopping", ctypes.c_double), ("theta", ctypes.c_double),
                ("fermi_energy", ctypes.c_double), ("beta", ctypes.c_double), ("relax_time", ctypes.c_double),
                ("relax_time_offdiag", ctypes.c_double), ("ic_rescue", ctypes.c_int32), ("eq_form", ctypes.c_int32),
                ("split_dispersion", ctypes.c_double), ("perp_current", ctypes.c_int32), ("pad2", ctypes.c_int32)]


class Pulse(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("pad", ctypes.c_int32), ("amp_a", ctypes.c_double),
                ("amp_b", ctypes.c_double), ("n_cycles", ctypes.c_double), ("sigma", ctypes.c_double),
                ("t0", ctypes.c_double), ("delay", ctypes.c_double), ("cep", ctypes.c_double),
                ("coef", ctypes.POINTER(ctypes.c_double)), ("n_intervals", ctypes.c_int64),
                ("s_t0", ctypes.c_double), ("s_dt", ctypes.c_double), ("s_t_last", ctypes.c_double)]


class Time(ctypes.Structure):
    _fields_ = [("t_begin", ctypes.c_double), ("t_end", ctypes.c_double), ("n_measurements", ctypes.c_int32),
                ("n_subdivisions", ctypes.c_int32)]


class Solver(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("proposal", ctypes.c_int32), ("rtol", ctypes.c_double),
                ("atol", ctypes.c_double), ("max_fail", ctypes.c_int32), ("magnus_substeps", ctypes.c_int32),
                ("max_steps", ctypes.c_int32), ("pad", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("attempts", ctypes.c_int64), ("rejects", ctypes.c_int64),
                ("failed_trajectories", ctypes.c_int64), ("rhs_evals", ctypes.c_int64)]


_libs: dict = {}


def build(quiet: bool = True) -> None:
    """Compile both oracle libraries with oracle/Makefile."""
    out = subprocess.run(["make", "-C", _HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)


def lib(fast: bool = False) -> ctypes.CDLL:
    name = "libhhg_oracle_fast.so" if fast else "libhhg_oracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    L.or_pulse_A.restype = ctypes.c_double
    L.or_pulse_A.argtypes = [ctypes.POINTER(Pulse), ctypes.c_double]
    L.or_gauss_legendre.argtypes = [ctypes.c_int, dp, dp]
    L.or_spline_coefficients.restype = ctypes.c_int
    L.or_spline_coefficients.argtypes = [dp, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
    L.or_rkf78_tableau.argtypes = [dp, dp, dp, dp]
    L.or_fermi_pair.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp, dp]
    d3 = [ctypes.c_double] * 3
    L.or_initial_state.argtypes = [ctypes.POINTER(Model), *d3, dp]
    L.or_rhs.argtypes = [ctypes.POINTER(Model), ctypes.POINTER(Pulse), *d3, dp, dp, ctypes.c_double, dp]
    L.or_vertex.restype = ctypes.c_double
    L.or_vertex.argtypes = [ctypes.POINTER(Model), *d3, ctypes.c_double, dp]
    L.or_magnus_moments.argtypes = [ctypes.POINTER(Pulse), ctypes.c_double, ctypes.c_double, dp]
    L.or_magnus_moments_step.argtypes = [ctypes.c_double, ctypes.c_double, dp, ctypes.c_double, dp]
    L.or_trajectory.restype = ctypes.c_int
    L.or_trajectory.argtypes = [ctypes.POINTER(Model), ctypes.POINTER(Pulse), ctypes.POINTER(Time),
                                ctypes.POINTER(Solver), *d3, dp, ctypes.POINTER(Stats)]
    L.or_trajectory_attempts.restype = ctypes.c_int
    L.or_trajectory_attempts.argtypes = [ctypes.POINTER(Model), ctypes.POINTER(Pulse), ctypes.POINTER(Time),
                                         ctypes.POINTER(Solver), *d3, ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(Stats)]
    L.or_occupations.restype = ctypes.c_int
    L.or_occupations.argtypes = [ctypes.POINTER(Model), ctypes.POINTER(Pulse), ctypes.POINTER(Time),
                                 ctypes.POINTER(Solver), *d3, dp, ctypes.POINTER(Stats)]
    L.or_current_density.restype = ctypes.c_int64
    L.or_current_density.argtypes = [ctypes.POINTER(Model), ctypes.POINTER(Pulse), ctypes.POINTER(Time),
                                     ctypes.POINTER(Solver), dp, dp, dp, dp, ctypes.c_int64, dp, ctypes.POINTER(Stats),
                                     ctypes.c_int]
    _libs[name] = L
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def model(kind=MODEL_DIRAC2D, v=1.0, mass=0.1, hopping=1.0, theta=0.0, fermi_energy=0.0, beta=0.0,
          relax=RELAX_NONE, relax_time=0.0, relax_time_offdiag=0.0, ic_rescue=0, split_dispersion=0.0,
          eq_form=0, perp_current=0) -> Model:
    """eq_form 1 evaluates the equilibrium-relaxation targets in the HIP
    kernels' one-exponential form (a CPU check of that form); 0 is the
    reference's.  perp_current 1 (honeycomb) appends the reference's j_perp
    to j."""
    return Model(kind, relax, v, mass, hopping, theta, fermi_energy, beta, relax_time, relax_time_offdiag,
                 ic_rescue, eq_form, split_dispersion, perp_current, 0)


def pulse(kind=PULSE_COS2, amp_a=0.5, amp_b=0.0, n_cycles=8.0, sigma=0.0, t0=0.0, delay=0.0, cep=0.0) -> Pulse:
    return Pulse(kind, 0, amp_a, amp_b, n_cycles, sigma, t0, delay, cep)


def spline_coefficients(samples, dt: float, slope_begin: float = 0.0, slope_end: float = 0.0) -> np.ndarray:
    """(n - 1, 4) coefficients of the clamped cubic spline (slope form)."""
    y = np.ascontiguousarray(samples, dtype=np.float64)
    coef = np.empty(4 * max(len(y) - 1, 0))
    if lib().or_spline_coefficients(_dp(y), len(y), dt, slope_begin, slope_end, _dp(coef)) != 0:
        raise ValueError("bad spline arguments")
    return coef.reshape(-1, 4)


def spline_pulse(samples, t0: float, dt: float, slope_begin: float = 0.0, slope_end: float = 0.0,
                 t_cut=None) -> Pulse:
    """A sampled pulse through its clamped cubic spline, A = 0 outside
    [t0, t_cut] (t_cut None: the last sample time)."""
    coef = np.ascontiguousarray(spline_coefficients(samples, dt, slope_begin, slope_end).ravel())
    last = t0 + (len(samples) - 1) * dt
    p = Pulse(PULSE_SPLINE)
    p.coef, p.n_intervals = _dp(coef), len(samples) - 1
    p.s_t0, p.s_dt = t0, dt
    p.s_t_last = last if t_cut is None or t_cut <= t0 else min(t_cut, last)
    p._keep = coef  # the C side reads the coefficients through p.coef
    return p


def time_grid(t_begin, t_end, n_measurements, n_subdivisions=50) -> Time:
    return Time(t_begin, t_end, n_measurements, n_subdivisions)


def solver(method=METHOD_RKF78, rtol=1e-10, atol=1e-10, proposal=PROPOSAL_RESTART, max_fail=500,
           magnus_substeps=1, max_steps=0) -> Solver:
    return Solver(method, proposal, rtol, atol, max_fail, magnus_substeps, max_steps, 0)


def pulse_A(p: Pulse, t) -> np.ndarray:
    L = lib()
    t = np.atleast_1d(np.asarray(t, dtype=np.float64))
    return np.array([L.or_pulse_A(ctypes.byref(p), float(x)) for x in t])


def gauss_legendre(n: int):
    x = np.empty(n)
    w = np.empty(n)
    lib().or_gauss_legendre(n, _dp(x), _dp(w))
    return x, w


def rkf78_tableau():
    c, a, b8, b7 = np.empty(13), np.empty(169), np.empty(13), np.empty(13)
    lib().or_rkf78_tableau(_dp(c), _dp(a), _dp(b8), _dp(b7))
    return c, a.reshape(13, 13), b8, b7


def initial_state(m: Model, kx: float, ky: float, kz: float = 0.0) -> np.ndarray:
    s = np.empty(3)
    lib().or_initial_state(ctypes.byref(m), kx, ky, kz, _dp(s))
    return s


def rhs(m: Model, p: Pulse, kx, ky, s_ref, s, t, kz: float = 0.0) -> np.ndarray:
    """ds/dt; s_ref is the initial state (relaxation toward it) or the state at
    the start of the step (equilibrium relaxation)."""
    out = np.empty(3)
    si = np.ascontiguousarray(s_ref, dtype=np.float64)
    ss = np.ascontiguousarray(s, dtype=np.float64)
    lib().or_rhs(ctypes.byref(m), ctypes.byref(p), kx, ky, kz, _dp(si), _dp(ss), t, _dp(out))
    return out


def vertex(m: Model, kx, ky, A, s, kz: float = 0.0) -> float:
    ss = np.ascontiguousarray(s, dtype=np.float64)
    return lib().or_vertex(ctypes.byref(m), kx, ky, kz, A, _dp(ss))


def trajectory(m: Model, p: Pulse, tc: Time, sv: Solver, kx: float, ky: float, kz: float = 0.0):
    out = np.empty((tc.n_measurements + 1, 3))
    st = Stats()
    r = lib().or_trajectory(ctypes.byref(m), ctypes.byref(p), ctypes.byref(tc), ctypes.byref(sv), kx, ky, kz,
                            _dp(out), ctypes.byref(st))
    if r < 0:
        raise ValueError("unsupported model/solver combination")
    return out, st, r


def attempts_per_interval(m: Model, p: Pulse, tc: Time, sv: Solver, kx, ky, kz=None, fast: bool = False):
    """Attempted steps of every k-point in every sampling interval, int32
    [n_k][N] (one trajectory per k-point, serial)."""
    L = lib(fast)
    kx = np.atleast_1d(np.asarray(kx, dtype=np.float64))
    ky = np.atleast_1d(np.asarray(ky, dtype=np.float64))
    kz = np.zeros_like(kx) if kz is None else np.atleast_1d(np.asarray(kz, dtype=np.float64))
    out = np.zeros((len(kx), tc.n_measurements), dtype=np.int32)
    for i in range(len(kx)):
        r = L.or_trajectory_attempts(ctypes.byref(m), ctypes.byref(p), ctypes.byref(tc), ctypes.byref(sv),
                                     float(kx[i]), float(ky[i]), float(kz[i]),
                                     out[i].ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), None)
        if r < 0:
            raise ValueError("unsupported model/solver combination")
    return out


def occupations(m: Model, p: Pulse, tc: Time, sv: Solver, kx: float, ky: float, kz: float, fast: bool = False):
    """(lower, upper) band occupations [N+1][2] of one 3D pi-flux point with
    equilibrium relaxation (PiFlux::evolve_occupation_numbers, PiFlux.cpp:222-318)."""
    out = np.empty((tc.n_measurements + 1, 2))
    st = Stats()
    r = lib(fast).or_occupations(ctypes.byref(m), ctypes.byref(p), ctypes.byref(tc), ctypes.byref(sv), kx, ky, kz,
                             _dp(out), ctypes.byref(st))
    if r < 0:
        raise ValueError("occupations need the 3D pi-flux lattice with equilibrium relaxation and RKF78")
    return out, st, r


def current_density(m: Model, p: Pulse, tc: Time, sv: Solver, kx, ky, w, n_threads: int = 0,
                    fast: bool = False, kz=None):
    """j(t_i), i = 0..N, and the solver statistics.  Deterministic for any
    thread count (per-chunk partial rows summed in chunk order)."""
    kx = np.ascontiguousarray(kx, dtype=np.float64)
    ky = np.ascontiguousarray(ky, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    kzp = None
    if kz is not None:
        kz = np.ascontiguousarray(kz, dtype=np.float64)
        kzp = _dp(kz)
    split = m.kind == MODEL_PIFLUX3D and m.split_dispersion > 0
    perp = m.kind == MODEL_HONEYCOMB and m.perp_current == 1
    j = np.zeros((2 if split or perp else 1) * (tc.n_measurements + 1))
    st = Stats()
    r = lib(fast).or_current_density(ctypes.byref(m), ctypes.byref(p), ctypes.byref(tc), ctypes.byref(sv),
                                     _dp(kx), _dp(ky), kzp, _dp(w), len(kx), _dp(j), ctypes.byref(st), n_threads)
    if r < 0:
        raise ValueError("unsupported model/solver combination")
    return j, st


def fermi_pair(form: int, c: float, x: float, beta: float):
    """(P - M, P M) of P = f(c + x), M = f(c - x): form 0 the reference's two
    Fermi functions, 1 the HIP kernels' one-exponential form."""
    pm, pp = ctypes.c_double(), ctypes.c_double()
    lib().or_fermi_pair(form, c, x, beta, ctypes.byref(pm), ctypes.byref(pp))
    return pm.value, pp.value


def magnus_moments(p: Pulse, t0: float, h: float) -> np.ndarray:
    out = np.zeros(4)
    lib().or_magnus_moments(ctypes.byref(p), t0, h, _dp(out))
    return out


def magnus_moments_step(kz: float, kappa: float, mom, h: float, s) -> np.ndarray:
    mom = np.ascontiguousarray(mom, dtype=np.float64)
    out = np.array(s, dtype=np.float64)
    lib().or_magnus_moments_step(kz, kappa, _dp(mom), h, _dp(out))
    return out


# --------------------------------------------------------------------------
# The reference's tabulated lasers (ExperimentParams.hpp, DoubleLaser.hpp,
# ExperimentalLaser.cpp) in its units
# --------------------------------------------------------------------------

REF_HBAR = 6.582119569509065698e-13   # meV s, GlobalDefinitions.hpp:29
EXP_DT = 0.03335555                   # ps, ExperimentParams.hpp:5
EXP_N_EXTRA, EXP_N, EXP_BEGIN_ZERO, EXP_ZERO_BUFF = 16, 201, 8, 8   # ExperimentParams.hpp:6-9
EXP_LASER_END = 6.67111 + (EXP_N_EXTRA + EXP_BEGIN_ZERO + EXP_ZERO_BUFF) * EXP_DT   # ps, :11
EXP_PHOTON_ENERGY = 6.091420442472572  # meV, :12


def reference_double_laser_samples(gauss, photon_energy, E0, model_ratio, n_cycles, delay_ps, active=0):
    """(samples, dt, t_max) of DoubleLaser<Cosine|Gauss>Laser (DoubleLaser.hpp:26-73):
    A of field 1.6 E0 and B of 0.7 E0 at t - shift, 1000 samples on [0, 1.5 x duration]."""
    e_ph = EXP_PHOTON_ENERGY * photon_energy
    duration = 2 * np.pi * (3.0 if gauss else 1.0) * n_cycles
    t_max = 1.5 * duration
    dt = t_max / 999
    shift = photon_energy * delay_ps * EXP_PHOTON_ENERGY / (1e12 * REF_HBAR)
    amp = lambda f: model_ratio * 1e11 * f / e_ph
    if gauss:
        sig = np.pi * n_cycles / np.sqrt(2 * np.log(20.0))
        pa = pulse(PULSE_GAUSS, amp_a=amp(1.6 * E0), sigma=sig, t0=0.5 * duration)
        pb = pulse(PULSE_GAUSS, amp_a=amp(0.7 * E0), sigma=sig, t0=0.5 * duration)
    else:
        pa = pulse(PULSE_COS2, amp_a=amp(1.6 * E0), n_cycles=n_cycles, t0=0.0)
        pb = pulse(PULSE_COS2, amp_a=amp(0.7 * E0), n_cycles=n_cycles, t0=0.0)
    t = np.arange(1000) * dt
    A, B = pulse_A(pa, t), pulse_A(pb, t - shift)
    return (A + B if active == 0 else (A if active == 1 else B)), dt, t_max


def experiment_vector_potential(E, dt):
    """ExperimentalLaser's vector_potential (ExperimentalLaser.cpp:28-58): eight zeros,
    A_i = -sum_{j<=i} E_j dt, then over N_extra - 1 steps the quartic with the last
    value, slope -E_last and half-curvature (E_{N-2} - E_{N-1})/(2 dt) that reaches 0
    with zero slope (solved here as a 5x5 system), then zeros."""
    E = np.asarray(E, dtype=np.float64)
    A = np.concatenate([np.zeros(EXP_BEGIN_ZERO), -np.cumsum(E) * dt])
    last, slope, half_curv = A[-1], -E[-1], 0.5 * (E[-2] - E[-1]) / dt
    e = (EXP_N_EXTRA - 1) * dt
    M = np.array([[1, 0, 0, 0, 0], [0, 1, 0, 0, 0], [0, 0, 1, 0, 0], [1, e, e ** 2, e ** 3, e ** 4],
                  [0, 1, 2 * e, 3 * e ** 2, 4 * e ** 3]], dtype=np.float64)
    c = np.linalg.solve(M, [last, slope, half_curv, 0.0, 0.0])
    t = np.arange(1, EXP_N_EXTRA - 1) * dt
    tail = c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * c[4])))
    return np.concatenate([A, tail, np.zeros(EXP_ZERO_BUFF + 2)])


def reference_experimental_samples(photon_energy, E0, model_ratio, shift_ps, E_A, E_B, active=0):
    """(samples, dt, t_end) of ExperimentalLaser::compute_spline (ExperimentalLaser.cpp:60-257)."""
    e_ph = photon_energy * EXP_PHOTON_ENERGY
    to_u = e_ph / (1e12 * REF_HBAR)
    t_end = 1.5 * EXP_LASER_END * to_u
    dt = e_ph * (EXP_DT / (1e12 * REF_HBAR))
    end_u = e_ph * (EXP_LASER_END / (1e12 * REF_HBAR))
    shift = shift_ps * to_u
    sa = spline_pulse(experiment_vector_potential(E_A, EXP_DT), 0.0, dt)
    sb = spline_pulse(experiment_vector_potential(E_B, EXP_DT), 0.0, dt)
    n = int(t_end / dt) + 2
    t = dt * np.arange(n)
    if shift >= 0:
        a = np.where(t <= end_u, pulse_A(sa, t), 0.0)
        tb = t - shift
        b = np.where((tb >= 0) & (tb <= end_u), pulse_A(sb, tb), 0.0)
    else:
        ta = t + shift
        a = np.where((ta >= 0) & (ta <= end_u), pulse_A(sa, ta), 0.0)
        b = np.where(t <= end_u, pulse_A(sb, t), 0.0)
    v = a + b if active == 0 else (a if active == 1 else b)
    return (0.1 / REF_HBAR) * E0 * model_ratio * v, dt, t_end


# --------------------------------------------------------------------------
# Quadrature grids (restating the reference's GL mappings)
# --------------------------------------------------------------------------

def dirac_gauss(n_z: int, n_kappa: int, max_k: float):
    """(k_z, kappa, w) of the Gauss-Legendre (k_z, kappa) quadrature of
    sources/hhg_gauss.cpp:71-133: k_z = max_k x_z over the n_z-point rule, kappa =
    kappa_max (x_r + 1) / 2 over the n_kappa-point rule, kappa_max =
    sqrt(max_k^2 - k_z^2); weight with the Jacobians max_k and kappa_max / 2 times
    the integrand weight k_z kappa / |k| (hhg_gauss.cpp:24-26)."""
    gz, wz = gauss_legendre(n_z)
    gk, wk = gauss_legendre(n_kappa)
    kz = np.repeat(max_k * gz, n_kappa)
    kmax = np.sqrt(max_k * max_k - kz * kz)
    ka = 0.5 * kmax * (np.tile(gk, n_z) + 1.0)
    w = max_k * np.repeat(wz, n_kappa) * 0.5 * kmax * np.tile(wk, n_z) * kz * ka / np.sqrt(kz * kz + ka * ka)
    return kz, ka, w


def dirac_cylinder(n_z: int, n_kappa: int, max_k: float):
    """(k_z, kappa) points of the 3D Dirac cone with the field along z.
    k_z on the uniform grid of DiracSystem::compute_current_density
    (sources/HHG/Systems/DiracSystem.cpp:280-306): (z - n_z/2) dz, z = 1..n_z-1
    except n_z/2, dz = 2 max_k / n_z, weight dz; kappa in (0, kappa_max(k_z)),
    kappa_max = sqrt(max_k^2 - k_z^2) (DiracSystem.cpp:259-263), by n_kappa-point
    Gauss-Legendre where the reference refines an adaptive trapezoid; the
    integrand weight k_z kappa / |k| (DiracSystem.cpp:279)."""
    gk, wk = gauss_legendre(n_kappa)
    dz = 2.0 * max_k / n_z
    kz_all, ka_all, w_all = [], [], []
    for z in range(1, n_z):
        if z == n_z // 2:
            continue
        kz = (z - n_z // 2) * dz
        kmax = np.sqrt(max_k * max_k - kz * kz)
        ka = 0.5 * kmax * (gk + 1.0)
        kz_all.append(np.full(n_kappa, kz))
        ka_all.append(ka)
        w_all.append(dz * 0.5 * kmax * wk * kz * ka / np.sqrt(kz * kz + ka * ka))
    return np.concatenate(kz_all), np.concatenate(ka_all), np.concatenate(w_all)


def piflux_continuum(z_range: int = 32, n_coarse: int = 4):
    """(kx, ky, kz, w) of the 3D pi-flux continuum-limit quadrature
    (PiFlux::current_density_continuum_limit, sources/HHG/Systems/PiFlux.cpp:794-863):
    k_z = +-pi x_z over the z_range positive nodes of the 2 z_range-point rule,
    weight w_z (the +z node first); (kx, ky) by improved_xy_integral
    (PiFlux.cpp:665-792), four Gauss patches on [0, pi] x [0, pi/2] split at
    edge = 0.35 pi, a 2N-point rule per side (N_fine = 8 N_coarse):
    [edge, pi - edge] x [edge, pi/2] with (N_fine, N_fine/2),
    [0, pi] x [0, edge] with (3 N_coarse, N_coarse),
    [0, edge] x [edge, pi/2] and [pi - edge, pi] x [edge, pi/2] with (N_coarse, N_coarse).
    Within a patch x is the slow index, nodes ascending."""
    edge = 0.35 * np.pi
    C = n_coarse
    F = 8 * C
    patches = [(edge, np.pi - edge, F, edge, 0.5 * np.pi, F // 2), (0.0, np.pi, 3 * C, 0.0, edge, C),
               (0.0, edge, C, edge, 0.5 * np.pi, C), (np.pi - edge, np.pi, C, edge, 0.5 * np.pi, C)]
    xy = []
    for xl, xh, nx, yl, yh, ny in patches:
        gx, wx = gauss_legendre(2 * nx)
        gy, wy = gauss_legendre(2 * ny)
        X = 0.5 * (xh - xl) * gx + 0.5 * (xh + xl)
        Y = 0.5 * (yh - yl) * gy + 0.5 * (yh + yl)
        W = np.outer(wx * (0.5 * (xh - xl) * 0.5 * (yh - yl)), wy).ravel()
        xy.append((np.repeat(X, 2 * ny), np.tile(Y, 2 * nx), W))
    px = np.concatenate([a[0] for a in xy])
    py = np.concatenate([a[1] for a in xy])
    pw = np.concatenate([a[2] for a in xy])
    gz, wz = gauss_legendre(2 * z_range)
    out = [], [], [], []
    for z in range(z_range, 2 * z_range):
        for sgn in (1.0, -1.0):
            out[0].append(px)
            out[1].append(py)
            out[2].append(np.full(len(px), sgn * np.pi * gz[z]))
            out[3].append(wz[z] * pw)
    return tuple(np.concatenate(a) for a in out)


def piflux_uniform(n: int):
    """(kx, ky, kz, w) of PiFlux::current_per_energy's grid (PiFlux.cpp:954-1035):
    (x, y, z) pi/n for i = z n^2 + y n + x, each followed by its -z copy, weight 1/n^3."""
    i = np.arange(n ** 3)
    z, y, x = i // (n * n), (i // n) % n, i % n
    d = np.pi / n
    kx = np.repeat(x * d, 2)
    ky = np.repeat(y * d, 2)
    kz = np.stack([d * z, -(d * z)], axis=1).ravel()
    return kx, ky, kz, np.full(2 * n ** 3, 1.0 / n ** 3)


def piflux_occupation(n: int):
    """(kx, ky, kz, w) of PiFlux::compute_occupation_numbers' slice (PiFlux.cpp:909-952):
    (x pi/n, pi/2, z pi/n) for i = x n + z, weight 1/n^2 (unused by the reference)."""
    i = np.arange(n * n)
    d = np.pi / n
    return (i // n) * d, np.full(n * n, 0.5 * np.pi), d * (i % n), np.full(n * n, 1.0 / n ** 2)


def occupation_maps(lower_upper, A, n: int, t_begin: float, delta: float):
    """The reference's occupation maps [N+1][n][n] (lower, upper) from the per-point
    occupations [n*n][N+1][2] of the piflux_occupation grid: point (x, z) at sample
    i goes to column (z - round(A(t_i) / dz)) mod n (coordinate_shift,
    PiFlux.cpp:918-921, 940-946); t_i accumulates t_begin + delta + ... as there."""
    occ = np.asarray(lower_upper)
    n_s = occ.shape[1]
    lower = np.zeros((n_s, n, n))
    upper = np.zeros((n_s, n, n))
    dz = np.pi / n
    x = np.arange(n * n) // n
    z = np.arange(n * n) % n
    for i in range(n_s):
        shift = int(np.round(A[i] / dz)) if np.isfinite(A[i]) else 0
        col = ((z - shift) % n + n) % n
        lower[i, x, col] = occ[:, i, 0]
        upper[i, x, col] = occ[:, i, 1]
    return lower, upper


def gl_box(nx: int, ny: int, x0: float, x1: float, y0: float, y1: float):
    """Tensor GL grid on [x0,x1]x[y0,y1], x-major (x index slowest)."""
    gx, wx = gauss_legendre(nx)
    gy, wy = gauss_legendre(ny)
    X = 0.5 * (x1 - x0) * gx + 0.5 * (x1 + x0)
    Y = 0.5 * (y1 - y0) * gy + 0.5 * (y1 + y0)
    kx = np.repeat(X, ny)
    ky = np.tile(Y, nx)
    w = np.outer(wx * 0.5 * (x1 - x0), wy * 0.5 * (y1 - y0)).ravel()
    return kx, ky, w


def gl_hex(nx: int, ny: int):
    """Hexagonal-BZ GL grid of sources/HHG/Systems/Honeycomb.cpp:229-268:
    x in [-2pi/3, 2pi/3], y in [-ymax(x), ymax(x)], ymax = 4pi/(3 sqrt 3) - |x|/sqrt 3."""
    gx, wx = gauss_legendre(nx)
    gy, wy = gauss_legendre(ny)
    X = (2 * np.pi / 3) * gx
    ymax = 4 * np.pi / (3 * np.sqrt(3)) - np.abs(X) / np.sqrt(3)
    kx = np.repeat(X, ny)
    ky = (ymax[:, None] * gy[None, :]).ravel()
    w = ((2 * np.pi / 3) * wx[:, None] * ymax[:, None] * wy[None, :]).ravel()
    return kx, ky, w


# --------------------------------------------------------------------------
# Spectrum
# --------------------------------------------------------------------------

def trapezoidal_spectrum(j: np.ndarray, t_begin: float, t_end: float, n_measurements: int, zero_padding: int = 8):
    """Restates TrapezoidalFFT::compute (sources/HHG/Fourier/TrapezoidalFFT.cpp:11-49)
    applied to j zero-padded to zero_padding*(N+1) samples (sources/hhg.cpp:277-282),
    including the reference's conventions: the upper-end correction reads the last
    element of the padded input (zero after padding), the DFT term is scaled by
    delta_t twice, and the phase factor uses exp(+i w t_begin) with FFTW's
    forward (exp(-i...)) transform.  Returns (frequencies, real, imag)."""
    delta = (t_end - t_begin) / n_measurements
    x = np.zeros(zero_padding * len(j))
    x[:len(j)] = j
    n_input = len(x)
    Y = np.fft.rfft(x)  # FFTW r2c forward: sum x_n exp(-2 pi i k n / N)
    freq = 2 * np.arange(len(Y)) * np.pi / (n_input * delta)
    theta = freq * delta
    small = np.abs(theta) < 5.684341886080802e-14
    th = np.where(small, 1.0, theta)
    W = np.where(small, 1.0, 2.0 * (1.0 - np.cos(th)) / (th * th))
    a0 = np.where(small, -0.5 + 0j, (-(1.0 - np.cos(th)) + 1j * (th - np.sin(th))) / (th * th))
    corr = a0 * x[0] + np.conj(a0) * np.exp(1j * freq * (t_end - t_begin)) * x[n_input - 1]
    out = delta * W * np.exp(1j * freq * t_begin) * Y + corr
    out = delta * out
    return freq, out.real.copy(), out.imag.copy()


def welch_spectrum(j: np.ndarray, t_begin: float, t_end: float, n_measurements: int):
    """sources/hhg.cpp:264-275 for continuous lasers: the first N samples times
    the Welch window 1 - (i - N/2)^2 / (N/2)^2 (WelchWindow.hpp:11-13), numpy's
    real FFT (the r2c transform of FFT.cpp:8-42), frequencies i / Delta."""
    N = n_measurements
    i = np.arange(N)
    x = np.asarray(j[:N], dtype=np.float64) * (1.0 - (i - 0.5 * N) ** 2 / (0.25 * N * N))
    X = np.fft.rfft(x)
    delta = (t_end - t_begin) / N
    return np.arange(N // 2 + 1) / delta, X.real, X.imag
