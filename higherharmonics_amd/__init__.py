"""MI355X HigherHarmonics Bloch hot path: Python binding of libhhg_amd.so.

The library's C ABI (``include/hhg_amd.h``) is the drop-in boundary; this
module binds it with ctypes and mirrors the reference's host interface:

* :class:`TimeIntegrationConfig` — sources/HHG/TimeIntegrationConfig.hpp:7-15
* :class:`CosineLaser`, :class:`GaussLaser`, :class:`DoubleLaser` —
  sources/HHG/Laser/{CosineLaser,GaussLaser,DoubleLaser}.hpp, with
  ``laser_function(t)`` as Laser.hpp:32
* :class:`DiracSystem2D`, :class:`PiFlux2D`, :class:`Honeycomb` with
  ``compute_current_density(laser, time_config, rank, n_ranks)`` as
  sources/HHG/Systems/{DiracSystem,PiFlux,Honeycomb}.hpp
* :class:`Plan` — a reusable multi-job plan (the Dispatcher analogue)
* :func:`trapezoidal_spectrum` — sources/HHG/Fourier/TrapezoidalFFT.cpp:11-49

Every solve runs on the GPU through the HIP kernels; there is no CPU path.
A missing library raises :class:`HHGError` instead of falling back.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhhg_amd.so")

# ---- enums (include/hhg_amd.h) ----
SUCCESS, ERROR_INVALID_ARGUMENT, ERROR_HIP, ERROR_TOLERANCE, ERROR_OUT_OF_MEMORY, ERROR_UNSUPPORTED, \
    ERROR_NO_DEVICE, ERROR_NONFINITE = range(8)
MODEL_DIRAC2D, MODEL_PIFLUX2D, MODEL_HONEYCOMB, MODEL_DIRAC3D, MODEL_PIFLUX3D = 0, 1, 2, 3, 4
RELAX_NONE, RELAX_TO_INITIAL, RELAX_T2, RELAX_EQUILIBRIUM = 0, 1, 2, 3
PULSE_COS2, PULSE_GAUSS, PULSE_DOUBLE_GAUSS, PULSE_DOUBLE_COS2, PULSE_CONTINUOUS, PULSE_SPLINE = 0, 1, 2, 3, 4, 5
KGRID_GL_BOX, KGRID_GL_HEX, KGRID_EXPLICIT, KGRID_DIRAC_CYLINDER, KGRID_PIFLUX_CONTINUUM, KGRID_PIFLUX_UNIFORM = 0, 1, 2, 3, 4, 5
KGRID_PIFLUX_OCCUPATION, KGRID_DIRAC_GAUSS = 6, 7
METHOD_RKF78, METHOD_MAGNUS4, METHOD_MAGNUS_MOMENTS = 0, 1, 2
PROPOSAL_RESTART, PROPOSAL_CARRY, PROPOSAL_LEARNED = 0, 1, 2
INFER_FIRST, INFER_SCHEDULED = 0, 1

_STATUS_NAMES = {0: "SUCCESS", 1: "INVALID_ARGUMENT", 2: "HIP", 3: "TOLERANCE", 4: "OUT_OF_MEMORY",
                 5: "UNSUPPORTED", 6: "NO_DEVICE", 7: "NONFINITE"}


class HHGError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


# ---- ABI structs ----
class ModelDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("relax", ctypes.c_int32), ("v", ctypes.c_double),
                ("mass", ctypes.c_double), ("hopping", ctypes.c_double), ("theta", ctypes.c_double),
                ("fermi_energy", ctypes.c_double), ("beta", ctypes.c_double), ("relax_time", ctypes.c_double),
                ("relax_time_offdiag", ctypes.c_double), ("split_dispersion", ctypes.c_double),
                ("perp_current", ctypes.c_int32), ("pad", ctypes.c_int32)]


class PulseDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("pad", ctypes.c_int32), ("amp_a", ctypes.c_double),
                ("amp_b", ctypes.c_double), ("n_cycles", ctypes.c_double), ("sigma", ctypes.c_double),
                ("t0", ctypes.c_double), ("delay", ctypes.c_double), ("cep", ctypes.c_double),
                ("samples", ctypes.POINTER(ctypes.c_double)), ("n_samples", ctypes.c_int64),
                ("sample_t0", ctypes.c_double), ("sample_dt", ctypes.c_double), ("slope_begin", ctypes.c_double),
                ("slope_end", ctypes.c_double), ("t_cut", ctypes.c_double)]


class TimeDesc(ctypes.Structure):
    _fields_ = [("t_begin", ctypes.c_double), ("t_end", ctypes.c_double), ("n_measurements", ctypes.c_int32),
                ("n_subdivisions", ctypes.c_int32)]


class KGridDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("x0", ctypes.c_double), ("x1", ctypes.c_double), ("y0", ctypes.c_double), ("y1", ctypes.c_double),
                ("kx", ctypes.POINTER(ctypes.c_double)), ("ky", ctypes.POINTER(ctypes.c_double)),
                ("w", ctypes.POINTER(ctypes.c_double)), ("n_points", ctypes.c_int64),
                ("kz", ctypes.POINTER(ctypes.c_double))]


class JobDesc(ctypes.Structure):
    _fields_ = [("model", ModelDesc), ("pulse", PulseDesc), ("time", TimeDesc), ("kgrid", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


class SolverDesc(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("proposal", ctypes.c_int32), ("rtol", ctypes.c_double),
                ("atol", ctypes.c_double), ("max_fail", ctypes.c_int32), ("magnus_substeps", ctypes.c_int32),
                ("mlp", ctypes.c_void_p), ("mlp_infer_every", ctypes.c_int32), ("fallback_after", ctypes.c_int32),
                ("max_steps", ctypes.c_int32), ("infer_policy", ctypes.c_int32)]


class LaunchDesc(ctypes.Structure):
    _fields_ = [("waves_per_block", ctypes.c_int32), ("queue_chunk", ctypes.c_int32),
                ("persistent", ctypes.c_int32), ("group_by_cost", ctypes.c_int32),
                ("count_wave_attempts", ctypes.c_int32), ("block_reduce", ctypes.c_int32),
                ("separate_launches", ctypes.c_int32), ("loop_only", ctypes.c_int32)]


class PlanDesc(ctypes.Structure):
    _fields_ = [("n_kgrids", ctypes.c_int32), ("n_jobs", ctypes.c_int32), ("kgrids", ctypes.POINTER(KGridDesc)),
                ("jobs", ctypes.POINTER(JobDesc)), ("solver", SolverDesc), ("launch", LaunchDesc),
                ("rank", ctypes.c_int32), ("n_ranks", ctypes.c_int32), ("device", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("trajectories", ctypes.c_int64), ("attempts", ctypes.c_int64), ("rejects", ctypes.c_int64),
                ("failed_trajectories", ctypes.c_int64), ("nonfinite_trajectories", ctypes.c_int64),
                ("proposal_fallbacks", ctypes.c_int64), ("wave_attempts", ctypes.c_int64),
                ("accumulator_overflows", ctypes.c_int64), ("solve_ms", ctypes.c_double),
                ("reduce_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# Every symbol include/hhg_amd.h declares (checked by the ABI test).
EXPORTED_SYMBOLS = [
    "hhg_abi_version", "hhg_last_error_string", "hhg_plan_create", "hhg_plan_output_size", "hhg_plan_job_offsets",
    "hhg_plan_execute", "hhg_plan_execute_device", "hhg_plan_autotune", "hhg_plan_capture_graph",
    "hhg_plan_destroy", "hhg_trajectories", "hhg_gauss_legendre", "hhg_kgrid_size", "hhg_kgrid_build",
    "hhg_pulse_evaluate", "hhg_trapezoidal_spectrum", "hhg_plan_execute_accumulators", "hhg_accumulators_add",
    "hhg_accumulators_to_double", "hhg_plan_set_step_log", "hhg_plan_get_step_log", "hhg_mlp_param_count",
    "hhg_mlp_create", "hhg_mlp_destroy", "hhg_mlp_set_normalization", "hhg_mlp_get_params", "hhg_mlp_set_params",
    "hhg_mlp_forward", "hhg_mlp_forward_mfma", "hhg_mlp_gradients", "hhg_mlp_train", "hhg_welch_spectrum",
    "hhg_spline_coefficients", "hhg_reference_double_laser", "hhg_reference_experimental_laser_size",
    "hhg_reference_experimental_laser", "hhg_occupations", "hhg_plan_job_failures", "hhg_plan_last_stats",
    "hhg_trajectories_noalloc", "hhg_plan_solve_launches",
]

_lib: Optional[ctypes.CDLL] = None
_DP = ctypes.POINTER(ctypes.c_double)


def lib() -> ctypes.CDLL:
    """The loaded libhhg_amd.so.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HHGError(ERROR_NO_DEVICE, f"{LIB_PATH} missing: run `python -m higherharmonics_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    L.hhg_abi_version.restype = ctypes.c_int32
    L.hhg_last_error_string.restype = ctypes.c_char_p
    L.hhg_plan_create.argtypes = [ctypes.POINTER(PlanDesc), ctypes.POINTER(P)]
    L.hhg_plan_output_size.restype = ctypes.c_int64
    L.hhg_plan_output_size.argtypes = [P]
    L.hhg_plan_job_offsets.argtypes = [P, ctypes.POINTER(ctypes.c_int64)]
    L.hhg_plan_execute.argtypes = [P, _DP, ctypes.POINTER(Stats)]
    L.hhg_plan_execute_accumulators.argtypes = [P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(Stats)]
    L.hhg_plan_execute_device.argtypes = [P, _DP, P]
    L.hhg_plan_last_stats.argtypes = [P, ctypes.POINTER(Stats)]
    L.hhg_plan_autotune.argtypes = [P, ctypes.POINTER(LaunchDesc)]
    L.hhg_plan_solve_launches.argtypes = [P]
    L.hhg_plan_solve_launches.restype = ctypes.c_int32
    L.hhg_plan_capture_graph.argtypes = [P]
    L.hhg_plan_destroy.argtypes = [P]
    L.hhg_trajectories.argtypes = [P, ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, _DP]
    L.hhg_trajectories_noalloc.argtypes = [P, ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, _DP]
    L.hhg_occupations.argtypes = [P, ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, _DP]
    L.hhg_plan_job_failures.argtypes = [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.hhg_gauss_legendre.argtypes = [ctypes.c_int32, _DP, _DP]
    L.hhg_kgrid_size.restype = ctypes.c_int64
    L.hhg_kgrid_size.argtypes = [ctypes.POINTER(KGridDesc)]
    L.hhg_kgrid_build.argtypes = [ctypes.POINTER(KGridDesc), _DP, _DP, _DP, _DP]
    L.hhg_pulse_evaluate.argtypes = [ctypes.POINTER(PulseDesc), _DP, _DP, ctypes.c_int64]
    L.hhg_spline_coefficients.argtypes = [_DP, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          _DP]
    L.hhg_reference_double_laser.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_int32, _DP,
                                             ctypes.POINTER(PulseDesc), _DP]
    L.hhg_reference_experimental_laser_size.restype = ctypes.c_int64
    L.hhg_reference_experimental_laser_size.argtypes = [ctypes.c_double]
    L.hhg_reference_experimental_laser.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int32, _DP, _DP, ctypes.c_int32, _DP,
                                                   ctypes.c_int64, ctypes.POINTER(PulseDesc), _DP]
    L.hhg_trapezoidal_spectrum.argtypes = [_DP, ctypes.POINTER(TimeDesc), ctypes.c_int32, _DP, _DP, _DP]
    L.hhg_welch_spectrum.argtypes = [_DP, ctypes.POINTER(TimeDesc), _DP, _DP, _DP]
    L.hhg_accumulators_add.restype = None
    L.hhg_accumulators_add.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                       ctypes.c_int64]
    L.hhg_accumulators_to_double.restype = None
    L.hhg_accumulators_to_double.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_int64, _DP]
    FP = ctypes.POINTER(ctypes.c_float)
    L.hhg_plan_set_step_log.argtypes = [P, ctypes.c_int64, ctypes.c_int32]
    L.hhg_plan_get_step_log.argtypes = [P, FP, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    L.hhg_mlp_param_count.restype = ctypes.c_int64
    L.hhg_mlp_create.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(P)]
    L.hhg_mlp_destroy.argtypes = [P]
    L.hhg_mlp_set_normalization.argtypes = [P, FP, FP, FP, FP]
    L.hhg_mlp_get_params.argtypes = [P, FP]
    L.hhg_mlp_set_params.argtypes = [P, FP]
    L.hhg_mlp_forward.argtypes = [P, FP, ctypes.c_int64, FP]
    L.hhg_mlp_forward_mfma.argtypes = [P, FP, ctypes.c_int64, FP]
    L.hhg_mlp_gradients.argtypes = [P, FP, FP, ctypes.c_int64, FP, FP]
    L.hhg_mlp_train.argtypes = [P, FP, FP, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                ctypes.c_uint64, FP]
    for name in ("hhg_plan_set_step_log", "hhg_plan_get_step_log", "hhg_mlp_create", "hhg_mlp_destroy",
                 "hhg_mlp_set_normalization", "hhg_mlp_get_params", "hhg_mlp_set_params", "hhg_mlp_forward",
                 "hhg_mlp_forward_mfma", "hhg_mlp_gradients", "hhg_mlp_train"):
        getattr(L, name).restype = ctypes.c_int
    for name in ("hhg_plan_create", "hhg_plan_job_offsets", "hhg_plan_execute", "hhg_plan_execute_accumulators",
                 "hhg_plan_execute_device", "hhg_plan_last_stats", "hhg_plan_autotune", "hhg_plan_capture_graph", "hhg_plan_destroy",
                 "hhg_trajectories", "hhg_trajectories_noalloc", "hhg_occupations", "hhg_plan_job_failures", "hhg_gauss_legendre",
                 "hhg_kgrid_build", "hhg_pulse_evaluate",
                 "hhg_trapezoidal_spectrum", "hhg_welch_spectrum"):
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _check(status: int, allow=()) -> int:
    if status != SUCCESS and status not in allow:
        msg = lib().hhg_last_error_string()
        raise HHGError(status, msg.decode() if msg else "")
    return status


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_DP)


# --------------------------------------------------------------------------
# Reference-shaped host objects
# --------------------------------------------------------------------------

@dataclass
class TimeIntegrationConfig:
    """{t_begin, t_end, n_measurements, n_subdivisions} (TimeIntegrationConfig.hpp:7-15)."""
    t_begin: float
    t_end: float
    n_measurements: int
    n_subdivisions: int = 50

    def measure_every(self) -> float:
        return (self.t_end - self.t_begin) / self.n_measurements

    def dt(self) -> float:
        return self.measure_every() / self.n_subdivisions

    def times(self) -> np.ndarray:
        return self.t_begin + np.arange(self.n_measurements + 1) * self.measure_every()

    def desc(self) -> TimeDesc:
        return TimeDesc(self.t_begin, self.t_end, self.n_measurements, self.n_subdivisions)


class Laser:
    """A(t) = amplitude * envelope(t) * sin(t + cep) (Laser.hpp:32-54)."""
    t_begin: float
    t_end: float

    def desc(self) -> PulseDesc:  # pragma: no cover - overridden
        raise NotImplementedError

    def laser_function(self, t) -> np.ndarray:
        t = np.ascontiguousarray(np.atleast_1d(np.asarray(t, dtype=np.float64)))
        A = np.empty_like(t)
        d = self.desc()
        _check(lib().hhg_pulse_evaluate(ctypes.byref(d), _dp(t), _dp(A), len(t)))
        return A


class CosineLaser(Laser):
    """cos^2 envelope on (t0, t0 + 2 pi n) (CosineLaser.cpp:22-26, CosineLaser.hpp:13-15);
    ``begin_shift`` is the reference's window begin (pi * t0_offset)."""

    def __init__(self, amplitude: float, n_cycles: float, begin_shift: float = 0.0, cep: float = 0.0):
        self.amplitude, self.n_cycles, self.cep = amplitude, n_cycles, cep
        self.t_begin = begin_shift
        self.t_end = begin_shift + 2 * math.pi * n_cycles

    def desc(self) -> PulseDesc:
        return PulseDesc(PULSE_COS2, 0, self.amplitude, 0.0, self.n_cycles, 0.0, self.t_begin, 0.0, self.cep)


class ContinuousLaser(Laser):
    """Envelope 1 (ContinuousLaser.cpp:7-9): A(t) = amplitude sin(t + cep)."""

    def __init__(self, amplitude: float, cep: float = 0.0):
        self.amplitude, self.cep = amplitude, cep

    def desc(self) -> PulseDesc:
        return PulseDesc(PULSE_CONTINUOUS, 0, self.amplitude, 0.0, 0.0, 0.0, 0.0, 0.0, self.cep)


class GaussLaser(Laser):
    """Gaussian envelope exp(-(t-c)^2 / 2 sigma^2) (GaussLaser.cpp:26-28).  ``window``
    is the half width of the time window in units of sigma."""

    def __init__(self, amplitude: float, sigma: float, center: float = 0.0, cep: float = 0.0, window: float = 5.0):
        self.amplitude, self.sigma, self.center, self.cep = amplitude, sigma, center, cep
        self.t_begin = center - window * sigma
        self.t_end = center + window * sigma

    @staticmethod
    def reference_sigma(n_cycles: float) -> float:
        """sigma of the reference's GaussLaser(n_cycles) (GaussLaser.cpp:7)."""
        return math.pi * n_cycles / math.sqrt(2 * math.log(1 / 0.05))

    def desc(self) -> PulseDesc:
        return PulseDesc(PULSE_GAUSS, 0, self.amplitude, 0.0, 0.0, self.sigma, self.center, 0.0, self.cep)


class DoubleLaser(Laser):
    """A/B pair: A(t) + B(t - delay), B the same base pulse (DoubleLaser.hpp:63-67)
    with its own amplitude, evaluated analytically."""

    def __init__(self, base: Laser, amplitude_b: float, delay: float):
        self.base, self.amplitude_b, self.delay = base, amplitude_b, delay
        self.t_begin = min(base.t_begin, base.t_begin + delay)
        self.t_end = max(base.t_end, base.t_end + delay)

    def desc(self) -> PulseDesc:
        d = self.base.desc()
        d.kind = PULSE_DOUBLE_GAUSS if d.kind == PULSE_GAUSS else PULSE_DOUBLE_COS2
        d.amp_b = self.amplitude_b
        d.delay = self.delay
        return d


class SplineLaser(Laser):
    """A sampled pulse through the clamped cubic spline the reference builds with
    Boost's cardinal_cubic_b_spline (Laser.hpp:8-11, 32-37): samples[i] = A(t0 + i dt),
    end slopes (the reference passes 0, 0), A = 0 outside [t0, t_cut] (t_cut None:
    the last sample time).  The samples are read when a plan is created."""

    def __init__(self, samples, t0: float, dt: float, slope_begin: float = 0.0, slope_end: float = 0.0,
                 t_cut: Optional[float] = None):
        self.samples = np.ascontiguousarray(samples, dtype=np.float64)
        self.t0, self.dt, self.slope_begin, self.slope_end = t0, dt, slope_begin, slope_end
        self.t_begin = t0
        self.t_end = t0 + (len(self.samples) - 1) * dt if t_cut is None else t_cut
        self.t_cut = t_cut

    def desc(self) -> PulseDesc:
        d = PulseDesc(PULSE_SPLINE)
        d.samples, d.n_samples = _dp(self.samples), len(self.samples)
        d.sample_t0, d.sample_dt = self.t0, self.dt
        d.slope_begin, d.slope_end = self.slope_begin, self.slope_end
        d.t_cut = self.t_cut if self.t_cut is not None else self.t0
        return d


ACTIVE_BOTH, ACTIVE_A, ACTIVE_B = 0, 1, 2


def _spline_from_desc(samples: np.ndarray, d: PulseDesc, amp: float) -> SplineLaser:
    laser = SplineLaser(samples, d.sample_t0, d.sample_dt, d.slope_begin, d.slope_end, t_cut=d.t_cut)
    laser.momentum_amplitude = amp
    return laser


def reference_double_laser(gauss: bool, photon_energy: float, E0: float, model_ratio: float, n_cycles: float,
                           delay_ps: float, active: int = ACTIVE_BOTH) -> SplineLaser:
    """The reference's DoubleLaser<CosineLaser | GaussLaser> (DoubleLaser.hpp:8-74) in its
    units: photon energy in units of the experiment's 6.0914 meV, E0 in MV/cm, delay in ps."""
    samples = np.empty(1000)
    d, amp = PulseDesc(), ctypes.c_double()
    _check(lib().hhg_reference_double_laser(int(gauss), photon_energy, E0, model_ratio, n_cycles, delay_ps, active,
                                            _dp(samples), ctypes.byref(d), ctypes.byref(amp)))
    return _spline_from_desc(samples, d, amp.value)


def reference_experimental_laser(photon_energy: float, E0: float, model_ratio: float, shift_ps: float, E_A, E_B,
                                 active: int = ACTIVE_BOTH) -> SplineLaser:
    """The reference's ExperimentalLaser (ExperimentalLaser.cpp:8-257) from measured fields
    E_A, E_B [MV/cm] of its experiment layout (201 samples 0.03335555 ps apart)."""
    ea = np.ascontiguousarray(E_A, dtype=np.float64)
    eb = np.ascontiguousarray(E_B, dtype=np.float64)
    if len(ea) != len(eb):
        raise ValueError("E_A and E_B differ in length")
    n = lib().hhg_reference_experimental_laser_size(photon_energy)
    if n < 2:
        raise ValueError("photon_energy must be > 0")
    samples = np.empty(n)
    d, amp = PulseDesc(), ctypes.c_double()
    _check(lib().hhg_reference_experimental_laser(photon_energy, E0, model_ratio, shift_ps, active, _dp(ea), _dp(eb),
                                                  len(ea), _dp(samples), n, ctypes.byref(d), ctypes.byref(amp)))
    return _spline_from_desc(samples, d, amp.value)


def spline_coefficients(samples, dt: float, slope_begin: float = 0.0, slope_end: float = 0.0) -> np.ndarray:
    """(n - 1, 4) coefficients c of the clamped cubic spline, A(t_i + x dt) = c0 + x (c1 + x (c2 + x c3))."""
    y = np.ascontiguousarray(samples, dtype=np.float64)
    coef = np.empty(4 * max(len(y) - 1, 0))
    _check(lib().hhg_spline_coefficients(_dp(y), len(y), dt, slope_begin, slope_end, _dp(coef)))
    return coef.reshape(-1, 4)


@dataclass
class KGrid:
    kind: int
    nx: int = 0
    ny: int = 0
    x0: float = 0.0
    x1: float = 0.0
    y0: float = 0.0
    y1: float = 0.0
    points: Optional[tuple] = None  # (kx, ky, w) for explicit grids

    @staticmethod
    def box(nx, ny, x0, x1, y0, y1) -> "KGrid":
        return KGrid(KGRID_GL_BOX, nx, ny, x0, x1, y0, y1)

    @staticmethod
    def hexagonal(nx, ny) -> "KGrid":
        return KGrid(KGRID_GL_HEX, nx, ny)

    @staticmethod
    def dirac_cylinder(n_z, n_kappa, max_k) -> "KGrid":
        """(k_z, kappa) points of the 3D Dirac cone (kind KGRID_DIRAC_CYLINDER)."""
        return KGrid(KGRID_DIRAC_CYLINDER, n_z, n_kappa, 0.0, max_k)

    @staticmethod
    def dirac_gauss(n_z: int, n_kappa: int, max_k: float) -> "KGrid":
        """(k_z, kappa) of the 3D Dirac cone by Gauss-Legendre in both (kind KGRID_DIRAC_GAUSS)."""
        return KGrid(KGRID_DIRAC_GAUSS, n_z, n_kappa, 0.0, max_k)

    @staticmethod
    def piflux_continuum(z_range=32, n_coarse=4) -> "KGrid":
        """(x, y, z) points of the 3D pi-flux lattice (kind KGRID_PIFLUX_CONTINUUM)."""
        return KGrid(KGRID_PIFLUX_CONTINUUM, z_range, n_coarse)

    @staticmethod
    def piflux_uniform(n: int) -> "KGrid":
        """(x, y, +-z) pi/n of the 3D pi-flux split current, weight 1/n^3 (kind KGRID_PIFLUX_UNIFORM)."""
        return KGrid(KGRID_PIFLUX_UNIFORM, n)

    @staticmethod
    def piflux_occupation(n: int) -> "KGrid":
        """(x pi/n, pi/2, z pi/n), i = x n + z, of the 3D pi-flux occupation maps (kind KGRID_PIFLUX_OCCUPATION)."""
        return KGrid(KGRID_PIFLUX_OCCUPATION, n)

    @staticmethod
    def explicit(kx, ky, w, kz=None) -> "KGrid":
        arrs = tuple(np.ascontiguousarray(a, dtype=np.float64) for a in (kx, ky, w))
        if kz is not None:
            arrs = arrs + (np.ascontiguousarray(kz, dtype=np.float64),)
        return KGrid(KGRID_EXPLICIT, points=arrs)

    def desc(self) -> KGridDesc:
        d = KGridDesc(self.kind, self.nx, self.ny, 0, self.x0, self.x1, self.y0, self.y1, None, None, None, 0, None)
        if self.kind == KGRID_EXPLICIT:
            kx, ky, w = self.points[:3]
            d.kx, d.ky, d.w, d.n_points = _dp(kx), _dp(ky), _dp(w), len(kx)
            if len(self.points) > 3:
                d.kz = _dp(self.points[3])
        return d

    def build3(self):
        """(kx, ky, kz, w); kz = 0 for 2D grids."""
        d = self.desc()
        n = lib().hhg_kgrid_size(ctypes.byref(d))
        kx, ky, kz, w = np.empty(n), np.empty(n), np.empty(n), np.empty(n)
        _check(lib().hhg_kgrid_build(ctypes.byref(d), _dp(kx), _dp(ky), _dp(kz), _dp(w)))
        return kx, ky, kz, w

    def build(self):
        kx, ky, _, w = self.build3()
        return kx, ky, w


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class MLP:
    """The step-proposal network on one device (8 -> 64 -> 64 -> 64 -> 2,
    ReLU; BF16 weights and activations, FP32 accumulation, FP32 masters)."""
    N_IN, N_OUT = 8, 2

    def __init__(self, device: int = 0, seed: int = 1):
        h = ctypes.c_void_p()
        _check(lib().hhg_mlp_create(device, seed, ctypes.byref(h)))
        self._h = h
        self.n_params = lib().hhg_mlp_param_count()

    @property
    def handle(self):
        return self._h

    def set_normalization(self, in_mean, in_std, out_mean, out_std) -> None:
        arr = [np.ascontiguousarray(a, dtype=np.float32) for a in (in_mean, in_std, out_mean, out_std)]
        self._norm = arr
        _check(lib().hhg_mlp_set_normalization(self._h, *[_fp(a) for a in arr]))

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        _check(lib().hhg_mlp_get_params(self._h, _fp(out)))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().hhg_mlp_set_params(self._h, _fp(p)))

    def forward(self, x, mfma: bool = False) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((len(x), self.N_OUT), dtype=np.float32)
        fn = lib().hhg_mlp_forward_mfma if mfma else lib().hhg_mlp_forward
        _check(fn(self._h, _fp(x), len(x), _fp(y)))
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.n_params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _check(lib().hhg_mlp_gradients(self._h, _fp(x), _fp(t), len(x), _fp(g), _fp(loss)))
        return g, float(loss[0])

    def train(self, x, t, epochs=4, batch=1024, lr=2e-3, seed=7) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        _check(lib().hhg_mlp_train(self._h, _fp(x), _fp(t), len(x), epochs, batch, lr, seed, _fp(losses)))
        return losses

    def save(self, path) -> None:
        """Parameters and normalisation to an .npz file (plain arrays)."""
        norm = getattr(self, "_norm", None)
        if norm is None:
            raise ValueError("MLP.save: normalisation not set")
        np.savez(path, params=self.params(), in_mean=norm[0], in_std=norm[1], out_mean=norm[2], out_std=norm[3])

    @classmethod
    def load(cls, path, device: int = 0) -> "MLP":
        with np.load(path, allow_pickle=False) as z:
            net = cls(device=device)
            net.set_normalization(z["in_mean"], z["in_std"], z["out_mean"], z["out_std"])
            net.set_params(z["params"])
        return net

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().hhg_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Solver:
    method: int = METHOD_RKF78
    rtol: float = 1e-10
    atol: float = 1e-10
    proposal: int = PROPOSAL_RESTART
    max_fail: int = 500
    magnus_substeps: int = 1
    mlp: Optional["MLP"] = None
    infer_every: int = 1
    fallback_after: int = 0
    max_steps: int = 0
    infer_policy: int = INFER_FIRST

    def desc(self) -> SolverDesc:
        return SolverDesc(self.method, self.proposal, self.rtol, self.atol, self.max_fail, self.magnus_substeps,
                          self.mlp.handle if self.mlp is not None else None, self.infer_every, self.fallback_after,
                          self.max_steps, self.infer_policy)


@dataclass
class Job:
    model: ModelDesc
    laser: Laser
    time_config: TimeIntegrationConfig
    kgrid: int = 0


class Plan:
    """A compiled set of jobs on one device (the reference Dispatcher's role).

    ``execute()`` returns one j(t) array per job."""

    def __init__(self, jobs: Sequence[Job], kgrids: Sequence[KGrid], solver: Solver = Solver(),
                 rank: int = 0, n_ranks: int = 1, device: int = 0, waves_per_block: int = 0,
                 queue_chunk: int = 0, persistent: int = 1, group_by_cost: int = 0, count_wave_attempts: int = 0,
                 block_reduce: int = 0, separate_launches: int = 0, loop_only: int = 0):
        L = lib()
        self._solver = solver  # keeps the network alive as long as the plan
        self._kgrids = list(kgrids)  # keep explicit arrays alive
        kg = (KGridDesc * len(kgrids))(*[g.desc() for g in kgrids])
        jd = (JobDesc * len(jobs))(*[JobDesc(j.model, j.laser.desc(), j.time_config.desc(), j.kgrid, 0)
                                     for j in jobs])
        d = PlanDesc(len(kgrids), len(jobs), kg, jd, solver.desc(),
                     LaunchDesc(waves_per_block, queue_chunk, persistent, group_by_cost, count_wave_attempts,
                                block_reduce, separate_launches, loop_only),
                     rank, n_ranks, device, 0)
        h = ctypes.c_void_p()
        _check(L.hhg_plan_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.n_jobs = len(jobs)
        self.n_out = L.hhg_plan_output_size(h)
        off = (ctypes.c_int64 * (len(jobs) + 1))()
        _check(L.hhg_plan_job_offsets(h, off))
        self.offsets = list(off)
        self._n_samples0 = jobs[0].time_config.n_measurements + 1 if jobs else 0  # job 0's samples per k-point
        self.stats = Stats()

    def execute_flat(self, allow_failures: bool = False) -> np.ndarray:
        j = np.empty(self.n_out)
        allow = (ERROR_TOLERANCE, ERROR_NONFINITE) if allow_failures else ()
        self.status = _check(lib().hhg_plan_execute(self._h, _dp(j), ctypes.byref(self.stats)), allow)
        return j

    def execute(self, allow_failures: bool = False):
        j = self.execute_flat(allow_failures)
        return [j[self.offsets[i]:self.offsets[i + 1]] for i in range(self.n_jobs)]

    def execute_device(self, j_dev_ptr: int, stream: int = 0) -> None:
        """Enqueues a solve that leaves j in device memory at j_dev_ptr (n_out doubles)."""
        _check(lib().hhg_plan_execute_device(self._h, ctypes.cast(j_dev_ptr, _DP), stream or None))

    def last_stats(self, allow_failures: bool = False) -> "Stats":
        """Waits for the last execution and returns its statistics (raises on its failures)."""
        st = Stats()
        allow = (ERROR_TOLERANCE, ERROR_NONFINITE) if allow_failures else ()
        self.status = _check(lib().hhg_plan_last_stats(self._h, ctypes.byref(st)), allow)
        return st

    def execute_accumulators(self) -> np.ndarray:
        acc = np.empty(2 * self.n_out, dtype=np.uint64)
        _check(lib().hhg_plan_execute_accumulators(
            self._h, acc.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(self.stats)))
        return acc

    def set_step_log(self, capacity: int, stride: int = 16) -> None:
        _check(lib().hhg_plan_set_step_log(self._h, capacity, stride))
        self._log_cap = capacity

    def step_log(self) -> np.ndarray:
        """Records of the last execute: [n][10] = 8 features, log(h_opt/Delta) (the step the
        error model of the interval's first accepted step allows), log(attempts)."""
        cap = getattr(self, "_log_cap", 0)
        rec = np.empty((max(cap, 1), 10), dtype=np.float32)
        n = ctypes.c_int64()
        _check(lib().hhg_plan_get_step_log(self._h, _fp(rec), cap, ctypes.byref(n)))
        return rec[:n.value].copy()

    def autotune(self) -> dict:
        ld = LaunchDesc()
        _check(lib().hhg_plan_autotune(self._h, ctypes.byref(ld)))
        return {"waves_per_block": ld.waves_per_block, "queue_chunk": ld.queue_chunk}

    @property
    def solve_launches(self) -> int:
        """Solve-kernel launches per execution (1 for a fused pair of groups)."""
        return int(lib().hhg_plan_solve_launches(self._h))

    def capture_graph(self) -> None:
        _check(lib().hhg_plan_capture_graph(self._h))

    def trajectories(self, k_index) -> np.ndarray:
        """Bloch vectors [n_sel][N+1][3] of job 0 for the given k indices."""
        sel = np.ascontiguousarray(k_index, dtype=np.int64)
        n = self._n_samples0
        out = np.empty((len(sel), n, 3))
        _check(lib().hhg_trajectories(self._h, sel.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(sel),
                                      _dp(out)))
        return out

    def trajectories_noalloc(self, k_index) -> np.ndarray:
        """As trajectories(), through the entry point that allocates no device memory."""
        sel = np.ascontiguousarray(k_index, dtype=np.int64)
        out = np.empty((len(sel), self._n_samples0, 3))
        _check(lib().hhg_trajectories_noalloc(self._h, sel.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                              len(sel), _dp(out)))
        return out

    def job_failures(self):
        """(failed, nonfinite) trajectory counts per job of the last execution."""
        f = np.zeros(self.n_jobs, dtype=np.int64)
        n = np.zeros(self.n_jobs, dtype=np.int64)
        _check(lib().hhg_plan_job_failures(self._h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                           n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return f, n

    def occupations(self, k_index) -> np.ndarray:
        """Band occupations [n_sel][N+1][2] = (lower, upper) of job 0 (3D pi-flux
        lattice, equilibrium relaxation, RKF78) for the given k indices."""
        sel = np.ascontiguousarray(k_index, dtype=np.int64)
        out = np.empty((len(sel), self._n_samples0, 2))
        _check(lib().hhg_occupations(self._h, sel.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(sel),
                                     _dp(out)))
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().hhg_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def accumulators_to_double(acc: np.ndarray) -> np.ndarray:
    n = len(acc) // 2
    j = np.empty(n)
    lib().hhg_accumulators_to_double(np.ascontiguousarray(acc).ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n,
                                     _dp(j))
    return j


def accumulators_add(total: np.ndarray, part: np.ndarray) -> None:
    lib().hhg_accumulators_add(total.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                               np.ascontiguousarray(part).ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                               len(total) // 2)


# --------------------------------------------------------------------------
# Systems with the reference's compute_current_density shape
# --------------------------------------------------------------------------

class _System:
    kind = MODEL_DIRAC2D

    def __init__(self, fermi_energy=0.0, beta=0.0, relax=RELAX_NONE, relax_time=0.0, theta=0.0, solver=None,
                 kgrid: Optional[KGrid] = None):
        self.fermi_energy, self.beta, self.relax, self.relax_time, self.theta = (fermi_energy, beta, relax,
                                                                                 relax_time, theta)
        self.solver = solver or Solver()
        self.kgrid = kgrid or self.default_grid()

    def default_grid(self) -> KGrid:  # pragma: no cover - overridden
        raise NotImplementedError

    def model_desc(self) -> ModelDesc:  # pragma: no cover - overridden
        raise NotImplementedError

    def compute_current_density(self, laser: Laser, time_config: TimeIntegrationConfig, rank: int = 0,
                                n_ranks: int = 1) -> np.ndarray:
        """j(t_i), i = 0..N (sources/HHG/Systems/*::compute_current_density)."""
        with Plan([Job(self.model_desc(), laser, time_config, 0)], [self.kgrid], self.solver, rank, n_ranks) as p:
            out = p.execute()[0]
            self.last_stats = p.stats.as_dict()
        if getattr(self, "perp_current", False):
            # j, then j_perp; warn where j_perp does not vanish (Honeycomb.cpp:328-334)
            n = len(out) // 2
            out, self.last_j_perp = out[:n], out[n:]
            bad = (np.abs(self.last_j_perp) > 1e-6 * np.abs(out)) & (np.abs(self.last_j_perp) > 1e-6)
            if bad.any():
                import warnings
                i = int(np.argmax(np.where(bad, np.abs(self.last_j_perp), 0.0)))
                warnings.warn(f"j_perp = {self.last_j_perp[i]} at sample {i} (j_par = {out[i]}); "
                              f"{int(bad.sum())} samples exceed 1e-6")
        return out


class DiracSystem2D(_System):
    """Massive 2D Dirac cone d = (v qx, v qy, Delta), q = k - A e_theta."""
    kind = MODEL_DIRAC2D

    def __init__(self, v=1.0, mass=0.1, k_max=3.0, nx=128, ny=64, **kw):
        self.v, self.mass, self.k_max, self.nx, self.ny = v, mass, k_max, nx, ny
        super().__init__(**kw)

    def default_grid(self) -> KGrid:
        return KGrid.box(self.nx, self.ny, -self.k_max, self.k_max, -self.k_max, self.k_max)

    def model_desc(self) -> ModelDesc:
        return ModelDesc(MODEL_DIRAC2D, self.relax, self.v, self.mass, 0.0, self.theta, self.fermi_energy,
                         self.beta, self.relax_time)


class PiFlux2D(_System):
    """PiFlux square lattice d = 2t (cos(kx - A), cos ky, 0) over the full BZ."""
    kind = MODEL_PIFLUX2D

    def __init__(self, hopping=1.0, nx=512, ny=512, **kw):
        self.hopping, self.nx, self.ny = hopping, nx, ny
        super().__init__(**kw)

    def default_grid(self) -> KGrid:
        return KGrid.box(self.nx, self.ny, -math.pi, math.pi, -math.pi, math.pi)

    def model_desc(self) -> ModelDesc:
        return ModelDesc(MODEL_PIFLUX2D, self.relax, 0.0, 0.0, self.hopping, self.theta, self.fermi_energy,
                         self.beta, self.relax_time)


class Honeycomb(_System):
    """Honeycomb nearest-neighbour model over the hexagonal BZ (Honeycomb.cpp:225-336).
    perp_current: the job's output is j, then the reference's consistency
    check j_perp (Honeycomb.cpp:328-334), N + 1 samples each."""
    kind = MODEL_HONEYCOMB

    def __init__(self, hopping=1.0, nx=128, ny=128, perp_current=False, **kw):
        self.hopping, self.nx, self.ny, self.perp_current = hopping, nx, ny, perp_current
        super().__init__(**kw)

    def default_grid(self) -> KGrid:
        return KGrid.hexagonal(self.nx, self.ny)

    def model_desc(self) -> ModelDesc:
        return ModelDesc(MODEL_HONEYCOMB, self.relax, 0.0, 0.0, self.hopping, self.theta, self.fermi_energy,
                         self.beta, self.relax_time, 0.0, 0.0, int(bool(self.perp_current)))


class DiracSystem3D(_System):
    """The reference's 3D Dirac cone in its band frame (DiracSystem.cpp:138-174),
    field along z, on the (k_z, kappa) cylinder of DiracSystem.cpp:259-306;
    lengths in hbar omega_L / v_F."""
    kind = MODEL_DIRAC3D

    def __init__(self, max_k=10.0, n_z=200, n_kappa=64, **kw):
        self.max_k, self.n_z, self.n_kappa = max_k, n_z, n_kappa
        super().__init__(**kw)

    def default_grid(self) -> KGrid:
        return KGrid.dirac_cylinder(self.n_z, self.n_kappa, self.max_k)

    def model_desc(self) -> ModelDesc:
        return ModelDesc(MODEL_DIRAC3D, self.relax, 0.0, 0.0, 0.0, 0.0, self.fermi_energy, self.beta,
                         self.relax_time)


class PiFlux3D(_System):
    """The reference's 3D pi-flux lattice d = 2t (cos x, cos y, cos(z - A))
    (PiFlux.cpp:73-111) on its continuum-limit quadrature (PiFlux.cpp:665-863);
    relax=RELAX_EQUILIBRIUM relaxes toward the instantaneous equilibrium with
    the diagonal time relax_time and the off-diagonal time relax_time_offdiag
    (PiFlux.cpp:113-220)."""
    kind = MODEL_PIFLUX3D

    def __init__(self, hopping=1.0, z_range=32, n_coarse=4, relax_time_offdiag=0.0, **kw):
        self.hopping, self.z_range, self.n_coarse = hopping, z_range, n_coarse
        self.relax_time_offdiag = relax_time_offdiag
        super().__init__(**kw)

    def default_grid(self) -> KGrid:
        return KGrid.piflux_continuum(self.z_range, self.n_coarse)

    def model_desc(self) -> ModelDesc:
        return ModelDesc(MODEL_PIFLUX3D, self.relax, 0.0, 0.0, self.hopping, 0.0, self.fermi_energy, self.beta,
                         self.relax_time, self.relax_time_offdiag)


# --------------------------------------------------------------------------
# Host utilities
# --------------------------------------------------------------------------

def gauss_legendre(n: int):
    x, w = np.empty(n), np.empty(n)
    _check(lib().hhg_gauss_legendre(n, _dp(x), _dp(w)))
    return x, w


def trapezoidal_spectrum(j: np.ndarray, time_config: TimeIntegrationConfig, zero_padding: int = 8):
    """(frequencies, real, imag) of the end-corrected Fourier integral of j."""
    j = np.ascontiguousarray(j, dtype=np.float64)
    n = (zero_padding * (time_config.n_measurements + 1)) // 2 + 1
    f, re, im = np.empty(n), np.empty(n), np.empty(n)
    td = time_config.desc()
    _check(lib().hhg_trapezoidal_spectrum(_dp(j), ctypes.byref(td), zero_padding, _dp(f), _dp(re), _dp(im)))
    return f, re, im


def welch_spectrum(j: np.ndarray, time_config: TimeIntegrationConfig):
    """(frequencies, real, imag) of a continuous-laser run: Welch window on the
    first N samples and a real-to-complex DFT (sources/hhg.cpp:264-275)."""
    j = np.ascontiguousarray(j, dtype=np.float64)
    n = time_config.n_measurements // 2 + 1
    f, re, im = np.empty(n), np.empty(n), np.empty(n)
    td = time_config.desc()
    _check(lib().hhg_welch_spectrum(_dp(j), ctypes.byref(td), _dp(f), _dp(re), _dp(im)))
    return f, re, im
