"""The C ABI library on the CPU: it loads, exports every declared symbol, and
its host utilities (Gauss-Legendre nodes, BZ grids, pulses, spectrum, input
validation) agree with the oracle and the reference's tables.  No kernel runs."""
import ctypes
import json
import math
import os
import re

import numpy as np
import pytest

import oracle as O
from tests.problems import SMALL_CASES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol(hhg):
    header = open(os.path.join(ROOT, "include", "hhg_amd.h")).read()
    declared = set(re.findall(r"\b(hhg_[a-z0-9_]+)\s*\(", header))
    assert declared == set(hhg.EXPORTED_SYMBOLS)
    L = hhg.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert L.hhg_abi_version() == 5


def test_gauss_legendre_matches_reference_tables(hhg):
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "thread_gauss_subset.json")))
    for n_str, tab in fx["orders"].items():
        n = int(n_str)
        x, w = hhg.gauss_legendre(n)
        ref_x = np.array([float(v) for v in tab["abscissa"]])
        ref_w = np.array([float(v) for v in tab["weights"]])
        # container<N> holds the N/2 positive nodes in ascending order
        assert np.max(np.abs(x[n // 2:] - ref_x)) < 4e-16
        assert np.max(np.abs(w[n // 2:] - ref_w) / ref_w) < 1e-14
        assert np.allclose(x[:n // 2], -ref_x[::-1], rtol=0, atol=4e-16)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 30, 64, 255, 1024])
def test_gauss_legendre_exactness_and_numpy(hhg, n):
    x, w = hhg.gauss_legendre(n)
    xn, wn = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xn)) < 1e-14
    # numpy's weights lose up to ~1e-9 relative near the ends at large n
    assert np.max(np.abs(w - wn) / wn) < 1e-8
    ox, ow = O.gauss_legendre(n)
    assert np.max(np.abs(x - ox)) < 4e-16 and np.max(np.abs(w - ow) / ow) < 1e-13
    assert abs(w.sum() - 2.0) < 1e-13
    # exact for polynomials of degree 2n-1: int_{-1}^{1} x^{2m} = 2/(2m+1)
    for m in range(0, n):
        assert abs(np.dot(w, x ** (2 * m)) - 2.0 / (2 * m + 1)) < 1e-13


def test_kgrids_match_oracle_and_areas(hhg):
    kx, ky, w = hhg.KGrid.box(32, 16, -3, 3, -2, 2).build()
    okx, oky, ow = O.gl_box(32, 16, -3, 3, -2, 2)
    assert np.allclose(kx, okx, atol=1e-15) and np.allclose(ky, oky, atol=1e-15)
    assert np.allclose(w, ow, rtol=1e-14)
    assert abs(w.sum() - 24.0) < 1e-12
    hx, hy, hw = hhg.KGrid.hexagonal(24, 24).build()
    ox, oy, owh = O.gl_hex(24, 24)
    assert np.allclose(hx, ox, atol=1e-15) and np.allclose(hy, oy, atol=1e-15) and np.allclose(hw, owh, rtol=1e-14)
    # hexagonal BZ of the unit honeycomb lattice: area 8 pi^2 / (3 sqrt 3).  The
    # reference's x-GL crosses the |x| kink of y_max at x = 0 (Honeycomb.cpp:231),
    # so the area converges as 1/n^2 rather than exponentially.
    area = 8 * math.pi ** 2 / (3 * math.sqrt(3))
    errs = [abs(hhg.KGrid.hexagonal(n, n).build()[2].sum() - area) for n in (24, 48, 96)]
    assert errs[0] < 1e-2 and 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5


def test_dirac_cylinder_grid_matches_oracle(hhg):
    """(k_z, kappa) points and weights of the 3D Dirac cone's grid, product vs oracle."""
    for n_z, n_kappa, M in [(16, 8, 3.0), (33, 5, 1.5)]:
        kx, ky, w = hhg.KGrid.dirac_cylinder(n_z, n_kappa, M).build()
        okx, oky, ow = O.dirac_cylinder(n_z, n_kappa, M)
        assert len(kx) == (n_z - 2) * n_kappa == len(okx)
        assert np.array_equal(kx, okx) and np.allclose(ky, oky, rtol=1e-15, atol=0) and np.allclose(w, ow, rtol=1e-14)


def test_piflux_continuum_grid_matches_oracle(hhg):
    """(x, y, z) points and weights of the 3D pi-flux continuum-limit grid,
    product vs oracle, including the reference's default 32 x 4; the third
    coordinate is zero for 2D grids and the explicit grid carries one."""
    for z_range, n_coarse in [(4, 1), (6, 2), (32, 4)]:
        kx, ky, kz, w = hhg.KGrid.piflux_continuum(z_range, n_coarse).build3()
        okx, oky, okz, ow = O.piflux_continuum(z_range, n_coarse)
        assert len(kx) == len(okx) == 2 * z_range * (2 * 64 * n_coarse ** 2 + 20 * n_coarse ** 2)
        for a, b in ((kx, okx), (ky, oky), (kz, okz), (w, ow)):
            assert np.allclose(a, b, rtol=1e-14, atol=0)
    assert np.all(hhg.KGrid.box(4, 3, -1.0, 1.0, -1.0, 1.0).build3()[2] == 0.0)
    e = hhg.KGrid.explicit([0.1, 0.2], [0.3, 0.4], [1.0, 2.0], kz=[0.5, -0.5]).build3()
    assert np.array_equal(e[2], [0.5, -0.5])


def test_continuous_pulse_matches_oracle(hhg):
    t = np.linspace(-20.0, 20.0, 401)
    A = hhg.ContinuousLaser(0.9, cep=0.2).laser_function(t)
    assert np.max(np.abs(A - O.pulse_A(O.pulse(O.PULSE_CONTINUOUS, amp_a=0.9, cep=0.2), t))) < 1e-15


@pytest.mark.parametrize("pr", SMALL_CASES, ids=lambda p: p.name)
def test_pulse_matches_oracle(hhg, pr):
    t = np.linspace(pr.time[0] - 1.0, pr.time[1] + 1.0, 997)
    A = pr.h_laser(hhg).laser_function(t)
    A0 = O.pulse_A(pr.o_pulse(), t)
    assert np.max(np.abs(A - A0)) < 1e-15


def test_cos2_pulse_is_the_configured_cos2_form(hhg):
    """The reference's cos^2 window (CosineLaser.cpp:22-26) with begin shift -T/2
    is A0 cos^2(pi t/T) sin t on [-T/2, T/2], T = 2 pi n (configuration 1)."""
    T = 16 * math.pi
    las = hhg.CosineLaser(0.5, 8.0, -T / 2)
    t = np.linspace(-T / 2 + 1e-9, T / 2 - 1e-9, 4001)
    assert np.max(np.abs(las.laser_function(t) - 0.5 * np.cos(np.pi * t / T) ** 2 * np.sin(t))) < 1e-15
    assert np.all(las.laser_function(np.array([-T / 2 - 1, T / 2 + 1])) == 0.0)


def test_welch_spectrum_matches_oracle_and_direct_sum(hhg):
    """Continuous-laser spectrum: Welch window and r2c DFT, against numpy's FFT
    and a direct sum on a small N."""
    rng = np.random.default_rng(9)
    for N in (64, 250):
        tc = hhg.TimeIntegrationConfig(-N * 0.05, N * 0.05, N, 50)
        j = rng.standard_normal(N + 1)
        f, re, im = hhg.welch_spectrum(j, tc)
        of, ore, oim = O.welch_spectrum(j, tc.t_begin, tc.t_end, N)
        assert np.allclose(f, of, rtol=1e-15) and np.allclose(re, ore, atol=1e-12) and np.allclose(im, oim, atol=1e-12)
        i = np.arange(N)
        x = j[:N] * (1 - (i - N / 2) ** 2 / (N / 2) ** 2)
        k = np.arange(N // 2 + 1)
        direct = (x[None, :] * np.exp(-2j * np.pi * k[:, None] * i[None, :] / N)).sum(1)
        assert np.allclose(re + 1j * im, direct, atol=1e-11)


def test_spectrum_matches_oracle(hhg):
    rng = np.random.default_rng(7)
    tc = hhg.TimeIntegrationConfig(-3.0, 5.0, 255, 50)
    j = rng.standard_normal(256)
    f, re, im = hhg.trapezoidal_spectrum(j, tc)
    f0, re0, im0 = O.trapezoidal_spectrum(j, -3.0, 5.0, 255)
    assert np.allclose(f, f0, rtol=1e-15)
    scale = np.max(np.hypot(re0, im0))
    assert np.max(np.abs(re - re0)) < 1e-13 * scale
    assert np.max(np.abs(im - im0)) < 1e-13 * scale


@pytest.mark.parametrize("mutate,status", [
    (lambda j, s: setattr(j.model, "kind", 7), 1),
    (lambda j, s: setattr(j.time, "n_measurements", 0), 1),
    (lambda j, s: setattr(j.time, "n_measurements", 10), 5),        # Delta > 0.25
    (lambda j, s: setattr(j.model, "theta", 0.3), 5),               # lattice with theta != 0
    (lambda j, s: setattr(s, "rtol", -1.0), 1),
    (lambda j, s: (setattr(s, "method", 1), setattr(j.model, "relax", 2), setattr(j.model, "relax_time", 3.0)), 5),
    (lambda j, s: setattr(j, "kgrid", 3), 1),
    # non-finite inputs are refused before any device work
    (lambda j, s: setattr(j.pulse, "amp_a", float("nan")), 1),
    (lambda j, s: setattr(j.pulse, "cep", float("inf")), 1),
    (lambda j, s: setattr(j.pulse, "delay", float("nan")), 1),
    (lambda j, s: setattr(j.model, "hopping", float("nan")), 1),
    (lambda j, s: setattr(j.time, "t_end", float("inf")), 1),
    (lambda j, s: setattr(s, "rtol", float("inf")), 1),
])
def test_plan_rejects_invalid_descriptions(hhg, mutate, status):
    from tests.problems import HONEYCOMB_SMALL as pr
    jd = hhg.JobDesc(pr.h_model(hhg), pr.h_laser(hhg).desc(), pr.h_time(hhg).desc(), 0, 0)
    sv = pr.h_solver(hhg).desc()
    mutate(jd, sv)
    kg = (hhg.KGridDesc * 1)(pr.h_grid(hhg).desc())
    jobs = (hhg.JobDesc * 1)(jd)
    d = hhg.PlanDesc(1, 1, kg, jobs, sv, hhg.LaunchDesc(0, 0, 1, 0), 0, 1, 0, 0)
    h = ctypes.c_void_p()
    st = hhg.lib().hhg_plan_create(ctypes.byref(d), ctypes.byref(h))
    assert st == status
    assert hhg.lib().hhg_last_error_string()


def test_plan_rejects_nonfinite_spline_sample(hhg):
    """A NaN among a sampled pulse's samples is refused at plan creation
    (INVALID_ARGUMENT naming the sample), not carried into the trajectories."""
    from tests.problems import HONEYCOMB_SMALL as pr
    samples = np.sin(np.linspace(0.0, 6.0, 200)) * 0.3
    samples[77] = np.nan
    laser = hhg.SplineLaser(samples, 0.0, 0.05)
    with pytest.raises(hhg.HHGError) as ei:
        hhg.Plan([hhg.Job(pr.h_model(hhg), laser, hhg.TimeIntegrationConfig(0.0, 9.0, 64, 50), 0)], [pr.h_grid(hhg)],
                 hhg.Solver())
    assert ei.value.status == hhg.ERROR_INVALID_ARGUMENT
    assert "77" in str(ei.value)


def test_gauss_legendre_1024_against_mpmath(hhg):
    """N = 1024 is beyond the reference tables (they stop at 960)."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    n = 1024
    x, w = hhg.gauss_legendre(n)
    for i in (0, 1, 100, 511):
        z = mp.mpf(x[i])
        for _ in range(4):
            p = mp.legendre(n, z)
            dp = n * (mp.legendre(n - 1, z) - z * p) / (1 - z * z)
            z = z - p / dp
        dp = n * (mp.legendre(n - 1, z) - z * mp.legendre(n, z)) / (1 - z * z)
        wt = 2 / ((1 - z * z) * dp * dp)
        assert abs(float(mp.mpf(x[i]) - z)) < 1e-16
        assert abs(float((mp.mpf(w[i]) - wt) / wt)) < 5e-14


def test_cpp_driver_rejects_unknown_laser(hhg, tmp_path):
    import subprocess
    cfg = tmp_path / "bad.cfg"
    cfg.write_text("system_type Dirac2D\nlaser_type sawtooth\n")
    exe = os.path.join(os.path.dirname(hhg.LIB_PATH), "hhg_run")
    r = subprocess.run([exe, str(cfg)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "not recognized" in r.stderr


@pytest.mark.parametrize("kind,relax,proposal,method,perp", [
    (4, 3, 1, 0, 0),  # equilibrium relaxation with the carried proposal
    (4, 2, 0, 0, 0),  # the 3D pi-flux lattice has no T2 dephasing
    (0, 3, 0, 0, 0),  # equilibrium relaxation of another model
    (4, 0, 0, 2, 0),  # the moment Magnus step is the 3D Dirac cone's
    (0, 0, 0, 0, 1),  # j_perp is the honeycomb's
    (2, 0, 0, 1, 1),  # j_perp needs RKF78
    (2, 0, 0, 0, 2),  # perp_current is 0 or 1
], ids=["equilibrium-carry", "piflux3d-t2", "dirac2d-equilibrium", "piflux3d-moments", "dirac2d-perp",
        "honeycomb-magnus-perp", "honeycomb-perp-2"])
def test_plan_rejects_unsupported_combinations(hhg, kind, relax, proposal, method, perp):
    """Validation runs on the host before any device call and names the reason."""
    m = hhg.ModelDesc(kind, relax, 1.0, 0.1, 0.6, 0.0, 0.0, 0.0, 10.0, 4.0, 0.0, perp)
    job = hhg.Job(m, hhg.ContinuousLaser(0.5), hhg.TimeIntegrationConfig(0.0, 6.0, 64, 50), 0)
    grid = hhg.KGrid.piflux_continuum(2, 1) if kind == 4 else hhg.KGrid.box(4, 4, -1.0, 1.0, -1.0, 1.0)
    with pytest.raises(hhg.HHGError) as ei:
        hhg.Plan([job], [grid], hhg.Solver(method=method, proposal=proposal))
    assert ei.value.status == hhg.ERROR_UNSUPPORTED


def test_spline_pulse_matches_oracle(hhg):
    """The product's clamped spline (second-derivative form) equals the
    oracle's (slope form) coefficient for coefficient, and its host evaluation
    equals the oracle's A(t), including the cut-off beyond t_cut."""
    rng = np.random.default_rng(8)
    for n in (2, 5, 1000):
        y = rng.normal(size=n)
        d0, d1 = rng.normal(size=2)
        c = hhg.spline_coefficients(y, 0.05, d0, d1)
        co = O.spline_coefficients(y, 0.05, d0, d1)
        assert np.allclose(c, co, rtol=0, atol=1e-12 * max(1.0, np.abs(co).max()))
    y = np.sin(np.linspace(0.0, 6.0, 300)) ** 2
    laser = hhg.SplineLaser(y, 0.3, 0.02, 0.1, -0.2, t_cut=5.0)
    p = O.spline_pulse(y, 0.3, 0.02, 0.1, -0.2, t_cut=5.0)
    t = np.linspace(0.0, 7.0, 2001)
    assert np.max(np.abs(laser.laser_function(t) - O.pulse_A(p, t))) < 1e-14
    assert np.all(laser.laser_function([0.29, 5.01, 6.5]) == 0.0)


def _synthetic_fields():
    """Measured-field stand-ins of the reference experiment's shape (201 samples,
    MV/cm): two chirped, Gaussian-modulated carriers."""
    i = np.arange(201)
    ea = 1.3 * np.exp(-((i - 90) / 22.0) ** 2) * np.sin(0.33 * i + 0.002 * i ** 2)
    eb = 0.6 * np.exp(-((i - 110) / 30.0) ** 2) * np.sin(0.3 * i + 1.0)
    return ea, eb


@pytest.mark.parametrize("gauss", [0, 1], ids=["cos2", "gauss"])
@pytest.mark.parametrize("active", [0, 1, 2], ids=["both", "A", "B"])
def test_reference_double_laser_table_matches_oracle(hhg, gauss, active):
    """DoubleLaser<Cosine|Gauss>Laser's 1000-sample table in the reference's units."""
    laser = hhg.reference_double_laser(gauss, 1.1, 0.8, 0.05, 2.0, 0.35, active)
    y, dt, t_max = O.reference_double_laser_samples(gauss, 1.1, 0.8, 0.05, 2.0, 0.35, active)
    assert len(laser.samples) == 1000 and abs(laser.dt - dt) < 1e-15 * dt and laser.t_end == t_max
    assert np.allclose(laser.samples, y, rtol=0, atol=1e-13 * np.abs(y).max())
    assert abs(laser.momentum_amplitude - 0.05 * 1e11 * 0.8 / (O.EXP_PHOTON_ENERGY * 1.1)) < 1e-12
    t = np.linspace(0.0, 1.1 * t_max, 777)
    ref = O.pulse_A(O.spline_pulse(y, 0.0, dt, t_cut=t_max), t)
    assert np.allclose(laser.laser_function(t), ref, rtol=0, atol=1e-12 * np.abs(y).max())


@pytest.mark.parametrize("shift", [0.0, 1.7, -0.9])
@pytest.mark.parametrize("active", [0, 1, 2], ids=["both", "A", "B"])
def test_reference_experimental_laser_table_matches_oracle(hhg, shift, active):
    """ExperimentalLaser's table from measured fields (synthetic stand-ins of the
    same shape): vector potential, quartic tail, spline, shift and scale."""
    ea, eb = _synthetic_fields()
    laser = hhg.reference_experimental_laser(0.9, 1.3, 0.02, shift, ea, eb, active)
    y, dt, t_end = O.reference_experimental_samples(0.9, 1.3, 0.02, shift, ea, eb, active)
    assert len(laser.samples) == len(y) and abs(laser.dt / dt - 1) < 1e-15 and abs(laser.t_end / t_end - 1) < 1e-15
    assert np.allclose(laser.samples, y, rtol=0, atol=1e-12 * np.abs(y).max())
    # the tail returns A to zero with zero slope, so the table ends in zeros
    assert np.all(laser.samples[-3:] == 0.0)


def test_experiment_vector_potential_is_the_field_integral():
    """-dA/dt = E at the measured samples (first differences of the oracle's
    vector potential), and the quartic tail meets 0 with zero slope."""
    ea, _ = _synthetic_fields()
    A = O.experiment_vector_potential(ea, O.EXP_DT)
    assert len(A) == 201 + 16 + 8 + 8
    assert np.allclose(-(A[8:209] - A[7:208]) / O.EXP_DT, ea, rtol=0, atol=1e-12)
    assert np.all(A[:8] == 0.0) and np.all(A[223:] == 0.0)


def test_piflux_uniform_grid_matches_oracle(hhg):
    for n in (1, 3, 8):
        kx, ky, kz, w = hhg.KGrid.piflux_uniform(n).build3()
        okx, oky, okz, ow = O.piflux_uniform(n)
        for a, b in ((kx, okx), (ky, oky), (kz, okz), (w, ow)):
            assert np.array_equal(a, b)


def test_dirac_gauss_grid_matches_oracle(hhg):
    for n in ((8, 6), (9, 5), (64, 48)):
        kx, ky, kz, w = hhg.KGrid.dirac_gauss(*n, 2.5).build3()
        ox, oy, ow = O.dirac_gauss(*n, 2.5)
        for a, b in ((kx, ox), (ky, oy), (w, ow)):
            assert np.array_equal(a, b)
        assert not np.any(kz)


def test_piflux_occupation_grid_matches_oracle(hhg):
    for n in (1, 4, 9):
        kx, ky, kz, w = hhg.KGrid.piflux_occupation(n).build3()
        okx, oky, okz, ow = O.piflux_occupation(n)
        for a, b in ((kx, okx), (ky, oky), (kz, okz), (w, ow)):
            assert np.array_equal(a, b)


def test_plan_rejects_split_current_elsewhere(hhg):
    """The split current is the 3D pi-flux lattice's, with RKF78."""
    import math
    m = hhg.ModelDesc(0, 0, 1.0, 0.1, 0.6, 0.0, 0.0, 0.0, 0.0, 0.0, 0.5 * math.sqrt(3))
    job = hhg.Job(m, hhg.ContinuousLaser(0.5), hhg.TimeIntegrationConfig(0.0, 6.0, 64, 50), 0)
    with pytest.raises(hhg.HHGError) as ei:
        hhg.Plan([job], [hhg.KGrid.box(4, 4, -1.0, 1.0, -1.0, 1.0)], hhg.Solver())
    assert ei.value.status == hhg.ERROR_UNSUPPORTED
    m4 = hhg.ModelDesc(4, 0, 1.0, 0.1, 0.6, 0.0, 0.0, 0.0, 0.0, 0.0, 0.5 * math.sqrt(3))
    job = hhg.Job(m4, hhg.ContinuousLaser(0.5), hhg.TimeIntegrationConfig(0.0, 6.0, 64, 50), 0)
    with pytest.raises(hhg.HHGError) as ei:
        hhg.Plan([job], [hhg.KGrid.piflux_uniform(2)], hhg.Solver(method=1))
    assert ei.value.status == hhg.ERROR_UNSUPPORTED


def test_per_point_entry_points_reject_null_arguments(hhg):
    """hhg_trajectories and hhg_occupations report a missing plan or buffer as
    INVALID_ARGUMENT with a message, before any device work."""
    import ctypes
    L = hhg.lib()
    sel = np.zeros(1, dtype=np.int64)
    out = np.zeros(8)
    for fn in (L.hhg_trajectories, L.hhg_occupations):
        st = fn(None, sel.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 1,
                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert st == hhg.ERROR_INVALID_ARGUMENT
        assert L.hhg_last_error_string()


def test_plan_calls_are_traced_with_roctx_ranges(hhg):
    """Plan execution, graph launch, autotuning and capture open roctx ranges
    (rocprofv3 --marker-trace shows them): the library imports the range calls."""
    import subprocess
    out = subprocess.run(["nm", "-D", hhg.LIB_PATH], capture_output=True, text=True, check=True).stdout
    imported = {line.split()[-1] for line in out.splitlines() if line.split()[:1] == ["U"] or " U " in line}
    assert {"roctxRangePushA", "roctxRangePop"} <= imported
