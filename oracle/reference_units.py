"""The reference's configuration keys and units, restated for the oracle.

TEST INFRASTRUCTURE ONLY (the oracle's side of the reference-units driver):
``tests/`` and bench.py's CPU-baseline and audit legs use it; product code
never does.  Restates, from the reference's sources:

* DiracSystem (sources/HHG/Systems/DiracSystem.cpp:45-66, DiracSystem.hpp:30-32):
  E_F and 1/T in units of the photon energy, v_F only through the laser ratio
  hbar v_F / E_ph, band width = max |k|, relaxation rate 1e15 hbar / (tau E_ph);
* PiFlux (PiFlux.cpp:56-70): hopping W / sqrt 12, lattice constant
  sqrt3 hbar v_F / (E_ph W), both relaxation rates (off-diagonal <= 0: the
  diagonal one, PiFluxDispatcher.cpp:18-34);
* Honeycomb (Honeycomb.cpp:83-92): hopping W / 6, lattice constant
  4 hbar v_F / (E_ph W);
* Laser amplitude ratio 1e11 E0 / E_ph (Laser.cpp:24-25), the time windows of
  the dispatchers (continuous [-n pi, n pi]; cos^2 from pi t0_offset, the
  lattices one cycle longer) and N = n_cycles 1024 (hhg.cpp:87-88); the
  tabulated DoubleLaser of the pi-flux lattice with the photon energy in units
  of the experiment's (Dispatcher.cpp:10-19); the pi-flux GaussLaser over
  [t0_offset, t0_offset + 6 pi n] (GaussLaser.cpp:9-12, GaussLaser.hpp:13-16);
* split_current (hhg.cpp:174-219, PiFlux.cpp:954-1035): the uniform n_z grid with
  its rescue and the sqrt3/2 cut; with it or occupations the tabulated pulses
  stop at 0.75 t_end (PiFluxDispatcher.cpp:44-104).
"""
from __future__ import annotations

import math

import numpy as np

from . import (MODEL_DIRAC3D, MODEL_HONEYCOMB, MODEL_PIFLUX3D, METHOD_MAGNUS_MOMENTS, PULSE_CONTINUOUS, PULSE_COS2,
               PULSE_GAUSS, RELAX_EQUILIBRIUM, RELAX_TO_INITIAL, EXP_PHOTON_ENERGY, REF_HBAR, dirac_cylinder, dirac_gauss,
               gl_hex,
               model, piflux_continuum, piflux_uniform, pulse, reference_double_laser_samples, solver, spline_pulse,
               time_grid)

K_B = 0.08617333262145177434  # meV / K, GlobalDefinitions.hpp:30


def parse(text: str) -> dict:
    """key value lines, '#' comments (the reference's InputFileReader format)."""
    out = {}
    for line in text.splitlines():
        line = line.split("#", 1)[0].split()
        if len(line) >= 2:
            out[line[0]] = line[1]
    return out


def setup(cfg: dict, grid_override: dict | None = None) -> dict:
    """Oracle model, pulse, time grid, solver and (kx, ky, kz, w) of a reference
    configuration (system_type Dirac | PiFlux | Honeycomb)."""
    f = lambda k, d: float(cfg.get(k, d))
    system = cfg.get("system_type", "Dirac")
    T, E_F, v_F, W = f("T", 0.0), f("E_F", 0.0), f("v_F", 1.5e5), f("band_width", 10.0)
    E0, e_in = f("field_amplitude", 1.0), f("photon_energy", 5.25)
    tau, tau_o = f("diagonal_relaxation_time", -1.0), f("offdiagonal_relaxation_time", -1.0)
    laser = cfg.get("laser_type", "continuous")
    n_cyc = int(f("n_laser_cycles", 8))
    t0_off = f("t0_offset", 0.0)
    N = int(f("n_measurements", n_cyc * 1024))
    g = dict(cfg)
    g.update(grid_override or {})
    tabulated = laser.startswith("dcos") or laser.startswith("dgauss")
    split = system == "PiFlux" and cfg.get("split_current", "no") != "no"
    t_scale = 0.75 if (split or cfg.get("occupations", "no") != "no") else 1.0
    e_sys = e_in * EXP_PHOTON_ENERGY if (tabulated and system != "Dirac") else e_in
    beta = e_sys / (K_B * T) if T > 0 else 0.0
    rate = lambda t_fs: 1e15 * REF_HBAR / (t_fs * e_sys)
    sv = solver(rtol=1e-8, atol=1e-12)
    if system == "Dirac":
        ratio = REF_HBAR * v_F / e_sys
        rule = dirac_gauss if g.get("z_rule", "uniform") == "gauss" else dirac_cylinder
        kx, ky, w = rule(int(g.get("n_z", 64)), int(g.get("n_kappa", 32)), W)
        kz = np.zeros_like(kx)
        if tau > 0:
            m = model(MODEL_DIRAC3D, fermi_energy=E_F / e_sys, beta=beta, relax=RELAX_TO_INITIAL,
                      relax_time=1.0 / rate(tau))
        else:
            m = model(MODEL_DIRAC3D, fermi_energy=E_F / e_sys, beta=beta)
            sv = solver(method=METHOD_MAGNUS_MOMENTS)
    elif system == "PiFlux":
        ratio = math.sqrt(3) * REF_HBAR * v_F / (e_sys * W)
        if split:
            kx, ky, kz, w = piflux_uniform(int(g.get("n_z", 64)))
            kw = dict(hopping=W / math.sqrt(12), fermi_energy=E_F / e_sys, beta=beta, ic_rescue=2,
                      split_dispersion=0.5 * math.sqrt(3))
        else:
            kx, ky, kz, w = piflux_continuum(int(g.get("z_range", 32)), int(g.get("n_coarse", 4)))
            kw = dict(hopping=W / math.sqrt(12), fermi_energy=E_F / e_sys, beta=beta, ic_rescue=1)
        if tau > 0:
            kw.update(relax=RELAX_EQUILIBRIUM, relax_time=1.0 / rate(tau),
                      relax_time_offdiag=1.0 / rate(tau_o if tau_o > 0 else tau))
        m = model(MODEL_PIFLUX3D, **kw)
    elif system == "Honeycomb":
        ratio = 4 * REF_HBAR * v_F / (e_sys * W)
        n_k = int(g.get("N_k", 64))
        kx, ky, w = gl_hex(2 * n_k, 2 * n_k)
        kz = np.zeros_like(kx)
        kw = dict(hopping=W / 6.0, fermi_energy=E_F / e_sys, beta=beta)
        if tau > 0:
            kw.update(relax=RELAX_EQUILIBRIUM, relax_time=1.0 / rate(tau))
        m = model(MODEL_HONEYCOMB, **kw)
    else:
        raise ValueError(system)
    A0 = ratio * 1e11 * E0 / e_in
    if laser == "continuous":
        p, t_b, t_e = pulse(PULSE_CONTINUOUS, amp_a=A0), -n_cyc * math.pi, n_cyc * math.pi
    elif laser == "cosine":
        t_b = math.pi * t0_off
        t_e = t_b + 2 * math.pi * n_cyc + (0.0 if system == "Dirac" else 2 * math.pi)
        p = pulse(PULSE_COS2, amp_a=A0, n_cycles=float(n_cyc), t0=t_b)
    elif tabulated and system == "PiFlux":
        gauss = laser.startswith("dgauss")
        active = 1 if laser.endswith("A") else (2 if laser.endswith("B") else 0)
        y, dt, t_max = reference_double_laser_samples(gauss, e_in, E0, ratio, float(n_cyc), t0_off, active)
        p, t_b, t_e = spline_pulse(y, 0.0, dt, t_cut=t_max), 0.0, t_scale * t_max
    elif laser == "gauss" and system == "PiFlux":
        base = 6 * math.pi * n_cyc
        p = pulse(PULSE_GAUSS, amp_a=A0, sigma=math.pi * n_cyc / math.sqrt(2 * math.log(20.0)), t0=t0_off + 0.5 * base)
        t_b, t_e = t0_off, t_scale * (t0_off + base)
    else:
        raise ValueError(laser)
    return dict(model=m, pulse=p, time=time_grid(t_b, t_e, N, 50), solver=sv, kx=kx, ky=ky, kz=kz, w=w,
                momentum_amplitude=A0)
