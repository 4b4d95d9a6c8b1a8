"""Shared problem definitions for the parity tests: one description feeds
both the oracle (oracle.*) and the HIP path (higherharmonics_amd.*)."""
import math
from dataclasses import dataclass, field

import numpy as np

import oracle as O


@dataclass
class Problem:
    name: str
    model: dict
    pulse: dict
    time: tuple            # (t_begin, t_end, N, n_sub)
    grid: tuple            # ("box", nx, ny, x0, x1, y0, y1), ("hex", nx, ny), ("cyl", n_z, n_kappa, max_k)
                           # or ("pfc", z_range, n_coarse)
    solver: dict = field(default_factory=lambda: dict(method=0, rtol=1e-10, atol=1e-10, proposal=0))

    # ---- oracle side ----
    def o_model(self):
        m = dict(self.model)
        # the reference drivers' rescues of cos k belong to their grids
        if self.grid[0] == "pfc" and m.get("kind") == O.MODEL_PIFLUX3D:
            m["ic_rescue"] = 1
        if self.grid[0] in ("pfu", "pfo") and m.get("kind") == O.MODEL_PIFLUX3D:
            m["ic_rescue"] = 2
        return O.model(**m)

    def o_pulse(self):
        if self.pulse.get("kind") == O.PULSE_SPLINE:
            p = dict(self.pulse)
            del p["kind"]
            return O.spline_pulse(**p)
        return O.pulse(**self.pulse)

    def o_time(self):
        return O.time_grid(*self.time)

    def o_solver(self, **over):
        d = dict(self.solver)
        d.update(over)
        return O.solver(**d)

    def o_grid(self):
        kx, ky, _, w = self.o_grid3()
        return kx, ky, w

    def o_grid3(self):
        """(kx, ky, kz, w); kz = 0 for the 2D grids."""
        if self.grid[0] == "pfc":
            return O.piflux_continuum(*self.grid[1:])
        if self.grid[0] == "pfu":
            return O.piflux_uniform(*self.grid[1:])
        if self.grid[0] == "pfo":
            return O.piflux_occupation(*self.grid[1:])
        if self.grid[0] == "box":
            kx, ky, w = O.gl_box(*self.grid[1:])
        elif self.grid[0] == "cyl":
            kx, ky, w = O.dirac_cylinder(*self.grid[1:])
        elif self.grid[0] == "dgl":
            kx, ky, w = O.dirac_gauss(*self.grid[1:])
        else:
            kx, ky, w = O.gl_hex(*self.grid[1:])
        return kx, ky, np.zeros_like(kx), w

    # ---- HIP side ----
    def h_model(self, H):
        m = self.model
        return H.ModelDesc(m.get("kind", 0), m.get("relax", 0), m.get("v", 1.0), m.get("mass", 0.1),
                           m.get("hopping", 1.0), m.get("theta", 0.0), m.get("fermi_energy", 0.0),
                           m.get("beta", 0.0), m.get("relax_time", 0.0), m.get("relax_time_offdiag", 0.0),
                           m.get("split_dispersion", 0.0), m.get("perp_current", 0))

    def h_laser(self, H):
        p = self.pulse
        kind = p.get("kind", 0)
        if kind == O.PULSE_SPLINE:
            q = dict(p)
            del q["kind"]
            return H.SplineLaser(**q)

        class _L(H.Laser):
            def desc(_self):
                return H.PulseDesc(kind, 0, p.get("amp_a", 0.5), p.get("amp_b", 0.0), p.get("n_cycles", 8.0),
                                   p.get("sigma", 0.0), p.get("t0", 0.0), p.get("delay", 0.0), p.get("cep", 0.0))
        return _L()

    def h_time(self, H):
        return H.TimeIntegrationConfig(*self.time)

    def h_grid(self, H):
        if self.grid[0] == "box":
            return H.KGrid.box(*self.grid[1:])
        if self.grid[0] == "cyl":
            return H.KGrid.dirac_cylinder(*self.grid[1:])
        if self.grid[0] == "dgl":
            return H.KGrid.dirac_gauss(*self.grid[1:])
        if self.grid[0] == "pfc":
            return H.KGrid.piflux_continuum(*self.grid[1:])
        if self.grid[0] == "pfu":
            return H.KGrid.piflux_uniform(*self.grid[1:])
        if self.grid[0] == "pfo":
            return H.KGrid.piflux_occupation(*self.grid[1:])
        return H.KGrid.hexagonal(*self.grid[1:])

    def h_solver(self, H, **over):
        d = dict(self.solver)
        d.update(over)
        return H.Solver(method=d.get("method", 0), rtol=d.get("rtol", 1e-10), atol=d.get("atol", 1e-10),
                        proposal=d.get("proposal", 0), max_fail=d.get("max_fail", 500),
                        magnus_substeps=d.get("magnus_substeps", 1), max_steps=d.get("max_steps", 0))

    def h_plan(self, H, **kw):
        solver_over = kw.pop("solver", {})
        return H.Plan([H.Job(self.h_model(H), self.h_laser(H), self.h_time(H), 0)], [self.h_grid(H)],
                      self.h_solver(H, **solver_over), **kw)


def oracle_j(pr: Problem, n_threads=0, **solver_over):
    kx, ky, kz, w = pr.o_grid3()
    return O.current_density(pr.o_model(), pr.o_pulse(), pr.o_time(), pr.o_solver(**solver_over), kx, ky, w,
                             n_threads=n_threads, kz=kz)


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


PI = math.pi
T1 = 16 * PI  # 8 cycles

# Configuration 1 (full): Dirac-2D, v=1, Delta=0.1, cos^2 A0=0.5, 8 cycles,
# T2 = 10 cycles, GL 128 x 64 on |k| <= 3, RKF78 1e-10, 4096 samples.
CONFIG1 = Problem(
    "config1_dirac_128x64",
    model=dict(kind=O.MODEL_DIRAC2D, v=1.0, mass=0.1, relax=O.RELAX_T2, relax_time=20 * PI),
    pulse=dict(kind=O.PULSE_COS2, amp_a=0.5, n_cycles=8.0, t0=-T1 / 2),
    time=(-T1 / 2, T1 / 2, 4096, 50),
    grid=("box", 128, 64, -3.0, 3.0, -3.0, 3.0),
)

# Reduced cases (seconds on the oracle)
DIRAC_SMALL = Problem(
    "dirac_t2_small",
    model=dict(kind=O.MODEL_DIRAC2D, v=1.0, mass=0.1, relax=O.RELAX_T2, relax_time=20 * PI),
    pulse=dict(kind=O.PULSE_COS2, amp_a=0.5, n_cycles=2.0, t0=-2 * PI),
    time=(-2 * PI, 2 * PI, 512, 50),
    grid=("box", 24, 16, -3.0, 3.0, -3.0, 3.0),
)

DIRAC_COHERENT = Problem(
    "dirac_coherent_small",
    model=dict(kind=O.MODEL_DIRAC2D, v=1.0, mass=0.1),
    pulse=dict(kind=O.PULSE_COS2, amp_a=0.5, n_cycles=2.0, t0=-2 * PI, cep=0.3),
    time=(-2 * PI, 2 * PI, 512, 50),
    grid=("box", 24, 16, -3.0, 3.0, -3.0, 3.0),
)

DIRAC_RELAX_INIT = Problem(
    "dirac_relax_to_initial",
    model=dict(kind=O.MODEL_DIRAC2D, v=1.0, mass=0.2, relax=O.RELAX_TO_INITIAL, relax_time=5.0, theta=0.4,
               fermi_energy=0.05, beta=20.0),
    pulse=dict(kind=O.PULSE_GAUSS, amp_a=0.4, sigma=PI, t0=0.0),
    time=(-5 * PI, 5 * PI, 640, 50),
    grid=("box", 16, 16, -2.5, 2.5, -2.5, 2.5),
)

PIFLUX_SMALL = Problem(
    "piflux_double_gauss",
    model=dict(kind=O.MODEL_PIFLUX2D, hopping=1.0),
    pulse=dict(kind=O.PULSE_DOUBLE_GAUSS, amp_a=0.6, amp_b=0.6, sigma=PI, t0=0.0, delay=2 * PI, cep=0.7),
    time=(-5 * PI, 2 * PI + 5 * PI, 700, 50),
    grid=("box", 24, 24, -PI, PI, -PI, PI),
)

HONEYCOMB_SMALL = Problem(
    "honeycomb_cos2",
    model=dict(kind=O.MODEL_HONEYCOMB, hopping=1.0),
    pulse=dict(kind=O.PULSE_COS2, amp_a=0.8, n_cycles=2.0, t0=0.0, cep=1.1),
    time=(0.0, 4 * PI + 2 * PI, 600, 50),
    grid=("hex", 20, 20),
)

HONEYCOMB_DOUBLE_COS2 = Problem(
    "honeycomb_double_cos2",
    model=dict(kind=O.MODEL_HONEYCOMB, hopping=1.0, fermi_energy=0.1, beta=8.0),
    pulse=dict(kind=O.PULSE_DOUBLE_COS2, amp_a=0.5, amp_b=0.3, n_cycles=1.5, t0=0.0, delay=PI),
    time=(0.0, 3 * PI + PI + 2 * PI, 600, 50),
    grid=("hex", 16, 16),
)

# cos^2 envelope whose support edges (corners of A'') fall inside sampling
# intervals; strong field and coarse sampling make whole-interval steps across
# a corner visibly wrong
HONEYCOMB_CORNERS = Problem(
    "honeycomb_cos2_corners",
    model=dict(kind=O.MODEL_HONEYCOMB, hopping=1.0),
    pulse=dict(kind=O.PULSE_COS2, amp_a=1.69, n_cycles=2.0, t0=0.05, cep=0.3),
    time=(0.0, 6 * PI, 192, 50),
    grid=("hex", 12, 12),
)

# the reference's 3D Dirac cone (band frame, field along z) on its (k_z, kappa) grid
DIRAC3D_CONTINUOUS = Problem(
    "dirac3d_continuous",
    model=dict(kind=O.MODEL_DIRAC3D),
    pulse=dict(kind=O.PULSE_CONTINUOUS, amp_a=0.6),
    time=(-4 * PI, 4 * PI, 1024, 50),
    grid=("cyl", 16, 8, 4.0),
)

DIRAC3D_COS2_RELAX = Problem(
    "dirac3d_cos2_relax",
    model=dict(kind=O.MODEL_DIRAC3D, fermi_energy=0.2, beta=10.0, relax=O.RELAX_TO_INITIAL, relax_time=15.0),
    pulse=dict(kind=O.PULSE_COS2, amp_a=0.8, n_cycles=2.0, t0=0.0),
    time=(0.0, 4 * PI, 512, 50),
    grid=("cyl", 12, 6, 3.0),
)

# the same two on the Gauss-Legendre (k_z, kappa) rule of hhg_gauss.cpp:71-133
DIRAC3D_GAUSS_CONTINUOUS = Problem("dirac3d_gauss_continuous", DIRAC3D_CONTINUOUS.model, DIRAC3D_CONTINUOUS.pulse,
                                   DIRAC3D_CONTINUOUS.time, ("dgl", 12, 8, 4.0))
DIRAC3D_GAUSS_COS2_RELAX = Problem("dirac3d_gauss_cos2_relax", DIRAC3D_COS2_RELAX.model, DIRAC3D_COS2_RELAX.pulse,
                                   DIRAC3D_COS2_RELAX.time, ("dgl", 10, 6, 3.0))

# the reference's 3D pi-flux lattice (field along z) on its continuum-limit
# quadrature, z_range = 4, N_coarse = 1 (1184 points)
PIFLUX3D_COS2 = Problem(
    "piflux3d_cos2",
    model=dict(kind=O.MODEL_PIFLUX3D, hopping=0.6),
    pulse=dict(kind=O.PULSE_COS2, amp_a=0.9, n_cycles=2.0, t0=0.0),
    time=(0.0, 4 * PI, 256, 50),
    grid=("pfc", 4, 1),
)

PIFLUX3D_EQUILIBRIUM = Problem(
    "piflux3d_equilibrium",
    model=dict(kind=O.MODEL_PIFLUX3D, hopping=0.6, fermi_energy=0.3, beta=6.0, relax=O.RELAX_EQUILIBRIUM,
               relax_time=12.0, relax_time_offdiag=5.0),
    pulse=dict(kind=O.PULSE_CONTINUOUS, amp_a=0.7),
    time=(-2 * PI, 2 * PI, 256, 50),
    grid=("pfc", 4, 1),
)

# the reference's honeycomb relaxation toward the instantaneous equilibrium
HONEYCOMB_EQUILIBRIUM = Problem(
    "honeycomb_equilibrium",
    model=dict(kind=O.MODEL_HONEYCOMB, hopping=1.0, fermi_energy=0.4, beta=3.0, relax=O.RELAX_EQUILIBRIUM,
               relax_time=6.0),
    pulse=dict(kind=O.PULSE_COS2, amp_a=1.2, n_cycles=2.0, t0=0.0, cep=0.2),
    time=(0.0, 6 * PI, 384, 50),
    grid=("hex", 12, 12),
)

HONEYCOMB_EQUILIBRIUM_CONT = Problem(
    "honeycomb_equilibrium_continuous",
    model=dict(kind=O.MODEL_HONEYCOMB, hopping=1.0, fermi_energy=0.2, beta=5.0, relax=O.RELAX_EQUILIBRIUM,
               relax_time=3.0),
    pulse=dict(kind=O.PULSE_CONTINUOUS, amp_a=0.9),
    time=(0.7, 0.7 + 4 * PI, 256, 50),
    grid=("hex", 12, 12),
)

# the reference's split current of the 3D pi-flux lattice (PiFlux.cpp:954-1035):
# uniform grid N = 8 (1024 points, including the rescued x = y = N/2 points),
# cut sqrt(3)/2
PIFLUX3D_SPLIT = Problem(
    "piflux3d_split_current",
    model=dict(kind=O.MODEL_PIFLUX3D, hopping=0.6, fermi_energy=0.2, beta=4.0, split_dispersion=0.5 * math.sqrt(3)),
    pulse=dict(kind=O.PULSE_COS2, amp_a=0.9, n_cycles=2.0, t0=0.0),
    time=(0.0, 4 * PI, 256, 50),
    grid=("pfu", 8),
)

# the occupation slice (x, pi/2, z) pi/8 of the relaxing 3D pi-flux lattice: the
# x = pi/2 column has cos x = cos y = 0, where the persistent rescue of the
# occupation run acts (PiFlux.cpp:909-952, 252-257)
PIFLUX3D_OCCUPATION = Problem(
    "piflux3d_occupation",
    model=dict(kind=O.MODEL_PIFLUX3D, hopping=0.6, fermi_energy=0.1, beta=4.0, relax=O.RELAX_EQUILIBRIUM,
               relax_time=5.0, relax_time_offdiag=3.0),
    pulse=dict(kind=O.PULSE_COS2, amp_a=1.5, n_cycles=1.0, t0=0.0),
    time=(0.0, 2 * PI, 40, 50),
    grid=("pfo", 8),
    solver=dict(rtol=1e-8, atol=1e-12),
)

# the same slice under a continuous pulse that starts at A = -0.9: the shifted
# frame is updated to A(t_begin) before the integration (PiFlux.cpp:285)
PIFLUX3D_OCCUPATION_CONT = Problem(
    "piflux3d_occupation_continuous",
    model=dict(PIFLUX3D_OCCUPATION.model),
    pulse=dict(kind=O.PULSE_CONTINUOUS, amp_a=1.5),
    time=(-2.5, -2.5 + 2 * PI, 40, 50),
    grid=("pfo", 8),
    solver=dict(rtol=1e-8, atol=1e-12),
)

# a sampled A/B cos^2 pulse through its clamped cubic spline over 1.5 x the
# window as the reference's DoubleLaser table (DoubleLaser.hpp:48-73; it takes
# 1000 samples, 300 here make the spline error visible in j)
def _double_cos2_samples(n=300, n_cycles=2.0, amp_a=1.1, amp_b=0.5, delay=2.5):
    t_max = 1.5 * 2 * PI * n_cycles
    t = np.arange(n) * (t_max / (n - 1))
    p = O.pulse(O.PULSE_DOUBLE_COS2, amp_a=amp_a, amp_b=amp_b, n_cycles=n_cycles, t0=0.0, delay=delay)
    return O.pulse_A(p, t), t_max / (n - 1)


_SPL_Y, _SPL_DT = _double_cos2_samples()
HONEYCOMB_SPLINE = Problem(
    "honeycomb_spline_double_cos2",
    model=dict(kind=O.MODEL_HONEYCOMB, hopping=1.0),
    pulse=dict(kind=O.PULSE_SPLINE, samples=_SPL_Y, t0=0.0, dt=_SPL_DT),
    time=(0.0, 6 * PI, 384, 50),
    grid=("hex", 12, 12),
)

SMALL_CASES = [DIRAC_SMALL, DIRAC_COHERENT, DIRAC_RELAX_INIT, PIFLUX_SMALL, HONEYCOMB_SMALL, HONEYCOMB_DOUBLE_COS2]
