#!/usr/bin/env python3
"""Throughput of the Bloch-equation hot path on one MI355X.

Workload (one "step" = one full solve of the configuration):
  2D massive Dirac d = (v(kx - A), v ky, Delta), v = 1, Delta = 0.1, on a
  1024 x 512 Gauss-Legendre grid over |kx|, |ky| <= 3 (524,288 k-points);
  cos^2 pulse A0 = 0.5, 8 optical cycles, T2 = 10 cycles transverse
  dephasing; 4096 sampling intervals; FP64 RKF78 at rtol = atol = 1e-10 with
  the reference's controller restart at Delta/50 every interval.

Two legs are timed on the same workload: the plain controller (restart at
Delta/50 every interval, as the reference) and the learned proposal (a BF16
network, trained on the GPU from step logs of a carried-controller solve on a
coarser grid, proposes each interval's first step; the embedded-error
controller still judges every step).  Both are checked against a
tight-tolerance solve (rtol = atol = 1e-13, first step Delta/500 every
interval) of the same grid; ``value`` is the learned leg's trajectories/s
when its j(t) error is within 1e-9, else the plain leg's.

Prints one JSON line: trajectories/s, the FP64 roofline of the solve kernel
(algorithmic FLOPs from the attempted-step counter over the HIP-event time of
the kernel on its own stream), and the CPU oracle timed on a bounded sample
of the same workload on the host cores.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--small]
"""
import argparse
import gzip
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64, vector and MFMA

# Algorithmic FLOPs (convention: + - * = 1, FMA = 2, / sqrt = 1, each
# sin/cos/exp = 1).  RKF78 step algebra: 55 tableau FMAs x 3 components +
# stage updates + 8th-order sum + embedded error + controller.
F_STEP = 450.0
F_RHS = {"dirac_t2": 47.0, "dirac": 23.0, "piflux": 26.0, "honeycomb": 34.0}
F_VERTEX = {"dirac": 3.0, "piflux": 4.0, "honeycomb": 20.0}
F_MAGNUS = 85.0
# One network inference per lane and interval: 2 x (8*64 + 2*64*64 + 64*2) FLOPs
# (BF16 MFMA hidden layers, FP32 output layer), not counted as FP64 work.
F_MLP = 2.0 * (8 * 64 + 2 * 64 * 64 + 64 * 2)

# HBM bytes per solve launch of this workload from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch, scaled by 2.0 and 1.0, the factors a plain 1 GiB copy
# reads): plain 27,889.1 x 2.0 + 1,050,496 x 1.0, learned 14,680.2 x 2.0 + 1,050,496 x 1.0;
# profiles/solve_kernel_dirac_t2_1024x512_counters.md and
# profiles/solve_kernel_learned_dirac_t2_1024x512_counters.md.
TRAFFIC_BYTES = (27889.1 * 2.0 + 1050496.0 * 1.0) * 1024
TRAFFIC_BYTES_LEARNED = (14680.2 * 2.0 + 1050496.0 * 1.0) * 1024
# Hardware FP64 FLOP per solve launch from the same counter runs (SQ_INSTS_VALU_FMA_F64 x 2 +
# MUL + ADD + TRANS, x 64 lanes): the rocprof view of the FP64 rate beside the algorithmic one.
HW_FP64_FLOP = (2 * 64123772928 + 15535849472 + 4631150592 + 2147508224) * 64
HW_FP64_FLOP_LEARNED = (2 * 16208044032 + 4093788160 + 1477033984 + 536895488) * 64


def workload(small: bool):
    T = 16 * math.pi
    nx, ny = (256, 128) if small else (1024, 512)
    return dict(T=T, nx=nx, ny=ny, N=4096, n_sub=50, v=1.0, mass=0.1, A0=0.5, n_cycles=8.0, T2=20 * math.pi,
                rtol=1e-10, atol=1e-10)


def make_plan(H, wl, method=0, relax=2, proposal=0, magnus_substeps=5, mlp=None, rtol=None, infer_every=1, n_sub=None,
              **launch):
    T = wl["T"]
    md = H.ModelDesc(H.MODEL_DIRAC2D, relax, wl["v"], wl["mass"], 0.0, 0.0, 0.0, 0.0, wl["T2"] if relax else 0.0)
    laser = H.CosineLaser(wl["A0"], wl["n_cycles"], -T / 2)
    tc = H.TimeIntegrationConfig(-T / 2, T / 2, wl["N"], n_sub or wl["n_sub"])
    grid = H.KGrid.box(wl["nx"], wl["ny"], -3.0, 3.0, -3.0, 3.0)
    tol = rtol or wl["rtol"]
    solver = H.Solver(method=method, rtol=tol, atol=tol if rtol else wl["atol"], proposal=proposal,
                      magnus_substeps=magnus_substeps, mlp=mlp, infer_every=infer_every)
    return H.Plan([H.Job(md, laser, tc, 0)], [grid], solver, **launch)


def train_proposal(H, wl):
    """The step-proposal network trained on accepted-step logs of the same
    configuration on a 256 x 128 grid of the same box (every 64th interval of
    every k-point, about 2.1M records per training round): the first from a
    carried-controller solve, the second adds a learned-proposal solve with the
    first network."""
    from higherharmonics_amd import learn
    coarse = dict(wl, nx=256, ny=128)
    t0 = time.perf_counter()
    net, rec, hist = learn.train_on_solver_logs(
        lambda sv: make_plan(H, coarse, proposal=sv.proposal, mlp=sv.mlp, persistent=1),
        capacity=1 << 21, stride=64, rounds=2, epochs=6, batch=1024)
    return net, {"records": int(len(rec)), "seconds": time.perf_counter() - t0, "rounds": len(hist),
                 "epochs": len(hist[-1]), "loss_first": float(hist[-1][0]), "loss_last": float(hist[-1][-1]),
                 "params": net.n_params,
                 "log_source": "carried-controller then learned-proposal solve, 256x128 grid of the same box, "
                               "stride 64"}


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def time_plan(plan, steps, warmup):
    for _ in range(warmup):
        plan.execute()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass
    solve_ms = []
    t0 = time.perf_counter()
    j = None
    for _ in range(steps):
        j = plan.execute()[0]
        solve_ms.append(plan.stats.solve_ms)
    wall = time.perf_counter() - t0
    return wall, float(np.mean(solve_ms)), plan.stats.as_dict(), j


def cpu_baseline(wl, target_s=10.0):
    """The oracle (optimised build) on a seeded random sample of the workload's
    k-points, OpenMP over the host cores (capped at 16)."""
    import oracle as O
    T = wl["T"]
    kx, ky, w = O.gl_box(wl["nx"], wl["ny"], -3.0, 3.0, -3.0, 3.0)
    m = O.model(O.MODEL_DIRAC2D, v=wl["v"], mass=wl["mass"], relax=O.RELAX_T2, relax_time=wl["T2"])
    p = O.pulse(O.PULSE_COS2, wl["A0"], n_cycles=wl["n_cycles"], t0=-T / 2)
    tc = O.time_grid(-T / 2, T / 2, wl["N"], wl["n_sub"])
    sv = O.solver(rtol=wl["rtol"], atol=wl["atol"])
    threads = max(1, min(16, os.cpu_count() or 1))
    rng = np.random.default_rng(20261203)
    # pilot large enough that every thread runs several k-chunks of 64
    pilot = rng.choice(len(kx), size=threads * 256, replace=False)
    t0 = time.perf_counter()
    O.current_density(m, p, tc, sv, kx[pilot], ky[pilot], w[pilot], n_threads=threads, fast=True)
    per = (time.perf_counter() - t0) / len(pilot)
    n = int(min(len(kx), max(len(pilot), target_s / max(per, 1e-9))))
    sel = rng.choice(len(kx), size=n, replace=False)
    t0 = time.perf_counter()
    O.current_density(m, p, tc, sv, kx[sel], ky[sel], w[sel], n_threads=threads, fast=True)
    wall = time.perf_counter() - t0
    return {"value": n / wall, "unit": "trajectories/s", "cores": threads, "kind": "port",
            "sample": f"{n} seeded-random k-points of the {wl['nx']}x{wl['ny']} grid (seed 20261203), "
                      f"oracle -O3 -ffast-math build, OpenMP {threads} threads, {wall:.1f} s"}


def config1_full(H, threads=None):
    """The Dirac-2D workload on the 128 x 64 grid (8,192 trajectories), the
    reference's own CPU-sized case, in full on both sides: the plain GPU kernel
    (mean of three executions after one warmup) and the oracle's optimised build
    on the host cores, with j(t) compared between them."""
    import oracle as O
    wl = dict(workload(False), nx=128, ny=64)
    plan = make_plan(H, wl)
    _, ms, st, j_gpu = time_plan(plan, 3, 1)
    plan.close()
    T = wl["T"]
    kx, ky, w = O.gl_box(wl["nx"], wl["ny"], -3.0, 3.0, -3.0, 3.0)
    m = O.model(O.MODEL_DIRAC2D, v=wl["v"], mass=wl["mass"], relax=O.RELAX_T2, relax_time=wl["T2"])
    p = O.pulse(O.PULSE_COS2, wl["A0"], n_cycles=wl["n_cycles"], t0=-T / 2)
    tc = O.time_grid(-T / 2, T / 2, wl["N"], wl["n_sub"])
    sv = O.solver(rtol=wl["rtol"], atol=wl["atol"])
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    j_cpu, _ = O.current_density(m, p, tc, sv, kx, ky, w, n_threads=threads, fast=True)
    cpu_s = time.perf_counter() - t0
    K = wl["nx"] * wl["ny"]
    return {"trajectories": K, "gpu_plain_kernel_ms": ms, "gpu_trajectories_per_s": K / (ms * 1e-3),
            "gpu_attempts": st["attempts"], "cpu_s": cpu_s, "cpu_trajectories_per_s": K / cpu_s,
            "cpu_threads": threads, "gpu_over_cpu": (K / (ms * 1e-3)) / (K / cpu_s),
            "rel_l2_gpu_vs_cpu": rel_l2(j_gpu, j_cpu)}


def sweep_extras(H):
    """Reduced-size sweeps of configurations 3-5 (same structure, fewer
    k-points per job): one plan each, autotuned, graph-captured, timed once
    after one warmup."""
    from higherharmonics_amd import workloads as W
    res = {}
    cases = [("config3_piflux_delay_sweep_128x128", W.piflux_sweep(n=128)),
             ("config4_honeycomb_16_pulses_192x192", W.honeycomb_cos2(n=192)),
             ("config5_mixed_64_honeycomb_64_piflux_64x64", W.mixed_sweep(n=64))]
    for name, jobs in cases:
        print(f"[bench] sweep {name}: {len(jobs)} jobs, {W.trajectories(jobs)} trajectories", file=sys.stderr,
              flush=True)
        plan = W.build_plan(H, jobs, persistent=1)
        tuned = plan.autotune()
        plan.capture_graph()
        plan.execute()
        t0 = time.perf_counter()
        plan.execute()
        wall = time.perf_counter() - t0
        st = plan.stats.as_dict()
        plan.close()
        res[name] = {"jobs": len(jobs), "trajectories": W.trajectories(jobs),
                     "samples_per_job": [min(j["time"][2] for j in jobs), max(j["time"][2] for j in jobs)],
                     "trajectories_per_s": W.trajectories(jobs) / wall, "wall_s": wall, "solve_ms": st["solve_ms"],
                     "attempts": st["attempts"], "rejects": st["rejects"], "launch": tuned}
    return res


def config3_full_leg(H):
    """Configuration 3 at full size: PiFlux-2D on the 512 x 512 grid, 32 A/B
    delays x 4 amplitudes of the double Gaussian (33.5M trajectories, 10,405 to
    15,360 samples per job), plain RKF78, one persistent launch over all 128 jobs;
    one execution (no warmup: the run takes tens of seconds), the launch shape
    the autotuner picks on the reduced sweep of the same jobs."""
    from higherharmonics_amd import workloads as W
    small = W.piflux_sweep(n=128)
    plan = W.build_plan(H, small, persistent=1)
    tuned = plan.autotune()
    plan.close()
    jobs = W.piflux_sweep()
    K = W.trajectories(jobs)
    print(f"[bench] configuration 3 in full: {len(jobs)} jobs, {K} trajectories, launch {tuned}", file=sys.stderr,
          flush=True)
    plan = W.build_plan(H, jobs, persistent=1, waves_per_block=tuned["waves_per_block"],
                        queue_chunk=tuned["queue_chunk"])
    t0 = time.perf_counter()
    plan.execute()
    wall = time.perf_counter() - t0
    st = plan.stats.as_dict()
    plan.close()
    return {"jobs": len(jobs), "trajectories": K,
            "samples_per_job": [min(j["time"][2] for j in jobs), max(j["time"][2] for j in jobs)],
            "trajectories_per_s": K / wall, "wall_s": wall, "solve_ms": st["solve_ms"], "attempts": st["attempts"],
            "rejects": st["rejects"], "fp64_tflops": st["attempts"] * (F_STEP + 13 * F_RHS["piflux"]) /
            (st["solve_ms"] * 1e-3) / 1e12, "launch": tuned}


def config4_leg(H, n=768, n_pulses=16, log_n=96):
    """Honeycomb over the hexagonal BZ (GL n x n), n_pulses seeded cos^2 pulses
    (10 cycles, A0 ~ U[0.1, 1.5], CEP ~ U[0, 2 pi]), 512 samples per cycle:
    plain vs learned proposal (with and without step-count grouping), each
    against a tight-tolerance solve of the same jobs.  The network is trained
    on step logs of a carried-controller solve of the same pulses on a
    log_n x log_n grid."""
    from higherharmonics_amd import learn
    from higherharmonics_amd import workloads as W
    jobs = W.honeycomb_cos2(n=n, n_pulses=n_pulses)
    K = W.trajectories(jobs)
    small = W.honeycomb_cos2(n=log_n, n_pulses=n_pulses)

    def solver(**kw):
        return H.Solver(rtol=1e-10, atol=1e-10, **kw)

    # log stride so that the records cover the whole time window of every job
    cap = 1 << 21
    stride = max(1, -(-sum(j["grid"][1] * j["grid"][2] * j["time"][2] for j in small) // cap))
    t0 = time.perf_counter()
    net, rec, hist = learn.train_on_solver_logs(lambda sv: W.build_plan(H, small, sv, persistent=1), capacity=cap,
                                                stride=stride, rounds=2)
    losses = hist[-1]
    train_s = time.perf_counter() - t0
    res = {"workload": f"honeycomb hex BZ GL {n}x{n}, {n_pulses} cos^2 pulses (10 cycles + 1), 512 samples/cycle",
           "trajectories": K, "training": {"records": int(len(rec)), "log_stride": stride, "seconds": train_s,
                                            "loss_first": float(losses[0]), "loss_last": float(losses[-1])}}
    js = {}
    for name, sv, kw in [("plain", solver(), {}),
                         ("learned", solver(proposal=H.PROPOSAL_LEARNED, mlp=net), {}),
                         ("learned_grouped", solver(proposal=H.PROPOSAL_LEARNED, mlp=net), {"group_by_cost": 1}),
                         ("carry", solver(proposal=H.PROPOSAL_CARRY), {})]:
        print(f"[bench] config 4 {name}", file=sys.stderr, flush=True)
        plan = W.build_plan(H, jobs, sv, persistent=1, **kw)
        plan.execute()
        t1 = time.perf_counter()
        js[name] = plan.execute()
        wall = time.perf_counter() - t1
        st = plan.stats.as_dict()
        plan.close()
        res[name] = {"trajectories_per_s": K / wall, "wall_s": wall, "kernel_ms": st["solve_ms"],
                     "attempts": st["attempts"], "rejects": st["rejects"], "fallbacks": st["proposal_fallbacks"]}
    print("[bench] config 4 tight reference", file=sys.stderr, flush=True)
    tight_jobs = [dict(j, time=(j["time"][0], j["time"][1], j["time"][2], 500)) for j in jobs]
    plan = W.build_plan(H, tight_jobs, H.Solver(rtol=1e-13, atol=1e-13), persistent=1)
    jt = plan.execute()
    plan.close()
    for name in js:
        res[name]["rel_l2_err_vs_tight_max_over_pulses"] = max(rel_l2(a, b) for a, b in zip(js[name], jt))
    res["learned_vs_plain"] = res["learned"]["trajectories_per_s"] / res["plain"]["trajectories_per_s"]
    res["tight_reference"] = "same kernel, restart controller at Delta/500, rtol=atol=1e-13"
    net.close()
    return res


def config5_leg(H, n=256, log_n=32):
    """64 honeycomb (double cos^2) + 64 PiFlux (double Gaussian) seeded pulse
    configurations on n x n grids in one plan: plain vs learned proposal with
    step-count grouping, both autotuned and graph-captured.  The network is
    trained on step logs of the same jobs on log_n x log_n grids.  Audit: a
    seeded random 1 % of every job's k-points solved with the learned proposal,
    the plain controller and a tight tolerance."""
    from higherharmonics_amd import learn
    from higherharmonics_amd import workloads as W
    jobs = W.mixed_sweep(n=n)
    K = W.trajectories(jobs)
    small = W.mixed_sweep(n=log_n)
    cap = 1 << 21
    stride = max(1, -(-sum(j["grid"][1] * j["grid"][2] * j["time"][2] for j in small) // cap))
    t0 = time.perf_counter()
    net, rec, hist = learn.train_on_solver_logs(lambda sv: W.build_plan(H, small, sv, persistent=1), capacity=cap,
                                                stride=stride, rounds=2)
    losses = hist[-1]
    res = {"workload": f"64 honeycomb double-cos^2 + 64 PiFlux double-Gaussian pulses, GL {n}x{n} each, "
                       f"256 samples/cycle", "jobs": len(jobs), "trajectories": K,
           "training": {"records": int(len(rec)), "log_stride": stride, "seconds": time.perf_counter() - t0,
                        "loss_first": float(losses[0]), "loss_last": float(losses[-1])}}
    for name, sv, kw in [("plain", H.Solver(), {}),
                         ("learned_grouped", H.Solver(proposal=H.PROPOSAL_LEARNED, mlp=net), {"group_by_cost": 1})]:
        print(f"[bench] config 5 {name}", file=sys.stderr, flush=True)
        plan = W.build_plan(H, jobs, sv, persistent=1, **kw)
        tuned = plan.autotune()
        plan.capture_graph()
        plan.execute()
        t1 = time.perf_counter()
        plan.execute()
        wall = time.perf_counter() - t1
        st = plan.stats.as_dict()
        plan.close()
        res[name] = {"trajectories_per_s": K / wall, "wall_s": wall, "attempts": st["attempts"],
                     "rejects": st["rejects"], "fallbacks": st["proposal_fallbacks"], "launch": tuned}
    res["learned_vs_plain"] = res["learned_grouped"]["trajectories_per_s"] / res["plain"]["trajectories_per_s"]
    print("[bench] config 5 audit", file=sys.stderr, flush=True)
    sub = W.audit_subset(H, jobs, 0.01)
    tight = [dict(j, time=(j["time"][0], j["time"][1], j["time"][2], 500)) for j in sub]
    out = {}
    for name, js, sv in [("learned", sub, H.Solver(proposal=H.PROPOSAL_LEARNED, mlp=net)), ("plain", sub, H.Solver()),
                         ("tight", tight, H.Solver(rtol=1e-13, atol=1e-13))]:
        plan = W.build_plan(H, js, sv, persistent=1)
        out[name] = plan.execute()
        plan.close()
    res["audit"] = {"k_points": W.trajectories(sub), "fraction": 0.01,
                    "max_rel_l2_learned_vs_plain": max(rel_l2(a, b) for a, b in zip(out["learned"], out["plain"])),
                    "max_rel_l2_learned_vs_tight": max(rel_l2(a, b) for a, b in zip(out["learned"], out["tight"])),
                    "max_rel_l2_plain_vs_tight": max(rel_l2(a, b) for a, b in zip(out["plain"], out["tight"]))}
    net.close()
    return res


# The reference's own systems in its units, run through the reference-units
# driver (hhg_run) as a drop-in for its dispatchers: the configurations in
# higherharmonics_amd/configs (the reference's params/params.config parameters
# with 4 laser cycles, and its params/dcos.config), with the grid overrides of
# the small audit runs.
REF_LEGS = [
    ("dirac3d_continuous_moment_magnus", "dirac_params.cfg", "n_z 12\nn_kappa 6\n", ""),
    ("piflux3d_cos2_equilibrium_relaxation", "piflux_params.cfg", "z_range 2\nn_coarse 1\n", ""),
    ("honeycomb_cos2_equilibrium_relaxation", "honeycomb_params.cfg", "N_k 6\n", ""),
    ("piflux3d_dcos_spline_pulse", "piflux_dcos.cfg", "z_range 2\nn_coarse 1\n", ""),
    # PiFlux::current_per_energy on the uniform 32^3 x 2 grid (split_current.json.gz);
    # the audit grid is odd: an even one holds the point where all three cosines
    # vanish, on which the relaxation starts on the Dirac point and its two oracle
    # builds already differ by percent
    ("piflux3d_split_current_equilibrium", "piflux_params.cfg", "n_z 5\n", "split_current only\nn_z 32\n"),
]
# Algorithmic FLOPs of the reference systems' right-hand sides (same convention):
# pi-flux 3D 4t m x s with cos(z - A): 15; its equilibrium relaxation (|m|, two
# Fermi factors, band-frame phase, targets, 3 x 7 update): 100 more; honeycomb
# RHS 34 plus its equilibrium state at k - A and the 3 x 3 update: 25 more.
# One moment-Magnus step of the 3D Dirac cone (generator polynomials, Rodrigues
# rotation): 175 per k-point and interval.
F_RHS_REF = {"piflux3d": 15.0, "piflux3d_eq": 115.0, "honeycomb_eq": 59.0}
F_MOMENT_STEP = 175.0


def reference_models_leg(H, repeats=3, cpu_s=5.0):
    """Each reference system through hhg_run: trajectories/s from the kernel
    time of `repeats` executions of one plan (the first not counted), an audit
    of the same physics on a small grid against the oracle, and the oracle on
    a seeded sample of the full grid on the host cores."""
    import subprocess
    import tempfile
    import oracle as O
    from oracle import reference_units as R
    exe = os.path.join(os.path.dirname(H.LIB_PATH), "hhg_run")
    threads = max(1, min(16, os.cpu_count() or 1))
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        def run(text, name):
            # each run in its own folder: the split current is written beside the output
            folder = os.path.join(tmp, name)
            os.makedirs(folder)
            cfg = os.path.join(folder, "run.cfg")
            out = os.path.join(folder, "current_density.json")
            with open(cfg, "w") as fh:
                fh.write(text)
            r = subprocess.run([exe, cfg, out], capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(r.stderr)
            if "split_current only" in text:
                with gzip.open(os.path.join(folder, "split_current.json.gz")) as fh:
                    d = json.load(fh)
                d["current_density_time"] = d["dirac_current"] + d["non_dirac_current"]
                return d
            with open(out) as fh:
                return json.load(fh)

        for name, cfg_file, small, extra in REF_LEGS:
            print(f"[bench] reference model {name}", file=sys.stderr, flush=True)
            with open(os.path.join(ROOT, "higherharmonics_amd", "configs", cfg_file)) as fh:
                base = fh.read() + extra
            d = run(base + f"timing_repeats {repeats}\n", name)
            runs = d["solve_ms_runs"][1:] or d["solve_ms_runs"]
            ms = float(np.mean(runs))
            n_traj = d["trajectories"]
            cfg = R.parse(base)
            st = R.setup(cfg)
            N = st["time"].n_measurements
            if cfg["system_type"] == "Dirac":
                flops = n_traj * N * F_MOMENT_STEP
            else:
                key = ("piflux3d_eq" if float(cfg["diagonal_relaxation_time"]) > 0 else "piflux3d") \
                    if cfg["system_type"] == "PiFlux" else "honeycomb_eq"
                flops = d["attempted_steps"] * (F_STEP + 13 * F_RHS_REF[key])
            # audit: the same configuration on a small grid, driver vs oracle
            a = run(base + small, name + "_audit")
            sa = R.setup(R.parse(base + small))
            j_ref, _ = O.current_density(sa["model"], sa["pulse"], sa["time"], sa["solver"], sa["kx"], sa["ky"],
                                         sa["w"], kz=sa["kz"], n_threads=threads)
            j_fast, _ = O.current_density(sa["model"], sa["pulse"], sa["time"], sa["solver"], sa["kx"], sa["ky"],
                                          sa["w"], kz=sa["kz"], n_threads=threads, fast=True)
            audit = rel_l2(np.asarray(a["current_density_time"]), j_ref)
            # the rounding sensitivity of the configuration itself: the oracle's
            # strict build against its -ffast-math build
            spread = rel_l2(j_fast, j_ref)
            # CPU: the optimised oracle on a seeded sample of the full grid
            rng = np.random.default_rng(20261203)
            pilot = rng.choice(len(st["kx"]), size=min(len(st["kx"]), threads * 64), replace=False)
            t0 = time.perf_counter()
            O.current_density(st["model"], st["pulse"], st["time"], st["solver"], st["kx"][pilot], st["ky"][pilot],
                              st["w"][pilot], kz=st["kz"][pilot], n_threads=threads, fast=True)
            per = (time.perf_counter() - t0) / len(pilot)
            n = int(min(len(st["kx"]), max(len(pilot), cpu_s / max(per, 1e-9))))
            sel = rng.choice(len(st["kx"]), size=n, replace=False)
            t0 = time.perf_counter()
            O.current_density(st["model"], st["pulse"], st["time"], st["solver"], st["kx"][sel], st["ky"][sel],
                              st["w"][sel], kz=st["kz"][sel], n_threads=threads, fast=True)
            cpu = n / (time.perf_counter() - t0)
            res[name] = {"trajectories": n_traj, "samples": N + 1, "kernel_ms": ms,
                         "trajectories_per_s": n_traj / (ms * 1e-3), "attempts": d["attempted_steps"],
                         "rejects": d.get("rejects"), "fp64_tflops": flops / (ms * 1e-3) / 1e12,
                         "audit_rel_l2_vs_oracle": audit, "audit_trajectories": a["trajectories"],
                         "audit_oracle_strict_vs_fast": spread, "audit_bar": max(1e-9, 10 * spread),
                         "cpu_trajectories_per_s": cpu, "cpu_sample": n, "cpu_threads": threads,
                         "gpu_over_cpu": n_traj / (ms * 1e-3) / cpu}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--small", action="store_true", help="256x128 grid (profiling runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--extras", action="store_true", help="also time Magnus-4 and the carry ablation")
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--profile-leg", choices=["plain", "learned"], default=None,
                    help="time only this leg (profiler runs)")
    ap.add_argument("--net-out", default=None, help="save the trained network (.npz)")
    ap.add_argument("--net-in", default=None, help="--profile-leg learned: load the network instead of training")
    ap.add_argument("--infer-every", type=int, default=1,
                    help="learned leg: network proposal every this many sampling intervals")
    ap.add_argument("--sweeps", action="store_true", help="also time reduced configurations 3-5")
    ap.add_argument("--config3-full", action="store_true",
                    help="only configuration 3 at full size (33.5M trajectories, one launch; one JSON object)")
    ap.add_argument("--config4", action="store_true",
                    help="also run the honeycomb learned-vs-plain configuration (16 pulses on 768 x 768)")
    ap.add_argument("--config5", action="store_true",
                    help="also run the mixed 128-pulse sweep, plain vs learned + grouping, with a 1%% audit")
    ap.add_argument("--reference-models-only", action="store_true",
                    help="only the reference systems' leg (one JSON object)")
    ap.add_argument("--reference-models", action="store_true",
                    help="also run the reference's 3D Dirac, 3D pi-flux and honeycomb systems in its units")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 8 * n}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU; runs on several GPUs are not supported",
              file=sys.stderr)
        sys.exit(2)

    import higherharmonics_amd as H
    if args.reference_models_only:
        print(json.dumps({"reference_models": reference_models_leg(H)}))
        return
    if args.config3_full:
        print(json.dumps({"config3_full": config3_full_leg(H)}))
        return
    wl = workload(args.small)
    K = wl["nx"] * wl["ny"]
    f_traj_rhs = F_STEP + 13 * F_RHS["dirac_t2"]
    vertex = K * (wl["N"] + 1) * F_VERTEX["dirac"]

    if args.profile_leg:
        # one leg only, for profiler runs: no tight reference, no CPU baseline
        net = None
        if args.profile_leg == "learned":
            net = H.MLP.load(args.net_in) if args.net_in else train_proposal(H, wl)[0]
        plan = make_plan(H, wl, proposal=H.PROPOSAL_LEARNED if net else 0, mlp=net, infer_every=args.infer_every,
                         persistent=1)
        wall, ms, st, _ = time_plan(plan, args.steps, args.warmup)
        plan.close()
        print(json.dumps({"profile_leg": args.profile_leg, "trajectories_per_s": args.steps * K / wall,
                          "kernel_ms": ms, "attempts": st["attempts"]}))
        return

    print(f"[bench] Dirac-2D {wl['nx']}x{wl['ny']} N={wl['N']} RKF78 1e-10 restart (plain), T2 dephasing",
          file=sys.stderr, flush=True)
    plan = make_plan(H, wl, persistent=1)
    tuned = plan.autotune() if args.autotune else None
    wall_p, ms_p, st_p, j_plain = time_plan(plan, args.steps, args.warmup)
    plan.close()
    flops_p = st_p["attempts"] * f_traj_rhs + vertex

    print("[bench] step logs + proposal network training", file=sys.stderr, flush=True)
    net, train = train_proposal(H, wl)
    if args.net_out:
        net.save(args.net_out)
    print(f"[bench] learned proposal solve ({train['records']} records, loss {train['loss_first']:.3g} -> "
          f"{train['loss_last']:.3g})", file=sys.stderr, flush=True)
    plan = make_plan(H, wl, proposal=H.PROPOSAL_LEARNED, mlp=net, infer_every=args.infer_every, persistent=1)
    wall_l, ms_l, st_l, j_learned = time_plan(plan, args.steps, args.warmup)
    plan.close()
    flops_l = st_l["attempts"] * f_traj_rhs + vertex
    mlp_flops = K * ((wl["N"] + args.infer_every - 1) // args.infer_every) * F_MLP
    infer_sweep = {}
    if args.extras:
        for m in (1, 2, 4, 8, 16):
            if m == args.infer_every:
                continue
            p = make_plan(H, wl, proposal=H.PROPOSAL_LEARNED, mlp=net, infer_every=m, persistent=1)
            w2, sm, st, jm = time_plan(p, max(1, args.steps // 2), 1)
            p.close()
            infer_sweep[m] = {"trajectories_per_s": max(1, args.steps // 2) * K / w2, "kernel_ms": sm,
                              "attempts": st["attempts"], "j": jm}

    # the reference: rtol = atol = 1e-13 and a first step of Delta/500 every
    # interval, a step sequence of its own (at 1e-13 the Delta/50 restart
    # sequence would repeat the plain leg's steps exactly on this workload)
    print("[bench] tight-tolerance reference solve (rtol = atol = 1e-13, restart at Delta/500)", file=sys.stderr,
          flush=True)
    plan = make_plan(H, wl, rtol=1e-13, n_sub=500, persistent=1)
    j_tight = plan.execute()[0]
    plan.close()
    err_p, err_l = rel_l2(j_plain, j_tight), rel_l2(j_learned, j_tight)
    for m, r in infer_sweep.items():
        r["rel_l2_err_vs_tight"] = rel_l2(r.pop("j"), j_tight)
    net.close()

    v_plain = args.steps * K / wall_p
    v_learned = args.steps * K / wall_l
    use_learned = err_l <= 1e-9
    value = v_learned if use_learned else v_plain
    solve_ms, flops = (ms_l, flops_l) if use_learned else (ms_p, flops_p)
    achieved = flops / (solve_ms * 1e-3) / 1e12
    out = {
        "metric": "k-point trajectories/s at <=1e-9 rel j(t) err; learned-proposal vs plain RKF78",
        "value": value,
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * (wall_l if use_learned else wall_p) / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic pulse, Gauss-Legendre k-grid)",
        "config": {"workload": f"Dirac-2D {wl['nx']}x{wl['ny']} GL, cos^2 A0=0.5 8 cycles, T2=10 cycles, "
                               f"4096 samples, RKF78 rtol=atol=1e-10", "trajectories": K,
                   "value_leg": "learned" if use_learned else "plain (learned leg missed 1e-9)",
                   "launch": tuned},
        "legs": {
            "plain": {"trajectories_per_s": v_plain, "kernel_ms": ms_p, "attempts": st_p["attempts"],
                      "rejects": st_p["rejects"], "rel_l2_err_vs_tight": err_p,
                      "fp64_tflops": flops_p / (ms_p * 1e-3) / 1e12},
            "learned": {"trajectories_per_s": v_learned, "kernel_ms": ms_l, "attempts": st_l["attempts"],
                        "rejects": st_l["rejects"], "fallbacks": st_l["proposal_fallbacks"],
                        "rel_l2_err_vs_tight": err_l, "fp64_tflops": flops_l / (ms_l * 1e-3) / 1e12,
                        "mlp_bf16_tflops": mlp_flops / (ms_l * 1e-3) / 1e12, "infer_every": args.infer_every,
                        "training": train},
            "learned_infer_every_sweep": infer_sweep or None,
            "learned_vs_plain": v_learned / v_plain,
            "tight_reference": "same kernel, restart controller at Delta/500, rtol=atol=1e-13",
        },
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "compute_unit": "FP64 vector ALU (MI355X's FP64 peak is 78.6 TFLOP/s for vector and matrix alike)",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": (TRAFFIC_BYTES_LEARNED if use_learned else TRAFFIC_BYTES) if not args.small else None,
                     "kernel_ms": solve_ms, "algorithmic_flop_per_launch": flops,
                     "hardware_fp64_tflops": None if args.small else
                     (HW_FP64_FLOP_LEARNED if use_learned else HW_FP64_FLOP) / (solve_ms * 1e-3) / 1e12,
                     "kernel": "learned-proposal solve" if use_learned else "plain solve"},
    }
    if args.extras:
        extras = {}
        for name, kw in [("magnus4_T2inf_dt_T20480", dict(method=1, relax=0, magnus_substeps=5)),
                         ("rkf78_T2inf", dict(method=0, relax=0)),
                         ("rkf78_carry_ablation", dict(method=0, relax=2, proposal=1))]:
            p = make_plan(H, wl, persistent=1, **kw)
            w2, sm, st, _ = time_plan(p, max(1, args.steps // 2), 1)
            p.close()
            extras[name] = {"trajectories_per_s": max(1, args.steps // 2) * K / w2, "kernel_ms": sm,
                            "attempts": st["attempts"]}
        out["extras"] = extras
    if args.sweeps:
        out["sweeps"] = sweep_extras(H)
    if args.config4:
        out["config4_honeycomb_learned"] = config4_leg(H)
    if args.config5:
        out["config5_mixed_learned"] = config5_leg(H)
    if args.reference_models:
        out["reference_models"] = reference_models_leg(H)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl)
        print("[bench] 128 x 64 grid in full on GPU and CPU", file=sys.stderr, flush=True)
        out["config1_full"] = config1_full(H)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
