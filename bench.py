#!/usr/bin/env python3
"""Throughput of the Bloch-equation hot path on one MI355X.

Headline workload (one "step" = one full solve of all 16 jobs):
  honeycomb lattice d = t sum_delta e^{i(k - A).delta}, t = a = 1, over the
  hexagonal Brillouin zone on a 768 x 768 Gauss-Legendre grid (589,824
  k-points); 16 seeded cos^2 pulses (10 cycles plus one cycle of free
  evolution, A0 ~ U[0.1, 1.5], CEP ~ U[0, 2 pi], seed 20261203), 512 samples
  per cycle (5,632 intervals); FP64 RKF78 at rtol = atol = 1e-10.
  9,437,184 trajectories per step.

Legs on that workload: the learned proposal (a BF16 network trained on the
GPU from step logs of the same pulses on a 96 x 96 grid proposes every
trajectory's first step; later intervals start from the controller's carried
step; the embedded-error controller judges every step), the plain controller
(the reference's restart at Delta/50 every interval) and the carried
controller (the non-learned alternative).  ``value`` is the learned leg's trajectories/s when its j(t),
on a seeded 1 % of every pulse's k-points, is within 1e-9 relative L2 of the
CPU oracle at rtol = atol = 1e-13; else the plain leg's.

Prints one JSON line: trajectories/s, the FP64 roofline of the value leg's
solve kernel (algorithmic FLOPs from the attempted-step counter over the
HIP-event time of the kernel on its own stream), and the CPU oracle timed on
a bounded sample of the same workload on the host cores.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
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

PEAK_FP64_TFLOPS = 78.6  # MI355X vector FP64

# Algorithmic FLOPs (convention: + - * = 1, FMA = 2, / sqrt = 1, each
# sin/cos/exp = 1).  RKF78 step algebra: 55 tableau FMAs x 3 components +
# stage updates + 8th-order sum + embedded error + controller.
F_STEP = 450.0
# Honeycomb right-hand side as the phase-table kernels evaluate it: sin and cos
# of (kx - A)/2 by angle addition from the tabulated (sin A/2, cos A/2) (6),
# w = 2 t (Re g, -Im g) from them with a = 2t cos, b = 2t sin: a (cos + 2 cy)
# - b sin and b (2 cy - 2 cos) (9), w x s with w_z = 0 (5).
F_RHS = {"dirac_t2": 47.0, "dirac": 23.0, "piflux": 26.0, "honeycomb": 20.0}
F_VERTEX = {"dirac": 3.0, "piflux": 4.0, "honeycomb": 20.0}
F_MAGNUS = 85.0
# One network inference per lane: 2 x (8*64 + 2*64*64 + 64*2) FLOPs
# (BF16 MFMA hidden layers, FP32 output layer), not counted as FP64 work.
F_MLP = 2.0 * (8 * 64 + 2 * 64 * 64 + 64 * 2)

HEADLINE = dict(n=768, n_pulses=16, per_cycle=512, log_n=96)
# Measured per launch of the headline's learned solve kernel by rocprofv3 --pmc
# (profiles/solve_kernel_learned_honeycomb_768x768_16_pulses_counters.md):
# HBM bytes (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, KiB per dispatch, the factors
# a plain 1 GiB copy reads) and the hardware FP64 FLOP (SQ_INSTS_VALU_FMA_F64 x 2
# + MUL + ADD + TRANS, x 64 lanes).
TRAFFIC_BYTES_HEADLINE = (271527.3125 * 2.0 + 26376256.0) * 1024
HW_FP64_FLOP_HEADLINE = (2 * 239417339958 + 93898906842 + 30749147136 + 32467362) * 64


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


def oracle_job(O, job):
    """Oracle descriptions (model, pulse, time) of one workload job."""
    m, p = job["model"], job["pulse"]
    model = O.model(kind=m.get("kind", 0), v=m.get("v", 1.0), mass=m.get("mass", 0.1), hopping=m.get("hopping", 1.0),
                    relax=m.get("relax", 0), relax_time=m.get("relax_time", 0.0))
    pulse = O.pulse(p.get("kind", 0), p.get("amp_a", 0.5), amp_b=p.get("amp_b", 0.0), n_cycles=p.get("n_cycles", 8.0),
                    sigma=p.get("sigma", 0.0), t0=p.get("t0", 0.0), delay=p.get("delay", 0.0), cep=p.get("cep", 0.0))
    return model, pulse, O.time_grid(*job["time"])


def run_leg(H, W, jobs, solver, steps, warmup, **launch):
    """Times `steps` executions of one plan after `warmup` untimed ones: host
    wall clock per solve and the mean HIP-event time of the solve kernel."""
    plan = W.build_plan(H, jobs, solver, persistent=1, **launch)
    try:
        for _ in range(warmup):
            plan.execute()
        kernel_ms = []
        t0 = time.perf_counter()
        js = None
        for _ in range(steps):
            js = plan.execute()
            kernel_ms.append(plan.stats.solve_ms)
        wall = time.perf_counter() - t0
        st = plan.stats.as_dict()
    finally:
        plan.close()
    return {"wall_s": wall, "kernel_ms": float(np.mean(kernel_ms)), "stats": st, "j": js}


def train_on_jobs(H, W, jobs, rtol=1e-10):
    """The step-proposal network trained on step logs of the jobs (typically the
    workload's pulses on a coarser grid): a carried-controller round, then a
    round that adds the logs of a learned-proposal solve with that network."""
    from higherharmonics_amd import learn
    cap = 1 << 21
    stride = max(1, -(-sum(j["grid"][1] * j["grid"][2] * j["time"][2] for j in jobs) // cap))
    t0 = time.perf_counter()
    net, rec, hist = learn.train_on_solver_logs(
        lambda sv: W.build_plan(H, jobs, H.Solver(rtol=rtol, atol=rtol, proposal=sv.proposal, mlp=sv.mlp),
                                persistent=1), capacity=cap, stride=stride, rounds=2)
    return net, {"records": int(len(rec)), "log_stride": stride, "seconds": time.perf_counter() - t0,
                 "rounds": len(hist), "loss_first": float(hist[-1][0]), "loss_last": float(hist[-1][-1]),
                 "params": net.n_params}


def oracle_audit(H, W, jobs, solvers, frac=0.01, threads=None):
    """Each solver's j(t) on a seeded random `frac` of every job's k-points
    (explicit grids with the original weights, the same points on both
    sides) against the CPU oracle at rtol = atol = 1e-13 with the carried step
    (whole-interval steps, stopping at the envelope corners).  Returns the
    largest relative L2 error over the jobs per solver."""
    import oracle as O
    sub = W.audit_subset(H, jobs, frac)
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    gpu = {}
    for name, sv in solvers.items():
        plan = W.build_plan(H, sub, sv, persistent=1)
        try:
            gpu[name] = plan.execute()
        finally:
            plan.close()
    t0 = time.perf_counter()
    ref = []
    for i, job in enumerate(sub):
        m, p, tc = oracle_job(O, job)
        g = job["grid"]
        j, _ = O.current_density(m, p, tc, O.solver(rtol=1e-13, atol=1e-13, proposal=O.PROPOSAL_CARRY),
                                 g[1], g[2], g[3], n_threads=threads, fast=False)
        ref.append(j)
        print(f"[bench] oracle audit job {i + 1}/{len(sub)}: {len(g[1])} k-points, "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    cpu_s = time.perf_counter() - t0
    out = {"k_points": W.trajectories(sub), "fraction": frac, "oracle": "oracle (strict IEEE build), "
           "rtol = atol = 1e-13, carried step", "oracle_seconds": cpu_s, "oracle_threads": threads}
    for name, js in gpu.items():
        out[f"max_rel_l2_{name}_vs_oracle"] = max(rel_l2(a, b) for a, b in zip(js, ref))
    return out


def cpu_baseline_jobs(jobs, target_s=10.0, seed=20261203):
    """The oracle's optimised build with the reference's controller (restart at
    Delta/50 every interval, rtol = atol = 1e-10) on a seeded random sample of
    the workload's (pulse, k-point) trajectories, OpenMP over the host cores
    (capped at 16)."""
    import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    g = jobs[0]["grid"]
    kx, ky, w = O.gl_hex(g[1], g[2]) if g[0] == "hex" else O.gl_box(*g[1:])
    rng = np.random.default_rng(seed)
    sv = O.solver(rtol=1e-10, atol=1e-10)

    def run(n_per_job):
        t0 = time.perf_counter()
        for job in jobs:
            m, p, tc = oracle_job(O, job)
            sel = rng.choice(len(kx), size=n_per_job, replace=False)
            O.current_density(m, p, tc, sv, kx[sel], ky[sel], w[sel], n_threads=threads, fast=True)
        return time.perf_counter() - t0

    pilot = max(64, threads * 16)
    per = run(pilot) / (pilot * len(jobs))
    n = int(max(pilot, min(len(kx), target_s / max(per, 1e-9) / len(jobs))))
    wall = run(n)
    total = n * len(jobs)
    return {"value": total / wall, "unit": "trajectories/s", "cores": threads, "kind": "port",
            "sample": f"{n} seeded-random k-points of each of the {len(jobs)} pulses on the {g[1]}x{g[2]} grid "
                      f"({total} trajectories, seed {seed}), oracle -O3 -ffast-math build with the reference's "
                      f"restart controller, OpenMP {threads} threads, {wall:.1f} s"}


def leg_summary(leg, K, f_attempt, vertex_flop):
    st = leg["stats"]
    flop = st["attempts"] * f_attempt + vertex_flop
    out = {"trajectories_per_s": K * leg["steps"] / leg["wall_s"], "kernel_ms": leg["kernel_ms"],
           "attempts": st["attempts"], "rejects": st["rejects"],
           "fp64_tflops": flop / (leg["kernel_ms"] * 1e-3) / 1e12, "algorithmic_flop": flop}
    if st.get("proposal_fallbacks"):
        out["fallbacks"] = st["proposal_fallbacks"]
    if st.get("wave_attempts"):
        out["wave_attempts"] = st["wave_attempts"]
    return out


def headline(H, steps, warmup, audit_frac=0.01, per_cycle=None):
    """Configuration 4: learned (value leg), plain and carried legs, the
    network's training record, the oracle audit and a full-grid tight solve
    of the same kernel as a second check."""
    from higherharmonics_amd import workloads as W
    cfg = dict(HEADLINE, per_cycle=per_cycle or HEADLINE["per_cycle"])
    jobs = W.honeycomb_cos2(n=cfg["n"], n_pulses=cfg["n_pulses"], per_cycle=cfg["per_cycle"])
    K = W.trajectories(jobs)
    small = W.honeycomb_cos2(n=cfg["log_n"], n_pulses=cfg["n_pulses"], per_cycle=cfg["per_cycle"])
    f_attempt = F_STEP + 13 * F_RHS["honeycomb"]
    vertex = sum((j["time"][2] + 1) * j["grid"][1] * j["grid"][2] for j in jobs) * F_VERTEX["honeycomb"]
    print(f"[bench] headline: honeycomb hex {cfg['n']}^2, {cfg['n_pulses']} cos^2 pulses, {cfg['per_cycle']} "
          f"samples/cycle, {K} trajectories; training the proposal network", file=sys.stderr, flush=True)
    net, train = train_on_jobs(H, W, small)
    sv = {"learned": H.Solver(proposal=H.PROPOSAL_LEARNED, mlp=net), "plain": H.Solver(),
          "carry": H.Solver(proposal=H.PROPOSAL_CARRY)}
    legs = {}
    for name, k, w in [("learned", steps, warmup), ("carry", 3, 1), ("plain", 2, 1)]:
        print(f"[bench] headline leg {name}: {w} warmup + {k} timed solves", file=sys.stderr, flush=True)
        legs[name] = dict(run_leg(H, W, jobs, sv[name], k, w), steps=k)
    print("[bench] headline learned leg with per-block LDS accumulation (block_reduce 2): 1 warmup + 2 timed solves",
          file=sys.stderr, flush=True)
    blk = dict(run_leg(H, W, jobs, sv["learned"], 2, 1, block_reduce=2), steps=2)
    print(f"[bench] oracle audit on {audit_frac:.2%} of every pulse's k-points", file=sys.stderr, flush=True)
    audit = oracle_audit(H, W, jobs, sv, audit_frac)
    print("[bench] full-grid tight solve (same kernel, rtol = atol = 1e-13, restart at Delta/500)", file=sys.stderr,
          flush=True)
    tight_jobs = [dict(j, time=(j["time"][0], j["time"][1], j["time"][2], 500)) for j in jobs]
    tight = run_leg(H, W, tight_jobs, H.Solver(rtol=1e-13, atol=1e-13), 1, 0)["j"]
    net.close()
    out = {"workload": f"honeycomb hex BZ GL {cfg['n']}x{cfg['n']}, {cfg['n_pulses']} seeded cos^2 pulses "
                       f"(10 cycles + 1, A0 ~ U[0.1, 1.5], CEP ~ U[0, 2 pi]), {cfg['per_cycle']} samples/cycle, "
                       f"RKF78 rtol=atol=1e-10", "trajectories": K, "jobs": len(jobs),
           "samples_per_job": jobs[0]["time"][2] + 1, "training": train, "audit": audit, "legs": {}}
    for name, leg in legs.items():
        s = leg_summary(leg, K, f_attempt, vertex)
        s["rel_l2_vs_oracle_tight_audit"] = audit[f"max_rel_l2_{name}_vs_oracle"]
        s["rel_l2_vs_same_kernel_tight_full_grid"] = max(rel_l2(a, b) for a, b in zip(leg["j"], tight))
        out["legs"][name] = s
    b = leg_summary(blk, K, f_attempt, vertex)
    b["bitwise_equal_to_learned_leg"] = all(np.array_equal(x, y) for x, y in zip(blk["j"], legs["learned"]["j"]))
    b["hbm_write_kib_per_launch"] = {"block_reduce_2": 737608.0, "per_wave_atomics": 26376256.0,
                                     "source": "rocprofv3 --pmc WRITE_SIZE, one launch each"}
    out["learned_block_accumulation"] = b
    L = out["legs"]
    out["learned_vs_plain"] = L["learned"]["trajectories_per_s"] / L["plain"]["trajectories_per_s"]
    out["learned_vs_carry"] = L["learned"]["trajectories_per_s"] / L["carry"]["trajectories_per_s"]
    out["_wall_learned"] = legs["learned"]["wall_s"]
    return out


def uneven_steps_leg(H, n=768, n_pulses=16, per_cycle=32, log_n=96, audit_frac=0.01):
    """The headline's pulses sampled at 32 per cycle: whole-interval steps are
    error-limited and the busiest k-points take about twice the attempts of
    the quietest, so a wave issues as many attempts per interval as its
    busiest lane.  Legs: plain, carried, learned (first-interval proposal),
    the network proposing every interval's first step, and learned with
    step-count grouping by the network's cost estimate (64 cost classes in
    grid order, and a full sort by the estimate); each counts the attempts
    its waves issue; oracle audit of the learned and carried legs."""
    from higherharmonics_amd import workloads as W
    jobs = W.honeycomb_cos2(n=n, n_pulses=n_pulses, per_cycle=per_cycle)
    small = W.honeycomb_cos2(n=log_n, n_pulses=n_pulses, per_cycle=per_cycle)
    K = W.trajectories(jobs)
    f_attempt = F_STEP + 13 * F_RHS["honeycomb"]
    vertex = sum((j["time"][2] + 1) * j["grid"][1] * j["grid"][2] for j in jobs) * F_VERTEX["honeycomb"]
    print(f"[bench] uneven-step leg: {per_cycle} samples/cycle, {K} trajectories", file=sys.stderr, flush=True)
    net, train = train_on_jobs(H, W, small)
    legs = {}
    every = H.Solver(proposal=H.PROPOSAL_LEARNED, mlp=net, infer_policy=H.INFER_SCHEDULED)
    for name, sv, kw in [("plain", H.Solver(), {}), ("carry", H.Solver(proposal=H.PROPOSAL_CARRY), {}),
                         ("learned", H.Solver(proposal=H.PROPOSAL_LEARNED, mlp=net), {}),
                         ("learned_every_interval", every, {}),
                         ("learned_grouped", H.Solver(proposal=H.PROPOSAL_LEARNED, mlp=net), {"group_by_cost": 1}),
                         ("learned_grouped_full_sort", H.Solver(proposal=H.PROPOSAL_LEARNED, mlp=net),
                          {"group_by_cost": 1 << 30})]:
        print(f"[bench] uneven-step leg {name}", file=sys.stderr, flush=True)
        leg = dict(run_leg(H, W, jobs, sv, 3, 1, count_wave_attempts=1, **kw), steps=3)
        legs[name] = leg_summary(leg, K, f_attempt, vertex)
    audit = oracle_audit(H, W, jobs, {"learned": H.Solver(proposal=H.PROPOSAL_LEARNED, mlp=net),
                                      "learned_every_interval": every,
                                      "carry": H.Solver(proposal=H.PROPOSAL_CARRY)}, audit_frac)
    net.close()
    return {"workload": f"headline pulses on hex {n}x{n} at {per_cycle} samples/cycle, RKF78 rtol=atol=1e-10",
            "trajectories": K, "training": train, "legs": legs, "audit": audit,
            "learned_vs_carry": legs["learned"]["trajectories_per_s"] / legs["carry"]["trajectories_per_s"],
            "learned_every_interval_vs_carry": legs["learned_every_interval"]["trajectories_per_s"] /
            legs["carry"]["trajectories_per_s"],
            "learned_grouped_vs_carry": legs["learned_grouped"]["trajectories_per_s"] /
            legs["carry"]["trajectories_per_s"],
            "wave_attempts_learned_grouped_over_carry": legs["learned_grouped"]["wave_attempts"] /
            legs["carry"]["wave_attempts"],
            "learned_grouped_full_sort_vs_carry": legs["learned_grouped_full_sort"]["trajectories_per_s"] /
            legs["carry"]["trajectories_per_s"],
            "wave_attempts_learned_every_interval_over_carry": legs["learned_every_interval"]["wave_attempts"] /
            legs["carry"]["wave_attempts"],
            "rejects_learned_every_interval_over_carry": legs["learned_every_interval"]["rejects"] /
            max(1, legs["carry"]["rejects"])}


def config2_leg(H, steps=3):
    """BASELINE configuration 2 (Dirac-2D 1024 x 512, T2 dephasing, 4096
    samples): plain, carried and learned legs (network trained on step logs of
    the 256 x 128 grid of the same box) and a tight solve of the same kernel."""
    wl = workload(False)
    K = wl["nx"] * wl["ny"]
    f_attempt = F_STEP + 13 * F_RHS["dirac_t2"]
    vertex = K * (wl["N"] + 1) * F_VERTEX["dirac"]
    print("[bench] configuration 2 (Dirac-2D 1024x512)", file=sys.stderr, flush=True)
    net, train = train_proposal(H, wl)
    out = {"workload": "Dirac-2D 1024x512 GL, cos^2 A0=0.5 8 cycles, T2=10 cycles, 4096 samples, RKF78 1e-10",
           "trajectories": K, "training": train, "legs": {}}
    js = {}
    for name, kw in [("plain", {}), ("carry", dict(proposal=H.PROPOSAL_CARRY)),
                     ("learned", dict(proposal=H.PROPOSAL_LEARNED, mlp=net))]:
        plan = make_plan(H, wl, persistent=1, **kw)
        wall, ms, st, js[name] = time_plan(plan, steps, 1)
        plan.close()
        leg = {"wall_s": wall, "kernel_ms": ms, "stats": st, "steps": steps}
        out["legs"][name] = leg_summary(leg, K, f_attempt, vertex)
    plan = make_plan(H, wl, rtol=1e-13, n_sub=500, persistent=1)
    j_tight = plan.execute()[0]
    plan.close()
    net.close()
    for name in js:
        out["legs"][name]["rel_l2_vs_same_kernel_tight"] = rel_l2(js[name], j_tight)
    L = out["legs"]
    out["learned_vs_plain"] = L["learned"]["trajectories_per_s"] / L["plain"]["trajectories_per_s"]
    out["learned_vs_carry"] = L["learned"]["trajectories_per_s"] / L["carry"]["trajectories_per_s"]
    return out


def config5_leg(H, n=256, log_n=32, audit_frac=0.01):
    """64 honeycomb (double cos^2) + 64 PiFlux (double Gaussian) seeded pulse
    configurations on n x n grids in one plan.  Legs: the plain controller in
    one fused launch over both groups and in one launch per group, the carried
    controller, the learned proposal, and the learned proposal with step-count
    grouping; every plan autotuned and graph-captured, one warmup and one
    timed execution.  The network is trained on step logs of the same jobs on
    log_n x log_n grids.  Audit: a seeded 1 % of every job's k-points against
    the CPU oracle at rtol = atol = 1e-13."""
    from higherharmonics_amd import workloads as W
    jobs = W.mixed_sweep(n=n)
    K = W.trajectories(jobs)
    small = W.mixed_sweep(n=log_n)
    print(f"[bench] configuration 5: {len(jobs)} jobs, {K} trajectories; training the proposal network",
          file=sys.stderr, flush=True)
    net, train = train_on_jobs(H, W, small)
    res = {"workload": f"64 honeycomb double-cos^2 + 64 PiFlux double-Gaussian pulses, GL {n}x{n} each, "
                       f"256 samples/cycle, RKF78 rtol=atol=1e-10", "jobs": len(jobs), "trajectories": K,
           "training": train, "legs": {}}
    learned = H.Solver(proposal=H.PROPOSAL_LEARNED, mlp=net)
    for name, sv, kw in [("plain", H.Solver(), {}), ("plain_separate_launches", H.Solver(), {"separate_launches": 1}),
                         ("carry", H.Solver(proposal=H.PROPOSAL_CARRY), {}), ("learned", learned, {}),
                         ("learned_grouped", learned, {"group_by_cost": 1})]:
        print(f"[bench] configuration 5 leg {name}", file=sys.stderr, flush=True)
        plan = W.build_plan(H, jobs, sv, persistent=1, count_wave_attempts=1, **kw)
        try:
            tuned = plan.autotune()
            plan.capture_graph()
            plan.execute()
            t1 = time.perf_counter()
            plan.execute()
            wall = time.perf_counter() - t1
            st = plan.stats.as_dict()
            launches = plan.solve_launches
        finally:
            plan.close()
        res["legs"][name] = {"trajectories_per_s": K / wall, "wall_s": wall, "solve_ms": st["solve_ms"],
                             "solve_launches": launches, "attempts": st["attempts"], "rejects": st["rejects"],
                             "wave_attempts": st["wave_attempts"], "fallbacks": st["proposal_fallbacks"],
                             "launch": tuned}
    L = res["legs"]
    res["learned_vs_plain"] = L["learned"]["trajectories_per_s"] / L["plain"]["trajectories_per_s"]
    res["learned_vs_carry"] = L["learned"]["trajectories_per_s"] / L["carry"]["trajectories_per_s"]
    res["fused_vs_separate_launches"] = L["plain"]["trajectories_per_s"] / \
        L["plain_separate_launches"]["trajectories_per_s"]
    print(f"[bench] configuration 5 oracle audit on {audit_frac:.2%} of every job's k-points", file=sys.stderr,
          flush=True)
    res["audit"] = oracle_audit(H, W, jobs, {"learned": learned, "plain": H.Solver(),
                                             "carry": H.Solver(proposal=H.PROPOSAL_CARRY)}, audit_frac)
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
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the 128 x 64 CPU/GPU case")
    ap.add_argument("--audit-frac", type=float, default=0.01, help="fraction of every pulse's k-points in the audit")
    ap.add_argument("--headline-only", action="store_true", help="only the headline (no configuration 2, no "
                    "uneven-step leg)")
    ap.add_argument("--profile-leg", choices=["plain", "learned", "carry"], default=None,
                    help="only time this leg of the headline (profiler runs: no audit, no CPU)")
    ap.add_argument("--net-out", default=None, help="save the headline's trained network (.npz)")
    ap.add_argument("--block-reduce", type=int, default=0, choices=[0, 1, 2],
                    help="--profile-leg: the plan's block_reduce (2 = per-block LDS accumulation)")
    ap.add_argument("--net-in", default=None, help="--profile-leg learned: load the network instead of training")
    ap.add_argument("--sweeps", action="store_true", help="also time reduced configurations 3-5")
    ap.add_argument("--config3-full", action="store_true",
                    help="only configuration 3 at full size (33.5M trajectories, one launch; one JSON object)")
    ap.add_argument("--config5-only", action="store_true",
                    help="only the mixed 128-pulse sweep (configuration 5; one JSON object)")
    ap.add_argument("--uneven-only", action="store_true",
                    help="only the uneven-step leg (the headline's pulses at 32 samples per cycle; one JSON object)")
    ap.add_argument("--profile-config5", choices=["plain", "plain_separate_launches", "carry"], default=None,
                    help="only time this configuration-5 leg (autotuned, graph-captured; profiler runs)")
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
        print("bench.py measures one GPU; runs on several GPUs are not supported", file=sys.stderr)
        sys.exit(2)

    import higherharmonics_amd as H
    if args.reference_models_only:
        print(json.dumps({"reference_models": reference_models_leg(H)}))
        return
    if args.config3_full:
        print(json.dumps({"config3_full": config3_full_leg(H)}))
        return
    if args.profile_config5:
        from higherharmonics_amd import workloads as W
        jobs = W.mixed_sweep()
        leg = args.profile_config5
        sv = H.Solver(proposal=H.PROPOSAL_CARRY) if leg == "carry" else H.Solver()
        plan = W.build_plan(H, jobs, sv, persistent=1, separate_launches=int(leg == "plain_separate_launches"))
        tuned = plan.autotune()
        plan.capture_graph()
        for _ in range(args.warmup):
            plan.execute()
        t0 = time.perf_counter()
        ms = []
        for _ in range(args.steps):
            plan.execute()
            ms.append(plan.stats.solve_ms)
        wall = time.perf_counter() - t0
        print(json.dumps({"profile_config5": leg, "solve_launches": plan.solve_launches, "launch": tuned,
                          "trajectories_per_s": args.steps * W.trajectories(jobs) / wall, "graph_ms": ms}))
        plan.close()
        return
    if args.config5_only:
        print(json.dumps({"config5_mixed_sweep": config5_leg(H, audit_frac=args.audit_frac)}))
        return
    if args.uneven_only:
        print(json.dumps({"uneven_steps_32_per_cycle": uneven_steps_leg(H, audit_frac=args.audit_frac)}))
        return
    if args.profile_leg:
        from higherharmonics_amd import workloads as W
        jobs = W.honeycomb_cos2(n=HEADLINE["n"], n_pulses=HEADLINE["n_pulses"], per_cycle=HEADLINE["per_cycle"])
        net = None
        if args.profile_leg == "learned" and args.net_in:
            net = H.MLP.load(args.net_in)
        elif args.profile_leg == "learned":
            small = W.honeycomb_cos2(n=HEADLINE["log_n"], n_pulses=HEADLINE["n_pulses"],
                                     per_cycle=HEADLINE["per_cycle"])
            net, _ = train_on_jobs(H, W, small)
        if args.net_out and net is not None:
            net.save(args.net_out)
        sv = {"learned": H.Solver(proposal=H.PROPOSAL_LEARNED, mlp=net), "plain": H.Solver(),
              "carry": H.Solver(proposal=H.PROPOSAL_CARRY)}[args.profile_leg]
        leg = run_leg(H, W, jobs, sv, args.steps, args.warmup, block_reduce=args.block_reduce)
        print(json.dumps({"profile_leg": args.profile_leg, "trajectories_per_s": args.steps * W.trajectories(jobs) /
                          leg["wall_s"], "kernel_ms": leg["kernel_ms"], "attempts": leg["stats"]["attempts"]}))
        return

    hl = headline(H, args.steps, args.warmup, args.audit_frac)
    L = hl["legs"]
    use_learned = L["learned"]["rel_l2_vs_oracle_tight_audit"] <= 1e-9
    vleg = "learned" if use_learned else "plain"
    hl.pop("_wall_learned")
    K = hl["trajectories"]
    achieved = L[vleg]["fp64_tflops"]
    kernel_ms = L[vleg]["kernel_ms"]
    out = {
        "metric": "k-point trajectories/s at <=1e-9 rel j(t) err; learned-proposal vs plain RKF78",
        "value": L[vleg]["trajectories_per_s"],
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * K / L[vleg]["trajectories_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic seeded pulses, Gauss-Legendre k-grid)",
        "config": {"workload": hl["workload"], "trajectories": K,
                   "value_leg": vleg if use_learned else "plain (learned leg missed 1e-9 against the oracle)",
                   "error_vs_oracle": L[vleg]["rel_l2_vs_oracle_tight_audit"],
                   "error_criterion": "max over the 16 pulses of the relative L2 error of j(t) on a seeded "
                                      f"{args.audit_frac:.2%} of the k-points against the CPU oracle at "
                                      "rtol = atol = 1e-13"},
        "legs": hl["legs"],
        "learned_vs_plain": hl["learned_vs_plain"],
        "learned_vs_carry": hl["learned_vs_carry"],
        "headline_detail": {k: v for k, v in hl.items() if k not in ("legs", "learned_vs_plain", "learned_vs_carry")},
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": TRAFFIC_BYTES_HEADLINE if use_learned else None,
                     "kernel": f"{vleg} solve kernel (honeycomb, cos^2)", "kernel_ms": kernel_ms,
                     "algorithmic_flop_per_launch": L[vleg]["algorithmic_flop"],
                     "hardware_fp64_tflops": (HW_FP64_FLOP_HEADLINE / (kernel_ms * 1e-3) / 1e12
                                              if use_learned and HW_FP64_FLOP_HEADLINE else None)},
    }
    if not args.headline_only:
        out["uneven_steps_32_per_cycle"] = uneven_steps_leg(H, audit_frac=args.audit_frac)
        out["config2_dirac_1024x512"] = config2_leg(H)
        out["config5_mixed_sweep"] = config5_leg(H, audit_frac=args.audit_frac)
    if args.sweeps:
        out["sweeps"] = sweep_extras(H)
    if args.reference_models:
        out["reference_models"] = reference_models_leg(H)
    if not args.no_cpu:
        print("[bench] CPU baseline (oracle, reference controller) on a sample of the headline workload",
              file=sys.stderr, flush=True)
        from higherharmonics_amd import workloads as W
        out["cpu_baseline"] = cpu_baseline_jobs(
            W.honeycomb_cos2(n=HEADLINE["n"], n_pulses=HEADLINE["n_pulses"], per_cycle=HEADLINE["per_cycle"]))
        print("[bench] 128 x 64 Dirac grid in full on GPU and CPU", file=sys.stderr, flush=True)
        out["config1_full"] = config1_full(H)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
