"""Bounded controller loops and failure reporting on the GPU, through the C
ABI: the forward-progress guard and the per-interval step cap against the
oracle, accumulator overflow counted apart from non-finite trajectories, and
the statistics of a device-side execution on a caller's stream."""
import ctypes
import math

import numpy as np
import pytest

import oracle as O
from tests.problems import Problem, oracle_j, rel_l2
from tests.test_oracle import NO_PROGRESS

pytestmark = pytest.mark.gpu


def _print_case(name, **extra):
    print(f"[case] {name} {extra}", flush=True)


def test_step_that_cannot_advance_t_fails_on_gpu(hhg):
    """At t ~ 1e10 the controller accepts only steps below half an ulp of t;
    every lane fails at its first such step (odeint would accept them
    forever), the kernel finishes and the counts, attempts and j equal the
    oracle's.  max_steps = 1000 bounds the interval a second way."""
    c = NO_PROGRESS
    pr = Problem("no_progress", model=dict(c["model"]), pulse=dict(c["pulse"]), time=c["time"],
                 grid=("box", 8, 8, 0.3, 0.5, 0.1, 0.3),
                 solver=dict(method=0, rtol=1e-10, atol=1e-10, proposal=0, max_steps=1000))
    _print_case(pr.name, t_begin=pr.time[0], v=pr.model["v"])
    j_ref, st_ref = oracle_j(pr)
    assert st_ref.failed_trajectories == 64
    with pr.h_plan(hhg) as plan:
        j = plan.execute(allow_failures=True)[0]
        st = plan.stats
        assert plan.status == hhg.ERROR_TOLERANCE
    print(f"  attempts gpu={st.attempts} oracle={st_ref.attempts} failed={st.failed_trajectories}", flush=True)
    assert st.failed_trajectories == 64 and st.nonfinite_trajectories == 0
    assert st.attempts == st_ref.attempts and st.attempts <= 64 * 50
    assert rel_l2(j, j_ref) <= 1e-13


def test_step_cap_per_interval_on_gpu(hhg):
    """The restart controller needs at least 4 accepted steps per interval at
    n_sub = 50: max_steps = 3 fails every trajectory after exactly 3 attempts
    (the oracle's count), and a cap that is never reached leaves j and the
    attempt count bitwise unchanged."""
    base = Problem("step_cap", model=dict(kind=O.MODEL_DIRAC2D, v=1.0, mass=0.1),
                   pulse=dict(kind=O.PULSE_COS2, amp_a=0.2, n_cycles=2.0, t0=-2 * math.pi),
                   time=(-2 * math.pi, 2 * math.pi, 400, 50), grid=("box", 16, 8, -3.0, 3.0, -3.0, 3.0),
                   solver=dict(method=0, rtol=1e-6, atol=1e-6, proposal=0))
    _print_case(base.name, max_steps=3)
    capped = Problem(base.name, base.model, base.pulse, base.time, base.grid, dict(base.solver, max_steps=3))
    _, st_ref = oracle_j(capped)
    assert st_ref.failed_trajectories == 128 and st_ref.attempts == 3 * 128
    with capped.h_plan(hhg) as plan:
        plan.execute(allow_failures=True)
        assert plan.status == hhg.ERROR_TOLERANCE
        assert plan.stats.failed_trajectories == 128 and plan.stats.attempts == 3 * 128
    with base.h_plan(hhg) as plan:
        j0 = plan.execute()[0]
        a0 = plan.stats.attempts
    loose = Problem(base.name, base.model, base.pulse, base.time, base.grid, dict(base.solver, max_steps=64))
    with loose.h_plan(hhg) as plan:
        j1 = plan.execute()[0]
        assert plan.stats.attempts == a0
    assert np.array_equal(j0, j1)


def test_accumulator_overflow_is_counted_apart(hhg):
    """Weights of 1e15 push the sample sums beyond the exact accumulators'
    range (2^46): the lanes holding them are counted as accumulator
    overflows, no trajectory is counted non-finite or failed, and the call
    reports NONFINITE."""
    from tests.problems import DIRAC_SMALL as pr
    _print_case("overflow", weight=1e15)
    kx, ky, w = pr.o_grid()
    grid = hhg.KGrid.explicit(kx[:128], ky[:128], np.full(128, 1e15))
    with hhg.Plan([hhg.Job(pr.h_model(hhg), pr.h_laser(hhg), pr.h_time(hhg), 0)], [grid], pr.h_solver(hhg)) as plan:
        plan.execute(allow_failures=True)
        st = plan.stats
        assert plan.status == hhg.ERROR_NONFINITE
    print(f"  overflows={st.accumulator_overflows}", flush=True)
    assert st.accumulator_overflows > 0
    assert st.nonfinite_trajectories == 0 and st.failed_trajectories == 0


def _hip():
    """The HIP runtime the library itself links (already loaded with it)."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return hip


def test_device_execute_reports_through_last_stats(hhg):
    """hhg_plan_execute_device on a caller's non-blocking stream returns at
    once; hhg_plan_last_stats and hhg_plan_job_failures wait for that stream's
    work and report what hhg_plan_execute reports (TOLERANCE, per-job counts),
    and the device j equals the host execute's bitwise.  A trajectory
    launch in between does not touch the execution's counts."""
    pr = Problem("forced_failure_device", model=dict(kind=O.MODEL_DIRAC2D, v=1.0, mass=0.1),
                 pulse=dict(kind=O.PULSE_COS2, amp_a=0.5, n_cycles=2.0, t0=-2 * math.pi),
                 time=(0.3, 4.3, 64, 50), grid=("box", 24, 16, -3.0, 3.0, -3.0, 3.0),
                 solver=dict(method=0, rtol=1e-300, atol=1e-300, proposal=0, max_fail=1))
    _print_case(pr.name)
    hip = _hip()
    with pr.h_plan(hhg) as plan:
        j_host = plan.execute(allow_failures=True)[0]
        st_host = plan.stats
        n = plan.n_out
        buf, stream = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(buf), 8 * n) == 0
        assert hip.hipStreamCreateWithFlags(ctypes.byref(stream), 1) == 0
        try:
            plan.execute_device(buf.value, stream.value)
            st = plan.last_stats(allow_failures=True)
            assert plan.status == hhg.ERROR_TOLERANCE
            assert st.failed_trajectories == st_host.failed_trajectories == 24 * 16
            assert st.attempts == st_host.attempts
            plan.trajectories(np.arange(4))  # its own counters
            failed, nonfinite = plan.job_failures()
            assert list(failed) == [24 * 16] and list(nonfinite) == [0]
            assert hip.hipStreamSynchronize(stream) == 0
            j_dev = np.empty(n)
            assert hip.hipMemcpy(j_dev.ctypes.data, buf, 8 * n, 2) == 0
        finally:
            hip.hipStreamDestroy(stream)
            hip.hipFree(buf)
    assert np.array_equal(j_dev, j_host)
