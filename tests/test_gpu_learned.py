"""The step-proposal network on the GPU: gradients against the CPU FP64
reference, the solve kernel's MFMA inference against the per-sample forward
pass, training on the solver's own step logs, and the learned solve's j(t)
within the parity criterion while taking fewer steps than the plain
controller."""
import math

import numpy as np
import pytest

import oracle as O
from tests import mlp_reference as R
from tests.problems import Problem, oracle_j, rel_l2

pytestmark = pytest.mark.gpu


def _net(hhg, seed=3):
    net = hhg.MLP(seed=seed)
    rng = np.random.default_rng(seed)
    in_mean, in_std = rng.normal(size=8).astype(np.float32), (0.5 + rng.random(8)).astype(np.float32)
    out_mean, out_std = rng.normal(size=2).astype(np.float32), (0.5 + rng.random(2)).astype(np.float32)
    net.set_normalization(in_mean, in_std, out_mean, out_std)
    return net, (in_mean, in_std, out_mean, out_std)


def test_mlp_gradients_match_cpu_fp64(hhg):
    print("[case] mlp gradients n=256", flush=True)
    net, norm = _net(hhg)
    rng = np.random.default_rng(11)
    p = (rng.normal(size=net.n_params) * 0.2).astype(np.float32)
    net.set_params(p)
    x = rng.normal(size=(256, 8)).astype(np.float32) * norm[1] + norm[0]
    t = rng.normal(size=(256, 2)).astype(np.float32)
    g_gpu, loss_gpu = net.gradients(x, t)
    loss_ref, g_ref, _ = R.loss_and_grad(p, x, t, *norm, bf16_round=True)
    assert abs(loss_gpu - loss_ref) <= 1e-5 * max(1.0, loss_ref)
    assert np.linalg.norm(g_gpu - g_ref) <= 1e-4 * np.linalg.norm(g_ref)
    assert np.array_equal(net.params(), p)


def test_mlp_forward_paths_agree(hhg):
    print("[case] mlp forward thread vs mfma n=1000", flush=True)
    net, norm = _net(hhg, seed=5)
    rng = np.random.default_rng(12)
    net.set_params((rng.normal(size=net.n_params) * 0.2).astype(np.float32))
    x = rng.normal(size=(1000, 8)).astype(np.float32) * norm[1] + norm[0]
    y_thread = net.forward(x)
    y_mfma = net.forward(x, mfma=True)
    _, _, y_ref = R.loss_and_grad(net.params(), x, np.zeros((1000, 2)), *norm, bf16_round=True)
    scale = np.max(np.abs(y_ref))
    assert np.max(np.abs(y_thread - y_ref)) <= 1e-3 * scale
    assert np.max(np.abs(y_mfma - y_ref)) <= 1e-2 * scale
    # BF16 activation flips aside, the two GPU paths agree closely
    assert np.median(np.abs(y_mfma - y_thread)) <= 1e-4 * scale


def test_mlp_training_reduces_loss(hhg):
    print("[case] mlp training synthetic regression", flush=True)
    rng = np.random.default_rng(13)
    x = rng.normal(size=(20000, 8)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2], np.tanh(x[:, 3] - x[:, 4])], axis=1).astype(np.float32)
    net = hhg.MLP(seed=9)
    net.set_normalization(np.zeros(8), np.ones(8), t.mean(0), t.std(0))
    losses = net.train(x, t, epochs=8, batch=512, lr=3e-3)
    # the first epoch's mean already includes part of the descent
    assert losses[-1] < 0.2 * losses[0]
    y = net.forward(x[:2000])
    assert np.mean((y - t[:2000]) ** 2) < 0.05 * np.mean((t[:2000] - t[:2000].mean(0)) ** 2) + 1e-3


LEARN_PR = Problem(
    "honeycomb_strong_field",
    model=dict(kind=O.MODEL_HONEYCOMB, hopping=1.0),
    pulse=dict(kind=O.PULSE_COS2, amp_a=1.2, n_cycles=3.0, t0=0.0, cep=0.4),
    time=(0.0, 6 * math.pi + 2 * math.pi, 512, 50),
    grid=("hex", 24, 24),
)


def test_learned_proposal_parity_and_fewer_steps(hhg):
    """Trained on the solver's own step logs (carried controller, then its own
    learned-proposal solve), the network proposes
    each interval's first step; the controller still judges every step.  j(t)
    stays within 1e-9 of the tight oracle and the solve needs far fewer
    attempts than the plain (restart) controller."""
    from higherharmonics_amd import learn
    pr = LEARN_PR
    print(f"[case] learned proposal {pr.name}", flush=True)
    def make_plan(sv):
        return hhg.Plan([hhg.Job(pr.h_model(hhg), pr.h_laser(hhg), pr.h_time(hhg), 0)], [pr.h_grid(hhg)], sv)

    net, rec, hist = learn.train_on_solver_logs(make_plan, capacity=1 << 18, stride=2, rounds=2, epochs=4, batch=512)
    assert len(rec) > 10000 and np.all(np.isfinite(rec))
    assert all(h[-1] <= h[0] for h in hist)
    j_tight, _ = oracle_j(pr, rtol=1e-13, atol=1e-13)
    with pr.h_plan(hhg) as plain:
        j_plain = plain.execute()[0]
        att_plain = plain.stats.attempts
    for group in (0, 1):
        solver = hhg.Solver(proposal=hhg.PROPOSAL_LEARNED, mlp=net)
        with hhg.Plan([hhg.Job(pr.h_model(hhg), pr.h_laser(hhg), pr.h_time(hhg), 0)], [pr.h_grid(hhg)], solver,
                      group_by_cost=group) as plan:
            j = plan.execute()[0]
            st = plan.stats
        print(f"  group={group} err_learned={rel_l2(j, j_tight):.2e} err_plain={rel_l2(j_plain, j_tight):.2e} "
              f"attempts learned={st.attempts} plain={att_plain} fallbacks={st.proposal_fallbacks}", flush=True)
        assert rel_l2(j, j_tight) <= 1e-9
        assert st.attempts < 0.5 * att_plain
        assert st.failed_trajectories == 0


def _constant_net(hhg, log_h_ratio):
    """A network whose outputs are constant: all weights zero, output bias =
    (log_h_ratio, 0) with identity normalisation."""
    net = hhg.MLP(seed=1)
    net.set_normalization(np.zeros(8), np.ones(8), np.zeros(2), np.ones(2))
    p = np.zeros(net.n_params, dtype=np.float32)
    p[-2] = log_h_ratio
    net.set_params(p)
    return net


def test_nonfinite_proposal_falls_back_to_plain_controller(hhg):
    """A network that proposes NaN restarts the interval at Delta/n_sub.
    Scheduled inference asks in every interval: every (k, interval) is counted
    as a fallback and the step sequence is the plain solve's (same attempts;
    j equal to rounding: the plain solve takes its stage pulses from the
    restart sequence's table rows).  First-interval inference falls back only
    there and the carried step starts the others."""
    from tests.problems import HONEYCOMB_SMALL as pr
    print(f"[case] nan proposal {pr.name}", flush=True)
    net = _constant_net(hhg, np.nan)
    with pr.h_plan(hhg) as plain:
        j_plain = plain.execute()[0]
        st_plain = plain.stats
    assert st_plain.rejects == 0
    n_k = len(pr.o_grid()[0])
    for policy in (hhg.INFER_SCHEDULED, hhg.INFER_FIRST):
        solver = hhg.Solver(proposal=hhg.PROPOSAL_LEARNED, mlp=net, infer_policy=policy)
        with hhg.Plan([hhg.Job(pr.h_model(hhg), pr.h_laser(hhg), pr.h_time(hhg), 0)], [pr.h_grid(hhg)],
                      solver) as plan:
            j = plan.execute()[0]
            st = plan.stats
        if policy == hhg.INFER_SCHEDULED:
            assert st.proposal_fallbacks == n_k * pr.time[2]
            assert st.attempts == st_plain.attempts
            assert rel_l2(j, j_plain) <= 1e-13
        else:
            assert st.proposal_fallbacks == n_k
            assert st.attempts < st_plain.attempts and st.failed_trajectories == 0
    net.close()


COARSE_PR = Problem(
    "honeycomb_coarse_sampling",
    model=dict(kind=O.MODEL_HONEYCOMB, hopping=1.0),
    pulse=dict(kind=O.PULSE_COS2, amp_a=1.5, n_cycles=4.0, t0=0.0, cep=0.0),
    time=(0.0, 10 * math.pi, 160, 50),  # Delta = 0.196: a whole-interval step fails the error test
    grid=("hex", 12, 12),
)


def test_rejected_proposals_fall_back_after_n(hhg):
    """A network that always proposes 4.5 Delta (clipped to the interval) on a coarse time
    grid: the controller rejects the proposal where the interval is too long,
    the interval goes back to the plain controller after fallback_after
    rejections (0: the controller shrinks the step itself), and j stays within
    the parity criterion of the tight oracle."""
    pr = COARSE_PR
    print(f"[case] fallback_after {pr.name}", flush=True)
    j_tight, _ = oracle_j(pr, rtol=1e-13, atol=1e-13)
    net = _constant_net(hhg, math.log(4.5))
    for fallback_after in (0, 1):
        solver = hhg.Solver(proposal=hhg.PROPOSAL_LEARNED, mlp=net, fallback_after=fallback_after,
                            infer_policy=hhg.INFER_SCHEDULED)
        with hhg.Plan([hhg.Job(pr.h_model(hhg), pr.h_laser(hhg), pr.h_time(hhg), 0)], [pr.h_grid(hhg)],
                      solver) as plan:
            j = plan.execute()[0]
            st = plan.stats
        print(f"  fallback_after={fallback_after} fallbacks={st.proposal_fallbacks} rejects={st.rejects} "
              f"err={rel_l2(j, j_tight):.2e}", flush=True)
        if fallback_after:
            assert 0 < st.proposal_fallbacks <= st.rejects
        else:  # the controller shrinks the rejected proposal itself
            assert st.proposal_fallbacks == 0 and st.rejects > 0
        assert st.failed_trajectories == 0
        assert rel_l2(j, j_tight) <= 1e-9
    net.close()


def test_learned_mixed_plan_grouped_graph(hhg):
    """Configuration-5 structure at small size with the learned proposal:
    honeycomb and PiFlux jobs in one plan, network trained on the plan's own
    carried-controller step logs, step-count grouping, autotuned, graph
    replays.  Every job within 1e-9 of the tight oracle (at rtol 1e-11, see
    below); grouping only
    reorders k-points, so j equals the ungrouped solve up to the rounding of
    the per-wave FP64 sums (1e-13); graph replays are bitwise equal;
    infer_every = 3 stays within the criterion too."""
    from higherharmonics_amd import learn
    from higherharmonics_amd import workloads as W
    from tests.test_gpu_parity import _oracle_job
    jobs = W.mixed_sweep(n=12, n_each=2, per_cycle=64, cos2_cycles=2, sigma_cycles=0.5)
    print(f"[case] learned mixed sweep {len(jobs)} jobs", flush=True)
    net, _, _ = learn.train_on_solver_logs(lambda sv: W.build_plan(hhg, jobs, sv), capacity=1 << 18, stride=1,
                                           rounds=2, epochs=4, batch=512)
    tight = []
    for j in jobs:
        from tests.problems import Problem as P
        pr = P("job", model=dict(j["model"]), pulse=dict(j["pulse"]), time=tuple(j["time"]), grid=tuple(j["grid"]))
        tight.append(oracle_j(pr, rtol=1e-13, atol=1e-13)[0])
    # 64 samples per cycle and fields up to A0 = 2 make whole-interval steps
    # error-limited: at rtol 1e-10 the learned (and the carried) solve lands at
    # 2e-9 of the tight oracle where the plain controller's short restart steps
    # land at 7e-10; rtol 1e-11 brings the learned solve inside 1e-9
    def learned(**kw):
        return hhg.Solver(proposal=hhg.PROPOSAL_LEARNED, mlp=net, rtol=1e-11, atol=1e-11, **kw)

    with W.build_plan(hhg, jobs, learned()) as plan:
        ungrouped = plan.execute()
    with W.build_plan(hhg, jobs, learned(), group_by_cost=1) as plan:
        grouped = plan.execute()
        plan.autotune()
        plan.capture_graph()
        replays = [plan.execute() for _ in range(3)]
    for a, b, t in zip(grouped, ungrouped, tight):
        # same per-k trajectories; only the FP64 sums inside each wave see other
        # k-points: 64-term sums round at ~1e-14 of sum |w g|, which exceeds |j|
        # tenfold where j changes sign, so the two groupings agree to ~1e-13
        assert rel_l2(a, b) <= 1e-12
        assert rel_l2(a, t) <= 1e-9
    for r in replays:
        for a, b in zip(r, grouped):
            assert np.array_equal(a, b)
    with W.build_plan(hhg, jobs, learned(infer_every=3, infer_policy=hhg.INFER_SCHEDULED)) as plan:
        for a, t in zip(plan.execute(), tight):
            assert rel_l2(a, t) <= 1e-9
    net.close()


def test_step_log_selection_and_targets(hhg):
    """Step logs of a carried-controller solve: exactly the (k, interval)
    pairs with (i + 7 (k mod 64)) mod stride = 0 are logged, features 0 and 1
    are the k-point, and the targets are log(h_opt / Delta), bounded by the
    error-model factor on a step of at most Delta, and log(attempts) >= 0."""
    from tests.problems import DIRAC_SMALL as pr
    print(f"[case] step log {pr.name}", flush=True)
    stride = 5
    with pr.h_plan(hhg, solver=dict(proposal=hhg.PROPOSAL_CARRY)) as plan:
        plan.set_step_log(1 << 20, stride)
        plan.execute()
        rec = plan.step_log()
    kx, ky, _ = pr.o_grid()
    N = pr.time[2]
    i = np.arange(1, N + 1)
    lanes = np.arange(len(kx)) % 64
    expect = int(sum(np.count_nonzero((i + 7 * l) % stride == 0) for l in lanes))
    assert len(rec) == expect
    assert np.all(np.isfinite(rec))
    pts = set(zip(np.float32(kx).tolist(), np.float32(ky).tolist()))
    assert all((a, b) in pts for a, b in zip(rec[:, 0].tolist(), rec[:, 1].tolist()))
    assert np.all(rec[:, 9] >= 0.0)
    # h_opt = h 0.9 max(err, 1e-16)^(-1/8) with the accepted step h <= Delta
    assert np.all(rec[:, 8] <= math.log(0.9 * 100.0) + 1e-5)


def test_learned_dirac_t2_matches_tight_oracle(hhg):
    """The learned proposal on the 2D Dirac model with T2 dephasing (the
    template of the Dirac workloads): trained on the solver's own step logs,
    both inference policies (first interval; every interval) stay within 1e-9
    of the oracle at rtol = atol = 1e-13 and need far fewer attempts than the
    plain controller; the first-interval policy never needs more attempts
    than the carried step."""
    from higherharmonics_amd import learn
    from tests.problems import DIRAC_SMALL as pr
    print(f"[case] learned proposal {pr.name}", flush=True)

    def make_plan(sv):
        return hhg.Plan([hhg.Job(pr.h_model(hhg), pr.h_laser(hhg), pr.h_time(hhg), 0)], [pr.h_grid(hhg)], sv)

    net, rec, _ = learn.train_on_solver_logs(make_plan, capacity=1 << 18, stride=2, rounds=2, epochs=4, batch=512)
    j_tight, _ = oracle_j(pr, rtol=1e-13, atol=1e-13)
    with pr.h_plan(hhg) as plain:
        plain.execute()
        att_plain = plain.stats.attempts
    with pr.h_plan(hhg, solver=dict(proposal=hhg.PROPOSAL_CARRY)) as carry:
        j_carry = carry.execute()[0]
        att_carry = carry.stats.attempts
    for policy in (hhg.INFER_FIRST, hhg.INFER_SCHEDULED):
        with make_plan(hhg.Solver(proposal=hhg.PROPOSAL_LEARNED, mlp=net, infer_policy=policy)) as plan:
            j = plan.execute()[0]
            st = plan.stats
        print(f"  policy={policy} err={rel_l2(j, j_tight):.2e} carry err={rel_l2(j_carry, j_tight):.2e} "
              f"attempts learned={st.attempts} carry={att_carry} plain={att_plain}", flush=True)
        assert rel_l2(j, j_tight) <= 1e-9
        assert st.attempts < 0.5 * att_plain
        assert st.failed_trajectories == 0
        if policy == hhg.INFER_FIRST:
            assert st.attempts <= att_carry
    net.close()


def _oracle_wave_attempts(att):
    """Sum over waves of 64 consecutive k-points and over intervals of the
    busiest lane's attempts, from per-k per-interval attempt counts."""
    m = -(-len(att) // 64) * 64
    a = np.zeros((m, att.shape[1]), dtype=np.int64)
    a[:len(att)] = att
    return int(a.reshape(-1, 64, att.shape[1]).max(axis=1).sum())


def test_wave_attempts_and_network_first_step_on_error_limited_steps(hhg):
    """Honeycomb under strong seeded cos^2 fields sampled 32 times per cycle:
    whole-interval steps are error-limited and the attempts per interval
    differ across k-points and over time.  The carried controller's counts
    equal the oracle's per trajectory (up to a rare rounding-level decision),
    so the kernel's wave-attempt counter (per wave and interval, the busiest
    lane's attempts) must equal the sum formed from the oracle's per-interval
    attempts.  The learned proposal (first interval) and the network in every
    interval are within 1e-9 of the tight oracle; the first needs at most the
    carried controller's attempts."""
    from higherharmonics_amd import learn
    from higherharmonics_amd import workloads as W
    from tests.problems import Problem as P
    jobs = W.honeycomb_cos2(n=48, n_pulses=2, per_cycle=32, A0_range=(1.2, 1.5))
    print(f"[case] error-limited steps on {len(jobs)} coarse honeycomb jobs", flush=True)
    probs = [P("job", model=dict(j["model"]), pulse=dict(j["pulse"]), time=tuple(j["time"]), grid=tuple(j["grid"]))
             for j in jobs]
    expect_waves, expect_att = 0, 0
    for pr in probs:
        kx, ky, _ = pr.o_grid()
        att = O.attempts_per_interval(pr.o_model(), pr.o_pulse(), pr.o_time(),
                                      pr.o_solver(proposal=O.PROPOSAL_CARRY), kx, ky, fast=True)
        expect_waves += _oracle_wave_attempts(att)
        expect_att += int(att.sum())
    with W.build_plan(hhg, jobs, hhg.Solver(proposal=hhg.PROPOSAL_CARRY), count_wave_attempts=1) as plan:
        j_carry = plan.execute()
        st_c = plan.stats
    print(f"  carry: attempts {st_c.attempts} (oracle {expect_att}), wave attempts {st_c.wave_attempts} "
          f"(oracle {expect_waves}), rejects {st_c.rejects}", flush=True)
    # the GPU's stage pulses come from its own tables: a rounding difference can
    # move one controller decision in millions
    assert abs(st_c.attempts - expect_att) <= 1e-5 * expect_att
    assert abs(st_c.wave_attempts - expect_waves) <= 1e-4 * expect_waves
    net, _, _ = learn.train_on_solver_logs(lambda sv: W.build_plan(hhg, jobs, sv), capacity=1 << 18, stride=1,
                                           rounds=2, epochs=4, batch=512)
    runs = {}
    for policy in (hhg.INFER_FIRST, hhg.INFER_SCHEDULED):
        sv = hhg.Solver(proposal=hhg.PROPOSAL_LEARNED, mlp=net, infer_policy=policy)
        with W.build_plan(hhg, jobs, sv, count_wave_attempts=1) as plan:
            runs[policy] = (plan.execute(), plan.stats)
    net.close()
    tight = [oracle_j(pr, rtol=1e-13, atol=1e-13)[0] for pr in probs]
    err_c = max(rel_l2(a, t) for a, t in zip(j_carry, tight))
    print(f"  carry err={err_c:.2e}", flush=True)
    assert err_c <= 1e-9
    for policy, (js, st) in runs.items():
        err = max(rel_l2(a, t) for a, t in zip(js, tight))
        print(f"  learned policy={policy}: attempts {st.attempts}, wave attempts {st.wave_attempts}, rejects "
              f"{st.rejects}, err={err:.2e}", flush=True)
        assert err <= 1e-9 and st.failed_trajectories == 0
    # the first-interval policy runs the carried controller after the first
    # interval
    assert runs[hhg.INFER_FIRST][1].attempts <= 1.02 * st_c.attempts
    st_s = runs[hhg.INFER_SCHEDULED][1]
    print(f"  network every interval / carry: rejects {st_s.rejects / st_c.rejects:.3f}, "
          f"wave attempts {st_s.wave_attempts / st_c.wave_attempts:.3f}", flush=True)


def test_network_proposals_cut_rejections_of_the_carried_step(hhg):
    """What the trained network gains over the method without it, on the
    error-limited workload of the uneven-step measurement at reduced size:
    seeded cos^2 pulses (A0 in [0.1, 1.5]) on the hexagonal 96 x 96 grid at
    32 samples per cycle, the network trained on the step logs of the same
    pulses on a 48 x 48 grid.  With the network proposing the first step of
    every interval, fewer than half of the carried step's rejections remain
    and the waves issue fewer attempts; the solve stays within 1e-9 of a tight
    solve and the plain-controller fallback never triggers."""
    from higherharmonics_amd import learn
    from higherharmonics_amd import workloads as W
    jobs = W.honeycomb_cos2(n=96, n_pulses=4, per_cycle=32)
    small = W.honeycomb_cos2(n=48, n_pulses=4, per_cycle=32)
    print(f"[case] network vs carried step, {len(jobs)} honeycomb jobs at 32 samples/cycle", flush=True)
    net, rec, _ = learn.train_on_solver_logs(lambda sv: W.build_plan(hhg, small, sv),
                                             capacity=1 << 20, stride=2, rounds=2)
    with W.build_plan(hhg, jobs, hhg.Solver(proposal=hhg.PROPOSAL_CARRY), count_wave_attempts=1) as plan:
        plan.execute()
        st_c = plan.stats
    sv = hhg.Solver(proposal=hhg.PROPOSAL_LEARNED, mlp=net, infer_policy=hhg.INFER_SCHEDULED)
    with W.build_plan(hhg, jobs, sv, count_wave_attempts=1) as plan:
        js = plan.execute()
        st_s = plan.stats
    with W.build_plan(hhg, jobs, hhg.Solver(rtol=1e-13, atol=1e-13)) as plan:
        tight = plan.execute()
    net.close()
    err = max(rel_l2(a, t) for a, t in zip(js, tight))
    print(f"  carry: rejects {st_c.rejects}, wave attempts {st_c.wave_attempts}; network every interval: rejects "
          f"{st_s.rejects}, wave attempts {st_s.wave_attempts}, fallbacks {st_s.proposal_fallbacks}, "
          f"err vs tight {err:.2e} ({len(rec)} training records)", flush=True)
    assert err <= 1e-9 and st_s.failed_trajectories == 0
    assert st_s.rejects < 0.5 * st_c.rejects
    assert st_s.wave_attempts < st_c.wave_attempts
