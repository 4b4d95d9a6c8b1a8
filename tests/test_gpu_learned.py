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
    """Trained on its own carried-controller step logs, the network proposes
    each interval's first step; the controller still judges every step.  j(t)
    stays within 1e-9 of the tight oracle and the solve needs far fewer
    attempts than the plain (restart) controller."""
    from higherharmonics_amd import learn
    pr = LEARN_PR
    print(f"[case] learned proposal {pr.name}", flush=True)
    rec = learn.collect_step_log(lambda solver: pr.h_plan(hhg, solver=dict(proposal=solver.proposal)),
                                 capacity=1 << 18, stride=2)
    assert len(rec) > 10000 and np.all(np.isfinite(rec))
    net, losses = learn.fit_proposal(rec, epochs=4, batch=512)
    assert losses[-1] <= losses[0]
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
