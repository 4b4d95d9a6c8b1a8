"""Training the step-proposal network from the solver's own step logs.

1. :func:`collect_step_log` runs a plan with step logging on: every logged
   record holds the 8 network features at an interval start and, as targets,
   log(h_opt / Delta) and log(attempts).  h_opt = h 0.9 max(err, 1e-16)^(-1/8)
   of the interval's first accepted step (size h, error ratio err) is the step
   the controller's error model allows there; since err scales as h^8 it does
   not depend on the step that happened to be tried.
2. :func:`fit_proposal` sets the normalisation from the records and trains
   the network on the GPU (hand-written forward/backward/Adam kernels).
3. :func:`train_on_solver_logs` starts from the carried controller's logs and
   adds, per further round, the logs of a learned-proposal solve with the
   current network: the states the learned policy itself visits (its step
   history feeds two of the features), labelled by the same error model.
4. The trained network drives ``PROPOSAL_LEARNED`` solves, where it proposes
   each interval's first step; the embedded-error controller still accepts
   or rejects every step and hands an interval back to the plain controller
   after ``fallback_after`` rejections of the proposed step.
"""
from __future__ import annotations

from typing import Callable, Optional

import numpy as np

# Floors of the input standard deviations: the step-history features (log of
# the carried step, log10 of the last error ratio) are constant in some logs
# (every step at the growth cap); a floor of a quarter log-unit keeps rounding
# noise from turning into many standard deviations at inference.
IN_STD_FLOOR = np.array([1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 0.25, 0.25])


def collect_step_log(make_plan: Callable[..., object], capacity: int = 1 << 21, stride: int = 8,
                     solver: Optional[object] = None) -> np.ndarray:
    """Step-log records [n][10] of one solve.  make_plan(solver) builds the plan
    of the training workload for the given solver (default: the carried
    controller)."""
    import higherharmonics_amd as H
    plan = make_plan(solver or H.Solver(proposal=H.PROPOSAL_CARRY))
    try:
        plan.set_step_log(capacity, stride)
        plan.execute()
        rec = plan.step_log()
    finally:
        plan.close()
    ok = np.all(np.isfinite(rec), axis=1)
    return rec[ok]


def normalization(records: np.ndarray):
    x, t = records[:, :8].astype(np.float64), records[:, 8:10].astype(np.float64)
    in_mean, in_std = x.mean(axis=0), np.maximum(x.std(axis=0), IN_STD_FLOOR)
    out_mean, out_std = t.mean(axis=0), np.maximum(t.std(axis=0), 1e-3)
    return in_mean, in_std, out_mean, out_std


def fit_proposal(records: np.ndarray, device: int = 0, epochs: int = 6, batch: int = 1024, lr: float = 2e-3,
                 seed: int = 1, mlp: Optional[object] = None):
    """A network trained on the records; returns (network, per-epoch losses)."""
    import higherharmonics_amd as H
    net = mlp or H.MLP(device=device, seed=seed)
    net.set_normalization(*normalization(records))
    losses = net.train(records[:, :8], records[:, 8:10], epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    return net, losses


def train_on_solver_logs(make_plan: Callable[..., object], capacity: int = 1 << 21, stride: int = 8, rounds: int = 2,
                         epochs: int = 6, batch: int = 1024, lr: float = 2e-3, seed: int = 1, device: int = 0):
    """Round 1 trains on the carried controller's logs; every further round
    appends the logs of a learned-proposal solve with the current network and
    retrains from scratch on all records.  Returns (network, records,
    per-round epoch losses)."""
    import higherharmonics_amd as H
    rec = collect_step_log(make_plan, capacity, stride)
    net, losses = fit_proposal(rec, device, epochs, batch, lr, seed)
    history = [losses]
    for _ in range(1, rounds):
        more = collect_step_log(make_plan, capacity, stride,
                                solver=H.Solver(proposal=H.PROPOSAL_LEARNED, mlp=net))
        rec = np.concatenate([rec, more])
        old = net
        net, losses = fit_proposal(rec, device, epochs, batch, lr, seed)
        old.close()
        history.append(losses)
    return net, rec, history
