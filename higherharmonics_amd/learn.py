"""Training the step-proposal network from the solver's own step logs.

1. :func:`collect_step_log` runs a plan with the carried controller and step
   logging on: every logged record holds the 8 network features at an
   interval start and, as targets, log(h / Delta) of the controller's own
   proposal after the interval's first accepted step and log(attempts).
2. :func:`fit_proposal` sets the normalisation from the records and trains
   the network on the GPU (hand-written forward/backward/Adam kernels).
3. The trained network drives ``PROPOSAL_LEARNED`` solves, where it proposes
   each interval's first step; the embedded-error controller still accepts
   or rejects every step and hands an interval back to the plain controller
   after ``fallback_after`` rejections of the proposed step.
"""
from __future__ import annotations

from typing import Callable, Optional

import numpy as np


def collect_step_log(make_plan: Callable[..., object], capacity: int = 1 << 21, stride: int = 8) -> np.ndarray:
    """Step-log records [n][10] of one carried-controller solve.  make_plan(solver)
    builds the plan of the training workload for the given solver."""
    import higherharmonics_amd as H
    plan = make_plan(H.Solver(proposal=H.PROPOSAL_CARRY))
    try:
        plan.set_step_log(capacity, stride)
        plan.execute()
        rec = plan.step_log()
    finally:
        plan.close()
    ok = np.all(np.isfinite(rec), axis=1)
    return rec[ok]


def normalization(records: np.ndarray, floor: float = 1e-3):
    x, t = records[:, :8].astype(np.float64), records[:, 8:10].astype(np.float64)
    in_mean, in_std = x.mean(axis=0), np.maximum(x.std(axis=0), floor)
    out_mean, out_std = t.mean(axis=0), np.maximum(t.std(axis=0), floor)
    return in_mean, in_std, out_mean, out_std


def fit_proposal(records: np.ndarray, device: int = 0, epochs: int = 6, batch: int = 1024, lr: float = 2e-3,
                 seed: int = 1, mlp: Optional[object] = None):
    """A network trained on the records; returns (network, per-epoch losses)."""
    import higherharmonics_amd as H
    net = mlp or H.MLP(device=device, seed=seed)
    net.set_normalization(*normalization(records))
    losses = net.train(records[:, :8], records[:, 8:10], epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    return net, losses
