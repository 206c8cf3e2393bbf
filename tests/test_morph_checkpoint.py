"""The checkpoint statistic of the surrogate-assisted morphology loop
(CPU): with the solver's K window means drawn as a mean field plus
independent noise, the noise-corrected error recovers the surrogate's own
error, which the raw error overstates."""
import math

import numpy as np
import pytest
import torch

from xbeach_gpu_amd.morph import checkpoint_error


@pytest.mark.parametrize("model_err,noise,K", [(0.03, 0.2, 12), (0.06, 0.2, 24), (0.0, 0.15, 48), (0.05, 0.0, 4)])
def test_noise_corrected_error_recovers_the_model_error(model_err, noise, K):
    g = np.random.default_rng(0)
    n = 200_000
    mu = 1.0 + g.random(n)                                   # the mean field
    dev = g.standard_normal(n)
    P = mu * (1.0 + model_err * dev)                         # the surrogate: relative error model_err
    W = mu[None, :] * (1.0 + noise * g.standard_normal((K, n)))  # window means of the solver
    F, F2 = W.mean(0), (W ** 2).mean(0)
    t = lambda a: torch.from_numpy(a)
    wet = torch.ones(n, dtype=torch.bool)
    err, raw = checkpoint_error(t(P), t(F), t(F2), K, wet)
    true = np.linalg.norm(P - mu) / np.linalg.norm(mu)
    assert err == pytest.approx(true, abs=0.004)
    assert raw >= err
    if noise > 0:
        # the raw error carries the K-window noise on top
        assert raw == pytest.approx(math.sqrt(true ** 2 + noise ** 2 / K), rel=0.05)


def test_dry_cells_are_left_out():
    P = torch.tensor([1.0, 5.0])
    F = torch.tensor([1.0, 0.0])
    err, raw = checkpoint_error(P, F, F * F, 4, torch.tensor([True, False]))
    assert err == 0.0 and raw == 0.0
