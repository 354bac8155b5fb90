"""The CPU restatement of the learned-h network (oracle/mlp.py), pinned before it is used to
check the GPU kernels: its gradient against central finite differences of its own loss, and
its bfloat16 rounding against hand-computed bit patterns."""

import numpy as np

from helpers import check


def _random_model(rng):
    from oracle import mlp

    p = rng.normal(0.0, 0.3, mlp.N_PARAMS)
    norm = np.concatenate([rng.normal(0, 1, 8), rng.uniform(0.5, 2.0, 8), [0.01, 0.2]])
    return p, norm


def test_gradient_matches_finite_differences(rng):
    from oracle import mlp

    p, norm = _random_model(rng)
    feat = rng.normal(0.0, 1.0, (37, 8))
    target = rng.normal(0.0, 0.3, 37)
    loss, g = mlp.loss_and_grad(p, norm, feat, target)
    # every bias and a sample of the weights, central differences (O(eps^2) truncation)
    idx = np.concatenate([np.arange(mlp.OFF_B1, mlp.OFF_W2), np.arange(mlp.OFF_B2, mlp.N_PARAMS),
                          rng.choice(mlp.OFF_B1, 40, replace=False),
                          mlp.OFF_W2 + rng.choice(mlp.K_HID * mlp.K_HID, 60, replace=False)])  # fmt: skip
    eps = 1e-5
    fd = np.empty(len(idx))
    for n, k in enumerate(idx):
        a, b = p.copy(), p.copy()
        a[k] += eps
        b[k] -= eps
        fd[n] = (mlp.loss_and_grad(a, norm, feat, target)[0] - mlp.loss_and_grad(b, norm, feat, target)[0]) / (2 * eps)
    scale = np.abs(g).max()
    check("loss gradient = finite differences (rel. to max |g|)", fd / scale, g[idx] / scale, atol=1e-8)
    check("loss is the mean squared error of the normalised target", loss,
          float(np.mean((mlp.forward(p, norm, feat) - target) ** 2) / norm[17] ** 2), rtol=1e-12)  # fmt: skip


def test_bf16_rounding():
    from oracle import mlp

    vals = np.array([1.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8, -2.5, 1.0 + 2.0**-7 + 2.0**-9, 0.0], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2.0**-6, -2.5, 1.0 + 2.0**-7, 0.0])
    check("bfloat16 round to nearest even", want, mlp.bf16_round(vals))
