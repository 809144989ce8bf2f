"""Pins the CPU reference of the surrogate network (oracle/mlp_ref.py) by
central finite differences of its own loss (exact FP64 path, no BF16)."""
import numpy as np

from oracle import mlp_ref as R


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    hidden, nf = 64, 8
    dims = [2 * nf, hidden, hidden, hidden, R.OUT_PAD]
    n_params = sum(dims[l + 1] * dims[l] + dims[l + 1] for l in range(4))
    params = rng.normal(size=n_params) * 0.1  # moderate log-variances keep the loss O(1)
    fourier = rng.normal(size=2 * nf)
    n = 32
    p = np.stack([rng.uniform(0.05, 0.4, n), rng.uniform(2, 20, n)], axis=1)
    lam = rng.normal(size=(n, 3))
    status = (rng.uniform(size=n) < 0.2).astype(np.uint32)
    args = (fourier, p, lam, status, (0.05, 0.4, 2.0, 20.0), (0.0, 0.0, -5.0), (0.1, 0.1, 3.0))
    loss, g, _ = R.forward_backward(params, *args, hidden=hidden, bf16_store=False)
    idx = rng.choice(n_params, 60, replace=False)
    # include the head (log-variance and escape rows) explicitly
    head = n_params - (16 * hidden + 16)
    idx = np.concatenate([idx, head + 3 * hidden + np.arange(5), head + 4 * hidden + np.arange(5)])
    h = 1e-6
    for k in idx:
        pp, pm = params.copy(), params.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.forward_backward(pp, *args, hidden=hidden, bf16_store=False)[0]
              - R.forward_backward(pm, *args, hidden=hidden, bf16_store=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])
