"""The surrogate's CPU oracle (oracle/surrogate_ref.py) against central
finite differences of its own float64 loss: pins the hand-written
backward pass (convolution window and strides, GELU derivative, the
per-scene / per-query split of the first MLP layer, bias sums)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import surrogate_ref as sr  # noqa: E402


def _problem(seed=0, grid=16, n=2):
    cfg = sr.Config(grid, 2, (2, 3, 3, 4), coords=6, width=5, outputs=3, queries=2)
    rng = np.random.default_rng(seed)
    params = [rng.normal(0, 0.5, sh) for sh in cfg.shapes()]
    maps = rng.uniform(0, 1, (n, grid, grid, 2))
    coords = rng.uniform(0, 1, (n, 2, 6))
    labels = rng.normal(0, 1, (n, 2, 3))
    return cfg, params, maps, coords, labels


def test_oracle_gradient_matches_finite_differences():
    cfg, params, maps, coords, labels = _problem()
    loss, grads = sr.loss_and_grad(cfg, params, maps, coords, labels, emulate_bf16=False)
    rng = np.random.default_rng(1)
    h = 1e-6
    for k, p in enumerate(params):
        for _ in range(3):
            idx = tuple(rng.integers(0, s) for s in p.shape)
            old = p[idx]
            p[idx] = old + h
            lp, _ = sr.loss_and_grad(cfg, params, maps, coords, labels, emulate_bf16=False)
            p[idx] = old - h
            lm, _ = sr.loss_and_grad(cfg, params, maps, coords, labels, emulate_bf16=False)
            p[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grads[k][idx]) <= 1e-6 * max(1.0, abs(fd)), (k, idx, fd, grads[k][idx])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-40], dtype=np.float64)
    # 1 + 2^-8 is a tie: to even (1.0); 1 + 3 * 2^-8 ties up to 1 + 2^-6
    assert list(sr.bf16(x)[:4]) == [1.0, 1.0, 1.0 + 2 ** -6, -2.5]
