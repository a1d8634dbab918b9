"""The surrogate's BF16 matrix-core network against its CPU oracle
(oracle/surrogate_ref.py, which emulates the BF16 operand rounding):
forward outputs, the loss and the gradient of every parameter tensor,
on configurations that exercise multi-tile, transposed and split-K
products; plus a training run that lowers the loss."""
import os
import sys

import numpy as np
import pytest

from rfsyn_amd.driver import Surrogate

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import surrogate_ref as sr  # noqa: E402

pytestmark = pytest.mark.gpu

CONFIGS = {
    # grid, C0, channels, coords, width, outputs, queries, batch
    "tiny": (16, 2, (4, 8, 8, 16), 6, 32, 6, 3, 3),
    "wide": (32, 3, (8, 16, 16, 32), 6, 96, 70, 4, 4),
    # the network of the surrogate workload: 64^3 voxels, 3 input channels,
    # 16-32-64-128 channels, 78 coordinate features, width 512, 8 probes x
    # 64 frequencies (real and imaginary), two scenes
    "c5": (64, 3, (16, 32, 64, 128), 78, 512, 128, 8, 2),
}


# Forward tolerance relative to the largest output.  At the surrogate
# workload's size (64^3 voxels, 16-32-64-128 channels) the oracle's own
# outputs move by about 2e-3 under an FP32-level (1e-7 relative) change of
# the parameters: such a change flips BF16 roundings of a few dozen
# activations of the deeper convolutions, and the flips cascade.  The
# device, which accumulates in FP32 in its own order, is held to five times
# that spread there; the parameter gradients move by at most 3e-3 under the
# same change, within their 2e-2 bound.
FWD_TOL = {"tiny": 1e-3, "wide": 1e-3, "c5": 1e-2}


def _setup(gpu, name, seed=3):
    grid, c0, ch, co, w, o, q, b = CONFIGS[name]
    net = Surrogate(gpu, grid, c0, ch, co, w, o, q, b, seed=seed)
    cfg = sr.Config(grid, c0, ch, co, w, o, q)
    rng = np.random.default_rng(seed)
    maps = rng.uniform(0, 1, (b, grid, grid, c0)).astype(np.float32)
    coords = rng.uniform(0, 1, (b, q, co)).astype(np.float32)
    labels = rng.normal(0, 0.3, (b, q, o)).astype(np.float32)
    return net, cfg, maps, coords, labels


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_forward_and_gradient_match_oracle(gpu, name):
    net, cfg, maps, coords, labels = _setup(gpu, name)
    flat = net.params()
    params = cfg.unflatten(flat.astype(np.float64))
    y_ref, _ = sr.forward(cfg, params, maps, coords)
    y = net.predict(maps, coords)
    assert np.max(np.abs(y - y_ref)) <= FWD_TOL[name] * np.max(np.abs(y_ref)) + 1e-6
    loss, g = net.gradient(maps, coords, labels)
    loss_ref, g_ref = sr.loss_and_grad(cfg, params, maps, coords, labels)
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    gg = cfg.unflatten(g.astype(np.float64))
    for k, (a, b) in enumerate(zip(gg, g_ref)):
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert err <= 2e-2, (k, err)
    net.close()


def test_params_round_trip_and_determinism(gpu):
    net, cfg, maps, coords, labels = _setup(gpu, "tiny")
    p = net.params()
    rng = np.random.default_rng(5)
    q = (p + rng.normal(0, 1e-2, p.shape)).astype(np.float32)
    net.set_params(q)
    assert np.array_equal(net.params(), q)
    l1, g1 = net.gradient(maps, coords, labels)
    l2, g2 = net.gradient(maps, coords, labels)
    assert l1 == l2 and np.array_equal(g1, g2)
    net.close()


def test_training_lowers_the_loss(gpu):
    """A fixed target that depends on the voxels and the query: Adam on the
    device cuts the loss well below its start."""
    grid, c0, ch, co, w, o, q, b = CONFIGS["tiny"]
    net = Surrogate(gpu, grid, c0, ch, co, w, o, q, 8, seed=11)
    rng = np.random.default_rng(2)
    maps = rng.uniform(0, 1, (8, grid, grid, c0)).astype(np.float32)
    coords = rng.uniform(0, 1, (8, q, co)).astype(np.float32)
    t = maps.mean(axis=(1, 2))                       # [8, c0]
    labels = np.stack([np.outer(coords[s].sum(axis=1), np.resize(t[s], o)) for s in range(8)]).astype(np.float32)
    first = net.train_step(maps, coords, labels, 3e-3)
    for _ in range(150):
        last = net.train_step(maps, coords, labels, 3e-3)
    assert last < 0.1 * first, (first, last)
    net.close()
