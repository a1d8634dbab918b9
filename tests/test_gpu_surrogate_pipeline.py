"""Surrogate pipeline on small scenes: GPU FDTD labels, device training,
and inference with FDTD spot checks and fallback."""
import numpy as np
import pytest

from rfsyn_amd import surrogate_pipeline as sp
from rfsyn_amd.driver import Surrogate

pytestmark = pytest.mark.gpu

N, NPML, STEPS = 32, 4, 384


def _data(gpu, lo, count):
    return sp.label_scenes(gpu, list(range(lo, lo + count)), STEPS, batch=64, n=N, npml=NPML)


def _net(gpu, batch):
    return Surrogate(gpu, N, 3, (8, 16, 16, 32), sp.N_COORDS, 64, 2 * sp.N_FREQS, sp.N_PROBES, batch, seed=4)


def test_labels_are_deterministic_and_finite(gpu):
    m1, c1, l1, _ = _data(gpu, 0, 6)
    m2, c2, l2, _ = _data(gpu, 0, 6)
    assert np.all(np.isfinite(l1)) and np.max(np.abs(l1)) > 0
    assert np.array_equal(l1, l2) and np.array_equal(m1, m2) and np.array_equal(c1, c2)


def test_training_gains_over_untrained_and_mean(gpu):
    mtr, ctr, ltr, _ = _data(gpu, 0, 192)
    mte, cte, lte, _ = _data(gpu, 1000, 64)
    scale = float(np.std(ltr))
    net = _net(gpu, 64)
    before = sp.rel_l2(sp.predict(net, mte, cte, 64) * scale, lte)
    sp.train(net, mtr, ctr, ltr / scale, 60, lr=1e-3, batch=64)
    after = sp.rel_l2(sp.predict(net, mte, cte, 64) * scale, lte)
    fit = sp.rel_l2(sp.predict(net, mtr, ctr, 64) * scale, ltr)
    mean = sp.rel_l2(np.broadcast_to(ltr.mean(axis=0), ltr.shape), ltr)
    net.close()
    assert np.median(after) < 0.5 * np.median(before), (np.median(before), np.median(after))
    assert np.median(fit) < np.median(mean), (np.median(fit), np.median(mean))


def test_spot_check_fallback_returns_fdtd_labels(gpu):
    """With a zero threshold every batch fails its spot check and is
    replaced by the FDTD labels, bit for bit."""
    idx = list(range(2000, 2040))
    m, c, lab, _ = sp.label_scenes(gpu, idx, STEPS, batch=64, n=N, npml=NPML)
    net = _net(gpu, 16)
    out, t, fallbacks, errs = sp.spot_checked_inference(gpu, net, idx, m, c, STEPS, 16, share=0.1,
                                                        threshold=0.0, n=N, npml=NPML)
    net.close()
    assert fallbacks == 3 and len(errs) == 2 + 2 + 1
    assert np.array_equal(out, lab)
