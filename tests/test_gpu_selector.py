"""Learned format selector on the GPU.

* the device forward/backward (k_mlp.hip) against the CPU FP64 gradient
  oracle (orc_mlp_loss_grad, itself pinned by finite differences);
* device training lowers the loss and fits a separable problem;
* the exhaustive autotuner returns finite times for the feasible formats and
  its best choice is at least as fast as the reference heuristic's.
"""

import numpy as np
import pytest

from annarchy_amd import selector as S
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_mlp_gradient_matches_oracle(gpu):
    rng = np.random.default_rng(0)
    n = 300
    X = rng.normal(0, 1, (n, 16)).astype(np.float32)
    T = rng.random((n, 12))
    T /= T.sum(axis=1, keepdims=True)
    sel = S.MLPSelector(S.init_params(3))
    loss, g = sel.gradient(X, T)
    ref_loss, ref_g = orc.mlp_loss_grad(X.astype(np.float64), T, sel.params.astype(np.float64))
    assert abs(loss - ref_loss) < 1e-5 * max(1.0, abs(ref_loss))
    scale = np.max(np.abs(ref_g))
    assert np.max(np.abs(g - ref_g)) < 2e-5 * scale


def test_mlp_training_fits_and_predicts(gpu):
    rng = np.random.default_rng(1)
    n = 600
    X = rng.normal(0, 1, (n, 16))
    labels = (X[:, 0] > 0).astype(int) * 3 + (X[:, 1] > 0).astype(int)  # 4 classes of 12
    times = np.full((n, 12), 2.0)
    times[np.arange(n), labels] = 1.0
    sel = S.MLPSelector()
    curve = sel.fit(X, times, epochs=800, lr=5e-3)
    assert curve[-1] < 0.5 * curve[0]
    pred = sel.predict(X)
    assert np.mean(pred == labels) > 0.95


def test_autotuner_vs_heuristic(gpu):
    for spec in S.corpus_specs(12, seed=5)[:6]:
        csr = S.generate(spec)
        t = S.autotune(csr, iters=3, repeats=1)
        # CSR and SELL classes always run; the others exactly when feasible
        assert np.isfinite(t[1:9]).all()
        assert all(np.isfinite(t[c]) == S.feasible(csr, c) for c in range(len(S.CLASSES)))
        h = S.heuristic_class(csr)
        assert t.min() <= t[h]


def test_committed_selector_device_predict_matches_oracle(gpu):
    # the shipped network (annarchy_amd/data) on every corpus matrix: device
    # FP32 inference against the FP64 oracle forward pass
    d = np.load(S.DATA_DIR / "selector_corpus.npz")
    sel = S.MLPSelector.load()
    dev = sel.predict(d["features"])
    logits = orc.mlp_forward(sel._norm(d["features"]).astype(np.float64), sel.params.astype(np.float64))
    ref = np.argmax(logits, axis=1)
    top2 = np.sort(logits, axis=1)[:, -2:]
    near_tie = (top2[:, 1] - top2[:, 0]) < 1e-4
    assert np.all((dev == ref) | near_tie)
    assert np.mean(dev == ref) > 0.99
    # select() on a fresh matrix returns a configuration that runs
    csr = S.generate(dict(n=20000, k=60, kind="lognormal", seed=77))
    cfg = S.select(csr)
    assert cfg in S.CLASSES
