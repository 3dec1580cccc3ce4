"""Learned format selector on the GPU.

* the device forward/backward (k_mlp.hip) against the CPU FP64 gradient
  oracle (orc_mlp_loss_grad, itself pinned by finite differences);
* device training lowers the loss and fits a separable problem;
* the exhaustive autotuner times exactly the feasible configurations;
* on held-out corpus matrices tuned live the trained network's picks are
  within 25% of the best configuration on average and better than the
  reference decision tree's;
* storage_format "auto" (the default) applies the network's pick to random
  rate connectivity and the reference's rules elsewhere, and the network
  still equals the oracle.
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


def test_autotuner_times_exactly_the_feasible_configurations(gpu):
    for spec in S.corpus_specs(12, seed=5)[:6]:
        csr = S.generate(spec)
        t = S.autotune(csr, iters=3, repeats=1)
        # CSR and SELL classes always run; the others exactly when feasible
        assert np.isfinite(t[1:9]).all()
        assert all(np.isfinite(t[c]) == S.feasible(csr, c) for c in range(len(S.CLASSES)))


def test_selector_regret_on_held_out_matrices_live(gpu):
    # held-out corpus matrices (never trained on), regenerated from their
    # seeds and tuned exhaustively now: the committed MLP's picks cost at
    # most 25% more than the best configuration on average, and less than
    # the reference decision tree's picks
    import json

    d = np.load(S.DATA_DIR / "selector_corpus.npz")
    specs = [json.loads(x) for x in d["specs"]]
    heur = np.array([s["heuristic"] for s in specs])
    groups = np.array([1 if s.get("large") else 0 for s in specs])
    _, te = S.holdout_split(d["times"], heur, groups=groups)
    te = [i for i in te if groups[i] == 0][::10][:10]
    sel = S.MLPSelector.load()
    T, mlp, ref = [], [], []
    for i in te:
        csr = S.generate(specs[i])
        T.append(S.autotune(csr, iters=5, repeats=2))
        mlp.append(int(sel.predict(S.features(csr, specs[i]["seed"] & 0xFFFF))[0]))
        ref.append(int(heur[i]))
    T = np.array(T)
    r_mlp = S.regret(T, S.with_fallback(T, np.array(mlp)))
    r_ref = S.regret(T, S.with_fallback(T, np.array(ref)))
    print(f"live held-out regret: MLP {r_mlp.mean():.3f}, reference tree {r_ref.mean():.3f}")
    assert r_mlp.mean() <= 0.25
    assert r_mlp.mean() < r_ref.mean()


def test_auto_storage_format(gpu):
    # storage_format "auto" (the default): the MLP's configuration for random
    # rate connectivity, dense for all-to-all, csr kept for one-to-one and
    # plastic projections; the simulation still equals the oracle
    import annarchy_amd as ah
    from annarchy_amd import connectivity as cn

    n = 3000
    net = ah.Network(dt=1.0, seed=3)
    pop = net.create(n, ah.RateNoisy)
    src = net.create(50, ah.Neuron(equations="r = 1", name="RateOne"))
    p_rand = net.connect(pop, pop, "exc").fixed_probability(0.05, cn.Normal(0.0, 0.05))
    p_all = net.connect(src, pop, "exc").all_to_all(0.01)
    p_one = net.connect(pop, pop, "exc").one_to_one(0.1)
    p_fix = net.connect(pop, pop, "exc").fixed_number_pre(20, 0.02, storage_format="csr", lanes=8)
    net.compile()
    assert p_rand.selected_format in S.CLASSES
    assert p_rand.get_format()["format"] == p_rand.selected_format["fmt"]
    assert p_all.selected_format == dict(fmt="dense") and p_all.get_format()["format"] == "dense"
    assert p_one.selected_format is None and p_one.get_format()["format"] == "csr"
    assert p_fix.selected_format is None and p_fix.get_format()["format"] == "csr"
    steps = 30
    r0 = np.random.default_rng(0).random(n)
    pop.r = r0
    src.r = 1.0
    noise = np.random.default_rng(1).normal(0.0, 0.1, (steps, n))
    pop.inject_random(0, noise)
    net.run(steps)
    # oracle: the four projections summed as one CSR (the same synapses)
    rows, cols, ws = [], [], []
    for pr in (p_rand, p_one, p_fix):
        c = pr.csr
        rows.append(np.repeat(np.arange(c.n_rows), np.diff(c.row_ptr)))
        cols.append(c.col_idx.astype(np.int64))
        ws.append(c.w if c.w is not None else np.full(c.nnz, c.w_const))
    x_in = 0.01 * 1.0 * 50  # all-to-all input from the constant source population
    got = pop.r.reshape(-1)
    r = r0.copy()
    import scipy.sparse as sp
    W = sp.csr_matrix((np.concatenate(ws), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    for k in range(steps):
        s_ = W @ r + x_in
        r = r + (s_ - r + noise[k]) / 10.0
        r = np.maximum(r, 0.0)
    np.testing.assert_allclose(got, r, rtol=1e-10, atol=1e-12)


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
