"""Selector host logic (no GPU): the reference decision tree, features,
corpus generation, soft targets and regret, and the gain of the committed
trained network over the decision tree on the held-out corpus matrices."""

import json

import numpy as np
import pytest

from annarchy_amd import connectivity as conn
from annarchy_amd import selector as S
from oracle import oracle as orc


def _csr_from_degrees(deg, n_cols, seed=0):
    rng = np.random.default_rng(seed)
    return conn.random_in_degree(n_cols, len(deg), rng, np.asarray(deg), conn.Normal(0, 1, rng=rng))


def test_heuristic_is_the_reference_tree():
    # core/Projection.py:752-787: density >= 0.6 -> dense; GPU: mean row <= 128 -> sell, else csr
    dense = _csr_from_degrees(np.full(100, 70), 100)
    assert S.heuristic_class(dense) == "dense"
    short = _csr_from_degrees(np.full(1000, 128), 10000)
    assert S.heuristic_class(short) == S.HEURISTIC_SELL
    long_ = _csr_from_degrees(np.full(1000, 129), 10000)
    assert S.heuristic_class(long_) == S.HEURISTIC_CSR
    # the mean is taken over non-empty rows (LIL rows)
    deg = np.zeros(1000, dtype=int)
    deg[::2] = 200
    assert S.heuristic_class(_csr_from_degrees(deg, 10000)) == S.HEURISTIC_CSR


def test_features_are_deterministic_and_finite():
    for spec in S.corpus_specs(12, seed=3):
        c = S.generate(spec)
        f1, f2 = S.features(c, 1), S.features(c, 1)
        assert f1.shape == (16,) and np.all(np.isfinite(f1)) and np.array_equal(f1, f2)
        assert abs(f1[1] - np.log10(c.nnz)) < 1e-12


def test_corpus_specs_cover_the_kinds_and_sizes():
    specs = S.corpus_specs(2000)
    kinds = {s["kind"] for s in specs}
    assert kinds == set(S.KINDS)
    n = np.array([s["n"] for s in specs])
    assert n.min() >= 1000 and n.max() <= 1_000_000
    # log-uniform: about a third per decade
    dec = np.histogram(np.log10(n), bins=[3, 4, 5, 6])[0] / len(n)
    assert np.all(np.abs(dec - 1 / 3) < 0.05)
    assert max(s["n"] * s["k"] for s in specs) <= 2e6 * 1.0001


def test_generated_patterns():
    banded = S.generate(dict(n=5000, k=20, kind="banded", seed=1, spread=1.0))
    rows = np.repeat(np.arange(banded.n_rows), np.diff(banded.row_ptr))
    d = np.minimum(np.abs(banded.col_idx - rows), banded.n_rows - np.abs(banded.col_idx - rows))
    assert d.max() <= 20
    block = S.generate(dict(n=4096, k=64, kind="block", seed=2, block=8, fill=1.0))
    assert S._tile_stats(block, np.random.default_rng(0), 8) > 0.99
    for c in (banded, block):
        for r in range(0, c.n_rows, 97):
            seg = c.col_idx[c.row_ptr[r]:c.row_ptr[r + 1]]
            assert np.all(np.diff(seg) > 0)


def test_soft_targets_and_regret():
    t = np.array([[1.0, 1.05, 2.0, np.inf], [3.0, 1.5, 1.5, 1.6]])
    p = S.soft_targets(t)
    assert np.allclose(p.sum(axis=1), 1.0) and p[0, 3] == 0.0
    assert p[0, 0] > p[0, 1] > p[0, 2]
    assert p[1, 1] == pytest.approx(p[1, 2])
    r = S.regret(t, np.array([1, 0]))
    assert r == pytest.approx([0.05, 1.0])


def test_feasibility_guards_padded_formats():
    skew = _csr_from_degrees(np.r_[np.full(999, 2), 5000], 10000)
    assert not S.feasible(skew, 9)          # ELLR would pad every row to 5000
    assert S.feasible(skew, 5)              # CSR always
    block = S.generate(dict(n=2048, k=32, kind="block", seed=3, block=8, fill=1.0))
    assert S.feasible(block, 10) and S.feasible(block, 11)


def _corpus():
    d = np.load(S.DATA_DIR / "selector_corpus.npz")
    specs = [json.loads(s) for s in d["specs"]]
    return d["features"], d["times"], specs, np.array([s["heuristic"] for s in specs])


def test_corpus_features_reproduce():
    # the stored features are those of the seeded matrices the specs describe
    X, T, specs, h = _corpus()
    small = [i for i, s in enumerate(specs) if s["n"] * s["k"] < 3e4][:12]
    assert len(small) >= 6
    for i in small:
        csr = S.generate(specs[i])
        assert csr.nnz == specs[i]["nnz"]
        np.testing.assert_allclose(S.features(csr, specs[i]["seed"] & 0xFFFF), X[i], rtol=1e-12, atol=1e-12)
        hc = S.heuristic_class(csr)
        assert (-1 if hc == "dense" else hc) == h[i]


def test_trained_selector_gain_over_heuristic():
    # held-out 20% of the 2000-matrix corpus and of the 40 matrices at the
    # rate benchmark's scale (nnz 1e7..1e8); the committed network's choices
    # are recomputed with the FP64 oracle forward pass
    X, T, specs, h = _corpus()
    groups = np.array([1 if s.get("large") else 0 for s in specs])
    assert (groups == 0).sum() == 2000 and (groups == 1).sum() == 40
    m = np.load(S.DATA_DIR / "selector_mlp.npz")
    tr, te = S.holdout_split(T, h, groups=groups)
    assert len(te) == 408 and (groups[te] == 1).sum() == 8
    Xn = ((X[te] - m["mean"]) / m["std"]).astype(np.float32).astype(np.float64)
    choice = S.with_fallback(T[te], np.argmax(orc.mlp_forward(Xn, m["params"].astype(np.float64)), axis=1))
    rows = np.arange(len(te))
    best = T[te].min(axis=1)
    mlp = T[te][rows, choice] / best - 1.0
    heur = T[te][rows, S.with_fallback(T[te], h[te])] / best - 1.0
    # within 10% of exhaustive tuning, and far better than the decision tree,
    # on both sets
    assert np.mean(mlp) < 0.10
    assert np.mean(mlp) < 0.1 * np.mean(heur)
    assert T[te][rows, choice].sum() < 0.5 * T[te][rows, h[te]].sum()
    big = groups[te] == 1
    assert np.mean(mlp[big]) < 0.10
    assert np.mean(mlp[big]) < 0.25 * np.mean(heur[big])
    # the committed evaluation record agrees
    res = json.loads((S.DATA_DIR / "selector_results.json").read_text())
    assert abs(res["mlp_seed0"]["mean_regret"] - np.mean(mlp)) < 0.01
    assert abs(res["mlp_seed0"]["large"]["mean_regret"] - np.mean(mlp[big])) < 0.01


def test_large_set_specs():
    # the matrices at the rate benchmark's scale: nnz 1e7..1e8, fixed and
    # lognormal in-degree, N 1e5..1e6
    sp = S.large_specs()
    assert len(sp) == 40 and {s["kind"] for s in sp} == {"fixed", "lognormal"}
    nnz = np.array([s["n"] * s["k"] for s in sp])
    assert nnz.min() >= 1e7 * 0.999 and nnz.max() <= 1e8 * 1.001
    assert all(1e5 <= s["n"] <= 1e6 for s in sp)
