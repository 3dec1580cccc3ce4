"""Host connectivity construction (numpy), replaying the reference connectors.

The reference builds connectivity in Cython before C++ sees it
(``cython_ext/Connector.pyx``): per post-synaptic neuron it draws the pre
ranks with the network's ``numpy.random.Generator`` and stores a LIL, which
``init_from_lil`` turns into the CSR forward view
(``ProjectionGenerator.py:743-797``, ``include/CSRMatrix.hpp:142-193``).
The functions below make the same numpy calls in the same order, so a given
seed produces the same synapses, and return the CSR arrays directly (bulk
ingest; no per-dendrite vectors for 1e8 synapses).

Weights: a number gives one constant weight (``core/Projection.py:823-839``);
a :class:`RandomDistribution` is sampled once per dendrite, after the pre
ranks of that dendrite, with ``get_values(size)`` (``Connector.pyx:322-328``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


# ---------------------------------------------------------------------------
# Random distributions (core/Random.py)
# ---------------------------------------------------------------------------
class RandomDistribution:
    def get_values(self, shape):
        raise NotImplementedError

    def get_list_values(self, size):
        return list(self.get_values(size))

    def get_value(self):
        return self.get_values(1)[0]


class Uniform(RandomDistribution):
    def __init__(self, min: float, max: float, rng: np.random.Generator | None = None):
        self.min, self.max = min, max
        self.rng = rng if rng is not None else np.random.default_rng()

    def get_values(self, shape):
        return self.rng.uniform(self.min, self.max, shape)


class DiscreteUniform(RandomDistribution):
    def __init__(self, min: int, max: int, rng: np.random.Generator | None = None):
        self.min, self.max = min, max
        self.rng = rng if rng is not None else np.random.default_rng()

    def get_values(self, shape):
        return self.rng.integers(self.min, self.max + 1, shape)


class Normal(RandomDistribution):
    def __init__(self, mu: float, sigma: float, min: float | None = None, max: float | None = None,
                 rng: np.random.Generator | None = None):
        self.mu, self.sigma, self.min, self.max = mu, sigma, min, max
        self.rng = rng if rng is not None else np.random.default_rng()

    def get_values(self, shape):
        data = self.rng.normal(self.mu, self.sigma, shape)
        if self.min is not None:
            data[data < self.min] = self.min
        if self.max is not None:
            data[data > self.max] = self.max
        return data


class LogNormal(RandomDistribution):
    def __init__(self, mu: float, sigma: float, min: float | None = None, max: float | None = None,
                 rng: np.random.Generator | None = None):
        self.mu, self.sigma, self.min, self.max = mu, sigma, min, max
        self.rng = rng if rng is not None else np.random.default_rng()

    def get_values(self, shape):
        data = self.rng.lognormal(self.mu, self.sigma, shape)
        if self.min is not None:
            data[data < self.min] = self.min
        if self.max is not None:
            data[data > self.max] = self.max
        return data


# ---------------------------------------------------------------------------
# CSR container
# ---------------------------------------------------------------------------
@dataclass
class CSR:
    """Forward view: rows = post ranks (all of them), columns = pre ranks."""

    n_rows: int
    n_cols: int
    row_ptr: np.ndarray      # int64 [n_rows + 1]
    col_idx: np.ndarray      # int32 [nnz]
    w: np.ndarray | None     # float64 [nnz] or None for a single weight
    w_const: float = 0.0

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    def weights(self) -> np.ndarray:
        if self.w is None:
            return np.full(self.nnz, self.w_const)
        return self.w

    def dense(self) -> np.ndarray:
        d = np.zeros((self.n_rows, self.n_cols))
        rows = np.repeat(np.arange(self.n_rows), np.diff(self.row_ptr))
        np.add.at(d, (rows, self.col_idx), self.weights())
        return d

    def post_ranks(self) -> np.ndarray:
        return np.nonzero(np.diff(self.row_ptr))[0].astype(np.int32)

    def to_lil(self):
        """(post_ranks, pre_ranks per dendrite, weights per dendrite)."""
        post = self.post_ranks()
        pre = [self.col_idx[self.row_ptr[r]:self.row_ptr[r + 1]] for r in post]
        w = self.weights()
        ww = [w[self.row_ptr[r]:self.row_ptr[r + 1]] for r in post]
        return post, pre, ww


def _assemble(n_rows, n_cols, rows_pre, rows_w, single_w, w_const):
    lens = np.zeros(n_rows, dtype=np.int64)
    for r, pre in rows_pre.items():
        lens[r] = len(pre)
    row_ptr = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(lens, out=row_ptr[1:])
    nnz = int(row_ptr[-1])
    col = np.empty(nnz, dtype=np.int32)
    w = None if single_w else np.empty(nnz, dtype=np.float64)
    for r, pre in rows_pre.items():
        col[row_ptr[r]:row_ptr[r + 1]] = pre
        if w is not None:
            w[row_ptr[r]:row_ptr[r + 1]] = rows_w[r]
    return CSR(n_rows, n_cols, row_ptr, col, w, float(w_const) if single_w else 0.0)


def _weights_for(weights, size):
    if isinstance(weights, RandomDistribution):
        return np.asarray(weights.get_values(size), dtype=np.float64)
    return None


def fixed_probability(pre_ranks, post_ranks, n_pre_total, n_post_total, rng, probability,
                      weights, allow_self_connections) -> CSR:
    """Connector.pyx:306-345 -- per post rank: rng.random(N_pre) < p, drop the
    self-connection unless allowed, sort."""
    pre_ranks = np.asarray(pre_ranks)
    single = not isinstance(weights, RandomDistribution)
    rows_pre, rows_w = {}, {}
    n_pre = len(pre_ranks)
    for r_post in list(np.asarray(post_ranks).tolist()):
        random_values = rng.random(n_pre)
        tmp = pre_ranks[random_values < probability]
        if not allow_self_connections:
            tmp = tmp[tmp != r_post]
        r = np.sort(tmp)
        if r.size == 0:
            continue
        rows_pre[r_post] = r
        if not single:
            rows_w[r_post] = _weights_for(weights, r.size)
    return _assemble(n_post_total, n_pre_total, rows_pre, rows_w, single, weights if single else 0.0)


def fixed_number_pre(pre_ranks, post_ranks, n_pre_total, n_post_total, rng, number, weights,
                     allow_self_connections) -> CSR:
    """Connector.pyx:347-380 -- per post rank: rng.choice(pre, number,
    replace=False), redrawn while it contains the post rank (unless allowed),
    sorted."""
    pre_list = list(np.asarray(pre_ranks).tolist())
    single = not isinstance(weights, RandomDistribution)
    rows_pre, rows_w = {}, {}
    for r_post in list(np.asarray(post_ranks).tolist()):
        r = rng.choice(pre_list, size=number, replace=False)
        if len(r) == 0:
            continue
        if not allow_self_connections:
            while r_post in list(r):
                r = rng.choice(pre_list, size=number, replace=False)
        r = np.sort(r)
        rows_pre[r_post] = r
        if not single:
            rows_w[r_post] = _weights_for(weights, number)
    return _assemble(n_post_total, n_pre_total, rows_pre, rows_w, single, weights if single else 0.0)


def all_to_all(pre_ranks, post_ranks, n_pre_total, n_post_total, weights,
               allow_self_connections) -> CSR:
    """Connector.pyx all_to_all: every pre rank, minus self unless allowed."""
    pre_ranks = np.asarray(pre_ranks)
    single = not isinstance(weights, RandomDistribution)
    rows_pre, rows_w = {}, {}
    for r_post in list(np.asarray(post_ranks).tolist()):
        r = pre_ranks if allow_self_connections else pre_ranks[pre_ranks != r_post]
        if r.size == 0:
            continue
        rows_pre[r_post] = np.sort(r)
        if not single:
            rows_w[r_post] = _weights_for(weights, r.size)
    return _assemble(n_post_total, n_pre_total, rows_pre, rows_w, single, weights if single else 0.0)


def one_to_one(pre_ranks, post_ranks, n_pre_total, n_post_total, weights) -> CSR:
    """Connector.pyx one_to_one: post rank k receives pre rank k of the views."""
    pre_ranks = np.asarray(pre_ranks)
    post_ranks = np.asarray(post_ranks)
    single = not isinstance(weights, RandomDistribution)
    rows_pre, rows_w = {}, {}
    for k, r_post in enumerate(post_ranks.tolist()):
        if k >= len(pre_ranks):
            break
        rows_pre[r_post] = np.array([pre_ranks[k]])
        if not single:
            rows_w[r_post] = _weights_for(weights, 1)
    return _assemble(n_post_total, n_pre_total, rows_pre, rows_w, single, weights if single else 0.0)


def random_in_degree(n_pre, n_post, rng, degrees, weights) -> CSR:
    """Bulk sparse pattern with a prescribed in-degree per post neuron (used
    for the lognormal in-degree benchmark): per post rank, ``degrees[i]``
    distinct pre ranks from ``rng.choice(n_pre, k, replace=False)``, sorted,
    then the dendrite's weights."""
    single = not isinstance(weights, RandomDistribution)
    degrees = np.asarray(degrees, dtype=np.int64)
    row_ptr = np.zeros(n_post + 1, dtype=np.int64)
    np.cumsum(degrees, out=row_ptr[1:])
    nnz = int(row_ptr[-1])
    col = np.empty(nnz, dtype=np.int32)
    w = None if single else np.empty(nnz, dtype=np.float64)
    for i in range(n_post):
        k = int(degrees[i])
        if k == 0:
            continue
        c = np.sort(rng.choice(n_pre, size=k, replace=False))
        col[row_ptr[i]:row_ptr[i + 1]] = c
        if w is not None:
            w[row_ptr[i]:row_ptr[i + 1]] = weights.get_values(k)
    return CSR(n_post, n_pre, row_ptr, col, w, float(weights) if single else 0.0)


def from_arrays(n_rows, n_cols, row_ptr, col_idx, w=None, w_const=0.0) -> CSR:
    """Bulk CSR ingest (validated)."""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
    if row_ptr.shape != (n_rows + 1,) or row_ptr[0] != 0 or np.any(np.diff(row_ptr) < 0):
        raise ValueError("row_ptr must be monotonic, start at 0 and have n_rows + 1 entries")
    if col_idx.shape != (int(row_ptr[-1]),):
        raise ValueError("col_idx must hold row_ptr[-1] entries")
    if col_idx.size and (col_idx.min() < 0 or col_idx.max() >= n_cols):
        raise ValueError("column index out of range")
    if w is not None:
        w = np.ascontiguousarray(w, dtype=np.float64)
        if w.shape != col_idx.shape:
            raise ValueError("w must hold one weight per synapse")
    return CSR(n_rows, n_cols, row_ptr, col_idx, w, float(w_const))
