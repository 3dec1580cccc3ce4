"""Connectivity drawn on the device (ahip_proj_generate).

* the rows equal a numpy restatement of the same Philox streams
  (tests/philox_ref.py, pinned by the Random123 known answers): the first
  `number` distinct candidates for fixed_number_pre, the geometric walk for
  fixed_probability, Uniform weights;
* invariants of the reference's connectors (Connector.pyx:306-380): exactly
  `number` synapses per row, ascending distinct columns, no self-connection
  unless allowed, ranks inside the (view) ranges; p = 1 is all-to-all and
  p = 0 empty;
* statistics: row counts of fixed_probability and column use of
  fixed_number_pre are those of the binomial / uniform distributions;
* a COBA network on device-generated connectivity is bit-exact with the
  oracle run on the same connectivity.
"""

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import connectivity as conn
from oracle import oracle as orc
import philox_ref as ref

pytestmark = pytest.mark.gpu


def _net(n, seed=3, spiking=False):
    net = ah.Network(dt=1.0, seed=seed)
    pop = net.create(n, ah.COBA if spiking else ah.RateNoisy)
    return net, pop


def _seed_of(proj_net_seed, nth=0):
    # the generator seed is the network generator's next 63-bit draw
    rng = np.random.default_rng(proj_net_seed)
    s = None
    for _ in range(nth + 1):
        s = int(rng.integers(1 << 63))
    return s


def _check_rows(csr, rows_expected, lo, hi, no_self):
    for r in range(csr.n_rows):
        seg = csr.col_idx[csr.row_ptr[r]:csr.row_ptr[r + 1]]
        assert np.all(np.diff(seg) > 0)
        if seg.size:
            assert seg[0] >= lo and seg[-1] < hi
        if no_self:
            assert r not in seg
    for r, exp in rows_expected.items():
        np.testing.assert_array_equal(csr.col_idx[csr.row_ptr[r]:csr.row_ptr[r + 1]], exp)


def test_fixed_number_pre_matches_stream(gpu):
    n, k = 3000, 50
    net, pop = _net(n)
    proj = net.connect(pop, pop, "exc").fixed_number_pre(k, conn.Uniform(-0.5, 0.5), generator="device")
    net.compile()
    csr = proj.csr
    seed = _seed_of(3)
    assert np.array_equal(np.diff(csr.row_ptr), np.full(n, k))
    rows = {r: ref.fixed_number_pre_row(seed, r, n, k, no_self=True) for r in (0, 1, 17, 999, 2999)}
    _check_rows(csr, rows, 0, n, no_self=True)
    for r in (0, 17, 2999):
        np.testing.assert_allclose(csr.w[csr.row_ptr[r]:csr.row_ptr[r + 1]], ref.uniform_weights(seed, r, k, -0.5, 0.5),
                                   rtol=0, atol=1e-15)
    np.testing.assert_array_equal(proj.w, csr.w)


def test_fixed_number_pre_views_and_self(gpu):
    n = 2000
    net, pop = _net(n, seed=9)
    pe = net.connect(pop[:1600], pop, "exc").fixed_number_pre(30, 0.6, generator="device")
    pi = net.connect(pop[1600:], pop[100:900], "exc").fixed_number_pre(400, 1.0, generator="device")
    net.compile()
    seed_e, seed_i = _seed_of(9, 0), _seed_of(9, 1)
    # a view of the same population is another pre/post object: self-connections allowed
    _check_rows(pe.csr, {5: ref.fixed_number_pre_row(seed_e, 5, 1600, 30)}, 0, 1600, no_self=False)
    cnt = np.diff(pi.csr.row_ptr)
    assert np.all(cnt[100:900] == 400) and cnt[:100].sum() == 0 and cnt[900:].sum() == 0
    _check_rows(pi.csr, {100: ref.fixed_number_pre_row(seed_i, 100, 400, 400, pre_begin=1600)}, 1600, 2000, False)
    assert pe.csr.w is None and pe.csr.w_const == 0.6


def test_fixed_probability_matches_stream_and_statistics(gpu):
    n, p = 4000, 0.03
    net, pop = _net(n, seed=4)
    proj = net.connect(pop, pop, "exc").fixed_probability(p, 1.0, generator="device")
    net.compile()
    seed = _seed_of(4)
    rows = {r: ref.fixed_probability_row(seed, r, n, p, no_self=True) for r in (0, 3, 2024, 3999)}
    _check_rows(proj.csr, rows, 0, n, no_self=True)
    cnt = np.diff(proj.csr.row_ptr)
    mean, var = (n - 1) * p, (n - 1) * p * (1 - p)
    assert abs(cnt.mean() - mean) < 4 * np.sqrt(var / n)
    assert abs(cnt.var() / var - 1.0) < 0.1


def test_fixed_probability_limits(gpu):
    net, pop = _net(300, seed=6)
    a = net.connect(pop, pop, "exc").fixed_probability(1.0, 1.0, generator="device")
    b = net.connect(pop, pop, "exc").fixed_probability(0.0, 1.0, generator="device")
    net.compile()
    assert a.csr.nnz == 300 * 299 and b.csr.nnz == 0
    _check_rows(a.csr, {7: np.delete(np.arange(300), 7)}, 0, 300, no_self=True)


def test_fixed_number_pre_column_use_is_uniform(gpu):
    n_post, n_pre, k = 20000, 1000, 100
    net = ah.Network(dt=1.0, seed=12)
    src = net.create(n_pre, ah.RateNoisy)
    dst = net.create(n_post, ah.RateNoisy)
    proj = net.connect(src, dst, "exc").fixed_number_pre(k, 1.0, generator="device")
    net.compile()
    use = np.bincount(proj.csr.col_idx, minlength=n_pre)
    expected = n_post * k / n_pre
    chi2 = np.sum((use - expected) ** 2 / expected)
    # chi-square with 999 degrees of freedom (without-replacement sampling only lowers it)
    assert chi2 < 999 + 5 * np.sqrt(2 * 999)


def test_device_connectivity_is_deterministic(gpu):
    csrs = []
    for _ in range(2):
        net, pop = _net(1500, seed=21)
        proj = net.connect(pop, pop, "exc").fixed_number_pre(64, conn.Normal(0, 0.1), generator="device")
        net.compile()
        csrs.append(proj.csr)
    assert np.array_equal(csrs[0].col_idx, csrs[1].col_idx) and np.array_equal(csrs[0].w, csrs[1].w)


def test_coba_on_device_connectivity_bit_exact(gpu):
    steps = 800
    net = ah.Network(dt=0.1, seed=31)
    P = net.create(4000, ah.COBA)
    ce = net.connect(P[:3200], P, "exc").fixed_probability(0.02, 0.6, generator="device")
    ci = net.connect(P[3200:], P, "inh").fixed_number_pre(16, 6.7, generator="device")
    P.v = ah.Uniform(-60.0, -50.0, rng=net.rng)
    mon = net.monitor(P, ["spike"], capacity=steps)
    net.compile(graph_batch=40)
    P.I = 20.0
    v0 = P.v.reshape(-1).copy()
    net.run(steps)
    projs = [orc.csrc_projection(ce.csr, 0), orc.csrc_projection(ci.csr, 1)]
    out = orc.spiking_run(0, P.size, 0.1, v0, steps, projs, refractory=np.asarray(P.refractory, dtype=np.int32))
    counts, idx = mon.raw_spikes()
    assert counts.sum() > 100
    np.testing.assert_array_equal(counts, out["counts"])
    np.testing.assert_array_equal(idx, out["idx"])
