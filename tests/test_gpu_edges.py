"""Edge cases of the step on the GPU: empty projections in every storage
format (the target must read 0, not the previous step's sums), an empty
projection sharing its target with a non-empty one, empty spiking and plastic
projections, and steps without any spike (the reference's tests build the
same cases with empty dendrites, tests_cpp/test_Matrix*_WithEmptyRows.cpp,
and silent populations)."""

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import connectivity as conn
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FORMATS = [("csr", {}), ("csr_vec", dict(lanes=16)), ("ell", {}), ("ellr", {}), ("sell", dict(lanes=4, sigma=64)),
           ("bsr", dict(chunk=4)), ("dense", {}), ("csr_panel", dict(lanes=8, panel_cols=256, chunk=128)),
           ("sell_panel", dict(lanes=2, sigma=8, panel_cols=300, chunk=128))]


def _empty(n_post, n_pre):
    return conn.from_arrays(n_pre, n_post, np.zeros(n_post + 1, np.int64), np.zeros(0, np.int32), np.zeros(0))


@pytest.mark.parametrize("fmt,kw", FORMATS, ids=[f[0] for f in FORMATS])
def test_empty_projection_writes_zero(gpu, fmt, kw):
    n = 300
    net = ah.Network(dt=1.0, seed=3)
    pop = net.create(n, ah.RateNoisy)
    net.connect(pop, pop, "exc").from_csr(_empty(n, n), storage_format=fmt, **kw)
    net.compile()
    pop.r = np.random.default_rng(1).random(n)
    pop.set("_sum_exc", 5.0)
    net.run(2)
    np.testing.assert_array_equal(pop.get("_sum_exc"), 0.0)


@pytest.mark.parametrize("fmt,kw", FORMATS, ids=[f[0] for f in FORMATS])
def test_empty_then_full_projection_same_target(gpu, fmt, kw):
    n = 256
    rng = np.random.default_rng(5)
    full = conn.random_in_degree(n, n, rng, rng.integers(0, 30, n), conn.Normal(0.0, 0.2, rng=rng))
    r0 = rng.random(n)
    net = ah.Network(dt=1.0, seed=3)
    pop = net.create(n, ah.RateNoisy)
    net.connect(pop, pop, "exc").from_csr(_empty(n, n), storage_format=fmt, **kw)
    net.connect(pop, pop, "exc").from_csr(full, storage_format="csr")
    net.compile()
    pop.r = r0
    pop.set("_sum_exc", 5.0)
    net.step()
    np.testing.assert_allclose(pop.get("_sum_exc").reshape(-1), orc.spmv_csr(full.row_ptr, full.col_idx, full.w, r0),
                               rtol=1e-12, atol=1e-14)


def test_empty_spiking_and_plastic_projections(gpu):
    steps = 300
    net = ah.Network(dt=0.1, seed=13)
    P = net.create(1000, ah.COBA)
    ce = net.connect(P[:800], P, "exc").fixed_probability(0.05, 0.6)
    ci = net.connect(P[800:], P, "inh").fixed_probability(0.05, 6.7)
    net.connect(P, P, "exc").fixed_probability(0.0, 1.0)                       # no synapse
    net.connect(P, P, "inh", synapse=ah.STDP()).fixed_probability(0.0, 0.5)    # plastic, no synapse
    P.v = ah.Uniform(-60.0, -50.0, rng=net.rng)
    mon = net.monitor(P, ["spike"], capacity=steps)
    net.compile(graph_batch=25)
    P.I = 20.0
    v0 = P.v.reshape(-1).copy()
    net.run(steps)
    ref = orc.spiking_run(0, P.size, 0.1, v0, steps, [orc.csrc_projection(ce.csr, 0), orc.csrc_projection(ci.csr, 1)],
                          refractory=np.asarray(P.refractory, dtype=np.int32))
    counts, idx = mon.raw_spikes()
    assert counts.sum() > 50
    np.testing.assert_array_equal(counts, ref["counts"])
    np.testing.assert_array_equal(idx, ref["idx"])


def test_silent_network_steps(gpu):
    """No neuron ever fires: compaction, scatter, plastic events and the spike
    monitor all see empty lists."""
    net = ah.Network(dt=0.1, seed=2)
    P = net.create(2000, ah.COBA)
    proj = net.connect(P, P, "exc", synapse=ah.STDP()).fixed_number_pre(20, 0.5)
    net.connect(P, P, "inh").fixed_number_pre(20, 1.0)
    mon = net.monitor(P, ["spike"], capacity=200)
    net.compile(graph_batch=20)
    P.I = 0.0
    P.v = -60.0
    net.run(200)
    counts, idx = mon.raw_spikes()
    assert counts.sum() == 0 and idx.size == 0
    np.testing.assert_array_equal(proj.w, 0.5)
    np.testing.assert_array_equal(proj.last_event, -10000)
    np.testing.assert_allclose(P.v, -60.0)
