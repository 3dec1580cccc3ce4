"""Synaptic delays on the GPU.

Known answers of the reference tests, run through the same Python calls:
* tests_python/RateSynapse/test_RateDelays.py:168-400 -- uniform delay 10 ms
  (and 5 ms after a change) and per-synapse delays [[3], [5], [2]] on the
  input sequence r_t = -1, 0, 2, 5, 9, 14, 20, ...;
* tests_python/SpikeSynapse/test_SpikingTransmission.py:86-253 -- 5 emitters
  firing at t = 1 reach g_exc at t = 4 with a 3 ms delay, and at t = 2, 3, 4
  with delays [[1,2,3,2,1],[3,2,1,2,3]].
Random networks against the oracle's delay queues (orc_rate_run_delayed,
orc_spiking_run with uniform and per-synapse delays): rates to 1e-9, spike
lists bit for bit, also through captured graphs.
"""

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import connectivity as conn
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

DelayIn = ah.Neuron(equations="r = r + t : init = -1", name="DelayIn")
DelayOut = ah.Neuron(equations="r = sum(ff)", name="DelayOut")


def _rate_kat_net(delays):
    net = ah.Network(dt=1.0, seed=5)
    p1 = net.create(3, DelayIn)
    p2 = net.create(3, DelayOut)
    proj = net.connect(p1, p2, "ff").fixed_number_pre(1, 1.0, delays=delays)
    net.compile()
    net.reset()
    return net, p1, p2, proj


def test_rate_uniform_delay_kat(gpu):
    net, p1, p2, proj = _rate_kat_net(10.0)
    np.testing.assert_allclose(proj.delay, 10.0)
    net.simulate(10)
    np.testing.assert_allclose(p2.get("_sum_ff"), -1.0)
    net.simulate(1)
    np.testing.assert_allclose(p2.get("_sum_ff"), -1.0)
    net.simulate(4)
    np.testing.assert_allclose(p2.get("_sum_ff"), 9.0)


def test_rate_modified_uniform_delay_kat(gpu):
    net, p1, p2, proj = _rate_kat_net(10.0)
    proj.delay = 5.0
    np.testing.assert_allclose(proj.delay, 5.0)
    net.simulate(5)
    np.testing.assert_allclose(p2.get("_sum_ff"), -1.0)
    net.simulate(1)
    np.testing.assert_allclose(p2.get("_sum_ff"), -1.0)
    net.simulate(4)
    np.testing.assert_allclose(p2.get("_sum_ff"), 9.0)


def test_rate_nonuniform_delay_kat(gpu):
    net, p1, p2, proj = _rate_kat_net(conn.DiscreteUniform(1, 5, rng=np.random.default_rng(3)))
    proj.delay = [[3], [5], [2]]
    net.simulate(5)
    np.testing.assert_allclose(p2.get("_sum_ff"), [0.0, -1.0, 2.0])
    # reset, then the same delay for every synapse (test_modified_nonuniform_delay)
    net.reset()
    proj.delay = [[3.0], [3.0], [3.0]]
    net.simulate(10)
    np.testing.assert_allclose(p2.get("_sum_ff"), [20.0, 20.0, 20.0])


def _spike_kat_net(delays):
    emit = ah.Neuron(spike="t==1", name="Emit")
    recv = ah.Neuron(equations="g_exc = 0", spike="g_exc>30", name="Recv")
    net = ah.Network(dt=1.0)
    inp = net.create(5, emit)
    out = net.create(2, recv)
    proj = net.connect(inp, out, "exc").all_to_all(1.0, delays=delays)
    mon = net.monitor(out, ["g_exc"], capacity=10)
    net.compile()
    net.reset()
    return net, proj, mon


def test_spiking_uniform_delay_kat(gpu):
    net, proj, mon = _spike_kat_net(3.0)
    net.simulate(6)
    np.testing.assert_array_equal(mon.get("g_exc"), [[0, 0], [0, 0], [0, 0], [0, 0], [5, 5], [0, 0]])


def test_spiking_nonuniform_delay_kat(gpu):
    net, proj, mon = _spike_kat_net(conn.DiscreteUniform(2, 10, rng=np.random.default_rng(4)))
    proj.delay = [[1.0, 2.0, 3.0, 2.0, 1.0], [3.0, 2.0, 1.0, 2.0, 3.0]]
    net.reset()
    net.simulate(5)
    np.testing.assert_array_equal(mon.get("g_exc"), [[0, 0], [0, 0], [2, 1], [2, 2], [1, 2]])


def _random_csr(n, k, seed):
    rng = np.random.default_rng(seed)
    return conn.random_in_degree(n, n, rng, np.full(n, k), conn.Normal(0.0, 0.9 / np.sqrt(k), rng=rng)), rng


@pytest.mark.parametrize("case", ["uniform", "per_synapse"])
@pytest.mark.parametrize("graph", [False, True])
def test_rate_network_delays_match_oracle(gpu, case, graph):
    n, steps = 1500, 60
    csr, rng = _random_csr(n, 40, 21)
    if case == "uniform":
        delays, max_delay, kw = 4.0, 4, dict(delay=4)
    else:
        d = rng.integers(1, 7, csr.nnz)
        delays, max_delay, kw = d.astype(float), 6, dict(delays=d)
    r0 = rng.random(n)
    noise = rng.normal(0.0, 0.1, (steps, n))
    net = ah.Network(dt=1.0, seed=2)
    pop = net.create(n, ah.RateNoisy)
    net.connect(pop, pop, "exc").from_csr(csr, storage_format="csr", lanes=8, delays=delays)
    net.compile(graph_batch=12 if graph else 0)
    net.set_graph(graph, 12)
    pop.r = r0
    pop.inject_random(0, noise)
    net.run(steps)
    # the queue starts at zero (no reset), as after the reference's initialisation
    r_ref, _, _ = orc.rate_run_delayed(csr.row_ptr, csr.col_idx, csr.w, r0, 10.0, 1.0, noise, steps, max_delay, **kw)
    got = pop.r.reshape(-1)
    assert np.max(np.abs(got - r_ref)) / np.max(np.abs(r_ref)) < 1e-9


def _coba_delay_net(d_exc, d_inh, steps):
    net = ah.Network(dt=0.1, seed=13)
    P = net.create(3000, ah.COBA)
    ce = net.connect(P[:2400], P, "exc").fixed_probability(0.02, 0.6, delays=d_exc)
    ci = net.connect(P[2400:], P, "inh").fixed_probability(0.02, 6.7, delays=d_inh)
    P.v = ah.Uniform(-60.0, -50.0, rng=net.rng)
    mon = net.monitor(P, ["spike"], capacity=steps)
    net.compile(graph_batch=25)
    P.I = 20.0
    return net, P, ce, ci, mon


@pytest.mark.parametrize("case", ["uniform", "per_synapse"])
def test_coba_delays_bit_exact_vs_oracle(gpu, case):
    steps = 1500
    if case == "uniform":
        net, P, ce, ci, mon = _coba_delay_net(0.5, 0.3, steps)   # 5 and 3 steps
        pe = orc.csrc_projection(ce.csr, 0, delay=5)
        pi = orc.csrc_projection(ci.csr, 1, delay=3)
    else:
        net, P, ce, ci, mon = _coba_delay_net(conn.Uniform(0.1, 0.8, rng=np.random.default_rng(8)), 0.3,
                                              steps)
        d_exc = np.rint(np.asarray(ce.delay) / 0.1).astype(int)
        assert d_exc.min() >= 1 and d_exc.max() <= 8 and len(np.unique(d_exc)) > 4
        pe = orc.csrc_projection(ce.csr, 0, delays=d_exc)
        pi = orc.csrc_projection(ci.csr, 1, delay=3)
    v0 = P.v.reshape(-1).copy()
    net.run(steps)
    ref = orc.spiking_run(0, P.size, 0.1, v0, steps, [pe, pi], refractory=np.asarray(P.refractory, dtype=np.int32))
    counts, idx = mon.raw_spikes()
    assert counts.sum() > 100
    np.testing.assert_array_equal(counts, ref["counts"])
    np.testing.assert_array_equal(idx, ref["idx"])
    assert np.array_equal(P.v.reshape(-1), ref["v"])
