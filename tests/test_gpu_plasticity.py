"""Plastic spiking projections (generated pre / post event kernels) on the GPU.

* known answers of tests_python/SpikeSynapse/test_SpikingSynapse.py:
  test_PreSpike (w -= 1 : min=0 on every pre spike: 5 -> 1 -> 0) and
  test_PostSpike (w += 1 * constant : max=10 on every post spike: 5 -> 10);
* STDP (models/Synapses.py:232-310) and the t_pre / t_post rule of
  test_TimeDependentUpdate against the oracle (orc_plastic_run) fed with the
  spike trains the GPU run recorded: post neurons whose spikes do not depend
  on g, so every synapse's trajectory is fixed by the two trains.  FP64:
  weights and traces to 1e-12 (device exp against the host libm), the
  conductance added per step to 1e-12 (device atomics add in any order);
  FP32 to 1e-5.
"""

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import connectivity as conn
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SpkNeuron = ah.Neuron(parameters="v = 1", equations="mp = g_exc", spike="v > 0", name="SpkKAT")


def _kat_net(synapse):
    net = ah.Network(dt=1.0, seed=1)
    pop = net.create(3, SpkNeuron)
    proj = net.connect(pop, pop, "exc", synapse=synapse).all_to_all(5.0)
    net.compile()
    net.reset()
    return net, proj


def test_prespike_kat(gpu):
    syn = ah.Synapse(parameters="g = 1.0", pre_spike="w -= 1.0 : min=0.0", name="BoundedPre")
    net, proj = _kat_net(syn)
    np.testing.assert_allclose(proj.w, 5.0)
    net.simulate(5)
    np.testing.assert_allclose(proj.w, 1.0)
    net.simulate(5)
    np.testing.assert_allclose(proj.w, 0.0)
    np.testing.assert_allclose(proj.g, 1.0)


def test_postspike_kat(gpu):
    syn = ah.Synapse(parameters="constant = 1.0", post_spike="w += 1.0 * constant : max=10.0",
                     name="BoundedPost")
    net, proj = _kat_net(syn)
    np.testing.assert_allclose(proj.w, 5.0)
    net.simulate(5)
    np.testing.assert_allclose(proj.w, 10.0)
    net.simulate(5)
    np.testing.assert_allclose(proj.w, 10.0)
    # a per-synapse parameter set from the host changes the increment
    proj.w = 1.0
    proj.constant = 0.5
    net.simulate(4)
    np.testing.assert_allclose(proj.w, 3.0)


# Poisson trains; the post neurons read g_exc (reset after every step) but
# do not depend on it
PoissonPre = ah.Neuron(parameters="rates = 40.0", equations="p = Uniform(0.0, 1.0) * 1000.0 / dt",
                       spike="p < rates", name="PoissonPre")
PoissonPostG = ah.Neuron(parameters="rates = 30.0", equations="p = Uniform(0.0, 1.0) * 1000.0 / dt\nmp = g_exc",
                         spike="p < rates", name="PoissonPostG")
STDP_A = ah.STDP(tau_plus=20.0, tau_minus=20.0, A_plus=0.01, A_minus=0.0105, w_min=0.0, w_max=1.0)
STDP_A_PRM = [20.0, 20.0, 0.01, 0.0105, 0.0, 1.0]


TIME_RULE = ah.Synapse(
    parameters="""
        tau_pre = 10.0 : projection
        tau_post = 10.0 : projection
        cApre = 0.01 : projection
        cApost = 0.0105 : projection
        wmax = 0.01 : projection
    """,
    pre_spike="""
        g_target += w
        w = clip(w - cApost * exp((t_post - t)/tau_post) , 0.0 , wmax)
    """,
    post_spike="""
        w = clip(w + cApre * exp((t_pre - t)/tau_pre) , 0.0 , wmax)
    """,
    name="TimeDependent",
)


def _plastic_run(synapse, weights, steps=400, dt=0.5, precision="double", graph=True, delay=0.0, seed=11,
                 n_pre=300, n_post=200, p=0.2):
    net = ah.Network(dt=dt, seed=seed, precision=precision)
    pre = net.create(n_pre, PoissonPre)
    post = net.create(n_post, PoissonPostG)
    proj = net.connect(pre, post, "exc", synapse=synapse).fixed_probability(p, weights, delays=delay)
    m_pre = net.monitor(pre, ["spike"], capacity=steps)
    m_post = net.monitor(post, ["spike", "g_exc"], capacity=steps)
    net.compile(graph_batch=25 if graph else 0)
    net.set_graph(graph, 25)
    w0 = proj.w.copy()
    net.run(steps)
    pc, pi = m_pre.raw_spikes()
    qc, qi = m_post.raw_spikes()
    assert pc.sum() > 100 and qc.sum() > 100
    return net, proj, w0, (pc, pi, qc, qi), m_post.get("g_exc")


@pytest.mark.parametrize("graph", [True, False])
def test_stdp_matches_oracle(gpu, graph):
    net, proj, w0, trains, g = _plastic_run(STDP_A, conn.Uniform(0.2, 0.8, rng=np.random.default_rng(4)),
                                            graph=graph)
    csr = proj.csr
    ref = orc.plastic_run(0, csr.row_ptr, csr.col_idx, csr.n_cols, w0, STDP_A_PRM, 0.5, *trains)
    assert np.max(np.abs(ref["w"] - w0)) > 1e-3      # the weights did move
    np.testing.assert_allclose(proj.w, ref["w"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(proj.x, ref["x"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(proj.y, ref["y"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(proj.last_event, ref["last_event"])
    np.testing.assert_allclose(g, ref["g"], rtol=0, atol=1e-12)


def test_time_dependent_rule_matches_oracle(gpu):
    net, proj, w0, trains, g = _plastic_run(TIME_RULE, conn.Uniform(0.0, 0.01, rng=np.random.default_rng(5)))
    csr = proj.csr
    ref = orc.plastic_run(1, csr.row_ptr, csr.col_idx, csr.n_cols, w0, [10.0, 10.0, 0.01, 0.0105, 0.01], 0.5,
                          *trains)
    assert np.max(np.abs(ref["w"] - w0)) > 1e-4
    np.testing.assert_allclose(proj.w, ref["w"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(g, ref["g"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(proj.last_event, -10000)


def test_stdp_uniform_delay_matches_oracle(gpu):
    net, proj, w0, trains, g = _plastic_run(STDP_A, 0.5, delay=2.0, steps=300)
    csr = proj.csr
    assert proj.max_delay == 4
    ref = orc.plastic_run(0, csr.row_ptr, csr.col_idx, csr.n_cols, w0, STDP_A_PRM, 0.5, *trains, delay=4)
    np.testing.assert_allclose(proj.w, ref["w"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(g, ref["g"], rtol=0, atol=1e-12)


def test_stdp_fp32(gpu):
    net, proj, w0, trains, g = _plastic_run(STDP_A, conn.Uniform(0.2, 0.8, rng=np.random.default_rng(6)),
                                            precision="float", steps=300)
    csr = proj.csr
    ref = orc.plastic_run(0, csr.row_ptr, csr.col_idx, csr.n_cols, w0, STDP_A_PRM, 0.5, *trains)
    np.testing.assert_allclose(proj.w, ref["w"], rtol=0, atol=1e-5)
    np.testing.assert_allclose(g, ref["g"], rtol=0, atol=1e-4)


def test_plastic_attributes_and_errors(gpu):
    net = ah.Network(dt=1.0, seed=2)
    a = net.create(50, PoissonPre)
    b = net.create(40, PoissonPostG)
    proj = net.connect(a, b, "exc", synapse=STDP_A).fixed_number_pre(10, 0.3)
    proj.A_plus = 0.05                   # before compile: the initial value
    net.compile()
    assert proj.A_plus == 0.05 and proj.tau_plus == 20.0
    assert proj.x.shape == (proj.nb_synapses,) and np.all(proj.x == 0.0)
    np.testing.assert_array_equal(proj.last_event, -10000)
    xs = np.linspace(0.0, 1.0, proj.nb_synapses)
    proj.x = xs
    np.testing.assert_array_equal(proj.x, xs)
    np.testing.assert_allclose(proj.w, 0.3)
    with pytest.raises(ah._lib.AhipError):
        proj.delay = np.arange(proj.nb_synapses) % 3 + 1.0   # per-synapse delays: not with plasticity
    net.run(50)
    w = proj.w
    assert np.all(w >= 0.0) and np.all(w <= 1.0)
