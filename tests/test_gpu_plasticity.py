"""Plastic spiking projections (generated pre / post event kernels) on the GPU.

* known answers of tests_python/SpikeSynapse/test_SpikingSynapse.py:
  test_PreSpike (w -= 1 : min=0 on every pre spike: 5 -> 1 -> 0) and
  test_PostSpike (w += 1 * constant : max=10 on every post spike: 5 -> 10);
* STDP (models/Synapses.py:232-310), STP (:131-230) and the t_pre / t_post
  rule of test_TimeDependentUpdate against the oracle (orc_plastic_run) fed
  with the spike trains the GPU run recorded: post neurons whose spikes do
  not depend on g, so every synapse's trajectory is fixed by the two trains.
  FP64: weights and traces to 1e-12 (device exp against the host libm), the
  conductance added per step to 1e-12 (device atomics add in any order);
  FP32 to 1e-5;
* rate-coded update_synapse: test_RateDefaultSynapseModels (Hebb, Oja, IBCM
  transmit w * pre.r), and the three rules against orc_rate_plastic_run with
  sinusoidal pre rates (relative 1e-10: device sin against numpy), with the
  update period of enable_learning and with learning disabled;
* continuous spiking update (the pattern of test_SpikingContinuousUpdate):
  a hand-derived weight trajectory of traces set by pre / post events and
  decremented every step.
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
    np.testing.assert_allclose(proj.dendrite(0).w, [5.0, 5.0])
    net.simulate(5)
    np.testing.assert_allclose(proj.dendrite(0).w, [1.0, 1.0])
    net.simulate(5)
    np.testing.assert_allclose(proj.dendrite(0).w, [0.0, 0.0])
    np.testing.assert_allclose(proj.w, 0.0)
    np.testing.assert_allclose(proj.g, 1.0)
    # the dendrite view: pre ranks of post neuron 1 (no self-connection), its size
    assert proj.dendrite(1).pre_ranks == [0, 2] and proj.dendrite(1).size == 2
    assert proj.post_ranks == [0, 1, 2] and proj.nb_dendrites == 3


def test_postspike_kat(gpu):
    syn = ah.Synapse(parameters="constant = 1.0", post_spike="w += 1.0 * constant : max=10.0",
                     name="BoundedPost")
    net, proj = _kat_net(syn)
    np.testing.assert_allclose(proj.dendrite(0).w, [5.0, 5.0])
    net.simulate(5)
    np.testing.assert_allclose(proj.dendrite(0).w, [10.0, 10.0])
    net.simulate(5)
    np.testing.assert_allclose(proj.dendrite(0).w, [10.0, 10.0])
    # a per-synapse parameter set from the host changes the increment
    proj.w = 1.0
    proj.constant = 0.5
    proj.dendrite(2).constant = 0.25
    net.simulate(4)
    np.testing.assert_allclose(proj.dendrite(0).w, [3.0, 3.0])
    np.testing.assert_allclose(proj.dendrite(2).w, [2.0, 2.0])


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


def test_stdp_disable_learning_keeps_weights(gpu):
    net, proj, w0, trains, g = None, None, None, None, None
    net = ah.Network(dt=0.5, seed=11)
    pre = net.create(300, PoissonPre)
    post = net.create(200, PoissonPostG)
    proj = net.connect(pre, post, "exc", synapse=STDP_A).fixed_probability(0.2, 0.4)
    m_pre = net.monitor(pre, ["spike"], capacity=200)
    m_post = net.monitor(post, ["spike"], capacity=200)
    net.compile(graph_batch=25)
    proj.disable_learning()
    net.run(200)
    np.testing.assert_array_equal(proj.w, 0.4)
    csr = proj.csr
    ref = orc.plastic_run(0, csr.row_ptr, csr.col_idx, csr.n_cols, 0.4, STDP_A_PRM, 0.5, *m_pre.raw_spikes(),
                          *m_post.raw_spikes())
    # the traces evolve as with learning (they do not depend on w)
    np.testing.assert_allclose(proj.x, ref["x"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(proj.y, ref["y"], rtol=0, atol=1e-12)


STP_A = ah.STP(tau_rec=100.0, tau_facil=50.0, U=0.3)


def test_stp_matches_oracle(gpu):
    net, proj, w0, trains, g = _plastic_run(STP_A, conn.Uniform(0.5, 1.5, rng=np.random.default_rng(9)))
    csr = proj.csr
    ref = orc.plastic_run(2, csr.row_ptr, csr.col_idx, csr.n_cols, w0, [100.0, 50.0, 0.3], 0.5, *trains,
                          x0=np.ones(csr.nnz), y0=np.full(csr.nnz, 0.3))
    np.testing.assert_array_equal(proj.w, w0)
    np.testing.assert_allclose(proj.x, ref["x"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(proj.u, ref["y"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(g, ref["g"], rtol=0, atol=1e-12)
    assert proj.U.shape == (csr.nnz,) and np.all(proj.U == 0.3)


# -- rate-coded update_synapse ----------------------------------------------------

RNeuron = ah.Neuron(equations="r=sum(Hebb)+sum(Oja)+sum(IBCM)", name="RDefault")


def test_rate_default_synapse_models_kat(gpu):
    net = ah.Network(dt=1.0, seed=1)
    inp = net.create(1, RNeuron)
    out = net.create(1, RNeuron)
    projs = [net.connect(inp, out, name, synapse=syn).all_to_all(0.1)
             for name, syn in (("Hebb", ah.Hebb()), ("Oja", ah.Oja()), ("IBCM", ah.IBCM()))]
    net.compile()
    net.reset()
    net.disable_learning()
    inp.r = 1.0
    net.simulate(1)
    for name in ("Hebb", "Oja", "IBCM"):
        np.testing.assert_allclose(out.get("_sum_" + name), 0.1)
    for p in projs:
        np.testing.assert_array_equal(p.w, 0.1)


SinRate = ah.Neuron(parameters="phase = 0.0", equations="r = 0.5 + 0.5*sin(0.1*t + phase)", name="SinRate")
SumRate = ah.Neuron(equations="r = sum(exc)", name="SumRate")
RULES = {"hebb": (0, ah.Hebb(eta=0.0005), [0.0005]), "oja": (1, ah.Oja(eta=0.01, alpha=1.0), [0.01, 1.0]),
         "ibcm": (2, ah.IBCM(eta=0.01, tau=50.0), [0.01, 50.0])}


def _rate_plastic(rule, steps=120, precision="double", period=None, learning=True):
    kind, syn, prm = RULES[rule]
    net = ah.Network(dt=1.0, seed=3, precision=precision)
    pre = net.create(100, SinRate)
    post = net.create(50, SumRate)
    proj = net.connect(pre, post, "exc", synapse=syn).fixed_probability(
        0.3, conn.Uniform(0.05, 0.2, rng=np.random.default_rng(7)))
    mon = net.monitor(post, ["r"], capacity=steps)
    phase = np.linspace(0.0, 2.0 * np.pi, 100)
    pre.phase = phase
    net.compile(graph_batch=20)
    if period is not None:
        proj.enable_learning(period=period)
    if not learning:
        proj.disable_learning()
    w0 = proj.w.copy()
    net.run(steps)
    t = np.arange(steps, dtype=np.float64)
    seq = 0.5 + 0.5 * np.sin(0.1 * t[:, None] + phase[None, :])
    csr = proj.csr
    ref = orc.rate_plastic_run(kind, csr.row_ptr, csr.col_idx, w0, prm, 1.0, 0.0, seq,
                               period=int(period) if period else 1)
    return proj, w0, mon.get("r"), ref


@pytest.mark.parametrize("rule", ["hebb", "oja", "ibcm"])
def test_rate_rules_match_oracle(gpu, rule):
    proj, w0, r, ref = _rate_plastic(rule)
    assert np.max(np.abs(ref["w"] - w0)) > 1e-3
    np.testing.assert_allclose(proj.w, ref["w"], rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(r, ref["sum"], rtol=1e-10, atol=1e-14)
    if rule == "ibcm":
        np.testing.assert_allclose(proj.theta, ref["theta"], rtol=1e-10)


def test_rate_update_period_matches_oracle(gpu):
    proj, w0, r, ref = _rate_plastic("oja", period=3.0)
    np.testing.assert_allclose(proj.w, ref["w"], rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(r, ref["sum"], rtol=1e-10, atol=1e-14)


def test_rate_disable_learning(gpu):
    proj, w0, r, ref = _rate_plastic("ibcm", steps=30, learning=False)
    np.testing.assert_array_equal(proj.w, w0)
    np.testing.assert_array_equal(proj.theta, 0.0)     # rate-coded: no update_synapse at all


def test_rate_oja_fp32(gpu):
    proj, w0, r, ref = _rate_plastic("oja", precision="float")
    np.testing.assert_allclose(proj.w, ref["w"], rtol=1e-4, atol=1e-6)


# -- continuous spiking update ------------------------------------------------------

def test_spiking_continuous_update_kat(gpu):
    pre_n = ah.Neuron(equations="v = v + 1.0\nmp = t", spike="v == 3.0", reset="v = 0.0", name="Every3")
    post_n = ah.Neuron(equations="v = v + 1.0\nmp = t\nq = g_exc", spike="v == 5.0", reset="v = 0.0",
                       name="Every5")
    syn = ah.Synapse(
        pre_spike="pre_trace = pre.mp",
        post_spike="post_trace = post.mp",
        equations="""
            post_trace = post_trace - 1 : min=0.0
            pre_trace = pre_trace - 1 : min=0.0
            w = post_trace + pre_trace
        """, name="Traces")
    net = ah.Network(dt=1.0)
    a = net.create(1, pre_n)
    b = net.create(1, post_n)
    proj = net.connect(a, b, "exc", synapse=syn).all_to_all(0.0)
    mon = net.monitor(b, ["g_exc"], capacity=11)
    net.compile()
    ws = []
    for _ in range(11):
        net.simulate(1)
        ws.append(proj.w[0])
    # pre spikes at 2, 5, 8 (handled at 3, 6, 9, setting pre_trace = t-1);
    # post spikes at 4, 9 (post_trace = t after the update); both traces
    # decrease by 1 per step, w = their sum
    np.testing.assert_array_equal(ws, [0, 0, 0, 1, 0, 3, 6, 4, 2, 7, 14])
    g = mon.get("g_exc")[:, 0]
    np.testing.assert_array_equal(np.flatnonzero(g), [6, 9])
    np.testing.assert_array_equal(g[[6, 9]], [3.0, 2.0])
