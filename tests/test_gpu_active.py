"""Population enable()/disable() and Projection.transmission on the GPU.

Semantics of the reference (no test of its own covers them; the known answers
below are derived by hand from its generated code):
* an inactive population skips its update, spike gather, firing-rate window
  and delay-queue push (Population/OpenMPGenerator.py:454, 859, 972, 1125,
  1190): its variables stay frozen, its `spiked` list keeps the last gathered
  spikes (re-delivered to its efferent projections and re-recorded by its
  spike monitors every step, BaseTemplates.py:95-108), and its delay deque
  keeps its entries while the clock runs on;
* compute_psp, pre/post events and update_synapse run only when the
  projection transmits and its post population is active
  (Projection/OpenMP/CSR.py:258, 835-918); the sums of an active population
  are cleared every step (OpenMPTemplates.py:266-276), those of an inactive
  one are kept.
"""

import numpy as np
import pytest

import annarchy_amd as ah
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

DelayIn = ah.Neuron(equations="r = r + t : init = -1", name="DelayIn")
DelayOut = ah.Neuron(equations="r = sum(ff)", name="DelayOut")
Every3 = ah.Neuron(equations="v = v + 1.0\nmp = t", spike="v == 3.0", reset="v = 0.0", name="Every3")
Emit = ah.Neuron(spike="t==1", name="Emit")
Recv = ah.Neuron(equations="g_exc = 0", spike="g_exc>30", name="Recv")
PoissonPre = ah.Neuron(parameters="rates = 40.0", equations="p = Uniform(0.0, 1.0) * 1000.0 / dt",
                       spike="p < rates", name="PoissonPre")
PoissonPostG = ah.Neuron(parameters="rates = 30.0", equations="p = Uniform(0.0, 1.0) * 1000.0 / dt\nmp = g_exc",
                         spike="p < rates", name="PoissonPostG")
STDP_A = ah.STDP(tau_plus=20.0, tau_minus=20.0, A_plus=0.01, A_minus=0.0105, w_min=0.0, w_max=1.0)
STDP_A_PRM = [20.0, 20.0, 0.01, 0.0105, 0.0, 1.0]


def _r(t):
    """r of DelayIn at the end of step t (r = r + t from -1)."""
    return -1.0 + t * (t + 1) / 2.0


def _rate_net(delay, graph):
    net = ah.Network(dt=1.0, seed=5)
    p1 = net.create(3, DelayIn)
    p2 = net.create(3, DelayOut)
    proj = net.connect(p1, p2, "ff").fixed_number_pre(1, 1.0, delays=delay)
    net.compile()
    net.set_graph(graph, 3)
    net.reset()
    return net, p1, p2, proj


def _steps(net, pop, name, t0, t1):
    out = []
    for _ in range(t0, t1):
        net.simulate(1)
        out.append(float(np.asarray(pop.get(name)).reshape(-1)[0]))
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_disabled_pre_freezes_delay_queue(gpu, graph):
    # delay 3: the sum of step t is the third latest push before t; the pushes
    # of the disabled steps 6..9 are missing, so the queue resumes after them
    net, p1, p2, _ = _rate_net(3.0, graph)
    net.simulate(6)
    p1.disable()
    assert not p1.enabled
    assert _steps(net, p2, "_sum_ff", 6, 10) == [5.0, 5.0, 5.0, 5.0]
    np.testing.assert_array_equal(p1.r, _r(5))
    p1.enable()
    assert _steps(net, p2, "_sum_ff", 10, 14) == [5.0, 9.0, 14.0, _r(5) + 10]
    np.testing.assert_array_equal(p1.r, _r(5) + 10 + 11 + 12 + 13)


@pytest.mark.parametrize("graph", [False, True])
def test_disabled_post_keeps_its_sum(gpu, graph):
    net, p1, p2, _ = _rate_net(3.0, graph)
    net.simulate(6)
    p2.disable()
    assert _steps(net, p2, "_sum_ff", 6, 10) == [_r(2)] * 4
    np.testing.assert_array_equal(p2.r, _r(2))
    np.testing.assert_array_equal(p1.r, _r(9))
    p2.enable()
    assert _steps(net, p2, "_sum_ff", 10, 12) == [_r(7), _r(8)]
    np.testing.assert_array_equal(p2.r, _r(8))


def test_transmission_off_reads_zero(gpu):
    net, p1, p2, proj = _rate_net(3.0, False)
    net.simulate(6)
    assert proj.transmission
    proj.transmission = False
    assert not proj.transmission
    assert _steps(net, p2, "_sum_ff", 6, 9) == [0.0, 0.0, 0.0]
    proj.transmission = True
    # the pre population kept pushing into its queue meanwhile
    assert _steps(net, p2, "_sum_ff", 9, 11) == [_r(6), _r(7)]


def test_transmission_set_before_compile(gpu):
    net = ah.Network(dt=1.0)
    p1 = net.create(3, DelayIn)
    p2 = net.create(3, DelayOut)
    proj = net.connect(p1, p2, "ff").fixed_number_pre(1, 1.0)
    proj.transmission = False
    p1.disable()
    net.compile()
    net.simulate(3)
    np.testing.assert_array_equal(p2.r, 0.0)
    np.testing.assert_array_equal(p1.r, -1.0)   # never updated: the initial value
    p1.enable()
    proj.transmission = True
    net.simulate(1)
    np.testing.assert_array_equal(p1.r, 2.0)    # r = -1 + t at t = 3
    np.testing.assert_array_equal(p2.r, -1.0)   # r of p1 at the end of step 2


def _spike_net(delay):
    net = ah.Network(dt=1.0)
    inp = net.create(5, Emit)
    out = net.create(2, Recv)
    proj = net.connect(inp, out, "exc").all_to_all(1.0, delays=delay)
    g = net.monitor(out, ["g_exc"], capacity=12)
    sp = net.monitor(inp, ["spike"], capacity=12)
    net.compile()
    net.reset()
    return net, inp, out, proj, g, sp


def test_disabled_spiking_pre_with_delay(gpu):
    # the spikes of step 1 reach g_exc 3 pushes later: the four disabled steps
    # 2..5 push nothing, so they arrive at step 8 instead of 4
    net, inp, _, _, g, sp = _spike_net(3.0)
    net.simulate(2)
    inp.disable()
    net.simulate(4)
    inp.enable()
    net.simulate(4)
    expected = np.zeros((10, 2))
    expected[8] = 5.0
    np.testing.assert_array_equal(g.get("g_exc"), expected)
    # the spike list of step 1 stays the population's list while it is
    # disabled, and its monitor records it at every such step
    spikes = sp.get("spike")
    assert spikes == {i: [1, 2, 3, 4, 5] for i in range(5)}


def test_disabled_spiking_pre_redelivers_last_spikes(gpu):
    net, inp, _, _, g, _ = _spike_net(0.0)
    net.simulate(2)
    inp.disable()
    net.simulate(4)
    inp.enable()
    net.simulate(4)
    expected = np.zeros((10, 2))
    expected[2:7] = 5.0   # steps 2..6: the list of step 1 (gathered again only at step 6)
    np.testing.assert_array_equal(g.get("g_exc"), expected)


def test_spiking_transmission_off(gpu):
    net, inp, _, proj, g, _ = _spike_net(3.0)
    proj.transmission = False
    net.simulate(6)
    np.testing.assert_array_equal(g.get("g_exc"), np.zeros((6, 2)))


@pytest.mark.parametrize("graph", [False, True])
def test_disable_shifts_a_deterministic_neuron(gpu, graph):
    net = ah.Network(dt=1.0)
    pop = net.create(4, Every3)
    pop.compute_firing_rate(6.0)
    sp = net.monitor(pop, ["spike"], capacity=40)
    net.compile()
    net.set_graph(graph, 4)
    net.simulate(10)
    r_frozen = pop.r.copy()
    pop.disable()
    net.simulate(5)
    np.testing.assert_array_equal(pop.v, 1.0)
    np.testing.assert_array_equal(pop.mp, 9.0)
    np.testing.assert_array_equal(pop.r, r_frozen)
    pop.enable()
    net.simulate(10)
    np.testing.assert_array_equal(pop.v, 2.0)
    np.testing.assert_array_equal(pop.mp, 24.0)
    assert sp.get("spike") == {i: [2, 5, 8, 16, 19, 22] for i in range(4)}
    # the window (18, 24] holds the spikes of steps 19 and 22
    np.testing.assert_allclose(pop.r, 1000.0 / 6.0 * 2, rtol=1e-15)


def test_firing_rate_window_expires_while_disabled(gpu):
    # spikes at 2, 5, 8; disabled at 9..14; at step 15 the 6 ms window
    # (9, 15] holds no spike: the reference drops spike times <= t - 6 at its
    # next active step, so r = 0 although only one active step passed
    net = ah.Network(dt=1.0)
    pop = net.create(3, Every3)
    pop.compute_firing_rate(6.0)
    net.compile()
    net.simulate(9)
    np.testing.assert_allclose(pop.r, 1000.0 / 6.0 * 2, rtol=1e-15)
    pop.disable()
    net.simulate(6)
    np.testing.assert_allclose(pop.r, 1000.0 / 6.0 * 2, rtol=1e-15)
    pop.enable()
    net.simulate(1)
    np.testing.assert_array_equal(pop.r, 0.0)
    net.simulate(2)   # v = 1, 2, 3 at steps 15, 16, 17: a spike at 17
    np.testing.assert_allclose(pop.r, 1000.0 / 6.0, rtol=1e-15)


def test_plastic_projection_transmission_off(gpu):
    # no pre/post events and no update while transmission is off; afterwards
    # the projection learns from the later spikes alone, as the oracle's STDP
    # started at that step from the untouched state
    steps_off, steps_on, dt = 60, 200, 0.5
    net = ah.Network(dt=dt, seed=3)
    pre = net.create(200, PoissonPre)
    post = net.create(100, PoissonPostG)
    proj = net.connect(pre, post, "exc", synapse=STDP_A).fixed_probability(0.2, 0.5)
    m_pre = net.monitor(pre, ["spike"], capacity=steps_off + steps_on)
    m_post = net.monitor(post, ["spike", "g_exc"], capacity=steps_off + steps_on)
    net.compile()
    w0 = proj.w.copy()
    proj.transmission = False
    net.run(steps_off)
    np.testing.assert_array_equal(proj.w, w0)
    np.testing.assert_array_equal(proj.last_event, -10000)
    proj.transmission = True
    net.run(steps_on)
    pc, pi = m_pre.raw_spikes()
    qc, qi = m_post.raw_spikes()
    g = m_post.get("g_exc")
    np.testing.assert_array_equal(g[:steps_off], 0.0)
    # the first step with transmission delivers the pre spikes of the step
    # before it (the last one switched off), whose post spikes had no event:
    # the oracle starts at that step with its post spikes removed
    s0 = steps_off - 1
    a, b = int(pc[:s0].sum()), int(qc[:steps_off].sum())
    q_counts = qc[s0:].copy()
    q_counts[0] = 0
    csr = proj.csr
    ref = orc.plastic_run(0, csr.row_ptr, csr.col_idx, csr.n_cols, w0, STDP_A_PRM, dt, pc[s0:], pi[a:],
                          q_counts, qi[b:], t0=s0)
    assert pc[s0] > 0 and np.max(np.abs(ref["w"] - w0)) > 1e-3
    np.testing.assert_allclose(proj.w, ref["w"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(proj.last_event, ref["last_event"])
    np.testing.assert_allclose(g[s0:], ref["g"], rtol=0, atol=1e-12)
