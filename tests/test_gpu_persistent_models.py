"""Persistent steps of other spiking models against the per-kernel step.

The persistent kernel splices the same generated statements as the
per-kernel update, with the neuron's values held in registers, the
touched local slots and conductances in LDS and the refractory periods in
registers.  These models exercise what COBA does not: a local random
variable (Philox draws keyed by neuron, variable and step), local
parameters, exponential Euler, a refractory period given by a local
variable, a reset of two variables, a projection into a conductance the
equations do not name by its target alone.  Every run: spike lists and
every state variable bit-identical between the two step forms after 1000
steps (five graph batches of 200).
"""

import numpy as np
import pytest

import annarchy_amd as ah

pytestmark = pytest.mark.gpu

Izhikevich = ah.Neuron(parameters="""
    a = 0.02
    b = 0.2
    c = -65.0
    d = 8.0
    I_ext = 5.0
    sigma = 2.0     : population
    tau_e = 5.0     : population
    tau_i = 10.0    : population
""", equations="""
    I = I_ext + g_exc - g_inh + sigma * Normal(0.0, 1.0)
    dv/dt = 0.04 * v * v + 5.0 * v + 140.0 - u + I
    du/dt = a * (b * v - u)
    tau_e * dg_exc/dt = -g_exc
    tau_i * dg_inh/dt = -g_inh
""", spike="v >= 30.0", reset="""
    v = c
    u += d
""", name="IzhikevichNoisy")

LIFExp = ah.Neuron(parameters="""
    tau_m = 10.0    : population
    E_L = -65.0     : population
    V_th = -52.0    : population
    V_r = -67.0     : population
    tau_ref = 2.5
    tau_s = 3.0     : population
    I_b = 14.0      : population
""", equations="""
    tau_m * dv/dt = E_L - v + g_exc - g_inh + I_b : exponential
    tau_s * dg_exc/dt = -g_exc : exponential
    tau_s * dg_inh/dt = -g_inh : exponential
""", spike="v > V_th", reset="v = V_r", refractory="tau_ref", name="LIFExpRefVar")


def _net(model, persistent, n=3000, ne=2400, seed=11, precision="double", init=None):
    net = ah.Network(dt=0.25, seed=seed, precision=precision)
    P = net.create(n, model)
    net.connect(P[:ne], P, "exc").fixed_probability(0.03, 0.5)
    net.connect(P[ne:], P, "inh").fixed_probability(0.03, 2.0)
    mon = net.monitor(P, ["spike"], capacity=1500)
    net.compile(graph_batch=200)
    if init:
        init(P, net)
    net.set_persistent(persistent)
    return net, P, mon


def _izh_init(P, net):
    rng = np.random.default_rng(3)
    P.v = rng.uniform(-70.0, -60.0, P.size)
    P.u = 0.2 * np.asarray(P.v).reshape(-1)
    P.I_ext = rng.uniform(2.0, 8.0, P.size)


def _lif_init(P, net):
    rng = np.random.default_rng(5)
    P.v = rng.uniform(-65.0, -53.0, P.size)
    P.tau_ref = rng.uniform(1.0, 4.0, P.size)


@pytest.mark.parametrize("model,init,names,precision", [
    (Izhikevich, _izh_init, ["v", "u", "g_exc", "g_inh"], "double"),
    (Izhikevich, _izh_init, ["v", "u", "g_exc", "g_inh"], "float"),
    (LIFExp, _lif_init, ["v", "g_exc", "g_inh"], "double"),
])
def test_persistent_models_equal_per_kernel(gpu, model, init, names, precision):
    out = {}
    for persistent in (True, False):
        net, P, mon = _net(model, persistent, precision=precision, init=init)
        assert net.uses_persistent() is persistent
        net.run(1000)
        counts, idx = mon.raw_spikes()
        out[persistent] = (counts, idx, [np.asarray(P.get(n)).copy() for n in names],
                           np.asarray(P.last_spike).copy() if hasattr(P, "last_spike") else None)
    a, b = out[True], out[False]
    assert a[0].sum() > 500, "network is silent"
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for x, y, n in zip(a[2], b[2], names):
        assert np.array_equal(x, y), n
