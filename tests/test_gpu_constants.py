"""Network constants on the GPU (core/Constant.py:16-120, core/Network.py:
482-512): equations read a named value that set() changes after compile,
through ahip_net_set_constant (the set_<constant> of the generated module).

Known answers by hand: r = r + k accumulates n*k, the exponential-Euler
solution of tau*dv/dt = -v + 1 from 0 is 1 - exp(-dt/tau)^n, a counter with
threshold th fires every th steps, and a rate synapse w = w + eta moves by
eta per step.
"""

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import network as ahnet

pytestmark = pytest.mark.gpu

ConstStep = ah.Neuron(equations="r = r + k_c", name="ConstStep")
ConstHidden = ah.Neuron(parameters="k_c = 0.5", equations="r = r + k_c", name="ConstHidden")
ConstTau = ah.Neuron(equations="tau_c * dv/dt = -v + 1.0 : exponential\nr = v", name="ConstTau")
ConstSpike = ah.Neuron(equations="v = v + 1.0", spike="v >= th_c", reset="v = 0.0", name="ConstSpike")
ConstSyn = ah.Synapse(equations="w = w + eta_c", name="ConstSyn")
DelayOut = ah.Neuron(equations="r = sum(ff)", name="DelayOut")


def test_constant_set_after_compile(gpu):
    net = ah.Network(dt=1.0)
    k = net.constant("k_c", 2.0)
    a = net.create(3, ConstStep)
    b = net.create(3, ConstHidden)
    net.compile()
    assert net.get_constant("k_c") is k
    net.simulate(5)
    np.testing.assert_array_equal(a.r, 10.0)
    np.testing.assert_array_equal(b.r, 2.5)       # its own parameter hides the constant
    k.set(3.0)
    assert k.value == 3.0
    net.simulate(5)
    np.testing.assert_array_equal(a.r, 25.0)
    np.testing.assert_array_equal(b.r, 5.0)
    # the C-ABI reports which objects mirror the constant
    import ctypes as C

    from annarchy_amd import _lib
    n = C.c_int32()
    _lib.check(_lib.lib().ahip_net_set_constant(net._h, b"k_c", 1.0, C.byref(n)))
    assert n.value == 1


def test_constant_in_exponential_step_size(gpu):
    net = ah.Network(dt=1.0)
    tau = net.constant("tau_c", 10.0)
    pop = net.create(2, ConstTau)
    net.compile()
    net.simulate(7)
    np.testing.assert_allclose(pop.v, 1.0 - np.exp(-0.1) ** 7, rtol=0, atol=1e-14)
    tau.set(20.0)   # the hoisted step size follows the new value
    net.simulate(3)
    v7 = 1.0 - np.exp(-0.1) ** 7
    np.testing.assert_allclose(pop.v, 1.0 - (1.0 - v7) * np.exp(-0.05) ** 3, rtol=0, atol=1e-14)


@pytest.mark.parametrize("graph", [False, True])
def test_constant_threshold(gpu, graph):
    net = ah.Network(dt=1.0)
    th = net.constant("th_c", 3.0)
    pop = net.create(2, ConstSpike)
    mon = net.monitor(pop, ["spike"], capacity=30)
    net.compile()
    net.set_graph(graph, 4)
    net.simulate(9)
    th.set(2.0)
    net.simulate(6)
    # v = 1, 2, 3 -> spikes at 2, 5, 8 with 3; then every 2 steps: 10, 12, 14
    assert mon.get("spike") == {0: [2, 5, 8, 10, 12, 14], 1: [2, 5, 8, 10, 12, 14]}


def test_constant_in_synapse(gpu):
    net = ah.Network(dt=1.0)
    eta = net.constant("eta_c", 0.25)
    a = net.create(4, ConstStep)
    b = net.create(2, DelayOut)
    net.constant("k_c", 0.0)
    proj = net.connect(a, b, "ff", synapse=ConstSyn).all_to_all(1.0, storage_format="csr")
    net.compile()
    assert "eta_c" not in proj._syn_info()
    net.simulate(4)
    np.testing.assert_array_equal(proj.w, 2.0)
    eta.set(-0.5)
    net.simulate(2)
    np.testing.assert_array_equal(proj.w, 1.0)


@pytest.fixture
def module_constant():
    c = ah.Constant("k_c", 1.5)
    yield c
    ahnet._GLOBAL_CONSTANTS.pop("k_c", None)


def test_module_level_constant_reaches_networks(gpu, module_constant):
    net1 = ah.Network(dt=1.0)
    net2 = ah.Network(dt=1.0)
    p1 = net1.create(1, ConstStep)
    p2 = net2.create(1, ConstStep)
    net1.compile()
    net2.compile()
    net1.simulate(2)
    net2.simulate(2)
    np.testing.assert_array_equal(p1.r, 3.0)
    module_constant.set(0.5)            # every network that copied it
    net2.get_constant("k_c").set(4.0)   # this network only
    net1.simulate(2)
    net2.simulate(2)
    np.testing.assert_array_equal(p1.r, 4.0)
    np.testing.assert_array_equal(p2.r, 11.0)
