"""Equation-language known answers on the GPU, from the reference's tests:
* tests_python/Common/test_ITE.py:25-112 -- one-line, nested multi-line and
  in-ODE if-then-else statements and the ite() function, 10 steps;
* tests_python/Common/test_BuiltinFunctions.py -- modulo, integer power and
  clip recorded by a monitor.
"""

import numpy as np
import pytest

import annarchy_amd as ah

pytestmark = pytest.mark.gpu

SimpleITE = ah.Neuron(equations="r = if (t is 1): 1.0 else: 0.0", name="SimpleITE")
NestedITE = ah.Neuron(parameters="state = 1.0", equations="""
    r = if (t <= 5):
            0
        else:
            if (t >= 7):
                2
            else:
                1
""", name="NestedITE")
ITEinODE = ah.Neuron(equations="dr/dt = if t > 5: 1 else: 0", name="ITEinODE")
SimpleITE2 = ah.Neuron(equations="r = ite(t is 1, 1.0, 0.0)", name="SimpleITE2")
NestedITE2 = ah.Neuron(parameters="state = 2.0", equations="r = ite(state >= 2, 2, ite(state == 1, 1, 0))",
                       name="NestedITE2")
ITEinODE2 = ah.Neuron(equations="dr/dt = ite((t > 5) and (t <20), 1, 0)", name="ITEinODE2")
BuiltinFuncs = ah.Neuron(parameters="base = 2.0", equations="""
    r = modulo(t,3)
    pr = power(base,3)
    clip_below = clip(-2, -1, 1)
    clip_within = clip(0, -1, 1)
    clip_above = clip(2, -1, 1)
""", name="BuiltinFuncs")


def test_if_then_else_kat(gpu):
    net = ah.Network(dt=1.0)
    pops = [net.create(1, n) for n in (SimpleITE, NestedITE, ITEinODE, SimpleITE2, NestedITE2, ITEinODE2)]
    m = net.monitor(pops[0], ["r"], capacity=10)
    m2 = net.monitor(pops[1], ["r"], capacity=10)
    net.compile()
    net.simulate(10)
    np.testing.assert_allclose([p.r[0] for p in pops], [0.0, 2.0, 4.0, 0.0, 2.0, 4.0])
    # the whole trajectories: 1 only at t = 1; 0 up to 5, 1 at 6, then 2
    np.testing.assert_array_equal(m.get("r")[:, 0], [0, 1, 0, 0, 0, 0, 0, 0, 0, 0])
    np.testing.assert_array_equal(m2.get("r")[:, 0], [0, 0, 0, 0, 0, 0, 1, 2, 2, 2])


def test_builtin_functions_kat(gpu):
    net = ah.Network(dt=1.0)
    pop = net.create(1, BuiltinFuncs)
    mon = net.monitor(pop, ["r", "pr", "clip_below", "clip_within", "clip_above"], capacity=10)
    net.compile()
    net.simulate(10)
    np.testing.assert_array_equal(mon.get("r", keep=True)[:, 0], [0, 1, 2, 0, 1, 2, 0, 1, 2, 0])
    np.testing.assert_array_equal(mon.get("pr", keep=True)[:, 0], 8.0)
    np.testing.assert_array_equal(mon.get("clip_below", keep=True)[:, 0], -1.0)
    np.testing.assert_array_equal(mon.get("clip_within", keep=True)[:, 0], 0.0)
    np.testing.assert_array_equal(mon.get("clip_above", keep=True)[:, 0], 1.0)
