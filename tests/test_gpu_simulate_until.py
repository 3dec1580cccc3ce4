"""simulate_until on the GPU: the known answers of the reference's
tests_python/Common/test_SimulateUntil.py (a counter r += 1 with the stop
condition "r > 5.0 : any", one population started at r in {-4, -3, -2}):
8 steps for that population alone or with 'and', 6 with 'or'; plus the
"spiked : any" condition and a network without any stop condition."""

import numpy as np
import pytest

import annarchy_amd as ah

pytestmark = pytest.mark.gpu

Counter = ah.Neuron(equations="r += 1", name="UntilCounter")


def _net(cond1, cond2):
    net = ah.Network(dt=1.0)
    p1 = net.create(15, Counter, stop_condition=cond1)
    p2 = net.create(15, Counter, stop_condition=cond2)
    p2.r = ah.DiscreteUniform(-4, -2, rng=np.random.default_rng(56789))
    net.compile()
    return net, p1, p2


def test_single_population(gpu):
    net, p1, p2 = _net(None, "r > 5.0 : any")
    assert net.simulate_until(max_duration=15.0, population=p2) == 8.0
    assert net.get_current_step() == 8


@pytest.mark.parametrize("op, steps", [("or", 6.0), ("and", 8.0)])
def test_two_conditions(gpu, op, steps):
    net, p1, p2 = _net("r > 5.0 : any", "r > 5.0 : any")
    assert net.simulate_until(max_duration=15.0, population=[p1, p2], operator=op) == steps


def test_all_neurons(gpu):
    # every neuron of p2 exceeds 5 once the ones started at -4 do: 10 steps
    net, p1, p2 = _net(None, "r > 5.0 : all")
    assert net.simulate_until(max_duration=15.0, population=p2) == 10.0


def test_max_duration(gpu):
    net = ah.Network(dt=1.0)
    q = net.create(15, Counter, stop_condition="r > 5.0 : any")
    net.compile()
    assert net.simulate_until(max_duration=3.0, population=q) == 3.0   # not true within 3 steps


def test_no_condition_runs_the_full_duration(gpu):
    net = ah.Network(dt=1.0)
    p = net.create(4, Counter)
    net.compile()
    assert net.simulate_until(max_duration=12.0, population=p) == 12.0
    np.testing.assert_array_equal(p.r, 12.0)


def test_spiked_any(gpu):
    net = ah.Network(dt=1.0)
    emit = ah.Neuron(spike="t==1", name="Emit")
    p = net.create(5, emit, stop_condition="spiked : any")
    net.compile()
    assert net.simulate_until(max_duration=10.0, population=p) == 2.0   # the spike of step 1
