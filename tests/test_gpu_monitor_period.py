"""Monitor periods, start=False, pause() and resume() on the GPU: the known
answers of the reference's tests_python/Neuron/test_Monitor.py
(test_PopulationMonitor1D, a neuron with r = t): a 10 ms period over 100
steps records r = 0, 10, ..., 90; a monitor started after 10 steps, or
paused for the first 10, holds r = 10 ... 19 after 10 more."""

import numpy as np
import pytest

import annarchy_amd as ah

pytestmark = pytest.mark.gpu

RT = ah.Neuron(equations="r = t", name="RateT")


def _net():
    net = ah.Network(dt=1.0)
    pop = net.create(3, RT)
    return net, pop


def test_period_kat(gpu):
    net, pop = _net()
    m = net.monitor(pop, "r", capacity=20, period=10.0)
    net.compile()
    net.simulate(100)
    np.testing.assert_array_equal(m.get("r"), np.repeat(np.arange(0.0, 100.0, 10.0)[:, None], 3, axis=1))


def test_period_offset(gpu):
    net, pop = _net()
    m = net.monitor(pop, "r", capacity=20, period=10.0, period_offset=3.0)
    net.compile()
    net.simulate(35)
    np.testing.assert_array_equal(m.get("r")[:, 0], [3.0, 13.0, 23.0, 33.0])


def test_start_false_kat(gpu):
    net, pop = _net()
    m = net.monitor(pop, "r", capacity=30, start=False)
    net.compile()
    net.simulate(10)
    m.start()
    net.simulate(10)
    np.testing.assert_array_equal(m.get("r")[:, 0], np.arange(10.0, 20.0))


@pytest.mark.parametrize("graph", [False, True])
def test_pause_resume_kat(gpu, graph):
    net, pop = _net()
    m = net.monitor(pop, "r", capacity=30)
    net.compile()
    net.set_graph(graph, 4)
    m.pause()
    net.simulate(10)
    m.resume()
    net.simulate(10)
    np.testing.assert_array_equal(m.get("r")[:, 0], np.arange(10.0, 20.0))
    # rows continue after a pause in the middle, and after a clear
    net.simulate(3)
    m.pause()
    net.simulate(5)
    m.resume()
    net.simulate(2)
    np.testing.assert_array_equal(m.get("r")[:, 0], [20.0, 21.0, 22.0, 28.0, 29.0])


def test_period_with_pause(gpu):
    net, pop = _net()
    m = net.monitor(pop, "r", capacity=30, period=4.0)
    net.compile()
    net.simulate(9)     # rows at 0, 4, 8
    m.pause()
    net.simulate(6)     # 9 .. 14
    m.resume()
    net.simulate(6)     # 15 .. 20: rows at 16, 20
    np.testing.assert_array_equal(m.get("r")[:, 0], [0.0, 4.0, 8.0, 16.0, 20.0])


Every3 = ah.Neuron(equations="v = v + 1.0\nmp = t", spike="v == 3.0", reset="v = 0.0", name="Every3")


@pytest.mark.parametrize("graph", [False, True])
def test_spike_monitor_pause_and_start(gpu, graph):
    # spikes at 2, 5, 8, 11, 14, 17; the first monitor is paused over 6..11,
    # the second starts at 6; the paused one hands the fused recording over
    net = ah.Network(dt=1.0)
    pop = net.create(2, Every3)
    m1 = net.monitor(pop, ["spike"], capacity=30)
    m2 = net.monitor(pop, ["spike"], capacity=30, start=False)
    net.compile()
    net.set_graph(graph, 4)
    net.simulate(6)
    m1.pause()
    m2.start()
    net.simulate(6)
    m1.resume()
    net.simulate(6)
    assert m1.get("spike") == {0: [2, 5, 14, 17], 1: [2, 5, 14, 17]}
    assert m2.get("spike") == {0: [8, 11, 14, 17], 1: [8, 11, 14, 17]}
    # after a clear, recording goes on from the current step
    net.simulate(3)
    assert m1.get("spike") == {0: [20], 1: [20]}
