"""compute_firing_rate(window) on the GPU (core/Population.py:700-723,
Population/OpenMPGenerator.py:562-618): r = 1000/window x the spikes of the
last window/dt steps, updated after every spike gather.

* known answer: a counter neuron firing every 2 steps (spikes at 2, 4, 6, ...)
  with a 10 ms window at dt = 1 ms holds r = 100 x (spikes in (t-10, t]);
* a Poisson population: r recorded every step equals, exactly, 1000/window
  times the sliding-window count of the spikes recorded in the same run;
* r stays 0 without compute_firing_rate, and reset() empties the history.
"""

import numpy as np
import pytest

import annarchy_amd as ah

pytestmark = pytest.mark.gpu


def test_counter_neuron_rate_kat(gpu):
    net = ah.Network(dt=1.0)
    pop = net.create(1, ah.models.CounterNeuron)
    pop.compute_firing_rate(10.0)
    mon = net.monitor(pop, ["r"], capacity=30)
    net.compile()
    net.run(30)
    r = mon.get("r")[:, 0]
    spikes = np.arange(2, 30, 2)
    expected = [100.0 * np.count_nonzero((spikes > t - 10) & (spikes <= t)) for t in range(30)]
    np.testing.assert_array_equal(r, expected)
    assert r[-1] == 500.0


def test_poisson_rate_equals_window_count(gpu):
    steps, n, window, dt = 1500, 1000, 50.0, 0.1
    W = int(window / dt)
    net = ah.Network(dt=dt, seed=4)
    pop = net.create(n, ah.poisson_neuron(50.0))
    mon = net.monitor(pop, ["spike", "r"], capacity=steps)
    net.compile(graph_batch=50)
    pop.compute_firing_rate(window)
    net.run(steps)
    counts, idx = mon.raw_spikes()
    rec = np.zeros((steps, n), dtype=np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)])
    for t in range(steps):
        rec[t, idx[starts[t]:starts[t + 1]]] = 1
    csum = np.cumsum(rec, axis=0)
    win = csum - np.vstack([np.zeros((W, n), np.int64), csum[:-W]])
    r = mon.get("r")
    np.testing.assert_array_equal(r, (1000.0 / window) * win)
    assert abs(r[-1].mean() - 50.0) < 5.0


def test_rate_off_and_reset(gpu):
    net = ah.Network(dt=1.0)
    a = net.create(1, ah.models.CounterNeuron)
    b = net.create(1, ah.models.CounterNeuron)
    net.compile()
    b.compute_firing_rate(10.0)
    net.run(20)
    assert float(a.r[0]) == 0.0 and float(b.r[0]) == 500.0
    net.reset()
    net.run(3)
    assert float(b.r[0]) == 100.0          # only the spike of step 2 since the reset
