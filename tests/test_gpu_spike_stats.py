"""Spike-train statistics of a device-recorded Poisson population: 1000
neurons at 50 Hz for 500 ms hold a mean rate within 3% of 50 Hz and
inter-spike intervals whose pooled coefficient of variation is that of
the discretised Poisson process (geometric intervals, CV = sqrt(1 - p) with
p = rate x dt), with consistent raster and histogram counts."""

import numpy as np
import pytest

import annarchy_amd as ah

pytestmark = pytest.mark.gpu


def test_poisson_statistics(gpu):
    net = ah.Network(dt=0.1, seed=12)
    pop = net.create(1000, ah.poisson_neuron(50.0))
    m = net.monitor(pop, ["spike"], capacity=5000)
    net.compile()
    net.simulate(500.0)
    spikes = m.get("spike", keep=True)
    n = sum(len(v) for v in spikes.values())
    rate = n / 1000 / 0.5
    assert abs(rate - 50.0) < 1.5
    assert abs(ah.Monitor.mean_fr(spikes, 500.0) - rate) < 1e-9
    isi = np.array(m.inter_spike_interval(spikes))
    assert abs(isi.mean() - 20.0) < 1.0             # 1 / 50 Hz in ms
    # pooled intervals: geometric with p = 50 Hz x 0.1 ms, CV = sqrt(1 - p)
    assert abs(isi.std() / isi.mean() - np.sqrt(1.0 - 0.005)) < 0.03
    # per neuron (~24 intervals each) the estimate is biased low
    cv = np.mean(m.coefficient_of_variation(spikes))
    assert 0.85 < cv < 1.02
    t, r = m.raster_plot(spikes)
    assert t.size == n and r.size == n and t.max() < 500.0
    h = m.histogram(spikes, bins=10.0)
    assert h.sum() == n
