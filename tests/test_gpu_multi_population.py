"""Steps with many populations.

* Spike compaction: every population's update runs first, then one launch
  compacts the spike lists of all updated spiking populations (8 per launch,
  so 11 populations take two).  Eleven clock populations of different sizes
  (up to three 4096-neuron chunks each), each neuron firing when its counter
  reaches the population's period: every spike monitor equals the
  trajectory computed on the host, step by step, over 60 steps in graphs of
  16 and an eager remainder.
* psp sums completed by the rate update: a neuron with three summed targets,
  each fed by one sliced-tile projection.  Two are completed by the update
  from the panel partials (the most one update takes), the third by its
  reduction launch; r and all three sums are bit-identical to the same
  network in which monitors of the sums keep every reduction launch.
"""

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import connectivity as conn

pytestmark = pytest.mark.gpu

Clock = ah.Neuron(parameters="period = 5.0 : population", equations="v = v + 1.0",
                  spike="v >= period", reset="v = 0.0", name="Clock")


def test_compaction_of_many_populations(gpu):
    steps = 60
    sizes = [5, 300, 4096, 4097, 9000, 64, 1, 12288, 700, 2500, 8200]
    net = ah.Network(dt=1.0)
    pops, mons, v0s, periods = [], [], [], []
    rng = np.random.default_rng(4)
    for k, n in enumerate(sizes):
        P = net.create(n, Clock)
        pops.append(P)
        mons.append(net.monitor(P, ["spike"], capacity=steps))
        periods.append(float(3 + k % 5))
    net.compile(graph_batch=16)
    for P, per in zip(pops, periods):
        P.period = per
        v0 = rng.integers(0, int(per), P.size).astype(np.float64)
        P.v = v0
        v0s.append(v0)
    net.run(steps)
    for P, mon, v0, per in zip(pops, mons, v0s, periods):
        v = v0.copy()
        exp_counts, exp_idx = [], []
        for _ in range(steps):
            v += 1.0
            fired = np.flatnonzero(v >= per)
            v[fired] = 0.0
            exp_counts.append(fired.size)
            exp_idx.append(fired)
        counts, idx = mon.raw_spikes()
        np.testing.assert_array_equal(counts, exp_counts)
        np.testing.assert_array_equal(idx, np.concatenate(exp_idx))
        np.testing.assert_array_equal(np.asarray(P.v).reshape(-1), v)


Rate3 = ah.Neuron(parameters="tau = 10.0 : population", equations="""
    tau * dr/dt = -r + sum(exc) - sum(inh) + 0.5 * sum(mod) : min = 0.0
""", name="Rate3")


def test_three_panel_sums_one_update(gpu):
    n, steps = 2500, 30
    rng = np.random.default_rng(9)
    csrs = [conn.random_in_degree(n, n, rng, np.full(n, 80, np.int64),
                                  conn.Normal(0.0, 0.1, rng=rng)) for _ in range(3)]
    r0 = rng.random(n)
    out = []
    for watch in (False, True):
        net = ah.Network(dt=1.0, seed=2)
        P = net.create(n, Rate3)
        for target, csr in zip(("exc", "inh", "mod"), csrs):
            net.connect(P, P, target).from_csr(csr, storage_format="sell_panel", lanes=2, sigma=8,
                                               panel_cols=900, chunk=512)
        mon = net.monitor(P, ["_sum_exc", "_sum_inh", "_sum_mod"] if watch else ["r"], capacity=steps)
        net.compile(graph_batch=8)
        P.r = r0
        net.run(steps)
        out.append([np.asarray(P.get(a)).reshape(-1).copy() for a in ("r", "_sum_exc", "_sum_inh", "_sum_mod")])
    assert np.abs(out[0][0]).max() > 0.0
    for x, y, name in zip(out[0], out[1], ("r", "_sum_exc", "_sum_inh", "_sum_mod")):
        assert np.array_equal(x, y), name
