"""Persistent steps of a small spiking network: one launch of one workgroup
per batch of steps (scatter into conductances held in LDS, the generated
neuron update, ascending spike list, spike recording).

* the 4000-neuron COBA network of the parity configuration: spike lists, v,
  g_exc and g_inh bit-identical between the persistent and the per-kernel
  step, and equal to the oracle, over 2000 steps (graph batches of 200 and
  an eager remainder);
* two projections with different weights into one conductance (scatter
  rounds in creation order): bit-exact against the oracle;
* FP32: the persistent and the per-kernel step are bit-identical;
* the delivery's forms (each spiking wave delivers its own spikes in steps
  of at most 64 spikes; above that the workgroup delivers the list: one lane
  group per item, or a scan of the column ranges above 64 items, in a
  strongly driven network) and column ranges read from global memory when
  they do not fit in LDS (4500 neurons);
* a spike monitor paused and resumed, and a network the path does not take
  (a Poisson input population) -- uses_persistent() says which.
"""

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import models
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _coba(n=4000, ne=3200, seed=7, precision="double", extra=False, persistent=True, graph_batch=200, I=20.0):
    net = ah.Network(dt=0.1, seed=seed, precision=precision)
    P = net.create(n, ah.COBA)
    projs = [net.connect(P[:ne], P, "exc").fixed_probability(0.02, 0.6),
             net.connect(P[ne:], P, "inh").fixed_probability(0.02, 6.7)]
    if extra:
        # a second projection into g_exc with another weight: its events
        # follow those of the first projection (creation order)
        projs.append(net.connect(P[ne:ne + 400], P, "exc").fixed_probability(0.05, 0.37))
    P.v = ah.Uniform(-60.0, -50.0, rng=net.rng)
    mon = net.monitor(P, ["spike"], capacity=3000)
    net.compile(graph_batch=graph_batch)
    P.I = I
    net.set_persistent(persistent)
    return net, P, projs, mon


def _state(P, mon):
    counts, idx = mon.raw_spikes()
    return counts, idx, P.v.reshape(-1).copy(), P.g_exc.reshape(-1).copy(), P.g_inh.reshape(-1).copy()


def _oracle(P, projs, v0, steps, I=20.0):
    po = [orc.csrc_projection(p.csr, 0 if p.target == "exc" else 1) for p in projs]
    return orc.spiking_run(0, P.size, 0.1, v0, steps, po, refractory=np.asarray(P.refractory, np.int32),
                           params={"I": I})


@pytest.mark.parametrize("extra", [False, True])
def test_persistent_equals_per_kernel_and_oracle(gpu, extra):
    steps = 2050   # 10 graph batches of 200 + 50 eager steps
    runs = {}
    for persistent in (True, False):
        net, P, projs, mon = _coba(extra=extra, persistent=persistent)
        assert net.uses_persistent() is persistent
        v0 = P.v.reshape(-1).copy()
        net.run(steps)
        runs[persistent] = _state(P, mon)
    a, b = runs[True], runs[False]
    assert a[0].sum() > 1000, "network is silent"
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    ref = _oracle(P, projs, v0, steps)
    assert np.array_equal(a[0], ref["counts"]) and np.array_equal(a[1], ref["idx"])
    assert np.array_equal(a[2], ref["v"]) and np.array_equal(a[3], ref["g_exc"]) and np.array_equal(a[4], ref["g_inh"])


@pytest.mark.parametrize("n,ne,I", [(4500, 3600, 20.0), (4000, 3200, 80.0)])
def test_persistent_scatter_forms(gpu, n, ne, I):
    steps = 600
    runs = {}
    for persistent in (True, False):
        net, P, projs, mon = _coba(n=n, ne=ne, I=I, persistent=persistent)
        assert net.uses_persistent() is persistent
        v0 = P.v.reshape(-1).copy()
        net.run(steps)
        runs[persistent] = _state(P, mon)
    a, b = runs[True], runs[False]
    if I > 20.0:
        # steps above 64 spikes: the workgroup delivers the list (the scan
        # form above 64 items); at most 64: the spiking waves deliver
        assert a[0].max() > 64 and a[0].min() <= 64
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    ref = _oracle(P, projs, v0, steps, I=I)
    assert np.array_equal(a[0], ref["counts"]) and np.array_equal(a[1], ref["idx"])
    assert np.array_equal(a[2], ref["v"])


def test_persistent_fp32_equals_per_kernel(gpu):
    runs = {}
    for persistent in (True, False):
        net, P, projs, mon = _coba(precision="float", persistent=persistent)
        assert net.uses_persistent() is persistent
        net.run(1000)
        runs[persistent] = _state(P, mon)
    for x, y in zip(runs[True], runs[False]):
        assert np.array_equal(x, y)


def test_persistent_monitor_pause_and_other_networks(gpu):
    net, P, projs, mon = _coba()
    v0 = P.v.reshape(-1).copy()
    net.run(300)
    mon.pause()
    assert net.uses_persistent()
    net.run(200)
    mon.resume()
    net.run(300)
    counts, idx = mon.raw_spikes()
    ref = _oracle(P, projs, v0, 800)
    keep = np.r_[np.arange(300), np.arange(500, 800)]
    starts = np.concatenate([[0], np.cumsum(ref["counts"])])
    ref_idx = np.concatenate([ref["idx"][starts[u]:starts[u + 1]] for u in keep])
    assert np.array_equal(counts, ref["counts"][keep])
    assert np.array_equal(idx, ref_idx)
    # a second population (Poisson input): the per-kernel step
    net2 = ah.Network(dt=0.1, seed=1)
    A = net2.create(200, ah.COBA)
    Q = net2.create(50, models.poisson_neuron(20.0))
    net2.connect(Q, A, "exc").fixed_probability(0.1, 0.5)
    net2.compile()
    assert not net2.uses_persistent()


def test_persistent_probe_continues_the_run(gpu):
    # the launch probe runs real persistent launches of a graph batch each:
    # the simulation continues exactly as an ordinary run would
    net, P, projs, mon = _coba(graph_batch=100)
    v0 = P.v.reshape(-1).copy()
    net.run(250)
    ms = net.time_kernel(7, 0, 3)
    assert ms > 0.0
    net.run(50)
    counts, idx = mon.raw_spikes()
    ref = _oracle(P, projs, v0, 600)
    assert np.array_equal(counts, ref["counts"])
    assert np.array_equal(idx, ref["idx"])
    net2 = ah.Network(dt=0.1, seed=1)
    A = net2.create(200, ah.COBA)
    Q = net2.create(50, models.poisson_neuron(20.0))
    net2.connect(Q, A, "exc").fixed_probability(0.1, 0.5)
    net2.compile()
    with pytest.raises(Exception, match="persistent"):
        net2.time_kernel(7, 0, 1)
