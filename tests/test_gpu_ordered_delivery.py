"""Spike delivery of projections with per-synapse weights.

The reference adds the events of a step projection by projection, spike by
spike in ascending pre-synaptic rank, down each CSRC column
(generator/Projection/OpenMP/CSR.py:870-885).  With distinct weights the
rounding of g_target depends on that order.  The default ordered delivery
appends each event's key to its post neuron's bucket and adds every bucket in
ascending key, so g, v and the spike lists equal the oracle's bit for bit:

* the shape of the 400k-neuron benchmark network at a tenth of its size
  (40,000 COBA neurons, fixed_number_pre 800 E + 200 I drawn on the device
  with per-synapse Normal / Uniform weights, Poisson 5 Hz into 10% with
  injected draws), 1000 steps (100 ms): spike lists, v, g_exc and g_inh
  bit-exact;
* buckets longer than the in-register sort (64 events per post neuron in one
  step) with weights whose sum depends on the order, and two projections
  into one conductance: g_exc bit-exact, and different from the sum in
  descending order;
* the atomic mode: g within 1e-12 relative of the oracle over 100 steps (the
  stated error of arbitrary-order device atomics);
* FP32: population rate and ISI CV within 2% of the FP64 oracle.
"""

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import connectivity as cn
from annarchy_amd import models
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

N, NE, M = 40_000, 32_000, 4_000
SCALE = 80.0 / 1000.0
WE, WI = 0.6 * SCALE, 6.7 * SCALE


def _c2_shape(steps, seed=21, precision="double", delivery="ordered", record=("spike",), graph_batch=50):
    net = ah.Network(dt=0.1, seed=seed, precision=precision, spike_delivery=delivery)
    P = net.create(N, ah.COBA)
    Q = net.create(M, models.poisson_neuron(5.0))
    ce = net.connect(P[:NE], P, "exc").fixed_number_pre(800, cn.Normal(WE, WE / 4), generator="device")
    ci = net.connect(P[NE:], P, "inh").fixed_number_pre(200, cn.Uniform(0.5 * WI, 1.5 * WI), generator="device")
    cq = net.connect(Q, P[:M], "exc").one_to_one(0.6)
    P.v = ah.Uniform(-60.0, -50.0, rng=net.rng)
    mon = net.monitor(P, list(record), capacity=steps)
    net.compile(graph_batch=graph_batch)
    P.I = 20.0
    unif = np.random.default_rng(seed + 1).random((steps, M))
    Q.inject_random(0, unif)
    return net, P, (ce, ci, cq), mon, unif


def _oracle(P, projs, v0, steps, unif, trace=False):
    ce, ci, cq = projs
    po = [orc.csrc_projection(ce.csr, 0), orc.csrc_projection(ci.csr, 1),
          orc.csrc_projection(cq.csr, 0, pre_is_input=True)]
    return orc.spiking_run(0, N, 0.1, v0, steps, po, refractory=np.asarray(P.refractory, np.int32),
                           poisson=dict(M=M, rates=5.0, uniforms=unif), trace=trace, spike_cap=N * 20)


def test_c2_shape_per_synapse_weights_bit_exact(gpu):
    steps = 1000
    net, P, projs, mon, unif = _c2_shape(steps)
    ce, ci, _ = projs
    # distinct weights: the order of the additions matters
    assert np.unique(ce.csr.w[:10000]).size > 9000 and np.unique(ci.csr.w[:10000]).size > 9000
    assert ce.csr.nnz == N * 800 and ci.csr.nnz == N * 200
    assert net.ordered_targets() == 2
    v0 = P.v.reshape(-1).copy()
    net.run(steps)
    ref = _oracle(P, projs, v0, steps, unif)
    counts, idx = mon.raw_spikes()
    print(f"{int(counts.sum())} spikes in {steps} steps, max {int(counts.max())} per step")
    assert counts.sum() > 2000, "network is silent"
    np.testing.assert_array_equal(counts, ref["counts"])
    np.testing.assert_array_equal(idx, ref["idx"])
    assert np.array_equal(P.v.reshape(-1), ref["v"])
    assert np.array_equal(P.g_exc.reshape(-1), ref["g_exc"])
    assert np.array_equal(P.g_inh.reshape(-1), ref["g_inh"])


def test_long_buckets_and_projection_order(gpu):
    # 64 Poisson inputs fire together at step 0 (injected draws), all-to-all
    # into 5 neurons with weights spread over six decades: every bucket holds
    # 64 + 40 events in step 1 (beyond the 16 sorted in registers).  A second
    # projection (first 40 inputs, one weight 1e-3) feeds the same g_exc.
    Mi, Np, steps = 64, 5, 4
    rng = np.random.default_rng(3)
    net = ah.Network(dt=0.1, seed=2)
    P = net.create(Np, ah.COBA)
    Q = net.create(Mi, models.poisson_neuron(5.0))
    w = 10.0 ** rng.uniform(-4, 2, Mi * Np)
    csr1 = cn.from_arrays(Np, Mi, np.arange(Np + 1, dtype=np.int64) * Mi, np.tile(np.arange(Mi, dtype=np.int32), Np), w)
    csr2 = cn.from_arrays(Np, Mi, np.arange(Np + 1, dtype=np.int64) * 40, np.tile(np.arange(40, dtype=np.int32), Np),
                          None, w_const=1e-3)
    p1 = net.connect(Q, P, "exc").from_csr(csr1)
    p2 = net.connect(Q, P, "exc").from_csr(csr2)
    mon = net.monitor(P, ["g_exc"], capacity=steps)
    net.compile(graph_batch=2)
    P.Vt = 1e9
    P.v = -60.0
    unif = np.ones((steps, Mi))
    unif[0, :] = 0.0
    Q.inject_random(0, unif)
    net.run(steps)
    po = [orc.csrc_projection(p1.csr, 0, pre_is_input=True), orc.csrc_projection(p2.csr, 0, pre_is_input=True)]
    ref = orc.spiking_run(0, Np, 0.1, np.full(Np, -60.0), steps, po, params=dict(Vt=1e9),
                          poisson=dict(M=Mi, rates=5.0, uniforms=unif), trace=True)
    ge = mon.get("g_exc")
    assert np.array_equal(ge, ref["trace_gexc"])
    # the order is what makes it exact: the sum in descending pre-synaptic
    # order rounds differently for at least one neuron
    fwd = np.zeros(Np)
    rev = np.zeros(Np)
    W = w.reshape(Np, Mi)
    for i in range(Np):
        for j in range(Mi):
            fwd[i] += W[i, j]
        for j in range(40):
            fwd[i] += 1e-3
        for j in range(40):
            rev[i] += 1e-3
        for j in reversed(range(Mi)):
            rev[i] += W[i, j]
    assert np.array_equal(ge[1], fwd)
    assert not np.array_equal(fwd, rev)


def test_atomic_delivery_error_bound(gpu):
    # arbitrary-order device atomics: each addition rounds once, so g differs
    # from the ordered sum by a few ulp of the bucket's sum; over 100 steps
    # g_exc stays within 1e-12 relative of the oracle
    steps = 100
    net, P, projs, mon, unif = _c2_shape(steps, delivery="atomic", record=("g_exc",))
    assert net.ordered_targets() == 0
    v0 = P.v.reshape(-1).copy()
    net.run(steps)
    ref = _oracle(P, projs, v0, steps, unif, trace=True)
    ge = mon.get("g_exc")
    scale = np.maximum(np.abs(ref["trace_gexc"]), 1e-3)
    err = float(np.max(np.abs(ge - ref["trace_gexc"]) / scale))
    print(f"atomic delivery: max relative g_exc error over {steps} steps {err:.2e}")
    assert err < 1e-12


def _spike_stats(counts, idx, n, dt_ms):
    t = np.repeat(np.arange(len(counts)), counts)
    rate = len(idx) / n / (len(counts) * dt_ms * 1e-3)
    order = np.lexsort((t, idx))
    ids, ts = idx[order], t[order]
    cvs = []
    starts = np.flatnonzero(np.r_[True, ids[1:] != ids[:-1]])
    for a, b in zip(starts, np.r_[starts[1:], len(ids)]):
        if b - a > 2:
            isi = np.diff(ts[a:b]) * dt_ms
            cvs.append(np.std(isi) / np.mean(isi))
    return rate, float(np.mean(cvs))


def test_c2_shape_fp32_rate_and_isi_cv(gpu):
    # FP32 state and weights on the GPU against the FP64 oracle from the same
    # initial state and the same Poisson draws, 2 s: population rate and mean
    # ISI CV within 2%
    steps = 20000
    net, P, projs, mon, unif = _c2_shape(steps, precision="float", graph_batch=100)
    v0 = P.v.reshape(-1).copy()
    net.run(steps)
    counts, idx = mon.raw_spikes()
    ref = _oracle(P, projs, v0, steps, unif)
    r32, cv32 = _spike_stats(counts, idx, N, 0.1)
    r64, cv64 = _spike_stats(ref["counts"], ref["idx"], N, 0.1)
    print(f"FP32 rate {r32:.4f} Hz / FP64 {r64:.4f} Hz; ISI CV {cv32:.4f} / {cv64:.4f}")
    assert r64 > 0.5
    assert abs(r32 / r64 - 1.0) < 0.02
    assert abs(cv32 / cv64 - 1.0) < 0.02


def test_prepared_graph_survives_eager_steps(gpu):
    # Regression: a graph captured by prepare_graph() must still address live
    # scatter tables after an eager run of fewer steps than the graph batch
    # (the sequence of the benchmark: prepare, warm up eagerly, replay).
    # 4000 COBA neurons with per-synapse Normal weights (ordered delivery),
    # graph of 20 steps: 7 eager steps, then 60 replayed, then 5 eager;
    # spike lists, v and g bit-exact against the oracle.
    steps_a, steps_b, steps_c, G = 7, 60, 5, 20
    steps = steps_a + steps_b + steps_c
    net = ah.Network(dt=0.1, seed=5)
    P = net.create(4000, ah.COBA)
    ce = net.connect(P[:3200], P, "exc").fixed_probability(0.02, cn.Normal(0.6, 0.15))
    ci = net.connect(P[3200:], P, "inh").fixed_probability(0.02, cn.Uniform(3.0, 9.0))
    P.v = ah.Uniform(-60.0, -50.0, rng=net.rng)
    mon = net.monitor(P, ["spike"], capacity=steps)
    net.compile(graph_batch=G)
    P.I = 20.0
    assert net.ordered_targets() == 2
    v0 = P.v.reshape(-1).copy()
    net.prepare_graph()
    net.run(steps_a)
    net.run(steps_b)
    net.run(steps_c)
    po = [orc.csrc_projection(ce.csr, 0), orc.csrc_projection(ci.csr, 1)]
    ref = orc.spiking_run(0, 4000, 0.1, v0, steps, po, refractory=np.asarray(P.refractory, np.int32))
    counts, idx = mon.raw_spikes()
    assert counts.sum() > 100, "network is silent"
    np.testing.assert_array_equal(counts, ref["counts"])
    np.testing.assert_array_equal(idx, ref["idx"])
    assert np.array_equal(P.v.reshape(-1), ref["v"])
    assert np.array_equal(P.g_exc.reshape(-1), ref["g_exc"])


@pytest.mark.parametrize("n_in", [3, 4, 5, 64])
def test_buckets_delivered_by_the_update(gpu, n_in):
    # no monitor reads g_exc after compute_psp, so the post population's
    # update adds the buckets: n_in inputs fire together at step 0 into 5
    # neurons (buckets of n_in + min(n_in, 40) events in step 1, in-register
    # sort up to 4, repeated minimum beyond); v, g_exc bit-exact after 3 steps
    Np, steps = 5, 3
    rng = np.random.default_rng(7 + n_in)
    net = ah.Network(dt=0.1, seed=2)
    P = net.create(Np, ah.COBA)
    Q = net.create(n_in, models.poisson_neuron(5.0))
    w = 10.0 ** rng.uniform(-4, 2, n_in * Np)
    k2 = min(n_in, 40)
    csr1 = cn.from_arrays(Np, n_in, np.arange(Np + 1, dtype=np.int64) * n_in,
                          np.tile(np.arange(n_in, dtype=np.int32), Np), w)
    csr2 = cn.from_arrays(Np, n_in, np.arange(Np + 1, dtype=np.int64) * k2,
                          np.tile(np.arange(k2, dtype=np.int32), Np), None, w_const=1e-3)
    p1 = net.connect(Q, P, "exc").from_csr(csr1)
    p2 = net.connect(Q, P, "exc").from_csr(csr2)
    net.compile(graph_batch=2)
    assert net.ordered_targets() == 1
    P.Vt = 1e9
    P.v = -60.0
    unif = np.ones((steps, n_in))
    unif[0, :] = 0.0
    Q.inject_random(0, unif)
    net.run(steps)
    po = [orc.csrc_projection(p1.csr, 0, pre_is_input=True), orc.csrc_projection(p2.csr, 0, pre_is_input=True)]
    ref = orc.spiking_run(0, Np, 0.1, np.full(Np, -60.0), steps, po, params=dict(Vt=1e9),
                          poisson=dict(M=n_in, rates=5.0, uniforms=unif))
    assert np.array_equal(P.g_exc.reshape(-1), ref["g_exc"])
    assert np.array_equal(P.v.reshape(-1), ref["v"])
