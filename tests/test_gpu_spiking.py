"""Spiking path on the GPU: known answers of the reference tests and
bit-exact agreement with the CPU oracle.

* tests_python/Neuron/test_SpikingNeuron.py -- v trajectories with and
  without refractory period;
* tests_python/Neuron/test_Monitor.py:294-333 -- spike times [4, 6, 8] and,
  with refractory=3, [2, 7];
* tests_python/SpikeSynapse/test_SpikingTransmission.py:14-78 -- 5 emitters
  firing at t=1, all-to-all w=1 -> g_exc == 5 at step 2 only;
* the COBA network (4000 neurons, p=0.02, constant weights) against the
  oracle over 3000 steps: spike lists, v and g_exc bit for bit (FMA
  contraction off on both sides; one weight per projection makes the
  atomic scatter order-independent);
* Poisson drive with injected uniforms, bit for bit;
* compaction: ascending, complete, no duplicates.
"""

import ctypes as C

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import _lib, models
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_counter_neuron_trajectory(gpu):
    net = ah.Network(dt=1.0)
    p1 = net.create(3, models.CounterNeuron)
    p2 = net.create(3, models.CounterNeuronRefractory)
    net.compile()
    v1, v2 = [], []
    for _ in range(7):
        v1.append(float(p1.v[0]))
        v2.append(float(p2.v[0]))
        net.simulate(1)
    v1.append(float(p1.v[0]))
    v2.append(float(p2.v[0]))
    assert v1[:6] == [0.0, 1.0, 2.0, 1.0, 2.0, 1.0]
    assert v2 == [0.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 2.0]


def test_monitor_spike_times(gpu):
    n2 = ah.Neuron(equations="v = v + 1.0", spike="v == 5.0", reset="v = 3.0", name="Mon2")
    n3 = ah.Neuron(equations="v = v + 1.0", spike="v == 3.0", reset="v = 1.0 ", refractory=3.0,
                   name="Mon3")
    net = ah.Network(dt=1.0)
    pop3 = net.create(3, n2)
    pop4 = net.create(3, n3)
    m3 = net.monitor(pop3, ["v", "spike"], capacity=20)
    m4 = net.monitor(pop4, ["v", "spike"], capacity=20)
    net.compile()
    net.simulate(10)
    assert m3.get("spike")[0] == [4, 6, 8]
    v4 = m4.get("v")
    expected = [1, 2, 1, 1, 1, 1, 2, 1, 1, 1]
    np.testing.assert_array_equal(v4[:, 0], expected)
    assert m4.get("spike")[1] == [2, 7]


def test_spike_transmission_known_answer(gpu):
    emit = ah.Neuron(spike="t==1", name="Emit")
    recv = ah.Neuron(equations="g_exc = 0", spike="g_exc>30", name="Recv")
    net = ah.Network(dt=1.0)
    inp = net.create(5, emit)
    out = net.create(2, recv)
    net.connect(inp, out, "exc").all_to_all(1.0)
    mon = net.monitor(out, ["g_exc"], capacity=10)
    net.compile()
    net.simulate(5)
    g = mon.get("g_exc")
    np.testing.assert_array_equal(g, [[0, 0], [0, 0], [5, 5], [0, 0], [0, 0]])


def _coba_net(N=4000, seed=7, p=0.02, graph_batch=0, steps_cap=1000, precision="double",
              record=("v", "g_exc", "spike")):
    net = ah.Network(dt=0.1, seed=seed, precision=precision)
    P = net.create(N, ah.COBA)
    ne = int(0.8 * N)
    Pe, Pi = P[:ne], P[ne:]
    ce = net.connect(Pe, P, "exc").fixed_probability(p, 0.6)
    ci = net.connect(Pi, P, "inh").fixed_probability(p, 6.7)
    P.v = ah.Uniform(-60.0, -50.0, rng=net.rng)
    mon = net.monitor(P, list(record), capacity=steps_cap)
    net.compile(graph_batch=graph_batch)
    return net, P, ce, ci, mon


def _oracle_coba(P, ce, ci, v0, steps, trace=True):
    projs = [orc.csrc_projection(ce.csr, 0), orc.csrc_projection(ci.csr, 1)]
    ref = np.asarray(P.refractory, dtype=np.int32)
    return orc.spiking_run(0, P.size, 0.1, v0, steps, projs, refractory=ref, trace=trace)


def test_coba_bit_exact_vs_oracle(gpu):
    steps = 3000  # 300 ms
    net, P, ce, ci, mon = _coba_net(graph_batch=50, steps_cap=steps)
    # regression: population-wide values written at set-up (zeroed allocations
    # followed by the fill of the initial values) must all have landed
    assert P.Vr == -60.0 and P.El == -60.0 and P.Vt == -50.0
    P.I = 20.0
    v0 = P.v.reshape(-1).copy()
    assert np.all(P.refractory == 50)
    net.run(steps)
    ref = _oracle_coba(P, ce, ci, v0, steps)
    counts, idx = mon.raw_spikes()
    assert counts.sum() > 100, "network is silent"
    np.testing.assert_array_equal(counts, ref["counts"])
    np.testing.assert_array_equal(idx, ref["idx"])
    v = mon.get("v", keep=True)
    ge = mon.get("g_exc", keep=True)
    assert np.array_equal(v, ref["trace_v"])
    assert np.array_equal(ge, ref["trace_gexc"])
    assert np.array_equal(P.g_inh.reshape(-1), ref["g_inh"])
    last = np.empty(P.size, dtype=np.int64)
    _lib.check(_lib.lib().ahip_pop_get_last_spike(net._h, P.id, _lib.ptr(last, C.c_int64), P.size))
    np.testing.assert_array_equal(last, ref["last_spike"])


def test_compaction_ascending_complete(gpu):
    net, P, ce, ci, mon = _coba_net(N=9000, seed=3, graph_batch=10, steps_cap=300)
    net.run(300)
    counts, idx = mon.raw_spikes()
    pos = 0
    for c in counts:
        seg = idx[pos:pos + c]
        assert np.all(np.diff(seg) > 0) and (c == 0 or (seg[0] >= 0 and seg[-1] < 9000))
        pos += c
    assert counts.sum() > 0


def test_poisson_drive_bit_exact(gpu):
    N, M, steps = 3000, 300, 400
    net = ah.Network(dt=0.1, seed=11)
    P = net.create(N, ah.COBA)
    Q = net.create(M, models.poisson_neuron(50.0))
    ce = net.connect(P[:2400], P, "exc").fixed_probability(0.02, 0.6)
    ci = net.connect(P[2400:], P, "inh").fixed_probability(0.02, 6.7)
    cq = net.connect(Q, P[:M], "exc").one_to_one(2.0)
    P.v = ah.Uniform(-60.0, -50.0, rng=net.rng)
    mon = net.monitor(P, ["spike"], capacity=steps)
    monq = net.monitor(Q, ["spike"], capacity=steps)
    net.compile(graph_batch=20)
    P.I = 0.0
    v0 = P.v.reshape(-1).copy()
    unif = np.random.default_rng(5).random((steps, M))
    Q.inject_random(0, unif)
    net.run(steps)
    projs = [orc.csrc_projection(ce.csr, 0), orc.csrc_projection(ci.csr, 1),
             orc.csrc_projection(cq.csr, 0, pre_is_input=True)]
    ref = orc.spiking_run(0, N, 0.1, v0, steps, projs, refractory=np.asarray(P.refractory),
                          params=dict(I=0.0), poisson=dict(M=M, rates=50.0, uniforms=unif))
    qc, _ = monq.raw_spikes()
    np.testing.assert_array_equal(qc, ref["input_counts"])
    counts, idx = mon.raw_spikes()
    np.testing.assert_array_equal(counts, ref["counts"])
    np.testing.assert_array_equal(idx, ref["idx"])
    assert np.array_equal(P.v.reshape(-1), ref["v"])


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


def test_coba_fp32_rate_and_isi_cv_vs_fp64(gpu):
    # A long chaotic run cannot agree spike for spike in FP32.  The north-star
    # criterion is the population rate and the ISI coefficient of variation
    # within 2% (core/Monitor.py:1172-1240 define both statistics).  Over 5 s
    # one trajectory's rate varies by about 1.5% with its initial state alone
    # (FP64 oracle runs from independent v0 draws), so the 2% bar is applied
    # to the means over six initial states, FP32 on the GPU against the FP64
    # oracle from the same states.
    steps = 50000  # 5 s
    r32, cv32, r64, cv64 = [], [], [], []
    for k in range(6):
        net, P, ce, ci, mon = _coba_net(graph_batch=100, steps_cap=steps, precision="float",
                                        record=("spike",))
        P.I = 20.0
        v0 = np.random.default_rng(900 + k).uniform(-60.0, -50.0, P.size)
        P.v = v0
        net.run(steps)
        counts, idx = mon.raw_spikes()
        ref = _oracle_coba(P, ce, ci, v0, steps, trace=False)
        net.destroy()
        a, b = _spike_stats(counts, idx, P.size, 0.1)
        c, d = _spike_stats(ref["counts"], ref["idx"], P.size, 0.1)
        r32.append(a), cv32.append(b), r64.append(c), cv64.append(d)
    print(f"rate FP32 {np.round(r32, 3)} / FP64 {np.round(r64, 3)} Hz; "
          f"ISI CV {np.round(cv32, 4)} / {np.round(cv64, 4)}")
    assert min(r64) > 5.0
    assert abs(np.mean(r32) / np.mean(r64) - 1.0) < 0.02
    assert abs(np.mean(cv32) / np.mean(cv64) - 1.0) < 0.02
