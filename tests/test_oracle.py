"""The CPU oracle pinned against the reference's golden vectors, known answers
and closed forms (no GPU).

Containers: tests_cpp golden layouts (tests/golden/container_layouts.json).
Rate psp: the dense product W.r (tests_python/RateSynapse/test_RateDelays.py).
Rate update: closed-form geometric decay / fixed point of the explicit Euler
recurrence.  Spiking: test_SpikingNeuron / test_Monitor trajectories,
test_SpikingTransmission's 5 x w=1, the free-membrane solution of the LIF and
the conductance decay.  MLP: finite differences.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from annarchy_amd import connectivity as conn
from oracle import oracle as orc

GOLD = json.loads((Path(__file__).parent / "golden" / "container_layouts.json").read_text())


@pytest.mark.parametrize("m", ["A", "B"])
def test_csr_golden(m):
    g = GOLD[m]
    rb, ci = orc.csr_from_lil(g["num_rows"], g["post_ranks"], g["pre_ranks"])
    flat_vals = np.concatenate([np.asarray(v, float) for v in g["values"]])
    assert np.array_equal(flat_vals, g["csr"])
    # rows are indexed by post rank: empty rows have empty ranges
    lens = np.diff(rb)
    for k, r in enumerate(g["post_ranks"]):
        assert lens[r] == len(g["pre_ranks"][k])
    assert lens.sum() == len(g["csr"])
    assert np.array_equal(ci, np.concatenate(g["pre_ranks"]))


@pytest.mark.parametrize("m", ["A", "B"])
def test_dense_golden(m):
    g = GOLD[m]
    d = orc.dense_from_lil(g["num_rows"], g["num_cols"], g["post_ranks"], g["pre_ranks"], g["values"], True)
    assert np.array_equal(d, g["dense_row_major"])
    if "dense_col_major" in g:
        d = orc.dense_from_lil(g["num_rows"], g["num_cols"], g["post_ranks"], g["pre_ranks"], g["values"], False)
        assert np.array_equal(d, g["dense_col_major"])


@pytest.mark.parametrize("m", ["A", "B"])
def test_ell_golden(m):
    g = GOLD[m]
    maxnzr, col, val, rl = orc.ell_from_lil(g["pre_ranks"], g["values"], row_major=True)
    assert maxnzr == 4
    assert np.array_equal(val, g["ell_row_major"])
    assert np.all(col[val == 0] == np.iinfo(np.int32).max)
    assert list(rl) == [len(x) for x in g["pre_ranks"]]
    _, col, val, _ = orc.ell_from_lil(g["pre_ranks"], g["values"], row_major=False)
    assert np.array_equal(val, g["ell_col_major"])


@pytest.mark.parametrize("m", ["A", "B"])
def test_sell_golden(m):
    g = GOLD[m]
    rp, col, val, mask = orc.sell_from_lil(g["num_rows"], g["post_ranks"], g["pre_ranks"], g["values"], 3, True)
    assert np.array_equal(val, g["sell3_row_major"])
    assert list(rp) == [0, 12, 18]
    assert np.all(col[mask == 0] == 0)


@pytest.mark.parametrize("m", ["A", "B"])
def test_bsr_golden(m):
    g = GOLD[m]
    brp, bci, val = orc.bsr_from_lil(g["num_rows"], g["num_cols"], 2, g["post_ranks"], g["pre_ranks"], g["values"])
    assert np.array_equal(val, g["bsr2"])
    assert brp[-1] == len(bci)


def test_csrc_inverse_brute_force():
    rng = np.random.default_rng(0)
    for trial in range(5):
        nr, nc = rng.integers(1, 60, size=2)
        dense = (rng.random((nr, nc)) < 0.2)
        post = [r for r in range(nr) if dense[r].any()]
        pre = [np.nonzero(dense[r])[0] for r in post]
        if not post:
            continue
        rb, ci = orc.csr_from_lil(int(nr), post, pre)
        col_ptr, row_idx, inv = orc.csrc_inverse(int(nr), int(nc), rb, ci)
        for c in range(nc):
            rows = row_idx[col_ptr[c]:col_ptr[c + 1]]
            assert list(rows) == list(np.nonzero(dense[:, c])[0])  # ascending post ranks
            for q in range(col_ptr[c], col_ptr[c + 1]):
                assert ci[inv[q]] == c and rb[row_idx[q]] <= inv[q] < rb[row_idx[q] + 1]


def test_csrc_inverse_fast_equals_restatement():
    """The counting-sort view used for large networks returns the arrays of
    the map-based restatement (empty rows and columns included)."""
    rng = np.random.default_rng(4)
    for trial in range(6):
        nr, nc = int(rng.integers(1, 300)), int(rng.integers(1, 300))
        csr = conn.random_in_degree(nc, nr, rng, rng.integers(0, min(nc, 40) + 1, nr), 1.0)
        a = orc.csrc_inverse(nr, nc, csr.row_ptr, csr.col_idx, fast=False)
        b = orc.csrc_inverse(nr, nc, csr.row_ptr, csr.col_idx, fast=True)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def test_spmv_dense_product():
    rng = np.random.default_rng(1)
    csr = conn.random_in_degree(500, 300, rng, rng.integers(0, 40, 300), conn.Normal(0, 1, rng=rng))
    r = rng.random(500)
    got = orc.spmv_csr(csr.row_ptr, csr.col_idx, csr.w, r, nthreads=4)
    np.testing.assert_allclose(got, csr.dense() @ r, rtol=1e-12, atol=1e-12)
    # += semantics: target accumulates
    got2 = orc.spmv_csr(csr.row_ptr, csr.col_idx, csr.w, r, target=np.ones(300))
    np.testing.assert_allclose(got2, csr.dense() @ r + 1.0, rtol=1e-12)


def test_rate_run_closed_forms():
    n, steps, tau, dt = 50, 40, 10.0, 1.0
    empty = conn.from_arrays(n, n, np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0))
    r0 = np.linspace(0.1, 2.0, n)
    r, _, tr = orc.rate_run(empty.row_ptr, empty.col_idx, np.zeros(1), r0, tau, dt, None, steps, trace=True)
    k = np.arange(1, steps + 1)[:, None]
    np.testing.assert_allclose(tr, r0[None, :] * (1 - dt / tau) ** k, rtol=1e-12)
    # self-excitation w=0.5 of each neuron on itself: fixed point of r = r + dt/tau(-r + 0.5 r) is 0,
    # with constant noise input c: r* = 2c
    eye = conn.from_arrays(n, n, np.arange(n + 1, dtype=np.int64), np.arange(n, dtype=np.int32), np.full(n, 0.5))
    c = 0.3
    noise = np.full((2000, n), c)
    r, _, _ = orc.rate_run(eye.row_ptr, eye.col_idx, eye.w, np.zeros(n), tau, dt, noise, 2000)
    np.testing.assert_allclose(r, 2 * c, rtol=1e-12)
    # bound r >= 0 (min=0.0): a negative drive clips at exactly zero
    r, _, _ = orc.rate_run(empty.row_ptr, empty.col_idx, np.zeros(1), r0, tau, dt, np.full((5, n), -100.0), 5)
    assert np.all(r == 0.0)


def test_counter_neuron_known_answers():
    # test_SpikingNeuron.py test_v: spike at v == 3, reset 1
    out = orc.spiking_run(1, 3, 1.0, np.zeros(3), 6, params=dict(Vt=3.0, Vr=1.0), trace=True)
    assert list(out["trace_v"][:, 0]) == [1.0, 2.0, 1.0, 2.0, 1.0, 2.0]
    # test_v_ref: refractory 3 steps
    out = orc.spiking_run(1, 3, 1.0, np.zeros(3), 7, params=dict(Vt=3.0, Vr=1.0),
                          refractory=np.full(3, 3), trace=True)
    assert list(out["trace_v"][:, 0]) == [1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 2.0]
    # test_Monitor.py:294-333: spikes at [4, 6, 8] (v == 5, reset 3) and [2, 7] with refractory
    out = orc.spiking_run(1, 3, 1.0, np.zeros(3), 10, params=dict(Vt=5.0, Vr=3.0))
    steps = [s for s in range(10) if out["counts"][s] > 0]
    assert steps == [4, 6, 8]
    out = orc.spiking_run(1, 3, 1.0, np.zeros(3), 10, params=dict(Vt=3.0, Vr=1.0), refractory=np.full(3, 3))
    steps = [s for s in range(10) if out["counts"][s] > 0]
    assert steps == [2, 7]
    assert list(out["last_spike"]) == [7, 7, 7]


def test_spike_transmission_known_answer():
    # 5 inputs firing once at step 1 into 2 neurons, all-to-all w = 1
    M, steps = 5, 5
    unif = np.ones((steps, M))
    unif[1, :] = 0.0
    csr = conn.all_to_all(np.arange(M), np.arange(2), M, 2, 1.0, True)
    proj = orc.csrc_projection(csr, 0, pre_is_input=True)
    out = orc.spiking_run(1, 2, 1.0, np.zeros(2), steps, [proj], params=dict(Vt=1e9),
                          poisson=dict(M=M, rates=1.0, uniforms=unif), trace=True)
    assert list(out["input_counts"]) == [0, 5, 0, 0, 0]
    assert out["trace_gexc"][1].tolist() == [0.0, 0.0]
    assert out["trace_gexc"][2].tolist() == [5.0, 5.0]


def test_lif_free_membrane_and_decay():
    # no synapses: explicit Euler towards El + I with factor (1 - dt/tau)
    N, dt, steps = 4, 0.1, 60
    v0 = np.array([-60.0, -58.0, -55.0, -52.0])
    out = orc.spiking_run(0, N, dt, v0, steps, params=dict(I=5.0), trace=True,
                          g_exc=np.full(N, 0.0))
    vinf = -60.0 + 5.0
    k = np.arange(1, steps + 1)[:, None]
    np.testing.assert_allclose(out["trace_v"], vinf + (v0 - vinf)[None, :] * (1 - dt / 20.0) ** k, rtol=1e-12)
    # conductances decay geometrically and are not masked by refractoriness
    out = orc.spiking_run(0, N, dt, np.full(N, -70.0), steps, params=dict(I=0.0),
                          g_exc=np.full(N, 2.0), g_inh=np.full(N, 1.0), refractory=np.full(N, 50))
    np.testing.assert_allclose(out["g_exc"], 2.0 * (1 - dt / 5.0) ** steps, rtol=1e-12)
    np.testing.assert_allclose(out["g_inh"], 1.0 * (1 - dt / 10.0) ** steps, rtol=1e-12)


def _coba_constant_g_closed_form(v0, ge, gi, dt, k, El=-60.0, Ee=0.0, Ei=-80.0, I=20.0, tau=20.0):
    """Explicit Euler of tau dv/dt = El - v + gE (Ee - v) + gI (Ei - v) + I at
    constant conductances: v_k = v* + (v0 - v*) (1 - dt (1 + gE + gI) / tau)^k
    with v* = (El + gE Ee + gI Ei + I) / (1 + gE + gI)."""
    vs = (El + ge * Ee + gi * Ei + I) / (1.0 + ge + gi)
    return vs + (v0 - vs) * (1.0 - dt * (1.0 + ge + gi) / tau) ** k


def test_coba_conductance_coupled_membrane_closed_form():
    # Population/OpenMPGenerator.py:870-997: the COBA membrane is driven by
    # g_exc (Erev_exc = 0) and g_inh (Erev_inh = -80).  Infinite conductance
    # time constants hold g exactly (-g/inf = -0), so v follows the geometric
    # closed form.  The gI-only and gE-only rows fail if the reversal
    # potentials are swapped or a coupling term has the wrong sign.
    N, dt, steps = 4, 0.1, 80
    v0 = np.array([-70.0, -62.0, -58.0, -75.0])
    ge = np.array([0.0, 0.5, 0.0, 0.3])
    gi = np.array([0.0, 0.0, 0.7, 1.2])
    out = orc.spiking_run(0, N, dt, v0, steps, params=dict(tau_exc=np.inf, tau_inh=np.inf, Vt=1e9),
                          g_exc=ge, g_inh=gi, trace=True)
    np.testing.assert_array_equal(out["g_exc"], ge)
    np.testing.assert_array_equal(out["g_inh"], gi)
    k = np.arange(1, steps + 1)[:, None]
    np.testing.assert_allclose(out["trace_v"], _coba_constant_g_closed_form(v0, ge, gi, dt, k), rtol=1e-12)
    # the fixed points themselves: gE-only pulls towards Erev_exc, gI-only
    # below El + I towards Erev_inh
    vinf = orc.spiking_run(0, N, dt, v0, 20000, params=dict(tau_exc=np.inf, tau_inh=np.inf, Vt=1e9),
                           g_exc=ge, g_inh=gi)["v"]
    np.testing.assert_allclose(vinf, [-40.0, -40.0 / 1.5, -96.0 / 1.7, (-40.0 - 96.0) / 2.5], rtol=1e-10)


def test_coba_inhibitory_projection_reaches_membrane():
    # one input spike at step 0 through an inhibitory (target 1) projection of
    # weight 0.4: g_inh = 0.4 from step 1 on (held by tau_inh = inf), g_exc
    # stays 0, and v follows the gI-only closed form from v_1
    dt, steps = 0.1, 40
    unif = np.ones((steps, 1))
    unif[0, 0] = 0.0
    csr = conn.all_to_all(np.arange(1), np.arange(1), 1, 1, 0.4, True)
    proj = orc.csrc_projection(csr, 1, pre_is_input=True)
    v0 = np.array([-55.0])
    out = orc.spiking_run(0, 1, dt, v0, steps, [proj], params=dict(tau_exc=np.inf, tau_inh=np.inf, Vt=1e9),
                          poisson=dict(M=1, rates=1.0, uniforms=unif), trace=True)
    assert out["g_exc"][0] == 0.0 and out["g_inh"][0] == 0.4
    # step 0 runs free (v_1 = trace_v[0]); the spike of step 0 is delivered by
    # step 1's compute_psp, so trace_v[k] is k steps of gI = 0.4 from v_1
    v1 = _coba_constant_g_closed_form(v0, 0.0, 0.0, dt, 1)
    np.testing.assert_allclose(out["trace_v"][0], v1, rtol=1e-14)
    k = np.arange(1, steps)[:, None]
    np.testing.assert_allclose(out["trace_v"][1:], _coba_constant_g_closed_form(v1, 0.0, 0.4, dt, k), rtol=1e-12)


def test_lif_refractory_period():
    # strong drive: every spike is followed by exactly 50 refractory steps at v = Vr
    out = orc.spiking_run(0, 1, 0.1, np.array([-50.5]), 400, params=dict(I=100.0),
                          refractory=np.array([50]), trace=True)
    spikes = [s for s in range(400) if out["counts"][s]]
    assert len(spikes) >= 3
    for a, b in zip(spikes, spikes[1:]):
        assert np.all(out["trace_v"][a:a + 51, 0] == -60.0)
        assert b - a > 50


def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(0)
    dims = (16, 64, 64, 12)
    n_par = 64 * 16 + 64 + 64 * 64 + 64 + 12 * 64 + 12
    P = rng.normal(0, 0.3, n_par)
    X = rng.normal(0, 1, (20, 16))
    y = rng.integers(0, 12, 20)
    loss, G = orc.mlp_loss_grad(X, y, P, dims)
    assert n_par == 6028
    for k in rng.choice(n_par, 40, replace=False):
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        fd = (orc.mlp_loss_grad(X, y, Pp, dims, grad=False)[0] - orc.mlp_loss_grad(X, y, Pm, dims, grad=False)[0]) / (2 * h)
        assert abs(fd - G[k]) < 1e-6 + 1e-4 * abs(fd)


def test_mlp_forward_closed_form_and_loss():
    # permutation weights make the logits a closed form that a transposed or
    # shifted parameter block would break: hidden unit o < 16 copies input
    # (5*o) % 16, layer 2 is the identity with bias -0.5, logit c reads
    # hidden unit (c+3) % 16 and has bias c
    d_in, h1, h2, d_out = 16, 64, 64, 12
    W1 = np.zeros((h1, d_in))
    for o in range(16):
        W1[o, (5 * o) % 16] = 1.0
    W2 = np.eye(h2)
    b2 = np.full(h2, -0.5)
    W3 = np.zeros((d_out, h2))
    for c in range(d_out):
        W3[c, (c + 3) % 16] = 1.0
    b3 = np.arange(d_out, dtype=float)
    P = np.concatenate([W1.ravel(), np.zeros(h1), W2.ravel(), b2, W3.ravel(), b3])
    X = np.random.default_rng(4).normal(0, 1, (9, d_in))
    z = orc.mlp_forward(X, P)
    for c in range(d_out):
        src = (5 * ((c + 3) % 16)) % 16
        np.testing.assert_allclose(z[:, c], np.maximum(np.maximum(X[:, src], 0) - 0.5, 0) + c, rtol=0, atol=1e-15)
    # the cross-entropy of these logits is the loss of the gradient oracle
    y = np.random.default_rng(5).integers(0, d_out, 9)
    zs = z - z.max(axis=1, keepdims=True)
    ce = np.mean(np.log(np.exp(zs).sum(axis=1)) - zs[np.arange(9), y])
    loss, _ = orc.mlp_loss_grad(X, y, P, grad=False)
    assert abs(loss - ce) < 1e-12


# ---- delays -------------------------------------------------------------------
def _kat_input_sequence(nsteps):
    # input neuron `r = r + t : init = -1` (dt = 1): r at the end of step s
    r, out = -1.0, []
    for s in range(nsteps):
        r = r + s
        out.append(r)
    return np.array(out)


def test_rate_uniform_delay_kat():
    # tests_python/RateSynapse/test_RateDelays.py:180-320: three posts, one
    # pre each (w = 1), after reset() the queue holds the initial value -1
    seq = np.repeat(_kat_input_sequence(20)[:, None], 3, axis=1)
    rp, ci, w = np.array([0, 1, 2, 3]), np.array([2, 0, 1]), np.ones(3)
    for d, (a, b, c) in ((10, (9, 10, 14)), (5, (4, 5, 9))):
        out, _ = orc.delayed_psp_run(rp, ci, w, seq, d, delay=d, hist=np.full((d, 3), -1.0))
        np.testing.assert_array_equal(out[a], -1.0)   # "the first steps: initial value"
        np.testing.assert_array_equal(out[b], -1.0)   # "first queue value, t = 0"
        np.testing.assert_array_equal(out[c], 9.0)    # "t = 4"


def test_rate_nonuniform_delay_kat():
    # test_RateDelays.py:330-400: delays [[3], [5], [2]] -> [0, -1, 2] after 5
    # steps; [[3], [3], [3]] -> 20 after 10 steps
    seq = np.repeat(_kat_input_sequence(20)[:, None], 3, axis=1)
    rp, ci, w = np.array([0, 1, 2, 3]), np.array([0, 1, 2]), np.ones(3)
    out, _ = orc.delayed_psp_run(rp, ci, w, seq, 5, delays=[3, 5, 2], hist=np.full((5, 3), -1.0))
    np.testing.assert_array_equal(out[4], [0.0, -1.0, 2.0])
    out, _ = orc.delayed_psp_run(rp, ci, w, seq, 5, delays=[3, 3, 3], hist=np.full((5, 3), -1.0))
    np.testing.assert_array_equal(out[9], [20.0, 20.0, 20.0])


def test_rate_delay_one_is_no_delay_and_queue_is_a_shift():
    rng = np.random.default_rng(1)
    csr = conn.random_in_degree(50, 40, rng, np.full(40, 7), conn.Normal(0, 1, rng=rng))
    r0 = rng.random(50)
    seq = rng.random((12, 50))
    # delay 1 on every synapse reads the previous step's rates
    h0 = np.zeros((4, 50)); h0[0] = r0
    out, _ = orc.delayed_psp_run(csr.row_ptr, csr.col_idx, csr.w, seq, 4, delays=np.ones(csr.nnz, int), hist=h0)
    prev = np.vstack([r0, seq[:-1]])
    for s in range(12):
        np.testing.assert_allclose(out[s], csr.dense() @ prev[s], rtol=1e-12, atol=1e-12)
    # mixed delays: each synapse reads the rates d steps back
    d = rng.integers(1, 5, csr.nnz)
    out, _ = orc.delayed_psp_run(csr.row_ptr, csr.col_idx, csr.w, seq, 4, delays=d, hist=np.zeros((4, 50)))
    rows = np.repeat(np.arange(40), np.diff(csr.row_ptr))
    for s in range(4, 12):
        ref = np.zeros(40)
        np.add.at(ref, rows, csr.w * seq[s - d, csr.col_idx])
        np.testing.assert_allclose(out[s], ref, rtol=1e-12, atol=1e-12)
    # the recurrent run: delay 1 with the queue holding r0 is orc_rate_run;
    # per-synapse delays all equal to 3 are the uniform delay 3
    noise = rng.normal(0, 0.1, (30, 40))
    sq = conn.random_in_degree(40, 40, rng, np.full(40, 6), conn.Normal(0, 0.3, rng=rng))
    r0 = rng.random(40)
    h = np.zeros((1, 40)); h[0] = r0
    a = orc.rate_run_delayed(sq.row_ptr, sq.col_idx, sq.w, r0, 10.0, 1.0, noise, 30, 1, delay=1, hist=h)[0]
    b = orc.rate_run(sq.row_ptr, sq.col_idx, sq.w, r0, 10.0, 1.0, noise, 30)[0]
    np.testing.assert_array_equal(a, b)
    c = orc.rate_run_delayed(sq.row_ptr, sq.col_idx, sq.w, r0, 10.0, 1.0, noise, 30, 3,
                             delays=np.full(sq.nnz, 3))[0]
    e = orc.rate_run_delayed(sq.row_ptr, sq.col_idx, sq.w, r0, 10.0, 1.0, noise, 30, 3, delay=3)[0]
    np.testing.assert_array_equal(c, e)
    assert not np.array_equal(c, b)


def _emitter_projection(n_pre, n_post, w, delays=None, delay=1):
    csr = conn.all_to_all(np.arange(n_pre), np.arange(n_post), n_pre, n_post, w, True)
    return csr, orc.csrc_projection(csr, 0, pre_is_input=True, delay=delay, delays=delays)


def _emit_at_step1(steps, n):
    # Poisson rule p = 1000*u/dt < rates with rates = 1: u = 0 spikes, u = 1 does not
    u = np.ones((steps, n))
    u[1] = 0.0
    return dict(M=n, rates=1.0, uniforms=u)


def test_spiking_uniform_delay_kat():
    # test_SpikingTransmission.py:86-163: 5 emitters fire at t == 1, all-to-all
    # w = 1 with delay 3 -> g_exc == 5 at step 4 only.  Receivers: COBA with
    # tau_exc = dt, so g_exc returns to 0 in every update (`g_exc = 0`).
    csr, pr = _emitter_projection(5, 2, 1.0, delay=3)
    ref = orc.spiking_run(0, 2, 1.0, np.full(2, -60.0), 6, [pr], params=dict(tau_exc=1.0, I=0.0),
                          poisson=_emit_at_step1(6, 5), trace=True)
    np.testing.assert_array_equal(ref["trace_gexc"], [[0, 0], [0, 0], [0, 0], [0, 0], [5, 5], [0, 0]])


def test_spiking_nonuniform_delay_kat():
    # test_SpikingTransmission.py:166-253: delays [[1,2,3,2,1],[3,2,1,2,3]]
    csr, _ = _emitter_projection(5, 2, 1.0)
    d = np.array([[1, 2, 3, 2, 1], [3, 2, 1, 2, 3]]).ravel()   # forward CSR order (post-major)
    pr = orc.csrc_projection(csr, 0, pre_is_input=True, delays=d)
    ref = orc.spiking_run(0, 2, 1.0, np.full(2, -60.0), 5, [pr], params=dict(tau_exc=1.0, I=0.0),
                          poisson=_emit_at_step1(5, 5), trace=True)
    np.testing.assert_array_equal(ref["trace_gexc"], [[0, 0], [0, 0], [2, 1], [2, 2], [1, 2]])


def test_spmv_csr_f32_error_bound():
    """FP32 row sums (sequential, j ascending) stay within the classical bound
    n * 2^-24 * sum |w r| of the exact product of the FP32 inputs."""
    rng = np.random.default_rng(11)
    csr = conn.random_in_degree(400, 300, rng, rng.integers(0, 200, 300), conn.Normal(0, 1, rng=rng))
    w32 = csr.w.astype(np.float32)
    r32 = rng.random(400).astype(np.float32)
    got = orc.spmv_csr_f32(csr.row_ptr, csr.col_idx, w32, r32).astype(np.float64)
    for i in range(300):
        seg = slice(csr.row_ptr[i], csr.row_ptr[i + 1])
        terms = w32[seg].astype(np.float64) * r32[csr.col_idx[seg]].astype(np.float64)
        n = max(1, terms.size)
        assert abs(got[i] - terms.sum()) <= 2 * n * 2.0 ** -24 * np.abs(terms).sum() + 1e-30


def test_sweep_instance_closed_form():
    """One sweep instance at omega = 0: the input rates are the constant
    u_k = 1 + 0.5 sin(phase_k) from step 0's update on, so with self-weights
    beta the rate of an input row relaxes geometrically to g (W_in u)_i / (1 - beta):
    r_t = r* (1 - a^(t-1)), a = 1 - (1 - beta) dt / tau; rows beyond n_in stay 0."""
    n, n_in, beta, g, tau, dt, steps = 40, 25, 0.3, 0.7, 10.0, 1.0, 60
    rng = np.random.default_rng(2)
    w_in = rng.random((n_in, n_in)) * 0.1
    phase = rng.random(n_in) * 2 * np.pi
    eye = conn.from_arrays(n, n, np.arange(n + 1, dtype=np.int64), np.arange(n, dtype=np.int32), np.full(n, beta))
    r, tr = orc.sweep_instance_run(eye, w_in, phase, 0.0, g, tau, dt, np.zeros(n), steps, trace=True)
    u = 1.0 + 0.5 * np.sin(phase)
    r_star = np.zeros(n)
    r_star[:n_in] = g * (w_in @ u) / (1.0 - beta)
    a = 1.0 - (1.0 - beta) * dt / tau
    for t in (1, 5, 30, steps):
        np.testing.assert_allclose(tr[t - 1], r_star * (1.0 - a ** (t - 1)), rtol=1e-12, atol=1e-15)
    assert np.all(r[n_in:] == 0.0)
