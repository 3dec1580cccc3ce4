"""Pins of the oracle's plastic projection (orc_plastic_run) on the CPU.

Each expected value below is derived by hand from the event semantics of the
reference's generated code (Projection/OpenMPGenerator.py:1182-1230 and
:1575-1600, Projection/OpenMP/CSR.py:870-918): a pre spike of step s is
handled at step s+1 after decaying the event-driven traces to step s, with
_last_event = s+1; a post spike of step s is handled at step s after decaying
to s.  A wrong decay interval, a g += w placed after the weight change, a
missing dt or a swapped pre/post index changes one of the numbers.
"""

import math

import numpy as np

from oracle import oracle as orc

STDP_PRM = dict(tau_plus=20.0, tau_minus=15.0, A_plus=0.01, A_minus=0.012, w_min=0.0, w_max=1.0)


def _trains(steps, n, spikes):
    """spikes: {step: [ranks]} -> per-step counts and concatenated ranks."""
    counts = np.zeros(steps, dtype=np.int32)
    idx = []
    for s in range(steps):
        r = sorted(spikes.get(s, []))
        counts[s] = len(r)
        idx += r
    return counts, np.asarray(idx, dtype=np.int32)


def _stdp(w0, pre, post, steps, dt, prm=STDP_PRM, n_pre=1, n_post=1, row_ptr=(0, 1), col=(0,)):
    pc, pi = _trains(steps, n_pre, pre)
    qc, qi = _trains(steps, n_post, post)
    p = [prm[k] for k in ("tau_plus", "tau_minus", "A_plus", "A_minus", "w_min", "w_max")]
    return orc.plastic_run(0, row_ptr, col, n_pre, w0, p, dt, pc, pi, qc, qi)


def test_stdp_single_synapse_closed_form():
    dt, w0 = 0.5, 0.5
    tp, tm, Ap, Am = 20.0, 15.0, 0.01, 0.012
    # pre spikes at 2 and 6 (handled at 3 and 7), post spike at 10, pre at 14 (handled at 15)
    out = _stdp(w0, {2: [0], 6: [0], 14: [0]}, {10: [0]}, 20, dt)
    # step 7: x decays from 3 to 6 (three steps), then + A_plus
    x7 = Ap * (1.0 + math.exp(-3 * dt / tp))
    # step 10 (post): x decays from 7 to 10, y = -A_minus, w += x
    x10 = x7 * math.exp(-3 * dt / tp)
    w1 = w0 + x10
    # step 15 (pre): decays from 10 to 14; g gets w1 before w += y
    y15 = -Am * math.exp(-4 * dt / tm)
    w2 = w1 + y15
    x15 = x10 * math.exp(-4 * dt / tp) + Ap
    np.testing.assert_allclose(out["w"], [w2], rtol=0, atol=1e-15)
    np.testing.assert_allclose(out["x"], [x15], rtol=0, atol=1e-16)
    np.testing.assert_allclose(out["y"], [y15], rtol=0, atol=1e-16)
    assert out["last_event"][0] == 15
    g = out["g"][:, 0]
    np.testing.assert_allclose(g[[3, 7, 15]], [w0, w0, w1], rtol=0, atol=1e-15)
    assert np.count_nonzero(g) == 3


def test_stdp_post_before_pre_depresses():
    dt, w0 = 1.0, 0.3
    tm, Am = 15.0, 0.012
    # post at 4, pre at 9 (handled at 10, y decayed from 4 to 9): w -= A_minus e^{-5 dt/tau_minus}
    out = _stdp(w0, {9: [0]}, {4: [0]}, 12, dt)
    np.testing.assert_allclose(out["w"], [w0 - Am * math.exp(-5 * dt / tm)], rtol=0, atol=1e-15)
    # the same spike step on both sides: the pre event of step s+1 sees the post spike of step s
    out = _stdp(w0, {5: [0]}, {5: [0]}, 8, dt)
    np.testing.assert_allclose(out["w"], [w0 - Am], rtol=0, atol=1e-15)


def test_stdp_clipping():
    prm = dict(STDP_PRM, w_max=0.505, A_plus=0.5)
    out = _stdp(0.5, {1: [0], 8: [0]}, {3: [0]}, 10, 1.0, prm=prm)
    # potentiation clipped at w_max, the next pre event transmits w_max
    assert out["w"][0] <= 0.505
    assert out["g"][9, 0] == 0.505
    prm = dict(STDP_PRM, A_minus=5.0)
    out = _stdp(0.5, {6: [0]}, {2: [0]}, 9, 1.0, prm=prm)
    assert out["w"][0] == 0.0


def test_stdp_synapses_are_independent():
    """A 3 x 4 projection evolves each synapse as the 1 x 1 run of its own
    pre and post trains (catches row/column and index mix-ups)."""
    rng = np.random.default_rng(3)
    steps, n_pre, n_post = 60, 4, 3
    rows = [[0, 2, 3], [1, 2], [0, 1, 3]]
    row_ptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
    col = np.concatenate(rows)
    w0 = rng.uniform(0.2, 0.8, col.size)
    pre = {s: list(np.flatnonzero(rng.random(n_pre) < 0.15)) for s in range(steps)}
    post = {s: list(np.flatnonzero(rng.random(n_post) < 0.15)) for s in range(steps)}
    out = _stdp(w0, pre, post, steps, 0.5, n_pre=n_pre, n_post=n_post, row_ptr=row_ptr, col=col)
    g_sum = np.zeros((steps, n_post))
    for i in range(n_post):
        for j in range(row_ptr[i], row_ptr[i + 1]):
            c = col[j]
            one = _stdp(w0[j], {s: [0] for s in pre if c in pre[s]}, {s: [0] for s in post if i in post[s]},
                        steps, 0.5)
            assert one["w"][0] == out["w"][j] and one["x"][0] == out["x"][j] and one["y"][0] == out["y"][j]
            g_sum[:, i] += one["g"][:, 0]
    np.testing.assert_allclose(out["g"], g_sum, rtol=0, atol=1e-15)


def test_time_dependent_rule_closed_form():
    """t_pre / t_post rule of test_SpikingSynapse.py (test_TimeDependentUpdate)."""
    dt, w0 = 0.5, 0.001
    tau_pre, tau_post, cApre, cApost, wmax = 10.0, 10.0, 0.01, 0.0105, 0.01
    pc, pi = _trains(12, 1, {2: [0], 8: [0]})
    qc, qi = _trains(12, 1, {5: [0]})
    out = orc.plastic_run(1, [0, 1], [0], 1, w0, [tau_pre, tau_post, cApre, cApost, wmax], dt, pc, pi, qc, qi)
    # step 3: t_post = -10000 ms -> no change; step 5: + cApre e^{(2 - 5) dt / tau_pre}
    w1 = w0 + cApre * math.exp((2 - 5) * dt / tau_pre)
    # step 9: - cApost e^{(5 - 9) dt / tau_post}
    w2 = w1 - cApost * math.exp((5 - 9) * dt / tau_post)
    np.testing.assert_allclose(out["w"], [w2], rtol=0, atol=1e-16)
    np.testing.assert_allclose(out["g"][[3, 9], 0], [w0, w1], rtol=0, atol=1e-16)
    assert out["last_event"][0] == -10000          # no event-driven variable
    assert out["pre_last"][0] == 8 and out["post_last"][0] == 5


def test_stdp_uniform_delay_closed_form():
    """With a delay of 4 steps the spike of step 2 is handled at step 6; the
    trace then decays from 6 to the post spike of step 9."""
    pc, pi = _trains(12, 1, {2: [0]})
    qc, qi = _trains(12, 1, {9: [0]})
    p = [STDP_PRM[k] for k in ("tau_plus", "tau_minus", "A_plus", "A_minus", "w_min", "w_max")]
    out = orc.plastic_run(0, [0, 1], [0], 1, 0.5, p, 0.5, pc, pi, qc, qi, delay=4)
    assert np.flatnonzero(out["g"][:, 0]).tolist() == [6]
    np.testing.assert_allclose(out["w"], [0.5 + 0.01 * math.exp(-3 * 0.5 / 20.0)], rtol=0, atol=1e-15)


# -- rate-coded update_synapse (orc_rate_plastic_run) ---------------------------

def _one(rule, w0, a, steps, prm, dt=1.0, period=1, theta0=None):
    seq = np.full((steps, 1), a)
    return orc.rate_plastic_run(rule, [0, 1], [0], w0, prm, dt, a, seq, period=period, theta0=theta0)


def test_hebb_geometric_growth():
    """One synapse at a constant pre rate a: the psp of step t is w*a, the
    post rate of the same step, so every update multiplies w by
    1 + dt*eta*a^2."""
    a, w0, eta, dt, steps = 0.7, 0.2, 0.01, 0.5, 40
    out = _one(0, w0, a, steps, [eta], dt=dt)
    g = 1.0 + dt * eta * a * a
    np.testing.assert_allclose(out["w"], [w0 * g ** steps], rtol=1e-13)
    np.testing.assert_allclose(out["sum"][:, 0], w0 * g ** np.arange(steps) * a, rtol=1e-13)


def test_hebb_update_period():
    """period 3: updates at steps 0, 3, 6 with dt scaled by 3."""
    a, w0, eta = 0.7, 0.2, 0.01
    out = _one(0, w0, a, 9, [eta], period=3)
    np.testing.assert_allclose(out["w"], [w0 * (1.0 + 3 * eta * a * a) ** 3], rtol=1e-13)
    s = out["sum"][:, 0]
    assert s[1] == s[2] == s[3] and s[4] == s[6] and s[3] < s[4]


def test_oja_fixed_point():
    """Oja's rule: dw = eta a^2 w (1 - alpha w^2) for one synapse, so w
    settles at 1 / sqrt(alpha) from either side, whatever the rate a (for
    Euler steps small enough not to overshoot below 0, where w stays)."""
    alpha = 4.0
    for a in (0.5, 1.3):
        for w0 in (0.2, 0.9):
            out = _one(1, w0, a, 20000, [0.05, alpha])
            np.testing.assert_allclose(out["w"], [1.0 / np.sqrt(alpha)], rtol=1e-9)


def test_ibcm_threshold_relaxation():
    """eta = 0 keeps w and the post rate b = w*a fixed; theta relaxes as
    b^2 (1 - exp(-dt/tau)^(t+1))."""
    a, w0, tau, dt, steps = 0.5, 0.8, 25.0, 1.0, 60
    out = _one(2, w0, a, steps, [0.0, tau], dt=dt)
    b = w0 * a
    np.testing.assert_allclose(out["theta"], [b * b * (1.0 - np.exp(-dt / tau) ** steps)], rtol=1e-12)
    np.testing.assert_array_equal(out["w"], [w0])
    # with learning: below threshold (theta > post rate^2 is impossible at start) w grows while post.r > theta
    out = _one(2, w0, a, 5, [0.1, tau], dt=dt)
    assert out["w"][0] > w0


def test_stp_closed_form():
    """STP (Tsodyks-Markram): the first spike transmits w*U and leaves
    x = 1 - U, u = U + U(1 - U); between spikes x relaxes to 1 with tau_rec
    and u to U with tau_facil (from the handled step, the decay ending one
    step before the next handled spike)."""
    tau_rec, tau_facil, U, dt, w = 100.0, 50.0, 0.3, 1.0, 2.0
    pc, pi = _trains(30, 1, {4: [0], 14: [0]})
    qc, qi = _trains(30, 1, {})
    out = orc.plastic_run(2, [0, 1], [0], 1, w, [tau_rec, tau_facil, U], dt, pc, pi, qc, qi,
                          x0=[1.0], y0=[U])
    x1, u1 = 1.0 - U, U + U * (1.0 - U)
    # second spike handled at 15: decay from 5 to 14
    x2 = 1.0 + (x1 - 1.0) * math.exp(-9 * dt / tau_rec)
    u2 = U + (u1 - U) * math.exp(-9 * dt / tau_facil)
    g = out["g"][:, 0]
    assert np.flatnonzero(g).tolist() == [5, 15]
    np.testing.assert_allclose(g[[5, 15]], [w * U, w * u2 * x2], rtol=1e-15)
    np.testing.assert_allclose(out["x"], [x2 * (1.0 - u2)], rtol=1e-15)
    np.testing.assert_allclose(out["y"], [u2 + U * (1.0 - u2)], rtol=1e-15)
