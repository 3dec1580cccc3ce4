"""Rate-coded path on the GPU against the CPU oracle.

compute_psp (generator/Projection/OpenMP/CSR.py:178-196) in every storage
format, the rate-neuron update with injected noise (the oracle's
orc_rate_run), the FP32 + Kahan form, and ragged / empty-row connectivity.
Tolerances: FP64 psp sums differ from the sequential CPU loop only by the
summation order -> rtol 1e-12 on W.r; 100-step traces -> 1e-9 relative;
FP32 + Kahan psp -> 2e-6 relative to the FP64 oracle.
"""

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import connectivity as conn
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FORMATS = [
    dict(storage_format="csr", lanes=1),
    dict(storage_format="csr", lanes=8),
    dict(storage_format="csr", lanes=32),
    dict(storage_format="csr", lanes=64),
    dict(storage_format="csr_vec", lanes=4),
    dict(storage_format="csr_vec", lanes=16),
    dict(storage_format="csr_vec", lanes=64),
    dict(storage_format="ell"),
    dict(storage_format="ellr"),
    dict(storage_format="sell", lanes=1, sigma=1),
    dict(storage_format="sell", lanes=4, sigma=256),
    dict(storage_format="sell", lanes=16, sigma=64),
    dict(storage_format="bsr", chunk=2),
    dict(storage_format="bsr", chunk=4),
    dict(storage_format="bsr", chunk=8),
    dict(storage_format="dense"),
    dict(storage_format="csr_panel", lanes=16, panel_cols=512),
    dict(storage_format="csr_panel", lanes=4, panel_cols=256, chunk=100),
    dict(storage_format="sell_panel", lanes=1, panel_cols=300, chunk=64),
    dict(storage_format="sell_panel", lanes=2, panel_cols=512, chunk=128),
    dict(storage_format="sell_panel", lanes=4, panel_cols=4096, chunk=256),
    dict(storage_format="sell_panel", lanes=8, sigma=8, panel_cols=1000, chunk=192),
    dict(storage_format="sell_panel", lanes=4, sigma=8, panel_cols=2048, chunk=1024),
    dict(storage_format="sell_panel"),
]


def _rel_err(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def _net_one_step(csr, r0, fmt, precision="double"):
    net = ah.Network(dt=1.0, seed=3, precision=precision)
    pop = net.create(csr.n_rows, ah.RateNoisy)
    proj = net.connect(pop, pop, "exc")
    proj.from_csr(csr, **fmt)
    net.compile()
    pop.r = r0
    pop.inject_random(0, np.zeros((1, csr.n_rows)))
    net.step()
    return net, pop, proj


def _random_csr(n, k, seed, ragged=False):
    rng = np.random.default_rng(seed)
    if ragged:
        deg = np.minimum(n, np.maximum(0, np.round(rng.lognormal(np.log(k) - 0.5, 1.0, n)))).astype(np.int64)
        deg[rng.random(n) < 0.05] = 0  # empty rows
    else:
        deg = np.full(n, k, dtype=np.int64)
    return conn.random_in_degree(n, n, rng, deg, conn.Normal(0.0, 1.0 / np.sqrt(max(k, 1)), rng=rng))


@pytest.mark.parametrize("fmt", FORMATS, ids=lambda f: "-".join(f"{k}{v}" for k, v in f.items()))
@pytest.mark.parametrize("ragged", [False, True])
def test_psp_formats_match_oracle(gpu, fmt, ragged):
    n = 1500
    csr = _random_csr(n, 60, 11 + int(ragged), ragged)
    r0 = np.random.default_rng(5).random(n)
    net, pop, proj = _net_one_step(csr, r0, fmt)
    got = pop.get("_sum_exc").reshape(-1)
    ref = orc.spmv_csr(csr.row_ptr, csr.col_idx, csr.weights(), r0)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-13 * np.max(np.abs(ref)))
    # closed form: the dense product (tests_python/RateSynapse/test_RateDelays.py:99-153)
    np.testing.assert_allclose(got, csr.dense() @ r0, rtol=1e-10, atol=1e-12)


def test_psp_constant_weight_and_view_ranks(gpu):
    # one constant weight (core/Projection.py:823-839) from fixed_probability
    net = ah.Network(dt=1.0, seed=42)
    pop = net.create(800, ah.RateNoisy)
    proj = net.connect(pop[0:600], pop, "exc")
    proj.fixed_probability(0.1, 0.25)
    net.compile()
    r0 = np.linspace(0.0, 1.0, 800)
    pop.r = r0
    pop.inject_random(0, np.zeros((1, 800)))
    net.step()
    csr = proj.csr
    assert csr.w is None and csr.col_idx.max() < 600
    ref = orc.spmv_csr(csr.row_ptr, csr.col_idx, csr.weights(), r0)
    np.testing.assert_allclose(pop.get("_sum_exc").reshape(-1), ref, rtol=1e-12, atol=1e-14)


def test_rate_run_matches_oracle_100_steps(gpu):
    n, steps = 2000, 100
    csr = _random_csr(n, 100, 3)
    rng = np.random.default_rng(9)
    r0 = rng.random(n)
    noise = rng.normal(0.0, 0.1, (steps, n))
    net = ah.Network(dt=1.0, seed=1)
    pop = net.create(n, ah.RateNoisy)
    proj = net.connect(pop, pop, "exc")
    proj.from_csr(csr)
    mon = net.monitor(pop, ["r"], capacity=steps)
    net.compile(graph_batch=16)
    pop.r = r0
    pop.inject_random(0, noise)
    net.run(steps)
    trace = mon.get("r")
    r_ref, _, tr_ref = orc.rate_run(csr.row_ptr, csr.col_idx, csr.w, r0, 10.0, 1.0, noise, steps, trace=True)
    assert trace.shape == (steps, n)
    for s in range(steps):
        assert _rel_err(trace[s], tr_ref[s]) < 1e-9, s
    assert _rel_err(pop.r.reshape(-1), r_ref) < 1e-9


@pytest.mark.parametrize("fmt", [dict(storage_format="sell_panel", lanes=4, sigma=8, panel_cols=700, chunk=128),
                                 dict(storage_format="sell_panel", lanes=2, sigma=4, panel_cols=2000, chunk=64),
                                 dict(storage_format="csr_panel", lanes=8, panel_cols=700, chunk=128)],
                         ids=["sell_panel-3panels", "sell_panel-1panel", "csr_panel"])
def test_panel_psp_relaunch_with_new_rates(gpu, fmt):
    # the partial-sum buffer of the tiles is reused: every launch must read
    # the current rates and write every row, launch after launch
    n = 2000
    csr = _random_csr(n, 80, 31, ragged=True)
    net, pop, proj = _net_one_step(csr, np.zeros(n), fmt)
    rng = np.random.default_rng(17)
    for it in range(4):
        r = rng.random(n)
        pop.r = r
        proj.time_psp(1 + it)
        got = pop.get("_sum_exc").reshape(-1)
        ref = orc.spmv_csr(csr.row_ptr, csr.col_idx, csr.weights(), r)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-13 * np.max(np.abs(ref)))


def test_sell_panel_heavy_rows(gpu):
    # a few rows far longer than the rest get several lanes per row (up to a
    # whole wave); their lane sums are added by the row's first lane
    n = 3000
    rng = np.random.default_rng(23)
    deg = rng.integers(0, 20, n)
    deg[[3, 700, 701, 2999]] = [n, 2500, 900, 300]
    csr = conn.random_in_degree(n, n, rng, deg, conn.Normal(0.0, 0.05, rng=rng))
    r0 = rng.random(n)
    ref = orc.spmv_csr(csr.row_ptr, csr.col_idx, csr.weights(), r0)
    for fmt in (dict(storage_format="sell_panel", lanes=4, sigma=8, panel_cols=1000, chunk=64),
                dict(storage_format="sell_panel", lanes=2, sigma=4, panel_cols=3000, chunk=1024)):
        net, pop, proj = _net_one_step(csr, r0, fmt)
        got = pop.get("_sum_exc").reshape(-1)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-13 * np.max(np.abs(ref)))


@pytest.mark.parametrize("precision", ["double", "float"])
def test_sell_panel_large_lds_panels(gpu, precision):
    # panels of up to 160 KiB of LDS (20000 FP64 / 40000 FP32 columns, one
    # workgroup per CU), the layout of the 1e8-synapse benchmark
    n = 45000
    csr = _random_csr(n, 40, 31, ragged=True)
    r0 = np.random.default_rng(8).random(n)
    ref = orc.spmv_csr(csr.row_ptr, csr.col_idx, csr.weights(), r0)
    pc = 20479 if precision == "double" else 40959
    fmt = dict(storage_format="sell_panel", lanes=2, sigma=8, panel_cols=pc, chunk=2048)
    net, pop, proj = _net_one_step(csr, r0, fmt, precision)
    assert proj.get_format()["panel_cols"] > 8192   # more than 64 KiB of LDS in FP64
    got = pop.get("_sum_exc").reshape(-1)
    if precision == "double":
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-13 * np.max(np.abs(ref)))
    else:
        # FP32 with compensated sums: within 1e-5 of the FP64 oracle
        assert _rel_err(got, ref) < 1e-5


def test_sell_panel_graph_run_matches_oracle(gpu):
    n, steps = 3000, 48
    csr = _random_csr(n, 120, 6, ragged=True)
    rng = np.random.default_rng(10)
    r0 = rng.random(n)
    noise = rng.normal(0.0, 0.1, (steps, n))
    net = ah.Network(dt=1.0, seed=1)
    pop = net.create(n, ah.RateNoisy)
    net.connect(pop, pop, "exc").from_csr(csr, storage_format="sell_panel", lanes=4, sigma=8,
                                          panel_cols=1024, chunk=256)
    net.compile(graph_batch=16)
    pop.r = r0
    pop.inject_random(0, noise)
    net.run(steps)
    r_ref, _, _ = orc.rate_run(csr.row_ptr, csr.col_idx, csr.w, r0, 10.0, 1.0, noise, steps)
    assert _rel_err(pop.r.reshape(-1), r_ref) < 1e-9


@pytest.mark.parametrize("precision", ["double", "float"])
def test_panel_sums_completed_by_the_update(gpu, precision):
    # a panel layout's partial sums are added by the rate update itself
    # (no reduction launch) unless a monitor reads the sum right after
    # compute_psp: both networks give bit-identical r and sums (the same
    # panel order, compensated in FP32), equal to the oracle within 1e-9
    n, steps = 3000, 40
    csr = _random_csr(n, 120, 8, ragged=True)
    rng = np.random.default_rng(12)
    r0 = rng.random(n)
    noise = rng.normal(0.0, 0.1, (steps, n))
    out = []
    for watch_sum in (False, True):
        net = ah.Network(dt=1.0, seed=1, precision=precision)
        pop = net.create(n, ah.RateNoisy)
        net.connect(pop, pop, "exc").from_csr(csr, storage_format="sell_panel", lanes=2, sigma=8,
                                              panel_cols=700, chunk=512)
        mon = net.monitor(pop, ["_sum_exc"] if watch_sum else ["r"], capacity=steps)
        net.compile(graph_batch=16)
        pop.r = r0
        pop.inject_random(0, noise)
        net.run(steps)
        out.append((pop.r.reshape(-1).copy(), pop.get("_sum_exc").reshape(-1).copy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    if precision == "double":
        r_ref, _, _ = orc.rate_run(csr.row_ptr, csr.col_idx, csr.w, r0, 10.0, 1.0, noise, steps)
        assert _rel_err(out[0][0], r_ref) < 1e-9


def test_rate_graph_and_eager_agree_bitwise(gpu):
    n, steps = 1000, 40
    csr = _random_csr(n, 50, 4)
    rng = np.random.default_rng(2)
    r0 = rng.random(n)
    noise = rng.normal(0.0, 0.1, (steps, n))
    finals = []
    for batch, graph in ((8, True), (8, False), (40, True)):
        net = ah.Network(dt=1.0, seed=1)
        pop = net.create(n, ah.RateNoisy)
        net.connect(pop, pop, "exc").from_csr(csr)
        net.compile()
        net.set_graph(graph, batch)
        pop.r = r0
        pop.inject_random(0, noise)
        net.run(steps)
        finals.append(pop.r.reshape(-1).copy())
        assert net.time == steps
    assert np.array_equal(finals[0], finals[1]) and np.array_equal(finals[0], finals[2])


def test_fp32_kahan_psp_error(gpu):
    n = 3000
    csr = _random_csr(n, 500, 21)
    r0 = np.random.default_rng(8).random(n)
    ref = orc.spmv_csr(csr.row_ptr, csr.col_idx, csr.weights(), r0)
    for fmt in (dict(storage_format="csr", lanes=64), dict(storage_format="sell", lanes=4, sigma=64)):
        net, pop, proj = _net_one_step(csr, r0, fmt, precision="float")
        got = pop.get("_sum_exc").reshape(-1)
        # inputs are rounded to FP32 first: compare with the FP64 sum of the rounded inputs
        ref32 = orc.spmv_csr(csr.row_ptr, csr.col_idx, csr.weights().astype(np.float32).astype(np.float64),
                             r0.astype(np.float32).astype(np.float64))
        scale = np.abs(csr.weights()).max() * np.sqrt(500)
        assert np.max(np.abs(got - ref32)) / scale < 2e-7
        assert np.max(np.abs(got - ref)) / scale < 2e-6


def test_noise_generator_statistics(gpu):
    # on-device Normal(0, 0.1) draws: r = 0, tau = dt -> r(t+1) = max(0, rand)
    n = 200000
    net = ah.Network(dt=1.0, seed=123)
    pop = net.create(n, ah.RateNoisy)
    net.compile()
    pop.tau = 1.0
    net.step()
    r = pop.r.reshape(-1)
    # half-normal after the clip at 0: P(r == 0) = 1/2, E[r] = sigma*sqrt(2/pi)/2
    assert abs(np.mean(r == 0.0) - 0.5) < 0.01
    assert abs(np.mean(r) - 0.1 * np.sqrt(2 / np.pi) / 2) < 1e-3
    r2 = pop.r.reshape(-1)
    net.step()
    assert not np.array_equal(r2, pop.r.reshape(-1))
