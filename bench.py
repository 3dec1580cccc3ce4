"""Benchmark of the simulation step on one MI355X.

Default workload: the rate network of 100,000 neurons with one recurrent
projection of fixed in-degree 1000 (1e8 synapses, weights N(0, 1/sqrt(1000)),
int32 indices), tau dr/dt = -r + sum(exc) + N(0, 0.1), r >= 0, dt = 1 ms, FP64,
CSR.  A step = compute_psp (SpMV) + the fused neuron update with on-device
noise, captured in a HIP graph.  metric = synaptic events/s = nnz / step time.

roofline: the psp kernel's algorithmic bytes per launch
  nnz*(8+4) + (N+1)*4 + N*8 + 2*N*8
divided by its mean launch time from HIP events on the network stream.
cpu_baseline: the oracle's OpenMP CSR step (the reference's default scalar
loop) on the host cores, a bounded number of steps of the same network.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_c1(n, k, seed, lognormal=False):
    from annarchy_amd import connectivity as conn

    rng = np.random.default_rng(seed)
    if lognormal:
        deg = np.round(rng.lognormal(np.log(k) - 0.5, 1.0, n)).astype(np.int64)
        deg = np.clip(deg, 1, n)
    else:
        deg = np.full(n, k, dtype=np.int64)
    w = conn.Normal(0.0, 1.0 / np.sqrt(1000.0), rng=rng)
    return conn.random_in_degree(n, n, rng, deg, w)


KERNEL_OF_FORMAT = {"csr": "csr_rows", "sell": "sell_slices", "csr_panel": "tile_psp", "csr_vec": "csr_vec",
                    "ell": "ell_rows", "ellr": "ell_rows", "bsr": "bsr_rows", "dense": "dense_rows",
                    "tiles": "tile_psp", "sell_panel": "sell_panel_psp"}


def traffic_from_profiles(name):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(csr, seconds, threads):
    from oracle import oracle as orc

    n = csr.n_rows
    r = np.random.default_rng(1).random(n)
    noise = np.random.default_rng(2).normal(0.0, 0.1, (1, n))
    # warm-up step
    orc.rate_run(csr.row_ptr, csr.col_idx, csr.w, r, 10.0, 1.0, noise, 1, nthreads=threads)
    steps, t0 = 0, time.perf_counter()
    while True:
        orc.rate_run(csr.row_ptr, csr.col_idx, csr.w, r, 10.0, 1.0, noise, 1, nthreads=threads)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 1000:
            break
    return csr.nnz * steps / el, steps, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--neurons", type=int, default=100_000)
    ap.add_argument("--degree", type=int, default=1000)
    ap.add_argument("--lognormal", action="store_true")
    ap.add_argument("--precision", default="double", choices=["double", "float"])
    # default: sliced tiles (2048-row blocks, 1024 in FP32, x 20479-column
    # panels staged in 160 KiB of LDS, sorted 64-row slices, one lane per row,
    # 2 steps of 4 synapses in flight, 8 waves per workgroup): the fastest
    # layout of the sweep
    ap.add_argument("--format", default="sell_panel")
    ap.add_argument("--lanes", type=int, default=2,
                    help="csr/csr_vec/csr_panel: lanes per row; sell: threads per row; sell_panel: steps in flight")
    ap.add_argument("--sigma", type=int, default=1, help="sell: sorting window")
    ap.add_argument("--waves", type=int, default=8, help="sell_panel: waves per workgroup")
    ap.add_argument("--panel-cols", type=int, default=20479,
                    help="sell_panel: columns per panel at most (20479: 160 KiB of LDS in FP64)")
    ap.add_argument("--chunk", type=int, default=0,
                    help="sell_panel: rows per tile (0: 2048 in FP64, 1024 in FP32)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", default="c1", choices=["c0", "c1", "c2", "c3", "stdp"],
                    help="c1 (default): rate network 1e8 synapses; c0: COBA 4000 neurons; "
                         "c2: COBA 400k neurons / 4e8 synapses; c3: batched sweep of 256 instances")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the simulation step of this backend is a "
              "single-device path (multi-GPU sharding is a design note, not built).", file=sys.stderr)
        sys.exit(2)

    if args.config != "c1":
        print(json.dumps(globals()["bench_" + args.config](args)))
        return

    import annarchy_amd as ah

    t_build = time.perf_counter()
    csr = build_c1(args.neurons, args.degree, 1234, args.lognormal)
    log(f"connectivity: {csr.nnz} synapses in {time.perf_counter() - t_build:.1f} s")

    net = ah.Network(dt=1.0, seed=42, precision=args.precision)
    pop = net.create(args.neurons, ah.RateNoisy)
    proj = net.connect(pop, pop, "exc")
    chunk = args.chunk or (2048 if args.precision == "double" else 1024)
    proj.from_csr(csr, storage_format=args.format, lanes=args.lanes,
                  sigma=args.waves if args.format == "sell_panel" else args.sigma,
                  panel_cols=args.panel_cols, chunk=chunk)
    net.compile(allow_build=False)
    pop.r = np.random.default_rng(7).random(args.neurons)
    log("network uploaded")

    net.run(args.warmup)
    t0 = time.perf_counter()
    net.run(args.steps)  # ahip_run synchronizes the stream before returning
    el = time.perf_counter() - t0
    ms_step = el * 1000.0 / args.steps
    events = csr.nnz * args.steps / el

    import ctypes as C
    from annarchy_amd import _lib

    copy_ms = C.c_double()
    copy_bytes = 1 << 31
    _lib.check(_lib.lib().ahip_probe_copy(copy_bytes, 10, C.byref(copy_ms)))
    copy_gbs = 2.0 * copy_bytes / (copy_ms.value * 1e-3) / 1e9
    log(f"streaming copy of {copy_bytes >> 20} MiB: {copy_ms.value:.3f} ms, {copy_gbs:.0f} GB/s (read + write)")

    psp_ms = proj.time_psp(50)
    psp_bytes = proj.psp_bytes()
    achieved = psp_bytes / (psp_ms * 1e-3) / 1e9
    dtype = "f64" if args.precision == "double" else "f32"
    kname = KERNEL_OF_FORMAT.get(args.format, args.format)
    tag = f"{kname}_{dtype}_c1_{'lognormal' if args.lognormal else 'fixed'}{args.degree}"
    traffic = traffic_from_profiles(f"{tag}_traffic.txt")
    log(f"step {ms_step:.3f} ms, psp {psp_ms * 1e3:.1f} us, {achieved:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        v, nsteps, secs = cpu_baseline(csr, args.cpu_seconds, threads)
        cpu = {"value": v, "unit": "synaptic events/s", "cores": threads, "kind": "port",
               "sample": f"{nsteps} full steps (SpMV + update) of the same {csr.nnz}-synapse network, "
                         f"{secs:.1f} s, oracle CSR loop FP64 with OpenMP"}
        log(f"cpu baseline {v:.3e} events/s on {threads} threads")

    out = {
        "metric": "synaptic events/s",
        "value": events,
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded random connectivity, random-init rates)",
        "config": {
            "workload": f"rate network N={args.neurons}, in-degree {'lognormal mean ' if args.lognormal else ''}"
                        f"{args.degree}, nnz={csr.nnz}, dt=1ms, {args.format} {dtype}",
            "neurons": args.neurons,
            "synapses": csr.nnz,
            "format": proj.get_format(),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": f"{kname}{' + tile_reduce' if args.format in ('csr_panel', 'sell_panel') else ''}"
                      f" (compute_psp, {args.format})",
            "kernel_ms": psp_ms,
            "algorithmic_bytes": psp_bytes,
        },
        "cpu_baseline": cpu,
        "hbm_gbs_step": psp_bytes / (ms_step * 1e-3) / 1e9,
        "hbm_copy_gbs": copy_gbs,
    }
    print(json.dumps(out))


# ---------------------------------------------------------------------------
# Other workloads (parity configurations and chosen forms); each returns one
# JSON object of the same shape as the default line.
# ---------------------------------------------------------------------------
def _line(metric, value, unit, steps, warmup, ms_step, dtype, config, roofline, cpu, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded)", "config": config,
           "roofline": roofline, "cpu_baseline": cpu}
    out.update(extra)
    return out


def _coba(N, ne, conn_fn, steps_cap, seed=7, poisson=None, weights=(0.6, 6.7), I=20.0,
          precision="double"):
    import annarchy_amd as ah
    from annarchy_amd import models

    net = ah.Network(dt=0.1, seed=seed, precision=precision)
    P = net.create(N, ah.COBA)
    projs = conn_fn(net, P, ne)
    Q = None
    if poisson is not None:
        M, rate, w = poisson
        Q = net.create(M, models.poisson_neuron(rate))
        projs.append(net.connect(Q, P[:M], "exc").one_to_one(w))
    P.v = ah.Uniform(-60.0, -50.0, rng=net.rng)
    mon = net.monitor(P, ["spike"], capacity=steps_cap)
    net.compile(allow_build=False)
    P.I = I
    return net, P, Q, projs, mon


def _events(projs, P, counts, idx):
    """Delivered events = sum over spikes of the out-degree (core/Utils.py:223-239)."""
    out = np.zeros(P.size, dtype=np.int64)
    for pr in projs:
        if pr.pre is P or getattr(pr.pre, "population", None) is P:
            out += np.bincount(pr.csr.col_idx, minlength=P.size)[: P.size]
    return int(out[idx].sum())


def bench_c0(args):
    """COBA LIF, 4000 neurons (3200 E / 800 I), p = 0.02, 10,000 steps."""
    from oracle import oracle as orc

    steps = 10000
    def conn(net, P, ne):
        return [net.connect(P[:ne], P, "exc").fixed_probability(0.02, 0.6),
                net.connect(P[ne:], P, "inh").fixed_probability(0.02, 6.7)]
    net, P, _, projs, mon = _coba(4000, 3200, conn, steps)
    net.set_graph(True, 200)
    v0 = P.v.reshape(-1).copy()
    t0 = time.perf_counter()
    net.run(steps)
    el = time.perf_counter() - t0
    counts, idx = mon.raw_spikes()
    events = _events(projs, P, counts, idx)
    cpu = None
    if not args.no_cpu:
        po = [orc.csrc_projection(projs[0].csr, 0), orc.csrc_projection(projs[1].csr, 1)]
        t1 = time.perf_counter()
        ref = orc.spiking_run(0, 4000, 0.1, v0, steps, po, refractory=np.asarray(P.refractory, np.int32))
        cel = time.perf_counter() - t1
        exact = bool(np.array_equal(ref["counts"], counts) and np.array_equal(ref["idx"], idx))
        cpu = {"value": steps / cel, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"the full 10000-step run, single thread; spike trains bit-exact: {exact}"}
    return _line("steps/s", steps / el, "steps/s", steps, 0, el * 1e3 / steps, "f64",
                 {"workload": "COBA LIF N=4000 (3200E/800I) p=0.02, dt=0.1ms, 10000 steps"},
                 None, cpu, synaptic_events_per_s=events / el, spikes=int(counts.sum()),
                 mean_rate_hz=float(counts.sum() / 4000 / (steps * 1e-4)))


def bench_c2(args):
    """COBA LIF, 400k neurons (320k E / 80k I), fixed_number_pre 1000 (800 E +
    200 I) drawn on the device, 4e8 synapses with per-synapse weights,
    Poisson 5 Hz into 10%.  roofline: the longer of the population update of
    the 400k LIF neurons (60 B per neuron in FP64: v, g_exc, g_inh read and
    written, refractory, refractory_remaining read and written) and the fused
    scatter of the E and I projections (per spike 20 B, per synapse 4 + 3*s
    B: post index, weight, conductance atomic), each timed with HIP events
    inside 20 real steps.  cpu_baseline: the oracle's spiking step on the
    same connectivity (Poisson draws from numpy), a bounded number of steps."""
    from annarchy_amd import connectivity as cn

    steps = args.steps if args.steps != 200 else 2000
    N, ne = 400_000, 320_000
    scale = 80.0 / 1000.0  # same mean input conductance as the 4000-neuron benchmark
    prec = args.precision
    es = 8 if prec == "double" else 4
    we, wi = 0.6 * scale, 6.7 * scale

    def conn(net, P, ne_):
        # one weight array entry per synapse (Uniform(w, w) draws w exactly)
        return [net.connect(P[:ne_], P, "exc").fixed_number_pre(800, cn.Uniform(we, we), generator="device"),
                net.connect(P[ne_:], P, "inh").fixed_number_pre(200, cn.Uniform(wi, wi), generator="device")]
    t_b = time.perf_counter()
    net, P, Q, projs, mon = _coba(N, ne, conn, steps + args.warmup + 40, poisson=(N // 10, 5.0, 0.6),
                                  precision=prec)
    log(f"c2 network ({sum(p.nb_synapses for p in projs)} synapses, connectivity drawn on the device) "
        f"set up in {time.perf_counter() - t_b:.1f} s")
    v0 = P.v.reshape(-1).copy()
    net.set_graph(True, 100)
    net.run(args.warmup)
    t0 = time.perf_counter()
    net.run(steps)
    el = time.perf_counter() - t0
    counts, idx = mon.raw_spikes()
    c_t = counts[args.warmup:args.warmup + steps]
    idx_t = idx[int(counts[:args.warmup].sum()):int(counts[:args.warmup + steps].sum())]
    events = _events(projs[:2], P, c_t, idx_t)
    # kernel times inside 2 x 20 further steps
    n_done = args.warmup + steps
    t_upd = net.time_kernel(3, P.id, 20)
    t_sc = net.time_kernel(5, 0, 20)
    counts2, idx2 = mon.raw_spikes()
    starts = np.concatenate([[0], np.cumsum(counts2)])
    outdeg = np.zeros(N, dtype=np.int64)
    for pr in projs[:2]:
        outdeg += np.bincount(pr.csr.col_idx, minlength=N)[:N]
    # the scatter of step u sends the spikes of step u - 1
    sc_steps = range(n_done + 19, n_done + 39)
    sc_syn = sum(int(outdeg[idx2[starts[u]:starts[u + 1]]].sum()) for u in sc_steps) / 20.0
    sc_spk = sum(int(counts2[u]) for u in sc_steps) / 20.0
    b_sc = sc_syn * (4 + 3 * es) + sc_spk * 20
    b_upd = N * (6 * es + 12)
    kern = {"local_update": {"ms": t_upd, "bytes": b_upd},
            "scatter_multi": {"ms": t_sc, "bytes": b_sc, "syn_events": sc_syn}}
    dom = "local_update" if t_upd >= t_sc else "scatter_multi"
    ach = kern[dom]["bytes"] / (kern[dom]["ms"] * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": None, "kernel": dom}
    cpu = None
    if not args.no_cpu and prec == "double":
        from oracle import oracle as orc
        t1 = time.perf_counter()
        po = [orc.csrc_projection(projs[0].csr, 0), orc.csrc_projection(projs[1].csr, 1),
              orc.csrc_projection(projs[2].csr, 0, pre_is_input=True)]
        log(f"c2 oracle views built in {time.perf_counter() - t1:.1f} s")
        cs = 4000
        unif = np.random.default_rng(5).random((cs, N // 10))
        t1 = time.perf_counter()
        ref = orc.spiking_run(0, N, 0.1, v0, cs, po, refractory=np.asarray(P.refractory, np.int32),
                              poisson=dict(M=N // 10, rates=5.0, uniforms=unif), spike_cap=N * 50)
        cel = time.perf_counter() - t1
        ev_c = int(outdeg[ref["idx"]].sum())
        cpu = {"value": ev_c / cel, "unit": "synaptic events/s", "cores": 1, "kind": "port",
               "sample": f"{cs} steps of the same network from the same initial state (Poisson draws from "
                         f"numpy), single thread: {cel:.1f} s, {cel / cs * 1e3:.1f} ms/step"}
    return _line("synaptic events/s", events / el, "synaptic events/s", steps, args.warmup,
                 el * 1e3 / steps, "f64" if prec == "double" else "f32",
                 {"workload": "COBA LIF N=400000 fixed_number_pre 1000 (800E/200I) 4e8 synapses, "
                              "Poisson 5 Hz into 10%, dt=0.1ms", "precision": prec,
                  "connectivity": "drawn on the device (Philox stream)"},
                 roof, cpu, kernels=kern, spikes_per_step=float(c_t.mean()),
                 mean_rate_hz=float(c_t.sum() / N / (steps * 1e-4)))


def bench_stdp(args):
    """STDP (models/Synapses.py:232-310) on 5e7 plastic synapses: 100,000
    Poisson pre-synaptic neurons at 20 Hz -> 50,000 post-synaptic Poisson
    neurons at 10 Hz (their spikes do not depend on g), fixed_number_pre
    1000 drawn on the device, weights U(0.2, 0.8), dt = 1 ms.  metric =
    synaptic events/s: pre spikes x out-degree + post spikes x in-degree
    (core/Utils.py:223-239).  roofline: the pre-event and post-event kernels
    timed with HIP events inside 20 further steps each; algorithmic bytes per synapse
    event: pre 20 + 8*s (row index, w, x, y, _last_event read and written,
    the conductance atomic), post 24 + 6*s (state position, w, x, y,
    _last_event), s = bytes per value.  cpu_baseline: the oracle's plastic
    projection (orc_plastic_run) over the warm-up steps' recorded spike
    trains, single thread; its weights are compared with the GPU's."""
    import annarchy_amd as ah
    from annarchy_amd import connectivity as cn

    steps = args.steps if args.steps != 200 else 1000
    warmup = args.warmup if args.warmup != 20 else 200
    n_pre, n_post, k = 100_000, 50_000, 1000
    prec = args.precision
    es = 8 if prec == "double" else 4
    pre_n = ah.Neuron(parameters="rates = 40.0", equations="p = Uniform(0.0, 1.0) * 1000.0 / dt",
                      spike="p < rates", name="PoissonPre")
    post_n = ah.Neuron(parameters="rates = 30.0", equations="p = Uniform(0.0, 1.0) * 1000.0 / dt\nmp = g_exc",
                       spike="p < rates", name="PoissonPostG")
    stdp = ah.STDP(tau_plus=20.0, tau_minus=20.0, A_plus=0.01, A_minus=0.0105, w_min=0.0, w_max=1.0)
    t_b = time.perf_counter()
    net = ah.Network(dt=1.0, seed=17, precision=prec)
    pre = net.create(n_pre, pre_n)
    post = net.create(n_post, post_n)
    proj = net.connect(pre, post, "exc", synapse=stdp).fixed_number_pre(k, cn.Uniform(0.2, 0.8),
                                                                         generator="device")
    m_pre = net.monitor(pre, ["spike"], capacity=warmup + steps + 40)
    m_post = net.monitor(post, ["spike"], capacity=warmup + steps + 40)
    net.compile(allow_build=False)
    pre.rates = 20.0
    post.rates = 10.0
    net.set_graph(True, 50)
    log(f"stdp network ({proj.nb_synapses} plastic synapses) set up in {time.perf_counter() - t_b:.1f} s")
    w0 = proj.w.copy()
    net.run(warmup)
    w_warm = proj.w.copy()
    t0 = time.perf_counter()
    net.run(steps)
    el = time.perf_counter() - t0
    pc, pi = m_pre.raw_spikes()
    qc, qi = m_post.raw_spikes()
    pc, qc = pc[:warmup + steps], qc[:warmup + steps]
    pi, qi = pi[:int(pc.sum())], qi[:int(qc.sum())]
    csr = proj.csr
    outdeg = np.bincount(csr.col_idx, minlength=n_pre)
    n_pw = int(pc[:warmup].sum())
    ev_pre = int(outdeg[pi[n_pw:]].sum())
    ev_post = int(qc[warmup:].sum()) * k
    events = ev_pre + ev_post
    log(f"stdp: {steps} steps in {el:.3f} s, {events / steps:.3e} synapse events/step")
    # kernel times inside 2 x 20 further steps (HIP events around the kernel
    # in each step); algorithmic bytes of exactly those steps' events
    n_done = warmup + steps
    t_pre = proj.time_events("pre", 20)
    t_post = proj.time_events("post", 20)
    pc2, pi2 = m_pre.raw_spikes()
    qc2, qi2 = m_post.raw_spikes()
    starts = np.concatenate([[0], np.cumsum(pc2)])
    # pre events of step u handle the pre spikes of step u - 1
    pre_steps = range(n_done - 1, n_done + 19)
    syn_pre = sum(int(outdeg[pi2[starts[u]:starts[u + 1]]].sum()) for u in pre_steps) / 20.0
    spk_pre = sum(int(pc2[u]) for u in pre_steps) / 20.0
    spk_post = float(qc2[n_done + 20:n_done + 40].mean())
    syn_post = spk_post * k
    b_pre = syn_pre * (20 + 8 * es) + spk_pre * 20
    b_post = syn_post * (24 + 6 * es) + spk_post * 20
    # streaming-copy probe: the achievable HBM rate and the counter calibration
    import ctypes as C
    from annarchy_amd import _lib
    copy_ms = C.c_double()
    _lib.check(_lib.lib().ahip_probe_copy(1 << 31, 10, C.byref(copy_ms)))
    copy_gbs = 2.0 * (1 << 31) / (copy_ms.value * 1e-3) / 1e9
    kern = {"pre_event": {"ms": t_pre, "bytes": b_pre, "syn_events": syn_pre},
            "post_event": {"ms": t_post, "bytes": b_post, "syn_events": syn_post}}
    dom = "pre_event" if t_pre >= t_post else "post_event"
    ach = kern[dom]["bytes"] / (kern[dom]["ms"] * 1e-3) / 1e9
    traffic = traffic_from_profiles("stdp_events_f64_5e7syn_traffic.txt") if (prec == "double" and
                                                                                dom == "pre_event") else None
    roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": traffic, "kernel": dom}
    cpu, parity = None, None
    if not args.no_cpu and prec == "double":
        from oracle import oracle as orc
        n_qw = int(qc[:warmup].sum())
        t1 = time.perf_counter()
        ref = orc.plastic_run(0, csr.row_ptr, csr.col_idx, n_pre, w0, [20.0, 20.0, 0.01, 0.0105, 0.0, 1.0], 1.0,
                              pc[:warmup], pi[:n_pw], qc[:warmup], qi[:n_qw])
        cel = time.perf_counter() - t1
        ev_w = int(outdeg[pi[:n_pw]].sum()) + n_qw * k
        parity = float(np.max(np.abs(ref["w"] - w_warm)))
        cpu = {"value": ev_w / cel, "unit": "synaptic events/s", "cores": 1, "kind": "port",
               "sample": f"the {warmup} warm-up steps ({ev_w:.3e} synapse events) of the same network, "
                         f"single thread; max |w_gpu - w_oracle| = {parity:.2e}"}
    return _line("synaptic events/s", events / el, "synaptic events/s", steps, warmup, el * 1e3 / steps,
                 "f64" if prec == "double" else "f32",
                 {"workload": "STDP 100000 Poisson(20 Hz) -> 50000 Poisson(10 Hz), fixed_number_pre 1000 "
                              "(5e7 plastic synapses), dt=1ms", "precision": prec},
                 roof, cpu, kernels=kern, w_max_abs_diff_vs_oracle=parity, hbm_copy_gbs=copy_gbs,
                 pre_spikes_per_step=float(pc[warmup:].mean()), post_spikes_per_step=float(qc[warmup:].mean()))


def bench_c3(args):
    """Batched sweep: 256 instances, N_i ~ U{500..8000}, recurrent p=0.05,
    shared dense 2000 x 2000 input, gains linspace(0.1, 2.0)."""
    from annarchy_amd.batched import RateSweep, time_plain

    steps = args.steps if args.steps != 200 else 2000
    rng = np.random.default_rng(2026)
    sizes = rng.integers(500, 8001, 256)
    t_b = time.perf_counter()
    sw = RateSweep(sizes, n_in=2000, p=0.05, seed=3)
    log(f"c3 sweep: {sw.total} neurons, {sw.nnz} synapses, built in {time.perf_counter() - t_b:.0f} s")
    b = sw.batched(lanes=16)
    b.run(20)
    t0 = time.perf_counter()
    b.run(steps)
    el_b = time.perf_counter() - t0
    # streaming-copy probe: the achievable HBM rate, and the calibration of the
    # HBM byte counters of a profiled run
    import ctypes as C
    from annarchy_amd import _lib
    copy_ms = C.c_double()
    _lib.check(_lib.lib().ahip_probe_copy(1 << 31, 10, C.byref(copy_ms)))
    copy_gbs = 2.0 * (1 << 31) / (copy_ms.value * 1e-3) / 1e9
    info = b.info()
    ms_spmv = b.time_kernel(2, 20)
    ms_gemm = b.time_kernel(1, 50)
    log(f"batched {el_b * 1e3 / steps:.3f} ms/step, spmv {ms_spmv * 1e3:.0f} us, gemm {ms_gemm * 1e3:.1f} us")
    # plain per-instance FP64 loop, same initial state, same number of steps
    nets = sw.plain_networks()
    plain_steps = min(steps, 200)
    el_p = time_plain(nets, plain_steps)
    # error of the batched form against the FP64 plain form after `steps`
    for net, exc in nets:
        net.run(20 + steps - plain_steps)
    got = b.rates()
    worst, pop_err = 0.0, 0.0
    for i, (net, exc) in enumerate(nets):
        ref = exc.r.reshape(-1)
        worst = max(worst, float(np.max(np.abs(got[i] - ref)) / max(np.max(np.abs(ref)), 1e-3)))
        pop_err = max(pop_err, abs(float(got[i].mean() - ref.mean())) / max(abs(ref.mean()), 1e-9))
    sps_b = steps / el_b
    sps_p = plain_steps / el_p
    achieved = info["sparse_bytes"] / (ms_spmv * 1e-3) / 1e9
    return _line("steps/s (batched sweep)", sps_b, "steps/s", steps, 20, el_b * 1e3 / steps, "f32/bf16",
                 {"workload": "256 rate instances N~U{500..8000} p=0.05 + shared dense 2000x2000 input",
                  "neurons": sw.total, "synapses": sw.nnz},
                 {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS,
                  "traffic": traffic_from_profiles("batch_sell_f32_c3_sweep256_traffic.txt") if info["mode"] == 2
                  else None,
                  "kernel": {0: "batch_spmv_update", 1: "batch_spmv_lds", 2: "batch_sell"}[info["mode"]],
                  "kernel_ms": ms_spmv},
                 None, plain_steps_per_s=sps_p, speedup_vs_plain=sps_b / sps_p,
                 max_rel_rate_error_vs_fp64=worst, max_rel_population_rate_error=pop_err,
                 gemm={"ms": ms_gemm, "tflops": 3 * info["dense_flops"] / (ms_gemm * 1e-3) / 1e12,
                       "peak_bf16_tflops": 2500.0}, hbm_copy_gbs=copy_gbs)


if __name__ == "__main__":
    main()
