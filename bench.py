"""Benchmark of the simulation step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

prints ONE JSON line.  Its top level is the headline workload: the rate
network of 100,000 neurons with one recurrent projection of fixed in-degree
1000 (1e8 synapses, weights N(0, 1/sqrt(1000)), int32 indices),
tau dr/dt = -r + sum(exc) + N(0, 0.1), r >= 0, dt = 1 ms, FP64, sliced-tile
layout.  A step = compute_psp (SpMV) + the neuron update with on-device noise.
W warm-up steps, then K steps timed; the timed steps replay a HIP graph
captured, instantiated and uploaded before the timer starts (graph batch =
the largest divisor of K up to 200).  metric = synaptic events/s = nnz per
step / step time.

Under "configs" the line holds one record of the same shape per further
workload, each run in a child process of its own (this process never touches
the GPU):
  c1_f32, c1_lognormal  the headline network in FP32 (with its psp error
      against the float64 product) and with lognormal in-degrees;
  c0  COBA LIF, 4000 neurons, p = 0.02, 10,000 steps (parity run, bit-exact
      spike trains against the oracle);
  c2  COBA LIF, 400,000 neurons, fixed_number_pre 800 E + 200 I drawn on the
      device, 4e8 synapses with per-synapse weights, Poisson 5 Hz into 10%;
  c2_f32  the same in FP32, its population rate and ISI CV against FP64;
  stdp  STDP on 5e7 plastic synapses (100,000 Poisson pre neurons at 20 Hz
      -> 50,000 Poisson post neurons at 10 Hz, fixed_number_pre 1000);
  c3  256 rate-network instances of uneven size in one batched launch, plain
      per-instance FP64 loop beside it;
  c4  the learned storage-format selector: held-out regret against
      exhaustive tuning, timed live on a sample of held-out matrices.

roofline: the dominant kernel's algorithmic bytes (or flop) per launch over
its mean launch time from HIP events on the network stream, against the
MI355X peaks (HBM 8 TB/s; dense BF16 MFMA 2.5 PFLOP/s).
cpu_baseline: the oracle (a C++/OpenMP restatement of the reference's
generated step) timed on this host's cores, a bounded sample of the same
workload; the record states the threads used, the host CPU model, nproc and
the CPUs this process may use.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
SUB_CONFIGS = ("c1_f32", "c1_lognormal", "c0", "c2", "c2_f32", "stdp", "c3", "c4")
CHILD_TIMEOUT_S = {"c1": 240, "c1_f32": 180, "c1_lognormal": 180, "c0": 120, "c2": 240, "c2_f32": 240,
                   "stdp": 240, "c3": 300, "c4": 300}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def host_cpu() -> dict:
    """nproc, the CPUs this process may run on, the thread share it uses
    (OMP_NUM_THREADS when set, as on a one-GPU box's CPU share) and the model."""
    info = {"nproc": os.cpu_count() or 1, "affinity": len(os.sched_getaffinity(0))}
    share = os.environ.get("OMP_NUM_THREADS", "")
    info["share"] = int(share) if share.isdigit() and int(share) > 0 else info["affinity"]
    info["threads"] = max(1, min(info["affinity"], info["share"]))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    info["model"] = model
    return info


def graph_batch_for(steps: int, cap: int = 200) -> int:
    """Largest divisor of `steps` up to `cap` (every timed step is a replay)."""
    for b in range(min(cap, max(steps, 1)), 0, -1):
        if steps % b == 0:
            return b
    return 1


def traffic_from_profiles(name):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def roofline(bound, achieved, peak, unit, traffic, kernel, **extra):
    out = {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
           "traffic": traffic, "kernel": kernel}
    out.update(extra)
    return out


def record(metric, value, unit, steps, warmup, ms_step, dtype, config, roof, cpu, higher=True, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms_step, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype, "data": "synthetic (seeded)", "config": config, "roofline": roof,
           "cpu_baseline": cpu}
    out.update(extra)
    return out


def copy_probe():
    """Streaming copy of 2 GiB: the achievable HBM rate on this box."""
    import ctypes as C
    from annarchy_amd import _lib
    ms = C.c_double()
    _lib.check(_lib.lib().ahip_probe_copy(1 << 31, 10, C.byref(ms)))
    return 2.0 * (1 << 31) / (ms.value * 1e-3) / 1e9


# ---------------------------------------------------------------------------
# C1: the rate network (headline)
# ---------------------------------------------------------------------------
KERNEL_OF_FORMAT = {"csr": "csr_rows", "sell": "sell_slices", "csr_panel": "tile_psp", "csr_vec": "csr_vec",
                    "ell": "ell_rows", "ellr": "ell_rows", "bsr": "bsr_rows", "dense": "dense_rows",
                    "tiles": "tile_psp", "sell_panel": "sell_panel_psp"}


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


def c1_cpu_baseline(csr, seconds, threads):
    """Steps of the oracle's CSR loop (the reference's default scalar row
    loop) + rate update on `threads` OpenMP threads, for about `seconds`."""
    from oracle import oracle as orc

    n = csr.n_rows
    r = np.random.default_rng(1).random(n)
    noise = np.random.default_rng(2).normal(0.0, 0.1, (1, n))
    orc.rate_run(csr.row_ptr, csr.col_idx, csr.w, r, 10.0, 1.0, noise, 1, nthreads=threads)
    steps, t0 = 0, time.perf_counter()
    while True:
        orc.rate_run(csr.row_ptr, csr.col_idx, csr.w, r, 10.0, 1.0, noise, 1, nthreads=threads)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 1000:
            break
    return csr.nnz * steps / el, steps, el


def bench_c1(args):
    import annarchy_amd as ah

    t_build = time.perf_counter()
    csr = build_c1(args.neurons, args.degree, 1234, args.lognormal)
    log(f"c1: connectivity {csr.nnz} synapses in {time.perf_counter() - t_build:.1f} s")
    cpu = None
    if not args.no_cpu:
        h = host_cpu()
        v1, s1, e1 = c1_cpu_baseline(csr, args.cpu_seconds, 1)
        log(f"c1: cpu baseline 1 thread {v1:.3e} events/s")
        vn, sn, en = c1_cpu_baseline(csr, args.cpu_seconds, h["threads"])
        log(f"c1: cpu baseline {h['threads']} threads {vn:.3e} events/s")
        cpu = {"value": vn, "unit": "synaptic events/s", "cores": h["threads"], "kind": "port",
               "sample": f"{sn} full steps (SpMV + update) of the same {csr.nnz}-synapse network in "
                         f"{en:.1f} s, oracle CSR loop FP64, OpenMP on {h['threads']} threads (the CPU share of "
                         f"this process; nproc {h['nproc']}, affinity {h['affinity']}, {h['model']})",
               "one_thread": {"value": v1, "steps": s1, "seconds": e1},
               "host": h}

    net = ah.Network(dt=1.0, seed=42, precision=args.precision)
    pop = net.create(args.neurons, ah.RateNoisy)
    proj = net.connect(pop, pop, "exc")
    # chunk 0: the builder sizes the tiles to the device's workgroup slots
    # (FP64 1e5 rows: 1984-row tiles, 255 workgroups; FP32: 1024 rows)
    chunk = args.chunk
    proj.from_csr(csr, storage_format=args.format, lanes=args.lanes,
                  sigma=args.waves if args.format == "sell_panel" else args.sigma,
                  panel_cols=args.panel_cols, chunk=chunk)
    net.compile(allow_build=False)
    pop.r = np.random.default_rng(7).random(args.neurons)
    G = graph_batch_for(args.steps)
    net.set_graph(True, G)
    net.prepare_graph()
    log(f"c1: network uploaded, graph of {G} steps ready")

    net.run(args.warmup)
    t0 = time.perf_counter()
    net.run(args.steps)  # ahip_run synchronizes the stream before returning
    el = time.perf_counter() - t0
    ms_step = el * 1000.0 / args.steps
    events = csr.nnz * args.steps / el

    # FP32: the psp of one more step against the float64 product of the same
    # rates (the FP32 weights widened), as a relative max-norm error
    psp_err = None
    if args.precision == "float":
        import scipy.sparse as sp
        r_in = np.asarray(pop.r, dtype=np.float64).reshape(-1)
        net.run(1)
        got = np.asarray(pop.sum("exc"), dtype=np.float64).reshape(-1)
        w32 = np.asarray(csr.w, dtype=np.float32).astype(np.float64)
        ref = sp.csr_matrix((w32, csr.col_idx, csr.row_ptr), shape=(args.neurons, args.neurons)) @ r_in
        psp_err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        log(f"c1: FP32 psp error against the float64 product {psp_err:.2e}")
    copy_gbs = copy_probe()
    psp_ms = proj.time_psp(50)
    psp_bytes = proj.psp_bytes()
    achieved = psp_bytes / (psp_ms * 1e-3) / 1e9
    dtype = "f64" if args.precision == "double" else "f32"
    kname = KERNEL_OF_FORMAT.get(args.format, args.format)
    tag = f"{kname}_{dtype}_c1_{'lognormal' if args.lognormal else 'fixed'}{args.degree}"
    log(f"c1: step {ms_step:.3f} ms, psp {psp_ms * 1e3:.1f} us, {achieved:.0f} GB/s; copy {copy_gbs:.0f} GB/s")
    roof = roofline("hbm", achieved, HBM_PEAK_GBS, "GB/s", traffic_from_profiles(f"{tag}_traffic.txt"),
                    f"{kname}{' + tile_reduce' if args.format in ('csr_panel', 'sell_panel') else ''}"
                    f" (compute_psp, {args.format})", kernel_ms=psp_ms, algorithmic_bytes=psp_bytes)
    return record("synaptic events/s", events, "synaptic events/s", args.steps, args.warmup, ms_step, dtype,
                  {"workload": f"rate network N={args.neurons}, in-degree "
                               f"{'lognormal mean ' if args.lognormal else ''}{args.degree}, nnz={csr.nnz}, "
                               f"dt=1ms, {args.format} {dtype}",
                   "neurons": args.neurons, "synapses": csr.nnz, "format": proj.get_format(),
                   "graph_steps": G, "parallelism": "single GPU"},
                  roof, cpu, hbm_gbs_step=psp_bytes / (ms_step * 1e-3) / 1e9, hbm_copy_gbs=copy_gbs,
                  psp_max_rel_error_vs_fp64=psp_err)


def bench_c1_f32(args):
    """The headline network in FP32 (Kahan-compensated sums, 1024-row tiles),
    with the psp's error against the float64 product stated."""
    args.precision = "float"
    args.cpu_seconds = min(args.cpu_seconds, 3.0)
    return bench_c1(args)


def bench_c1_lognormal(args):
    """The headline network with lognormal in-degrees (mean 1000): uneven rows."""
    args.lognormal = True
    args.cpu_seconds = min(args.cpu_seconds, 3.0)
    return bench_c1(args)


# ---------------------------------------------------------------------------
# spiking workloads
# ---------------------------------------------------------------------------
def _coba(N, ne, conn_fn, steps_cap, seed=7, poisson=None, I=20.0, precision="double"):
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


def _outdeg(projs, P):
    out = np.zeros(P.size, dtype=np.int64)
    for pr in projs:
        if pr.pre is P or getattr(pr.pre, "population", None) is P:
            out += np.bincount(pr.csr.col_idx, minlength=P.size)[: P.size]
    return out


def _events(projs, P, idx):
    """Delivered events = sum over spikes of the out-degree (core/Utils.py:223-239)."""
    return int(_outdeg(projs, P)[idx].sum())


def _spiking_kernels(net, P, projs, outdeg, mon, es, n_steps_done):
    """Mean time of each kernel of a spiking step (HIP events inside 20 real
    eager steps per kernel) and its algorithmic bytes over exactly those
    steps.  Kernel bytes:
      update      N * (6 s + 12): v, g_exc, g_inh read and written, refractory
                  read, refractory_remaining read and written; when it
                  delivers the ordered conductances also N * 8 per
                  conductance (packed bucket heads), per event 4 + s (key,
                  weight) and per touched neuron 20 (bucket offset, head
                  cleared);
      scatter     per spike 20 B (list entry, column range), per synapse event
                  4 + 3 s (post rank, weight, conductance read and written) with
                  atomic delivery or one weight, and 4 + 16 + 4 + 8 (post rank,
                  packed bucket head read and written, key, weight) ordered;
      deliver     per touched post neuron 4 + 4 + 16 + 2 s (list entry, counter,
                  bucket range, g read and written) + per event 4 + s (key,
                  weight).
    """
    from annarchy_amd import _lib

    t = {}
    t["update"] = net.time_kernel(3, P.id, 20)
    t["compact"] = net.time_kernel(4, P.id, 20)
    ng = net.scatter_groups()
    t["scatter"] = sum(net.time_kernel(5, g, 20) for g in range(ng))
    no = net.ordered_targets()
    # ordered conductances are delivered by the post population's update
    # unless another kernel writes them: a separate launch only then
    sep = 0
    if no:
        try:
            t["deliver"] = net.time_kernel(6, 0, 20)
            sep = 1
        except _lib.AhipError:
            t["deliver"] = 0.0
    else:
        t["deliver"] = 0.0
    counts, idx = mon.raw_spikes()
    starts = np.concatenate([[0], np.cumsum(counts)])
    # the kernels above ran 20 + 20 + 20 ng + 20 (one delivery launch) steps
    # after n_steps_done; the scatter of step u sends the spikes of step
    # u - 1: average over the steps of the scatter and deliver probes
    first = n_steps_done + 40
    last = n_steps_done + 40 + 20 * (ng + sep)
    us = range(first - 1, min(last - 1, len(counts)))
    syn = np.mean([outdeg[idx[starts[u]:starts[u + 1]]].sum() for u in us]) if len(us) else 0.0
    spk = np.mean([counts[u] for u in us]) if len(us) else 0.0
    N = P.size
    per_syn = (4 + 16 + 4 + 8) if no else (4 + 3 * es)
    # touched post neurons: the expected number of distinct targets of `syn`
    # events spread over N neurons, N (1 - exp(-syn / N))
    touched = N * -np.expm1(-syn / N)
    # update: + per ordered conductance N * 8 (packed heads) when it
    # delivers, and per event 4 + s (key, weight), per touched neuron 20
    # (bucket offset, head cleared)
    deliver_in_update = (no * N * 8 + syn * (4 + es) + touched * 20) if (no and not sep) else 0.0
    b = {"update": N * (6 * es + 12) + deliver_in_update, "compact": N / 8 + spk * 4 + 8,
         "scatter": syn * per_syn + spk * 20,
         "deliver": (syn * (4 + es) + touched * (24 + 2 * es)) if sep else 0.0}
    kern = {k: {"ms": t[k], "bytes": float(b[k])} for k in t if t[k] > 0}
    return kern, float(syn), float(spk)


def bench_c0(args):
    """COBA LIF, 4000 neurons (3200 E / 800 I), p = 0.02, 10,000 steps: one
    graph-replayed parity run from the seeded initial state."""
    from oracle import oracle as orc

    steps = 10000
    h = host_cpu()

    def conn(net, P, ne):
        return [net.connect(P[:ne], P, "exc").fixed_probability(0.02, 0.6),
                net.connect(P[ne:], P, "inh").fixed_probability(0.02, 6.7)]
    G, probe_launches = 200, 5
    net, P, _, projs, mon = _coba(4000, 3200, conn, steps + G * probe_launches + 400)
    net.set_graph(True, G)
    net.prepare_graph()
    persistent = net.uses_persistent()
    v0 = P.v.reshape(-1).copy()
    t0 = time.perf_counter()
    net.run(steps)
    el = time.perf_counter() - t0
    log(f"c0: {steps} steps in {el:.3f} s ({'persistent steps' if persistent else 'per-kernel steps'})")
    outdeg = _outdeg(projs, P)
    es = 8
    done = steps
    kern = {}
    if persistent:
        # one launch runs G steps in one workgroup: state in and out of LDS
        # once, N (3 s + 4) each way; per step the spike list and monitor
        # entries (4 + 4 B per spike), last_spike (8 B), the column range
        # (16 B) per spike and row index (4 B) per event
        ms = net.time_kernel(7, 0, probe_launches)
        counts, idx = mon.raw_spikes()
        starts = np.concatenate([[0], np.cumsum(counts)])
        us = range(done - 1, done + G * probe_launches - 1)
        syn = float(np.mean([outdeg[idx[starts[u]:starts[u + 1]]].sum() for u in us]))
        spk = float(np.mean([counts[u] for u in us]))
        b = 2 * 4000 * (3 * es + 4) + G * (spk * (4 + 4 + 8 + 16) + syn * 4)
        kern["persistent"] = {"ms": ms, "bytes": float(b), "steps_per_launch": G, "us_per_step": ms * 1e3 / G}
        done += G * probe_launches
        log(f"c0: persistent launch {ms:.3f} ms / {G} steps = {ms * 1e3 / G:.2f} us per step")
    per_k, syn, spk = _spiking_kernels(net, P, projs, outdeg, mon, es, done)
    if not persistent:
        kern = per_k
    counts, idx = mon.raw_spikes()
    events = _events(projs, P, idx[:int(counts[:steps].sum())])
    dom = max(kern, key=lambda k: kern[k]["ms"])
    ach = kern[dom]["bytes"] / (kern[dom]["ms"] * 1e-3) / 1e9
    note = ("latency-bound: one workgroup runs 200 dependent steps per launch (scatter, update, spike list "
            "with barriers between them); the state stays in LDS" if persistent else
            "latency-bound: three launches of a few microseconds per step")
    roof = roofline("hbm", ach, HBM_PEAK_GBS, "GB/s", None, dom, kernel_ms=kern[dom]["ms"],
                    algorithmic_bytes=kern[dom]["bytes"], note=note)
    cpu = None
    if not args.no_cpu:
        po = [orc.csrc_projection(projs[0].csr, 0), orc.csrc_projection(projs[1].csr, 1)]
        times = []
        for _ in range(3):
            t1 = time.perf_counter()
            ref = orc.spiking_run(0, 4000, 0.1, v0, steps, po, refractory=np.asarray(P.refractory, np.int32))
            times.append(time.perf_counter() - t1)
        cel = float(np.median(times))
        exact = bool(np.array_equal(ref["counts"], counts[:steps]) and
                     np.array_equal(ref["idx"], idx[:int(counts[:steps].sum())]))
        cpu = {"value": steps / cel, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"the full {steps}-step run, single thread (the oracle's spiking step is the "
                         f"reference's single-thread loop), median of 3; spike trains bit-exact: {exact}",
               "host": h}
        log(f"c0: oracle {steps / cel:.0f} steps/s, bit-exact {exact}")
    return record("steps/s", steps / el, "steps/s", steps, 0, el * 1e3 / steps, "f64",
                  {"workload": "COBA LIF N=4000 (3200E/800I) p=0.02, dt=0.1ms, 10000 steps", "graph_steps": G,
                   "persistent_steps": persistent},
                  roof, cpu, synaptic_events_per_s=events / el, spikes=int(counts[:steps].sum()),
                  mean_rate_hz=float(counts[:steps].sum() / 4000 / (steps * 1e-4)), kernels=kern,
                  per_kernel_step_kernels=per_k if persistent else None)


def bench_c2(args):
    """COBA LIF, 400k neurons (320k E / 80k I), fixed_number_pre 1000 (800 E +
    200 I) drawn on the device, 4e8 synapses with per-synapse weights drawn
    from Uniform(0.8 w, 1.2 w), Poisson 5 Hz into 10%, ordered delivery."""
    from annarchy_amd import connectivity as cn

    steps = args.c2_steps
    warmup = 20
    N, ne = 400_000, 320_000
    scale = 80.0 / 1000.0  # same mean input conductance as the 4000-neuron network
    prec = args.precision
    es = 8 if prec == "double" else 4
    we, wi = 0.6 * scale, 6.7 * scale

    def conn(net, P, ne_):
        return [net.connect(P[:ne_], P, "exc").fixed_number_pre(800, cn.Uniform(0.8 * we, 1.2 * we),
                                                                generator="device"),
                net.connect(P[ne_:], P, "inh").fixed_number_pre(200, cn.Uniform(0.8 * wi, 1.2 * wi),
                                                                generator="device")]
    t_b = time.perf_counter()
    net, P, Q, projs, mon = _coba(N, ne, conn, steps + warmup + 200, poisson=(N // 10, 5.0, 0.6), precision=prec)
    log(f"c2: network ({sum(p.nb_synapses for p in projs)} synapses drawn on the device) set up in "
        f"{time.perf_counter() - t_b:.1f} s")
    v0 = P.v.reshape(-1).copy()
    G = graph_batch_for(steps, 100)
    net.set_graph(True, G)
    net.prepare_graph()
    net.run(warmup)
    t0 = time.perf_counter()
    net.run(steps)
    el = time.perf_counter() - t0
    counts, idx = mon.raw_spikes()
    c_t = counts[warmup:warmup + steps]
    idx_t = idx[int(counts[:warmup].sum()):int(counts[:warmup + steps].sum())]
    outdeg = _outdeg(projs[:2], P)
    events = int(outdeg[idx_t].sum())
    log(f"c2: {steps} steps in {el:.3f} s ({el / steps * 1e6:.1f} us/step), {events / el:.3e} events/s")
    kern, syn, spk = _spiking_kernels(net, P, projs[:2], outdeg, mon, es, warmup + steps)
    # every step the monitor recorded (warm-up, timed and probe steps): the
    # mean spikes and synaptic events per step, one scatter launch each
    c_all, i_all = mon.raw_spikes()
    all_steps = {"steps": int(len(c_all)), "spikes_per_step": float(c_all.mean()) if len(c_all) else 0.0,
                 "events_per_step": float(outdeg[i_all].sum()) / max(1, len(c_all))}
    dom = max(kern, key=lambda k: kern[k]["ms"])
    ach = kern[dom]["bytes"] / (kern[dom]["ms"] * 1e-3) / 1e9
    kname = {"scatter": "scatter_multi", "update": "local_update", "compact": "compact_spikes",
             "deliver": "deliver_ordered"}[dom]
    traffic = traffic_from_profiles(f"{kname}_{'f64' if prec == 'double' else 'f32'}_c2_coba400k_traffic.txt")
    roof = roofline("hbm", ach, HBM_PEAK_GBS, "GB/s", traffic, dom, kernel_ms=kern[dom]["ms"],
                    algorithmic_bytes=kern[dom]["bytes"])
    fp32_vs_fp64 = None
    if prec == "float":
        # the same seeded network in FP64 (connectivity, initial state and
        # Poisson draws are the same streams): population rate and mean ISI
        # CV over the same steps
        rate32, cv32 = _rate_and_cv(c_t, idx_t, N, 0.1)
        del net, P, Q, mon
        net64, P64, _, _, mon64 = _coba(N, ne, conn, steps + warmup + 200, poisson=(N // 10, 5.0, 0.6),
                                        precision="double")
        net64.set_graph(True, G)
        net64.run(warmup + steps)
        c64, i64 = mon64.raw_spikes()
        rate64, cv64 = _rate_and_cv(c64[warmup:warmup + steps],
                                    i64[int(c64[:warmup].sum()):int(c64[:warmup + steps].sum())], N, 0.1)
        fp32_vs_fp64 = {"rate_hz": [rate32, rate64], "rate_rel_diff": abs(rate32 - rate64) / rate64,
                        "isi_cv": [cv32, cv64], "isi_cv_rel_diff": abs(cv32 - cv64) / cv64}
        log(f"c2: FP32 rate {rate32:.3f} Hz / FP64 {rate64:.3f} Hz, ISI CV {cv32:.4f} / {cv64:.4f}")
        del net64, P64, mon64
    cpu = None
    if not args.no_cpu and prec == "double":
        from oracle import oracle as orc
        t1 = time.perf_counter()
        po = [orc.csrc_projection(projs[0].csr, 0), orc.csrc_projection(projs[1].csr, 1),
              orc.csrc_projection(projs[2].csr, 0, pre_is_input=True)]
        log(f"c2: oracle views built in {time.perf_counter() - t1:.1f} s")
        cs = 2000
        unif = np.random.default_rng(5).random((cs, N // 10))
        t1 = time.perf_counter()
        ref = orc.spiking_run(0, N, 0.1, v0, cs, po, refractory=np.asarray(P.refractory, np.int32),
                              poisson=dict(M=N // 10, rates=5.0, uniforms=unif), spike_cap=N * 50)
        cel = time.perf_counter() - t1
        ev_c = int(outdeg[ref["idx"]].sum())
        cpu = {"value": ev_c / cel, "unit": "synaptic events/s", "cores": 1, "kind": "port",
               "sample": f"{cs} steps of the same network from the same initial state (Poisson draws from numpy), "
                         f"single thread: {cel:.1f} s, {cel / cs * 1e3:.1f} ms/step", "host": host_cpu()}
        log(f"c2: oracle {ev_c / cel:.3e} events/s")
    return record("synaptic events/s", events / el, "synaptic events/s", steps, warmup, el * 1e3 / steps,
                  "f64" if prec == "double" else "f32",
                  {"workload": "COBA LIF N=400000 fixed_number_pre 1000 (800E/200I) 4e8 synapses with per-synapse "
                               "weights, Poisson 5 Hz into 10%, dt=0.1ms", "precision": prec,
                   "delivery": "ordered (the reference's summation order)", "graph_steps": G,
                   "connectivity": "drawn on the device (Philox stream)"},
                  roof, cpu, kernels=kern, all_recorded_steps=all_steps, spikes_per_step=float(c_t.mean()),
                  mean_rate_hz=float(c_t.sum() / N / (steps * 1e-4)), fp32_vs_fp64=fp32_vs_fp64)


def bench_c2_f32(args):
    """C2 in FP32, with its population rate and mean ISI CV against an FP64
    run of the same seeded network."""
    args.precision = "float"
    return bench_c2(args)


def _rate_and_cv(counts, idx, n, dt_ms):
    """Population rate (Hz) and mean ISI coefficient of variation over the
    neurons with at least two intervals."""
    counts = np.asarray(counts, dtype=np.int64)
    idx = np.asarray(idx, dtype=np.int64)
    rate = float(counts.sum()) / n / (len(counts) * dt_ms * 1e-3)
    steps = np.repeat(np.arange(len(counts)), counts)
    order = np.lexsort((steps, idx))
    nn, tt = idx[order], steps[order]
    same = np.diff(nn) == 0
    isi = np.diff(tt)[same].astype(np.float64)
    who = nn[1:][same]
    cnt = np.bincount(who, minlength=n)
    s1 = np.bincount(who, weights=isi, minlength=n)
    s2 = np.bincount(who, weights=isi * isi, minlength=n)
    ok = cnt >= 2
    mean = s1[ok] / cnt[ok]
    var = np.maximum(s2[ok] / cnt[ok] - mean * mean, 0.0)
    return rate, float(np.mean(np.sqrt(var) / mean)) if ok.any() else float("nan")


# ---------------------------------------------------------------------------
# C3: the batched sweep
# ---------------------------------------------------------------------------
_SWEEP = None


def _sweep_worker(args):
    """One instance of the sweep on one CPU core (parallel_run semantics:
    every instance is its own single-threaded simulation)."""
    from oracle import oracle as orc

    i, steps = args
    sw = _SWEEP
    t0 = time.perf_counter()
    orc.sweep_instance_run(sw.recurrent[i], sw.w_in, sw.phase[:, i], sw.omega, float(sw.gains[i]), sw.tau,
                           sw.dt, sw.r0[i], steps)
    return time.perf_counter() - t0


def c3_cpu_baseline(sw, steps, workers):
    """core/Network.py:751-926 (parallel_run): one process per instance,
    `workers` at a time, each advancing its instance `steps` steps through
    the oracle.  Forked before this process touches the GPU."""
    import multiprocessing as mp

    global _SWEEP
    _SWEEP = sw
    ctx = mp.get_context("fork")
    order = sorted(range(sw.n_inst), key=lambda i: -int(sw.recurrent[i].nnz))
    t0 = time.perf_counter()
    with ctx.Pool(workers) as pool:
        per = pool.map(_sweep_worker, [(i, steps) for i in order], chunksize=1)
    wall = time.perf_counter() - t0
    _SWEEP = None
    return steps / wall, wall, float(sum(per))


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
    log(f"stdp: {steps} steps in {el:.3f} s ({el / steps * 1e6:.1f} us/step), {events / steps:.3e} synapse "
        f"events/step")
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
    copy_gbs = copy_probe()
    kern = {"pre_event": {"ms": t_pre, "bytes": b_pre, "syn_events": syn_pre},
            "post_event": {"ms": t_post, "bytes": b_post, "syn_events": syn_post}}
    dom = "pre_event" if t_pre >= t_post else "post_event"
    ach = kern[dom]["bytes"] / (kern[dom]["ms"] * 1e-3) / 1e9
    traffic = traffic_from_profiles(f"{dom}_f64_stdp5e7_traffic.txt") if prec == "double" else None
    roof = roofline("hbm", ach, HBM_PEAK_GBS, "GB/s", traffic, dom, kernel_ms=kern[dom]["ms"],
                    algorithmic_bytes=kern[dom]["bytes"])
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
                         f"single thread; max |w_gpu - w_oracle| = {parity:.2e}", "host": host_cpu()}
        log(f"stdp: oracle {ev_w / cel:.3e} events/s, max |w_gpu - w_oracle| {parity:.2e}")
    return record("synaptic events/s", events / el, "synaptic events/s", steps, warmup, el * 1e3 / steps,
                 "f64" if prec == "double" else "f32",
                 {"workload": "STDP 100000 Poisson(20 Hz) -> 50000 Poisson(10 Hz), fixed_number_pre 1000 "
                              "(5e7 plastic synapses), dt=1ms", "precision": prec},
                 roof, cpu, kernels=kern, w_max_abs_diff_vs_oracle=parity, hbm_copy_gbs=copy_gbs,
                 pre_spikes_per_step=float(pc[warmup:].mean()), post_spikes_per_step=float(qc[warmup:].mean()))


def bench_c3(args):
    """Batched sweep: 256 instances, N_i ~ U{500..8000}, recurrent p = 0.05,
    shared dense 2000 x 2000 input, gains linspace(0.1, 2.0)."""
    steps = args.c3_steps
    rng = np.random.default_rng(2026)
    sizes = rng.integers(500, 8001, 256)
    t_b = time.perf_counter()
    from annarchy_amd.batched import RateSweep, time_plain
    sw = RateSweep(sizes, n_in=2000, p=0.05, seed=3)
    log(f"c3: sweep of {sw.total} neurons, {sw.nnz} synapses built in {time.perf_counter() - t_b:.0f} s")
    cpu = None
    if not args.no_cpu:
        h = host_cpu()
        cs = 20
        sps, wall, busy = c3_cpu_baseline(sw, cs, h["threads"])
        cpu = {"value": sps, "unit": "steps/s", "cores": h["threads"], "kind": "port",
               "sample": f"{cs} steps of all 256 instances, one single-threaded oracle process per instance, "
                         f"{h['threads']} at a time (parallel_run, core/Network.py:751-926): {wall:.1f} s wall, "
                         f"{busy:.1f} s of CPU work", "host": h}
        log(f"c3: cpu baseline {sps:.2f} steps/s on {h['threads']} processes")
    b = sw.batched(lanes=16)
    G = graph_batch_for(steps, 50)
    b.run(20, G)
    t0 = time.perf_counter()
    b.run(steps, G)
    el_b = time.perf_counter() - t0
    info = b.info()
    ms_spmv = b.time_kernel(2, 20)
    ms_gemm = b.time_kernel(1, 50)
    log(f"c3: batched {el_b * 1e3 / steps:.3f} ms/step, spmv {ms_spmv * 1e3:.0f} us, gemm {ms_gemm * 1e3:.1f} us "
        f"per launch of {info['chunk_steps']} steps")
    sps_b = steps / el_b
    sps_p = worst = pop_err = None
    plain_sample = None
    if not args.no_plain:
        # every 8th instance as a plain FP64 network of its own (CSR
        # everywhere): its time scaled to the 256 instances, and the batched
        # rates against its rates after the same steps
        plain_sample = list(range(0, sw.n_inst, 8))
        t_p = time.perf_counter()
        nets = sw.plain_networks(plain_sample)
        log(f"c3: {len(nets)} plain instances built in {time.perf_counter() - t_p:.0f} s")
        plain_steps = min(steps, 200)
        el_p = time_plain(nets, plain_steps) * sw.n_inst / len(nets)
        for net, exc in nets:
            net.run(20 + steps - plain_steps)
        got = b.rates()
        worst, pop_err = 0.0, 0.0
        for i, (net, exc) in zip(plain_sample, nets):
            ref = exc.r.reshape(-1)
            worst = max(worst, float(np.max(np.abs(got[i] - ref)) / max(np.max(np.abs(ref)), 1e-3)))
            pop_err = max(pop_err, abs(float(got[i].mean() - ref.mean())) / max(abs(ref.mean()), 1e-9))
        sps_p = plain_steps / el_p
        log(f"c3: plain instances {sps_p:.2f} steps/s (all 256, scaled), worst rate error {worst:.2e}")
    achieved = info["sparse_bytes"] / (ms_spmv * 1e-3) / 1e9
    # one GEMM launch covers the input psp of a whole chunk (graph batch) of steps
    gemm_flop = info["dense_flops"] * info["chunk_steps"]
    tflops = gemm_flop / (ms_gemm * 1e-3) / 1e12
    roof = roofline("hbm", achieved, HBM_PEAK_GBS, "GB/s",
                    traffic_from_profiles("batch_sell_f32_c3_sweep256_traffic.txt") if info["mode"] == 2 else None,
                    {0: "batch_spmv_update", 1: "batch_spmv_lds", 2: "batch_sell"}[info["mode"]],
                    kernel_ms=ms_spmv, algorithmic_bytes=info["sparse_bytes"])
    return record("steps/s (batched sweep)", sps_b, "steps/s", steps, 20, el_b * 1e3 / steps, "f32/bf16",
                  {"workload": "256 rate instances N~U{500..8000} p=0.05 + shared dense 2000x2000 input, 2000 steps",
                   "neurons": sw.total, "synapses": sw.nnz, "graph_steps": G},
                  roof, cpu, plain_steps_per_s=sps_p, speedup_vs_plain=sps_b / sps_p if sps_p else None,
                  plain_instances_compared=plain_sample,
                  max_rel_rate_error_vs_fp64=worst, max_rel_population_rate_error=pop_err,
                  gemm={"kernel": "batch_gemm", "ms_per_launch": ms_gemm, "steps_per_launch": info["chunk_steps"],
                        "ms_per_step": ms_gemm / info["chunk_steps"], "useful_flop_per_launch": gemm_flop,
                        "mfma_flop_per_launch": 3 * gemm_flop, "tflops": tflops,
                        "peak_bf16_tflops": BF16_PEAK_TFLOPS, "frac": tflops / BF16_PEAK_TFLOPS,
                        "mfma_frac": 3 * tflops / BF16_PEAK_TFLOPS})


# ---------------------------------------------------------------------------
# C4: the learned selector
# ---------------------------------------------------------------------------
def bench_c4(args):
    """Held-out regret of the trained selector against exhaustive tuning,
    timed live: held-out matrices of both corpus sets (N 1e3..1e6 with
    nnz <= 2e6, and the set at the rate benchmark's scale, nnz 1e7..1e8) are
    regenerated from their seeds, every feasible configuration is timed on
    the device (the exhaustive autotuner), and the MLP's and the reference
    heuristic's picks are compared with the best; then the same for the
    1e8-synapse rate network itself (in-degree fixed 1000)."""
    import json as js

    from annarchy_amd import selector as S
    from oracle import oracle as orc

    d = np.load(S.DATA_DIR / "selector_corpus.npz")
    specs = [js.loads(s) for s in d["specs"]]
    heur = np.array([s["heuristic"] for s in specs])
    groups = np.array([1 if s.get("large") else 0 for s in specs])
    _, te = S.holdout_split(d["times"], heur, groups=groups)
    small = [i for i in te if groups[i] == 0]
    large = [i for i in te if groups[i] == 1]
    pick = small[:: max(1, len(small) // args.c4_sample)][: args.c4_sample] + large[: args.c4_large]
    sel = S.MLPSelector.load()
    T, choice, hc, nnz, grp = [], [], [], [], []
    t0 = time.perf_counter()
    for i in pick:
        csr = S.generate(specs[i])
        f = S.features(csr, specs[i]["seed"] & 0xFFFF)
        choice.append(int(sel.predict(f)[0]))
        # the best of 5 means of 10 launches: small matrices' launches last
        # microseconds, and their timing varies between boxes
        T.append(S.autotune(csr, repeats=5))
        hc.append(int(heur[i]))
        nnz.append(csr.nnz)
        grp.append(int(groups[i]))
        del csr
    el = time.perf_counter() - t0
    T, grp = np.array(T), np.array(grp)
    choice = S.with_fallback(T, np.array(choice))
    hcs = S.with_fallback(T, np.array(hc))
    reg = S.regret(T, choice)
    reg_h = S.regret(T, hcs)
    # the rate network of the benchmark (1e8 synapses, in-degree fixed 1000)
    c1 = build_c1(100_000, 1000, 1234)
    c1_pick = S.select(c1, mode="mlp")
    c1_cls = S.CLASSES.index(c1_pick)
    t_c1 = S.autotune(c1)
    c1_h = S.heuristic_class(c1)
    c1_rec = {"workload": "rate network N=100000, in-degree fixed 1000, 1e8 synapses",
              "mlp_choice": c1_pick, "exhaustive_best": S.CLASSES[int(np.argmin(t_c1))],
              "mlp_regret": float(t_c1[c1_cls] / t_c1.min() - 1.0),
              "heuristic_choice": S.CLASSES[c1_h], "heuristic_regret": float(t_c1[c1_h] / t_c1.min() - 1.0),
              "times_ms": [None if not np.isfinite(x) else float(x) for x in t_c1]}
    log(f"c4: C1 matrix: MLP {c1_pick} regret {c1_rec['mlp_regret']:.3f}, heuristic {c1_rec['heuristic_regret']:.3f}")
    # roofline: the selected configuration's psp on the C1 matrix
    from annarchy_amd.models import RateNoisy
    from annarchy_amd.network import Network
    net = Network(dt=1.0, seed=0)
    pop = net.create(c1.n_rows, RateNoisy)
    proj = net.connect(pop, pop, "exc")
    proj.from_csr(c1, storage_format="csr", lanes=64)
    net.compile(allow_build=False)
    proj.set_format(**c1_pick)
    ms = proj.time_psp(50)
    pb = proj.psp_bytes()
    ach = pb / (ms * 1e-3) / 1e9
    # the default sliced-tile layout of the headline network has its HBM
    # counters committed (the same kernel on the same matrix)
    sp_default = c1_pick.get("fmt") == "sell_panel" and c1_pick.get("lanes") == 2 and c1_pick.get("sigma") == 8 \
        and c1_pick.get("chunk") == 2048 and c1_pick.get("panel_cols") == 20479
    roof = roofline("hbm", ach, HBM_PEAK_GBS, "GB/s",
                    traffic_from_profiles("sell_panel_psp_f64_rb2048_c1_fixed1000_traffic.txt") if sp_default else None,
                    f"psp of the selected configuration {c1_pick} on the 1e8-synapse rate network",
                    kernel_ms=ms, algorithmic_bytes=pb, matrix_nnz=int(c1.nnz))
    net.destroy()
    cpu = None
    if not args.no_cpu:
        # the selector's own work on the host: the 16 features of the 1e8-synapse matrix
        t1 = time.perf_counter()
        S.features(c1)
        tf = time.perf_counter() - t1
        n = c1.n_rows
        r = np.random.default_rng(1).random(n)
        noise = np.zeros((1, n))
        t1, steps = time.perf_counter(), 0
        while time.perf_counter() - t1 < 2.0:
            orc.rate_run(c1.row_ptr, c1.col_idx, c1.w, r, 10.0, 1.0, noise, 1, nthreads=1)
            steps += 1
        cel = time.perf_counter() - t1
        cpu = {"value": c1.nnz * steps / cel, "unit": "synaptic events/s", "cores": 1, "kind": "port",
               "sample": f"{steps} oracle CSR steps of the 1e8-synapse matrix, single thread (the reference's "
                         f"format for it, csr); feature extraction for the selector {tf:.2f} s",
               "host": host_cpu()}
    def stats(m):
        return {"n": int(m.sum()), "mlp_mean_regret": float(reg[m].mean()) if m.any() else None,
                "heuristic_mean_regret": float(reg_h[m].mean()) if m.any() else None}
    log(f"c4: {len(pick)} held-out matrices tuned in {el:.1f} s: MLP regret {reg.mean():.4f}, "
        f"heuristic {reg_h.mean():.3f}")
    return record("held-out regret vs exhaustive tuning", float(reg.mean()), "fraction", len(pick), 0,
                  el * 1e3 / len(pick), "f32",
                  {"workload": f"learned selector, {len(pick)} held-out matrices of the corpus (2000 matrices with "
                               "N 1e3..1e6, 40 with nnz 1e7..1e8), MLP 16-64-64-12, 12 configurations timed live"},
                  roof, cpu, higher=False, heuristic_regret=float(reg_h.mean()),
                  regret_median=float(np.median(reg)), regret_p90=float(np.quantile(reg, 0.9)),
                  total_time_ratio=float(T[np.arange(len(T)), choice].sum() / T.min(axis=1).sum()),
                  small=stats(grp == 0), large=stats(grp == 1), c1_matrix=c1_rec)


# ---------------------------------------------------------------------------
# driver
# ---------------------------------------------------------------------------
def run_child(cfg, args):
    cmd = [sys.executable, "-u", str(Path(__file__).resolve()), "--config", cfg, "--single",
           "--steps", str(args.steps), "--warmup", str(args.warmup)]
    if args.no_cpu:
        cmd.append("--no-cpu")
    log(f"bench: {cfg} ...")
    t0 = time.perf_counter()
    try:
        res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True,
                             timeout=CHILD_TIMEOUT_S.get(cfg, 300))
    except subprocess.TimeoutExpired:
        return {"error": f"{cfg} exceeded {CHILD_TIMEOUT_S.get(cfg, 300)} s"}
    lines = [ln for ln in res.stdout.splitlines() if ln.strip().startswith("{")]
    if res.returncode != 0 or not lines:
        return {"error": f"{cfg} exited with {res.returncode}"}
    out = json.loads(lines[-1])
    out["child_wall_s"] = time.perf_counter() - t0
    return out


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
    # 2 steps of 4 synapses in flight, 8 waves per workgroup)
    ap.add_argument("--format", default="sell_panel")
    ap.add_argument("--lanes", type=int, default=2,
                    help="csr/csr_vec/csr_panel: lanes per row; sell: threads per row; sell_panel: steps in flight")
    ap.add_argument("--sigma", type=int, default=1, help="sell: sorting window")
    ap.add_argument("--waves", type=int, default=8, help="sell_panel: waves per workgroup")
    ap.add_argument("--panel-cols", type=int, default=20479,
                    help="sell_panel: columns per panel at most (20479: 160 KiB of LDS in FP64)")
    ap.add_argument("--chunk", type=int, default=0,
                    help="sell_panel: rows per tile (0: 2048 in FP64, 1024 in FP32)")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", default="all", choices=["all", "c0", "c1", "c1_f32", "c1_lognormal", "c2", "c2_f32", "stdp", "c3", "c4"],
                    help="all (default): c1 headline + c0, c2, c3, c4 records, each in a child process")
    ap.add_argument("--single", action="store_true", help="run one config in this process")
    ap.add_argument("--c2-steps", type=int, default=10000)
    ap.add_argument("--c3-steps", type=int, default=2000)
    ap.add_argument("--no-plain", action="store_true",
                    help="c3: skip the plain per-instance FP64 networks (and the error and speedup they give)")
    ap.add_argument("--c4-sample", type=int, default=20)
    ap.add_argument("--c4-large", type=int, default=2)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only time the streaming copy of 2 GiB (the kernel that calibrates HBM counters)")
    args = ap.parse_args()
    if args.copy_probe:
        print(json.dumps({"copy_probe_GBps": copy_probe()}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every GPU box of this setup has one GPU, and the simulation "
              "step of this backend is a single-device path (sharding is a design note, not built).",
              file=sys.stderr)
        sys.exit(2)

    if args.single or args.config not in ("all",):
        if args.config == "all":
            args.config = "c1"
        print(json.dumps(globals()["bench_" + args.config](args)), flush=True)
        return

    head = run_child("c1", args)
    if "error" in head:
        print(json.dumps({"metric": "synaptic events/s", "value": None, "error": head["error"]}))
        sys.exit(1)
    head["configs"] = {cfg: run_child(cfg, args) for cfg in SUB_CONFIGS}
    print(json.dumps(head), flush=True)


if __name__ == "__main__":
    main()
