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


def traffic_from_profiles(name):
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get("hbm_bytes_per_launch")
    except Exception:
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
    ap.add_argument("--format", default="csr")
    ap.add_argument("--lanes", type=int, default=64)
    ap.add_argument("--sigma", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the simulation step of this backend is a "
              "single-device path (multi-GPU sharding is a design note, not built).", file=sys.stderr)
        sys.exit(2)

    import annarchy_amd as ah

    t_build = time.perf_counter()
    csr = build_c1(args.neurons, args.degree, 1234, args.lognormal)
    log(f"connectivity: {csr.nnz} synapses in {time.perf_counter() - t_build:.1f} s")

    net = ah.Network(dt=1.0, seed=42, precision=args.precision)
    pop = net.create(args.neurons, ah.RateNoisy)
    proj = net.connect(pop, pop, "exc")
    proj.from_csr(csr, storage_format=args.format, lanes=args.lanes, sigma=args.sigma)
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
    tag = f"c1_{'lognormal' if args.lognormal else 'fixed'}_{args.format}_{dtype}"
    traffic = traffic_from_profiles(f"{tag}_traffic.json")
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
            "format": args.format,
            "lanes_per_row": args.lanes,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "csr_rows (compute_psp)",
            "kernel_ms": psp_ms,
            "algorithmic_bytes": psp_bytes,
        },
        "cpu_baseline": cpu,
        "hbm_gbs_step": psp_bytes / (ms_step * 1e-3) / 1e9,
        "hbm_copy_gbs": copy_gbs,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
