"""Batched parameter sweep: many rate-network instances of uneven size.

The reference runs a sweep as independent networks, one process each
(``Network.parallel_run``, core/Network.py:751-926).  Two forms are provided:

* :meth:`RateSweep.plain_networks` -- the reference's shape on one GPU: one
  FP64 :class:`~annarchy_amd.network.Network` per instance (CSR everywhere,
  the reference default), stepped one instance after the other;
* :class:`BatchedSweep` -- every instance in one device state: the recurrent
  projections as one block-diagonal CSR (FP32, compensated sums) fused with
  the neuron update, and the input projection whose weights all instances
  share as one GEMM on MFMA per graph batch (its input rates depend on t
  alone, so the rows of the product are (step, instance) pairs; split-BF16
  operands, FP32 accumulation); one SpMV + update launch per step, all
  captured in a HIP graph.

Instance i:  input population (n_in neurons)  r = 1 + 0.5*sin(omega*t + phase)
             exc population (N_i neurons)  tau dr/dt = -r + sum(exc) + gain_i*sum(inp), r >= 0
             exc -> exc  Bernoulli(p) connectivity, w ~ N(0, g/sqrt(p*N_i))
             inp -> exc  W_in[:min(N_i, n_in), :] of one shared n_in x n_in matrix
"""

from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import _lib
from . import connectivity as conn
from .models import Neuron, RateGain
from .network import Network

SIN_INPUT = Neuron(
    parameters="""
        omega = 0.0628 : population
        phase = 0.0
    """,
    equations="""
        r = 1.0 + 0.5 * sin(omega * t + phase)
    """,
    name="SinInput",
)


def bernoulli_csr(n: int, p: float, rng: np.random.Generator, sigma: float) -> conn.CSR:
    """Exact Bernoulli(p) pattern over the n x n cells (self-connections
    included) drawn with geometric gaps, sorted row-major, weights N(0, sigma)."""
    cells = n * n
    expect = int(p * cells * 1.1 + 10 * math.sqrt(p * cells) + 16)
    pos = np.cumsum(rng.geometric(p, size=expect)) - 1
    while pos[-1] < cells:
        more = np.cumsum(rng.geometric(p, size=expect // 4 + 16)) + pos[-1]
        pos = np.concatenate([pos, more])
    pos = pos[pos < cells]
    rows = pos // n
    cols = (pos % n).astype(np.int32)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=row_ptr[1:])
    w = rng.normal(0.0, sigma, size=pos.size)
    return conn.CSR(n, n, row_ptr, cols, w)


class RateSweep:
    """Description of the sweep (connectivity, shared input weights, phases, gains)."""

    def __init__(self, sizes, n_in=2000, p=0.05, gains=None, dt=1.0, tau=10.0,
                 omega=2 * math.pi / 100.0, g=0.9, seed=0):
        self.sizes = np.asarray(sizes, dtype=np.int64)
        self.n_inst = len(self.sizes)
        self.n_in, self.p, self.dt, self.tau, self.omega = n_in, p, dt, tau, omega
        self.gains = np.linspace(0.1, 2.0, self.n_inst) if gains is None else np.asarray(gains, float)
        rng = np.random.default_rng(seed)
        self.recurrent = [bernoulli_csr(int(n), p, rng, g / math.sqrt(p * n)) for n in self.sizes]
        self.w_in = rng.normal(0.0, 1.0 / math.sqrt(n_in), size=(n_in, n_in))
        self.phase = 2 * math.pi * rng.random((n_in, self.n_inst))
        self.r0 = [rng.random(int(n)) * 0.5 for n in self.sizes]

    @property
    def nnz(self) -> int:
        return int(sum(c.nnz for c in self.recurrent))

    @property
    def total(self) -> int:
        return int(self.sizes.sum())

    def input_csr(self, i: int) -> conn.CSR:
        n = int(self.sizes[i])
        rows = min(n, self.n_in)
        row_ptr = np.zeros(n + 1, dtype=np.int64)
        row_ptr[1:rows + 1] = np.arange(1, rows + 1) * self.n_in
        row_ptr[rows + 1:] = rows * self.n_in
        col = np.tile(np.arange(self.n_in, dtype=np.int32), rows)
        return conn.CSR(n, self.n_in, row_ptr, col, self.w_in[:rows].reshape(-1).copy())

    # -- plain form ---------------------------------------------------------------
    def plain_networks(self, instances=None):
        nets = []
        for i in (range(self.n_inst) if instances is None else instances):
            net = Network(dt=self.dt, seed=i, precision="double")
            inp = net.create(self.n_in, SIN_INPUT)
            exc = net.create(int(self.sizes[i]), RateGain)
            net.connect(exc, exc, "exc").from_csr(self.recurrent[i])
            net.connect(inp, exc, "inp").from_csr(self.input_csr(i))
            net.compile(allow_build=False)
            inp.omega = self.omega
            inp.phase = self.phase[:, i]
            exc.tau = self.tau
            exc.gain = float(self.gains[i])
            exc.r = self.r0[i]
            nets.append((net, exc))
        return nets

    # -- batched form ---------------------------------------------------------------
    def batched(self, lanes: int = 16) -> "BatchedSweep":
        return BatchedSweep(self, lanes)


class BatchedSweep:
    def __init__(self, sweep: RateSweep, lanes: int = 16):
        L = _lib.lib()
        self.sweep = sweep
        sizes = np.ascontiguousarray(sweep.sizes, dtype=np.int32)
        h = C.c_void_p()
        self._check(L.ahip_batch_create(sweep.n_inst, _lib.ptr(sizes, C.c_int32), sweep.n_in,
                                        sweep.dt, sweep.tau, C.byref(h)))
        self._h = h
        offsets = np.concatenate([[0], np.cumsum(sweep.sizes)])
        row_ptr = np.zeros(sweep.total + 1, dtype=np.int64)
        cols, ws = [], []
        nnz = 0
        for i, c in enumerate(sweep.recurrent):
            row_ptr[offsets[i] + 1: offsets[i + 1] + 1] = c.row_ptr[1:] + nnz
            nnz += c.nnz
            cols.append(c.col_idx + np.int32(offsets[i]))
            ws.append(c.w)
        col = np.ascontiguousarray(np.concatenate(cols), dtype=np.int32)
        w = np.ascontiguousarray(np.concatenate(ws), dtype=np.float64)
        self._check(L.ahip_batch_set_recurrent(h, _lib.ptr(row_ptr, C.c_int64), _lib.ptr(col, C.c_int32),
                                               _lib.ptr(w, C.c_double), lanes))
        w_in = np.ascontiguousarray(sweep.w_in)
        ph = np.ascontiguousarray(sweep.phase)
        self._check(L.ahip_batch_set_input(h, _lib.ptr(w_in, C.c_double), _lib.ptr(ph, C.c_double), sweep.omega))
        g = np.ascontiguousarray(sweep.gains, dtype=np.float64)
        self._check(L.ahip_batch_set_gain(h, _lib.ptr(g, C.c_double)))
        r0 = np.ascontiguousarray(np.concatenate(sweep.r0))
        self._check(L.ahip_batch_set_rates(h, _lib.ptr(r0, C.c_double)))
        self.offsets = offsets

    @staticmethod
    def _check(rc):
        if rc != 0:
            raise _lib.AhipError(_lib.lib().ahip_batch_last_error().decode())

    def set_mode(self, mode: int):
        """2: sliced chunks, one lane per row (default when every instance has
        fewer than 16384 neurons); 1: instance rates staged in LDS, L lanes per
        row; 0: gathers from global memory."""
        self._check(_lib.lib().ahip_batch_set_mode(self._h, int(mode)))

    def run(self, steps: int, graph_batch: int = 50):
        self._check(_lib.lib().ahip_batch_run(self._h, int(steps), graph_batch))

    def rates(self):
        out = np.empty(self.sweep.total)
        self._check(_lib.lib().ahip_batch_get_rates(self._h, _lib.ptr(out, C.c_double)))
        return [out[self.offsets[i]:self.offsets[i + 1]] for i in range(self.sweep.n_inst)]

    def input_sums(self):
        out = np.empty((self.sweep.n_in, self.sweep.n_inst))
        self._check(_lib.lib().ahip_batch_get_input_sums(self._h, _lib.ptr(out, C.c_double)))
        return out

    def time_kernel(self, which: int, iters: int = 20) -> float:
        ms = C.c_double()
        self._check(_lib.lib().ahip_batch_time(self._h, which, iters, C.byref(ms)))
        return ms.value

    def info(self):
        nnz, total, sb, df = C.c_int64(), C.c_int32(), C.c_double(), C.c_double()
        self._check(_lib.lib().ahip_batch_info(self._h, C.byref(nnz), C.byref(total), C.byref(sb), C.byref(df)))
        mode, chunk = C.c_int32(), C.c_int32()
        self._check(_lib.lib().ahip_batch_get_mode(self._h, C.byref(mode)))
        self._check(_lib.lib().ahip_batch_chunk_steps(self._h, C.byref(chunk)))
        return dict(nnz=nnz.value, total=total.value, sparse_bytes=sb.value, dense_flops=df.value, mode=mode.value,
                    chunk_steps=chunk.value)

    def __del__(self):
        try:
            if self._h is not None and _lib._lib is not None:
                _lib.lib().ahip_batch_destroy(self._h)
                self._h = None
        except Exception:
            pass


def time_plain(nets, steps: int) -> float:
    """Wall time to advance every plain instance by `steps` (one after the other)."""
    t0 = time.perf_counter()
    for net, _ in nets:
        net.run(steps)
    return time.perf_counter() - t0
