"""ctypes front of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the annarchy_amd package.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# AHIP_ORACLE_LIB: another build of oracle.cpp (the sanitizer build of the
# CPU test that reruns the oracle's tests under ASan and UBSan)
LIB = Path(os.environ["AHIP_ORACLE_LIB"]) if os.environ.get("AHIP_ORACLE_LIB") else HERE / "build" / "liboracle.so"

_I32, _I64, _D = C.c_int32, C.c_int64, C.c_double
_PI32, _PI64, _PD = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_double)
_PI8 = C.POINTER(C.c_int8)


class OrcProj(C.Structure):
    _fields_ = [
        ("pre_is_input", _I32),
        ("target", _I32),
        ("num_cols", _I32),
        ("col_ptr", _PI64),
        ("row_idx", _PI32),
        ("w_col", _PD),
        ("w_const", _D),
        ("delay", _I32),
        ("delay_col", _PI32),
    ]


class OrcSpikingCfg(C.Structure):
    _fields_ = [
        ("model", _I32),
        ("N", _I32),
        ("dt", _D),
        ("t0", _I64),
        ("El", _D), ("Vr", _D), ("Erev_exc", _D), ("Erev_inh", _D), ("Vt", _D),
        ("tau", _D), ("tau_exc", _D), ("tau_inh", _D), ("I", _D),
        ("has_refractory", _I32),
        ("M", _I32),
        ("rates", _D),
        ("uniforms", _PD),
        ("n_proj", _I32),
        ("proj", C.POINTER(OrcProj)),
    ]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.orc_delayed_psp_run.argtypes = [_I32, _PI64, _PI32, _PD, _PI32, _I32, _I32, _I32, _PD, _PD,
                                          _I32, _PD]
        L.orc_delayed_psp_run.restype = None
        L.orc_rate_run_delayed.argtypes = [_I32, _PI64, _PI32, _PD, _PI32, _I32, _I32, _PD, _PD, _PD,
                                           _D, _D, _PD, _I32, _PD]
        L.orc_rate_run_delayed.restype = None
        L.orc_csr_from_lil.argtypes = [_I32, _I32, _PI32, _PI64, _PI32, _PI64, _PI32]
        L.orc_csr_from_lil.restype = _I64
        L.orc_csrc_inverse.argtypes = [_I32, _I32, _PI64, _PI32, _PI64, _PI32, _PI64]
        L.orc_csrc_inverse_fast.argtypes = [_I32, _I32, _PI64, _PI32, _PI64, _PI32, _PI64]
        L.orc_csrc_inverse_fast.restype = None
        L.orc_csrc_inverse.restype = None
        L.orc_ell_from_lil.argtypes = [_I32, _PI64, _PI32, _PD, _I32, _PI32, _PD, _PI32]
        L.orc_ell_from_lil.restype = _I32
        L.orc_sell_from_lil.argtypes = [_I32, _I32, _PI32, _PI64, _PI32, _PD, _I32, _I32, _PI64, _PI32, _PD, _PI8]
        L.orc_sell_from_lil.restype = _I64
        L.orc_bsr_from_lil.argtypes = [_I32, _I32, _I32, _I32, _PI32, _PI64, _PI32, _PD, _PI32, _PI32, _PD]
        L.orc_bsr_from_lil.restype = _I64
        L.orc_dense_from_lil.argtypes = [_I32, _I32, _I32, _PI32, _PI64, _PI32, _PD, _I32, _PD]
        L.orc_dense_from_lil.restype = None
        L.orc_spmv_csr.argtypes = [_I32, _PI32, _PI64, _PI32, _PD, _PD, _PD, _I32]
        L.orc_spmv_csr.restype = None
        L.orc_spmv_csr_f32.argtypes = [_I32, _PI64, _PI32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                       C.POINTER(C.c_float), _I32]
        L.orc_spmv_csr_f32.restype = None
        L.orc_rate_run.argtypes = [_I32, _PI64, _PI32, _PD, _PD, _PD, _D, _D, _PD, _I32, _PD, _I32]
        L.orc_rate_run.restype = None
        L.orc_spiking_run.argtypes = [C.POINTER(OrcSpikingCfg), _PD, _PD, _PD, _PI32, _PI32, _PI64,
                                      _I32, _PI32, _PI32, _I64, _PD, _PD, _PI32]
        L.orc_spiking_run.restype = _I64
        L.orc_sweep_instance_run.argtypes = [_I32, _PI64, _PI32, _PD, _I32, _PD, _PD, _D, _D, _D, _D,
                                             _PD, _I32, _I64, _PD]
        L.orc_sweep_instance_run.restype = None
        L.orc_plastic_run.argtypes = [_I32, _I32, _I32, _PI64, _PI32, _PD, _PD, _PD, _PI64, _PD, _D, _I64,
                                      _I32, _PI64, _PI32, _PI64, _PI32, _PI64, _PI64, _I32, _PD]
        L.orc_plastic_run.restype = None
        L.orc_rate_plastic_run.argtypes = [_I32, _I32, _I32, _PI64, _PI32, _PD, _PD, _PD, _D, _I32, _I32, _PD,
                                           _PD, _PD, _PD]
        L.orc_rate_plastic_run.restype = None
        L.orc_mlp_forward.argtypes = [_I32, _PD, _PD, _PD, _I32, _I32, _I32, _I32]
        L.orc_mlp_forward.restype = None
        L.orc_mlp_loss_grad.argtypes = [_I32, _PD, _PD, _PD, _PD, _I32, _I32, _I32, _I32]
        L.orc_mlp_loss_grad.restype = _D
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _lil_arrays(pre_lists):
    lens = np.array([len(x) for x in pre_lists], dtype=np.int64)
    ptr = np.zeros(len(pre_lists) + 1, dtype=np.int64)
    np.cumsum(lens, out=ptr[1:])
    cols = np.concatenate([np.asarray(x, dtype=np.int32) for x in pre_lists]) if len(pre_lists) else np.zeros(0, np.int32)
    return ptr, np.ascontiguousarray(cols, dtype=np.int32)


def _vals(val_lists):
    if val_lists is None:
        return None
    return np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in val_lists]))


# ---- containers -------------------------------------------------------------
def csr_from_lil(num_rows, post_ranks, pre_lists):
    post = np.ascontiguousarray(post_ranks, dtype=np.int32)
    ptr, cols = _lil_arrays(pre_lists)
    row_begin = np.zeros(num_rows + 1, dtype=np.int64)
    col_idx = np.zeros(max(1, cols.size), dtype=np.int32)
    nnz = lib().orc_csr_from_lil(num_rows, len(post), _p(post, _I32), _p(ptr, _I64), _p(cols, _I32),
                                 _p(row_begin, _I64), _p(col_idx, _I32))
    if nnz < 0:
        raise ValueError("LIL rows not ascending / not within num_rows")
    return row_begin, col_idx[:nnz]


def csrc_inverse(num_rows, num_cols, row_begin, col_idx, fast=None):
    """orc_csrc_inverse (the map-based restatement of CSRCMatrix.hpp:164-210)
    or, for large views (fast=None and nnz > 1e6, or fast=True), the
    counting sort orc_csrc_inverse_fast that returns the same arrays."""
    nnz = int(row_begin[-1])
    if fast is None:
        fast = nnz > 1_000_000
    col_ptr = np.zeros(num_cols + 1, dtype=np.int64)
    row_idx = np.zeros(max(1, nnz), dtype=np.int32)
    inv_idx = np.zeros(max(1, nnz), dtype=np.int64)
    rb = np.ascontiguousarray(row_begin, dtype=np.int64)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    fn = lib().orc_csrc_inverse_fast if fast else lib().orc_csrc_inverse
    fn(num_rows, num_cols, _p(rb, _I64), _p(ci, _I32), _p(col_ptr, _I64), _p(row_idx, _I32), _p(inv_idx, _I64))
    return col_ptr, row_idx[:nnz], inv_idx[:nnz]


def ell_from_lil(pre_lists, val_lists=None, row_major=True):
    ptr, cols = _lil_arrays(pre_lists)
    vals = _vals(val_lists)
    L = lib()
    maxnzr = L.orc_ell_from_lil(len(pre_lists), _p(ptr, _I64), _p(cols, _I32), None, int(row_major), None, None, None)
    col = np.zeros(max(1, maxnzr * len(pre_lists)), dtype=np.int32)
    val = np.zeros(col.size, dtype=np.float64)
    rl = np.zeros(max(1, len(pre_lists)), dtype=np.int32)
    L.orc_ell_from_lil(len(pre_lists), _p(ptr, _I64), _p(cols, _I32), _p(vals, _D), int(row_major),
                       _p(col, _I32), _p(val, _D), _p(rl, _I32))
    n = maxnzr * len(pre_lists)
    return maxnzr, col[:n], val[:n], rl[: len(pre_lists)]


def sell_from_lil(num_rows, post_ranks, pre_lists, val_lists=None, block_size=3, row_major=True):
    post = np.ascontiguousarray(post_ranks, dtype=np.int32)
    ptr, cols = _lil_arrays(pre_lists)
    vals = _vals(val_lists)
    L = lib()
    total = L.orc_sell_from_lil(num_rows, len(post), _p(post, _I32), _p(ptr, _I64), _p(cols, _I32), None,
                                block_size, int(row_major), None, None, None, None)
    nb = (num_rows + block_size - 1) // block_size
    row_ptr = np.zeros(nb + 1, dtype=np.int64)
    col = np.zeros(max(1, total), dtype=np.int32)
    val = np.zeros(max(1, total), dtype=np.float64)
    mask = np.zeros(max(1, total), dtype=np.int8)
    L.orc_sell_from_lil(num_rows, len(post), _p(post, _I32), _p(ptr, _I64), _p(cols, _I32), _p(vals, _D),
                        block_size, int(row_major), _p(row_ptr, _I64), _p(col, _I32), _p(val, _D),
                        _p(mask, C.c_int8))
    return row_ptr, col[:total], val[:total], mask[:total]


def bsr_from_lil(num_rows, num_cols, tile, post_ranks, pre_lists, val_lists=None):
    post = np.ascontiguousarray(post_ranks, dtype=np.int32)
    ptr, cols = _lil_arrays(pre_lists)
    vals = _vals(val_lists)
    L = lib()
    nt = L.orc_bsr_from_lil(num_rows, num_cols, tile, len(post), _p(post, _I32), _p(ptr, _I64),
                            _p(cols, _I32), None, None, None, None)
    nbr = (num_rows + tile - 1) // tile
    brp = np.zeros(nbr + 1, dtype=np.int32)
    bci = np.zeros(max(1, nt), dtype=np.int32)
    val = np.zeros(max(1, nt * tile * tile), dtype=np.float64)
    L.orc_bsr_from_lil(num_rows, num_cols, tile, len(post), _p(post, _I32), _p(ptr, _I64),
                       _p(cols, _I32), _p(vals, _D), _p(brp, _I32), _p(bci, _I32), _p(val, _D))
    return brp, bci[:nt], val[: nt * tile * tile]


def dense_from_lil(num_rows, num_cols, post_ranks, pre_lists, val_lists=None, row_major=True):
    post = np.ascontiguousarray(post_ranks, dtype=np.int32)
    ptr, cols = _lil_arrays(pre_lists)
    vals = _vals(val_lists)
    out = np.zeros(num_rows * num_cols, dtype=np.float64)
    lib().orc_dense_from_lil(num_rows, num_cols, len(post), _p(post, _I32), _p(ptr, _I64), _p(cols, _I32),
                             _p(vals, _D), int(row_major), _p(out, _D))
    return out


# ---- rate path ---------------------------------------------------------------
def spmv_csr(row_ptr, col_idx, w, r, target=None, nthreads=1):
    n = len(row_ptr) - 1
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    ww = np.ascontiguousarray(w, dtype=np.float64)
    rr = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(n) if target is None else np.array(target, dtype=np.float64)
    lib().orc_spmv_csr(n, None, _p(rp, _I64), _p(ci, _I32), _p(ww, _D), _p(rr, _D), _p(out, _D), nthreads)
    return out


def spmv_csr_f32(row_ptr, col_idx, w, r, nthreads=1):
    n = len(row_ptr) - 1
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    ww = np.ascontiguousarray(w, dtype=np.float32)
    rr = np.ascontiguousarray(r, dtype=np.float32)
    out = np.zeros(n, dtype=np.float32)
    lib().orc_spmv_csr_f32(n, _p(rp, _I64), _p(ci, _I32), _p(ww, C.c_float), _p(rr, C.c_float),
                           _p(out, C.c_float), nthreads)
    return out


def rate_run(row_ptr, col_idx, w, r0, tau, dt, noise, nsteps, trace=False, nthreads=1):
    n = len(row_ptr) - 1
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    ww = np.ascontiguousarray(w, dtype=np.float64)
    r = np.array(r0, dtype=np.float64)
    s = np.zeros(n)
    nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
    tr = np.zeros(nsteps * n) if trace else None
    lib().orc_rate_run(n, _p(rp, _I64), _p(ci, _I32), _p(ww, _D), _p(r, _D), _p(s, _D), tau, dt,
                       _p(nz, _D), nsteps, _p(tr, _D), nthreads)
    return r, s, (tr.reshape(nsteps, n) if trace else None)


def delayed_psp_run(row_ptr, col_idx, w, pre_seq, max_delay, delays=None, delay=1, hist=None):
    """psp of one projection per step for rates pre_seq[s] (end of step s);
    `hist` is the initial queue (max_delay x n_pre; zeros by default)."""
    n_post = len(row_ptr) - 1
    pre_seq = np.ascontiguousarray(pre_seq, dtype=np.float64)
    nsteps, n_pre = pre_seq.shape
    h = np.zeros((max_delay, n_pre)) if hist is None else np.array(hist, dtype=np.float64)
    dl = None if delays is None else np.ascontiguousarray(delays, dtype=np.int32)
    out = np.zeros((nsteps, n_post))
    lib().orc_delayed_psp_run(n_post, _p(np.ascontiguousarray(row_ptr, dtype=np.int64), _I64),
                              _p(np.ascontiguousarray(col_idx, dtype=np.int32), _I32),
                              _p(np.ascontiguousarray(w, dtype=np.float64), _D), _p(dl, _I32), int(delay),
                              int(max_delay), n_pre, _p(pre_seq, _D), _p(h, _D), nsteps, _p(out, _D))
    return out, h


def rate_run_delayed(row_ptr, col_idx, w, r0, tau, dt, noise, nsteps, max_delay, delays=None, delay=1,
                     hist=None, trace=False):
    n = len(row_ptr) - 1
    h = np.zeros((max_delay, n)) if hist is None else np.array(hist, dtype=np.float64)
    dl = None if delays is None else np.ascontiguousarray(delays, dtype=np.int32)
    r = np.array(r0, dtype=np.float64)
    s = np.zeros(n)
    nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
    tr = np.zeros(nsteps * n) if trace else None
    lib().orc_rate_run_delayed(n, _p(np.ascontiguousarray(row_ptr, dtype=np.int64), _I64),
                               _p(np.ascontiguousarray(col_idx, dtype=np.int32), _I32),
                               _p(np.ascontiguousarray(w, dtype=np.float64), _D), _p(dl, _I32), int(delay),
                               int(max_delay), _p(h, _D), _p(r, _D), _p(s, _D), tau, dt, _p(nz, _D), nsteps,
                               _p(tr, _D))
    return r, s, (tr.reshape(nsteps, n) if trace else None)


# ---- spiking path ----------------------------------------------------------------
COBA_DEFAULTS = dict(El=-60.0, Vr=-60.0, Erev_exc=0.0, Erev_inh=-80.0, Vt=-50.0, tau=20.0,
                     tau_exc=5.0, tau_inh=10.0, I=20.0)


def spiking_run(model, N, dt, v, nsteps, projections=(), refractory=None, g_exc=None, g_inh=None,
                params=None, poisson=None, t0=0, trace=False, spike_cap=None):
    """projections: list of dicts with keys pre_is_input, target (0 exc / 1 inh),
    num_cols, col_ptr, row_idx, w_col (or None) and w_const.
    poisson: dict(M=..., rates=..., uniforms=[nsteps, M]) or None."""
    p = dict(COBA_DEFAULTS)
    if params:
        p.update(params)
    keep = []
    projs = (OrcProj * max(1, len(projections)))()
    for k, pr in enumerate(projections):
        cp = np.ascontiguousarray(pr["col_ptr"], dtype=np.int64)
        ri = np.ascontiguousarray(pr["row_idx"], dtype=np.int32)
        wc = None if pr.get("w_col") is None else np.ascontiguousarray(pr["w_col"], dtype=np.float64)
        dc = None if pr.get("delay_col") is None else np.ascontiguousarray(pr["delay_col"], dtype=np.int32)
        keep += [cp, ri, wc, dc]
        projs[k] = OrcProj(int(pr.get("pre_is_input", 0)), int(pr["target"]), int(pr["num_cols"]),
                           _p(cp, _I64), _p(ri, _I32), _p(wc, _D), float(pr.get("w_const", 0.0)),
                           int(pr.get("delay", 1)), _p(dc, _I32))
    cfg = OrcSpikingCfg()
    cfg.model = model
    cfg.N = N
    cfg.dt = dt
    cfg.t0 = t0
    for k in COBA_DEFAULTS:
        setattr(cfg, k, float(p[k]))
    cfg.has_refractory = 1 if refractory is not None else 0
    unif = None
    if poisson is not None:
        cfg.M = int(poisson["M"])
        cfg.rates = float(poisson["rates"])
        unif = np.ascontiguousarray(poisson["uniforms"], dtype=np.float64)
        cfg.uniforms = _p(unif, _D)
    cfg.n_proj = len(projections)
    cfg.proj = projs
    v = np.array(v, dtype=np.float64)
    ge = np.zeros(N) if g_exc is None else np.array(g_exc, dtype=np.float64)
    gi = np.zeros(N) if g_inh is None else np.array(g_inh, dtype=np.float64)
    ref = np.zeros(N, dtype=np.int32) if refractory is None else np.ascontiguousarray(refractory, dtype=np.int32)
    rr = np.zeros(N, dtype=np.int32)
    last = np.full(N, -10000, dtype=np.int64)
    counts = np.zeros(nsteps, dtype=np.int32)
    cap = spike_cap if spike_cap is not None else max(1024, N * nsteps // 4)
    idx = np.zeros(cap, dtype=np.int32)
    tv = np.zeros(nsteps * N) if trace else None
    tg = np.zeros(nsteps * N) if trace else None
    in_counts = np.zeros(nsteps, dtype=np.int32)
    total = lib().orc_spiking_run(C.byref(cfg), _p(v, _D), _p(ge, _D), _p(gi, _D), _p(ref, _I32),
                                  _p(rr, _I32), _p(last, _I64), nsteps, _p(counts, _I32), _p(idx, _I32),
                                  cap, _p(tv, _D), _p(tg, _D), _p(in_counts, _I32))
    if total < 0:
        raise ValueError("spike buffer overflow; raise spike_cap")
    return dict(v=v, g_exc=ge, g_inh=gi, refractory_remaining=rr, last_spike=last, counts=counts,
                idx=idx[:total], trace_v=None if tv is None else tv.reshape(nsteps, N),
                trace_gexc=None if tg is None else tg.reshape(nsteps, N), input_counts=in_counts)


def csrc_projection(csr, target, pre_is_input=False, delay=1, delays=None):
    """Backward view of a connectivity.CSR as an oracle projection dict;
    `delays` (steps per synapse, forward CSR order) or a uniform `delay`."""
    col_ptr, row_idx, inv = csrc_inverse(csr.n_rows, csr.n_cols, csr.row_ptr, csr.col_idx)
    w_col = None if csr.w is None else np.asarray(csr.w)[inv]
    d_col = None if delays is None else np.asarray(delays, dtype=np.int32)[inv]
    return dict(pre_is_input=int(pre_is_input), target=target, num_cols=csr.n_cols, col_ptr=col_ptr,
                row_idx=row_idx, w_col=w_col, w_const=csr.w_const, delay=int(delay), delay_col=d_col)


def sweep_instance_run(csr, w_in, phase, omega, gain, tau, dt, r0, nsteps, trace=False):
    """One instance of the batched sweep in FP64 (rows 0..min(N, n_in)-1 of w_in)."""
    n = csr.n_rows
    n_in = w_in.shape[1]
    rp = np.ascontiguousarray(csr.row_ptr, dtype=np.int64)
    ci = np.ascontiguousarray(csr.col_idx, dtype=np.int32)
    ww = np.ascontiguousarray(csr.weights(), dtype=np.float64)
    wi = np.ascontiguousarray(w_in, dtype=np.float64)
    ph = np.ascontiguousarray(phase, dtype=np.float64)
    r = np.array(r0, dtype=np.float64)
    tr = np.zeros(nsteps * n) if trace else None
    lib().orc_sweep_instance_run(n, _p(rp, _I64), _p(ci, _I32), _p(ww, _D), n_in, _p(wi, _D), _p(ph, _D),
                                 omega, gain, tau, dt, _p(r, _D), nsteps, 0, _p(tr, _D))
    return r, (tr.reshape(nsteps, n) if trace else None)


# ---- MLP gradient oracle ------------------------------------------------------------
def mlp_loss_grad(X, targets, params, dims=(16, 64, 64, 12), grad=True):
    """targets: integer labels (n,) or target distributions (n, d_out)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    t = np.asarray(targets)
    if t.ndim == 1:
        t = np.eye(dims[-1])[t.astype(int)]
    y = np.ascontiguousarray(t, dtype=np.float64)
    P = np.ascontiguousarray(params, dtype=np.float64)
    G = np.zeros_like(P) if grad else None
    loss = lib().orc_mlp_loss_grad(X.shape[0], _p(X, _D), _p(y, _D), _p(P, _D), _p(G, _D), *dims)
    return loss, G


def mlp_forward(X, params, dims=(16, 64, 64, 12)):
    """FP64 logits of the selector MLP."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    P = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros((X.shape[0], dims[-1]))
    lib().orc_mlp_forward(X.shape[0], _p(X, _D), _p(P, _D), _p(out, _D), *dims)
    return out


def plastic_run(rule, row_ptr, col_idx, n_cols, w0, params, dt, pre_counts, pre_idx, post_counts, post_idx,
                t0=0, x0=None, y0=None, last_event=None, pre_last=None, post_last=None, delay=1):
    """orc_plastic_run: a plastic projection driven by the spike trains
    (per-step counts and concatenated ascending ranks) of its pre and post
    populations.  rule 0 STDP (params tau_plus, tau_minus, A_plus, A_minus,
    w_min, w_max), rule 1 the t_pre / t_post rule (tau_pre, tau_post, cApre,
    cApost, wmax); `delay` is the uniform delay in steps.  Returns w, x, y,
    last_event (forward order), the last spikes and g[steps, n_rows]
    (conductance added by each step's pre events)."""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
    n_rows = row_ptr.size - 1
    nnz = int(row_ptr[-1])
    steps = len(pre_counts)
    assert len(post_counts) == steps
    w = np.array(np.broadcast_to(np.asarray(w0, dtype=np.float64), (nnz,)))
    x = np.zeros(nnz) if x0 is None else np.array(x0, dtype=np.float64)
    y = np.zeros(nnz) if y0 is None else np.array(y0, dtype=np.float64)
    le = np.full(nnz, -10000, dtype=np.int64) if last_event is None else np.array(last_event, dtype=np.int64)
    pl = np.full(n_cols, -10000, dtype=np.int64) if pre_last is None else np.array(pre_last, dtype=np.int64)
    ql = np.full(n_rows, -10000, dtype=np.int64) if post_last is None else np.array(post_last, dtype=np.int64)
    pre_ptr = np.concatenate([[0], np.cumsum(pre_counts)]).astype(np.int64)
    post_ptr = np.concatenate([[0], np.cumsum(post_counts)]).astype(np.int64)
    pre_idx = np.ascontiguousarray(pre_idx, dtype=np.int32)
    post_idx = np.ascontiguousarray(post_idx, dtype=np.int32)
    prm = np.ascontiguousarray(params, dtype=np.float64)
    g = np.zeros((steps, n_rows))
    e32 = np.zeros(1, dtype=np.int32)
    lib().orc_plastic_run(rule, n_rows, n_cols, _p(row_ptr, C.c_int64), _p(col_idx, C.c_int32), _p(w, C.c_double),
                          _p(x, C.c_double), _p(y, C.c_double), _p(le, C.c_int64), _p(prm, C.c_double), dt, t0,
                          steps, _p(pre_ptr, C.c_int64), _p(pre_idx if pre_idx.size else e32, C.c_int32),
                          _p(post_ptr, C.c_int64), _p(post_idx if post_idx.size else e32, C.c_int32),
                          _p(pl, C.c_int64), _p(ql, C.c_int64), int(delay), _p(g, C.c_double))
    return {"w": w, "x": x, "y": y, "last_event": le, "pre_last": pl, "post_last": ql, "g": g}


def rate_plastic_run(rule, row_ptr, col_idx, w0, params, dt, r_pre0, r_pre_seq, period=1, theta0=None):
    """orc_rate_plastic_run: rule 0 Hebb (eta), 1 Oja (eta, alpha), 2 IBCM
    (eta, tau) on a rate projection into post neurons r = sum(exc), driven
    by the pre rates r_pre_seq[steps, n_cols] (r_pre0 before step 0).
    Returns w, theta, r_post and sum[steps, n_rows]."""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
    n_rows = row_ptr.size - 1
    r_pre_seq = np.ascontiguousarray(r_pre_seq, dtype=np.float64)
    steps, n_cols = r_pre_seq.shape
    w = np.array(np.broadcast_to(np.asarray(w0, dtype=np.float64), (int(row_ptr[-1]),)))
    theta = np.zeros(n_rows) if theta0 is None else np.array(theta0, dtype=np.float64)
    r0 = np.ascontiguousarray(np.broadcast_to(np.asarray(r_pre0, dtype=np.float64), (n_cols,)))
    prm = np.ascontiguousarray(params, dtype=np.float64)
    r_post = np.zeros(n_rows)
    sums = np.zeros((steps, n_rows))
    lib().orc_rate_plastic_run(rule, n_rows, n_cols, _p(row_ptr, C.c_int64), _p(col_idx, C.c_int32),
                               _p(w, C.c_double), _p(theta, C.c_double), _p(prm, C.c_double), dt, int(period),
                               steps, _p(r0, C.c_double), _p(r_pre_seq, C.c_double), _p(r_post, C.c_double),
                               _p(sums, C.c_double))
    return {"w": w, "theta": theta, "r_post": r_post, "sum": sums}
