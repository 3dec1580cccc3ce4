"""Storage-format selection for rate-coded projections.

Three selectors, all returning one of the twelve configurations in
:data:`CLASSES`:

* :func:`heuristic_class` -- the reference's decision tree
  (core/Projection.py:691-789): density >= 0.6 -> dense, else on the GPU
  mean row length <= 128 -> sell (thread per row, slices of 32 rows, no
  sorting, CUDA/SELL.py:230-246; here SELL with one thread per row), else
  csr (one block of 64 threads per row, CUDA/CSR.py:20-42; here 64 lanes);
* :func:`autotune` -- exhaustive: time every configuration on the device
  (the judge and the fallback);
* :class:`MLPSelector` -- a 16-64-64-12 MLP on the 16 connectivity
  statistics of :func:`features`, trained on the device (k_mlp.hip) from
  autotuner timings; the soft targets are exp(-(t_c/t_best - 1)/0.05).
"""

from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

from . import connectivity as conn

CLASSES = [
    dict(fmt="sell_panel", lanes=4, sigma=4, chunk=256, panel_cols=4096),
    dict(fmt="csr", lanes=4),
    dict(fmt="csr", lanes=8),
    dict(fmt="csr", lanes=16),
    dict(fmt="csr", lanes=32),
    dict(fmt="csr", lanes=64),
    dict(fmt="sell", lanes=1, sigma=1),
    dict(fmt="sell", lanes=4, sigma=1024),
    dict(fmt="sell", lanes=16, sigma=1024),
    dict(fmt="ellr"),
    dict(fmt="bsr", chunk=4),
    # sliced tiles sized for 1e7-1e8-synapse matrices: 2048-row tiles,
    # 20479-column panels (160 KiB of LDS in FP64), 8 waves per workgroup
    dict(fmt="sell_panel", lanes=2, sigma=8, chunk=2048, panel_cols=20479),
]
HEURISTIC_CSR = 5    # csr, 64 lanes per row
HEURISTIC_SELL = 6   # sell, thread per row, no sorting
N_FEATURES = 16
DATA_DIR = Path(__file__).resolve().parent / "data"


# ---------------------------------------------------------------------------
# Features
# ---------------------------------------------------------------------------
def _tile_stats(csr: conn.CSR, rng: np.random.Generator, tile: int, max_rows: int = 4096):
    """Occupied-tile count per stored synapse on a sample of block rows."""
    n = csr.n_rows
    nbr = (n + tile - 1) // tile
    pick = np.arange(nbr) if nbr * tile <= max_rows else rng.choice(nbr, max_rows // tile, replace=False)
    rows = (pick[:, None] * tile + np.arange(tile)[None, :]).reshape(-1)
    rows = rows[rows < n]
    beg, end = csr.row_ptr[rows], csr.row_ptr[rows + 1]
    lens = end - beg
    tot = int(lens.sum())
    if tot == 0:
        return 1.0
    idx = np.repeat(beg - np.cumsum(np.concatenate([[0], lens[:-1]])), lens) + np.arange(tot)
    cols = csr.col_idx[idx]
    rr = np.repeat(rows, lens)
    keys = (rr // tile).astype(np.int64) * ((csr.n_cols + tile - 1) // tile + 1) + cols // tile
    tiles = np.unique(keys).size
    return tot / (tiles * tile * tile)


def features(csr: conn.CSR, seed: int = 0) -> np.ndarray:
    """16 statistics of the connectivity (row lengths, bandwidth, blocking)."""
    rng = np.random.default_rng(seed)
    n, m = csr.n_rows, csr.n_cols
    deg = np.diff(csr.row_ptr).astype(np.float64)
    nnz = float(deg.sum())
    mean = max(deg.mean(), 1e-9)
    f = np.zeros(N_FEATURES)
    f[0] = math.log10(n)
    f[1] = math.log10(max(nnz, 1.0))
    f[2] = math.log10(mean)
    f[3] = deg.std() / mean
    f[4] = math.log10(max(deg.max(), 1.0) / mean)
    f[5] = deg.min() / mean
    f[6] = float(np.mean(deg == 0))
    f[7] = float(np.mean(deg < 8))
    f[8] = float(np.mean(deg >= 64))
    # bandwidth on a sample of synapses
    if nnz > 0:
        k = min(int(nnz), 200000)
        sel = rng.integers(0, int(nnz), k)
        rows = np.searchsorted(csr.row_ptr, sel, side="right") - 1
        dist = np.abs(csr.col_idx[sel].astype(np.int64) - rows)
        f[9] = float(dist.mean()) / max(m, 1)
        f[10] = math.log10(float(dist.max()) + 1.0) / max(math.log10(max(m, 2)), 1e-9)
        gaps = np.diff(np.sort(csr.col_idx[csr.row_ptr[rows[0]]:csr.row_ptr[rows[0] + 1]]).astype(np.float64))
        f[13] = math.log10(float(np.median(gaps)) + 1.0) if gaps.size else 0.0
    f[11] = _tile_stats(csr, rng, 4)
    f[12] = _tile_stats(csr, rng, 8)
    # padding of unsorted 64-row slices and of ELLPACK
    C64 = 64
    pad = np.zeros(((n + C64 - 1) // C64) * C64)
    pad[:n] = deg
    f[14] = math.log10(max(float(pad.reshape(-1, C64).max(axis=1).sum() * C64), 1.0) / max(nnz, 1.0))
    f[15] = math.log10(max(float(deg.max() * np.count_nonzero(deg)), 1.0) / max(nnz, 1.0))
    return f


def heuristic_class(csr: conn.CSR) -> int | str:
    """Reference decision tree (core/Projection.py:752-787, GPU branch)."""
    density = csr.nnz / float(csr.n_rows * csr.n_cols)
    if density >= 0.6:
        return "dense"
    rows = np.count_nonzero(np.diff(csr.row_ptr))
    avg = csr.nnz / max(rows, 1)
    return HEURISTIC_SELL if avg <= 128 else HEURISTIC_CSR


def feasible(csr: conn.CSR, cls: int, max_blowup: float = 8.0) -> bool:
    """Padded formats are skipped when they would store > max_blowup x nnz."""
    cfg = CLASSES[cls]
    nnz = max(csr.nnz, 1)
    deg = np.diff(csr.row_ptr)
    if cfg["fmt"] == "ellr":
        return deg.max() * np.count_nonzero(deg) <= max_blowup * nnz
    if cfg["fmt"] == "bsr":
        t = cfg["chunk"]
        dens = _tile_stats(csr, np.random.default_rng(0), t)
        return 1.0 / max(dens, 1e-9) <= max_blowup
    if cfg["fmt"] == "sell_panel":
        # one partial sum per (panel, row): worth it only with several
        # synapses per row and panel; bounds the host-side tile tables too
        panels = -(-csr.n_cols // cfg["panel_cols"])
        cells = panels * csr.n_rows
        return cells <= 3e7 and nnz >= 4 * cells
    return True


# ---------------------------------------------------------------------------
# Corpus of connectivity matrices
# ---------------------------------------------------------------------------
KINDS = ["fixed", "binomial", "lognormal", "powerlaw", "banded", "block"]
# the large set: the in-degree laws of the 1e8-synapse rate network
LARGE_KINDS = ["fixed", "lognormal"]


def _sorted_unique_rows(n, rows, cols):
    keys = rows.astype(np.int64) * (n + 1) + cols
    keys = np.unique(keys)
    rows = (keys // (n + 1)).astype(np.int64)
    cols = (keys % (n + 1)).astype(np.int32)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=row_ptr[1:])
    return row_ptr, cols


def _generate_large(spec: dict) -> conn.CSR:
    """Seeded square connectivity of 1e7-1e8 synapses: per row `deg` uniform
    columns (with replacement; duplicates are rare at these densities),
    sorted within the row; weights N(0, 1/sqrt(k))."""
    rng = np.random.default_rng(spec["seed"])
    n, k = int(spec["n"]), float(spec["k"])
    if spec["kind"] == "fixed":
        deg = np.full(n, int(round(k)), dtype=np.int64)
    else:
        deg = np.clip(np.round(rng.lognormal(math.log(k) - 0.5, 1.0, n)), 1, n).astype(np.int64)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(deg, out=row_ptr[1:])
    keys = np.repeat(np.arange(n, dtype=np.int64) * n, deg)
    keys += rng.integers(0, n, keys.size)
    keys.sort()
    cols = (keys % n).astype(np.int32)
    del keys
    w = rng.normal(0.0, 1.0 / math.sqrt(max(k, 1.0)), cols.size)
    return conn.CSR(n, n, row_ptr, cols, w)


def generate(spec: dict) -> conn.CSR:
    """Seeded square connectivity of one corpus entry."""
    if spec.get("large"):
        return _generate_large(spec)
    rng = np.random.default_rng(spec["seed"])
    n, k, kind = int(spec["n"]), float(spec["k"]), spec["kind"]
    if kind == "block":
        b = int(spec.get("block", 8))
        nb = (n + b - 1) // b
        m = max(1, int(round(k / (b * spec.get("fill", 0.8)))))
        m = min(m, nb)
        brows = np.repeat(np.arange(nb), m)
        bcols = rng.integers(0, nb, nb * m)
        r = (brows[:, None, None] * b + np.arange(b)[None, :, None]).repeat(b, axis=2).reshape(-1)
        c = (bcols[:, None, None] * b + np.arange(b)[None, None, :]).repeat(b, axis=1).reshape(-1)
        keep = (r < n) & (c < n) & (rng.random(r.size) < spec.get("fill", 0.8))
        row_ptr, cols = _sorted_unique_rows(n, r[keep], c[keep])
    else:
        if kind == "fixed" or kind == "banded":
            deg = np.full(n, int(round(k)), dtype=np.int64)
        elif kind == "binomial":
            deg = rng.binomial(n, min(1.0, k / n), n)
        elif kind == "lognormal":
            deg = np.round(rng.lognormal(math.log(k) - 0.5, 1.0, n)).astype(np.int64)
        else:
            a = spec.get("alpha", 2.5)
            kmin = k * (a - 2.0) / (a - 1.0)
            deg = np.round(kmin * (1.0 - rng.random(n)) ** (-1.0 / (a - 1.0))).astype(np.int64)
        deg = np.clip(deg, 0, n)
        rows = np.repeat(np.arange(n), deg)
        if kind == "banded":
            bw = max(int(round(k * spec.get("spread", 2.0))), 1)
            off = rng.integers(-bw, bw + 1, rows.size)
            cols = (rows + off) % n
        else:
            cols = rng.integers(0, n, rows.size)
        row_ptr, cols = _sorted_unique_rows(n, rows, cols)
    w = rng.normal(0.0, 1.0 / math.sqrt(max(k, 1.0)), cols.size)
    return conn.CSR(n, n, row_ptr, cols, w)


def corpus_specs(count: int = 2000, seed: int = 2024, nnz_max: float = 2e6):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(round(10 ** rng.uniform(3.0, 6.0)))
        kind = KINDS[i % len(KINDS)]
        k = 10 ** rng.uniform(math.log10(4), math.log10(512))
        k = min(k, nnz_max / n, n / 4)
        k = max(k, 2.0)
        spec = dict(n=n, k=k, kind=kind, seed=int(rng.integers(1 << 31)))
        if kind == "powerlaw":
            spec["alpha"] = float(rng.uniform(2.1, 3.0))
        if kind == "block":
            spec["block"] = int(rng.choice([4, 8, 16, 32]))
            spec["fill"] = float(rng.uniform(0.5, 1.0))
        if kind == "banded":
            spec["spread"] = float(rng.uniform(0.6, 4.0))
        specs.append(spec)
    return specs


def large_specs(count: int = 40, seed: int = 4048):
    """Matrices at the scale of the rate benchmark: N log-uniform in
    [1e5, 1e6], mean in-degree log-uniform in [100, 1000], nnz = N k kept in
    [1e7, 1e8], in-degree fixed or lognormal (sigma 1)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(round(10 ** rng.uniform(5.0, 6.0)))
        k = 10 ** rng.uniform(2.0, 3.0)
        k = float(np.clip(k, 1e7 / n, 1e8 / n))
        specs.append(dict(n=n, k=k, kind=LARGE_KINDS[i % 2], seed=int(rng.integers(1 << 31)), large=True))
    return specs


# ---------------------------------------------------------------------------
# Exhaustive autotuner (device)
# ---------------------------------------------------------------------------
def autotune(csr: conn.CSR, iters: int = 10, repeats: int = 2, classes=None) -> np.ndarray:
    """Mean psp launch time (ms) of every configuration; inf where infeasible."""
    from .models import RateNoisy
    from .network import Network

    classes = range(len(CLASSES)) if classes is None else classes
    net = Network(dt=1.0, seed=0)
    pop = net.create(csr.n_rows, RateNoisy)
    proj = net.connect(pop, pop, "exc")
    proj.from_csr(csr, storage_format="csr", lanes=64)
    net.compile(allow_build=False)
    pop.r = np.random.default_rng(1).random(csr.n_rows)
    times = np.full(len(CLASSES), np.inf)
    for c in classes:
        if not feasible(csr, c):
            continue
        proj.set_format(**CLASSES[c])
        times[c] = min(proj.time_psp(iters) for _ in range(repeats))
    net.destroy()
    return times


def soft_targets(times: np.ndarray, temperature: float = 0.05) -> np.ndarray:
    t = np.asarray(times, dtype=np.float64)
    best = np.min(t, axis=1, keepdims=True)
    z = -(t / best - 1.0) / temperature
    z[~np.isfinite(t)] = -np.inf
    z -= np.max(z, axis=1, keepdims=True)
    p = np.exp(z)
    return p / p.sum(axis=1, keepdims=True)


def regret(times: np.ndarray, choice: np.ndarray) -> np.ndarray:
    t = np.asarray(times)
    best = np.min(t, axis=1)
    return t[np.arange(len(t)), choice] / best - 1.0


# ---------------------------------------------------------------------------
# MLP selector (device training / inference)
# ---------------------------------------------------------------------------
N_PARAMS = 64 * 16 + 64 + 64 * 64 + 64 + 12 * 64 + 12


def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((16, 64), (64, 64), (64, 12)):
        parts.append(rng.normal(0, math.sqrt(2.0 / fan_in), fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


class MLPSelector:
    def __init__(self, params=None, mean=None, std=None):
        self.params = init_params() if params is None else np.asarray(params, dtype=np.float32)
        self.mean = np.zeros(N_FEATURES) if mean is None else np.asarray(mean)
        self.std = np.ones(N_FEATURES) if std is None else np.asarray(std)

    def _norm(self, X):
        return np.ascontiguousarray((np.asarray(X) - self.mean) / self.std, dtype=np.float32)

    def fit(self, X, times, epochs=3000, lr=3e-3, weight_decay=1e-4, seed=0):
        from . import _lib

        X = np.asarray(X, dtype=np.float64)
        self.mean = X.mean(axis=0)
        self.std = X.std(axis=0) + 1e-6
        Xn = self._norm(X)
        T = np.ascontiguousarray(soft_targets(times), dtype=np.float32)
        self.params = init_params(seed)
        curve = np.zeros(epochs, dtype=np.float32)
        P = np.ascontiguousarray(self.params)
        rc = _lib.lib().ahip_mlp_train(_lib.ptr(Xn, C.c_float), _lib.ptr(T, C.c_float), Xn.shape[0],
                                       _lib.ptr(P, C.c_float), epochs, lr, weight_decay,
                                       _lib.ptr(curve, C.c_float), None)
        _lib.check(rc)
        self.params = P
        return curve

    def gradient(self, X, T):
        """Loss and gradient of one device forward/backward pass (no update)."""
        from . import _lib

        Xn = np.ascontiguousarray(X, dtype=np.float32)
        Tn = np.ascontiguousarray(T, dtype=np.float32)
        P = np.ascontiguousarray(self.params, dtype=np.float32).copy()
        g = np.zeros(N_PARAMS, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _lib.check(_lib.lib().ahip_mlp_train(_lib.ptr(Xn, C.c_float), _lib.ptr(Tn, C.c_float), Xn.shape[0],
                                             _lib.ptr(P, C.c_float), 0, 0.0, 0.0, _lib.ptr(loss, C.c_float),
                                             _lib.ptr(g, C.c_float)))
        return float(loss[0]), g

    def predict(self, X) -> np.ndarray:
        from . import _lib

        Xn = self._norm(np.atleast_2d(X))
        out = np.zeros(Xn.shape[0], dtype=np.int32)
        P = np.ascontiguousarray(self.params, dtype=np.float32)
        _lib.check(_lib.lib().ahip_mlp_predict(_lib.ptr(Xn, C.c_float), Xn.shape[0], _lib.ptr(P, C.c_float),
                                               _lib.ptr(out, C.c_int32), None))
        return out

    def save(self, path):
        np.savez(path, params=self.params, mean=self.mean, std=self.std)

    @classmethod
    def load(cls, path=None):
        path = DATA_DIR / "selector_mlp.npz" if path is None else path
        d = np.load(path)
        return cls(d["params"], d["mean"], d["std"])


_DEFAULT: MLPSelector | None = None   # the committed network, loaded once


def select(csr: conn.CSR, mode: str = "mlp", selector: MLPSelector | None = None) -> dict:
    """Configuration for `csr`: mode "mlp" (trained network), "exhaustive"
    (autotuner) or "heuristic" (the reference's decision tree)."""
    if mode == "heuristic":
        c = heuristic_class(csr)
        return dict(fmt="dense") if c == "dense" else CLASSES[c]
    if mode == "exhaustive":
        return CLASSES[int(np.argmin(autotune(csr)))]
    global _DEFAULT
    if selector is None:
        if _DEFAULT is None:
            _DEFAULT = MLPSelector.load()
        selector = _DEFAULT
    c = int(selector.predict(features(csr))[0])
    if not feasible(csr, c):
        c = HEURISTIC_CSR
    return CLASSES[c]


# ---------------------------------------------------------------------------
# Corpus driver: generate (CPU threads) + autotune (device)
# ---------------------------------------------------------------------------
def _gen_task(spec):
    csr = generate(spec)
    h = heuristic_class(csr)
    spec = dict(spec, heuristic=-1 if h == "dense" else int(h), nnz=int(csr.nnz))
    return spec, csr, features(csr, spec["seed"] & 0xFFFF)


def build_corpus(count: int, out_path, workers: int = 8, seed: int = 2024, log=print, tuner=None,
                 start: int = 0, stop: int | None = None, time_limit: float | None = None, specs=None):
    """Autotune entries [start, stop) of the `count`-matrix corpus (or of
    `specs`); writes specs, features and times to `out_path` (npz) every 50
    matrices and at the end.  With `time_limit` (seconds) it stops early
    after the matrix that crosses it; the file then holds the prefix done
    so far."""
    import json
    import time
    from concurrent.futures import ThreadPoolExecutor

    specs = (corpus_specs(count, seed) if specs is None else specs)[start:stop]
    feats, times, kept = [], [], []
    t0 = time.perf_counter()

    def windows():
        # generation runs ahead in a bounded window of threads (numpy
        # releases the GIL in its heavy parts); device timing stays here
        with ThreadPoolExecutor(max_workers=workers) as pool:
            for w0 in range(0, len(specs), 2 * workers):
                yield from pool.map(_gen_task, specs[w0:w0 + 2 * workers])

    def save():
        np.savez(out_path, features=np.array(feats), times=np.array(times),
                 specs=np.array([json.dumps(s) for s in kept]))

    for i, (spec, csr, f) in enumerate(windows()):
        t = tuner(csr) if tuner is not None else autotune(csr)
        feats.append(f)
        times.append(t)
        kept.append(spec)
        elapsed = time.perf_counter() - t0
        if (i + 1) % 50 == 0 or i + 1 == len(specs) or spec.get("large"):
            save()
            log(f"corpus {start + i + 1}/{start + len(specs)}: {elapsed:.0f} s")
        if time_limit is not None and elapsed > time_limit:
            save()
            log(f"corpus stopped at {start + i + 1} after {elapsed:.0f} s")
            break
    return np.array(feats), np.array(times), kept


def holdout_split(T, heuristic, seed: int = 0, groups=None):
    """Indices (train, test) of the 80/20 split over the corpus entries that
    have a finite timing and a sparse heuristic choice, split within each
    group (small corpus, large set)."""
    T, heuristic = np.asarray(T), np.asarray(heuristic)
    groups = np.zeros(len(T), dtype=int) if groups is None else np.asarray(groups)
    ok = np.isfinite(T).any(axis=1) & (heuristic >= 0)
    tr, te = [], []
    for g in np.unique(groups[ok]):
        idx = np.flatnonzero(ok & (groups == g))
        perm = idx[np.random.default_rng(seed).permutation(len(idx))]
        ntr = int(0.8 * len(idx))
        tr.append(perm[:ntr])
        te.append(perm[ntr:])
    return np.concatenate(tr), np.concatenate(te)


def with_fallback(T, choice):
    """The choice per matrix, with the reference's CSR configuration where the
    chosen one is infeasible (inf timing), as select() does."""
    T, choice = np.asarray(T), np.asarray(choice)
    ok = np.isfinite(T[np.arange(len(T)), choice])
    return np.where(ok, choice, HEURISTIC_CSR)


def _regret_stats(Tt, choice):
    rows = np.arange(len(Tt))
    best = Tt.min(axis=1)
    reg = Tt[rows, choice] / best - 1.0
    return dict(n=int(len(Tt)), mean_regret=float(np.mean(reg)), median=float(np.median(reg)),
                p90=float(np.quantile(reg, 0.9)), total_time_ratio=float(Tt[rows, choice].sum() / best.sum()))


def evaluate(X, T, heuristic, save_path=None, seeds=(0, 1, 2), epochs=3000, log=print, groups=None,
             group_names=("small", "large")):
    """80/20 split (within each group): regret of the MLP (per seed), of the
    reference heuristic and of the best single configuration on the
    held-out matrices, over all of them and per group."""
    X, T, heuristic = np.asarray(X), np.asarray(T), np.asarray(heuristic)
    groups = np.zeros(len(T), dtype=int) if groups is None else np.asarray(groups)
    tr, te = holdout_split(T, heuristic, groups=groups)
    Tt = T[te]
    res = {"n_train": int(len(tr)), "n_test": int(len(te))}
    h = with_fallback(Tt, heuristic[te])
    res["heuristic"] = _regret_stats(Tt, h)
    rel = np.where(np.isfinite(T[tr]), T[tr] / T[tr].min(axis=1, keepdims=True), np.nan)
    single = int(np.nanargmin(np.nanmean(rel, axis=0)))
    res["best_single_class"] = dict(cls=single, **_regret_stats(Tt, with_fallback(Tt, np.full(len(te), single))))
    for gi in np.unique(groups[te]):
        m = groups[te] == gi
        res[f"heuristic_{group_names[gi]}"] = _regret_stats(Tt[m], h[m])
    for seed in seeds:
        sel = MLPSelector()
        curve = sel.fit(X[tr], T[tr], epochs=epochs, lr=3e-3, seed=seed)
        pr = with_fallback(Tt, sel.predict(X[te]))
        r = _regret_stats(Tt, pr)
        r.update(loss0=float(curve[0]), loss=float(curve[-1]),
                 train_regret=float(np.mean(regret(T[tr], with_fallback(T[tr], sel.predict(X[tr]))))))
        for gi in np.unique(groups[te]):
            m = groups[te] == gi
            r[group_names[gi]] = _regret_stats(Tt[m], pr[m])
        res[f"mlp_seed{seed}"] = r
        if seed == seeds[0] and save_path is not None:
            sel.save(save_path)
        log(f"seed {seed}: {r}")
    res["label_histogram"] = np.bincount(np.argmin(T, axis=1), minlength=len(CLASSES)).tolist()
    return res


# ---------------------------------------------------------------------------
# Command line: tune slices of the corpus on the device, merge, evaluate
# ---------------------------------------------------------------------------
def _load_parts(paths):
    import json

    F, T, specs = [], [], []
    for p in paths:
        d = np.load(p)
        F.append(d["features"])
        T.append(d["times"])
        specs += [json.loads(x) for x in d["specs"]]
    return np.concatenate(F), np.concatenate(T), specs


def main(argv=None):
    """python -m annarchy_amd.selector tune --set small|large --start A --stop B --out F.npz [--time-limit S]
    python -m annarchy_amd.selector evaluate --parts F1.npz F2.npz ... [--out-dir DIR]

    tune: exhaustive timings of a slice of the 2000-matrix corpus ("small",
    N 1e3..1e6, nnz <= 2e6) or of the 40-matrix set at the rate benchmark's
    scale ("large", nnz 1e7..1e8).  evaluate: merges the parts, trains the
    MLP on 80% of each set on the device and writes the corpus, the trained
    network and the held-out regrets (selector_corpus.npz, selector_mlp.npz,
    selector_results.json) into DIR (default: the package data folder)."""
    import argparse
    import json

    ap = argparse.ArgumentParser(prog="python -m annarchy_amd.selector")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("tune")
    t.add_argument("--set", choices=["small", "large"], default="small")
    t.add_argument("--start", type=int, default=0)
    t.add_argument("--stop", type=int, default=None)
    t.add_argument("--out", required=True)
    t.add_argument("--time-limit", type=float, default=None)
    t.add_argument("--workers", type=int, default=8)
    e = sub.add_parser("evaluate")
    e.add_argument("--parts", nargs="+", required=True)
    e.add_argument("--out-dir", default=str(DATA_DIR))
    e.add_argument("--epochs", type=int, default=3000)
    a = ap.parse_args(argv)
    if a.cmd == "tune":
        specs = corpus_specs(2000) if a.set == "small" else large_specs()
        build_corpus(len(specs), a.out, workers=a.workers, start=a.start, stop=a.stop, time_limit=a.time_limit,
                     specs=specs, log=lambda m: print(m, flush=True))
        return
    F, T, specs = _load_parts(a.parts)
    heur = np.array([s["heuristic"] for s in specs])
    groups = np.array([1 if s.get("large") else 0 for s in specs])
    out = Path(a.out_dir)
    out.mkdir(parents=True, exist_ok=True)
    np.savez(out / "selector_corpus.npz", features=F, times=T, specs=np.array([json.dumps(s) for s in specs]))
    res = evaluate(F, T, heur, save_path=out / "selector_mlp.npz", epochs=a.epochs, groups=groups,
                   log=lambda m: print(m, flush=True))
    (out / "selector_results.json").write_text(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
