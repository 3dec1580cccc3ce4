"""numpy restatement of Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and
of the device connectivity streams built on it (annarchy_amd/csrc/k_connect.hip),
used by the tests to recompute device-generated rows exactly."""

import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)
TAG_CAND = 0x434F4E4E
TAG_W = 0x57


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised over counters (arrays of uint32 values); returns 4 uint64 arrays."""
    c = [np.asarray(x, dtype=np.uint64) & _MASK for x in (c0, c1, c2, c3)]
    k0 = np.uint64(k0) & _MASK
    k1 = np.uint64(k1) & _MASK
    for _ in range(10):
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return c


def _key(seed, stream=0):
    seed = int(seed)
    return seed & 0xFFFFFFFF, ((seed >> 32) & 0xFFFFFFFF) ^ stream


def u01_53(a, b):
    v = ((np.asarray(a, dtype=np.uint64) << np.uint64(32)) | np.asarray(b, dtype=np.uint64)) >> np.uint64(11)
    return v.astype(np.float64) * (1.0 / 9007199254740992.0)


def candidates(seed, row, n, n_pre, pre_begin=0):
    """The first n candidate pre ranks of a row (fixed_number_pre stream)."""
    k = np.arange(n, dtype=np.uint64)
    k0, k1 = _key(seed)
    w = philox4x32_10(np.full(n, row), k >> np.uint64(2), np.zeros(n), np.full(n, TAG_CAND), k0, k1)
    word = np.choose((k & np.uint64(3)).astype(np.int64), w)
    return pre_begin + ((word * np.uint64(n_pre)) >> np.uint64(32)).astype(np.int64)


def fixed_number_pre_row(seed, row, n_pre, number, pre_begin=0, no_self=False):
    out, seen, k = [], set(), 0
    while len(out) < number:
        for c in candidates(seed, row, k + 4096, n_pre, pre_begin)[k:]:
            k += 1
            if no_self and c == row:
                continue
            if c not in seen:
                seen.add(int(c))
                out.append(int(c))
                if len(out) == number:
                    break
    return np.sort(np.array(out, dtype=np.int64))


def fixed_probability_row(seed, row, n_pre, p, pre_begin=0, no_self=False):
    if p >= 1.0:
        c = np.arange(pre_begin, pre_begin + n_pre)
        return c[c != row] if no_self else c
    if p <= 0.0:
        return np.zeros(0, dtype=np.int64)
    k0, k1 = _key(seed)
    lq = np.log1p(-p)
    out, j, k = [], -1, 0
    while True:
        r = philox4x32_10([row], [k], [1], [TAG_CAND], k0, k1)
        u = 1.0 - u01_53(r[0], r[1])[0]
        g = np.floor(np.log(u) / lq)
        k += 1
        if g >= n_pre:
            break
        j += int(g) + 1
        if j >= n_pre:
            break
        c = pre_begin + j
        if no_self and c == row:
            continue
        out.append(c)
    return np.array(out, dtype=np.int64)


def uniform_weights(seed, row, n, a, b):
    """Weights of synapses 0..n-1 of a row drawn as Uniform(a, b)."""
    k0, k1 = _key(seed, TAG_W)
    r = philox4x32_10(np.full(n, row), np.arange(n), np.zeros(n), np.zeros(n), k0, k1)
    return a + (b - a) * u01_53(r[0], r[1])
