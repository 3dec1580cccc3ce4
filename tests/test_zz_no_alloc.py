"""ahip_mlp_train_noalloc: the MLP training entry point without device
allocations (its scratch is the library's static device array), against
ahip_mlp_train on the same inputs.

The caller owns the device buffers: this test allocates them with the HIP
runtime the library itself links (/opt/rocm), through ctypes.  Two calls in
a row, the second with fewer samples (the array keeps the first call's
contents), and a gradient-only call: parameters and gradients bit-identical,
loss curves within 1e-6 relative (per-epoch sums in FP64 on the device
against FP64 on the host, in another order).  More than 2048 samples is
rejected with AHIP_EINVAL and a message.
"""

import ctypes as C

import numpy as np
import pytest

from annarchy_amd import _lib
from annarchy_amd import selector as S

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        _lib.lib()  # the library first: the same libamdhip64 is then reused
        self.h = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 4))) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0
        return p

    def zeros(self, n, dtype=np.float32):
        return self.upload(np.zeros(n, dtype=dtype))

    def download(self, p, n, dtype=np.float32):
        out = np.empty(n, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), D2H) == 0
        return out

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.0, 1.0, (n, 16)).astype(np.float32)
    t = rng.gamma(1.0, 1.0, (n, 12)) * (1.0 + 3.0 * (rng.random((n, 12)) < 0.1))
    T = (t / t.sum(axis=1, keepdims=True)).astype(np.float32)
    return X, T, S.init_params(seed)


def _reference(X, T, P, epochs, lr=3e-3, wd=1e-4):
    n = X.shape[0]
    P = P.copy()
    curve = np.zeros(max(epochs, 1), dtype=np.float32)
    grad = np.zeros(S.N_PARAMS, dtype=np.float32)
    _lib.check(_lib.lib().ahip_mlp_train(_lib.ptr(X, C.c_float), _lib.ptr(T, C.c_float), n, _lib.ptr(P, C.c_float),
                                         epochs, lr, wd, _lib.ptr(curve, C.c_float), _lib.ptr(grad, C.c_float)))
    return P, curve, grad


def _noalloc(hip, X, T, P, epochs, lr=3e-3, wd=1e-4):
    n = X.shape[0]
    dX, dT, dP = hip.upload(X), hip.upload(T), hip.upload(P)
    dC, dG = hip.zeros(max(epochs, 1)), hip.zeros(S.N_PARAMS)
    _lib.check(_lib.lib().ahip_mlp_train_noalloc(dX, dT, n, dP, epochs, lr, wd, dC, dG, None))
    hip.sync()
    return hip.download(dP, S.N_PARAMS), hip.download(dC, max(epochs, 1)), hip.download(dG, S.N_PARAMS)


@pytest.fixture
def hip(gpu):
    h = _Hip()
    yield h
    h.free()


def test_noalloc_training_equals_allocating_entry_point(hip):
    for n, epochs, seed in ((700, 40, 1), (300, 25, 2), (2048, 3, 3), (450, 0, 4)):
        X, T, P = _problem(n, seed)
        P_ref, c_ref, g_ref = _reference(X, T, P, epochs)
        P_na, c_na, g_na = _noalloc(hip, X, T, P, epochs)
        assert np.array_equal(P_na, P_ref), (n, epochs)
        assert np.array_equal(g_na, g_ref), (n, epochs)
        np.testing.assert_allclose(c_na, c_ref, rtol=1e-6)
        if epochs:
            assert c_na[-1] < c_na[0]


def test_noalloc_rejects_more_samples_than_its_scratch(hip):
    n = 2049
    X, T, P = _problem(n, 5)
    dX, dT, dP = hip.upload(X), hip.upload(T), hip.upload(P)
    rc = _lib.lib().ahip_mlp_train_noalloc(dX, dT, n, dP, 1, 1e-3, 0.0, None, None, None)
    assert rc == -1
    with pytest.raises(_lib.AhipError, match="2048"):
        _lib.check(rc)
