"""Error paths of the C-ABI: every refused call returns its AHIP_E* code and
leaves a message in ahip_last_error(); the network stays usable afterwards.

The reference returns false from init_from_lil / the container builders and
prints to std::cerr (ProjectionGenerator.py:743-797,
include/helper_functions.hpp:22-105); Python raises ANNarchyException
(intern/Messages.py:94-112).  Here the same conditions are status codes
and Python raises AhipError.  Non-finite weights are not refused (the
reference does not check them either): they propagate into the sums exactly
as its loop would.
"""

import ctypes as C

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import _lib
from annarchy_amd import connectivity as cn

pytestmark = pytest.mark.gpu

EINVAL, ESTATE = -1, -4


def _net():
    net = ah.Network(dt=1.0, seed=1)
    pre = net.create(6, ah.Neuron(equations="r = 1", name="RateOne"))
    post = net.create(4, ah.Neuron(equations="r = sum(exc)", name="SumRate"))
    proj = net.connect(pre, post, "exc").from_csr(cn.from_arrays(4, 6, np.array([0, 1, 2, 3, 4]),
                                                                 np.array([0, 1, 2, 3], dtype=np.int32),
                                                                 np.ones(4)))
    net.compile(allow_build=False)
    return net, pre, post, proj


def _init_csr(net, proj, n_rows, n_cols, row_ptr, col, w):
    L = _lib.lib()
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    ci = np.ascontiguousarray(col, dtype=np.int32)
    ww = np.ascontiguousarray(w, dtype=np.float64)
    return L.ahip_proj_init_csr(net._h, proj.id, n_rows, n_cols, _lib.ptr(rp, C.c_int64), _lib.ptr(ci, C.c_int32),
                                _lib.ptr(ww, C.c_double), 0.0, C.byref(_lib.opts("csr", 64)))


def _msg():
    return _lib.lib().ahip_last_error().decode()


def test_invalid_csr_is_refused(gpu):
    net, pre, post, proj = _net()
    ok_rp, ok_col = [0, 1, 2, 3, 4], [0, 1, 2, 3]
    assert _init_csr(net, proj, 4, 6, [1, 1, 2, 3, 4], ok_col, np.ones(4)) == EINVAL
    assert "row_ptr[0]" in _msg()
    assert _init_csr(net, proj, 4, 6, [0, 2, 1, 3, 4], ok_col, np.ones(4)) == EINVAL
    assert "monotonic" in _msg()
    assert _init_csr(net, proj, 4, 6, ok_rp, [0, 1, 6, 3], np.ones(4)) == EINVAL
    assert "out of range" in _msg()
    assert _init_csr(net, proj, 4, 6, ok_rp, [0, -1, 2, 3], np.ones(4)) == EINVAL
    assert "out of range" in _msg()
    assert _init_csr(net, proj, 3, 6, [0, 1, 2, 3], [0, 1, 2], np.ones(3)) == EINVAL
    assert "post size" in _msg()
    assert _init_csr(net, proj, 4, 5, ok_rp, ok_col, np.ones(4)) == EINVAL
    assert "pre size" in _msg()
    # the refused calls left the projection as it was (step 2 sums the rates
    # the pre-synaptic neurons reached in step 1)
    net.run(2)
    np.testing.assert_array_equal(post.sum("exc"), [1.0, 1.0, 1.0, 1.0])


def test_bad_names_sizes_and_handles(gpu):
    net, pre, post, proj = _net()
    L = _lib.lib()
    v = np.zeros(4)
    assert L.ahip_pop_set(net._h, post.id, b"nope", _lib.ptr(v, C.c_double), 4) == EINVAL
    assert "no attribute 'nope'" in _msg()
    assert L.ahip_pop_set(net._h, post.id, b"r", _lib.ptr(v, C.c_double), 3) == EINVAL
    assert "expects 4 values" in _msg()
    assert L.ahip_pop_set(net._h, 99, b"r", _lib.ptr(v, C.c_double), 4) == EINVAL
    assert L.ahip_run(net._h, -1) == EINVAL
    assert L.ahip_net_initialize(net._h, 1.0, 0) == ESTATE
    assert L.ahip_net_set_spike_delivery(net._h, 7) == EINVAL
    pid = C.c_int32()
    assert L.ahip_proj_create(net._h, pre.id, post.id, b"inh", 0, C.byref(pid)) == EINVAL
    assert "sum(inh)" in _msg()
    bad = _lib.opts("sell_panel", 3)
    assert L.ahip_proj_set_format(net._h, proj.id, C.byref(bad)) != 0
    # still usable
    net.run(2)
    np.testing.assert_array_equal(post.sum("exc"), [1.0, 1.0, 1.0, 1.0])


def test_non_finite_weights_propagate(gpu):
    net = ah.Network(dt=1.0, seed=1)
    pre = net.create(6, ah.Neuron(equations="r = 1", name="RateOne"))
    post = net.create(4, ah.Neuron(equations="r = sum(exc)", name="SumRate"))
    w = np.array([1.0, np.nan, 2.0, np.inf, 3.0, 4.0])
    net.connect(pre, post, "exc").from_csr(cn.from_arrays(4, 6, np.array([0, 2, 3, 4, 6]),
                                                          np.array([0, 1, 2, 3, 4, 5], dtype=np.int32), w))
    net.compile(allow_build=False)
    net.run(2)
    s = post.sum("exc")
    assert np.isnan(s[0]) and s[1] == 2.0 and np.isposinf(s[2]) and s[3] == 7.0
