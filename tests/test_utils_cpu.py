"""Delivered-events metric (core/Utils.py:208-280) on a hand-built projection:
counts checked against a brute-force count over the synapse list."""

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import connectivity as conn
from annarchy_amd import utils


def _proj():
    net = ah.Network(dt=1.0)
    pre = net.create(5, ah.poisson_neuron(10.0))
    post = net.create(4, ah.poisson_neuron(10.0))
    proj = net.connect(pre[1:], post, "exc")
    # rows = post ranks; columns are pre ranks of the base population
    row_ptr = np.array([0, 2, 2, 5, 6], dtype=np.int64)
    col_idx = np.array([1, 3, 1, 2, 4, 3], dtype=np.int32)
    proj.csr = conn.CSR(4, 5, row_ptr, col_idx, None, 1.0)
    return proj, row_ptr, col_idx


def test_nb_efferent_synapses():
    proj, _, _ = _proj()
    assert proj.nb_efferent_synapses() == {1: 2, 2: 1, 3: 2, 4: 1}


def test_delivered_spikes_brute_force():
    proj, row_ptr, col_idx = _proj()
    pre_ev = {0: [3], 1: [1, 5], 3: [2], 4: []}
    post_ev = {0: [1, 2, 3], 1: [4], 3: [9]}
    rows = np.repeat(np.arange(4), np.diff(row_ptr))
    eff = sum(len(pre_ev.get(int(c), [])) for c in col_idx if int(c) in (1, 2, 3, 4))
    aff = sum(len(post_ev.get(int(r), [])) for r in rows)
    assert utils.compute_delivered_efferent_spikes(proj, pre_ev) == eff == 6
    assert utils.compute_delivered_afferent_spikes(proj, post_ev) == aff == 7
    assert utils.compute_delivered_spikes(proj, pre_ev, post_ev) == 13
    assert utils.compute_delivered_spikes_per_second(proj, pre_ev, post_ev, 2.0, 1.0) == 6.5


def test_rate_projection_has_no_efferent_count():
    net = ah.Network()
    a = net.create(3, ah.RateNoisy)
    proj = net.connect(a, a, "exc")
    with pytest.raises(ValueError):
        proj.nb_efferent_synapses()
