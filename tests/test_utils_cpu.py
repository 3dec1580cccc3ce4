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


SPIKES = {0: [1, 4, 10], 1: [2, 3], 2: [], 3: [5, 6, 7, 9]}


def test_raster_plot():
    t, r = utils.raster_plot(SPIKES, dt=0.5)
    np.testing.assert_array_equal(t, 0.5 * np.array([1, 4, 10, 2, 3, 5, 6, 7, 9]))
    np.testing.assert_array_equal(r, [0, 0, 0, 1, 1, 3, 3, 3, 3])


def test_histogram_counts_every_spike_once():
    h = utils.histogram(SPIKES)
    assert h.sum() == 9 and h.size == 10            # steps 1 .. 10
    np.testing.assert_array_equal(h, [1, 1, 1, 1, 1, 1, 1, 0, 1, 1])
    h2 = utils.histogram(SPIKES, bins=2.0)          # bins of 2 steps from step 1
    np.testing.assert_array_equal(h2, [2, 2, 2, 1, 2])
    hp = utils.histogram(SPIKES, per_neuron=True)
    assert hp.shape == (4, 10) and hp.sum() == 9 and hp[3].sum() == 4


def test_isi_and_cv_by_hand():
    isi = utils.inter_spike_interval(SPIKES, per_neuron=True, dt=0.5)
    assert isi == {0: [1.5, 3.0], 1: [0.5], 3: [0.5, 0.5, 1.0]}
    assert utils.inter_spike_interval(SPIKES, ranks=[1]) == [1]
    cv = utils.coefficient_of_variation(SPIKES, per_neuron=True)
    # neuron 0: ISIs 3, 6 -> std 1.5 / mean 4.5; neuron 3: 1, 1, 2 -> sqrt(2)/3 / (4/3)
    assert set(cv) == {0, 3}
    assert abs(cv[0] - 1.5 / 4.5) < 1e-15
    assert abs(cv[3] - (np.sqrt(2.0) / 3.0) / (4.0 / 3.0)) < 1e-15
