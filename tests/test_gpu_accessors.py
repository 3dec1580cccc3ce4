"""Object accessors on the GPU: size_in_bytes of populations and projections
(the device memory they hold, bounded below by the arrays their layout
needs) and nb_efferent_synapses of a spiking projection (column lengths of
the connectivity, checked against a count over the synapse list)."""

import numpy as np
import pytest

import annarchy_amd as ah
from annarchy_amd import utils

pytestmark = pytest.mark.gpu

PoissonPre = ah.Neuron(parameters="rates = 40.0", equations="p = Uniform(0.0, 1.0) * 1000.0 / dt",
                       spike="p < rates", name="PoissonPre")
PoissonPostG = ah.Neuron(parameters="rates = 30.0", equations="p = Uniform(0.0, 1.0) * 1000.0 / dt\nmp = g_exc",
                         spike="p < rates", name="PoissonPostG")


def test_size_in_bytes_bounds(gpu):
    net = ah.Network(dt=1.0, seed=2)
    pop = net.create(1000, ah.RateNoisy)
    proj = net.connect(pop, pop, "exc").fixed_number_pre(50, 0.1, storage_format="csr")
    assert pop.size_in_bytes() == 0
    net.compile()
    local = [s for s in pop._slot_info().values() if s[1]]
    assert pop.size_in_bytes() >= len(local) * 1000 * 8
    nnz = proj.nb_synapses
    assert nnz == 50 * 1000
    assert proj.size_in_bytes() >= nnz * (8 + 4) + 1001 * 8


def test_nb_efferent_and_delivered_events(gpu):
    net = ah.Network(dt=0.5, seed=4)
    pre = net.create(300, PoissonPre)
    post = net.create(200, PoissonPostG)
    proj = net.connect(pre, post, "exc").fixed_probability(0.1, 0.2)
    m_pre = net.monitor(pre, ["spike"], capacity=200)
    m_post = net.monitor(post, ["spike"], capacity=200)
    net.compile()
    net.run(200)
    eff = proj.nb_efferent_synapses()
    col = proj.csr.col_idx
    for r in (0, 17, 299):
        assert eff.get(r, 0) == int(np.count_nonzero(col == r))
    assert sum(eff.values()) == proj.nb_synapses
    pre_ev, post_ev = m_pre.get("spike"), m_post.get("spike")
    brute = sum(len(pre_ev[int(c)]) for c in col)
    rows = np.repeat(np.arange(200), np.diff(proj.csr.row_ptr))
    brute += sum(len(post_ev[int(r)]) for r in rows)
    assert brute > 0
    assert utils.compute_delivered_spikes(proj, pre_ev, post_ev) == brute
