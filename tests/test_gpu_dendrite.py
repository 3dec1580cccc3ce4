"""Dendrite and synapse accessors on the GPU, in several storage formats: the
known answers of the reference's tests_python/RateSynapse/test_Dendrite.py
(test_DendriteDefaultSynapse: 5 -> 8 all_to_all with force_multiple_weights)
and the psp of weights set through them."""

import numpy as np
import pytest

import annarchy_amd as ah

pytestmark = pytest.mark.gpu

IndN = ah.Neuron(parameters="tau = 10", equations="r += 1/tau * t", name="IndN")
IndN2 = ah.Neuron(parameters="tau = 10: population", equations="r += 1/tau * t: init = 1.0", name="IndN2")
One = ah.Neuron(equations="r = 1", name="RateOne")
SumExc = ah.Neuron(equations="r = sum(exc)", name="SumRate")


@pytest.mark.parametrize("fmt", ["csr", "ell", "dense", "sell_panel"])
def test_dendrite_default_synapse_kat(gpu, fmt):
    net = ah.Network(dt=1.0)
    pop1 = net.create(5, IndN)
    pop2 = net.create(8, IndN2)
    proj = net.connect(pop1, pop2, "exc").all_to_all(1.0, storage_format=fmt, force_multiple_weights=True)
    net.compile()
    assert proj.dendrite(5).pre_ranks == [0, 1, 2, 3, 4]
    assert proj.dendrite(5).size == 5
    np.testing.assert_allclose(proj.dendrite(7).w, [1.0] * 5)
    proj.dendrite(6).w = 2.0
    np.testing.assert_allclose(proj.dendrite(6).w, [2.0] * 5)
    proj.dendrite(6)[2].w = 3.0
    np.testing.assert_allclose(proj.dendrite(6).w, [2.0, 2.0, 3.0, 2.0, 2.0])
    assert proj.dendrite(6)[2].pre_rank == 2
    np.testing.assert_allclose(proj.dendrite(7).w, [1.0] * 5)


@pytest.mark.parametrize("fmt", ["csr", "ellr", "sell", "sell_panel"])
def test_weights_set_through_dendrites_reach_the_psp(gpu, fmt):
    net = ah.Network(dt=1.0)
    pre = net.create(4, One)
    post = net.create(3, SumExc)
    proj = net.connect(pre, post, "exc").all_to_all(1.0, storage_format=fmt)
    net.compile()
    proj.dendrite(1).w = 2.0
    proj.dendrite(1)[0].w = 5.0
    net.simulate(2)
    np.testing.assert_allclose(post.r, [4.0, 5.0 + 3 * 2.0, 4.0])
