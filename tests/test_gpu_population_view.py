"""Individual neurons and population views on the GPU: the known answers of
the reference's tests_python/Neuron/test_IndividualNeuron.py (an 8x8
population, r += t/tau) -- attribute access through the union of three
one-neuron views and the row-major rank of coordinates -- plus a view set
before compile that the simulation then carries on."""

import numpy as np
import pytest

import annarchy_amd as ah

pytestmark = pytest.mark.gpu

IndN = ah.Neuron(parameters="tau = 10", equations="r += 1/tau * t", name="IndN")


def test_individual_neurons_kat(gpu):
    net = ah.Network(dt=1.0)
    pop = net.create((8, 8), IndN)
    net.compile()
    trio = pop[2, 2] + pop[3, 3] + pop[4, 4]
    np.testing.assert_allclose(trio.r, [0.0, 0.0, 0.0])
    trio.r = 1.0
    np.testing.assert_allclose((pop[2, 2] + pop[3, 3] + pop[4, 4]).r, [1.0, 1.0, 1.0])
    assert pop[2, 4].rank == 20   # 2 * 8 + 4
    r = pop.r.reshape(-1)
    assert np.count_nonzero(r) == 3 and r[18] == r[27] == r[36] == 1.0


def test_view_values_evolve(gpu):
    net = ah.Network(dt=1.0)
    pop = net.create((8, 8), IndN)
    pop[0:8].r = np.arange(8.0)   # before compile: the initial values of the first row
    net.compile()
    net.simulate(4)                # r += t/10 for t = 0..3
    r = pop.r.reshape(8, 8)
    np.testing.assert_allclose(r[0], np.arange(8.0) + 0.6, rtol=1e-15)
    np.testing.assert_allclose(r[1:], 0.6, rtol=1e-15)
    np.testing.assert_allclose(pop[1, 0:3].r, [0.6, 0.6, 0.6], rtol=1e-15)
