"""Neuron-update known answers on the GPU, from the reference's
tests_python/Neuron/test_NeuronUpdate.py: local and population-wide
equations (with random variables beside them), a population-wide parameter
in a local equation, and bounds given by population-wide parameters."""

import numpy as np
import pytest

import annarchy_amd as ah

pytestmark = pytest.mark.gpu

LocalEq = ah.Neuron(equations="""
    noise = Uniform(0,1)
    r = t
""", name="UpdLocal")
GlobalEq = ah.Neuron(equations="""
    noise = Uniform(0,1) : population
    glob_r = t : population
    r = t
""", name="UpdGlobal")
MixedEq = ah.Neuron(parameters="glob_par = 1.0: population", equations="""
    r = t + glob_par
""", name="UpdMixed")
BoundEq = ah.Neuron(parameters="""
    min_r=1.0: population
    max_r=3.0: population
""", equations="""
    r = t : min=min_r, max=max_r
""", name="UpdBound")


def test_neuron_update_kat(gpu):
    net = ah.Network(dt=1.0)
    loc = net.create(3, LocalEq)
    glo = net.create(3, GlobalEq)
    mix = net.create(3, MixedEq)
    bnd = net.create(3, BoundEq)
    mon = net.monitor(bnd, "r", capacity=5)
    net.compile()
    net.simulate(5)
    np.testing.assert_allclose(loc.r, [4.0, 4.0, 4.0])
    np.testing.assert_allclose(glo.glob_r, 4.0)
    np.testing.assert_allclose(mix.r, [5.0, 5.0, 5.0])
    np.testing.assert_allclose(mon.get("r")[:, 0], [1, 1, 2, 3, 3])
    # the draws: one per neuron (local) or one per population, in [0, 1)
    assert np.all((loc.noise >= 0.0) & (loc.noise < 1.0)) and len(set(loc.noise.tolist())) == 3
    assert 0.0 <= glo.noise < 1.0
