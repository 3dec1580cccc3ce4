"""Rate-coded transmission patterns on the GPU in every storage format: the
known answers of the reference's tests_python/RateSynapse/
test_RateTransmission.py (3x3 populations, r = 1 into r = sum(exc) through
one_to_one, all_to_all and fixed_number_pre(3) projections of weight 0.1),
plus the sum they deliver together, 0.1 + 0.9 + 0.3."""

import numpy as np
import pytest

import annarchy_amd as ah

pytestmark = pytest.mark.gpu

One = ah.Neuron(equations="r = 1", name="RateOne")
SumExc = ah.Neuron(equations="r = sum(exc)", name="SumRate")

FORMATS = ["csr", "csr_vec", "ell", "ellr", "sell", "bsr", "dense", "csr_panel", "sell_panel"]


@pytest.mark.parametrize("fmt", FORMATS)
def test_rate_transmission_kat(gpu, fmt):
    net = ah.Network(dt=1.0, seed=1)
    pop1 = net.create((3, 3), One)
    pop2 = net.create((3, 3), SumExc)
    o2o = net.connect(pop1, pop2, "exc").one_to_one(0.1, storage_format=fmt)
    a2a = net.connect(pop1, pop2, "exc").all_to_all(0.1, storage_format=fmt)
    fnp = net.connect(pop1, pop2, "exc").fixed_number_pre(3, 0.1, storage_format=fmt)
    net.compile()
    assert o2o.dendrite(3).pre_ranks == [3]
    np.testing.assert_allclose(o2o.dendrite(3).w, [0.1])
    assert a2a.dendrite(3).pre_ranks == list(range(9))
    np.testing.assert_allclose(a2a.dendrite(3).w, np.full(9, 0.1))
    assert [d.size for d in fnp.dendrites] == [3] * 9
    net.simulate(2)   # r of pop1 is 1 from the end of step 0 on
    np.testing.assert_allclose(pop2.r.reshape(-1), 0.1 + 0.9 + 0.3, rtol=1e-15)
    np.testing.assert_allclose(np.asarray(pop2.sum("exc")).reshape(-1), 0.1 + 0.9 + 0.3, rtol=1e-15)
