"""Division of the neuron equations by a population-wide value
(ahip::DivConst): the reciprocal steps of the IEEE double division once per
kernel, the quotient, residual and corrected quotient per neuron, and the
division itself outside the exponent range where the scaling steps do
nothing.  Checked bit for bit against the same division by a per-neuron
copy of the divisor (which the generated code leaves as a plain division)
and against numpy's IEEE division, over numerators across the whole double
range (signed zeros, subnormals, the 2^+-600 bounds of the short path,
infinities, NaN) and divisors inside and outside [2^-100, 2^100]."""

import numpy as np
import pytest

import annarchy_amd as ah

pytestmark = pytest.mark.gpu

DivEq = ah.Neuron(parameters="""
    c = 3.0 : population
    cl = 3.0
    x = 1.0
""", equations="""
    r = x / c
    r2 = x / cl
""", name="DivConstKAT")

DIVISORS = [3.0, 1.0 / 3.0, 0.1, -7.0, 2.0 ** -100, 2.0 ** 100, -(2.0 ** -100) * 1.5, 2.0 ** -101,
            2.0 ** 101, 1e-300, 1.7e300, 5e-324]


def _numerators(n=6000, seed=3):
    rng = np.random.default_rng(seed)
    m = rng.uniform(1.0, 2.0, n)
    e = rng.integers(-1074, 1024, n)
    x = np.ldexp(m, e) * rng.choice([-1.0, 1.0], n)
    edge = [0.0, -0.0, 5e-324, -5e-324, 2.0 ** -1022, 2.0 ** -600, -(2.0 ** -600), np.nextafter(2.0 ** -600, 0),
            2.0 ** 600, np.nextafter(2.0 ** 600, np.inf), -(2.0 ** 600), 1.7976931348623157e308, np.inf, -np.inf,
            np.nan, 1.0, -1.0, 3.0, 0.1]
    # normal-range values, the common case
    mid = rng.standard_normal(2000) * 10.0 ** rng.integers(-8, 8, 2000)
    return np.concatenate([edge, x, mid])


def _bits_equal(a, b):
    it = np.int64 if a.dtype == np.float64 else np.int32
    same = a.view(it) == b.view(it)
    return np.all(np.where(np.isnan(a) | np.isnan(b), np.isnan(a) & np.isnan(b), same))


@pytest.mark.parametrize("precision", ["double", "float"])
def test_division_by_population_value(gpu, precision):
    x = _numerators()
    net = ah.Network(dt=1.0, precision=precision)
    pops = [net.create(x.size, DivEq) for _ in DIVISORS]
    net.compile()
    dt = np.float64 if precision == "double" else np.float32
    for P, c in zip(pops, DIVISORS):
        P.x = x.astype(dt)
        P.c = c
        P.cl = np.full(x.size, c)
    net.simulate(1)
    for P, c in zip(pops, DIVISORS):
        r = np.asarray(P.r, dtype=dt).reshape(-1)
        r2 = np.asarray(P.r2, dtype=dt).reshape(-1)
        assert _bits_equal(r, r2), (precision, c)
        with np.errstate(all="ignore"):
            ref = x.astype(dt) / dt(c)
        if precision == "double":
            assert _bits_equal(r, ref), c
