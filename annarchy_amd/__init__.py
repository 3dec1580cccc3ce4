"""MI355X-native simulation backend for ANNarchy networks (HIP / CDNA4).

The Python objects mirror the reference API for the simulation step
(Network, Population, Projection, Monitor, compile(), simulate()); every step
runs in hand-written HIP kernels behind the C-ABI of include/annarchy_hip.h.
"""

from .connectivity import CSR, DiscreteUniform, LogNormal, Normal, RandomDistribution, Uniform
from .models import COBA, IBCM, STDP, STP, Hebb, Neuron, Oja, RateGain, RateNoisy, Synapse, poisson_neuron
from .network import Constant, Dendrite, Monitor, Network, Population, PopulationView, Projection

__all__ = [
    "CSR", "COBA", "Constant", "Dendrite", "DiscreteUniform", "LogNormal", "Monitor", "Network", "Neuron", "Normal",
    "Population", "PopulationView", "Projection", "RandomDistribution", "RateGain", "RateNoisy",
    "Hebb", "IBCM", "Oja", "STDP", "STP", "Synapse", "Uniform", "poisson_neuron",
]
