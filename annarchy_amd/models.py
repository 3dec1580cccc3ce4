"""Neuron type definitions (the ``Neuron`` object of the reference API).

``Neuron`` mirrors ``ANNarchy.core.Neuron`` (parameters / equations / spike /
reset / refractory strings); the analysis into C snippets happens in
:mod:`annarchy_amd.equations`, the kernel generation in
:mod:`annarchy_amd.codegen`.  The models below are the ones of the benchmark
networks, written exactly as ANNarchy users write them.
"""

from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class Synapse:
    """Spiking synapse type (``ANNarchy.core.Synapse``): parameters,
    event-driven equations and the pre_spike / post_spike statements."""
    parameters: str = ""
    equations: str = ""
    pre_spike: str | None = None
    post_spike: str | None = None
    name: str = "synapse"
    description: str = ""


def Hebb(eta=0.01) -> Synapse:
    """Hebbian rule of the reference's models/Synapses.py:33-61."""
    return Synapse(parameters=f"eta = {float(eta)!r} : projection",
                   equations="dw/dt = eta * pre.r * post.r : min=0.0, explicit", name="Hebb")


def Oja(eta=0.01, alpha=1.0) -> Synapse:
    """Oja's rule, models/Synapses.py:63-94."""
    return Synapse(parameters=f"eta = {float(eta)!r} : projection\nalpha = {float(alpha)!r} : projection",
                   equations="dw/dt = eta * ( pre.r * post.r - alpha * post.r^2 * w ) : min=0.0, explicit",
                   name="Oja")


def IBCM(eta=0.01, tau=2000.0) -> Synapse:
    """Intrator and Cooper (1992) rule, models/Synapses.py:96-129."""
    return Synapse(parameters=f"eta = {float(eta)!r} : projection\ntau = {float(tau)!r} : projection",
                   equations="""
                       tau * dtheta/dt + theta = post.r^2 : postsynaptic, exponential
                       dw/dt = eta * post.r * (post.r - theta) * pre.r : min=0.0, explicit
                   """, name="IBCM")


def STP(tau_rec=100.0, tau_facil=0.01, U=0.5) -> Synapse:
    """Short-term facilitation and depression (Tsodyks, Uziel and Markram
    2000), models/Synapses.py:131-230."""
    if tau_facil <= 0.0:
        raise ValueError("STP: tau_facil must be positive")
    return Synapse(
        parameters=f"tau_rec = {float(tau_rec)!r} : projection\ntau_facil = {float(tau_facil)!r} : projection\n"
                   f"U = {float(U)!r}",
        equations=f"""
            dx/dt = (1 - x)/tau_rec : init = 1.0, event-driven
            du/dt = (U - u)/tau_facil : init = {float(U)!r}, event-driven
        """,
        pre_spike="""
            g_target += w * u * x
            x *= (1 - u)
            u += U * (1 - u)
        """,
        name="STP",
    )


def STDP(tau_plus=20.0, tau_minus=20.0, A_plus=0.01, A_minus=0.01, w_min=0.0, w_max=1.0) -> Synapse:
    """Spike-timing dependent plasticity with exponential traces, the
    reference's models/Synapses.py:232-310 (online additive rule, traces x
    and y integrated exactly between events)."""
    return Synapse(
        parameters=f"""
            tau_plus = {float(tau_plus)!r} : projection
            tau_minus = {float(tau_minus)!r} : projection
            A_plus = {float(A_plus)!r} : projection
            A_minus = {float(A_minus)!r} : projection
            w_min = {float(w_min)!r} : projection
            w_max = {float(w_max)!r} : projection
        """,
        equations="""
            tau_plus  * dx/dt = -x : event-driven
            tau_minus * dy/dt = -y : event-driven
        """,
        pre_spike="""
            g_target += w
            x += A_plus * w_max
            w = clip(w + y, w_min , w_max)
        """,
        post_spike="""
            y -= A_minus * w_max
            w = clip(w + x, w_min , w_max)
        """,
        name="STDP",
    )


@dataclass(frozen=True)
class Neuron:
    parameters: str = ""
    equations: str = ""
    spike: str | None = None
    reset: str | None = None
    refractory: float | str | None = None
    name: str = "neuron"
    description: str = ""

    @property
    def spiking(self) -> bool:
        return self.spike is not None


# Conductance-based LIF of the COBA benchmark (Vogels & Abbott 2005 as used by
# Brette et al. 2007; ANNarchy's COBA example): dimensionless conductances in
# units of the leak conductance, I = 20 mV background drive.
COBA = Neuron(
    parameters="""
        El = -60.0          : population
        Vr = -60.0          : population
        Erev_exc = 0.0      : population
        Erev_inh = -80.0    : population
        Vt = -50.0          : population
        tau = 20.0          : population
        tau_exc = 5.0       : population
        tau_inh = 10.0      : population
        I = 20.0            : population
    """,
    equations="""
        tau * dv/dt = (El - v) + g_exc * (Erev_exc - v) + g_inh * (Erev_inh - v ) + I

        tau_exc * dg_exc/dt = - g_exc
        tau_inh * dg_inh/dt = - g_inh
    """,
    spike="v > Vt",
    reset="v = Vr",
    refractory=5.0,
    name="COBA",
)

# Rate neuron of the rate-network benchmark:
#   tau dr/dt = -r + sum(exc) + N(0, 0.1),  r >= 0
RateNoisy = Neuron(
    parameters="""
        tau = 10.0 : population
    """,
    equations="""
        tau * dr/dt = -r + sum(exc) + Normal(0.0, 0.1) : min=0.0
    """,
    name="RateNoisy",
)

# Leaky rate neuron with a per-instance input gain, used by the batched sweep:
#   tau dr/dt = -r + sum(exc) + gain * sum(in) ,  r >= 0
RateGain = Neuron(
    parameters="""
        tau = 10.0 : population
        gain = 1.0 : population
    """,
    equations="""
        tau * dr/dt = -r + sum(exc) + gain * sum(inp) : min=0.0
    """,
    name="RateGain",
)


def poisson_neuron(rates: float) -> Neuron:
    """PoissonPopulation with a constant rate (inputs/PoissonPopulation.py:171-185)."""
    return Neuron(
        parameters=f"rates = {float(rates)!r}",
        equations="p = Uniform(0.0, 1.0) * 1000.0 / dt",
        spike="p < rates",
        name="Poisson",
    )


# Known-answer neurons of tests_python/Neuron/test_SpikingNeuron.py
CounterNeuron = Neuron(
    equations="v = v + 1.0",
    spike="v == 3.0",
    reset="v = 1.0",
    name="Counter",
)
CounterNeuronRefractory = Neuron(
    parameters="Vt = 3.0",
    equations="v = v + 1.0",
    spike="v == Vt",
    reset="v = 1.0 ",
    refractory=3.0,
    name="CounterRef",
)
