"""Throughput of spiking projections: delivered events (core/Utils.py:208-280).

A delivered event is one synapse reached by one spike: every spike of a
pre-synaptic neuron counts its efferent synapses (the projection's column
length), every spike of a post-synaptic neuron its afferent synapses (its
row length).  The spike events are the {rank: [steps]} dictionaries of a
spike monitor.
"""

from __future__ import annotations

import numpy as np


def _ranks(obj) -> set:
    return {int(r) for r in np.asarray(obj.ranks).reshape(-1)}


def compute_delivered_efferent_spikes(proj, spike_events: dict) -> int:
    """Spikes of pre-synaptic neurons times their efferent synapses
    (core/Utils.py:223-239)."""
    eff = proj.nb_efferent_synapses()
    pre = _ranks(proj.pre)
    return int(sum(eff[r] * len(ts) for r, ts in spike_events.items() if r in pre and r in eff))


def compute_delivered_afferent_spikes(proj, spike_events: dict) -> int:
    """Spikes of post-synaptic neurons times their afferent synapses
    (core/Utils.py:242-254)."""
    post = _ranks(proj.post)
    return int(sum(proj.dendrite_size(r) * len(ts) for r, ts in spike_events.items() if r in post))


def compute_delivered_spikes(proj, pre_spike_events: dict, post_spike_events: dict) -> int:
    """Both directions (core/Utils.py:208-220)."""
    return (compute_delivered_efferent_spikes(proj, pre_spike_events)
            + compute_delivered_afferent_spikes(proj, post_spike_events))


def compute_delivered_spikes_per_second(proj, pre_spike_events: dict, post_spike_events: dict,
                                        time_in_seconds: float, scale_factor: float = 10**6) -> float:
    """Delivered events per second, divided by `scale_factor` (millions by
    default; core/Utils.py:257-277)."""
    n = compute_delivered_spikes(proj, pre_spike_events, post_spike_events)
    return (n / time_in_seconds) / scale_factor
