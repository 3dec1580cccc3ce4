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


# ---------------------------------------------------------------------------
# Spike-train statistics of a spike monitor's {rank: [steps]} dictionary
# (core/Monitor.py:1092-1240)
# ---------------------------------------------------------------------------

def raster_plot(spikes: dict, dt: float = 1.0) -> tuple:
    """(times in ms, ranks) of every spike, neuron by neuron."""
    ranks = [int(n) for n, ts in spikes.items() for _ in ts]
    times = [t for ts in spikes.values() for t in ts]
    return dt * np.array(times), np.array(ranks)


def histogram(spikes: dict, bins: float | None = None, per_neuron: bool = False,
              recording_window: tuple | None = None, dt: float = 1.0) -> np.ndarray:
    """Spike counts in bins of `bins` ms (one step by default) from the first
    to the last spike, or over `recording_window` (steps); per neuron rank
    when `per_neuron`."""
    bin_step = int((dt if bins is None else bins) / dt)
    if recording_window is None:
        nonempty = [ts for ts in spikes.values() if len(ts)]
        t_min = min(min(ts) for ts in nonempty)
        t_max = max(max(ts) for ts in nonempty)
    else:
        t_min, t_max = recording_window
    nb = int((t_max - t_min) / bin_step) + 1
    if per_neuron:
        out = np.zeros((max(int(k) for k in spikes) + 1, nb), dtype=np.int64)
        for n, ts in spikes.items():
            for t in ts:
                out[int(n), int((t - t_min) / float(bin_step))] += 1
        return out
    out = np.zeros(nb, dtype=np.int64)
    for ts in spikes.values():
        for t in ts:
            out[int((t - t_min) / float(bin_step))] += 1
    return out


def inter_spike_interval(spikes: dict, ranks=None, per_neuron: bool = False, dt: float = 1.0):
    """Intervals between consecutive spikes (ms) of every neuron with at
    least two spikes: {rank: [...]} or all of them in one list."""
    isi = {}
    for n, ts in spikes.items():
        if len(ts) < 2 or (ranks is not None and n not in ranks):
            continue
        isi[n] = list(np.diff(np.asarray(ts)) * dt)
    if per_neuron:
        return isi
    return [v for vals in isi.values() for v in vals]


def coefficient_of_variation(spikes: dict, ranks=None, per_neuron: bool = False, dt: float = 1.0):
    """std / mean of the intervals of every neuron with at least two of
    them: {rank: cv} or the list of them."""
    cv = {n: float(np.std(v) / np.mean(v))
          for n, v in inter_spike_interval(spikes, ranks, True, dt).items() if len(v) >= 2}
    return cv if per_neuron else list(cv.values())
