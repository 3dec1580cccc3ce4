"""Python driver of the backend: the counterpart of ANNarchy's Network /
Population / Projection / Monitor objects for the simulation step.

What it reproduces of the reference caller (core/Network.py, core/Population.py,
core/Projection.py, core/Simulate.py):
  * ``Network(dt, seed)`` owns ``numpy.random.default_rng(seed)``
    (core/Network.py:147) used by the connectors in projection order;
  * ``simulate(T)`` runs ``ceil(T/dt)`` steps (core/Simulate.py:63);
  * refractory periods in ms become ``(value/dt * ones).astype(int)`` steps
    (core/Population.py:669-692);
  * a projection whose pre and post differ allows self-connections
    (core/ConnectorMethods.py:287-288, :344-345);
  * monitors record targets after compute_psp and other variables after the
    update (OpenMPBaseTemplate.py:235-305).
Everything below compile() goes through the C-ABI of libannarchy_hip.so.
"""

from __future__ import annotations

import ctypes as C
import math
import os
import weakref
from pathlib import Path

import numpy as np

from . import _lib
from . import build as _build
from . import codegen
from . import connectivity as conn
from .models import Neuron, Synapse


# -- constants (core/Constant.py:16-120, core/Network.py:482-512) ------------
_GLOBAL_CONSTANTS: dict = {}
_NETWORKS = weakref.WeakSet()


class Constant(float):
    """A named value that every neuron and synapse equation of a network may
    read; ``set()`` changes it, also after compile.  Declared at the module
    level (``Constant(name, value)``), it is copied into every network created
    afterwards and ``set()`` reaches all of them; ``net.get_constant(name)``
    (or ``net.constant(name, value)``) is the copy of one network.  A model
    attribute of the same name hides the constant."""

    def __new__(cls, name: str, value: float, net: "Network | None" = None):
        return float.__new__(cls, float(value))

    def __init__(self, name: str, value: float, net: "Network | None" = None):
        self.name = name
        self.value = float(value)
        self._net = net
        if net is None:
            _GLOBAL_CONSTANTS[name] = self
        else:
            if net.compiled and name not in net._constants:
                raise RuntimeError("constants are added before compile()")
            net._constants[name] = self

    def __str__(self):
        return str(self.value)

    __repr__ = __str__

    def set(self, value: float) -> None:
        self.value = float(value)
        if self._net is None:
            for net in list(_NETWORKS):
                local = net._constants.get(self.name)
                if local is not None:
                    local.set(value)
        elif self._net.compiled:
            _lib.check(_lib.lib().ahip_net_set_constant(self._net._h, self.name.encode(), self.value, None))


def _constant_names(net) -> list:
    return sorted(net._constants) if net is not None else sorted(_GLOBAL_CONSTANTS)


def model_library(neuron: Neuron, allow_build: bool = True, constants=None, stop_condition=None,
                  incoming_targets=None) -> Path:
    """Path of the compiled model library of `neuron` (built on demand when
    hipcc is available and `allow_build`); `constants`: the names of the
    network constants its equations may read; `stop_condition`: the
    population's stop condition, compiled into the same library."""
    src = codegen.generate_from_text(neuron.name, neuron.parameters, neuron.equations,
                                     neuron.spike, neuron.reset, neuron.refractory, constants=constants,
                                     stop_condition=stop_condition, incoming_targets=incoming_targets)
    path = _build.model_lib_path(neuron.name, src.source)
    if path.exists():
        return path
    if not allow_build or os.environ.get("AHIP_NO_BUILD"):
        raise _lib.AhipError(f"model library {path.name} for {neuron.name} is not built")
    return _build.build_model(neuron.name, src.source)


def synapse_library(synapse: Synapse, allow_build: bool = True, constants=None) -> Path:
    """Path of the compiled event-kernel library of a spiking synapse type."""
    src = codegen.generate_synapse_from_text(synapse.name, synapse.parameters, synapse.equations,
                                             synapse.pre_spike, synapse.post_spike, constants=constants)
    path = _build.model_lib_path(synapse.name, src.source, "syn")
    if path.exists():
        return path
    if not allow_build or os.environ.get("AHIP_NO_BUILD"):
        raise _lib.AhipError(f"synapse library {path.name} for {synapse.name} is not built")
    return _build.build_model(synapse.name, src.source, kind="syn")


class Population:
    def __init__(self, net: "Network", geometry, neuron: Neuron, name: str | None = None,
                 stop_condition: str | None = None):
        self.net = net
        self.stop_condition = stop_condition
        self.geometry = (geometry,) if isinstance(geometry, int) else tuple(geometry)
        self.size = int(np.prod(self.geometry))
        self.neuron = neuron
        self.name = name or f"pop{len(net.populations)}"
        self.id = -1
        self._init = {}
        self._refractory = neuron.refractory
        self.ranks = np.arange(self.size)
        self.spiking = neuron.spiking
        self._attrs = None

    # -- attribute access (Population.py:375-538) --------------------------
    def __getattr__(self, name):
        if name.startswith("_") or name in ("net", "geometry", "size", "neuron", "id", "ranks",
                                             "spiking", "name", "stop_condition"):
            raise AttributeError(name)
        return self.get(name)

    def __setattr__(self, name, value):
        if name in ("net", "geometry", "size", "neuron", "name", "id", "ranks", "spiking",
                    "stop_condition") or name.startswith("_"):
            object.__setattr__(self, name, value)
        else:
            self.set(name, value)

    def _slot_info(self):
        if self._attrs is None:
            src = codegen.generate_from_text(self.neuron.name, self.neuron.parameters,
                                             self.neuron.equations, self.neuron.spike,
                                             self.neuron.reset, self.neuron.refractory,
                                             constants=_constant_names(self.net),
                                             incoming_targets=self.net._incoming_targets(self))
            # the slots mirroring network constants are not population attributes
            self._attrs = {s[0]: s for s in src.slots if s[3] != 4}
        return self._attrs

    def set(self, name, value):
        if name == "refractory":
            self._refractory = value
            if self.net.compiled:
                self._upload_refractory()
            return
        info = self._slot_info()
        if name not in info:
            raise AttributeError(f"population {self.name} has no attribute {name}")
        if not self.net.compiled:
            self._init[name] = value
            return
        local = info[name][1]
        vals = self._materialize(value, local)
        _lib.check(_lib.lib().ahip_pop_set(self.net._h, self.id, name.encode(),
                                           _lib.ptr(vals, C.c_double), vals.size))

    def get(self, name):
        if name == "refractory":
            if not self.net.compiled:
                return self._refractory
            out = np.empty(self.size, dtype=np.int32)
            _lib.check(_lib.lib().ahip_pop_get_int(self.net._h, self.id, b"refractory",
                                                   _lib.ptr(out, C.c_int32), self.size))
            return out
        info = self._slot_info()
        if name not in info:
            raise AttributeError(f"population {self.name} has no attribute {name}")
        if not self.net.compiled:
            v = self._init.get(name, info[name][2])
            return v
        local = info[name][1]
        n = self.size if local else 1
        out = np.empty(n, dtype=np.float64)
        _lib.check(_lib.lib().ahip_pop_get(self.net._h, self.id, name.encode(),
                                           _lib.ptr(out, C.c_double), n))
        # local attributes keep the population's shape (also for one neuron)
        return out.reshape(self.geometry) if local else float(out[0])

    def size_in_bytes(self) -> int:
        """Device memory held by the population (0 before compile)."""
        if not self.net.compiled:
            return 0
        b = C.c_int64()
        _lib.check(_lib.lib().ahip_pop_size_in_bytes(self.net._h, self.id, C.byref(b)))
        return b.value

    # -- enable / disable (core/Population.py:355-375) -------------------------
    @property
    def enabled(self) -> bool:
        return getattr(self, "_enabled", True)

    def enable(self) -> None:
        """(Re-)enables the population's computations after disable()."""
        self._enabled = True
        self._upload_active()

    def disable(self) -> None:
        """Freezes the population: no neuron update, no spike emission (the
        last spike list stays as it is) and the projections into it do not
        transmit, until enable()."""
        self._enabled = False
        self._upload_active()

    def _upload_active(self):
        if self.net.compiled:
            _lib.check(_lib.lib().ahip_pop_set_active(self.net._h, self.id, 1 if self.enabled else 0))

    def compute_firing_rate(self, window: float):
        """core/Population.py:700-723: spiking neurons hold in r their mean
        firing rate over the last `window` ms (Hz), updated every step."""
        if not self.spiking:
            raise ValueError("compute_firing_rate(): the neuron is already rate-coded")
        self._fr_window = float(window)
        if self.net.compiled:
            _lib.check(_lib.lib().ahip_pop_compute_firing_rate(self.net._h, self.id, self._fr_window))

    def reset(self, attributes=None):
        """Population.reset (core/Population.py:309-343): parameters and
        variables back to their initial values (those set before compile or
        the model's init), then the spiking state and the delay queues
        (refilled with the current values)."""
        info = self._slot_info()
        names = [s[0] for s in info.values() if s[3] in (0, 1)] if attributes is None else list(attributes)
        for name in names:
            self.set(name, self._init.get(name, info[name][2]))
        if self.spiking:
            _lib.check(_lib.lib().ahip_pop_reset_spiking(self.net._h, self.id))
        _lib.check(_lib.lib().ahip_pop_reset_delays(self.net._h, self.id))

    def _materialize(self, value, local):
        n = self.size if local else 1
        if isinstance(value, conn.RandomDistribution):
            vals = np.asarray(value.get_values(n), dtype=np.float64)
        else:
            vals = np.asarray(value, dtype=np.float64).reshape(-1)
            if vals.size == 1 and n > 1:
                vals = np.full(n, vals[0])
        if vals.size != n:
            raise ValueError(f"expected {n} values")
        return np.ascontiguousarray(vals)

    def _upload_refractory(self):
        if not self.spiking:
            return
        value = self._refractory if self._refractory is not None else 0.0
        if isinstance(value, str):
            return
        dt = self.net.dt
        if isinstance(value, conn.RandomDistribution):
            refs = (value.get_values(self.size) / dt).astype(int)
        elif isinstance(value, np.ndarray):
            refs = (value / dt).astype(int).reshape(self.size)
        else:
            refs = (value / dt * np.ones(self.size)).astype(int)
        refs = np.ascontiguousarray(refs, dtype=np.int32)
        _lib.check(_lib.lib().ahip_pop_set_int(self.net._h, self.id, b"refractory",
                                               _lib.ptr(refs, C.c_int32), self.size))

    def __getitem__(self, key):
        """pop[rank], pop[list or 1-D array of ranks], pop[start:stop:step] or,
        with the population's geometry, pop[i, j, ...] with ints or slices
        per dimension (core/Population.py:782-880; ranks in row-major order)."""
        if isinstance(key, tuple):
            if len(key) != len(self.geometry):
                raise IndexError(f"{len(key)} coordinates for a {len(self.geometry)}-D population")
            axes = [np.atleast_1d(np.arange(d)[k]) for k, d in zip(key, self.geometry)]
            grid = np.meshgrid(*axes, indexing="ij")
            ranks = np.ravel_multi_index(tuple(g.reshape(-1) for g in grid), self.geometry)
            return PopulationView(self, ranks)
        if isinstance(key, (int, np.integer)):
            if not -self.size <= int(key) < self.size:
                raise IndexError(f"rank {key} out of range")
            return PopulationView(self, np.array([int(key) % self.size]))
        return PopulationView(self, np.atleast_1d(self.ranks[key]))

    def sum(self, target: str):
        """Population.sum(target) (core/Population.py:617-635): the weighted
        sum of the rate-coded projections into the population through
        `target`, the conductance g_<target> of a spiking population, and
        zeros (with a warning) for a target no projection uses."""
        name = ("g_" if self.spiking else "_sum_") + target
        if name not in self._slot_info():
            import warnings
            warnings.warn(f"population {self.name} receives no projection with target {target}")
            return np.zeros(self.geometry)
        return self.get(name)

    def rank_from_coordinates(self, coord) -> int:
        """Rank of the neuron at these coordinates (row-major)."""
        return int(np.ravel_multi_index(tuple(int(c) for c in np.atleast_1d(coord)), self.geometry))

    def coordinates_from_rank(self, rank: int) -> tuple:
        """Coordinates of the neuron of this rank (row-major)."""
        return tuple(int(c) for c in np.unravel_index(int(rank), self.geometry))

    def neuron(self, *coord) -> "PopulationView":
        """The one-neuron view of a rank or of coordinates."""
        rank = coord[0] if len(coord) == 1 else self.rank_from_coordinates(coord)
        return self[int(rank)]

    # -- spikes ---------------------------------------------------------------
    def spikes(self) -> np.ndarray:
        out = np.empty(self.size, dtype=np.int32)
        n = C.c_int32()
        _lib.check(_lib.lib().ahip_pop_spikes(self.net._h, self.id, _lib.ptr(out, C.c_int32),
                                              self.size, C.byref(n)))
        return out[: n.value].copy()

    def inject_random(self, k: int, draws: np.ndarray | None):
        """Feed rand_k from a host stream [steps, size] (parity runs)."""
        if draws is None:
            _lib.check(_lib.lib().ahip_pop_inject_random(self.net._h, self.id, k, None, 0))
            return
        d = np.ascontiguousarray(draws, dtype=np.float64)
        steps = d.size // self.size
        _lib.check(_lib.lib().ahip_pop_inject_random(self.net._h, self.id, k,
                                                     _lib.ptr(d, C.c_double), steps))


class PopulationView:
    """A subset of a population's neurons (core/PopulationView.py): usable as
    the pre or post side of a projection and for attribute access."""

    _OWN = ("population", "ranks", "size", "geometry")

    def __init__(self, pop: Population, ranks: np.ndarray):
        object.__setattr__(self, "population", pop)
        object.__setattr__(self, "ranks", np.asarray(ranks))
        object.__setattr__(self, "size", len(self.ranks))
        object.__setattr__(self, "geometry", (self.size,))

    @property
    def spiking(self):
        return self.population.spiking

    @property
    def rank(self):
        """The rank of a one-neuron view; the ranks otherwise."""
        return int(self.ranks[0]) if self.size == 1 else self.ranks

    def __len__(self):
        return self.size

    def __add__(self, other):
        """The union of two views of the same population, in ascending rank
        order (core/PopulationView.py:320-333)."""
        if not isinstance(other, PopulationView) or other.population is not self.population:
            raise ValueError("only views of the same population can be added")
        return PopulationView(self.population, np.array(sorted(set(self.ranks.tolist()) | set(other.ranks.tolist()))))

    def get(self, name):
        """The values of a local attribute for the view's neurons; a
        population-wide attribute is returned as it is."""
        val = self.population.get(name)
        if np.ndim(val) == 0:
            return val
        return np.asarray(val).reshape(-1)[self.ranks]

    def set(self, values: dict):
        """Sets local attributes of the view's neurons (one value or one per
        neuron); population-wide attributes cannot be set from a view."""
        pop = self.population
        for name, value in values.items():
            info = pop._slot_info()
            if name not in info:
                raise AttributeError(f"population {pop.name} has no attribute {name}")
            if not info[name][1]:
                raise ValueError(f"{name} is population-wide: set it on the population")
            if isinstance(value, conn.RandomDistribution):
                value = value.get_values(self.size)
            v = np.asarray(value, dtype=np.float64).reshape(-1)
            if v.size not in (1, self.size):
                raise ValueError(f"{name}: one value or {self.size} values expected")
            cur = pop.get(name)
            if isinstance(cur, conn.RandomDistribution):   # an initial distribution not drawn yet
                cur = cur.get_values(pop.size)
            cur = np.array(cur, dtype=np.float64).reshape(-1)
            if cur.size != pop.size:
                cur = np.full(pop.size, float(cur.reshape(-1)[0]))
            cur[self.ranks] = v
            pop.set(name, cur)

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        if name in self.population._slot_info():
            return self.get(name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in PopulationView._OWN or name.startswith("_"):
            object.__setattr__(self, name, value)
        elif name in self.population._slot_info():
            self.set({name: value})
        else:
            object.__setattr__(self, name, value)


def _base(obj):
    return obj.population if isinstance(obj, PopulationView) else obj


class Projection:
    _OWN = ("net", "pre", "post", "target", "id", "synapse", "csr")

    def __init__(self, net: "Network", pre, post, target: str, synapse: Synapse | None = None):
        self.net = net
        self.pre, self.post, self.target = pre, post, target
        self.id = -1
        self.synapse = synapse
        self._connector = None
        self._generator = "numpy"
        self._delays = 0.0
        self._format = _lib.opts("csr", 64)
        self._auto_format = True
        self._selected = None
        self._syn_init = {}
        self._syn_attrs = None
        self.csr: conn.CSR | None = None

    # -- synaptic attributes of a plastic projection (Projection.py:1000-1200)
    def _syn_info(self) -> dict:
        if self._syn_attrs is None:
            if self.synapse is None:
                self._syn_attrs = {}
            else:
                src = codegen.generate_synapse_from_text(self.synapse.name, self.synapse.parameters,
                                                         self.synapse.equations, self.synapse.pre_spike,
                                                         self.synapse.post_spike,
                                                         constants=_constant_names(self.net))
                consts = set(src.desc.constant_names())
                self._syn_attrs = {a[0]: a for a in src.attrs if a[0] not in consts}
        return self._syn_attrs

    def __getattr__(self, name):
        if name.startswith("_") or name in Projection._OWN:
            raise AttributeError(name)
        if name in self._syn_info():
            return self.get(name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if not name.startswith("_") and name not in Projection._OWN and name in self._syn_info():
            self.set(name, value)
        else:
            object.__setattr__(self, name, value)

    def _syn_size(self, locality: str) -> int:
        if locality == "local":
            return self.nb_synapses
        return _base(self.post).size if locality == "semiglobal" else 1

    def get(self, name):
        """A synaptic attribute: one value per synapse (forward order), per
        post-synaptic neuron (``: postsynaptic``) or, for ``: projection``,
        one float."""
        if name == "w":
            return self.w
        info = self._syn_info()
        if name not in info:
            raise AttributeError(f"the projection has no synaptic attribute {name}")
        if not self.net.compiled:
            return self._syn_init.get(name, info[name][2])
        loc = info[name][1]
        n = self._syn_size(loc)
        out = np.empty(max(n, 1))
        if n == 0:
            return out[:0]
        _lib.check(_lib.lib().ahip_proj_syn_get(self.net._h, self.id, name.encode(), _lib.ptr(out, C.c_double), n))
        return float(out[0]) if loc == "global" else out

    def set(self, name, value):
        if name == "w":
            self.w = value
            return
        info = self._syn_info()
        if name not in info:
            raise AttributeError(f"the projection has no synaptic attribute {name}")
        if not self.net.compiled:
            self._syn_init[name] = value
            return
        loc = info[name][1]
        if isinstance(value, conn.RandomDistribution):
            value = value.get_values(self._syn_size(loc))
        v = np.ascontiguousarray(np.asarray(value, dtype=np.float64).reshape(-1))
        if loc == "global" and v.size != 1:
            raise ValueError(f"{name} is one value per projection")
        if self._syn_size(loc) == 0:
            return
        _lib.check(_lib.lib().ahip_proj_syn_set(self.net._h, self.id, name.encode(), _lib.ptr(v, C.c_double),
                                                v.size))

    # -- learning flags (core/Projection.py:1440-1510) ---------------------------
    def enable_learning(self, period: float | None = None, offset: float | None = None):
        """update_synapse every `period` ms (default every step) at `offset`
        ms into the period; weights change again."""
        dt = self.net.dt
        if offset is not None and period is None:
            raise ValueError("enable_learning(): an offset needs a period")
        if period is not None and offset is not None and offset >= period:
            raise ValueError("enable_learning(): the offset must be smaller than the period")
        n = int(period / dt) if period is not None else 1
        off = 0
        if offset is not None:
            rel = (self.net.time * dt) % period + offset
            off = int(int(rel % period) / dt)
        self._learning = (1, 1, max(n, 1), off)
        self._upload_learning()

    def disable_learning(self):
        """Rate-coded: no update_synapse at all; spiking: the weights stay
        fixed while the other synaptic variables keep evolving."""
        u, p, n, off = getattr(self, "_learning", (1, 1, 1, 0))
        self._learning = (0, p, n, off) if not _base(self.pre).spiking else (u, 0, n, off)
        self._upload_learning()

    def _upload_learning(self):
        if self.net.compiled and self.synapse is not None:
            u, p, n, off = self._learning
            _lib.check(_lib.lib().ahip_proj_set_learning(self.net._h, self.id, u, p, n, off))

    # -- projection flags (core/Projection.py:1066-1100) -------------------------
    @property
    def transmission(self) -> bool:
        """False: the projection delivers nothing (no psp, no pre/post events,
        no update_synapse); its target sum reads 0."""
        return getattr(self, "_transmission", True)

    @transmission.setter
    def transmission(self, value: bool):
        self._transmission = bool(value)
        if self.net.compiled:
            _lib.check(_lib.lib().ahip_proj_set_transmission(self.net._h, self.id, 1 if self._transmission else 0))

    @property
    def update(self) -> bool:
        """update_synapse on (the _update flag)."""
        return bool(getattr(self, "_learning", (1, 1, 1, 0))[0])

    @update.setter
    def update(self, value: bool):
        u, p, n, off = getattr(self, "_learning", (1, 1, 1, 0))
        self._learning = (1 if value else 0, p, n, off)
        self._upload_learning()

    @property
    def plasticity(self) -> bool:
        """Weights change (the _plasticity flag)."""
        return bool(getattr(self, "_learning", (1, 1, 1, 0))[1])

    @plasticity.setter
    def plasticity(self, value: bool):
        u, p, n, off = getattr(self, "_learning", (1, 1, 1, 0))
        self._learning = (u, 1 if value else 0, n, off)
        self._upload_learning()

    @property
    def last_event(self) -> np.ndarray:
        """Step of the last pre or post event of every synapse (forward order)."""
        out = np.empty(self.nb_synapses, dtype=np.int64)
        _lib.check(_lib.lib().ahip_proj_get_last_event(self.net._h, self.id, _lib.ptr(out, C.c_int64), out.size))
        return out

    # -- connectors (core/ConnectorMethods.py) ---------------------------------
    def _store(self, kind, *args, storage_format=None, **fmt):
        self._connector = (kind, args)
        # force_multiple_weights: one weight per synapse even when all are equal
        self._force_multi = bool(fmt.pop("force_multiple_weights", False))
        fmt.pop("storage_order", None)   # post_to_pre is the only order here
        # no format given, or "auto": chosen when the connectivity is known
        # (core/Projection.py:691-789; _select_format)
        self._auto_format = storage_format in (None, "auto") and not fmt
        if not self._auto_format:
            self._format = _lib.opts(storage_format if storage_format not in (None, "auto") else "csr",
                                     fmt.get("lanes", 0), fmt.get("chunk", 0), fmt.get("sigma", 1),
                                     fmt.get("panel_cols", 0))
        return self

    def _select_format(self):
        """storage_format "auto" (the default): the reference's
        _automatic_format_selection (core/Projection.py:691-789) with its
        decision tree for rate-coded random connectivity replaced by the
        learned selector.  All-to-all -> dense, one-to-one -> csr as there;
        spiking projections use the CSRC view whatever the format; plastic
        projections and per-synapse delays keep csr; every other rate-coded
        pattern gets the configuration the trained MLP predicts from its
        connectivity statistics (selector.select, mode "mlp")."""
        if not getattr(self, "_auto_format", True) or self.csr is None:
            return None
        kind = self._connector[0] if self._connector else None
        if _base(self.pre).spiking or self.synapse is not None:
            return None
        if not np.isscalar(self._delays):
            return None
        if kind == "all_to_all":
            return dict(fmt="dense")
        if kind == "one_to_one" or self.csr.nnz == 0:
            return None
        if self.csr.nnz >= 0.6 * float(self.csr.n_rows) * self.csr.n_cols:
            return dict(fmt="dense")
        from . import selector
        return selector.select(self.csr, mode="mlp")

    def fixed_probability(self, probability, weights, delays=0.0, allow_self_connections=False,
                          storage_format=None, generator="numpy", **fmt):
        """generator="numpy" replays the reference's numpy calls
        (Connector.pyx:306-345); "device" draws the same distribution on the
        GPU from a counter-based stream (ahip_proj_generate)."""
        self._delays = delays
        self._generator = generator
        return self._store("fixed_probability", probability, weights, allow_self_connections,
                           storage_format=storage_format, **fmt)

    def fixed_number_pre(self, number, weights, delays=0.0, allow_self_connections=False,
                         storage_format=None, generator="numpy", **fmt):
        """generator as for fixed_probability (Connector.pyx:347-380)."""
        self._delays = delays
        self._generator = generator
        if number > _sizes(self.pre):
            raise ValueError("fixed_number_pre: the number of pre-synaptic neurons exceeds the population size")
        return self._store("fixed_number_pre", number, weights, allow_self_connections,
                           storage_format=storage_format, **fmt)

    def all_to_all(self, weights, delays=0.0, allow_self_connections=False, storage_format=None, **fmt):
        self._delays = delays
        return self._store("all_to_all", weights, allow_self_connections,
                           storage_format=storage_format, **fmt)

    def one_to_one(self, weights, delays=0.0, storage_format=None, **fmt):
        self._delays = delays
        return self._store("one_to_one", weights, storage_format=storage_format, **fmt)

    def from_csr(self, csr: conn.CSR, storage_format=None, delays=0.0, **fmt):
        self._delays = delays
        return self._store("csr", csr, storage_format=storage_format, **fmt)

    # -- delays (core/Projection.py:640-690, :1280-1412) ------------------------
    def _delay_steps(self, value) -> np.ndarray:
        """Delay(s) in ms -> steps, round(d/dt) as the reference, at least 1
        (one step is the plain transmission)."""
        dt = self.net.dt
        if isinstance(value, conn.RandomDistribution):
            value = value.get_values(self.nb_synapses)
        elif isinstance(value, (list, tuple)) and value and isinstance(value[0], (list, tuple, np.ndarray)):
            value = np.concatenate([np.asarray(v, dtype=np.float64) for v in value])
        d = np.rint(np.asarray(value, dtype=np.float64).reshape(-1) / dt).astype(np.int64)
        return np.ascontiguousarray(np.maximum(d, 1), dtype=np.int32)

    def _upload_delays(self, value):
        steps = self._delay_steps(value)
        if steps.size not in (1, self.nb_synapses):
            raise ValueError(f"expected one delay or {self.nb_synapses}")
        if isinstance(value, conn.RandomDistribution) and getattr(value, "max", None) is not None:
            # the queue holds the distribution's largest delay (core/Projection.py:661-679)
            m = max(1, int(round(float(value.max) / self.net.dt)))
            _lib.check(_lib.lib().ahip_proj_set_max_delay(self.net._h, self.id, m))
        _lib.check(_lib.lib().ahip_proj_set_delays(self.net._h, self.id, _lib.ptr(steps, C.c_int32), steps.size))

    @property
    def delay(self):
        """Uniform delay in ms, or the delay of every synapse (forward order)."""
        if not self.net.compiled:
            return self._delays
        u, m = C.c_int32(), C.c_int32()
        if self.nb_synapses == 0:
            return self.net.dt
        out = np.empty(self.nb_synapses, dtype=np.int32)
        _lib.check(_lib.lib().ahip_proj_get_delays(self.net._h, self.id, C.byref(u), C.byref(m),
                                                   _lib.ptr(out, C.c_int32), out.size))
        return u.value * self.net.dt if u.value > 0 else out * self.net.dt

    @delay.setter
    def delay(self, value):
        self._delays = value
        if self.net.compiled:
            self._upload_delays(value)

    @property
    def max_delay(self) -> int:
        u, m = C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().ahip_proj_get_delays(self.net._h, self.id, C.byref(u), C.byref(m), None, 0))
        return m.value

    def _build_csr(self) -> conn.CSR:
        kind, args = self._connector
        pre_b, post_b = _base(self.pre), _base(self.post)
        pre_r, post_r = np.asarray(self.pre.ranks), np.asarray(self.post.ranks)
        allow_self = self.pre is not self.post
        rng = self.net.rng
        if kind == "fixed_probability":
            p, w, a = args
            return conn.fixed_probability(pre_r, post_r, pre_b.size, post_b.size, rng, p, w, a or allow_self)
        if kind == "fixed_number_pre":
            n, w, a = args
            return conn.fixed_number_pre(pre_r, post_r, pre_b.size, post_b.size, rng, n, w, a or allow_self)
        if kind == "all_to_all":
            w, a = args
            return conn.all_to_all(pre_r, post_r, pre_b.size, post_b.size, w, a or allow_self)
        if kind == "one_to_one":
            (w,) = args
            return conn.one_to_one(pre_r, post_r, pre_b.size, post_b.size, w)
        if kind == "csr":
            return args[0]
        raise ValueError(kind)

    @staticmethod
    def _rank_range(obj):
        r = np.asarray(obj.ranks)
        if r.size and not np.array_equal(r, np.arange(r[0], r[0] + r.size)):
            raise ValueError("device-generated connectivity needs contiguous rank ranges")
        return (int(r[0]), int(r[0] + r.size)) if r.size else (0, 0)

    def _generate_on_device(self):
        kind, args = self._connector
        if kind not in ("fixed_number_pre", "fixed_probability"):
            raise ValueError(f"{kind} has no device generator")
        spec = _lib.ConnectSpec()
        value, w, allow = args
        if kind == "fixed_number_pre":
            spec.kind, spec.number = 0, int(value)
        else:
            spec.kind, spec.probability = 1, float(value)
        spec.pre_begin, spec.pre_end = self._rank_range(self.pre)
        spec.post_begin, spec.post_end = self._rank_range(self.post)
        spec.allow_self = int(bool(allow) or self.pre is not self.post)
        if isinstance(w, conn.Uniform):
            spec.w_kind, spec.w_a, spec.w_b = 1, float(w.min), float(w.max)
        elif isinstance(w, conn.Normal) and w.min is None and w.max is None:
            spec.w_kind, spec.w_a, spec.w_b = 2, float(w.mu), float(w.sigma)
        elif isinstance(w, conn.RandomDistribution):
            raise ValueError("device-generated weights: a constant, Uniform or Normal without bounds")
        else:
            spec.w_kind, spec.w_a = 0, float(w)
        spec.seed = int(self.net.rng.integers(1 << 63))
        L = _lib.lib()
        _lib.check(L.ahip_proj_generate(self.net._h, self.id, C.byref(spec), C.byref(self._format)))
        nnz = C.c_int64()
        _lib.check(L.ahip_proj_nb_synapses(self.net._h, self.id, C.byref(nnz)))
        post_b, pre_b = _base(self.post), _base(self.pre)
        rp = np.zeros(post_b.size + 1, dtype=np.int64)
        col = np.zeros(max(1, nnz.value), dtype=np.int32)
        _lib.check(L.ahip_proj_get_csr(self.net._h, self.id, _lib.ptr(rp, C.c_int64), _lib.ptr(col, C.c_int32),
                                       col.size))
        if spec.w_kind == 0:
            self.csr = conn.CSR(post_b.size, pre_b.size, rp, col[:nnz.value], None, w_const=spec.w_a)
        else:
            wv = np.empty(nnz.value)
            _lib.check(L.ahip_proj_get_w(self.net._h, self.id, _lib.ptr(wv, C.c_double), wv.size))
            self.csr = conn.CSR(post_b.size, pre_b.size, rp, col[:nnz.value], wv)

    def _instantiate(self):
        L = _lib.lib()
        pre_b, post_b = _base(self.pre), _base(self.post)
        spiking = pre_b.spiking
        pid = C.c_int32()
        _lib.check(L.ahip_proj_create(self.net._h, pre_b.id, post_b.id, self.target.encode(),
                                      int(spiking), C.byref(pid)))
        self.id = pid.value
        if self.synapse is not None:
            lib_path = synapse_library(self.synapse, self.net._allow_build, _constant_names(self.net))
            _lib.check(L.ahip_proj_set_synapse(self.net._h, self.id, str(lib_path).encode()))
        if self._generator == "device":
            self._generate_on_device()
            self._selected = self._select_format()
            if self._selected is not None:
                sel = self._selected
                self._format = _lib.opts(sel["fmt"], sel.get("lanes", 0), sel.get("chunk", 0), sel.get("sigma", 1),
                                         sel.get("panel_cols", 0))
                _lib.check(L.ahip_proj_set_format(self.net._h, self.id, C.byref(self._format)))
        else:
            self.csr = self._build_csr()
            if getattr(self, "_force_multi", False) and self.csr.w is None:
                self.csr.w = np.full(self.csr.nnz, float(self.csr.w_const))
            self._selected = self._select_format()
            if self._selected is not None:
                sel = self._selected
                self._format = _lib.opts(sel["fmt"], sel.get("lanes", 0), sel.get("chunk", 0), sel.get("sigma", 1),
                                         sel.get("panel_cols", 0))
            c = self.csr
            w = c.w
            _lib.check(L.ahip_proj_init_csr(self.net._h, self.id, c.n_rows, c.n_cols,
                                            _lib.ptr(c.row_ptr, C.c_int64), _lib.ptr(c.col_idx, C.c_int32),
                                            _lib.ptr(w, C.c_double) if w is not None else None,
                                            c.w_const, C.byref(self._format)))
        if not (np.isscalar(self._delays) and float(self._delays) <= self.net.dt):
            self._upload_delays(self._delays)
        for name, value in self._syn_init.items():
            self.set(name, value)
        if hasattr(self, "_learning"):
            self._upload_learning()
        if not self.transmission:
            self.transmission = False

    # -- accessors ------------------------------------------------------------
    @property
    def nb_synapses(self) -> int:
        return self.csr.nnz if self.csr is not None else 0

    # dendrites: the synapses of one post-synaptic neuron (Projection.py:
    # dendrite(), post_ranks, nb_dendrites; the wrapper's post_rank(),
    # pre_rank(i), dendrite_size(i), NanoBind/Generator.py:193-204)
    @property
    def post_ranks(self) -> list:
        """Ranks of the post-synaptic neurons that receive synapses."""
        return [int(r) for r in np.flatnonzero(np.diff(self.csr.row_ptr))]

    @property
    def nb_dendrites(self) -> int:
        return len(self.post_ranks)

    def dendrite_size(self, post_rank: int) -> int:
        return int(self.csr.row_ptr[post_rank + 1] - self.csr.row_ptr[post_rank])

    def size_in_bytes(self) -> int:
        """Device memory held by the projection (0 before compile)."""
        if not self.net.compiled:
            return 0
        b = C.c_int64()
        _lib.check(_lib.lib().ahip_proj_size_in_bytes(self.net._h, self.id, C.byref(b)))
        return b.value

    def nb_efferent_synapses(self) -> dict:
        """{pre rank: number of its synapses} for every pre-synaptic neuron
        with at least one (CSRCMatrix.hpp nb_efferent_synapses; spiking
        projections, core/Projection.py:916-930)."""
        if not _base(self.pre).spiking:
            raise ValueError("nb_efferent_synapses() is defined for spiking projections")
        if self.csr is None:
            raise RuntimeError("nb_efferent_synapses() needs the connectivity (after compile())")
        counts = np.bincount(self.csr.col_idx, minlength=self.csr.n_cols)
        return {int(r): int(counts[r]) for r in np.flatnonzero(counts)}

    def pre_ranks(self, post_rank: int) -> list:
        rp = self.csr.row_ptr
        return [int(c) for c in self.csr.col_idx[rp[post_rank]:rp[post_rank + 1]]]

    @property
    def dendrites(self) -> list:
        """The dendrites of every post-synaptic neuron that receives synapses
        (core/Projection.py dendrites)."""
        return [Dendrite(self, r) for r in self.post_ranks]

    def dendrite(self, post_rank: int) -> "Dendrite":
        if self.dendrite_size(post_rank) == 0:
            raise IndexError(f"post-synaptic neuron {post_rank} receives no synapse")
        return Dendrite(self, post_rank)

    @property
    def w(self) -> np.ndarray:
        out = np.empty(self.nb_synapses)
        _lib.check(_lib.lib().ahip_proj_get_w(self.net._h, self.id, _lib.ptr(out, C.c_double), out.size))
        return out

    @w.setter
    def w(self, value):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=np.float64), (self.nb_synapses,)))
        _lib.check(_lib.lib().ahip_proj_set_w(self.net._h, self.id, _lib.ptr(v, C.c_double), v.size))

    def set_format(self, fmt="csr", lanes=0, chunk=0, sigma=1, panel_cols=0):
        self._format = _lib.opts(fmt, lanes, chunk, sigma, panel_cols)
        self._auto_format = False
        if self.net.compiled:
            _lib.check(_lib.lib().ahip_proj_set_format(self.net._h, self.id, C.byref(self._format)))

    @property
    def selected_format(self) -> dict | None:
        """The configuration storage_format "auto" chose (None: csr kept)."""
        return self._selected

    def get_format(self) -> dict:
        o = _lib.FormatOpts()
        _lib.check(_lib.lib().ahip_proj_get_format(self.net._h, self.id, C.byref(o)))
        return {"format": _lib.FORMAT_NAMES[o.format], "lanes": o.lanes_per_row, "chunk": o.chunk,
                "sigma": o.sigma, "panel_cols": o.panel_cols}

    def psp_bytes(self) -> float:
        b = C.c_double()
        _lib.check(_lib.lib().ahip_proj_psp_bytes(self.net._h, self.id, C.byref(b)))
        return b.value

    def time_events(self, kind="pre", iters=20) -> float:
        """Mean time (ms) of the plastic projection's pre-event, update or
        post-event kernel inside `iters` ordinary steps, which advance the
        simulation."""
        k = {"pre": 0, "update": 1, "post": 2}[kind]
        return self.net.time_kernel(k, self.id, iters)

    def time_psp(self, iters=20) -> float:
        ms = C.c_double()
        _lib.check(_lib.lib().ahip_proj_time_psp(self.net._h, self.id, iters, C.byref(ms)))
        return ms.value


class Dendrite:
    """The synapses of one post-synaptic neuron (core/Dendrite.py): w and the
    synaptic attributes restricted to its row, in pre-rank order."""

    def __init__(self, proj: Projection, post_rank: int):
        object.__setattr__(self, "proj", proj)
        object.__setattr__(self, "post_rank", int(post_rank))

    @property
    def pre_ranks(self) -> list:
        return self.proj.pre_ranks(self.post_rank)

    @property
    def size(self) -> int:
        return self.proj.dendrite_size(self.post_rank)

    def _slice(self):
        rp = self.proj.csr.row_ptr
        return slice(int(rp[self.post_rank]), int(rp[self.post_rank + 1]))

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        info = self.proj._syn_info()
        if name != "w" and name not in info:
            raise AttributeError(f"the dendrite has no attribute {name}")
        loc = "local" if name == "w" else info[name][1]
        v = self.proj.get(name)
        if loc == "local":
            return np.asarray(v)[self._slice()]
        return float(np.asarray(v).reshape(-1)[self.post_rank]) if loc == "semiglobal" else v

    def __setattr__(self, name, value):
        info = self.proj._syn_info()
        if name != "w" and name not in info:
            raise AttributeError(f"the dendrite has no attribute {name}")
        loc = "local" if name == "w" else info[name][1]
        if loc == "global":
            raise AttributeError(f"{name} is one value per projection: set it on the projection")
        v = np.array(self.proj.get(name), dtype=np.float64).reshape(-1)
        if v.size == 1 and loc == "local":
            v = np.full(self.proj.nb_synapses, float(v[0]))
        if loc == "local":
            v[self._slice()] = value
        else:
            v[self.post_rank] = value
        self.proj.set(name, v)

    @property
    def rank(self) -> int:
        return self.post_rank

    def get(self, name):
        return getattr(self, name)

    def set(self, values: dict):
        for name, value in values.items():
            setattr(self, name, value)

    def __len__(self):
        return self.size

    def __getitem__(self, index: int) -> "Synapse_":
        """The synapse at position `index` of the dendrite (pre-rank order)."""
        if not -self.size <= int(index) < self.size:
            raise IndexError(f"the dendrite has {self.size} synapses")
        return Synapse_(self, int(index) % self.size)


class Synapse_:
    """One synapse of a dendrite (core/Dendrite.py IndividualSynapse): its
    per-synapse attributes by name."""

    def __init__(self, dendrite: Dendrite, index: int):
        object.__setattr__(self, "dendrite", dendrite)
        object.__setattr__(self, "index", index)

    @property
    def pre_rank(self) -> int:
        return self.dendrite.pre_ranks[self.index]

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        v = getattr(self.dendrite, name)
        return float(v[self.index]) if np.ndim(v) else v

    def __setattr__(self, name, value):
        v = np.array(getattr(self.dendrite, name), dtype=np.float64).reshape(-1)
        v[self.index] = value
        setattr(self.dendrite, name, v)


def _sizes(obj) -> int:
    return obj.size


class Monitor:
    """Device-side recording (the reference records on the host every step,
    generator/Monitor/BaseTemplates.py:5-149; here each recorded attribute has
    a device buffer of `capacity` steps filled inside the captured step)."""

    def __init__(self, net: "Network", obj, variables, capacity: int = 10000, period: float | None = None,
                 period_offset: float | None = None, start: bool = True):
        self.net = net
        self.obj = obj
        self.variables = [variables] if isinstance(variables, str) else list(variables)
        self.capacity = capacity
        self._ids = {}
        # period / period_offset in ms (core/Monitor.py:120-145); spikes are
        # recorded every step
        self.period = float(period) if period is not None else net.dt
        self.period_offset = float(period_offset) if period_offset else 0.0
        if self.period_offset >= self.period:
            raise ValueError("Monitor(): period_offset must be smaller than period")
        self._recording = bool(start)
        # spikes: (first row, its step) of every recorded segment
        self._segments = []

    def _instantiate(self):
        pop = _base(self.obj)
        L = _lib.lib()
        period = max(1, int(round(self.period / self.net.dt)))
        # recorded steps: (t - offset) mod period == period_offset with
        # offset = the current step mod period (core/Monitor.py:318-332)
        phase = self.net.get_current_step() % period + int(self.period_offset / self.net.dt)
        for v in self.variables:
            mid = C.c_int32()
            _lib.check(L.ahip_monitor_create(self.net._h, pop.id, v.encode(), self.capacity, C.byref(mid)))
            self._ids[v] = mid.value
            if v != "spike" and (period != 1 or phase != 0):
                _lib.check(L.ahip_monitor_set_period(self.net._h, mid.value, period, phase))
            if not self._recording:
                _lib.check(L.ahip_monitor_pause(self.net._h, mid.value))
        self._segments = [(0, self.net.get_current_step())] if self._recording else []

    def pause(self):
        """Stops recording until resume() (core/Monitor.py:436-471)."""
        self._recording = False
        if self.net.compiled:
            for mid in self._ids.values():
                _lib.check(_lib.lib().ahip_monitor_pause(self.net._h, mid))

    def resume(self):
        """Records again, appending to what was recorded before."""
        if self.net.compiled and not self._recording:
            for v, mid in self._ids.items():
                if v == "spike":
                    self._segments.append((self._steps(mid), self.net.get_current_step()))
                _lib.check(_lib.lib().ahip_monitor_resume(self.net._h, mid))
        self._recording = True

    def start(self):
        """Starts a monitor created with start=False."""
        self.resume()

    def clear(self):
        """Empty every recorded buffer; recording restarts at the current step."""
        for mid in self._ids.values():
            _lib.check(_lib.lib().ahip_monitor_clear(self.net._h, mid))
        self._segments = [(0, self.net.get_current_step())] if self._recording else []

    def _row_steps(self, rows: int) -> np.ndarray:
        """The simulation step of every recorded spike row."""
        out = np.empty(rows, dtype=np.int64)
        for k, (r0, t0) in enumerate(self._segments):
            r1 = self._segments[k + 1][0] if k + 1 < len(self._segments) else rows
            out[r0:r1] = t0 + np.arange(max(0, min(r1, rows) - r0))
        return out

    def _steps(self, mid) -> int:
        s = C.c_int64()
        _lib.check(_lib.lib().ahip_monitor_steps(self.net._h, mid, C.byref(s)))
        return s.value

    def get(self, name: str, keep: bool = False):
        """Values [steps, size] or, for "spike", {rank: [time steps]}."""
        mid = self._ids[name]
        steps = self._steps(mid)
        pop = _base(self.obj)
        if name == "spike":
            counts, idx = self.raw_spikes()
            row_step = self._row_steps(steps)
            out = {int(r): [] for r in np.asarray(self.obj.ranks)}
            pos = 0
            for s in range(steps):
                for r in idx[pos: pos + counts[s]]:
                    if int(r) in out:
                        out[int(r)].append(int(row_step[s]))
                pos += counts[s]
            if not keep:
                _lib.check(_lib.lib().ahip_monitor_clear(self.net._h, mid))
                self._segments = [(0, self.net.get_current_step())] if self._recording else []
            return out
        out = np.empty(steps * pop.size)
        _lib.check(_lib.lib().ahip_monitor_get(self.net._h, mid, _lib.ptr(out, C.c_double), out.size))
        if not keep:
            _lib.check(_lib.lib().ahip_monitor_clear(self.net._h, mid))
        out = out.reshape(steps, pop.size)
        if isinstance(self.obj, PopulationView):
            out = out[:, self.obj.ranks]
        return out

    def raw_spikes(self):
        mid = self._ids["spike"]
        steps = self._steps(mid)
        counts = np.zeros(steps, dtype=np.int32)
        total = C.c_int64()
        _lib.check(_lib.lib().ahip_monitor_get_spikes(self.net._h, mid, _lib.ptr(counts, C.c_int32),
                                                      steps, None, 0, C.byref(total)))
        idx = np.zeros(max(1, total.value), dtype=np.int32)
        _lib.check(_lib.lib().ahip_monitor_get_spikes(self.net._h, mid, _lib.ptr(counts, C.c_int32),
                                                      steps, _lib.ptr(idx, C.c_int32), idx.size,
                                                      C.byref(total)))
        return counts, idx[: total.value]

    # -- analysis (core/Monitor.py:712-1330) ------------------------------------
    def _spike_data(self, spikes):
        if spikes is None:
            return self.get("spike", keep=True)
        return spikes["spike"] if "spike" in spikes else spikes

    def raster_plot(self, spikes: dict | None = None) -> tuple:
        from . import utils
        return utils.raster_plot(self._spike_data(spikes), self.net.dt)

    def histogram(self, spikes: dict | None = None, bins=None, per_neuron=False, recording_window=None):
        from . import utils
        return utils.histogram(self._spike_data(spikes), bins, per_neuron, recording_window, self.net.dt)

    def inter_spike_interval(self, spikes: dict | None = None, ranks=None, per_neuron: bool = False):
        from . import utils
        return utils.inter_spike_interval(self._spike_data(spikes), ranks, per_neuron, self.net.dt)

    def coefficient_of_variation(self, spikes: dict | None = None, ranks=None):
        from . import utils
        return utils.coefficient_of_variation(self._spike_data(spikes), ranks, False, self.net.dt)

    @staticmethod
    def mean_fr(spikes: dict, duration_ms: float) -> float:
        n = sum(len(v) for v in spikes.values())
        return n / len(spikes) / (duration_ms / 1000.0) if spikes else 0.0



class Network:
    def __init__(self, dt: float = 1.0, seed: int | None = None, precision: str = "double",
                 spike_delivery: str = "ordered"):
        """spike_delivery: "ordered" (default) adds the events of spiking
        projections with per-synapse weights in the reference's order
        (projection by projection, ascending pre-synaptic rank,
        generator/Projection/OpenMP/CSR.py:870-885), so conductances are
        bit-identical to it; "atomic" adds them with device atomics in
        arbitrary order (faster; one rounding of difference per event)."""
        if spike_delivery not in ("ordered", "atomic"):
            raise ValueError("spike_delivery must be 'ordered' or 'atomic'")
        self.dt = float(dt)
        self.seed = seed
        self.rng = np.random.default_rng(seed)
        self.precision = precision
        self.spike_delivery = spike_delivery
        self.populations: list[Population] = []
        self.projections: list[Projection] = []
        self.monitors: list[Monitor] = []
        self.compiled = False
        self._h = None
        # the module-level constants declared so far, copied (core/Network.py:502-512)
        self._constants = {}
        for c in list(_GLOBAL_CONSTANTS.values()):
            Constant(c.name, c.value, net=self)
        _NETWORKS.add(self)

    def _incoming_targets(self, pop) -> list:
        """Targets of the rate-coded projections into `pop`: the population
        holds a sum for each, read by its equations or not."""
        return sorted({pr.target for pr in self.projections
                       if _base(pr.post) is pop and not _base(pr.pre).spiking})

    def constant(self, name: str, value: float) -> Constant:
        """A constant of this network only (core/Network.py:482-500)."""
        return Constant(name, value, net=self)

    def get_constant(self, name: str) -> Constant:
        """This network's copy of the constant `name`."""
        if name not in self._constants:
            raise KeyError(f"no constant named {name}")
        return self._constants[name]

    def create(self, geometry, neuron: Neuron, name: str | None = None,
               stop_condition: str | None = None) -> Population:
        """Population(geometry, neuron, name, stop_condition): `stop_condition`
        ("expr : any|all", or "spiked : any|all") ends simulate_until()."""
        if self.compiled:
            raise RuntimeError("the network is already compiled")
        p = Population(self, geometry, neuron, name, stop_condition)
        self.populations.append(p)
        return p

    def connect(self, pre, post, target: str, synapse: Synapse | None = None) -> Projection:
        """Projection(pre, post, target, synapse): `synapse` is a spiking
        synapse type with pre_spike / post_spike events (plasticity)."""
        if self.compiled:
            raise RuntimeError("the network is already compiled")
        p = Projection(self, pre, post, target, synapse)
        self.projections.append(p)
        # a new target adds a sum slot to the post population
        _base(post)._attrs = None
        return p

    def monitor(self, obj, variables, capacity: int = 10000, period: float | None = None,
                period_offset: float | None = None, start: bool = True) -> Monitor:
        m = Monitor(self, obj, variables, capacity, period, period_offset, start)
        self.monitors.append(m)
        if self.compiled:
            m._instantiate()
        return m

    def compile(self, allow_build: bool = True, graph_batch: int = 0):
        """Build (or look up) the model libraries, create the device network,
        run the connectors and upload everything (Compiler._instantiate,
        generator/Compiler.py:796-1000)."""
        L = _lib.lib()
        self._allow_build = allow_build
        h = C.c_void_p()
        _lib.check(L.ahip_net_create(C.byref(h)))
        self._h = h
        prec = _lib.AHIP_F64 if self.precision in ("double", "float64", "f64") else _lib.AHIP_F32
        _lib.check(L.ahip_net_initialize(h, self.dt, prec))
        seed = self.seed if self.seed is not None else int(np.random.SeedSequence().entropy % (2**63))
        _lib.check(L.ahip_set_seed(h, C.c_uint64(seed)))
        if graph_batch:
            _lib.check(L.ahip_set_graph_batch(h, graph_batch))
        _lib.check(L.ahip_net_set_spike_delivery(
            h, _lib.AHIP_DELIVERY_ORDERED if self.spike_delivery == "ordered" else _lib.AHIP_DELIVERY_ATOMIC))
        for p in self.populations:
            p._attrs = None
            lib_path = model_library(p.neuron, allow_build, _constant_names(self), p.stop_condition,
                                     self._incoming_targets(p))
            pid = C.c_int32()
            _lib.check(L.ahip_pop_create(h, p.size, str(lib_path).encode(), C.byref(pid)))
            p.id = pid.value
        self.compiled = True
        # connectors first, then attribute initial values: the order in which
        # the reference draws from the shared generator (Compiler.py:796-1000)
        for proj in self.projections:
            proj._instantiate()
        for p in self.populations:
            for name, value in p._init.items():
                p.set(name, value)
            p._upload_refractory()
            if getattr(p, "_fr_window", None):
                p.compute_firing_rate(p._fr_window)
            if not p.enabled:
                p._upload_active()
            if p.stop_condition:
                cond, mode = codegen.parse_stop_condition(p.stop_condition)
                code = (1 if mode == "any" else 2) if cond is not None else (3 if mode == "any" else 4)
                _lib.check(L.ahip_pop_set_stop_condition(h, p.id, code))
        # constants: the libraries hold 0, the network writes the values
        for c in self._constants.values():
            _lib.check(L.ahip_net_set_constant(h, c.name.encode(), c.value, None))
        for m in self.monitors:
            m._instantiate()
        return self

    # -- simulation (core/Simulate.py) ----------------------------------------
    def simulate(self, duration: float):
        nb_steps = math.ceil(duration / self.dt)
        self.run(nb_steps)

    def run(self, nsteps: int):
        _lib.check(_lib.lib().ahip_run(self._h, int(nsteps)))

    def simulate_until(self, max_duration: float, population, operator: str = "and") -> float:
        """core/Simulate.py:173-229: runs at most ceil(max_duration/dt) steps
        and stops after the first step at which the stop conditions of the
        given population(s), combined with `operator` ('and' / 'or'), hold.
        Returns the number of steps divided by dt, as the reference does."""
        nb_steps = math.ceil(float(max_duration) / self.dt)
        pops = population if isinstance(population, list) else [population]
        ids = np.array([_base(p).id for p in pops], dtype=np.int32)
        done = C.c_int64()
        _lib.check(_lib.lib().ahip_run_until(self._h, nb_steps, _lib.ptr(ids, C.c_int32), ids.size,
                                             1 if operator == "and" else 0, C.byref(done)))
        return float(done.value) / self.dt

    def step(self):
        _lib.check(_lib.lib().ahip_step(self._h))

    @property
    def time(self) -> int:
        t = C.c_int64()
        _lib.check(_lib.lib().ahip_get_time(self._h, C.byref(t)))
        return t.value

    def get_current_step(self) -> int:
        return self.time

    def get_time(self) -> float:
        return self.time * self.dt

    def reset(self, populations: bool = True, projections: bool = False, monitors: bool = True):
        """Network.reset (core/Network.py:628-670): time 0, populations back
        to their initial values with their spiking state and delay queues,
        monitors emptied.  Projections keep their weights and synaptic
        variables (the reference's default projections=False)."""
        _lib.check(_lib.lib().ahip_set_time(self._h, 0))
        if populations:
            for p in self.populations:
                p.reset()
        if monitors:
            for m in self.monitors:
                m.clear()

    def time_kernel(self, what: int, index: int, iters: int = 20) -> float:
        """Mean time (ms) of one kernel inside `iters` ordinary eager steps
        (ahip_time_kernel; the steps advance the simulation)."""
        ms = C.c_double()
        _lib.check(_lib.lib().ahip_time_kernel(self._h, int(what), int(index), int(iters), C.byref(ms)))
        return ms.value

    def scatter_groups(self) -> int:
        n = C.c_int32()
        _lib.check(_lib.lib().ahip_scatter_groups(self._h, C.byref(n)))
        return n.value

    def ordered_targets(self) -> int:
        """Number of conductances delivered in the reference's order (one
        deliver kernel each per step)."""
        n = C.c_int32()
        _lib.check(_lib.lib().ahip_ordered_targets(self._h, C.byref(n)))
        return n.value

    def enable_learning(self, projections=None, period=None, offset=None):
        """Network.enable_learning: every (or the given) plastic projection."""
        for p in projections or self.projections:
            if p.synapse is not None:
                p.enable_learning(period, offset)

    def disable_learning(self, projections=None):
        for p in projections or self.projections:
            if p.synapse is not None:
                p.disable_learning()

    def set_graph(self, enabled: bool = True, batch: int = 0):
        _lib.check(_lib.lib().ahip_set_graph_enabled(self._h, int(enabled)))
        _lib.check(_lib.lib().ahip_set_graph_batch(self._h, int(batch)))

    def set_persistent(self, enabled: bool = True):
        """Persistent steps of a small single-population spiking network (one
        launch of one workgroup per batch of steps; on by default)."""
        _lib.check(_lib.lib().ahip_net_set_persistent(self._h, int(enabled)))

    def uses_persistent(self) -> bool:
        """Whether the next run() takes the persistent steps."""
        y = C.c_int32()
        _lib.check(_lib.lib().ahip_net_uses_persistent(self._h, C.byref(y)))
        return bool(y.value)

    def prepare_graph(self):
        """Capture, instantiate and upload the step graph now (simulates
        nothing), so that the next run() only replays it."""
        _lib.check(_lib.lib().ahip_graph_prepare(self._h))

    def destroy(self):
        """Release the device network now (instead of at garbage collection)."""
        if self._h is not None and _lib._lib is not None:
            _lib.lib().ahip_net_destroy(self._h)
        self._h = None
        self.compiled = False

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
