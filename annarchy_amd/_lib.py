"""ctypes binding of the C-ABI in include/annarchy_hip.h.

nanobind (the reference's binding, generator/NanoBind/*) is not part of this
image; the C-ABI is plain enough that ctypes is a complete binding.  The
library is loaded from annarchy_amd/lib/; a missing library is an error, never
a silent fallback.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .build import CORE_LIB

AHIP_F64 = 0
AHIP_F32 = 1
AHIP_DELIVERY_ORDERED = 0
AHIP_DELIVERY_ATOMIC = 1

FORMATS = {"csr": 0, "ell": 1, "ellr": 2, "sell": 3, "bsr": 4, "dense": 5, "csr_panel": 6, "csr_vec": 7,
           "sell_panel": 8}
FORMAT_NAMES = {v: k for k, v in FORMATS.items()}


class ConnectSpec(C.Structure):
    """ahip_connect_spec (include/annarchy_hip.h)."""
    _fields_ = [("kind", C.c_int32), ("number", C.c_int32), ("probability", C.c_double),
                ("pre_begin", C.c_int32), ("pre_end", C.c_int32), ("post_begin", C.c_int32),
                ("post_end", C.c_int32), ("allow_self", C.c_int32), ("w_kind", C.c_int32),
                ("w_a", C.c_double), ("w_b", C.c_double), ("seed", C.c_uint64)]


class AhipError(RuntimeError):
    pass


class FormatOpts(C.Structure):
    _fields_ = [
        ("format", C.c_int32),
        ("lanes_per_row", C.c_int32),
        ("chunk", C.c_int32),
        ("sigma", C.c_int32),
        ("panel_cols", C.c_int32),
    ]


_lib = None

_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_D = C.c_double
_PI32 = C.POINTER(C.c_int32)
_PI64 = C.POINTER(C.c_int64)
_PD = C.POINTER(C.c_double)
_PF = C.POINTER(C.c_float)

_SIGS = {
    "ahip_last_error": ([], C.c_char_p),
    "ahip_build_info": ([], C.c_char_p),
    "ahip_set_device": ([_I32], _I32),
    "ahip_device_count": ([_PI32], _I32),
    "ahip_probe_copy": ([_I64, _I32, _PD], _I32),
    "ahip_net_create": ([C.POINTER(_P)], _I32),
    "ahip_net_destroy": ([_P], _I32),
    "ahip_net_initialize": ([_P, _D, _I32], _I32),
    "ahip_set_seed": ([_P, C.c_uint64], _I32),
    "ahip_run": ([_P, _I64], _I32),
    "ahip_step": ([_P], _I32),
    "ahip_run_until": ([_P, _I64, _PI32, _I32, _I32, _PI64], _I32),
    "ahip_get_time": ([_P, _PI64], _I32),
    "ahip_set_time": ([_P, _I64], _I32),
    "ahip_set_graph_batch": ([_P, _I32], _I32),
    "ahip_set_graph_enabled": ([_P, _I32], _I32),
    "ahip_graph_prepare": ([_P], _I32),
    "ahip_pop_create": ([_P, _I32, C.c_char_p, _PI32], _I32),
    "ahip_pop_set": ([_P, _I32, C.c_char_p, _PD, _I64], _I32),
    "ahip_pop_get": ([_P, _I32, C.c_char_p, _PD, _I64], _I32),
    "ahip_pop_set_int": ([_P, _I32, C.c_char_p, _PI32, _I64], _I32),
    "ahip_pop_get_int": ([_P, _I32, C.c_char_p, _PI32, _I64], _I32),
    "ahip_pop_get_last_spike": ([_P, _I32, _PI64, _I64], _I32),
    "ahip_pop_spikes": ([_P, _I32, _PI32, _I32, _PI32], _I32),
    "ahip_pop_reset_spiking": ([_P, _I32], _I32),
    "ahip_pop_reset_delays": ([_P, _I32], _I32),
    "ahip_pop_compute_firing_rate": ([_P, _I32, _D], _I32),
    "ahip_pop_set_active": ([_P, _I32, _I32], _I32),
    "ahip_monitor_set_period": ([_P, _I32, _I32, _I64], _I32),
    "ahip_monitor_pause": ([_P, _I32], _I32),
    "ahip_monitor_resume": ([_P, _I32], _I32),
    "ahip_pop_set_stop_condition": ([_P, _I32, _I32], _I32),
    "ahip_net_set_constant": ([_P, C.c_char_p, _D, _PI32], _I32),
    "ahip_pop_size_in_bytes": ([_P, _I32, _PI64], _I32),
    "ahip_proj_size_in_bytes": ([_P, _I32, _PI64], _I32),
    "ahip_proj_set_transmission": ([_P, _I32, _I32], _I32),
    "ahip_pop_size": ([_P, _I32, _PI32], _I32),
    "ahip_pop_inject_random": ([_P, _I32, _I32, _PD, _I64], _I32),
    "ahip_proj_create": ([_P, _I32, _I32, C.c_char_p, _I32, _PI32], _I32),
    "ahip_proj_init_csr": ([_P, _I32, _I32, _I32, _PI64, _PI32, _PD, _D, C.POINTER(FormatOpts)], _I32),
    "ahip_proj_set_format": ([_P, _I32, C.POINTER(FormatOpts)], _I32),
    "ahip_proj_get_format": ([_P, _I32, C.POINTER(FormatOpts)], _I32),
    "ahip_proj_nb_synapses": ([_P, _I32, _PI64], _I32),
    "ahip_proj_get_w": ([_P, _I32, _PD, _I64], _I32),
    "ahip_proj_set_w": ([_P, _I32, _PD, _I64], _I32),
    "ahip_proj_set_synapse": ([_P, _I32, C.c_char_p], _I32),
    "ahip_proj_syn_get": ([_P, _I32, C.c_char_p, _PD, _I64], _I32),
    "ahip_proj_syn_set": ([_P, _I32, C.c_char_p, _PD, _I64], _I32),
    "ahip_proj_get_last_event": ([_P, _I32, _PI64, _I64], _I32),
    "ahip_proj_set_learning": ([_P, _I32, _I32, _I32, _I32, _I32], _I32),
    "ahip_time_kernel": ([_P, _I32, _I32, _I32, _PD], _I32),
    "ahip_scatter_groups": ([_P, _PI32], _I32),
    "ahip_net_set_spike_delivery": ([_P, _I32], _I32),
    "ahip_net_set_persistent": ([_P, _I32], _I32),
    "ahip_net_uses_persistent": ([_P, _PI32], _I32),
    "ahip_ordered_targets": ([_P, _PI32], _I32),
    "ahip_proj_psp_bytes": ([_P, _I32, _PD], _I32),
    "ahip_proj_time_psp": ([_P, _I32, _I32, _PD], _I32),
    "ahip_proj_set_launch": ([_P, _I32, _I32], _I32),
    "ahip_proj_set_delays": ([_P, _I32, _PI32, _I64], _I32),
    "ahip_proj_set_max_delay": ([_P, _I32, _I32], _I32),
    "ahip_proj_generate": ([_P, _I32, C.c_void_p, C.c_void_p], _I32),
    "ahip_proj_get_csr": ([_P, _I32, _PI64, _PI32, _I64], _I32),
    "ahip_proj_get_delays": ([_P, _I32, _PI32, _PI32, _PI32, _I64], _I32),
    "ahip_monitor_create": ([_P, _I32, C.c_char_p, _I64, _PI32], _I32),
    "ahip_monitor_steps": ([_P, _I32, _PI64], _I32),
    "ahip_monitor_get": ([_P, _I32, _PD, _I64], _I32),
    "ahip_monitor_get_spikes": ([_P, _I32, _PI32, _I64, _PI32, _I64, _PI64], _I32),
    "ahip_monitor_clear": ([_P, _I32], _I32),
    "ahip_batch_last_error": ([], C.c_char_p),
    "ahip_batch_create": ([_I32, _PI32, _I32, _D, _D, C.POINTER(_P)], _I32),
    "ahip_batch_destroy": ([_P], _I32),
    "ahip_batch_set_recurrent": ([_P, _PI64, _PI32, _PD, _I32], _I32),
    "ahip_batch_set_input": ([_P, _PD, _PD, _D], _I32),
    "ahip_batch_set_gain": ([_P, _PD], _I32),
    "ahip_batch_set_rates": ([_P, _PD], _I32),
    "ahip_batch_get_rates": ([_P, _PD], _I32),
    "ahip_batch_get_input_sums": ([_P, _PD], _I32),
    "ahip_batch_run": ([_P, _I64, _I32], _I32),
    "ahip_batch_time": ([_P, _I32, _I32, _PD], _I32),
    "ahip_batch_info": ([_P, _PI64, _PI32, _PD, _PD], _I32),
    "ahip_batch_chunk_steps": ([_P, _PI32], _I32),
    "ahip_batch_set_mode": ([_P, _I32], _I32),
    "ahip_batch_get_mode": ([_P, _PI32], _I32),
    "ahip_mlp_train": ([_PF, _PF, _I32, _PF, _I32, C.c_float, C.c_float, _PF, _PF], _I32),
    "ahip_mlp_train_noalloc": ([_P, _P, _I32, _P, _I32, C.c_float, C.c_float, _P, _P, _P], _I32),
    "ahip_mlp_predict": ([_PF, _I32, _PF, _PI32, _PF], _I32),
}


def declared_symbols() -> list[str]:
    return list(_SIGS)


def library_path() -> Path:
    return CORE_LIB


def lib():
    """Load libannarchy_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not path.exists():
        raise AhipError(f"{path} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
    L = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
    for name, (args, res) in _SIGS.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().ahip_last_error().decode(errors="replace")
        raise AhipError(f"annarchy_hip error {rc}: {msg}")


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def opts(fmt: str | int = "csr", lanes: int = 0, chunk: int = 0, sigma: int = 1,
         panel_cols: int = 0) -> FormatOpts:
    f = FORMATS[fmt] if isinstance(fmt, str) else int(fmt)
    return FormatOpts(f, lanes, chunk, sigma, panel_cols)
