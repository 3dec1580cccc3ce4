"""The C-ABI library loads and exports every entry point its headers declare
(no device calls), and the generated model libraries export their
descriptor.  CPU only."""

import ctypes as C
import re
from pathlib import Path

import pytest

from annarchy_amd import _lib
from annarchy_amd import build as B

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: str) -> list[str]:
    text = (ROOT / "include" / header).read_text()
    return sorted(set(re.findall(r"\b(ahip_\w+)\s*\(", text)) - {"ahip_model_describe", "ahip_synapse_describe"})


def test_core_library_exports_every_declared_symbol():
    L = C.CDLL(str(B.CORE_LIB))
    declared = _declared("annarchy_hip.h")
    assert len(declared) >= 40
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert sorted(_lib.declared_symbols()) == _declared("annarchy_hip.h")


def test_model_libraries_export_descriptor():
    libs = sorted(B.LIB_DIR.glob("libahip_model_*.so"))
    assert libs, "no model library built"
    for p in libs:
        L = C.CDLL(str(p))
        assert hasattr(L, "ahip_model_describe")


class _SynDesc(C.Structure):
    """Leading fields of ahip_synapse_desc (include/annarchy_hip_synapse.h)."""
    _fields_ = [("abi", C.c_int32), ("name", C.c_char_p), ("spiking", C.c_int32), ("n_attr", C.c_int32),
                ("attr_names", C.POINTER(C.c_char_p)), ("attr_locality", C.POINTER(C.c_int32)),
                ("attr_init", C.POINTER(C.c_double)), ("n_pre_deps", C.c_int32),
                ("pre_deps", C.POINTER(C.c_char_p)), ("n_post_deps", C.c_int32),
                ("post_deps", C.POINTER(C.c_char_p)), ("has_pre", C.c_int32), ("has_post", C.c_int32),
                ("has_update", C.c_int32)]


def _describe(syn):
    from annarchy_amd.network import synapse_library
    L = C.CDLL(str(synapse_library(syn)))
    L.ahip_synapse_describe.restype = C.POINTER(_SynDesc)
    return L.ahip_synapse_describe().contents


def test_synapse_library_descriptor():
    import annarchy_amd as ah
    d = _describe(ah.STDP())
    assert d.abi == 2 and d.name == b"STDP" and d.spiking == 1 and d.n_attr == 8
    names = [d.attr_names[k].decode() for k in range(d.n_attr)]
    assert names == ["tau_plus", "tau_minus", "A_plus", "A_minus", "w_min", "w_max", "x", "y"]
    assert [d.attr_locality[k] for k in range(d.n_attr)] == [2] * 6 + [0, 0]
    assert d.attr_init[0] == 20.0 and d.has_pre == 1 and d.has_post == 1 and d.has_update == 0
    d = _describe(ah.IBCM())
    assert d.spiking == 0 and d.has_update == 1 and d.has_pre == 0 and d.has_post == 0
    assert [d.attr_locality[k] for k in range(d.n_attr)] == [2, 2, 1]
    assert d.n_pre_deps == 1 and d.pre_deps[0] == b"r" and d.n_post_deps == 1 and d.post_deps[0] == b"r"


def test_error_path_without_device_calls():
    L = _lib.lib()
    rc = L.ahip_net_initialize(None, 1.0, 0)
    assert rc == -1
    assert b"null" in L.ahip_last_error()
    rc = L.ahip_pop_set(None, 0, b"v", None, 0)
    assert rc == -1
