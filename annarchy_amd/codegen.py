"""Splice analysed neuron equations into the CDNA4 update-kernel template.

The reference emits, per population, a CPU loop
(``generator/Population/OpenMPGenerator.py:765-1237``) or a CUDA pair of
kernels (``generator/Population/CUDATemplates.py:543-720``: update, then a
separate spike_gather that appends with atomics).  Here one kernel per model
does the whole per-neuron work of a step, one thread per neuron, 256 threads
per workgroup (4 waves):

    in_ref = refractory_remaining[i] > 0 ? 0 : 1         (if refractory)
    random draws (Philox, or the injected stream)
    ODE increments, switches (masked), bounds            (spliced text)
    refractory_remaining[i] -= (1 - in_ref)
    if (in_ref && <spike condition>) { reset; last_spike[i] = t;
                                       refractory_remaining[i] = refractory[i]; }
    wave ballot -> spike bit mask, per-block spike count

and a compaction kernel of the runtime turns the bit mask into the ascending
spike list.  Population-wide equations run first in a one-thread kernel, as
the reference runs them in ``#pragma omp single`` before the loop.
"""

from __future__ import annotations

import hashlib
import re

from .equations import DISTRIBUTIONS, NeuronDescription, analyse_neuron, equation_code

ABI = 3  # AHIP_MODEL_ABI in include/annarchy_hip_model.h

_PRELUDE = r"""// Generated neuron-model library for the MI355X simulation backend.
// Model: %(name)s
#include "ahip_device.hpp"
#include <cmath>
#include <cstdint>

namespace {

// Built-in functions of the equation language (parser/ParserTemplate.py:27-55).
template <typename T> __host__ __device__ inline T positive(T x) { return x > T(0) ? x : T(0); }
template <typename T> __host__ __device__ inline T negative(T x) { return x < T(0) ? x : T(0); }
template <typename T, typename A, typename B>
__host__ __device__ inline T clip(T x, A lo, B hi) { return x < T(lo) ? T(lo) : (x > T(hi) ? T(hi) : x); }
template <typename A, typename B> __host__ __device__ inline double power(A x, B y) { return pow((double)x, (double)y); }
template <typename A, typename B> __host__ __device__ inline A modulo(A x, B y) { return x %% y; }
template <typename C, typename A, typename B> __host__ __device__ inline A ite(C c, A a, B b) { return c ? a : (A)b; }

"""

_GLOBAL_KERNEL = r"""
template <typename T>
__global__ void global_update(ahip_pop_args a) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t t = ahip::sim_time(a);
    const T dt = (T)a.dt;
    (void)t; (void)dt;
%(global_refs)s
%(global_rand)s
%(global_eqs)s
}
"""

_LOCAL_KERNEL = r"""
template <typename T>
__global__ __launch_bounds__(AHIP_UPDATE_BLOCK) void local_update(ahip_pop_args a) {
    const int32_t i = blockIdx.x * AHIP_UPDATE_BLOCK + threadIdx.x;
    const int64_t t = ahip::sim_time(a);
    const T dt = (T)a.dt;
    (void)t; (void)dt;
%(global_loads)s
%(local_ptrs)s
    bool _spike = false;
    if (i < a.size) {
%(refrac_in)s
%(local_rand)s
%(local_eqs)s
%(refrac_dec)s
%(gather)s
    }
    (void)_spike;
%(publish)s
}

template <typename T>
int launch(const ahip_pop_args* a, void* stream) {
    hipStream_t s = (hipStream_t)stream;
%(launch_global)s
    const int nb = (a->size + AHIP_UPDATE_BLOCK - 1) / AHIP_UPDATE_BLOCK;
    if (nb > 0) hipLaunchKernelGGL(local_update<T>, dim3(nb), dim3(AHIP_UPDATE_BLOCK), 0, s, *a);
    return (int)hipGetLastError();
}

int launch_f64(const ahip_pop_args* a, void* s) { return launch<double>(a, s); }
int launch_f32(const ahip_pop_args* a, void* s) { return launch<float>(a, s); }

%(preloop_fn)s

const char* const kSlotNames[] = {%(slot_names)s};
const int32_t kSlotLocal[] = {%(slot_local)s};
const double kSlotInit[] = {%(slot_init)s};
const int32_t kSlotKind[] = {%(slot_kind)s};
const int32_t kRandDist[] = {%(rand_dist)s};
const double kRandArgs[] = {%(rand_args)s};

const ahip_model_desc kDesc = {
    %(abi)d, "%(name)s", %(spiking)d, %(n_slots)d, kSlotNames, kSlotLocal, kSlotInit, kSlotKind,
    %(n_rand)d, kRandDist, kRandArgs, %(has_refractory)d, %(has_global)d,
    launch_f64, launch_f32, %(preloop_ptr)s, "%(hash)s"};

}  // namespace

extern "C" const ahip_model_desc* ahip_model_describe(void) { return &kDesc; }
"""


def _indent(code: str, n: int) -> str:
    pad = " " * n
    return "\n".join(pad + line if line.strip() else line for line in code.split("\n"))


class ModelSource:
    """Generated HIP source of one neuron model plus its slot table."""

    def __init__(self, desc: NeuronDescription, source: str, slots: list, refractory):
        self.desc = desc
        self.source = source
        self.slots = slots  # (name, local, init, kind)
        self.refractory = refractory

    @property
    def name(self) -> str:
        return self.desc.name

    @property
    def hash(self) -> str:
        return hashlib.sha1(self.source.encode()).hexdigest()[:12]


def _c_name(name: str) -> str:
    return re.sub(r"\W", "_", name)


def generate(desc: NeuronDescription) -> ModelSource:
    """Build the HIP source of the update kernel of one neuron model."""
    slots = []  # (name, local, init, kind)
    for p in desc.parameters:
        slots.append((p.name, p.locality == "local", p.init, 0))
    for v in desc.variables:
        if any(s[0] == v.name for s in slots):
            continue
        slots.append((v.name, v.locality == "local", v.init, 1))
    if not desc.spiking:
        for tgt in desc.targets:
            slots.append(("_sum_" + tgt, True, 0.0, 2))
    preloops = []
    for v in desc.variables:
        if v.pre_loop:
            preloops.append(v.pre_loop)
            slots.append((v.pre_loop["name"], False, 0.0, 3))
    slot_of = {s[0]: k for k, s in enumerate(slots)}

    refractory = desc.refractory if desc.spiking else None
    has_ref = refractory is not None and refractory != 0 and refractory != 0.0

    # -- population-wide attributes -----------------------------------------
    global_names = [s[0] for s in slots if not s[1]]
    global_loads = "\n".join(
        f"    const T {_c_name(n)} = ((const T*)a.slot[{slot_of[n]}])[0];" for n in global_names)
    global_refs = "\n".join(
        f"    T& {_c_name(n)} = ((T*)a.slot[{slot_of[n]}])[0];" for n in global_names)
    local_names = [s[0] for s in slots if s[1]]
    local_ptrs = "\n".join(
        f"    T* __restrict__ {_c_name(n)} = (T*)a.slot[{slot_of[n]}];" for n in local_names)

    def subst(code: str) -> str:
        for r in desc.randoms:
            code = code.replace(r.name + "%(local_index)s", r.name)
            code = code.replace(r.name + "%(global_index)s", r.name)
        code = code.replace("in_ref%(local_index)s", "in_ref")
        code = code.replace("%(local_index)s", "[i]").replace("%(global_index)s", "")
        code = code.replace("%(semiglobal_index)s", "")
        return code

    def rand_code(locality: str, idx: str) -> str:
        out = []
        for k, r in enumerate(desc.randoms):
            if r.locality != locality:
                continue
            a0, a1 = r.args
            out.append(f"const T {r.name} = ahip::random_value<T>(a, {idx}, {k}, t, "
                       f"{DISTRIBUTIONS[r.dist]}, {a0!r}, {a1!r});")
        return "\n".join(out)

    global_eqs = subst(equation_code(desc, "global", None))
    has_global = bool(global_eqs.strip())
    mask = "in_ref" if has_ref else None
    local_eqs = subst(equation_code(desc, "local", mask))

    refrac_in = "const int in_ref = (a.refractory_remaining[i] > 0) ? 0 : 1;" if has_ref else ""
    refrac_dec = "a.refractory_remaining[i] -= (1 - in_ref);" if has_ref else ""
    gather = ""
    publish = ""
    if desc.spiking:
        if isinstance(refractory, str):
            rvar = refractory
            idx = "[i]" if any(s[0] == rvar and s[1] for s in slots) else ""
            refrac_set = f"a.refractory_remaining[i] = int({rvar}{idx}/dt);"
        elif has_ref:
            refrac_set = "a.refractory_remaining[i] = a.refractory[i];"
        else:
            refrac_set = ""
        reset = "\n".join(subst(r) for r in desc.spike_reset)
        check = "in_ref != 0 && " if has_ref else ""
        gather = (f"if ({check}({subst(desc.spike_cond)})) {{\n"
                  f"{_indent(reset, 4)}\n"
                  f"    a.last_spike[i] = t;\n"
                  f"    {refrac_set}\n"
                  f"    _spike = true;\n"
                  f"}}")
        publish = "    ahip::publish_spikes(a, _spike);"

    launch_global = ("    hipLaunchKernelGGL(global_update<T>, dim3(1), dim3(64), 0, s, *a);"
                     if has_global else "")

    if preloops:
        lines = ["void preloop(const double* g, double dt, double* out) {"]
        for n in global_names:
            lines.append(f"    const double {_c_name(n)} = g[{slot_of[n]}]; (void){_c_name(n)};")
        for pl in preloops:
            val = pl["value"].replace("%(global_index)s", "").replace("%(local_index)s", "")
            lines.append(f"    out[{slot_of[pl['name']]}] = {val};")
        lines.append("}")
        preloop_fn = "\n".join(lines)
        preloop_ptr = "preloop"
    else:
        preloop_fn = ""
        preloop_ptr = "nullptr"

    body = _PRELUDE % {"name": desc.name}
    if has_global:
        body += _GLOBAL_KERNEL % {
            "global_refs": global_refs,
            "global_rand": _indent(rand_code("global", "0"), 4),
            "global_eqs": _indent(global_eqs, 4),
        }
    src_wo_hash = _LOCAL_KERNEL % {
        "global_loads": global_loads,
        "local_ptrs": local_ptrs,
        "refrac_in": _indent(refrac_in, 8),
        "local_rand": _indent(rand_code("local", "i"), 8),
        "local_eqs": _indent(local_eqs, 8),
        "refrac_dec": _indent(refrac_dec, 8),
        "gather": _indent(gather, 8),
        "publish": publish,
        "launch_global": launch_global,
        "preloop_fn": preloop_fn,
        "slot_names": ", ".join(f'"{s[0]}"' for s in slots) or "nullptr",
        "slot_local": ", ".join(str(int(s[1])) for s in slots) or "0",
        "slot_init": ", ".join(repr(float(s[2])) for s in slots) or "0.0",
        "slot_kind": ", ".join(str(s[3]) for s in slots) or "0",
        "rand_dist": ", ".join(str(DISTRIBUTIONS[r.dist]) for r in desc.randoms) or "0",
        "rand_args": ", ".join(f"{r.args[0]!r}, {r.args[1]!r}" for r in desc.randoms) or "0.0",
        "abi": ABI,
        "name": desc.name,
        "spiking": int(desc.spiking),
        "n_slots": len(slots),
        "n_rand": len(desc.randoms),
        "has_refractory": int(has_ref),
        "has_global": int(has_global),
        "preloop_ptr": preloop_ptr,
        "hash": "%(hash)s",
    }
    h = hashlib.sha1((body + src_wo_hash).encode()).hexdigest()[:12]
    source = body + src_wo_hash.replace("%(hash)s", h)
    return ModelSource(desc, source, slots, refractory)


def generate_from_text(name: str, parameters: str = "", equations: str = "", spike=None,
                       reset=None, refractory=None) -> ModelSource:
    return generate(analyse_neuron(parameters, equations, spike, reset, refractory, name))
