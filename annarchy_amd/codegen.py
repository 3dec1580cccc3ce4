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


# ---------------------------------------------------------------------------
# Spiking synapse models (pre / post events)
# ---------------------------------------------------------------------------

SYNAPSE_ABI = 1  # AHIP_SYNAPSE_ABI in include/annarchy_hip_synapse.h

_SYN_KERNELS = r"""
#include "annarchy_hip_synapse.h"

namespace {

// Pre events (compute_psp of step t, spikes of step t-d): one wave per
// spike, lanes over the synapses of the pre neuron's column.  A synapse
// belongs to one pre neuron, so no two lanes touch the same state.
template <typename T>
__global__ __launch_bounds__(256) void pre_event(ahip_syn_args a) {
    const int lane = threadIdx.x & (ahip::kWave - 1);
    const int wid = threadIdx.x / ahip::kWave;
    const int64_t t = *a.t_base + a.step;
    const T dt = (T)a.dt;
    (void)dt;
    T* __restrict__ w = (T*)a.w;
    T* __restrict__ _g_post = (T*)a.g;
    int64_t* __restrict__ _last_event = a.last_event;
    const int64_t* __restrict__ _pre_last_spike = a.pre_last_spike;
    const int64_t* __restrict__ _post_last_spike = a.post_last_spike;
    (void)_last_event; (void)_pre_last_spike; (void)_post_last_spike;
%(globals)s
%(locals)s
    int64_t slot = 0;
    if (a.pre_depth > 1) {
        slot = (t - a.pre_delay) %% a.pre_depth;
        if (slot < 0) slot += a.pre_depth;
    }
    const int32_t n = a.pre_count[slot];
    const int32_t* __restrict__ list = a.pre_spiked + slot * a.pre_stride;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t s = (int64_t)blockIdx.x * 4 + wid; s < n; s += waves) {
        const int32_t pre = list[s];
        if (pre < 0 || pre >= a.n_cols) continue;
        const int64_t e = a.col_ptr[pre + 1];
        for (int64_t q = a.col_ptr[pre] + lane; q < e; q += ahip::kWave) {
            const int32_t post = a.row_idx[q];
            (void)post;
%(pre_body)s
        }
    }
}

// Post events (after the neuron updates of step t, spikes of step t): one
// wave per spiking post neuron, lanes over its row; state by column position.
template <typename T>
__global__ __launch_bounds__(256) void post_event(ahip_syn_args a) {
    const int lane = threadIdx.x & (ahip::kWave - 1);
    const int wid = threadIdx.x / ahip::kWave;
    const int64_t t = *a.t_base + a.step;
    const T dt = (T)a.dt;
    (void)dt;
    T* __restrict__ w = (T*)a.w;
    int64_t* __restrict__ _last_event = a.last_event;
    const int64_t* __restrict__ _pre_last_spike = a.pre_last_spike;
    const int64_t* __restrict__ _post_last_spike = a.post_last_spike;
    (void)_last_event; (void)_pre_last_spike; (void)_post_last_spike;
%(globals)s
%(locals)s
    const int32_t n = *a.post_count;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t s = (int64_t)blockIdx.x * 4 + wid; s < n; s += waves) {
        const int32_t i = a.post_spiked[s];
        if (i < 0 || i >= a.n_rows) continue;
        const int64_t e = a.row_ptr[i + 1];
        for (int64_t j = a.row_ptr[i] + lane; j < e; j += ahip::kWave) {
            const int64_t q = a.fwd2col[j];
            const int32_t pre = a.col_idx[j];
            const int32_t post = i;
            (void)pre; (void)post;
%(post_body)s
        }
    }
}

template <typename T, bool PRE>
int launch(const ahip_syn_args* a, void* stream) {
    const int32_t units = PRE ? a->n_cols : a->n_rows;
    int nb = (units + 3) / 4;
    if (nb > 512) nb = 512;
    if (nb < 1) nb = 1;
    if (PRE)
        hipLaunchKernelGGL(pre_event<T>, dim3(nb), dim3(256), 0, (hipStream_t)stream, *a);
    else
        hipLaunchKernelGGL(post_event<T>, dim3(nb), dim3(256), 0, (hipStream_t)stream, *a);
    return (int)hipGetLastError();
}

int pre_f64(const ahip_syn_args* a, void* s) { return launch<double, true>(a, s); }
int pre_f32(const ahip_syn_args* a, void* s) { return launch<float, true>(a, s); }
int post_f64(const ahip_syn_args* a, void* s) { return launch<double, false>(a, s); }
int post_f32(const ahip_syn_args* a, void* s) { return launch<float, false>(a, s); }

const char* const kAttrNames[] = {%(attr_names)s};
const int32_t kAttrLocal[] = {%(attr_local)s};
const double kAttrInit[] = {%(attr_init)s};

const ahip_synapse_desc kDesc = {
    %(abi)d, "%(name)s", %(n_attr)d, kAttrNames, kAttrLocal, kAttrInit, %(has_pre)d, %(has_post)d, %(pre_g)d,
    pre_f64, pre_f32, post_f64, post_f32, "%(hash)s"};

}  // namespace

extern "C" const ahip_synapse_desc* ahip_synapse_describe(void) { return &kDesc; }
"""


class SynapseSource:
    """Generated HIP source of one synapse model plus its attribute table."""

    def __init__(self, desc, source: str, attrs: list):
        self.desc = desc
        self.source = source
        self.attrs = attrs  # (name, local, init)

    @property
    def name(self) -> str:
        return self.desc.name


def generate_synapse(desc) -> SynapseSource:
    """Build the HIP source of the pre / post event kernels of a synapse model."""
    attrs = [(a.name, a.locality == "local", a.init) for a in desc.parameters + desc.variables]
    if len(attrs) > 16:
        raise ValueError("at most 16 synapse attributes besides w")
    slot = {a[0]: k for k, a in enumerate(attrs)}
    globals_ = "\n".join(f"    const T _s_{_c_name(n)} = (T)a.global[{slot[n]}];" for n, loc, _ in attrs if not loc)
    locals_ = "\n".join(f"    T* __restrict__ _s_{_c_name(n)} = (T*)a.local[{slot[n]}];" for n, loc, _ in attrs
                         if loc)

    def subst(code: str, t_text: str) -> str:
        code = code.replace("%(t)s", t_text)
        return code.replace("%(local_index)s", "[q]").replace("%(global_index)s", "")

    pre, post = [], []
    # Projection/OpenMPGenerator.py:1182-1230: event-driven decay to t-1,
    # _last_event = t, then g_target += w (the default transmission when the
    # model does not write g_target itself), then the pre_spike statements
    for _, code in desc.event_driven:
        pre.append(subst(code, "(t-1)"))
    if desc.event_driven:
        pre.append("_last_event[q] = t;")
    pre.append("unsafeAtomicAdd(_g_post + post, w[q]);")
    pre += [subst(c, "(t)") for c in desc.pre_code]
    if desc.post_code:
        # OpenMPGenerator.py:1575-1600: event-driven decay to t, _last_event = t
        for _, code in desc.event_driven:
            post.append(subst(code, "(t)"))
        if desc.event_driven:
            post.append("_last_event[q] = t;")
        post += [subst(c, "(t)") for c in desc.post_code]
    body = _PRELUDE % {"name": desc.name} + "}  // namespace\n" + _SYN_KERNELS % {
        "globals": globals_,
        "locals": locals_,
        "pre_body": _indent("\n".join(pre), 12),
        "post_body": _indent("\n".join(post), 12),
        "attr_names": ", ".join(f'"{n}"' for n, _, _ in attrs) or "nullptr",
        "attr_local": ", ".join(str(int(loc)) for _, loc, _ in attrs) or "0",
        "attr_init": ", ".join(repr(float(v)) for _, _, v in attrs) or "0.0",
        "abi": SYNAPSE_ABI,
        "name": desc.name,
        "n_attr": len(attrs),
        "has_pre": int(bool(pre)),
        "has_post": int(bool(post)),
        "pre_g": 1,
        "hash": "%(hash)s",
    }
    h = hashlib.sha1(body.encode()).hexdigest()[:12]
    return SynapseSource(desc, body.replace("%(hash)s", h), attrs)


def generate_synapse_from_text(name: str, parameters: str = "", equations: str = "", pre_spike=None,
                               post_spike=None) -> SynapseSource:
    from .equations import analyse_synapse
    return generate_synapse(analyse_synapse(parameters, equations, pre_spike, post_spike, name))
