"""Splice analysed neuron equations into the CDNA4 update-kernel template.

The reference emits, per population, a CPU loop
(``generator/Population/OpenMPGenerator.py:765-1237``) or a CUDA pair of
kernels (``generator/Population/CUDATemplates.py:543-720``: update, then a
separate spike_gather that appends with atomics).  Here one kernel per model
does the whole per-neuron work of a step, one thread per neuron, 256 threads
per workgroup (4 waves):

    ordered spike delivery into g_<target>               (when the runtime asks)
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

ABI = 6  # AHIP_MODEL_ABI in include/annarchy_hip_model.h

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
// power(x, a): x multiplied a times by itself, the exponent taken as an
// unsigned integer (OpenMPBaseTemplate.py:544-552; a = 0 gives 1 here)
template <typename A> struct ahip_real { typedef double type; };
template <> struct ahip_real<float> { typedef float type; };
template <typename A>
__host__ __device__ inline typename ahip_real<A>::type power(A x, unsigned int a) {
    typedef typename ahip_real<A>::type R;
    R res = (R)x;
    for (unsigned int i = 1; i < a; i++) res *= (R)x;
    return a == 0 ? R(1) : res;
}
// modulo(a, b) of the integer parts (OpenMPBaseTemplate.py:539)
template <typename A, typename B> __host__ __device__ inline long modulo(A x, B y) { return (long)x %% (long)y; }
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
    // this step's spike events of the ordered conductances (end of compute_psp)
    if (a.n_deliver > 0 && i < a.size) ahip::deliver_buckets<T>(a, i);
    asm volatile("" ::: "memory");  // the equations below read the delivered values
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


_STOP_KERNEL = r"""
namespace {
// stop_condition of the population (Population/OpenMPGenerator.py:488-555):
// the number of neurons for which the condition holds
template <typename T>
__global__ __launch_bounds__(AHIP_UPDATE_BLOCK) void stop_count(ahip_pop_args a, int32_t* count) {
    const int32_t i = blockIdx.x * AHIP_UPDATE_BLOCK + threadIdx.x;
    const int64_t t = ahip::sim_time(a);
    const T dt = (T)a.dt;
    (void)t; (void)dt;
%(global_loads)s
%(local_ptrs)s
    const bool c = i < a.size && (%(cond)s);
    const unsigned long long m = __ballot(c);
    if ((threadIdx.x & 63) == 0 && m != 0ull) atomicAdd(count, (int32_t)__popcll(m));
}

template <typename T>
int stop_launch(const ahip_pop_args* a, int32_t* count, void* stream) {
    const int nb = (a->size + AHIP_UPDATE_BLOCK - 1) / AHIP_UPDATE_BLOCK;
    if (nb > 0) hipLaunchKernelGGL(stop_count<T>, dim3(nb), dim3(AHIP_UPDATE_BLOCK), 0, (hipStream_t)stream, *a, count);
    return (int)hipGetLastError();
}
}  // namespace

extern "C" int ahip_model_stop_f64(const ahip_pop_args* a, int32_t* count, void* s) { return stop_launch<double>(a, count, s); }
extern "C" int ahip_model_stop_f32(const ahip_pop_args* a, int32_t* count, void* s) { return stop_launch<float>(a, count, s); }
"""


def parse_stop_condition(text: str | None) -> tuple[str | None, str]:
    """stop_condition "expr : any|all" (Extraction.py extract_stop_condition):
    (expression or None for "spiked", mode); mode is "any" unless ": all"."""
    if not text:
        return None, ""
    cond, mode = text, "any"
    if ":" in text:
        head, tail = text.rsplit(":", 1)
        if tail.strip() in ("any", "all"):
            cond, mode = head, tail.strip()
    cond = cond.strip()
    return (None if cond == "spiked" else cond), mode


def generate(desc: NeuronDescription, stop_condition: str | None = None) -> ModelSource:
    """Build the HIP source of the update kernel of one neuron model (and of
    the population's stop condition, when it has one)."""
    slots = []  # (name, local, init, kind)
    for p in desc.parameters:
        # kind 4: a network constant (written by ahip_net_set_constant)
        slots.append((p.name, p.locality == "local", p.init, 4 if "constant" in p.flags else 0))
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
    cond, _ = parse_stop_condition(stop_condition)
    if cond is not None:
        from .equations import _Translator
        src_wo_hash += _STOP_KERNEL % {"global_loads": global_loads, "local_ptrs": local_ptrs,
                                       "cond": subst(_Translator(desc, "double").condition(cond))}
    h = hashlib.sha1((body + src_wo_hash).encode()).hexdigest()[:12]
    source = body + src_wo_hash.replace("%(hash)s", h)
    return ModelSource(desc, source, slots, refractory)


def generate_from_text(name: str, parameters: str = "", equations: str = "", spike=None,
                       reset=None, refractory=None, constants=None, stop_condition=None,
                       incoming_targets=None) -> ModelSource:
    return generate(analyse_neuron(parameters, equations, spike, reset, refractory, name, constants=constants,
                                   incoming_targets=incoming_targets), stop_condition)


# ---------------------------------------------------------------------------
# Synapse models (pre / post events, continuous updates)
# ---------------------------------------------------------------------------

SYNAPSE_ABI = 2  # AHIP_SYNAPSE_ABI in include/annarchy_hip_synapse.h
SYN_CONSTANT = 3  # AHIP_SYN_CONSTANT: a network constant, one value per projection

_SYN_KERNELS = r"""
#include "annarchy_hip_synapse.h"

namespace {

__device__ inline bool update_due(const ahip_syn_args& a, int64_t t) {
    if (a.update_period <= 1) return true;
    int64_t m = (t - a.update_offset) %% a.update_period;
    return m == 0;
}

// Pre events (compute_psp of step t, spikes of step t-d): work item (spike,
// chunk c) goes to one wave, whose lanes take synapses c*64 + lane,
// (c + chunks)*64 + lane, ... of the pre neuron's column.  A synapse belongs
// to one pre neuron and one chunk, so no two lanes touch the same state.
template <typename T>
__global__ __launch_bounds__(256) void pre_event(ahip_syn_args a) {
    const int lane = threadIdx.x & (ahip::kWave - 1);
    const int wid = threadIdx.x / ahip::kWave;
    const int64_t t = *a.t_base + a.step;
    const T dt = (T)a.dt;
    const int32_t _plasticity = a.plasticity;
    const int64_t _rs = a.state_stride, _rs8 = a.state_stride8;
    T* __restrict__ w = (T*)a.w;
    T* __restrict__ _g_post = (T*)a.g;
    int64_t* __restrict__ _last_event = a.last_event;
    const int64_t* __restrict__ _pre_last_spike = a.pre_last_spike;
    const int64_t* __restrict__ _post_last_spike = a.post_last_spike;
    (void)dt; (void)_plasticity; (void)_last_event; (void)_pre_last_spike; (void)_post_last_spike;
%(ev_attrs)s
%(deps)s
    int64_t slot = 0;
    if (a.pre_depth > 1) {
        slot = (t - a.pre_delay) %% a.pre_depth;
        if (slot < 0) slot += a.pre_depth;
    }
    const int32_t n = a.pre_count[slot];
    const int32_t* __restrict__ list = a.pre_spiked + slot * a.pre_stride;
    const int32_t ch = a.pre_chunks;
    const int64_t items = (int64_t)n * ch;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t it = (int64_t)blockIdx.x * 4 + wid; it < items; it += waves) {
        const int64_t s = it / ch;
        const int32_t c = (int32_t)(it - s * ch);
        const int32_t pre = list[s];
        if (pre < 0 || pre >= a.n_cols) continue;
        const int64_t e = a.col_ptr[pre + 1];
        for (int64_t q = a.col_ptr[pre] + (int64_t)c * ahip::kWave + lane; q < e; q += (int64_t)ch * ahip::kWave) {
            const int32_t post = a.row_idx[q];
%(pre_body)s
        }
    }
}

// Post events (after the neuron updates of step t, spikes of step t): work
// items (spike, chunk) over the spiking post neuron's row as for the pre
// events; state by column position.
template <typename T>
__global__ __launch_bounds__(256) void post_event(ahip_syn_args a) {
    const int lane = threadIdx.x & (ahip::kWave - 1);
    const int wid = threadIdx.x / ahip::kWave;
    const int64_t t = *a.t_base + a.step;
    const T dt = (T)a.dt;
    const int32_t _plasticity = a.plasticity;
    const int64_t _rs = a.state_stride, _rs8 = a.state_stride8;
    T* __restrict__ w = (T*)a.w;
    int64_t* __restrict__ _last_event = a.last_event;
    const int64_t* __restrict__ _pre_last_spike = a.pre_last_spike;
    const int64_t* __restrict__ _post_last_spike = a.post_last_spike;
    (void)dt; (void)_plasticity; (void)_last_event; (void)_pre_last_spike; (void)_post_last_spike;
%(ev_attrs)s
%(deps)s
    const int32_t n = *a.post_count;
    const int32_t ch = a.post_chunks;
    const int64_t items = (int64_t)n * ch;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t it = (int64_t)blockIdx.x * 4 + wid; it < items; it += waves) {
        const int64_t s = it / ch;
        const int32_t c = (int32_t)(it - s * ch);
        const int32_t i = a.post_spiked[s];
        if (i < 0 || i >= a.n_rows) continue;
        const int64_t e = a.row_ptr[i + 1];
        for (int64_t j = a.row_ptr[i] + (int64_t)c * ahip::kWave + lane; j < e; j += (int64_t)ch * ahip::kWave) {
            const int64_t q = a.fwd2col ? a.fwd2col[j] : j;
            const int32_t pre = a.col_idx[j];
            const int32_t post = i;
            (void)pre; (void)post;
%(post_body)s
        }
    }
}

// update_synapse (Projection/OpenMP/CSR.py:833-865), once every
// update_period steps with dt scaled by the period
// (OpenMPGenerator.py:1722-1725, :1805).  Per-projection equations run in a
// one-thread kernel first.
template <typename T>
__global__ void update_global(ahip_syn_args a) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t t = *a.t_base + a.step;
    if (!update_due(a, t)) return;
    const T dt = (T)(a.dt * (double)(a.update_period > 1 ? a.update_period : 1));
    const int32_t _plasticity = a.plasticity;
    (void)dt; (void)_plasticity;
%(glob_regs)s
%(preloop)s
%(global_eqs)s
%(glob_store)s
}

// Then one wave per post row: the per-post-neuron equations (evaluated in
// every lane on the same values, stored by lane 0), then the per-synapse
// equations, one lane per synapse.
template <typename T>
__global__ __launch_bounds__(256) void update_local(ahip_syn_args a) {
    const int lane = threadIdx.x & (ahip::kWave - 1);
    const int wid = threadIdx.x / ahip::kWave;
    const int64_t t = *a.t_base + a.step;
    if (!update_due(a, t)) return;
    const T dt = (T)(a.dt * (double)(a.update_period > 1 ? a.update_period : 1));
    const int32_t _plasticity = a.plasticity;
    const int64_t _rs = a.state_stride;
    T* __restrict__ w = (T*)a.w;
    const int64_t* __restrict__ _pre_last_spike = a.pre_last_spike;
    const int64_t* __restrict__ _post_last_spike = a.post_last_spike;
    (void)dt; (void)_plasticity; (void)w; (void)_pre_last_spike; (void)_post_last_spike;
%(up_attrs)s
%(deps)s
%(preloop)s
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t row = (int64_t)blockIdx.x * 4 + wid; row < a.n_rows; row += waves) {
        const int32_t post = (int32_t)row;
        (void)post;
%(sg_load)s
%(semiglobal_eqs)s
%(sg_store)s
        const int64_t e = a.row_ptr[row + 1];
        for (int64_t j = a.row_ptr[row] + lane; j < e; j += ahip::kWave) {
            const int64_t q = a.fwd2col ? a.fwd2col[j] : j;
            const int32_t pre = a.col_idx[j];
            (void)q; (void)pre;
%(local_eqs)s
        }
    }
    (void)lane;
}

template <typename T>
int launch_events(const ahip_syn_args* a, void* stream, bool pre) {
    // work items (spike, chunk of 64 synapses): one pass over every neuron
    // of the population spiking, up to 4 workgroups per CU, then loops
    const int64_t units = (int64_t)(pre ? a->n_cols : a->n_rows) * (pre ? a->pre_chunks : a->post_chunks);
    int64_t nb = (units + 3) / 4;
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    if (pre)
        hipLaunchKernelGGL(pre_event<T>, dim3(nb), dim3(256), 0, (hipStream_t)stream, *a);
    else
        hipLaunchKernelGGL(post_event<T>, dim3(nb), dim3(256), 0, (hipStream_t)stream, *a);
    return (int)hipGetLastError();
}

template <typename T>
int launch_update(const ahip_syn_args* a, void* stream) {
    hipStream_t s = (hipStream_t)stream;
%(launch_global)s
%(launch_local)s
    return (int)hipGetLastError();
}

int pre_f64(const ahip_syn_args* a, void* s) { return launch_events<double>(a, s, true); }
int pre_f32(const ahip_syn_args* a, void* s) { return launch_events<float>(a, s, true); }
int post_f64(const ahip_syn_args* a, void* s) { return launch_events<double>(a, s, false); }
int post_f32(const ahip_syn_args* a, void* s) { return launch_events<float>(a, s, false); }
int update_f64(const ahip_syn_args* a, void* s) { return launch_update<double>(a, s); }
int update_f32(const ahip_syn_args* a, void* s) { return launch_update<float>(a, s); }

const char* const kAttrNames[] = {%(attr_names)s};
const int32_t kAttrLocality[] = {%(attr_loc)s};
const double kAttrInit[] = {%(attr_init)s};
const char* const kPreDeps[] = {%(pre_deps)s};
const char* const kPostDeps[] = {%(post_deps)s};

const ahip_synapse_desc kDesc = {
    %(abi)d, "%(name)s", %(spiking)d, %(n_attr)d, kAttrNames, kAttrLocality, kAttrInit,
    %(n_pre_deps)d, kPreDeps, %(n_post_deps)d, kPostDeps, %(has_pre)d, %(has_post)d, %(has_update)d,
    pre_f64, pre_f32, post_f64, post_f32, update_f64, update_f32, "%(hash)s"};

}  // namespace

extern "C" const ahip_synapse_desc* ahip_synapse_describe(void) { return &kDesc; }
"""


class SynapseSource:
    """Generated HIP source of one synapse model plus its attribute table."""

    def __init__(self, desc, source: str, attrs: list):
        self.desc = desc
        self.source = source
        self.attrs = attrs  # (name, locality, init)

    @property
    def name(self) -> str:
        return self.desc.name


def _syn_subst(code: str, t_text: str, sg_index: str) -> str:
    """Per-synapse values of synapse q sit at [q * _rs] (_last_event at
    [q * _rs8]): one record per synapse, or separate arrays with stride 1."""
    code = code.replace("%(t)s", t_text).replace("_last_event%(local_index)s", "_last_event[q*_rs8]")
    return (code.replace("%(local_index)s", "[q*_rs]").replace("%(global_index)s", "")
            .replace("%(semiglobal_index)s", sg_index))


def generate_synapse(desc) -> SynapseSource:
    """Build the HIP source of the event and update kernels of a synapse model."""
    from .equations import SYN_LOCALITY, synapse_update_code

    attrs = desc.attrs()
    if len(attrs) > 16:
        raise ValueError("at most 16 synapse attributes besides w")
    if len(desc.pre_deps) > 8 or len(desc.post_deps) > 8:
        raise ValueError("at most 8 pre- and 8 post-synaptic dependencies")
    slot = {a[0]: k for k, a in enumerate(attrs)}
    ptr = "    T* __restrict__ {p}_{n} = (T*)a.attr[{k}];"
    load = "    const T _s_{n} = ((const T*)a.attr[{k}])[0];"
    # event kernels: per-synapse and per-post-neuron attributes by pointer, per-projection by value
    ev = []
    for n, loc, _ in attrs:
        ev.append(load.format(n=_c_name(n), k=slot[n]) if loc == "global"
                  else ptr.format(p="_s", n=_c_name(n), k=slot[n]))
    # update kernel: semiglobal attributes in registers per row
    up, sg_load, sg_store = [], [], []
    for n, loc, _ in attrs:
        c = _c_name(n)
        if loc == "global":
            up.append(load.format(n=c, k=slot[n]))
        elif loc == "local":
            up.append(ptr.format(p="_s", n=c, k=slot[n]))
        else:
            up.append(ptr.format(p="_sp", n=c, k=slot[n]))
            sg_load.append(f"        T _s_{c} = _sp_{c}[post];")
            if any(v.name == n and v.cpp for v in desc.variables):
                sg_store.append(f"        if (lane == 0) _sp_{c}[post] = _s_{c};")
    glob_regs, glob_store = [], []
    for n, loc, _ in attrs:
        if loc != "global":
            continue
        c = _c_name(n)
        glob_regs.append(f"    T _s_{c} = ((T*)a.attr[{slot[n]}])[0];")
        if any(v.name == n and v.cpp for v in desc.variables):
            glob_store.append(f"    ((T*)a.attr[{slot[n]}])[0] = _s_{c};")
    deps = []
    for side, lst in (("pre", desc.pre_deps), ("post", desc.post_deps)):
        for k, n in enumerate(lst):
            deps.append(f"    const T* __restrict__ _d{side}_{n} = (const T*)a.{side}_attr[{k}];")
            deps.append(f"    const int32_t _st_{side}_{n} = a.{side}_attr_stride[{k}];")
            deps.append(f"    (void)_d{side}_{n}; (void)_st_{side}_{n};")
    preloop = "\n".join(f"    const T {n} = (T)({_syn_subst(v, '(t)', '')});" for n, v in desc.preloop)

    pre, post = [], []
    # Projection/OpenMPGenerator.py:1182-1230: event-driven decay to t-1,
    # _last_event = t, then g_target += <expr> (w when the model does not
    # write g_target), then the pre_spike statements
    for _, code in desc.event_driven:
        pre.append(_syn_subst(code, "(t-1)", "[post]"))
    if desc.event_driven:
        pre.append("_last_event[q*_rs8] = t;")
    g = _syn_subst(desc.g_code, "(t)", "[post]") if desc.g_code else "w[q*_rs]"
    pre.append(f"unsafeAtomicAdd(_g_post + post, (T)({g}));")
    pre += [_syn_subst(c, "(t)", "[post]") for c in desc.pre_code]
    if desc.post_code:
        # OpenMPGenerator.py:1575-1600: event-driven decay to t, _last_event = t
        for _, code in desc.event_driven:
            post.append(_syn_subst(code, "(t)", "[post]"))
        if desc.event_driven:
            post.append("_last_event[q*_rs8] = t;")
        post += [_syn_subst(c, "(t)", "[post]") for c in desc.post_code]
    global_eqs = _syn_subst(synapse_update_code(desc, "global"), "(t)", "")
    sg_eqs = _syn_subst(synapse_update_code(desc, "semiglobal"), "(t)", "")
    local_eqs = _syn_subst(synapse_update_code(desc, "local"), "(t)", "")
    has_global = bool(global_eqs.strip())
    has_rows = bool(sg_eqs.strip() or local_eqs.strip())
    launch_global = ("    hipLaunchKernelGGL(update_global<T>, dim3(1), dim3(64), 0, s, *a);"
                     if has_global else "")
    launch_local = ("    {\n        int nb = (a->n_rows + 3) / 4;\n        if (nb > 1024) nb = 1024;\n"
                    "        if (nb > 0) hipLaunchKernelGGL(update_local<T>, dim3(nb), dim3(256), 0, s, *a);\n    }"
                    if has_rows else "")
    body = _PRELUDE % {"name": desc.name} + "}  // namespace\n" + _SYN_KERNELS % {
        "ev_attrs": "\n".join(ev),
        "up_attrs": "\n".join(up),
        "deps": "\n".join(deps),
        "preloop": preloop,
        "sg_load": "\n".join(sg_load),
        "sg_store": "\n".join(sg_store),
        "glob_regs": "\n".join(glob_regs),
        "glob_store": "\n".join(glob_store),
        "global_eqs": _indent(global_eqs, 4),
        "semiglobal_eqs": _indent(sg_eqs, 8),
        "local_eqs": _indent(local_eqs, 12),
        "pre_body": _indent("\n".join(pre), 12),
        "post_body": _indent("\n".join(post), 12),
        "launch_global": launch_global,
        "launch_local": launch_local,
        "attr_names": ", ".join(f'"{n}"' for n, _, _ in attrs) or "nullptr",
        "attr_loc": ", ".join(str(SYN_CONSTANT if n in desc.constant_names() else SYN_LOCALITY[loc])
                              for n, loc, _ in attrs) or "0",
        "attr_init": ", ".join(repr(float(v)) for _, _, v in attrs) or "0.0",
        "pre_deps": ", ".join(f'"{n}"' for n in desc.pre_deps) or "nullptr",
        "post_deps": ", ".join(f'"{n}"' for n in desc.post_deps) or "nullptr",
        "n_pre_deps": len(desc.pre_deps),
        "n_post_deps": len(desc.post_deps),
        "abi": SYNAPSE_ABI,
        "name": desc.name,
        "spiking": int(desc.spiking),
        "n_attr": len(attrs),
        "has_pre": int(desc.spiking),
        "has_post": int(bool(post)),
        "has_update": int(has_global or has_rows),
        "hash": "%(hash)s",
    }
    h = hashlib.sha1(body.encode()).hexdigest()[:12]
    return SynapseSource(desc, body.replace("%(hash)s", h), attrs)


def generate_synapse_from_text(name: str, parameters: str = "", equations: str = "", pre_spike=None,
                               post_spike=None, constants=None) -> SynapseSource:
    from .equations import analyse_synapse
    return generate_synapse(analyse_synapse(parameters, equations, pre_spike, post_spike, name,
                                            constants=constants))
