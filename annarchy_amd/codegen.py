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

ABI = 10  # AHIP_MODEL_ABI in include/annarchy_hip_model.h

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
%(global_loads)s
%(div_consts)s
%(local_ptrs)s
    // the neuron's values the delivery below does not write, in registers
    // (their loads overlap the delivery's)
%(reg_pre)s
    // this step's spike events of the ordered conductances (end of compute_psp)
    if (a.n_deliver > 0 && i < a.size) ahip::deliver_buckets<T>(a, i);
    asm volatile("" ::: "memory");  // the equations below read the delivered values
%(reg_post)s
    bool _spike = false;
    if (i < a.size) {
%(refrac_in)s
%(local_rand)s
%(local_eqs)s
%(refrac_dec)s
%(gather)s
%(reg_out)s
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

%(persist_fn)s

const char* const kSlotNames[] = {%(slot_names)s};
const int32_t kSlotLocal[] = {%(slot_local)s};
const double kSlotInit[] = {%(slot_init)s};
const int32_t kSlotKind[] = {%(slot_kind)s};
const int32_t kRandDist[] = {%(rand_dist)s};
const double kRandArgs[] = {%(rand_args)s};

const ahip_model_desc kDesc = {
    %(abi)d, "%(name)s", %(spiking)d, %(n_slots)d, kSlotNames, kSlotLocal, kSlotInit, kSlotKind,
    %(n_rand)d, kRandDist, kRandArgs, %(has_refractory)d, %(has_global)d,
    launch_f64, launch_f32, %(preloop_ptr)s, "%(hash)s", %(persist_ptrs)s};

}  // namespace

extern "C" const ahip_model_desc* ahip_model_describe(void) { return &kDesc; }
"""


# Persistent steps of a small network (ahip_persist_args): one workgroup of
# 1024 threads runs `pa.steps` ordinary steps.  Thread k owns the neurons
# [k*per, k*per + per), so one block scan of the per-thread spike counts puts
# the spike list in ascending order.  The conductances the projections write
# live in LDS for the whole launch (the statements reach them through the
# same slot pointers), the spike list and the monitor's fill counts too, so
# no thread reads another thread's global writes within the launch.
_PERSIST_KERNEL = r"""
constexpr int kPersistItems = AHIP_PERSIST_ITEMS;
constexpr int kPersistWaves = AHIP_PERSIST_THREADS / 64;
static_assert(kPersistWaves == 16, "the wave-count scan covers one DPP row");
constexpr int kPersistGroup = 16;                                   // lanes per item, small scatters
constexpr int kPersistGroups = AHIP_PERSIST_THREADS / kPersistGroup;
constexpr int kPersistSlotCount = %(n_pslots)d;
constexpr bool kPersistRefractory = %(p_refr)s;

// Workgroup barrier ordering LDS only: within the launch no thread reads the
// global memory another thread writes (the spike list, the monitor and
// last_spike are only written), so the barriers need not wait for global
// stores as __syncthreads() does.
__device__ inline void persist_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// exclusive scan of one value per thread over the workgroup with one
// barrier (wave scans, then each thread adds the totals of the waves before
// it); total = sum.  `wtot` must not be rewritten before the next barrier.
template <typename V>
__device__ inline V persist_scan(V v, V* wtot, V& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    V x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const V y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wtot[wid] = x;
    persist_sync();
    V before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kPersistWaves; w++) {
        const V u = wtot[w];
        before += w < wid ? u : V(0);
        all += u;
    }
    total = all;
    return before + x - v;
}

// inclusive sum over the lanes of each row of 16 (lane l: lanes 16 (l / 16)
// .. l), row_shr DPP moves: lanes shifted in from outside the row add 0
__device__ inline uint32_t persist_row_scan(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);
    return x;
}

// the projection of item `pi` of round r (the pi-th projection in that round)
__device__ inline int32_t persist_round_proj(const ahip_persist_proj* proj, int32_t n_proj, int32_t r, int32_t pi) {
    int32_t c = 0;
    for (int32_t p = 0; p < n_proj; p++)
        if (proj[p].round == r) {
            if (c == pi) return p;
            c++;
        }
    return 0;
}

// loads through the connectivity pointers of the projection table: the
// table sits in LDS, so the compiler cannot tell that they address global
// memory and would emit flat loads (which an LDS barrier waits for)
template <typename V>
__device__ inline V persist_ld(const V* p, int64_t e) {
    return ((const __attribute__((address_space(1))) V*)p)[e];
}

// synapse range [bb, ee) of pre-synaptic rank `pre` in projection P
__device__ inline void persist_range(const ahip_persist_proj& P, const int32_t* cols, int32_t pre, int64_t& bb,
                                     int64_t& ee) {
    bb = ee = 0;
    if (pre >= P.n_cols) return;
    if (P.lds_cols >= 0) {
        bb = cols[P.lds_cols + pre];
        ee = cols[P.lds_cols + pre + 1];
    } else {
        bb = persist_ld(P.col_ptr, pre);
        ee = persist_ld(P.col_ptr, pre + 1);
    }
}

template <typename T>
__global__ __launch_bounds__(AHIP_PERSIST_THREADS) void persistent_steps(ahip_pop_args a_in, ahip_persist_args pa) {
    extern __shared__ unsigned char lds_raw[];
    __shared__ int32_t wsum[kPersistWaves];
    __shared__ uint32_t wlo[kPersistWaves], whi[kPersistWaves];
    __shared__ int64_t fill_s[2];
    __shared__ int32_t nspk_s;
    __shared__ ahip_persist_proj proj_s[AHIP_PERSIST_MAX_PROJ];   // read per lane: LDS, not kernel arguments
    const int32_t N = a_in.size;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    // LDS: the state the statements touch (slot by slot), the refractory
    // counters, the spike list, the column ranges and the scatter's item
    // tables
    T* lds = reinterpret_cast<T*>(lds_raw);
    int32_t* rr = reinterpret_cast<int32_t*>(lds + (size_t)kPersistSlotCount * N);
    int32_t* spk = rr + (kPersistRefractory ? N : 0);
    int32_t* cols = spk + N;   // column ranges of the projections (n_lds_cols entries)
    const size_t beg_at = ((size_t)(reinterpret_cast<unsigned char*>(cols + pa.n_lds_cols) - lds_raw) + 7) &
                          ~(size_t)7;
    int64_t* beg = reinterpret_cast<int64_t*>(lds_raw + beg_at);
    int32_t* off = reinterpret_cast<int32_t*>(beg + kPersistItems);
    int32_t* projq = off + kPersistItems + 1;
    const ahip_pop_args& a = a_in;
    for (int32_t i = tid; i < N; i += AHIP_PERSIST_THREADS) {
%(p_load)s
        if (kPersistRefractory) rr[i] = a.refractory_remaining[i];
    }
    if (tid < pa.n_proj) proj_s[tid] = pa.proj[tid];
    if (tid == 0) {
        nspk_s = *pa.spike_count;
        if (pa.rec_counts) {
            fill_s[0] = pa.rec_fill[0];
            fill_s[1] = pa.rec_fill[1];
        }
    }
    persist_sync();
    int32_t nspk = nspk_s;
    for (int32_t i = tid; i < nspk; i += AHIP_PERSIST_THREADS) spk[i] = pa.spiked[i];
    for (int32_t p = 0; p < pa.n_proj; p++) {
        const ahip_persist_proj& P = pa.proj[p];
        if (P.lds_cols < 0) continue;
        for (int32_t j = tid; j <= P.n_cols; j += AHIP_PERSIST_THREADS) cols[P.lds_cols + j] = (int32_t)P.col_ptr[j];
    }
    const int64_t t0 = *a.t_base + a.step;
    const T dt = (T)a.dt;
    (void)dt;
%(global_loads)s
%(div_consts)s
%(local_ptrs)s
    int32_t* __restrict__ refrac_rem = rr;
    (void)refrac_rem;
    // neuron i = k * AHIP_PERSIST_THREADS + tid, round k < per (consecutive
    // lanes, consecutive neurons)
    const int32_t per = (N + AHIP_PERSIST_THREADS - 1) / AHIP_PERSIST_THREADS;
    // the refractory periods (constant during the launch) of the first four
    // rounds in registers: no global load when a neuron spikes
    int32_t ref0 = 0, ref1 = 0, ref2 = 0, ref3 = 0;
    if (kPersistRefractory && a.refractory) {
        if (tid < N) ref0 = a.refractory[tid];
        if (tid + AHIP_PERSIST_THREADS < N) ref1 = a.refractory[tid + AHIP_PERSIST_THREADS];
        if (tid + 2 * AHIP_PERSIST_THREADS < N) ref2 = a.refractory[tid + 2 * AHIP_PERSIST_THREADS];
        if (tid + 3 * AHIP_PERSIST_THREADS < N) ref3 = a.refractory[tid + 3 * AHIP_PERSIST_THREADS];
    }
    (void)ref0; (void)ref1; (void)ref2; (void)ref3;
    persist_sync();

    // compute_psp of the spike list spk[0, n): projection rounds in creation
    // order, the whole workgroup on each round; ends at a barrier
    auto scatter_list = [&](int32_t n) {
        for (int32_t r = 0; r < pa.n_rounds; r++) {
            int32_t nr = pa.n_proj;
            if (pa.n_rounds > 1) {
                nr = 0;
                for (int32_t p = 0; p < pa.n_proj; p++) nr += proj_s[p].round == r;
            }
            const int32_t n_items = nr * n;
            if (n_items <= kPersistGroups) {
                // one group of lanes per (projection, spike) item: 64, 32 or
                // 16 lanes, the most that give every item a group
                const int gsh = n_items <= kPersistGroups / 4 ? 6 : n_items <= kPersistGroups / 2 ? 5 : 4;
                const int32_t gsz = 1 << gsh;
                const int32_t g = tid >> gsh, gl = tid & (gsz - 1);
                if (g < n_items) {
                    const ahip_persist_proj& P =
                        proj_s[pa.n_rounds > 1 ? persist_round_proj(proj_s, pa.n_proj, r, g / n) : g / n];
                    int64_t bb, ee;
                    persist_range(P, cols, spk[g %% n], bb, ee);
                    T* const gt = lds + (size_t)P.target * N;
                    const T w = (T)P.w;
                    for (int64_t e0 = bb + gl; e0 < ee; e0 += 4 * gsz) {
                        int32_t row[4];
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const int64_t e = e0 + u * gsz;
                            row[u] = e < ee ? persist_ld(P.row_idx, e) : -1;
                        }
#pragma unroll
                        for (int u = 0; u < 4; u++)
                            if (row[u] >= 0) atomicAdd(gt + row[u], w);
                    }
                }
                persist_sync();
                continue;
            }
            for (int32_t q0 = 0; q0 < n_items; q0 += kPersistItems) {
                const int32_t m = min(kPersistItems, n_items - q0);
                int32_t d = 0;
                if (tid < m) {
                    const int32_t q = q0 + tid;
                    const int32_t pi = pa.n_rounds > 1 ? persist_round_proj(proj_s, pa.n_proj, r, q / n) : q / n;
                    int64_t bb, ee;
                    persist_range(proj_s[pi], cols, spk[q %% n], bb, ee);
                    d = (int32_t)(ee - bb);
                    beg[tid] = bb;
                    projq[tid] = pi;
                }
                int32_t E;
                const int32_t o = persist_scan<int32_t>(d, wsum, E);
                if (tid < m) off[tid] = o;
                persist_sync();
                for (int32_t e = tid; e < E; e += AHIP_PERSIST_THREADS) {
                    int32_t lo = 0, hi = m - 1;
                    while (lo < hi) {
                        const int32_t mid = (lo + hi + 1) >> 1;
                        if (off[mid] <= e) lo = mid; else hi = mid - 1;
                    }
                    const ahip_persist_proj& P = proj_s[projq[lo]];
                    const int32_t row = persist_ld(P.row_idx, beg[lo] + (e - off[lo]));
                    atomicAdd(lds + (size_t)P.target * N + row, (T)P.w);
                }
                persist_sync();
            }
        }
    };

    for (int32_t s = 0; s < pa.steps; s++) {
        const int64_t t = t0 + s;
        // 1. compute_psp: the spikes of step t - 1
        scatter_list(nspk);
        // 2. neuron update (the statements of the per-kernel step)
        uint32_t bits = 0;
#pragma unroll 2
        for (int32_t k = 0; k < per; k++) {
            const int32_t i = k * AHIP_PERSIST_THREADS + tid;
            if (i >= N) break;
            bool _spike = false;
#define AHIP_REFRACTORY_I (k == 0 ? ref0 : k == 1 ? ref1 : k == 2 ? ref2 : k == 3 ? ref3 : a.refractory[i])
%(reg_in)s
%(refrac_in)s
%(local_rand)s
%(local_eqs)s
%(refrac_dec)s
%(gather)s
%(reg_out)s
            if (_spike) bits |= 1u << k;
#undef AHIP_REFRACTORY_I
        }
        // 3. ascending spike list (round by round, in lane order: the
        //    position of a spike is its lane's rank in the wave's ballot plus
        //    the counts of the waves before, a 16-lane DPP scan of the wave
        //    counts; four rounds per pass, one 16-bit count each), recording
        int64_t slot = -1, f = 0;
        if (pa.rec_counts) {
            slot = t - pa.rec_t_start;
            if (slot >= 0 && slot < pa.rec_capacity) f = fill_s[slot & 1];
            else slot = -1;
        }
        int32_t base = 0;
        for (int32_t g0 = 0; g0 < per; g0 += 4) {
            unsigned long long msk[4];
#pragma unroll
            for (int k = 0; k < 4; k++) msk[k] = __ballot((bits >> (g0 + k)) & 1u);
            const uint32_t clo = (uint32_t)__popcll(msk[0]) | ((uint32_t)__popcll(msk[1]) << 16);
            const uint32_t chi = (uint32_t)__popcll(msk[2]) | ((uint32_t)__popcll(msk[3]) << 16);
            if (lane == 0) {
                wlo[wid] = clo;
                whi[wid] = chi;
            }
            persist_sync();
            const uint32_t xlo = persist_row_scan(lane < kPersistWaves ? wlo[lane] : 0u);
            const uint32_t xhi = persist_row_scan(lane < kPersistWaves ? whi[lane] : 0u);
            const uint32_t blo = (uint32_t)__builtin_amdgcn_readlane((int)xlo, wid) - clo;
            const uint32_t bhi = (uint32_t)__builtin_amdgcn_readlane((int)xhi, wid) - chi;
            const uint32_t alo = (uint32_t)__builtin_amdgcn_readlane((int)xlo, kPersistWaves - 1);
            const uint32_t ahi = (uint32_t)__builtin_amdgcn_readlane((int)xhi, kPersistWaves - 1);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t before = ((k < 2 ? blo : bhi) >> (16 * (k & 1))) & 0xffffu;
                const uint32_t all = ((k < 2 ? alo : ahi) >> (16 * (k & 1))) & 0xffffu;
                if ((bits >> (g0 + k)) & 1u) {
                    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(msk[k] >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)msk[k], 0u));
                    const int32_t pos = base + (int32_t)(before + below);
                    const int32_t i = (g0 + k) * AHIP_PERSIST_THREADS + tid;
                    spk[pos] = i;
                    pa.spiked[pos] = i;
                    if (slot >= 0 && f + pos < pa.rec_idx_capacity) pa.rec_idx[f + pos] = i;
                }
                base += (int32_t)all;
            }
            if (g0 + 4 < per) persist_sync();   // wlo / whi are reused
        }
        if (tid == 0 && slot >= 0) {
            pa.rec_counts[slot] = base;
            fill_s[(slot + 1) & 1] = f + base;
        }
        nspk = base;
        persist_sync();   // the list is complete before the next step reads it
    }
    for (int32_t i = tid; i < N; i += AHIP_PERSIST_THREADS) {
%(p_store)s
        if (kPersistRefractory) a.refractory_remaining[i] = rr[i];
    }
    if (tid == 0) {
        *pa.spike_count = nspk;
        if (pa.rec_counts) {
            pa.rec_fill[0] = fill_s[0];
            pa.rec_fill[1] = fill_s[1];
        }
    }
}

template <typename T>
int persist(const ahip_pop_args* a, const ahip_persist_args* pa, size_t lds, void* stream) {
    if (a->size > 32 * AHIP_PERSIST_THREADS) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(persistent_steps<T>, dim3(1), dim3(AHIP_PERSIST_THREADS), lds, (hipStream_t)stream, *a, *pa);
    return (int)hipGetLastError();
}
int persist_f64(const ahip_pop_args* a, const ahip_persist_args* pa, size_t lds, void* s) {
    return persist<double>(a, pa, lds, s);
}
int persist_f32(const ahip_pop_args* a, const ahip_persist_args* pa, size_t lds, void* s) {
    return persist<float>(a, pa, lds, s);
}
const int32_t kPersistSlotsHost[] = {%(pslot_list)s};
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


def _u(name: str) -> str:
    """The C identifier of a model's attribute in the generated kernels."""
    return "u_" + _c_name(name)


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

    # user identifiers are emitted with a prefix (u_v, u_tau, ...), so no
    # model name can collide with the kernels' own identifiers (a, i, t, k,
    # N, ...); the parser's C text marks each of them with an index
    # placeholder, the pre-loop step sizes aside
    user_names = sorted({s[0] for s in slots} | {r.name for r in desc.randoms}, key=len, reverse=True)
    bare_names = sorted({pl["name"] for pl in preloops}, key=len, reverse=True)
    name_re = re.compile(r"(?<![\w.])(" + "|".join(re.escape(n) for n in user_names) + r")(?=%\()") \
        if user_names else None
    bare_re = re.compile(r"(?<![\w.])(" + "|".join(re.escape(n) for n in bare_names) + r")(?![\w])") \
        if bare_names else None

    def mangle(code: str) -> str:
        if name_re is not None:
            code = name_re.sub(lambda m: _u(m.group(1)), code)
        if bare_re is not None:
            code = bare_re.sub(lambda m: _u(m.group(1)), code)
        return code

    refractory = desc.refractory if desc.spiking else None
    has_ref = refractory is not None and refractory != 0 and refractory != 0.0

    # -- population-wide attributes -----------------------------------------
    global_names = [s[0] for s in slots if not s[1]]
    global_loads = "\n".join(
        f"    const T {_u(n)} = ((const T*)a.slot[{slot_of[n]}])[0];" for n in global_names)
    global_refs = "\n".join(
        f"    T& {_u(n)} = ((T*)a.slot[{slot_of[n]}])[0];" for n in global_names)
    local_names = [s[0] for s in slots if s[1]]
    local_ptrs = "\n".join(
        f"    T* __restrict__ {_u(n)} = (T*)a.slot[{slot_of[n]}];" for n in local_names)

    def subst(code: str) -> str:
        code = mangle(code)
        for r in desc.randoms:
            code = code.replace(_u(r.name) + "%(local_index)s", _u(r.name))
            code = code.replace(_u(r.name) + "%(global_index)s", _u(r.name))
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
            out.append(f"const T {_u(r.name)} = ahip::random_value<T>(a, {idx}, {k}, t, "
                       f"{DISTRIBUTIONS[r.dist]}, {a0!r}, {a1!r});")
        return "\n".join(out)

    global_eqs = subst(equation_code(desc, "global", None))
    has_global = bool(global_eqs.strip())
    mask = "in_ref" if has_ref else None
    local_eqs = subst(equation_code(desc, "local", mask))
    # divisions of the neuron equations by a population-wide value: the
    # reciprocal steps once per kernel (ahip::DivConst)
    div_names = [n for n in global_names
                 if re.search(r"(?<!/)/\s*" + re.escape(_u(n)) + r"(?![\w\[(])", local_eqs)]
    for n in div_names:
        local_eqs = re.sub(r"(?<!/)(/\s*)" + re.escape(_u(n)) + r"(?![\w\[(])", r"\g<1>" + _u(n) + "__dc",
                           local_eqs)
    div_consts = "\n".join(f"    const ahip::DivConst<T> {_u(n)}__dc = ahip::div_const<T>({_u(n)});"
                            for n in div_names)

    refrac_in = "const int in_ref = (a.refractory_remaining[i] > 0) ? 0 : 1;" if has_ref else ""
    refrac_dec = "a.refractory_remaining[i] -= (1 - in_ref);" if has_ref else ""
    gather = ""
    publish = ""
    if desc.spiking:
        if isinstance(refractory, str):
            rvar = refractory
            idx = "[i]" if any(s[0] == rvar and s[1] for s in slots) else ""
            refrac_set = f"a.refractory_remaining[i] = int({_u(rvar)}{idx}/dt);"
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
            lines.append(f"    const double {_u(n)} = g[{slot_of[n]}]; (void){_u(n)};")
        for pl in preloops:
            val = mangle(pl["value"]).replace("%(global_index)s", "").replace("%(local_index)s", "")
            lines.append(f"    out[{slot_of[pl['name']]}] = {val};")
        lines.append("}")
        preloop_fn = "\n".join(lines)
        preloop_ptr = "preloop"
    else:
        preloop_fn = ""
        preloop_ptr = "nullptr"

    # persistent steps: spiking models without population-wide equations.
    # LDS holds the local slots the statements write and the conductances
    # (g_<target>), in slot order; local slots they only read stay in global
    # memory (constant during the launch, L2-resident).  The statements see
    # the neuron's values in registers: loaded before them, stored after
    # them when written.
    if desc.spiking and not has_global:
        rr = lambda c: c.replace("a.refractory_remaining", "refrac_rem").replace("a.refractory[i]", "AHIP_REFRACTORY_I")
        parts = {"refrac_in": rr(refrac_in), "local_rand": rand_code("local", "i"), "local_eqs": rr(local_eqs),
                 "refrac_dec": rr(refrac_dec), "gather": rr(gather)}
        joined = "\n".join(parts.values())
        ref_pat = lambda c: r"(?<![\w.])" + re.escape(c) + r"\[i\]"
        writes = lambda c: re.search(ref_pat(c) + r"\s*(?:[-+*/]?=(?!=)|\+\+|--)", joined) is not None
        reads = lambda c: re.search(ref_pat(c), joined) is not None
        pslots = [n for n in local_names if n.startswith("g_") or writes(_u(n))]
        ro_slots = [n for n in local_names if n not in pslots and reads(_u(n))]
        persist_ptrs_code = "\n".join(
            [f"    T* __restrict__ {_u(n)} = lds + {k} * (size_t)N;" for k, n in enumerate(pslots)] +
            [f"    const T* __restrict__ {_u(n)} = (const T*)a.slot[{slot_of[n]}];" for n in ro_slots])
        reg_in, reg_out = [], []
        for c, ty in ([(_u(n), "T") for n in pslots + ro_slots] +
                      ([("refrac_rem", "int32_t")] if has_ref else [])):
            if not reads(c):
                continue
            reg_in.append(f"            {ty} {c}__r = {c}[i];")
            if writes(c):
                reg_out.append(f"            {c}[i] = {c}__r;")
            for key in parts:
                parts[key] = re.sub(ref_pat(c), c + "__r", parts[key])
        persist_fn = _PERSIST_KERNEL % {
            "n_pslots": len(pslots),
            "p_refr": "true" if has_ref else "false",
            "p_load": "\n".join(f"        lds[{k} * (size_t)N + i] = ((const T*)a.slot[{slot_of[n]}])[i];"
                                for k, n in enumerate(pslots)),
            "p_store": "\n".join(f"        ((T*)a.slot[{slot_of[n]}])[i] = lds[{k} * (size_t)N + i];"
                                 for k, n in enumerate(pslots)),
            "pslot_list": ", ".join(str(slot_of[n]) for n in pslots) or "0",
            # the population-wide values in vector registers for the whole
            # launch (the kernel's scalar registers are taken)
            "global_loads": "\n".join(
                [f"    T {_u(n)} = ((const T*)a.slot[{slot_of[n]}])[0];" for n in global_names] +
                [f"    ahip::in_vgpr({_u(n)});" for n in global_names]),
            "div_consts": "\n".join(
                [f"    ahip::DivConst<T> {_u(n)}__dc = ahip::div_const<T>({_u(n)});" for n in div_names] +
                [f"    ahip::in_vgpr({_u(n)}__dc.v);\n    ahip::in_vgpr({_u(n)}__dc.r);\n"
                 f"    ahip::in_vgpr({_u(n)}__dc.span);" for n in div_names]),
            "local_ptrs": persist_ptrs_code,
            "refrac_in": _indent(parts["refrac_in"], 12),
            "local_rand": _indent(parts["local_rand"], 12),
            "local_eqs": _indent(parts["local_eqs"], 12),
            "refrac_dec": _indent(parts["refrac_dec"], 12),
            "gather": _indent(parts["gather"], 12),
            "reg_in": "\n".join(reg_in),
            "reg_out": "\n".join(reg_out),
        }
        persist_ptrs = f"persist_f64, persist_f32, {len(pslots)}, kPersistSlotsHost"
    else:
        persist_fn = ""
        persist_ptrs = "nullptr, nullptr, 0, nullptr"

    # the per-kernel update: the statements work on register copies of the
    # neuron's values (X[i] -> X__r); the local slots other than the
    # conductances (g_*, which the ordered delivery may write) and the
    # refractory counter are loaded before the delivery, the conductances
    # after it; written values are stored after the statements
    lparts = {"refrac_in": refrac_in, "local_rand": rand_code("local", "i"), "local_eqs": local_eqs,
              "refrac_dec": refrac_dec, "gather": gather}
    for key in lparts:
        lparts[key] = lparts[key].replace("a.refractory_remaining[i]", "refrac_rem[i]")
    ljoined = "\n".join(lparts.values())
    lref = lambda c: r"(?<![\w.])" + re.escape(c) + r"\[i\]"
    lwrites = lambda c: re.search(lref(c) + r"\s*(?:[-+*/]?=(?!=)|\+\+|--)", ljoined) is not None
    # read: any reference other than the target of a plain assignment
    lreads = lambda c: re.search(lref(c) + r"(?!\s*=(?!=))", ljoined) is not None
    reg_pre, reg_post, reg_out = [], [], []
    kind_of = {s_[0]: s_[3] for s_ in slots}
    cands = [(_u(n), "T", f"{_u(n)}[i]", n.startswith("g_")) for n in local_names]
    if has_ref:
        cands.append(("refrac_rem", "int32_t", "a.refractory_remaining[i]", False))
    for n in local_names:
        # a psp sum the psp may leave in panel partials: completed (and
        # stored) by the update, read or not
        if kind_of.get(n) == 2 and not lreads(_u(n)):
            reg_pre.append(f"    if (a.n_psum > 0 && i < a.size) (void)ahip::psp_sum<T>(a, {slot_of[n]}, {_u(n)}, i);")
    for c, ty, mem, after in cands:
        if not lreads(c):
            if lwrites(c):
                # written without being read: stays a store
                for key in lparts:
                    lparts[key] = re.sub(lref(c), mem.replace("\\", "\\\\"), lparts[key])
            continue
        load = mem
        name = {_u(n): n for n in local_names}.get(c)
        if name is not None and kind_of.get(name) == 2:
            load = f"ahip::psp_sum<T>(a, {slot_of[name]}, {c}, i)"
        (reg_post if after else reg_pre).append(f"    {ty} {c}__r = i < a.size ? {load} : {ty}(0);")
        if lwrites(c):
            reg_out.append(f"        {mem} = {c}__r;")
        for key in lparts:
            lparts[key] = re.sub(lref(c), c + "__r", lparts[key])

    body = _PRELUDE % {"name": desc.name}
    if has_global:
        body += _GLOBAL_KERNEL % {
            "global_refs": global_refs,
            "global_rand": _indent(rand_code("global", "0"), 4),
            "global_eqs": _indent(global_eqs, 4),
        }
    src_wo_hash = _LOCAL_KERNEL % {
        "global_loads": global_loads,
        "div_consts": div_consts,
        "local_ptrs": local_ptrs,
        "refrac_in": _indent(lparts["refrac_in"], 8),
        "local_rand": _indent(lparts["local_rand"], 8),
        "local_eqs": _indent(lparts["local_eqs"], 8),
        "refrac_dec": _indent(lparts["refrac_dec"], 8),
        "gather": _indent(lparts["gather"], 8),
        "reg_pre": "\n".join(reg_pre),
        "reg_post": "\n".join(reg_post),
        "reg_out": "\n".join(reg_out),
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
        "persist_fn": persist_fn,
        "persist_ptrs": persist_ptrs,
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
