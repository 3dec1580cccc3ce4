// ahip_device.hpp -- device helpers shared by the core kernels and by the
// generated neuron-model libraries (CDNA4 / gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "annarchy_hip_model.h"

namespace ahip {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Counter-based generator: Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).  Key = (seed low,
// seed high ^ stream), counter = (neuron, random-variable index, step low,
// step high).  Every (neuron, variable, step) gets an independent draw, so a
// run does not depend on the launch shape.
// ---------------------------------------------------------------------------
struct u32x4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(M0, c.x, hi0, lo0);
        mulhilo(M1, c.z, hi1, lo1);
        u32x4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// 53-bit uniform in [0, 1) from two 32-bit words.
__host__ __device__ inline double u01_53(uint32_t a, uint32_t b) {
    uint64_t v = (((uint64_t)a << 32) | b) >> 11;
    return (double)v * (1.0 / 9007199254740992.0);
}

// Draw for (neuron i, variable k, step t): dist 0 = Uniform(a, b),
// 1 = Normal(a, b) by the Box-Muller transform (cosine branch).
__device__ inline double draw(uint64_t seed, uint32_t stream, int32_t i, int32_t k, int64_t t,
                              int32_t dist, double a, double b) {
    u32x4 c{(uint32_t)i, (uint32_t)k, (uint32_t)(uint64_t)t, (uint32_t)((uint64_t)t >> 32)};
    u32x4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32) ^ stream);
    if (dist == 0) {
        return a + (b - a) * u01_53(r.x, r.y);
    }
    double u1 = 1.0 - u01_53(r.x, r.y);  // (0, 1]
    double u2 = u01_53(r.z, r.w);
    double rad = sqrt(-2.0 * log(u1));
    return a + b * (rad * cospi(2.0 * u2));
}

// ---------------------------------------------------------------------------
// Wave reductions (64 lanes; butterfly over the lanes of a group of width W).
// ---------------------------------------------------------------------------
template <int W, typename T>
__device__ inline T group_sum(T v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

// The same butterfly over aligned groups of w lanes, w a power of two known
// only at run time (uniform across the wave).
__device__ inline double group_sum_rt(double v, int w) {
    for (int o = w / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Kahan-Babuska (Neumaier) compensated accumulator for FP32 psp sums.
struct Kahan {
    float s, c;
    __device__ inline void add(float x) {
        float t = s + x;
        if (fabsf(s) >= fabsf(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    __device__ inline float value() const { return s + c; }
};

// Compensated reduction over a lane group: pairs (s, c) are combined with a
// two-sum at every butterfly level.
template <int W>
__device__ inline float group_sum_kahan(float s, float c) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) {
        float so = __shfl_xor(s, o, W);
        float co = __shfl_xor(c, o, W);
        float t = s + so;
        float bp = t - s;
        float err = (s - (t - bp)) + (so - bp);
        s = t;
        c = c + co + err;
    }
    return s + c;
}

__device__ inline float group_sum_kahan_rt(float s, float c, int w) {
    for (int o = w / 2; o > 0; o >>= 1) {
        float so = __shfl_xor(s, o, kWave);
        float co = __shfl_xor(c, o, kWave);
        float t = s + so;
        float bp = t - s;
        float err = (s - (t - bp)) + (so - bp);
        s = t;
        c = c + co + err;
    }
    return s + c;
}

__device__ inline int64_t sim_time(const ahip_pop_args& a) { return *a.t_base + a.step; }

// The panel partial sums of row r of a sliced-tile psp, added in ascending
// panel order (FP32: compensated), the one sequence both the psp's own
// reduction and an update that completes the sum use.
__device__ inline double panel_sum(const double* __restrict__ part, int32_t n_panels, int64_t stride, int64_t r) {
    double s = 0.0;
    for (int32_t q = 0; q < n_panels; q++) s += part[(int64_t)q * stride + r];
    return s;
}
__device__ inline float panel_sum(const float* __restrict__ part, int32_t n_panels, int64_t stride, int64_t r) {
    Kahan k{0.0f, 0.0f};
    for (int32_t q = 0; q < n_panels; q++) k.add(part[(int64_t)q * stride + r]);
    return k.value();
}

// Psp sum slot `slot` of neuron i for the update: completed from the panel
// partials when the psp left it to the update (and stored, so the slot holds
// it after the step), else the slot's value.
template <typename T>
__device__ inline T psp_sum(const ahip_pop_args& a, int32_t slot, T* x, int32_t i) {
    for (int k = 0; k < a.n_psum && k < AHIP_MAX_PSUM; k++)
        if (a.psum_slot[k] == slot) {
            const T v = panel_sum(static_cast<const T*>(a.psum_part[k]), a.psum_panels[k], a.psum_stride[k], i);
            const T y = a.psum_acc[k] ? x[i] + v : v;
            x[i] = y;
            return y;
        }
    return x[i];
}

// Random variable k of neuron i at the current step: injected stream when one
// is set (row = t - rand_t0, wrapped into the rows available), else Philox.
template <typename T>
__device__ inline T random_value(const ahip_pop_args& a, int32_t i, int32_t k, int64_t t,
                                 int32_t dist, double pa, double pb) {
    const void* inj = a.rand_inject[k];
    if (inj) {
        int64_t row = t - a.rand_t0;
        int64_t n = a.rand_inject_steps;
        row %= n;
        if (row < 0) row += n;
        return ((const T*)inj)[row * (int64_t)a.size + i];
    }
    return (T)draw(a.seed, a.stream_id, i, k, t, dist, pa, pb);
}

// Spike flags of one update block: wave ballot -> one 64-bit word per wave,
// and the number of spikes of the block (256 threads = 4 waves).
__device__ inline void publish_spikes(const ahip_pop_args& a, bool spk) {
    __shared__ int32_t wave_cnt[AHIP_UPDATE_BLOCK / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    uint64_t m = __ballot(spk);
    const int64_t word = (int64_t)blockIdx.x * (AHIP_UPDATE_BLOCK / kWave) + wid;
    if (lane == 0) {
        if (word * kWave < a.size) a.spike_mask[word] = m;
        wave_cnt[wid] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t s = 0;
#pragma unroll
        for (int w = 0; w < AHIP_UPDATE_BLOCK / kWave; w++) s += wave_cnt[w];
        a.block_counts[blockIdx.x] = s;
    }
}

// ---- global-memory views -------------------------------------------------
// A pointer read from a table in memory is generic to the compiler, which
// then emits flat instructions (counted against the LDS wait counter as
// well); these views of pointers known to address device memory give
// global ones.
template <typename V>
using gptr = __attribute__((address_space(1))) V*;
template <typename V>
__device__ inline gptr<V> as_global(V* p) {
    return (gptr<V>)p;
}
template <typename V>
__device__ inline gptr<const V> as_global(const V* p) {
    return (gptr<const V>)p;
}

// ---- ordered spike delivery ----------------------------------------------
// A post neuron's events of one step are records (key, weight) of a bucket;
// adding their weights in ascending key (projection by projection in
// creation order, in ascending pre-synaptic rank) reproduces every rounding
// of the reference's CSRC loop g[post] += w
// (generator/Projection/OpenMP/CSR.py:870-885).  One record is one store of
// the scatter and one load of the delivery.
template <typename T>
struct BucketRec;
template <>
struct alignas(16) BucketRec<double> {
    uint32_t key, pad;
    double w;
};
template <>
struct alignas(8) BucketRec<float> {
    uint32_t key;
    float w;
};

static_assert(sizeof(BucketRec<double>) == 16 && sizeof(BucketRec<float>) == 8,
              "the runtime sizes the buckets by these records");

// record `at` of a bucket in device memory: one 16-byte (FP64) or 8-byte
// (FP32) store
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
__device__ inline void store_rec(BucketRec<double>* bucket, int64_t at, uint32_t key, double w) {
    const uint64_t bits = (uint64_t)__double_as_longlong(w);
    const u32x4_t v = {key, 0u, (uint32_t)bits, (uint32_t)(bits >> 32)};
    ((gptr<u32x4_t>)(reinterpret_cast<u32x4_t*>(bucket)))[at] = v;
}
__device__ inline void store_rec(BucketRec<float>* bucket, int64_t at, uint32_t key, float w) {
    const u32x2_t v = {key, (uint32_t)__float_as_uint(w)};
    ((gptr<u32x2_t>)(reinterpret_cast<u32x2_t*>(bucket)))[at] = v;
}

constexpr int kDeliverSort = 4;

// acc + the weights of records rec[b .. b + L) in ascending key.  Up to 4
// records are sorted in registers by a bitonic network on the keys (a
// neuron rarely receives more events in one step, and a small network keeps
// the update kernel's registers, hence its occupancy, low); a longer bucket
// is walked by repeated minimum, O(L^2) reads of the keys.
template <typename T>
__device__ inline T ordered_sum(T acc, const BucketRec<T>* __restrict__ rec, int64_t b, int32_t L) {
    if (L == 1) return acc + rec[b].w;
    if (L <= kDeliverSort) {
        uint32_t k[kDeliverSort];
        T v[kDeliverSort];
#pragma unroll
        for (int m = 0; m < kDeliverSort; m++) {
            const BucketRec<T> r = m < L ? rec[b + m] : BucketRec<T>{};
            k[m] = m < L ? r.key : 0xffffffffu;
            v[m] = m < L ? r.w : T(0);
        }
#pragma unroll
        for (int s = 2; s <= kDeliverSort; s <<= 1)
#pragma unroll
            for (int j = s >> 1; j > 0; j >>= 1)
#pragma unroll
                for (int i = 0; i < kDeliverSort; i++) {
                    const int l = i ^ j;
                    if (l > i) {
                        const bool up = (i & s) == 0;
                        const bool swap = up ? k[i] > k[l] : k[i] < k[l];
                        const uint32_t ka = k[i], kb = k[l];
                        const T va = v[i], vb = v[l];
                        k[i] = swap ? kb : ka;
                        k[l] = swap ? ka : kb;
                        v[i] = swap ? vb : va;
                        v[l] = swap ? va : vb;
                    }
                }
#pragma unroll
        for (int m = 0; m < kDeliverSort; m++)
            if (m < L) acc += v[m];
        return acc;
    }
    uint32_t last = 0;
    for (int32_t m = 0; m < L; m++) {
        uint32_t best = 0xffffffffu;
        int64_t at = b;
        for (int32_t q = 0; q < L; q++) {
            const uint32_t x = rec[b + q].key;
            if ((m == 0 || x > last) && x < best) {
                best = x;
                at = b + q;
            }
        }
        acc += rec[at].w;
        last = best;
    }
    return acc;
}

// The ordered conductances of this population, neuron i: add its buckets
// and clear their counts (called by the update kernel before the equations
// read the conductances).  The heads of every conductance are loaded before
// any bucket is summed.
template <typename T>
__device__ inline void deliver_buckets(const ahip_pop_args& a, int32_t i) {
    uint64_t h[AHIP_MAX_DELIVER], hn[AHIP_MAX_DELIVER];
#pragma unroll
    for (int k = 0; k < AHIP_MAX_DELIVER; k++) {
        h[k] = k < a.n_deliver ? a.deliver[k].head[i] : 0;
        hn[k] = k < a.n_deliver ? a.deliver[k].head[i + 1] : 0;
    }
#pragma unroll
    for (int k = 0; k < AHIP_MAX_DELIVER; k++) {
        const int64_t c = (int64_t)(h[k] & AHIP_HEAD_COUNT);
        if (c == 0) continue;
        const ahip_deliver_desc& d = a.deliver[k];
        const int64_t b = (int64_t)(h[k] >> AHIP_HEAD_SHIFT);
        d.head[i] = h[k] & ~AHIP_HEAD_COUNT;
        // a bucket never holds more events than the neuron's summed in-degree
        const int32_t L = (int32_t)min(c, (int64_t)(hn[k] >> AHIP_HEAD_SHIFT) - b);
        if (L <= 0) continue;
        T* g = static_cast<T*>(d.g);
        g[i] = ordered_sum<T>(g[i], static_cast<const BucketRec<T>*>(d.bucket), b, L);
    }
}

// Keep a value in vector registers (an empty asm that reads and writes it
// there): a loop that would otherwise reload spilled scalar registers every
// iteration reads the copy instead.  No arithmetic changes.
template <typename V>
__device__ inline void in_vgpr(V& x) {
    asm volatile("" : "+v"(x));
}

// ---------------------------------------------------------------------------
// Division by a population-wide value (constant within a kernel): x / d as
// the compiler's IEEE double division computes it (reciprocal, two
// Newton-Raphson steps, quotient, residual, corrected quotient, the last
// step with the operands scaled and the result fixed up only for extreme
// exponents), with the reciprocal steps done once per kernel instead of per
// neuron.  For d in [2^-100, 2^100] and |x| in [2^-600, 2^600) the scaling
// and the fix-up do nothing, so the three remaining operations give the
// division's own result (correctly rounded); other operands (zeros,
// infinities, NaN, extreme exponents) take the division itself.  Operands
// of other types divide as written.
// ---------------------------------------------------------------------------
template <typename T>
struct DivConst {
    T v;
    double r;        // the reciprocal of (double) v after the Newton-Raphson steps
    uint32_t span;   // 1200: the quotient steps serve |x| in [2^-600, 2^600); 0: never
};

template <typename T>
__device__ inline DivConst<T> div_const(T v) {
    DivConst<T> c;
    c.v = v;
    c.r = 0.0;
    c.span = 0u;
    const double d = (double)v;
    const double ad = __builtin_fabs(d);
    if (ad >= 0x1p-100 && ad <= 0x1p100) {
        const double r0 = __builtin_amdgcn_rcp(d);
        const double r1 = __builtin_fma(r0, __builtin_fma(-d, r0, 1.0), r0);
        c.r = __builtin_fma(r1, __builtin_fma(-d, r1, 1.0), r1);
        c.span = 1200u;
    }
    return c;
}

template <typename X, typename T>
__device__ inline auto operator/(X x, const DivConst<T>& c) -> decltype(x / c.v) {
    using R = decltype(x / c.v);
    if constexpr (std::is_same<R, double>::value) {
        const double xd = (double)x;
        // |x| in [2^-600, 2^600): biased exponent in [423, 1622], one
        // integer test on the high word (span 0 fails it for any x)
        const uint32_t ex = ((uint32_t)((uint64_t)__double_as_longlong(xd) >> 52)) & 0x7ffu;
        if (ex - 423u < c.span) {
            const double q = xd * c.r;
            return __builtin_fma(__builtin_fma(-(double)c.v, q, xd), c.r, q);
        }
        return xd / (double)c.v;
    } else {
        return x / c.v;
    }
}

}  // namespace ahip
