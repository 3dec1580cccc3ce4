// k_delay.hip -- synaptic delays on the device.
//
// The reference keeps deques of past values: `_delayed_<var>` for rate-coded
// populations (Population/OpenMPTemplates.py:181-220) and `_delayed_spike`
// for spiking ones (Population/OpenMPGenerator.py:433-450), both pushed to
// the front after the population update; a projection with delay d reads
// element d-1 (Projection/OpenMP/CSR.py:962-963, OpenMPGenerator.py:1238-1249).
// Here a queue of depth D is a ring indexed by the simulation step: slot
// t mod D holds the value at the end of step t, so element d-1 of the
// reference's deque at step t is slot (t - d) mod D.  Every kernel reads t
// from device memory (t_base + step), so captured graphs replay correctly.
//
// Per-synapse spiking delays follow the reference's ring of event lists
// (Projection/OpenMP/LIL.py:946-1017): a spike of step t-1 on a synapse with
// delay d is delivered at step t-1+d.  The ring holds, per delivery slot and
// post neuron, the number of events (one weight per projection: the weight is
// then added that many times, in sequence, as the reference adds it event by
// event) or the sum of their weights (one weight per synapse).
#include "ahip_device.hpp"
#include "internal.hpp"

namespace ahip {
namespace {

constexpr int kBlock = 256;

__device__ inline int64_t ring_slot(int64_t t, int32_t D) {
    const int64_t m = t % D;
    return m < 0 ? m + D : m;
}

// hist[(t mod D)][i] = src[i]
template <typename T>
__global__ __launch_bounds__(kBlock) void push_values(const T* __restrict__ src, T* __restrict__ hist,
                                                      int32_t n, int32_t D, const int64_t* __restrict__ t_base,
                                                      int32_t step) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t slot = ring_slot(*t_base + step, D);
    hist[slot * n + i] = src[i];
}

// dst[i] = hist[((t - d) mod D)][i]
template <typename T>
__global__ __launch_bounds__(kBlock) void gather_delayed(const T* __restrict__ hist, T* __restrict__ dst,
                                                         int32_t n, int32_t D, int32_t d,
                                                         const int64_t* __restrict__ t_base, int32_t step) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t slot = ring_slot(*t_base + step - d, D);
    dst[i] = hist[slot * n + i];
}

// every slot of the queue = src (reset)
template <typename T>
__global__ __launch_bounds__(kBlock) void fill_hist(const T* __restrict__ src, T* __restrict__ hist, int32_t n,
                                                    int32_t D) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)n * D) return;
    hist[i] = src[i % n];
}

// spike-list ring: slot (t mod D) = the step's list and count
__global__ __launch_bounds__(kBlock) void push_spikes(const int32_t* __restrict__ spiked,
                                                      const int32_t* __restrict__ count,
                                                      int32_t* __restrict__ ring, int32_t* __restrict__ ring_count,
                                                      int32_t n, int32_t D, const int64_t* __restrict__ t_base,
                                                      int32_t step) {
    const int64_t slot = ring_slot(*t_base + step, D);
    const int32_t c = *count;
    for (int32_t k = blockIdx.x * kBlock + threadIdx.x; k < c; k += gridDim.x * kBlock)
        ring[slot * n + k] = spiked[k];
    if (blockIdx.x == 0 && threadIdx.x == 0) ring_count[slot] = c;
}

// Rate psp with one delay per synapse (CSR rows, L lanes per row, forward
// order): y[row] = sum_j w[j] * hist[((t - delay[j]) mod D)][col[j]].
template <typename T, int L>
__global__ __launch_bounds__(kBlock) void csr_delayed(int32_t n_rows, int32_t n_cols, int32_t D,
                                                      const int64_t* __restrict__ row_ptr,
                                                      const int32_t* __restrict__ col,
                                                      const T* __restrict__ w, const int16_t* __restrict__ delay,
                                                      const T* __restrict__ hist, T* __restrict__ y, int32_t acc,
                                                      const int64_t* __restrict__ t_base, int32_t step) {
    const int64_t row = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / L;
    const int lane = threadIdx.x & (L - 1);
    if (row >= n_rows) return;
    const int64_t t = *t_base + step;
    const int64_t beg = row_ptr[row], end = row_ptr[row + 1];
    T s;
    if constexpr (sizeof(T) == 8) {
        double a = 0.0;
        for (int64_t j = beg + lane; j < end; j += L)
            a += w[j] * hist[ring_slot(t - delay[j], D) * n_cols + col[j]];
        s = group_sum<L>(a);
    } else {
        Kahan k{0.0f, 0.0f};
        for (int64_t j = beg + lane; j < end; j += L)
            k.add(w[j] * hist[ring_slot(t - delay[j], D) * n_cols + col[j]]);
        s = group_sum_kahan<L>(k.s, k.c);
    }
    if (lane == 0) {
        if (acc)
            y[row] += s;
        else
            y[row] = s;
    }
}

// Per-synapse spiking delays: the spikes of step t-1 (the current list) are
// binned into delivery slot (t - 1 + d) mod D.  One wave per spike.
template <typename T, bool SINGLE_W>
__global__ __launch_bounds__(kBlock) void ring_scatter(const int32_t* __restrict__ spiked,
                                                       const int32_t* __restrict__ count, int32_t n_cols,
                                                       int32_t n_post, int32_t D,
                                                       const int64_t* __restrict__ col_ptr,
                                                       const int32_t* __restrict__ row_idx,
                                                       const T* __restrict__ w_col,
                                                       const int16_t* __restrict__ delay_col,
                                                       void* __restrict__ ring,
                                                       const int64_t* __restrict__ t_base, int32_t step) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int32_t n = *count;
    const int64_t t = *t_base + step;
    const int64_t waves = (int64_t)gridDim.x * (kBlock / kWave);
    for (int64_t s = (int64_t)blockIdx.x * (kBlock / kWave) + wid; s < n; s += waves) {
        const int32_t pre = spiked[s];
        if (pre < 0 || pre >= n_cols) continue;
        const int64_t b = col_ptr[pre], e = col_ptr[pre + 1];
        for (int64_t syn = b + lane; syn < e; syn += kWave) {
            const int64_t slot = ring_slot(t - 1 + delay_col[syn], D);
            const int64_t at = slot * n_post + row_idx[syn];
            if (SINGLE_W)
                atomicAdd(static_cast<int32_t*>(ring) + at, 1);
            else
                unsafeAtomicAdd(static_cast<T*>(ring) + at, w_col[syn]);
        }
    }
}

// Delivery at step t: g[i] += the events of slot t mod D, then clear it.
template <typename T, bool SINGLE_W>
__global__ __launch_bounds__(kBlock) void ring_deliver(void* __restrict__ ring, T* __restrict__ g, T w_const,
                                                       int32_t n_post, int32_t D,
                                                       const int64_t* __restrict__ t_base, int32_t step) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_post) return;
    const int64_t at = ring_slot(*t_base + step, D) * n_post + i;
    if (SINGLE_W) {
        int32_t* r = static_cast<int32_t*>(ring);
        const int32_t c = r[at];
        if (c) {
            T v = g[i];
            for (int32_t k = 0; k < c; k++) v += w_const;
            g[i] = v;
            r[at] = 0;
        }
    } else {
        T* r = static_cast<T*>(ring);
        const T v = r[at];
        g[i] += v;
        r[at] = T(0);
    }
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

int launch_push_values(const ahip_net* net, const void* src, void* hist, int32_t n, int32_t D, int32_t step,
                       hipStream_t s) {
    if (n == 0) return 0;
    const int64_t* tb = net->t_dev.as<int64_t>();
    if (net->precision == AHIP_F64)
        hipLaunchKernelGGL(push_values<double>, dim3(blocks_for(n)), dim3(kBlock), 0, s, (const double*)src,
                           (double*)hist, n, D, tb, step);
    else
        hipLaunchKernelGGL(push_values<float>, dim3(blocks_for(n)), dim3(kBlock), 0, s, (const float*)src,
                           (float*)hist, n, D, tb, step);
    return (int)hipGetLastError();
}

int launch_gather_delayed(const ahip_net* net, const void* hist, void* dst, int32_t n, int32_t D, int32_t d,
                          int32_t step, hipStream_t s) {
    if (n == 0) return 0;
    const int64_t* tb = net->t_dev.as<int64_t>();
    if (net->precision == AHIP_F64)
        hipLaunchKernelGGL(gather_delayed<double>, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                           (const double*)hist, (double*)dst, n, D, d, tb, step);
    else
        hipLaunchKernelGGL(gather_delayed<float>, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                           (const float*)hist, (float*)dst, n, D, d, tb, step);
    return (int)hipGetLastError();
}

int launch_fill_hist(const ahip_net* net, const void* src, void* hist, int32_t n, int32_t D, hipStream_t s) {
    if (n == 0) return 0;
    if (net->precision == AHIP_F64)
        hipLaunchKernelGGL(fill_hist<double>, dim3(blocks_for((int64_t)n * D)), dim3(kBlock), 0, s,
                           (const double*)src, (double*)hist, n, D);
    else
        hipLaunchKernelGGL(fill_hist<float>, dim3(blocks_for((int64_t)n * D)), dim3(kBlock), 0, s,
                           (const float*)src, (float*)hist, n, D);
    return (int)hipGetLastError();
}

int launch_push_spikes(const ahip_net* net, const Pop& p, int32_t step, hipStream_t s) {
    unsigned nb = blocks_for(p.size);
    if (nb > 64) nb = 64;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(push_spikes, dim3(nb), dim3(kBlock), 0, s, p.spiked.as<int32_t>(),
                       p.spike_count.as<int32_t>(), p.hist.as<int32_t>(), p.hist_count.as<int32_t>(), p.size,
                       p.depth, net->t_dev.as<int64_t>(), step);
    return (int)hipGetLastError();
}

int launch_rate_psp_delayed(const ahip_net* net, const Proj& p, const Pop& pre, void* target, int32_t step,
                            hipStream_t s) {
    const RateLayout& L = p.rate;
    const int32_t acc = p.accumulate ? 1 : 0;
    const int64_t* tb = net->t_dev.as<int64_t>();
    const int lanes = L.lanes;
    const int64_t nb = ((int64_t)p.n_rows * lanes + kBlock - 1) / kBlock;
    if (nb == 0) return 0;
#define AHIP_DL_CASE(T, N)                                                                              \
    case N:                                                                                             \
        hipLaunchKernelGGL((csr_delayed<T, N>), dim3(nb), dim3(kBlock), 0, s, p.n_rows, p.n_cols,       \
                           pre.depth, L.ptr.as<int64_t>(), L.idx.as<int32_t>(), L.val.as<T>(),          \
                           p.delay_dev.as<int16_t>(), pre.hist.as<T>(), (T*)target, acc, tb, step);     \
        break;
    if (net->precision == AHIP_F64) {
        switch (lanes) {
            AHIP_DL_CASE(double, 1)
            AHIP_DL_CASE(double, 2)
            AHIP_DL_CASE(double, 4)
            AHIP_DL_CASE(double, 8)
            AHIP_DL_CASE(double, 16)
            AHIP_DL_CASE(double, 32)
            AHIP_DL_CASE(double, 64)
            default:
                return (int)hipErrorInvalidValue;
        }
    } else {
        switch (lanes) {
            AHIP_DL_CASE(float, 1)
            AHIP_DL_CASE(float, 2)
            AHIP_DL_CASE(float, 4)
            AHIP_DL_CASE(float, 8)
            AHIP_DL_CASE(float, 16)
            AHIP_DL_CASE(float, 32)
            AHIP_DL_CASE(float, 64)
            default:
                return (int)hipErrorInvalidValue;
        }
    }
#undef AHIP_DL_CASE
    return (int)hipGetLastError();
}

int launch_spike_ring(const ahip_net* net, const Proj& p, const Pop& pre, void* target, int32_t step,
                      hipStream_t s) {
    const int64_t* tb = net->t_dev.as<int64_t>();
    int64_t nb = ((int64_t)pre.size + 3) / 4;
    if (nb > 512) nb = 512;
    if (nb < 1) nb = 1;
    const unsigned nbd = blocks_for(p.n_rows);
    const int32_t D = p.max_delay;
    if (net->precision == AHIP_F64) {
        if (p.single_w) {
            hipLaunchKernelGGL((ring_scatter<double, true>), dim3(nb), dim3(kBlock), 0, s, pre.spiked.as<int32_t>(),
                               pre.spike_count.as<int32_t>(), p.n_cols, p.n_rows, D, p.col_ptr.as<int64_t>(),
                               p.row_idx.as<int32_t>(), (const double*)nullptr, p.delay_dev.as<int16_t>(),
                               p.ring.p, tb, step);
            if (nbd) hipLaunchKernelGGL((ring_deliver<double, true>), dim3(nbd), dim3(kBlock), 0, s, p.ring.p,
                                        (double*)target, p.w_const, p.n_rows, D, tb, step);
        } else {
            hipLaunchKernelGGL((ring_scatter<double, false>), dim3(nb), dim3(kBlock), 0, s, pre.spiked.as<int32_t>(),
                               pre.spike_count.as<int32_t>(), p.n_cols, p.n_rows, D, p.col_ptr.as<int64_t>(),
                               p.row_idx.as<int32_t>(), p.w_col.as<double>(), p.delay_dev.as<int16_t>(),
                               p.ring.p, tb, step);
            if (nbd) hipLaunchKernelGGL((ring_deliver<double, false>), dim3(nbd), dim3(kBlock), 0, s, p.ring.p,
                                        (double*)target, 0.0, p.n_rows, D, tb, step);
        }
    } else {
        if (p.single_w) {
            hipLaunchKernelGGL((ring_scatter<float, true>), dim3(nb), dim3(kBlock), 0, s, pre.spiked.as<int32_t>(),
                               pre.spike_count.as<int32_t>(), p.n_cols, p.n_rows, D, p.col_ptr.as<int64_t>(),
                               p.row_idx.as<int32_t>(), (const float*)nullptr, p.delay_dev.as<int16_t>(),
                               p.ring.p, tb, step);
            if (nbd) hipLaunchKernelGGL((ring_deliver<float, true>), dim3(nbd), dim3(kBlock), 0, s, p.ring.p,
                                        (float*)target, (float)p.w_const, p.n_rows, D, tb, step);
        } else {
            hipLaunchKernelGGL((ring_scatter<float, false>), dim3(nb), dim3(kBlock), 0, s, pre.spiked.as<int32_t>(),
                               pre.spike_count.as<int32_t>(), p.n_cols, p.n_rows, D, p.col_ptr.as<int64_t>(),
                               p.row_idx.as<int32_t>(), p.w_col.as<float>(), p.delay_dev.as<int16_t>(),
                               p.ring.p, tb, step);
            if (nbd) hipLaunchKernelGGL((ring_deliver<float, false>), dim3(nbd), dim3(kBlock), 0, s, p.ring.p,
                                        (float*)target, 0.0f, p.n_rows, D, tb, step);
        }
    }
    return (int)hipGetLastError();
}

}  // namespace ahip
