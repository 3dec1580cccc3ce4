// k_spike.hip -- spike compaction and event-driven spike propagation.
#include "ahip_device.hpp"
#include "internal.hpp"

#include <cstring>

namespace ahip {
namespace {

constexpr int kCompactThreads = 1024;
constexpr int kCompactPerThread = 4;
constexpr int kCompactChunk = kCompactThreads * kCompactPerThread;  // 4096 neurons
static_assert(kCompactChunk % AHIP_UPDATE_BLOCK == 0, "chunk must cover whole update blocks");

// Block-wide exclusive scan of one int per thread (1024 threads, 16 waves).
__device__ inline int32_t block_exclusive_scan(int32_t v, int32_t* lds_waves, int32_t& total) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    int32_t incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int32_t n = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += n;
    }
    if (lane == kWave - 1) lds_waves[wid] = incl;
    __syncthreads();
    if (wid == 0) {
        int32_t wv = lane < kCompactThreads / kWave ? lds_waves[lane] : 0;
        int32_t wi = wv;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            int32_t n = __shfl_up(wi, o, kWave);
            if (lane >= o) wi += n;
        }
        if (lane < kCompactThreads / kWave) lds_waves[lane] = wi - wv;  // exclusive wave offsets
        if (lane == kCompactThreads / kWave - 1) lds_waves[kCompactThreads / kWave] = wi;
    }
    __syncthreads();
    total = lds_waves[kCompactThreads / kWave];
    return lds_waves[wid] + incl - v;
}

// Spike compaction, ascending neuron order (the order of the CPU gather,
// Population/OpenMPGenerator.py:1124-1186 with a static schedule).
// Workgroup b covers neurons [b*4096, (b+1)*4096): its global offset is the
// sum of the spike counts of all update blocks before it (read from
// block_counts written by the update kernel), then a block-wide scan over the
// per-thread popcounts of the spike bit mask places every index.
// With a spike monitor (rec_counts != NULL) the kernel also appends the list
// to it: the monitor's index buffer holds the lists of the recorded steps
// back to back; fill[s & 1] is the number of indices before the list of
// recorded step s (written by step s - 1's launch, so no workgroup of this
// launch reads a value another one writes), the last workgroup writes
// counts[s] and fill[(s + 1) & 1].
struct SpikeRec {
    int32_t* counts;          // int32[capacity] spikes per recorded step, NULL: no monitor
    int32_t* idx;             // int32[idx_capacity]
    int64_t* fill;            // int64[2]
    int64_t idx_capacity;
    int64_t t_start;
    int64_t capacity;
};

// One population's compaction: its buffers and its first spike monitor.
struct CompactArgs {
    int32_t size;
    int32_t nb;                   // workgroups (chunks of kCompactChunk neurons)
    const uint64_t* mask;
    const int32_t* bcnt;
    int32_t* spiked;
    int32_t* count;
    SpikeRec rec;
};

// Chunk b of a population's compaction at step t_base + step.
__device__ inline void compact_chunk(const CompactArgs& c, int32_t b, const int64_t* __restrict__ t_base,
                                     int32_t step) {
    __shared__ int32_t lds_waves[kCompactThreads / kWave + 1];
    __shared__ int32_t lds_base[kCompactThreads / kWave];
    const int32_t size = c.size;
    const SpikeRec& rec = c.rec;
    const int32_t chunk0 = b * kCompactChunk;
    const int32_t ub0 = chunk0 / AHIP_UPDATE_BLOCK;  // first update block of the chunk
    // every global load first (they do not depend on one another): this
    // thread's mask word, its share of the block counts before the chunk,
    // the monitor's fill count
    const int32_t n0 = chunk0 + threadIdx.x * kCompactPerThread;
    const uint64_t word = n0 < size ? c.mask[n0 / 64] : 0ull;
    int32_t part = 0;
    for (int32_t k = threadIdx.x; k < ub0; k += kCompactThreads) part += c.bcnt[k];
    int64_t slot = -1, f = 0;
    if (rec.counts) {
        slot = *t_base + step - rec.t_start;
        if (slot >= 0 && slot < rec.capacity) f = rec.fill[slot & 1];
        else slot = -1;
    }
    // base = sum(bcnt[0..ub0))
    part = group_sum<kWave>(part);
    if ((threadIdx.x & (kWave - 1)) == 0) lds_base[threadIdx.x / kWave] = part;
    __syncthreads();
    int32_t base = 0;
#pragma unroll
    for (int w = 0; w < kCompactThreads / kWave; w++) base += lds_base[w];
    // this thread's 4 neurons
    uint32_t bits = 0;
    if (n0 < size) {
        bits = (uint32_t)(word >> (n0 % 64)) & 0xFu;
        const int32_t valid = size - n0;
        if (valid < kCompactPerThread) bits &= (1u << valid) - 1u;
    }
    int32_t total;
    const int32_t excl = block_exclusive_scan(__popc(bits), lds_waves, total);
    int32_t pos = base + excl;
#pragma unroll
    for (int k = 0; k < kCompactPerThread; k++)
        if (bits & (1u << k)) {
            c.spiked[pos] = n0 + k;
            if (slot >= 0 && f + pos < rec.idx_capacity) rec.idx[f + pos] = n0 + k;
            pos++;
        }
    // grand total: the last chunk knows base + its own total
    if (b == c.nb - 1 && threadIdx.x == 0) {
        *c.count = base + total;
        if (slot >= 0) {
            rec.counts[slot] = base + total;
            rec.fill[(slot + 1) & 1] = f + base + total;
        }
    }
}

// The compactions of up to kMaxCompact populations in one launch: workgroup
// ranges one after another, population k's chunks at [b0[k], b0[k + 1]).
constexpr int kMaxCompact = 8;
struct CompactBatch {
    CompactArgs c[kMaxCompact];
    int32_t b0[kMaxCompact + 1];
    int32_t n;
};

__global__ __launch_bounds__(kCompactThreads) void compact_spikes(const CompactBatch B,
                                                                  const int64_t* __restrict__ t_base, int32_t step) {
    const int32_t bid = __builtin_amdgcn_readfirstlane((int32_t)blockIdx.x);
    int32_t k = 0;
    while (k < B.n - 1 && bid >= B.b0[k + 1]) k++;
    compact_chunk(B.c[k], bid - B.b0[k], t_base, step);
}

// Event-driven propagation over the compacted spike list, CSRC backward view
// (generator/Projection/OpenMP/CSR.py:870-885; CUDA/CSR.py:575-657 uses one
// block per spike with a host-read spike count).  Here the grid is fixed and
// reads the count from device memory, so the step needs no host sync.  Work
// item (spike s, chunk c) goes to one wave: it reads the column range of the
// pre-synaptic neuron and, for every 256-synapse chunk c, c + chunks, ... of
// it, loads 4 post ranks (and weights) per lane before it issues the 4 device
// atomics into the post-synaptic conductances.  With one weight per
// projection all addends are equal, so the result does not depend on the
// order of the atomics.
template <typename T, bool SINGLE_W>
__global__ __launch_bounds__(256) void scatter_spikes(const int32_t* __restrict__ spiked,
                                                      const int32_t* __restrict__ count,
                                                      int32_t n_cols, int32_t chunks,
                                                      const int64_t* __restrict__ col_ptr,
                                                      const int32_t* __restrict__ row_idx,
                                                      const T* __restrict__ w_col, T w_const,
                                                      T* __restrict__ g) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int64_t items = (int64_t)(*count) * chunks;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t it = (int64_t)blockIdx.x * 4 + wid; it < items; it += waves) {
        const int64_t s = it / chunks;
        const int32_t c = (int32_t)(it - s * chunks);
        const int32_t pre = spiked[s];
        if (pre < 0 || pre >= n_cols) continue;
        const int64_t b = col_ptr[pre], e = col_ptr[pre + 1];
        for (int64_t base = b + (int64_t)c * 256; base < e; base += (int64_t)chunks * 256) {
            int32_t post[4];
            T v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int64_t j = base + lane + u * kWave;
                post[u] = j < e ? row_idx[j] : -1;
                v[u] = SINGLE_W ? w_const : (j < e ? w_col[j] : T(0));
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (post[u] >= 0) unsafeAtomicAdd(g + post[u], v[u]);
        }
    }
}

template <typename T>
int launch_scatter_t(const Proj& p, const Pop& pre, void* target, hipStream_t s) {
    // one wave per (spike, chunk of 256 synapses); a wave walks the chunks
    // c, c + chunks, ... of its spike.  The grid covers every spike of the
    // pre population in one pass up to two workgroups per CU, then loops.
    const int32_t ch = p.scatter_chunks;
    int64_t nb = ((int64_t)pre.size * ch + 3) / 4;
    if (nb > 512) nb = 512;
    if (nb < 1) nb = 1;
    if (p.single_w)
        hipLaunchKernelGGL((scatter_spikes<T, true>), dim3(nb), dim3(256), 0, s,
                           pre.spiked.as<int32_t>(), pre.spike_count.as<int32_t>(), p.n_cols, ch,
                           p.col_ptr.as<int64_t>(), p.row_idx.as<int32_t>(), (const T*)nullptr,
                           (T)p.w_const, (T*)target);
    else
        hipLaunchKernelGGL((scatter_spikes<T, false>), dim3(nb), dim3(256), 0, s,
                           pre.spiked.as<int32_t>(), pre.spike_count.as<int32_t>(), p.n_cols, ch,
                           p.col_ptr.as<int64_t>(), p.row_idx.as<int32_t>(), p.w_col.as<T>(),
                           (T)0, (T*)target);
    return (int)hipGetLastError();
}

// All projections of a scatter group in one launch.  Work items of
// descriptor k are (spike s, chunk c) for s < count_k, c < chunks_k; the
// waves of descriptor k take items lw, lw + waves_per, ... and proceed as
// scatter_spikes.  Descriptors of a group have distinct targets.
__device__ inline int64_t list_slot(const ScatterDesc& d, int64_t t) {
    if (d.depth <= 1) return 0;
    const int64_t m = (t - d.delay) % d.depth;
    return m < 0 ? m + d.depth : m;
}

// The descriptors of a group travel by value in the kernel arguments (no
// dependent load of a table before the spike counts); descriptor k owns waves
// [k * waves_per, (k + 1) * waves_per).
struct ScatterArgs {
    ScatterDesc d[kMaxScatter];
    int32_t nd;
    int32_t waves_per;
};

template <typename T>
__global__ __launch_bounds__(256) void scatter_multi(const ScatterArgs A, const int64_t* __restrict__ t_base,
                                                     int32_t step) {
    const int lane = threadIdx.x & (kWave - 1);
    // the wave's index, uniform: the descriptor is read from the kernel
    // arguments with scalar loads
    const int32_t gw = __builtin_amdgcn_readfirstlane((int32_t)blockIdx.x * 4 + (int32_t)(threadIdx.x / kWave));
    const int32_t k = gw / A.waves_per;
    if (k >= A.nd) return;
    const ScatterDesc& d = A.d[k];
    const int32_t lw = gw - k * A.waves_per;
    const int64_t sl = d.depth <= 1 ? 0 : list_slot(d, *t_base + step);
    const auto list = as_global(d.list + sl * d.list_stride);
    // the list entry of the wave's first item loads together with the count
    // (a list never holds more than n_cols entries)
    const int64_t s0 = lw / d.chunks;
    const int32_t pre0 = s0 < d.n_cols ? list[s0] : -1;
    const int64_t items = (int64_t)as_global(d.count)[sl] * d.chunks;
    for (int64_t it = lw; it < items; it += A.waves_per) {
        const int64_t s = it / d.chunks;
        const int32_t c = (int32_t)(it - s * d.chunks);
        const int32_t pre = it == lw ? pre0 : list[s];
        if (pre < 0 || pre >= d.n_cols) continue;
        const int64_t b = as_global(d.col_ptr)[pre], e = as_global(d.col_ptr)[pre + 1];
        const auto ri = as_global(d.row_idx);
        const T* wc = static_cast<const T*>(d.w_col);
        const auto wg = as_global(wc);
        const T wk = (T)d.w_const;
        if (d.head) {
            // ordered delivery: append (key gid_off + j, weight) to the post
            // neuron's bucket; each synapse fires at most once per step, so a
            // bucket never holds more entries than the neuron's summed in-degree
            for (int64_t base = b + (int64_t)c * 256; base < e; base += (int64_t)d.chunks * 256) {
                int32_t post[4];
                int64_t at[4], slot[4];
                T w[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int64_t j = base + lane + u * kWave;
                    post[u] = j < e ? ri[j] : -1;
                    w[u] = wc == nullptr ? wk : (j < e ? wg[j] : T(0));
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (post[u] >= 0) {
                        // one atomic returns both the bucket base and the slot
                        const uint64_t old = atomicAdd((unsigned long long*)as_global(d.head + post[u]), 1ull);
                        slot[u] = (int64_t)(old & AHIP_HEAD_COUNT);
                        at[u] = (int64_t)(old >> AHIP_HEAD_SHIFT) + slot[u];
                    }
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (post[u] >= 0 && at[u] < d.cap) {
                        store_rec(static_cast<BucketRec<T>*>(d.bucket), at[u],
                                  d.gid_off + (uint32_t)(base + lane + u * kWave), w[u]);
                        // the first event of a post neuron in this step lists it in its
                        // shard (targets delivered by their own launch)
                        if (slot[u] == 0 && d.touched) {
                            const int32_t sh = post[u] & (d.shards - 1);
                            const int32_t q = atomicAdd(d.scnt + sh * kShardStride, 1);
                            d.touched[(int64_t)sh * d.shard_cap + q] = post[u];
                        }
                    }
            }
            continue;
        }
        T* g = static_cast<T*>(d.g);
        for (int64_t base = b + (int64_t)c * 256; base < e; base += (int64_t)d.chunks * 256) {
            int32_t post[4];
            T v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int64_t j = base + lane + u * kWave;
                post[u] = j < e ? ri[j] : -1;
                v[u] = wc == nullptr ? wk : (j < e ? wg[j] : T(0));
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (post[u] >= 0) unsafeAtomicAdd((T*)as_global(g + post[u]), v[u]);
        }
    }
}

// One launch delivers every ordered conductance: one wave per shard of the
// touched post neurons of each target (the shards of a target follow one
// another); the wave delivers the shard's neurons and clears its counter.
template <typename T>
__device__ inline void deliver_one(const DeliverTarget& d, int32_t i) {
    const uint64_t h = d.head[i], hn = d.head[i + 1];
    d.head[i] = h & ~AHIP_HEAD_COUNT;
    const int64_t b = (int64_t)(h >> AHIP_HEAD_SHIFT);
    // a bucket never holds more events than the neuron's summed in-degree
    const int32_t L = (int32_t)min((int64_t)(h & AHIP_HEAD_COUNT), (int64_t)(hn >> AHIP_HEAD_SHIFT) - b);
    if (L <= 0) return;
    T* g = static_cast<T*>(d.g);
    g[i] = ordered_sum<T>(g[i], static_cast<const BucketRec<T>*>(d.bucket), b, L);
}

template <typename T>
__global__ __launch_bounds__(256) void deliver_ordered(const DeliverTarget* __restrict__ tg, int32_t n_tg,
                                                       int32_t n_waves) {
    const int lane = threadIdx.x & (kWave - 1);
    const int32_t w = blockIdx.x * 4 + threadIdx.x / kWave;
    if (w >= n_waves) return;
    int32_t k = 0;
    while (k < n_tg - 1 && w >= tg[k + 1].wave0) k++;
    const DeliverTarget& d = tg[k];
    const int32_t sh = w - d.wave0;
    int32_t* sc = d.scnt + sh * kShardStride;
    const int32_t n = *sc;
    const int32_t* list = d.touched + (int64_t)sh * d.shard_cap;
    for (int32_t q = lane; q < n; q += kWave) deliver_one<T>(d, list[q]);
    // every lane of the wave has read the count above
    if (lane == 0 && n) *sc = 0;
}

// Spike recorder of a further spike monitor of a population (the first one is
// filled by the compaction): appends the step's list to the monitor buffer.
// One workgroup; fill is double-buffered by recorded step as in compact_spikes.
__global__ __launch_bounds__(256) void record_spikes(const int32_t* __restrict__ spiked,
                                                     const int32_t* __restrict__ count,
                                                     const int64_t* __restrict__ t_base,
                                                     int32_t step, int64_t t_start,
                                                     int64_t capacity, int32_t* __restrict__ counts,
                                                     int32_t* __restrict__ idx,
                                                     int64_t idx_capacity,
                                                     int64_t* __restrict__ fill) {
    const int64_t slot = *t_base + step - t_start;
    if (slot < 0 || slot >= capacity) return;
    const int32_t n = *count;
    const int64_t f = fill[slot & 1];   // same double-buffered fill as compact_spikes
    for (int32_t k = threadIdx.x; k < n; k += 256)
        if (f + k < idx_capacity) idx[f + k] = spiked[k];
    if (threadIdx.x == 0) {
        counts[slot] = n;
        fill[(slot + 1) & 1] = f + n;
    }
}

// Mean firing rate over a sliding window of W steps (compute_firing_rate,
// Population/OpenMPGenerator.py:562-618: each spike's step is queued, steps
// <= t - W are dropped, r = 1000/window * queue size).  Here each neuron keeps
// a W-bit circular history, word-major (word k of all neurons contiguous), and
// a count: at step t the bit of step t - W at position t mod W leaves the
// window and the spike bit of step t takes its place.
template <typename T>
__global__ __launch_bounds__(256) void update_rate(int32_t size, int32_t W, const int64_t* __restrict__ t_base,
                                                   int32_t step, const uint64_t* __restrict__ mask,
                                                   uint32_t* __restrict__ bits, int32_t* __restrict__ count,
                                                   T* __restrict__ r, T rate, int32_t active) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= size) return;
    const int64_t t = *t_base + step;
    const int64_t pos = t % W;
    uint32_t* wp = bits + (pos >> 5) * (int64_t)size + i;
    const uint32_t b = 1u << (pos & 31);
    const uint32_t word = *wp;
    // an inactive population only lets step t - W leave the window (the
    // reference drops spike times <= t - W at its next active step); r stays
    const int32_t spiked = active ? (int32_t)((mask[i >> 6] >> (i & 63)) & 1ull) : 0;
    const int32_t old = (word & b) ? 1 : 0;
    *wp = spiked ? (word | b) : (word & ~b);
    const int32_t c = count[i] + spiked - old;
    count[i] = c;
    if (active) r[i] = rate * (T)c;
}

}  // namespace

int launch_update_rate(const ahip_net* net, const Pop& p, int32_t step, hipStream_t s, bool active) {
    if (p.fr_window <= 0) return 0;
    const unsigned nb = (unsigned)((p.size + 255) / 256);
    const int rs = p.slot_index("r");
    if (rs < 0) return (int)hipErrorInvalidValue;
    if (net->precision == AHIP_F64)
        hipLaunchKernelGGL(update_rate<double>, dim3(nb), dim3(256), 0, s, p.size, p.fr_window,
                           net->t_dev.as<int64_t>(), step, p.spike_mask.as<uint64_t>(), p.fr_bits.as<uint32_t>(),
                           p.fr_count.as<int32_t>(), p.slots[rs].as<double>(), p.fr_rate, active ? 1 : 0);
    else
        hipLaunchKernelGGL(update_rate<float>, dim3(nb), dim3(256), 0, s, p.size, p.fr_window,
                           net->t_dev.as<int64_t>(), step, p.spike_mask.as<uint64_t>(), p.fr_bits.as<uint32_t>(),
                           p.fr_count.as<int32_t>(), p.slots[rs].as<float>(), (float)p.fr_rate, active ? 1 : 0);
    return (int)hipGetLastError();
}

int launch_compact(const ahip_net* net, const std::vector<int32_t>& pops, const int64_t* t_base, int32_t step,
                   hipStream_t s) {
    for (size_t first = 0; first < pops.size(); first += kMaxCompact) {
        CompactBatch B;
        std::memset(&B, 0, sizeof(B));
        for (size_t q = first; q < pops.size() && B.n < kMaxCompact; q++) {
            const Pop& p = *net->pops[pops[q]];
            const Monitor* m = first_spike_monitor(net, p.id);
            CompactArgs& c = B.c[B.n];
            c.size = p.size;
            c.nb = (p.size + kCompactChunk - 1) / kCompactChunk;
            c.mask = p.spike_mask.as<uint64_t>();
            c.bcnt = p.block_counts.as<int32_t>();
            c.spiked = p.spiked.as<int32_t>();
            c.count = p.spike_count.as<int32_t>();
            if (m) {
                c.rec.counts = m->counts.as<int32_t>();
                c.rec.idx = m->data.as<int32_t>();
                c.rec.fill = m->fill.as<int64_t>();
                c.rec.idx_capacity = m->idx_capacity;
                c.rec.t_start = m->t_start;
                c.rec.capacity = m->capacity;
            }
            B.b0[B.n + 1] = B.b0[B.n] + c.nb;
            B.n++;
        }
        if (B.b0[B.n] == 0) continue;
        hipLaunchKernelGGL(compact_spikes, dim3((unsigned)B.b0[B.n]), dim3(kCompactThreads), 0, s, B, t_base, step);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
    }
    return 0;
}

int launch_scatter_group(const ahip_net* net, const ScatterGroup& g, int32_t step, hipStream_t s) {
    if (g.n == 0) return 0;
    if (g.n > kMaxScatter) return (int)hipErrorInvalidValue;
    ScatterArgs A;
    std::memset(&A, 0, sizeof(A));
    for (int32_t k = 0; k < g.n; k++) A.d[k] = g.descs[k];
    A.nd = g.n;
    A.waves_per = g.waves_per;
    if (net->precision == AHIP_F64)
        hipLaunchKernelGGL(scatter_multi<double>, dim3(g.blocks), dim3(256), 0, s, A, net->t_dev.as<int64_t>(),
                           step);
    else
        hipLaunchKernelGGL(scatter_multi<float>, dim3(g.blocks), dim3(256), 0, s, A, net->t_dev.as<int64_t>(),
                           step);
    return (int)hipGetLastError();
}

int launch_deliver(const ahip_net* net, hipStream_t s) {
    if (net->deliver_waves == 0) return 0;
    const int32_t n_tg = net->deliver_n_tg;
    const unsigned nb = (unsigned)((net->deliver_waves + 3) / 4);
    if (net->precision == AHIP_F64)
        hipLaunchKernelGGL(deliver_ordered<double>, dim3(nb), dim3(256), 0, s, net->deliver_targets.as<DeliverTarget>(),
                           n_tg, net->deliver_waves);
    else
        hipLaunchKernelGGL(deliver_ordered<float>, dim3(nb), dim3(256), 0, s, net->deliver_targets.as<DeliverTarget>(),
                           n_tg, net->deliver_waves);
    return (int)hipGetLastError();
}

int launch_spike_psp(const ahip_net* net, const Proj& p, const Pop& pre, void* target,
                     hipStream_t s) {
    if (p.nnz == 0) return 0;
    if (net->precision == AHIP_F64) return launch_scatter_t<double>(p, pre, target, s);
    return launch_scatter_t<float>(p, pre, target, s);
}

int launch_record_spikes(const Monitor& m, const Pop& p, const int64_t* t_base, int32_t step,
                         hipStream_t s) {
    hipLaunchKernelGGL(record_spikes, dim3(1), dim3(256), 0, s, p.spiked.as<int32_t>(),
                       p.spike_count.as<int32_t>(), t_base, step, m.t_start, m.capacity,
                       m.counts.as<int32_t>(), m.data.as<int32_t>(), m.idx_capacity,
                       m.fill.as<int64_t>());
    return (int)hipGetLastError();
}

}  // namespace ahip
