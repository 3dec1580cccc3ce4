// internal.hpp -- runtime structures and kernel launchers of libannarchy_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "annarchy_hip.h"
#include "annarchy_hip_model.h"
#include "annarchy_hip_synapse.h"

namespace ahip {

// Device allocation owned by the network (RAII, moved, never copied).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release();
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Device layout of a rate-coded projection in one storage format.
struct RateLayout {
    int32_t format = AHIP_FMT_CSR;
    int32_t lanes = 64;       // CSR lanes per row / SELL threads per row
    int32_t chunk = 64;       // SELL rows per slice / BSR tile
    int32_t sigma = 1;
    int32_t panel_cols = 0;
    int64_t stored = 0;       // stored entries including padding
    int32_t width = 0;        // ELL max row length
    int32_t n_units = 0;      // SELL slices / BSR block rows / panels
    DevBuf ptr;               // CSR row_ptr (int64) / SELL slice offsets (int64) / BSR block row ptr (int32)
    DevBuf idx;               // column indices (int32), padded formats included
    DevBuf val;               // weights (double or float)
    DevBuf rl;                // ELLR row lengths (int32)
    DevBuf perm;              // SELL: slice row -> post rank (int32)
    DevBuf aux;               // CSR_PANEL: per-row offsets in a tile (int32) / SELL_PANEL: slice lane map (int32)
    DevBuf tiles;             // SELL_PANEL: first slice of each tile (int32)
};

struct Pop {
    int32_t id = 0;
    int32_t size = 0;
    std::string lib_path;
    void* dl = nullptr;
    const ahip_model_desc* model = nullptr;
    std::vector<DevBuf> slots;
    std::vector<double> slot_host_global;   // host mirror of population-wide slots
    DevBuf refractory, refractory_remaining, last_spike, spike_mask, block_counts;
    DevBuf spiked, spike_count;             // int32[size], int32[1]
    DevBuf deliver;                         // ahip_deliver_desc[n_deliver]: ordered targets of the update
    int32_t n_deliver = 0;
    // delay queue (ring of `depth` steps, slot t mod depth = end of step t):
    // rate-coded: the psp variable, depth x size; spiking: spike lists,
    // depth x size int32 with counts int32[depth]
    int32_t depth = 1;
    DevBuf hist, hist_count;
    // mean firing rate over fr_window steps (0: off): a fr_window-bit history
    // per neuron (uint32 words, word-major) and the number of spikes in it
    // stop_condition: 0 none, 1 the model's condition for any neuron, 2 for
    // all neurons, 3 "spiked : any", 4 "spiked : all"
    int32_t stop_mode = 0;
    int (*stop_f64)(const ahip_pop_args*, int32_t*, void*) = nullptr;
    int (*stop_f32)(const ahip_pop_args*, int32_t*, void*) = nullptr;
    DevBuf stop_count;
    bool active = true;           // activate(false): no update, no spike gather
    DevBuf ring_tmp;              // inactive with a delay queue: staging of the ring rotation
    int32_t fr_window = 0;
    double fr_rate = 0.0;         // 1000 / window (ms)
    DevBuf fr_bits, fr_count;
    DevBuf rand_inject[AHIP_MAX_RAND];
    int64_t rand_inject_steps[AHIP_MAX_RAND] = {0};
    int64_t rand_t0 = 0;
    int32_t nblocks() const { return (size + AHIP_UPDATE_BLOCK - 1) / AHIP_UPDATE_BLOCK; }
    int32_t nwords() const { return (size + 63) / 64; }
    int slot_index(const std::string& name) const;
};

struct Proj {
    int32_t id = 0;
    int32_t pre = 0, post = 0;
    std::string target;
    bool spiking = false;
    int32_t n_rows = 0, n_cols = 0;
    int64_t nnz = 0;
    int32_t target_slot = -1;     // slot of post: _sum_<target> (rate) or g_<target> (spiking)
    int32_t pre_slot = -1;        // rate: slot "r" of pre
    bool accumulate = false;      // rate: add into the target instead of overwriting it
    bool defer_reduce = false;    // rate, panel layout: the post update adds the panel partial sums
    bool single_w = false;
    double w_const = 0.0;
    bool transmission = true;     // _transmission: compute_psp (and the learning rules) on / off
    // host copy of the forward view (for re-layout and get/set)
    std::vector<int64_t> h_row_ptr;
    std::vector<int32_t> h_col;
    std::vector<double> h_w;
    RateLayout rate;
    // spiking: backward (column) view
    DevBuf col_ptr;   // int64[n_cols+1]
    DevBuf row_idx;   // int32[nnz]
    DevBuf w_col;     // T[nnz] column order (absent when single_w)
    int32_t scatter_chunks = 1;   // waves per spike in the scatter (256 synapses each)
    // synaptic delays in steps (1 = the value of the previous step)
    int32_t delay_uniform = 1;    // the delay of every synapse when !delay_per_syn
    bool delay_per_syn = false;
    int32_t max_delay = 1;
    std::vector<int32_t> h_delay; // per synapse, forward order (delay_per_syn)
    DevBuf delay_dev;             // int16: rate -> forward order, spiking -> column order
    DevBuf ring;                  // spiking per-synapse delays: max_delay x n_rows events
    DevBuf xd;                    // rate uniform delay: the delayed pre-synaptic rates
    // plasticity: a generated synapse model and its state.  State order:
    // spiking -> column order (as w_col), rate -> forward order (CSR val).
    void* syn_dl = nullptr;
    std::string syn_path;
    const ahip_synapse_desc* syn = nullptr;
    std::vector<DevBuf> syn_attr;            // per attribute: nnz / n_rows / 1 values (a per-synapse
                                             // attribute of a spiking projection lives in syn_rec)
    // spiking: one record per synapse, column order: w, the per-synapse
    // attributes, _last_event (8-aligned); rec_off = byte offset of each field
    DevBuf syn_rec;
    int32_t rec_bytes = 0, rec_off_w = 0, rec_off_last = 0;
    std::vector<int32_t> rec_off;            // per attribute, -1 if not per synapse
    std::vector<int32_t> dep_pre, dep_post;  // slots of the pre.<x> / post.<x> attributes
    int32_t syn_update = 1, syn_plasticity = 1, syn_period = 1, syn_offset = 0;
    int32_t syn_pre_chunks = 1, syn_post_chunks = 1;   // waves per spike in the event kernels
    DevBuf last_event, fwd2col, row_ptr_dev, col_dev;
    std::vector<int64_t> h_inv_idx;
};

struct Monitor {
    int32_t id = 0;
    int32_t pop = 0;
    std::string attr;
    int32_t slot = -1;
    bool spikes = false;
    bool after_psp = false;       // targets are recorded right after compute_psp
    int64_t capacity = 0;
    // spikes: the step of the first recorded list.  Values: the first step of
    // the current recording segment at which (t - phase) mod period == 0;
    // sample k of the segment (step t_start + k * period) goes to row
    // seg_base + k.  A paused monitor records nothing.
    int64_t t_start = 0;
    int32_t period = 1;
    int64_t phase = 0;
    int64_t seg_base = 0;
    bool paused = false;
    DevBuf data;                  // values [capacity][size] (T), or spike indices
    DevBuf counts;                // spikes: int32[capacity]
    DevBuf fill;                  // spikes: int64[2], indices written before recorded step s in fill[s & 1]
    int64_t idx_capacity = 0;
};

}  // namespace ahip

namespace ahip {
// One spiking projection in a fused scatter launch (k_spike.hip:scatter_multi).
struct ScatterDesc {
    const int32_t* list;      // spike list of the pre population (ascending ranks)
    const int32_t* count;     // its length (device)
    const int64_t* col_ptr;   // CSRC backward view
    const int32_t* row_idx;
    const void* w_col;        // weights in column order, or null for w_const
    void* g;                  // post-synaptic conductance g_<target>
    double w_const;
    int64_t list_stride;      // uniform delay >= 2: the pre population's spike-list ring
    int32_t depth;            // ring depth, 1 = read `list` itself
    int32_t delay;            // list of step t - delay = ring slot (t - delay) mod depth
    int32_t n_cols;
    int32_t chunks;           // waves per spike (256 synapses each)
    // ordered delivery (cnt != NULL): the event of column-order synapse j is
    // appended as key gid_off + j to the bucket of its post neuron instead of
    // being added to g; deliver_ordered adds the bucket in ascending keys
    uint64_t* head;           // post neuron i: (bucket base << AHIP_HEAD_SHIFT) | events this step
    int64_t cap;              // bucket records
    void* bucket;             // records: key and weight (BucketRec of the precision)
    // post neurons with events this step, sharded by post & (shards - 1):
    // shard s lists its neurons at touched[s * shard_cap ..] and counts them
    // in scnt[s * kShardStride] (the delivery clears it)
    int32_t* touched;
    int32_t* scnt;
    int32_t shards, shard_cap;
    uint32_t gid_off;
};
constexpr int kMaxScatter = 8;
// One projection feeding an ordered target: its keys are [gid_off, gid_end),
// key k names column-order synapse k - gid_off of that projection.
// One ordered conductance in the merged delivery launch (conductances that
// no update can deliver): its touched post neurons of the step, sharded.
struct DeliverTarget {
    void* g;
    uint64_t* head;
    const void* bucket;       // BucketRec records
    const int32_t* touched;
    int32_t* scnt;
    int32_t shards, shard_cap;
    int32_t wave0;            // first wave of this target in the launch (one wave per shard)
    int32_t reserved;
};
constexpr int kShardStride = 16;     // one 64-byte line per shard counter
constexpr int kMaxShards = 1024;
// A conductance g_<target> fed by at least one spiking projection with
// per-synapse weights.  The reference adds the events projection by
// projection, spike by spike in ascending pre-synaptic rank
// (Projection/OpenMP/CSR.py:870-885); keys gid_off + j of the projections in
// creation order sort exactly in that order, so adding each post neuron's
// bucket in ascending keys reproduces every rounding of the reference.
struct OrderedTarget {
    void* g = nullptr;
    int32_t pop = 0, size = 0;
    int64_t cap = 0;              // bucket entries (the summed in-degree)
    std::vector<int32_t> projs;   // projection ids, creation order
    std::vector<int64_t> key_off; // key offset of each projection (its keys: key_off + column-order index)
    DevBuf head;                  // uint64[size + 1]: (bucket base << AHIP_HEAD_SHIFT) | count, count 0 between steps
    DevBuf bucket;                // [cap] records (key, weight): 16 B in FP64, 8 B in FP32
    int32_t shards = 1, shard_cap = 0;
    DevBuf touched;               // int32[shards * shard_cap]
    DevBuf scnt;                  // int32[shards * kShardStride]
    // delivered by the update kernel of the post population (no other writer
    // of the conductance, no monitor reading it right after compute_psp, at
    // most AHIP_MAX_DELIVER per population); otherwise by the merged
    // delivery launch after the scatters
    bool in_update = false;
};
// Spiking projections launched together: their targets are distinct, so the
// atomics of one projection never interleave with another's into the same
// conductance and every sum keeps the reference's projection order.
struct ScatterGroup {
    std::vector<ScatterDesc> descs;   // passed by value in the launch's arguments
    int32_t n = 0;
    int32_t waves_per = 1;            // waves of each descriptor
    int64_t blocks = 1;
};
}  // namespace ahip

struct ahip_net {
    int32_t device = 0;
    double dt = 1.0;
    int32_t precision = AHIP_F64;
    bool initialized = false;
    uint64_t seed = 0;
    int64_t t = 0;                 // host mirror of the simulation step
    hipStream_t stream = nullptr;
    ahip::DevBuf t_dev;            // int64 simulation step at the start of a batch
    std::vector<std::unique_ptr<ahip::Pop>> pops;
    std::vector<std::unique_ptr<ahip::Proj>> projs;
    std::vector<std::unique_ptr<ahip::Monitor>> mons;
    bool graph_enabled = true;
    int32_t graph_batch = 0;
    hipGraphExec_t graph_exec = nullptr;
    int32_t graph_steps = 0;
    bool dirty = true;             // structure changed since the last capture
    bool scatter_dirty = true;     // the fused scatter tables need rebuilding
    uint64_t scatter_gen = 0;       // bumped whenever the scatter tables are rebuilt
    uint64_t graph_scatter_gen = ~0ull;  // the tables generation the step graph was captured with
    // kernel-time probe: HIP events around one plastic kernel inside eager steps
    int32_t probe_proj = -1, probe_kind = -1;
    hipEvent_t probe_ev[2] = {nullptr, nullptr};
    std::vector<ahip::ScatterGroup> scatter_groups;
    std::vector<ahip::OrderedTarget> ordered;   // targets delivered in the reference's order
    ahip::DevBuf deliver_targets;               // DeliverTarget of the merged delivery launch
    int32_t deliver_n_tg = 0, deliver_waves = 0;
    int32_t delivery = AHIP_DELIVERY_ORDERED;
    bool persist_enabled = true;   // small single-population networks: one launch per batch
    size_t elem() const { return precision == AHIP_F64 ? sizeof(double) : sizeof(float); }
};

namespace ahip {

// ---- kernel launchers (return hipError_t as int) ---------------------------
// reduce = false: a sliced-tile layout leaves its panel partial sums to the
// post population's update (ahip_pop_args.psum_*)
int launch_rate_psp(const ahip_net* net, const Proj& p, const void* pre_r, void* target,
                    hipStream_t s, bool reduce = true);
int launch_spike_psp(const ahip_net* net, const Proj& p, const Pop& pre, void* target,
                     hipStream_t s);
// the spike compactions of the populations `pops` (one launch per 8)
int launch_compact(const ahip_net* net, const std::vector<int32_t>& pops, const int64_t* t_base, int32_t step,
                   hipStream_t s);
// The first spike monitor of a population (filled by its compaction kernel).
inline const Monitor* first_spike_monitor(const ahip_net* net, int32_t pop) {
    for (auto& mp : net->mons)
        if (mp->spikes && mp->pop == pop && !mp->paused) return mp.get();
    return nullptr;
}
int launch_update_rate(const ahip_net* net, const Pop& p, int32_t step, hipStream_t s, bool active = true);
int launch_scatter_group(const ahip_net* net, const ScatterGroup& g, int32_t step, hipStream_t s);
int launch_deliver(const ahip_net* net, hipStream_t s);
int launch_push_values(const ahip_net* net, const void* src, void* hist, int32_t n, int32_t D, int32_t step,
                       hipStream_t s);
int launch_gather_delayed(const ahip_net* net, const void* hist, void* dst, int32_t n, int32_t D, int32_t d,
                          int32_t step, hipStream_t s);
int launch_fill_hist(const ahip_net* net, const void* src, void* hist, int32_t n, int32_t D, hipStream_t s);
int launch_push_spikes(const ahip_net* net, const Pop& p, int32_t step, hipStream_t s);
int launch_rate_psp_delayed(const ahip_net* net, const Proj& p, const Pop& pre, void* target, int32_t step,
                            hipStream_t s);
int launch_spike_ring(const ahip_net* net, const Proj& p, const Pop& pre, void* target, int32_t step,
                      hipStream_t s);
int launch_gen_fixed_number_pre(int32_t post_begin, int32_t post_end, int32_t pre_begin, int32_t n_pre,
                                int32_t number, int32_t no_self, uint64_t seed, const int64_t* row_ptr,
                                int32_t* col, double* w, int32_t w_kind, double wa, double wb, hipStream_t s);
int launch_gen_fixed_probability(int32_t post_begin, int32_t post_end, int32_t pre_begin, int32_t n_pre,
                                 double p, int32_t no_self, uint64_t seed, int32_t pass, int64_t* counts,
                                 const int64_t* row_ptr, int32_t* col, double* w, int32_t w_kind, double wa,
                                 double wb, hipStream_t s);
int launch_advance_time(int64_t* t, int64_t n, hipStream_t s);
int launch_fill(void* dst, int64_t n, int32_t elem, double value, hipStream_t s);
int launch_fill_i32(int32_t* dst, int64_t n, int32_t value, hipStream_t s);
int launch_fill_i64(int64_t* dst, int64_t n, int64_t value, hipStream_t s);
int launch_record_values(const Monitor& m, const void* src, int32_t size, int32_t elem,
                         const int64_t* t_base, int32_t step, hipStream_t s);
int launch_record_spikes(const Monitor& m, const Pop& p, const int64_t* t_base, int32_t step,
                         hipStream_t s);
int launch_convert(const void* src, int32_t src_f64, void* dst, int32_t dst_f64, int64_t n,
                   hipStream_t s);
int launch_copy16(const void* src, void* dst, int64_t n16, hipStream_t s);

// Device state of the batched multi-instance rate sweep (k_batched.hip).
struct BatchDev {
    int32_t n_inst = 0, n_in = 0, total = 0;
    int32_t m_tiles = 0, n_tiles = 0, k_steps = 0;  // GEMM tiling (32 x 32 x 16): input neurons,
                                                    // instances per step, depth
    int32_t cap_steps = 0;           // steps per input chunk the X and Y buffers hold
    int32_t lanes = 16;
    double dt = 1.0, tau = 10.0, omega = 0.0;
    DevBuf row_ptr, col, w;          // block-diagonal CSR, FP32 weights
    DevBuf inst_of_row, local_of_row;
    DevBuf gain;                     // float [n_inst]
    DevBuf r[2];                     // float [total] ping-pong
    DevBuf sum_exc;                  // float [total], psp of the last step
    DevBuf w_hi, w_lo;               // W^T as bf16 B fragments [m_tiles][k_steps][64][8]
    DevBuf x_hi, x_lo;               // X^T of a chunk as bf16 A fragments [cap_steps*n_tiles][k_steps][64][8]
    DevBuf phase;                    // double [n_in][n_inst]
    DevBuf y;                        // Y^T of a chunk: float [cap_steps][n_tiles*32][m_tiles*32]
    DevBuf t_dev;                    // int64
    int32_t mode = 1;                // 0: gathers from global memory, 1: instance rates staged in LDS,
                                     // 2: sliced chunks (one lane per row, uint16 columns)
    int32_t n_chunks = 0, max_n = 0;
    DevBuf chunks;                   // int32 [n_chunks][3] = (instance, first row, end row)
    DevBuf inst_base;                // int64 [n_inst + 1]
    // mode 2: chunks of <= 512 rows, their 64-row slices and the sliced arrays
    int32_t n_schunks = 0;
    int64_t sl_stored = 0;
    DevBuf schunks, schunk_slices;   // int32 [n][3], int32 [n + 1]
    DevBuf sl_off, sl_row;           // int64 [slices + 1], int32 [slices][64] (-1: no row)
    DevBuf sl_col, sl_w;             // uint16 / float [sl_stored]
};
int batch_launch_input(const BatchDev& B, int32_t step0, int32_t n, hipStream_t s);
int batch_launch_gemm(const BatchDev& B, int32_t n, hipStream_t s);
int batch_launch_spmv(const BatchDev& B, int32_t parity, int32_t j, hipStream_t s);
int batch_launch_advance(const BatchDev& B, int64_t n, hipStream_t s);

// Sets the message returned by ahip_last_error(); returns `code`.
int report_error(int code, const char* msg);

// Format builders (host, from the forward CSR view).
int build_layout(const Proj& p, int32_t precision, const ahip_format_opts& o, RateLayout& out,
                 std::string& err);
double psp_bytes(const Proj& p, int32_t precision);

}  // namespace ahip
