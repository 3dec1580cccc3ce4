// runtime.cpp -- C-ABI of libannarchy_hip: network, populations, projections,
// monitors and the captured simulation step.
//
// The step order follows the reference's singleStep()
// (generator/Template/OpenMPBaseTemplate.py:217-325): compute_psp of every
// projection (rate SpMV or spike scatter), record targets, population updates
// (ODE + threshold + reset + refractory fused, random draws inside), spike
// compaction, record variables, t++.  Unlike the reference CUDA host loop
// (CUDABaseTemplate.py:474-589), nothing is read back per step: spike counts
// stay on the device, grids are sized for the worst case, and batches of
// steps are captured once in a HIP graph and replayed.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace ahip;

static int syn_field_read(ahip_net* net, Proj& p, int field, double* out);

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(AHIP_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),   \
                        __FILE__, __LINE__);                                                \
    } while (0)

#define LCHK(expr)                                                                          \
    do {                                                                                    \
        int _e = (expr);                                                                    \
        if (_e != 0)                                                                        \
            return fail(AHIP_EHIP, "kernel launch %s failed: %s", #expr,                    \
                        hipGetErrorString((hipError_t)_e));                                 \
    } while (0)

int dev_alloc(DevBuf& b, size_t bytes) {
    b.release();
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(AHIP_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    b.bytes = bytes;
    // The network's work runs on a non-blocking stream, which is not ordered
    // against the null stream a plain memset uses: wait for the clear here so
    // a later kernel on the network stream cannot be overtaken by it.
    HIPCHK(hipMemset(b.p, 0, bytes));
    HIPCHK(hipDeviceSynchronize());
    return 0;
}

Pop* get_pop(ahip_net* net, int32_t id) {
    if (!net || id < 0 || id >= (int32_t)net->pops.size()) return nullptr;
    return net->pops[id].get();
}
Proj* get_proj(ahip_net* net, int32_t id) {
    if (!net || id < 0 || id >= (int32_t)net->projs.size()) return nullptr;
    return net->projs[id].get();
}
Monitor* get_mon(ahip_net* net, int32_t id) {
    if (!net || id < 0 || id >= (int32_t)net->mons.size()) return nullptr;
    return net->mons[id].get();
}

// Drops the instantiated step graph: its kernel nodes hold the device
// addresses of the buffers and tables they were captured with, so any
// reallocation of one of them must come with a recapture.
void drop_graph(ahip_net* net) {
    net->dirty = true;
    if (net->graph_exec) {
        hipGraphExecDestroy(net->graph_exec);
        net->graph_exec = nullptr;
        net->graph_steps = 0;
    }
}

// Topology or buffers changed: rebuild the scatter tables and the graph.
void invalidate(ahip_net* net) {
    drop_graph(net);
    net->scatter_dirty = true;
}

bool layout_covers_all_rows(const Proj& p) {
    if (p.rate.format == AHIP_FMT_ELL || p.rate.format == AHIP_FMT_ELLR)
        return p.rate.n_units == p.n_rows;
    return true;
}

struct StepPlan {
    std::vector<int32_t> zero_before;   // per projection: 1 = fill target with 0 first
};

// Decide, per rate projection, whether it overwrites its target (first one
// into that target and its layout writes every row), or adds into it.
// Re-slot a ring indexed by step (slot k mod D) into a deeper one: the
// entries of steps [k_from, k_to) move to their new slots, the others of the
// new ring are zero.  Rings only grow, as the reference's queues
// (update_max_delay resizes them keeping their content).
static int regrow_ring(DevBuf& buf, size_t slot_bytes, int32_t d_old, int32_t d_new, int64_t k_from,
                       int64_t k_to) {
    DevBuf nb;
    int rc = dev_alloc(nb, (size_t)d_new * slot_bytes);
    if (rc) return rc;
    if (buf.p && d_old > 0) {
        for (int64_t k = k_from; k < k_to; k++) {
            const int64_t so = ((k % d_old) + d_old) % d_old, sn = ((k % d_new) + d_new) % d_new;
            HIPCHK(hipMemcpy((char*)nb.p + sn * slot_bytes, (char*)buf.p + so * slot_bytes, slot_bytes,
                             hipMemcpyDeviceToDevice));
        }
    }
    buf = std::move(nb);
    return 0;
}

// Delay queues: the depth of a population's queue is the largest delay that
// reads it (rate-coded: every delayed projection; spiking: the uniform
// delays, per-synapse delays keep their own event ring), and it never
// shrinks.  A new queue is zero (no spikes); a deeper one keeps the values of
// the last steps.
int ensure_delay_buffers(ahip_net* net) {
    const size_t es = net->elem();
    std::vector<int32_t> need(net->pops.size(), 1);
    for (auto& pp : net->projs) {
        Proj& p = *pp;
        if (!p.spiking)
            need[p.pre] = std::max(need[p.pre], p.max_delay);
        else if (!p.delay_per_syn)
            need[p.pre] = std::max(need[p.pre], p.max_delay);
    }
    for (auto& pp : net->pops) {
        Pop& p = *pp;
        const int32_t D = std::max(need[p.id], p.depth);
        if (D == p.depth && (D == 1 || p.hist.p)) continue;
        HIPCHK(hipStreamSynchronize(net->stream));
        const int32_t d_old = p.hist.p ? p.depth : 0;
        int rc;
        if (p.model->spiking) {
            if ((rc = regrow_ring(p.hist, (size_t)p.size * sizeof(int32_t), d_old, D, net->t - d_old, net->t))) return rc;
            if ((rc = regrow_ring(p.hist_count, sizeof(int32_t), d_old, D, net->t - d_old, net->t))) return rc;
        } else {
            if ((rc = regrow_ring(p.hist, (size_t)p.size * es, d_old, D, net->t - d_old, net->t))) return rc;
        }
        p.depth = D;
        invalidate(net);
    }
    for (auto& pp : net->projs) {
        Proj& p = *pp;
        if (!p.spiking && p.delay_uniform >= 2 && !p.delay_per_syn && !p.xd.p) {
            int rc = dev_alloc(p.xd, (size_t)p.n_cols * es);
            if (rc) return rc;
            invalidate(net);
        }
        if (p.spiking && p.delay_per_syn) {
            // pending events of steps [t, t + old depth) keep their delivery step
            const size_t slot = (size_t)p.n_rows * (p.single_w ? sizeof(int32_t) : es);
            const int32_t d_old = p.ring.p ? (int32_t)(p.ring.bytes / std::max<size_t>(slot, 1)) : 0;
            if (!p.ring.p || d_old < p.max_delay) {
                HIPCHK(hipStreamSynchronize(net->stream));
                int rc = regrow_ring(p.ring, slot, d_old, std::max(p.max_delay, d_old), net->t, net->t + d_old);
                if (rc) return rc;
                invalidate(net);
            }
        }
    }
    return 0;
}

// Synaptic state of a plastic projection (spiking: column order as w_col;
// rate-coded: forward order as the CSR weights): every attribute at its
// initial value, one per synapse / post neuron / projection by locality;
// _last_event = -10000 (the reference's initial value,
// Projection/CUDA/CSR.py:280-291); the forward row pointer and columns, and
// for spiking projections the forward -> column map.
int init_syn_state(ahip_net* net, Proj& p) {
    const ahip_synapse_desc* d = p.syn;
    const size_t es = net->elem();
    HIPCHK(hipStreamSynchronize(net->stream));
    p.syn_attr.clear();
    p.syn_attr.resize(d->n_attr);
    p.rec_off.assign(d->n_attr, -1);
    int rc;
    for (int32_t k = 0; k < d->n_attr; k++) {
        const int64_t n = d->attr_locality[k] == AHIP_SYN_LOCAL ? p.nnz
                          : d->attr_locality[k] == AHIP_SYN_SEMIGLOBAL ? p.n_rows : 1;
        if (p.spiking && d->attr_locality[k] == AHIP_SYN_LOCAL) continue;   // in the record
        if ((rc = dev_alloc(p.syn_attr[k], (size_t)n * es))) return rc;
        LCHK(launch_fill(p.syn_attr[k].p, n, (int32_t)es, d->attr_init[k], net->stream));
    }
    p.syn_rec.release();
    p.last_event.release();
    if (p.spiking) {
        // record of a synapse: w, the per-synapse attributes, _last_event
        int32_t off = (int32_t)es;
        p.rec_off_w = 0;
        for (int32_t k = 0; k < d->n_attr; k++)
            if (d->attr_locality[k] == AHIP_SYN_LOCAL) {
                p.rec_off[k] = off;
                off += (int32_t)es;
            }
        off = (off + 7) & ~7;
        p.rec_off_last = off;
        p.rec_bytes = off + 8;
        if ((rc = dev_alloc(p.syn_rec, (size_t)p.nnz * p.rec_bytes))) return rc;
        const int64_t chunk = 1 << 20;
        std::vector<char> buf((size_t)std::min<int64_t>(chunk, std::max<int64_t>(p.nnz, 1)) * p.rec_bytes);
        for (int64_t q0 = 0; q0 < p.nnz; q0 += chunk) {
            const int64_t m = std::min(chunk, p.nnz - q0);
            for (int64_t i = 0; i < m; i++) {
                char* r = buf.data() + (size_t)i * p.rec_bytes;
                const double wv = p.h_w[p.h_inv_idx[q0 + i]];
                auto put = [&](int32_t o, double v) {
                    if (es == 8) std::memcpy(r + o, &v, 8);
                    else { const float f = (float)v; std::memcpy(r + o, &f, 4); }
                };
                std::memset(r, 0, p.rec_bytes);
                put(p.rec_off_w, wv);
                for (int32_t k = 0; k < d->n_attr; k++)
                    if (p.rec_off[k] >= 0) put(p.rec_off[k], d->attr_init[k]);
                const int64_t le = -10000;
                std::memcpy(r + p.rec_off_last, &le, 8);
            }
            HIPCHK(hipMemcpy((char*)p.syn_rec.p + (size_t)q0 * p.rec_bytes, buf.data(), (size_t)m * p.rec_bytes,
                             hipMemcpyHostToDevice));
        }
        p.w_col.release();   // the weights live in the records
    } else {
        if ((rc = dev_alloc(p.last_event, (size_t)p.nnz * sizeof(int64_t)))) return rc;
        LCHK(launch_fill_i64(p.last_event.as<int64_t>(), p.nnz, -10000, net->stream));
    }
    p.fwd2col.release();
    if (p.spiking) {
        std::vector<int64_t> f2c(p.nnz);
        for (int64_t q = 0; q < p.nnz; q++) f2c[p.h_inv_idx[q]] = q;
        if ((rc = dev_alloc(p.fwd2col, (size_t)p.nnz * sizeof(int64_t)))) return rc;
        if (p.nnz) HIPCHK(hipMemcpy(p.fwd2col.p, f2c.data(), f2c.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    }
    // waves per spike: 64-synapse chunks of the longest column / row, at most 32
    int64_t max_row = 0;
    std::vector<int64_t> col_len(p.n_cols, 0);
    for (int32_t r = 0; r < p.n_rows; r++) max_row = std::max(max_row, p.h_row_ptr[r + 1] - p.h_row_ptr[r]);
    for (int64_t j = 0; j < p.nnz; j++) col_len[p.h_col[j]]++;
    const int64_t max_col = col_len.empty() ? 0 : *std::max_element(col_len.begin(), col_len.end());
    p.syn_pre_chunks = (int32_t)std::min<int64_t>(32, std::max<int64_t>(1, (max_col + 63) / 64));
    p.syn_post_chunks = (int32_t)std::min<int64_t>(32, std::max<int64_t>(1, (max_row + 63) / 64));
    if ((rc = dev_alloc(p.row_ptr_dev, p.h_row_ptr.size() * sizeof(int64_t)))) return rc;
    HIPCHK(hipMemcpy(p.row_ptr_dev.p, p.h_row_ptr.data(), p.h_row_ptr.size() * sizeof(int64_t),
                     hipMemcpyHostToDevice));
    if ((rc = dev_alloc(p.col_dev, (size_t)p.nnz * sizeof(int32_t)))) return rc;
    if (p.nnz) HIPCHK(hipMemcpy(p.col_dev.p, p.h_col.data(), p.nnz * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(net->stream));
    return 0;
}


int launch_syn_now(ahip_net* net, Proj& p, int kind, int32_t step);

// Pre events (kind 0), update_synapse (1) or post events (2) of a plastic
// projection at `step`, between the probe events when this kernel is probed.
int launch_syn(ahip_net* net, Proj& p, int kind, int32_t step) {
    const bool probe = net->probe_proj == p.id && net->probe_kind == kind;
    if (probe) HIPCHK(hipEventRecord(net->probe_ev[0], net->stream));
    int rc = launch_syn_now(net, p, kind, step);
    if (rc) return rc;
    if (probe) HIPCHK(hipEventRecord(net->probe_ev[1], net->stream));
    return 0;
}

int launch_syn_now(ahip_net* net, Proj& p, int kind, int32_t step) {
    Pop& pre = *net->pops[p.pre];
    Pop& post = *net->pops[p.post];
    ahip_syn_args a;
    std::memset(&a, 0, sizeof(a));
    if (p.spiking) {
        if (p.delay_uniform >= 2) {
            a.pre_spiked = pre.hist.as<int32_t>();
            a.pre_count = pre.hist_count.as<int32_t>();
            a.pre_stride = pre.size;
            a.pre_depth = pre.depth;
            a.pre_delay = p.delay_uniform;
        } else {
            a.pre_spiked = pre.spiked.as<int32_t>();
            a.pre_count = pre.spike_count.as<int32_t>();
            a.pre_depth = 1;
        }
        a.col_ptr = p.col_ptr.as<int64_t>();
        a.row_idx = p.row_idx.as<int32_t>();
        a.post_spiked = post.spiked.as<int32_t>();
        a.post_count = post.spike_count.as<int32_t>();
        a.fwd2col = p.fwd2col.as<int64_t>();
        a.g = post.slots[p.target_slot].p;
    }
    a.n_cols = p.n_cols;
    a.n_rows = p.n_rows;
    a.row_ptr = p.row_ptr_dev.as<int64_t>();
    a.col_idx = p.col_dev.as<int32_t>();
    a.pre_last_spike = pre.last_spike.as<int64_t>();
    a.post_last_spike = post.last_spike.as<int64_t>();
    if (p.spiking) {
        char* base = (char*)p.syn_rec.p;
        a.w = base + p.rec_off_w;
        a.last_event = (int64_t*)(base + p.rec_off_last);
        a.state_stride = p.rec_bytes / (int32_t)net->elem();
        a.state_stride8 = p.rec_bytes / 8;
        for (int32_t k = 0; k < p.syn->n_attr; k++)
            a.attr[k] = p.rec_off[k] >= 0 ? (void*)(base + p.rec_off[k]) : p.syn_attr[k].p;
    } else {
        a.w = p.rate.val.p;
        a.last_event = p.last_event.as<int64_t>();
        a.state_stride = 1;
        a.state_stride8 = 1;
        for (int32_t k = 0; k < p.syn->n_attr; k++) a.attr[k] = p.syn_attr[k].p;
    }
    for (size_t k = 0; k < p.dep_pre.size(); k++) {
        a.pre_attr[k] = pre.slots[p.dep_pre[k]].p;
        a.pre_attr_stride[k] = pre.model->slot_local[p.dep_pre[k]] ? 1 : 0;
    }
    for (size_t k = 0; k < p.dep_post.size(); k++) {
        a.post_attr[k] = post.slots[p.dep_post[k]].p;
        a.post_attr_stride[k] = post.model->slot_local[p.dep_post[k]] ? 1 : 0;
    }
    a.t_base = net->t_dev.as<int64_t>();
    a.step = step;
    a.dt = net->dt;
    a.pre_chunks = p.syn_pre_chunks;
    a.post_chunks = p.syn_post_chunks;
    a.plasticity = p.syn_plasticity;
    a.update_period = p.syn_period;
    a.update_offset = p.syn_offset;
    const bool f64 = net->precision == AHIP_F64;
    const ahip_synapse_desc* d = p.syn;
    ahip_syn_launch_fn fn = kind == 0 ? (f64 ? d->pre_f64 : d->pre_f32)
                            : kind == 1 ? (f64 ? d->update_f64 : d->update_f32)
                                        : (f64 ? d->post_f64 : d->post_f32);
    const int rc = fn(&a, (void*)net->stream);
    if (rc != 0)
        return fail(AHIP_EHIP, "%s of projection %d failed: %s",
                    kind == 0 ? "pre events" : kind == 1 ? "update_synapse" : "post events", p.id,
                    hipGetErrorString((hipError_t)rc));
    return 0;
}

// An inactive population does not push into its delay queue: the reference's
// deque keeps its lists while the clock runs on (Population/OpenMPGenerator.py:
// 440-455 inside `if (_active)`).  The ring here is indexed by step (slot
// u mod D = end of step u), so it is rotated by one slot per inactive step:
// slot (t - d) then still holds what deque entry d - 1 holds.
int rotate_ring(DevBuf& ring, size_t slot_bytes, int32_t D, DevBuf& tmp, hipStream_t s) {
    char* r = (char*)ring.p;
    char* t = (char*)tmp.p;
    if (!r || D <= 1) return 0;
    if (!t || tmp.bytes < slot_bytes * D) return fail(AHIP_ESTATE, "delay ring staging missing");
    HIPCHK(hipMemcpyAsync(t + slot_bytes, r, slot_bytes * (D - 1), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(t, r + slot_bytes * (D - 1), slot_bytes, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(r, t, slot_bytes * D, hipMemcpyDeviceToDevice, s));
    return 0;
}

// compute_psp of a projection runs when it transmits and its post-synaptic
// population is active (Projection/OpenMP/CSR.py: _transmission && post _active)
bool transmits(const ahip_net* net, const Proj& p) {
    return p.transmission && net->pops[p.post]->active;
}

// Ordered targets of the fused spike scatter (`targets`: the conductances
// fed by a projection with per-synapse weights).  Buffers of a previous
// build of the same conductance are kept when they are large enough.
int build_ordered_targets(ahip_net* net, const std::vector<const void*>& targets) {
    std::vector<OrderedTarget> old;
    old.swap(net->ordered);
    std::vector<const void*> seen;
    for (const void* g : targets) {
        if (std::find(seen.begin(), seen.end(), g) != seen.end()) continue;  // one target per conductance
        seen.push_back(g);
        OrderedTarget o;
        o.g = const_cast<void*>(g);
        int64_t total = 0;
        for (auto& pp : net->projs) {
            Proj& p = *pp;
            if (!(p.spiking && p.nnz != 0 && !p.delay_per_syn && !p.syn && transmits(net, p))) continue;
            if (net->pops[p.post]->slots[p.target_slot].p != g) continue;
            o.pop = p.post;
            o.projs.push_back(p.id);
            o.key_off.push_back(total);
            total += p.nnz;
        }
        if (total >= (int64_t)UINT32_MAX)
            return fail(AHIP_EINVAL, "ordered spike delivery keys 32-bit synapse indices: %lld synapses into one "
                        "conductance (use AHIP_DELIVERY_ATOMIC)", (long long)total);
        const Pop& post = *net->pops[o.pop];
        o.size = post.size;
        o.cap = total;
        std::vector<int64_t> boff(o.size + 1, 0);
        for (int32_t id : o.projs) {
            const Proj& p = *net->projs[id];
            for (int32_t i = 0; i < o.size; i++) boff[i + 1] += p.h_row_ptr[i + 1] - p.h_row_ptr[i];
        }
        int64_t max_in = 0;
        for (int32_t i = 0; i < o.size; i++) max_in = std::max(max_in, boff[i + 1]);
        if (max_in >= (int64_t)AHIP_HEAD_COUNT)
            return fail(AHIP_EINVAL, "ordered spike delivery counts events in %d bits: in-degree %lld into one "
                        "conductance (use AHIP_DELIVERY_ATOMIC)", AHIP_HEAD_SHIFT, (long long)max_in);
        for (int32_t i = 0; i < o.size; i++) boff[i + 1] += boff[i];
        // bucket records: key and weight, 16 B in FP64 and 8 B in FP32
        // (ahip::BucketRec of the precision)
        const size_t rec = net->precision == AHIP_F64 ? 16 : 8;
        // head[i] = bucket base << AHIP_HEAD_SHIFT, event count 0
        std::vector<uint64_t> head(o.size + 1);
        for (int32_t i = 0; i <= o.size; i++) head[i] = (uint64_t)boff[i] << AHIP_HEAD_SHIFT;
        for (OrderedTarget& q : old)
            if (q.g == g) {
                if (q.head.bytes >= head.size() * sizeof(uint64_t)) o.head = std::move(q.head);
                if (q.bucket.bytes >= (size_t)o.cap * rec) o.bucket = std::move(q.bucket);
                if (q.touched.bytes >= (size_t)o.size * sizeof(int32_t)) o.touched = std::move(q.touched);
            }
        // shards of at most 64 neurons each (one wave), a power of two up to kMaxShards
        o.shards = 1;
        while (o.shards < kMaxShards && (int64_t)o.shards * 64 < o.size) o.shards *= 2;
        o.shard_cap = (o.size + o.shards - 1) / o.shards;
        int rc;
        if (!o.head.p && (rc = dev_alloc(o.head, head.size() * sizeof(uint64_t)))) return rc;
        if (!o.bucket.p && (rc = dev_alloc(o.bucket, (size_t)o.cap * rec))) return rc;
        const size_t touched_n = (size_t)o.shards * o.shard_cap;
        if (o.touched.bytes < touched_n * sizeof(int32_t) && (rc = dev_alloc(o.touched, touched_n * sizeof(int32_t))))
            return rc;
        if ((rc = dev_alloc(o.scnt, (size_t)o.shards * kShardStride * sizeof(int32_t)))) return rc;
        HIPCHK(hipMemcpy(o.head.p, head.data(), head.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
        // the post population's update adds the buckets when no other kernel
        // writes the conductance and no monitor records it after compute_psp
        o.in_update = true;
        for (auto& pp : net->projs) {
            const Proj& p = *pp;
            if (net->pops[p.post]->slots[p.target_slot].p == g &&
                std::find(o.projs.begin(), o.projs.end(), p.id) == o.projs.end() && transmits(net, p))
                o.in_update = false;
        }
        for (auto& mp : net->mons) {
            const Monitor& m = *mp;
            if (m.after_psp && m.pop == o.pop && net->pops[o.pop]->slots[m.slot].p == g) o.in_update = false;
        }
        int32_t fused_here = 0;
        for (const OrderedTarget& q : net->ordered)
            if (q.in_update && q.pop == o.pop) fused_here++;
        if (fused_here >= AHIP_MAX_DELIVER) o.in_update = false;
        HIPCHK(hipDeviceSynchronize());
        net->ordered.push_back(std::move(o));
    }
    // the descriptors of every population's update
    for (auto& pp : net->pops) {
        Pop& p = *pp;
        std::vector<ahip_deliver_desc> dd;
        for (OrderedTarget& o : net->ordered) {
            if (!o.in_update || o.pop != p.id) continue;
            ahip_deliver_desc d;
            std::memset(&d, 0, sizeof(d));
            d.g = o.g;
            d.head = o.head.as<uint64_t>();
            d.bucket = o.bucket.p;
            dd.push_back(d);
        }
        p.n_deliver = (int32_t)dd.size();
        p.deliver.release();
        if (!dd.empty()) {
            int rc;
            if ((rc = dev_alloc(p.deliver, dd.size() * sizeof(ahip_deliver_desc)))) return rc;
            HIPCHK(hipMemcpy(p.deliver.p, dd.data(), dd.size() * sizeof(ahip_deliver_desc), hipMemcpyHostToDevice));
        }
    }
    // the merged delivery launch: one DeliverTarget per conductance, their
    // sources in one table (staged in LDS by the kernel)
    std::vector<DeliverTarget> tg;
    int32_t waves = 0;
    for (OrderedTarget& o : net->ordered) {
        if (o.in_update) continue;
        DeliverTarget d;
        d.g = o.g;
        d.head = o.head.as<uint64_t>();
        d.bucket = o.bucket.p;
        d.touched = o.touched.as<int32_t>();
        d.scnt = o.scnt.as<int32_t>();
        d.shards = o.shards;
        d.shard_cap = o.shard_cap;
        d.wave0 = waves;
        d.reserved = 0;
        tg.push_back(d);
        waves += o.shards;
    }
    net->deliver_targets.release();
    net->deliver_n_tg = (int32_t)tg.size();
    net->deliver_waves = waves;
    if (!tg.empty()) {
        int rc;
        if ((rc = dev_alloc(net->deliver_targets, tg.size() * sizeof(DeliverTarget)))) return rc;
        HIPCHK(hipMemcpy(net->deliver_targets.p, tg.data(), tg.size() * sizeof(DeliverTarget), hipMemcpyHostToDevice));
    }
    return 0;
}

int prepare(ahip_net* net) {
    if (!net->initialized) return fail(AHIP_ESTATE, "network not initialized");
    for (auto& pp : net->projs) {
        Proj& p = *pp;
        if (p.nnz < 0) return fail(AHIP_ESTATE, "projection %d has no connectivity", p.id);
    }
    int rc0 = ensure_delay_buffers(net);
    if (rc0) return rc0;
    for (auto& pp : net->pops) {
        Pop& p = *pp;
        if (p.active || p.depth <= 1 || !p.hist.p) continue;
        if (p.ring_tmp.bytes < p.hist.bytes) {
            drop_graph(net);  // the rotation's copy nodes address ring_tmp
            if ((rc0 = dev_alloc(p.ring_tmp, p.hist.bytes))) return rc0;
        }
    }
    // fused scatter groups: spiking projections in creation order, each in the
    // first group after the last one that already writes its target, so a
    // target receives the projections' sums in the reference's order.  A
    // target that some projection with per-synapse weights writes has no
    // such order (its atomic sum is order-dependent in any case), so its
    // projections share groups freely.
    if (net->scatter_dirty) {
        // the tables freed here are the ones a captured graph reads
        drop_graph(net);
        net->scatter_gen++;
        net->scatter_groups.clear();
        std::vector<std::vector<ScatterDesc>> groups;
        std::vector<int64_t> blocks;
        std::vector<std::vector<const void*>> targets;
        auto fused = [&](const Proj& p) {
            return p.spiking && p.nnz != 0 && !p.delay_per_syn && !p.syn && transmits(net, p);
        };
        std::vector<const void*> unordered;
        for (auto& pp : net->projs)
            if (fused(*pp) && !pp->single_w) unordered.push_back(net->pops[pp->post]->slots[pp->target_slot].p);
        // ordered delivery: every fused projection into such a target appends
        // its events to the target's buckets (single-weight ones included, so
        // the projections' sums keep their creation order)
        int rc_o = build_ordered_targets(net, net->delivery == AHIP_DELIVERY_ORDERED ? unordered
                                                                                     : std::vector<const void*>());
        if (rc_o) return rc_o;
        for (auto& pp : net->projs) {
            Proj& p = *pp;
            if (!fused(p)) continue;
            Pop& pre = *net->pops[p.pre];
            void* g = net->pops[p.post]->slots[p.target_slot].p;
            const bool ordered = std::find(unordered.begin(), unordered.end(), g) == unordered.end();
            size_t k = 0;
            for (size_t q = 0; q < groups.size() && ordered; q++)
                for (const void* t : targets[q])
                    if (t == g) k = q + 1;
            while (k < groups.size() && (int)groups[k].size() >= kMaxScatter) k++;
            if (k == groups.size()) {
                groups.emplace_back();
                targets.emplace_back();
                blocks.push_back(0);
            }
            ScatterDesc d;
            if (p.delay_uniform >= 2) {
                // the spike list of step t - delay from the pre population's ring
                d.list = pre.hist.as<int32_t>();
                d.count = pre.hist_count.as<int32_t>();
                d.list_stride = pre.size;
                d.depth = pre.depth;
                d.delay = p.delay_uniform;
            } else {
                d.list = pre.spiked.as<int32_t>();
                d.count = pre.spike_count.as<int32_t>();
                d.list_stride = 0;
                d.depth = 1;
                d.delay = 0;
            }
            d.col_ptr = p.col_ptr.as<int64_t>();
            d.row_idx = p.row_idx.as<int32_t>();
            d.w_col = p.single_w ? nullptr : p.w_col.p;
            d.g = g;
            d.w_const = p.w_const;
            d.n_cols = p.n_cols;
            d.chunks = p.scatter_chunks;
            d.head = nullptr;
            d.cap = 0;
            d.bucket = nullptr;
            d.touched = nullptr;
            d.scnt = nullptr;
            d.shards = 1;
            d.shard_cap = 0;
            d.gid_off = 0;
            for (OrderedTarget& o : net->ordered) {
                if (o.g != g) continue;
                uint32_t off = 0;
                for (size_t q = 0; q < o.projs.size(); q++)
                    if (o.projs[q] == p.id) off = (uint32_t)o.key_off[q];
                d.head = o.head.as<uint64_t>();
                d.cap = o.cap;
                d.bucket = o.bucket.p;
                // targets delivered by the post update need no touched list
                d.touched = o.in_update ? nullptr : o.touched.as<int32_t>();
                d.scnt = o.scnt.as<int32_t>();
                d.shards = o.shards;
                d.shard_cap = o.shard_cap;
                d.gid_off = off;
            }
            groups[k].push_back(d);
            targets[k].push_back(g);
            blocks[k] += ((int64_t)pre.size * p.scatter_chunks + 3) / 4;
        }
        // each descriptor of a group gets an equal share of the launch's
        // waves (up to 2048 in all)
        for (size_t q = 0; q < groups.size(); q++) {
            ScatterGroup sg;
            sg.descs = groups[q];
            sg.n = (int32_t)groups[q].size();
            const int64_t waves = 4 * std::max<int64_t>(1, std::min<int64_t>(512, blocks[q]));
            sg.waves_per = (int32_t)std::max<int64_t>(1, waves / sg.n);
            sg.blocks = ((int64_t)sg.n * sg.waves_per + 3) / 4;
            net->scatter_groups.push_back(std::move(sg));
        }
        net->scatter_dirty = false;
    }
    // pre-loop values (hoisted step sizes) from the host mirror of the globals
    for (auto& pp : net->pops) {
        Pop& p = *pp;
        if (!p.model->preloop) continue;
        std::vector<double> vals(p.slot_host_global);
        p.model->preloop(p.slot_host_global.data(), net->dt, vals.data());
        for (int32_t k = 0; k < p.model->n_slots; k++) {
            if (p.model->slot_kind[k] != 3) continue;
            double v = vals[k];
            p.slot_host_global[k] = v;
            if (net->precision == AHIP_F64) {
                HIPCHK(hipMemcpy(p.slots[k].p, &v, sizeof(double), hipMemcpyHostToDevice));
            } else {
                float f = (float)v;
                HIPCHK(hipMemcpy(p.slots[k].p, &f, sizeof(float), hipMemcpyHostToDevice));
            }
        }
    }
    return 0;
}

// Rate projections whose panel partial sums the post population's update
// adds instead of a reduction launch: a panel layout with rows = the post
// population, the only projection writing its target, an active rate-coded
// post population without a population-wide update kernel and without a
// monitor of the target right after compute_psp, at most AHIP_MAX_PSUM per
// population.
static void assign_deferred_reduce(ahip_net* net) {
    std::vector<int32_t> n_def(net->pops.size(), 0);
    for (auto& pp : net->projs) {
        Proj& p = *pp;
        p.defer_reduce = false;
        if (p.spiking || p.delay_per_syn || !transmits(net, p)) continue;
        const Pop& post = *net->pops[p.post];
        if (post.model->spiking || post.model->has_global_update) continue;
        const RateLayout& L = p.rate;
        if (L.format != AHIP_FMT_SELL_PANEL && L.format != AHIP_FMT_CSR_PANEL) continue;
        if ((int64_t)L.n_units * L.width == 0 || p.n_rows != post.size || !L.perm.p) continue;
        bool alone = true;
        for (auto& qq : net->projs)
            if (qq.get() != &p && qq->post == p.post && qq->target_slot == p.target_slot) alone = false;
        for (auto& mp : net->mons)
            if (mp->after_psp && mp->pop == p.post && mp->slot == p.target_slot) alone = false;
        if (!alone || n_def[p.post] >= AHIP_MAX_PSUM) continue;
        n_def[p.post]++;
        p.defer_reduce = true;
    }
}

// Which projections write first into their target.
void assign_accumulate(ahip_net* net, std::vector<int32_t>& zero_before) {
    zero_before.assign(net->projs.size(), 0);
    std::vector<std::pair<int32_t, int32_t>> seen;
    for (auto& pp : net->projs) {
        Proj& p = *pp;
        if (p.spiking) { p.accumulate = true; continue; }
        std::pair<int32_t, int32_t> key(p.post, p.target_slot);
        bool first = true;
        for (auto& k : seen)
            if (k == key) first = false;
        if (first) seen.push_back(key);
        if (!first) {
            p.accumulate = true;
        } else if (layout_covers_all_rows(p)) {
            p.accumulate = false;
        } else {
            p.accumulate = true;
            zero_before[p.id] = 1;
        }
    }
    assign_deferred_reduce(net);
}

ahip_pop_args make_args(ahip_net* net, Pop& p, int32_t step) {
    ahip_pop_args a;
    std::memset(&a, 0, sizeof(a));
    a.size = p.size;
    a.step = step;
    a.t_base = net->t_dev.as<int64_t>();
    a.dt = net->dt;
    for (int32_t k = 0; k < p.model->n_slots && k < AHIP_MAX_SLOTS; k++) a.slot[k] = p.slots[k].p;
    int64_t rows = 0;
    for (int k = 0; k < AHIP_MAX_RAND; k++) {
        a.rand_inject[k] = p.rand_inject[k].p;
        if (p.rand_inject[k].p) rows = rows == 0 ? p.rand_inject_steps[k] : std::min(rows, p.rand_inject_steps[k]);
    }
    a.rand_inject_steps = rows > 0 ? rows : 1;
    a.rand_t0 = p.rand_t0;
    a.seed = net->seed;
    a.stream_id = (uint32_t)p.id;
    a.refractory = p.refractory.as<int32_t>();
    a.refractory_remaining = p.refractory_remaining.as<int32_t>();
    a.last_spike = p.last_spike.as<int64_t>();
    a.spike_mask = p.spike_mask.as<uint64_t>();
    a.block_counts = p.block_counts.as<int32_t>();
    a.deliver = p.deliver.as<ahip_deliver_desc>();
    a.n_deliver = p.n_deliver;
    for (auto& pp : net->projs) {
        const Proj& q = *pp;
        if (!q.defer_reduce || q.post != p.id || a.n_psum >= AHIP_MAX_PSUM) continue;
        const int32_t k = a.n_psum++;
        a.psum_slot[k] = q.target_slot;
        a.psum_panels[k] = q.rate.n_units;
        a.psum_acc[k] = q.accumulate ? 1 : 0;
        a.psum_stride[k] = q.n_rows;
        a.psum_part[k] = q.rate.perm.p;
    }
    return a;
}


// Enqueue one simulation step (index `step` inside the current batch).
int enqueue_step(ahip_net* net, int32_t step, const std::vector<int32_t>& zero_before) {
    hipStream_t s = net->stream;
    const size_t es = net->elem();
    // 1. compute_psp: the spiking projections in fused scatter groups, then
    // the rate-coded ones
    for (size_t gi = 0; gi < net->scatter_groups.size(); gi++) {
        const bool probe = net->probe_kind == AHIP_PROBE_SCATTER && net->probe_proj == (int32_t)gi;
        if (probe) HIPCHK(hipEventRecord(net->probe_ev[0], s));
        LCHK(launch_scatter_group(net, net->scatter_groups[gi], step, s));
        if (probe) HIPCHK(hipEventRecord(net->probe_ev[1], s));
    }
    if (net->deliver_waves > 0) {
        const bool probe = net->probe_kind == AHIP_PROBE_DELIVER;
        if (probe) HIPCHK(hipEventRecord(net->probe_ev[0], s));
        LCHK(launch_deliver(net, s));
        if (probe) HIPCHK(hipEventRecord(net->probe_ev[1], s));
    }
    for (auto& pp : net->projs) {
        Proj& p = *pp;
        Pop& pre = *net->pops[p.pre];
        Pop& post = *net->pops[p.post];
        void* target = post.slots[p.target_slot].p;
        if (p.spiking) {
            if (!transmits(net, p)) continue;
            if (p.syn) {
                if (p.nnz > 0) {
                    int rc = launch_syn(net, p, 0, step);
                    if (rc) return rc;
                }
            } else if (p.delay_per_syn && p.nnz > 0) {
                LCHK(launch_spike_ring(net, p, pre, target, step, s));
            }
            continue;
        }
        if (!transmits(net, p)) {
            // the sums of an active population are still cleared at every
            // step (OpenMPTemplates.py:266-276): a first writer that does not
            // transmit leaves 0 in its target; an inactive one keeps its sums
            if (post.active && (zero_before[p.id] || !p.accumulate))
                LCHK(launch_fill(target, post.size, (int32_t)es, 0.0, s));
            continue;
        }
        if (zero_before[p.id]) LCHK(launch_fill(target, post.size, (int32_t)es, 0.0, s));
        if (p.delay_per_syn) {
            LCHK(launch_rate_psp_delayed(net, p, pre, target, step, s));
        } else if (p.delay_uniform >= 2) {
            LCHK(launch_gather_delayed(net, pre.hist.p, p.xd.p, pre.size, pre.depth, p.delay_uniform, step, s));
            LCHK(launch_rate_psp(net, p, p.xd.p, target, s, !p.defer_reduce));
        } else {
            LCHK(launch_rate_psp(net, p, pre.slots[p.pre_slot].p, target, s, !p.defer_reduce));
        }
    }
    // 2. record targets
    for (auto& mp : net->mons) {
        Monitor& m = *mp;
        if (!m.after_psp || m.paused) continue;
        Pop& p = *net->pops[m.pop];
        LCHK(launch_record_values(m, p.slots[m.slot].p, p.size, (int32_t)es,
                                  net->t_dev.as<int64_t>(), step, s));
    }
    // 3. neuron updates (+ spike flags) of every population, 4. the spike
    // compactions of the updated spiking populations in one launch (no
    // update reads another population's compaction), then each population's
    // firing-rate window and delay push
    std::vector<int32_t> compact;
    for (auto& pp : net->pops) {
        Pop& p = *pp;
        if (!p.active) {
            if (p.depth > 1 && p.hist.p) {
                const size_t slot = (size_t)p.size * (p.model->spiking ? sizeof(int32_t) : es);
                int rc = rotate_ring(p.hist, slot, p.depth, p.ring_tmp, s);
                if (!rc && p.model->spiking) rc = rotate_ring(p.hist_count, sizeof(int32_t), p.depth, p.ring_tmp, s);
                if (rc) return rc;
            }
            if (p.model->spiking && p.fr_window > 0) LCHK(launch_update_rate(net, p, step, s, false));
            continue;
        }
        ahip_pop_args a = make_args(net, p, step);
        ahip_launch_fn fn = net->precision == AHIP_F64 ? p.model->launch_f64 : p.model->launch_f32;
        const bool probe_u = net->probe_kind == AHIP_PROBE_UPDATE && net->probe_proj == p.id;
        if (probe_u) HIPCHK(hipEventRecord(net->probe_ev[0], s));
        int rc = fn(&a, (void*)s);
        if (rc != 0) return fail(AHIP_EHIP, "update of population %d failed: %s", p.id,
                                 hipGetErrorString((hipError_t)rc));
        if (probe_u) HIPCHK(hipEventRecord(net->probe_ev[1], s));
        if (p.model->spiking) compact.push_back(p.id);
    }
    if (!compact.empty()) {
        const bool probe_c = net->probe_kind == AHIP_PROBE_COMPACT &&
                             std::find(compact.begin(), compact.end(), net->probe_proj) != compact.end();
        if (probe_c) HIPCHK(hipEventRecord(net->probe_ev[0], s));
        // each population's first spike monitor is filled by its compaction
        LCHK(launch_compact(net, compact, net->t_dev.as<int64_t>(), step, s));
        if (probe_c) HIPCHK(hipEventRecord(net->probe_ev[1], s));
    }
    for (auto& pp : net->pops) {
        Pop& p = *pp;
        if (!p.active) continue;
        if (p.model->spiking && p.fr_window > 0) LCHK(launch_update_rate(net, p, step, s));
        // delay queues: push this step's values (after the update and the
        // compaction, as the reference's update_delay)
        if (p.depth > 1) {
            if (p.model->spiking) {
                LCHK(launch_push_spikes(net, p, step, s));
            } else {
                const int32_t rs = p.slot_index("r");
                if (rs >= 0) LCHK(launch_push_values(net, p.slots[rs].p, p.hist.p, p.size, p.depth, step, s));
            }
        }
    }
    // plastic projections, after every population update: update_synapse,
    // then the post events (OpenMPBaseTemplate.py:217-325)
    for (auto& pp : net->projs) {
        Proj& p = *pp;
        if (!p.syn || p.nnz == 0 || !transmits(net, p)) continue;
        if (p.syn->has_update && p.syn_update) {
            int rc = launch_syn(net, p, 1, step);
            if (rc) return rc;
        }
        if (p.spiking && p.syn->has_post) {
            int rc = launch_syn(net, p, 2, step);
            if (rc) return rc;
        }
    }
    // 5. record variables and spikes
    for (auto& mp : net->mons) {
        Monitor& m = *mp;
        if (m.after_psp || m.paused) continue;
        Pop& p = *net->pops[m.pop];
        if (m.spikes) {
            // an inactive population keeps its last spike list, which its
            // monitors record again (its compaction does not run)
            if (first_spike_monitor(net, m.pop) != &m || !p.active)
                LCHK(launch_record_spikes(m, p, net->t_dev.as<int64_t>(), step, s));
        } else
            LCHK(launch_record_values(m, p.slots[m.slot].p, p.size, (int32_t)es,
                                      net->t_dev.as<int64_t>(), step, s));
    }
    return 0;
}

// Persistent steps (ahip_persist_args, annarchy_hip_model.h): a network of
// one active spiking population whose projections all start and end in it,
// each with one weight, no delay and no synapse model, recorded at most by
// one spike monitor, small enough for the state its statements touch, its
// conductances and its spike list to sit in one workgroup's LDS, runs a
// whole batch of steps in one launch of one workgroup.
constexpr size_t kPersistLdsMax = 160 * 1024 - 1024;   // dynamic LDS, the kernel's static part aside

bool persist_plan(const ahip_net* net, ahip_persist_args& pa, size_t& lds, int32_t& pop_id) {
    if (!net->persist_enabled || net->pops.size() != 1) return false;
    if (net->probe_kind >= 0 && net->probe_kind != AHIP_PROBE_PERSIST) return false;
    const Pop& p = *net->pops[0];
    const ahip_persist_fn fn = net->precision == AHIP_F64 ? p.model->persist_f64 : p.model->persist_f32;
    if (!fn || !p.model->spiking || !p.active || p.depth > 1 || p.fr_window > 0 || p.n_deliver > 0) return false;
    if (p.size > 32 * AHIP_PERSIST_THREADS) return false;
    std::memset(&pa, 0, sizeof(pa));
    const int32_t n_ps = p.model->persist_n_slots;
    std::vector<int32_t> round_of(n_ps, -1);   // per LDS slot: the last round that writes it
    for (auto& pp : net->projs) {
        const Proj& q = *pp;
        if (!q.spiking || q.pre != 0 || q.post != 0) return false;
        if (q.syn || q.delay_per_syn || q.delay_uniform > 1 || !q.single_w) return false;
        if (!transmits(net, q) || q.nnz == 0) continue;
        if (pa.n_proj >= AHIP_PERSIST_MAX_PROJ) return false;
        int32_t tgt = -1;
        for (int32_t k = 0; k < n_ps; k++)
            if (p.model->persist_slots[k] == q.target_slot) tgt = k;
        if (tgt < 0) return false;
        ahip_persist_proj& d = pa.proj[pa.n_proj++];
        d.col_ptr = q.col_ptr.as<int64_t>();
        d.row_idx = q.row_idx.as<int32_t>();
        d.w = q.w_const;
        d.n_cols = q.n_cols;
        d.target = tgt;
        // projections into one conductance add their events in creation order
        d.round = round_of[tgt] + 1;
        round_of[tgt] = d.round;
        pa.n_rounds = std::max(pa.n_rounds, d.round + 1);
    }
    const Monitor* rec = nullptr;
    for (auto& mp : net->mons) {
        const Monitor& m = *mp;
        if (m.paused) continue;
        if (!m.spikes || m.pop != 0 || rec) return false;
        rec = &m;
    }
    // LDS: the slots, the refractory counters, the spike list, the column
    // ranges of the projections when they fit (else read from col_ptr), then
    // (8-byte aligned) the scatter's item tables
    const size_t es = net->elem(), n = (size_t)p.size;
    const size_t fixed = (size_t)n_ps * n * es + (p.model->has_refractory ? n * sizeof(int32_t) : 0) +
                         n * sizeof(int32_t);
    const size_t tables = (size_t)AHIP_PERSIST_ITEMS * sizeof(int64_t) +
                          (size_t)(2 * AHIP_PERSIST_ITEMS + 1) * sizeof(int32_t);
    int32_t n_cols = 0;
    bool narrow = true;   // column offsets fit in int32
    for (int32_t k = 0; k < pa.n_proj; k++) n_cols += pa.proj[k].n_cols + 1;
    for (auto& pp : net->projs) narrow &= pp->nnz < (int64_t)INT32_MAX;
    const auto total = [&](int32_t cols) { return ((fixed + (size_t)cols * sizeof(int32_t) + 7) & ~(size_t)7) + tables; };
    if (narrow && total(n_cols) <= kPersistLdsMax) {
        pa.n_lds_cols = n_cols;
        for (int32_t k = 0, o = 0; k < pa.n_proj; k++) {
            pa.proj[k].lds_cols = o;
            o += pa.proj[k].n_cols + 1;
        }
    } else {
        pa.n_lds_cols = 0;
        for (int32_t k = 0; k < pa.n_proj; k++) pa.proj[k].lds_cols = -1;
    }
    lds = total(pa.n_lds_cols);
    if (lds > kPersistLdsMax) return false;
    pa.spiked = p.spiked.as<int32_t>();
    pa.spike_count = p.spike_count.as<int32_t>();
    if (rec) {
        pa.rec_counts = rec->counts.as<int32_t>();
        pa.rec_idx = rec->data.as<int32_t>();
        pa.rec_fill = rec->fill.as<int64_t>();
        pa.rec_idx_capacity = rec->idx_capacity;
        pa.rec_t_start = rec->t_start;
        pa.rec_capacity = rec->capacity;
    }
    pop_id = 0;
    return true;
}

int enqueue_batch(ahip_net* net, int32_t nsteps, const std::vector<int32_t>& zb) {
    ahip_persist_args pa;
    size_t lds = 0;
    int32_t pid = -1;
    if (nsteps > 0 && persist_plan(net, pa, lds, pid)) {
        Pop& p = *net->pops[pid];
        pa.steps = nsteps;
        ahip_pop_args a = make_args(net, p, 0);
        const ahip_persist_fn fn = net->precision == AHIP_F64 ? p.model->persist_f64 : p.model->persist_f32;
        const bool probe = net->probe_kind == AHIP_PROBE_PERSIST;
        if (probe) HIPCHK(hipEventRecord(net->probe_ev[0], net->stream));
        const int rc = fn(&a, &pa, lds, (void*)net->stream);
        if (probe && rc == 0) HIPCHK(hipEventRecord(net->probe_ev[1], net->stream));
        if (rc != 0) return fail(AHIP_EHIP, "persistent steps of population %d failed: %s", p.id,
                                 hipGetErrorString((hipError_t)rc));
        LCHK(launch_advance_time(net->t_dev.as<int64_t>(), nsteps, net->stream));
        return 0;
    }
    for (int32_t j = 0; j < nsteps; j++) {
        int rc = enqueue_step(net, j, zb);
        if (rc) return rc;
    }
    LCHK(launch_advance_time(net->t_dev.as<int64_t>(), nsteps, net->stream));
    return 0;
}

// Captures `steps` steps with the scatter tables and buffers that prepare()
// left in place (it must run first); only the previous executable graph is
// dropped, so the tables stay valid for as long as the new graph lives.
int build_graph(ahip_net* net, int32_t steps, const std::vector<int32_t>& zb) {
    drop_graph(net);
    HIPCHK(hipStreamBeginCapture(net->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_batch(net, steps, zb);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(net->stream, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(AHIP_EHIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&net->graph_exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        net->graph_exec = nullptr;
        return fail(AHIP_EHIP, "graph instantiation failed: %s", hipGetErrorString(e));
    }
    net->graph_steps = steps;
    net->graph_scatter_gen = net->scatter_gen;
    net->dirty = false;
    return 0;
}

// A captured graph launches only with the scatter tables it was captured with.
int check_graph_tables(const ahip_net* net) {
    if (net->graph_scatter_gen != net->scatter_gen || net->scatter_dirty)
        return fail(AHIP_ESTATE, "internal: the step graph was captured with scatter tables that were rebuilt since");
    return 0;
}

int32_t auto_batch(ahip_net* net) {
    if (net->graph_batch > 0) return net->graph_batch;
    // keep a graph around a few thousand nodes
    int64_t nodes = 1;
    for (auto& p : net->pops) nodes += p->model->spiking ? 2 : 1;
    nodes += (int64_t)net->projs.size() * 2 + (int64_t)net->mons.size();
    int64_t b = 2000 / std::max<int64_t>(1, nodes);
    return (int32_t)std::max<int64_t>(1, std::min<int64_t>(b, 200));
}

template <typename T>
void to_host_t(const void* dev, double* out, int64_t n) {
    std::vector<T> tmp(n);
    hipMemcpy(tmp.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost);
    for (int64_t i = 0; i < n; i++) out[i] = (double)tmp[i];
}

}  // namespace

// The message of ahip_last_error() for failures detected in other
// translation units; returns `code`.
int ahip::report_error(int code, const char* msg) {
    g_err = msg;
    return code;
}

void ahip::DevBuf::release() {
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
}

int ahip::Pop::slot_index(const std::string& name) const {
    for (int32_t k = 0; k < model->n_slots; k++)
        if (name == model->slot_names[k]) return k;
    return -1;
}

extern "C" {

const char* ahip_last_error(void) { return g_err.c_str(); }

const char* ahip_build_info(void) { return "annarchy_hip 0.1 gfx950 wave64"; }

int ahip_probe_copy(int64_t bytes, int32_t iters, double* avg_ms) {
    if (bytes < 16 || iters <= 0 || !avg_ms) return fail(AHIP_EINVAL, "bad copy probe arguments");
    const int64_t n16 = bytes / 16;
    DevBuf a, b;
    int rc;
    if ((rc = dev_alloc(a, n16 * 16))) return rc;
    if ((rc = dev_alloc(b, n16 * 16))) return rc;
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    LCHK(launch_copy16(a.p, b.p, n16, s));
    HIPCHK(hipEventRecord(e0, s));
    for (int32_t i = 0; i < iters; i++) LCHK(launch_copy16(a.p, b.p, n16, s));
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(s);
    *avg_ms = (double)ms / iters;
    return 0;
}

int ahip_device_count(int32_t* count) {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    *count = n;
    return 0;
}

int ahip_set_device(int32_t device) {
    HIPCHK(hipSetDevice(device));
    return 0;
}

int ahip_net_create(ahip_net** out) {
    if (!out) return fail(AHIP_EINVAL, "null output pointer");
    ahip_net* net = new ahip_net();
    hipError_t e = hipGetDevice(&net->device);
    if (e != hipSuccess) {
        delete net;
        return fail(AHIP_EHIP, "no HIP device: %s", hipGetErrorString(e));
    }
    e = hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete net;
        return fail(AHIP_EHIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    *out = net;
    return 0;
}

int ahip_net_destroy(ahip_net* net) {
    if (!net) return 0;
    hipSetDevice(net->device);
    hipStreamSynchronize(net->stream);
    invalidate(net);
    net->mons.clear();
    net->projs.clear();
    for (auto& p : net->pops) {
        for (auto& s : p->slots) s.release();
        // the model library stays mapped: kernels may still be referenced by HIP
    }
    net->pops.clear();
    net->t_dev.release();
    hipStreamDestroy(net->stream);
    delete net;
    return 0;
}

int ahip_net_initialize(ahip_net* net, double dt, int32_t precision) {
    if (!net) return fail(AHIP_EINVAL, "null network");
    if (!(dt > 0.0)) return fail(AHIP_EINVAL, "dt must be positive");
    if (precision != AHIP_F64 && precision != AHIP_F32) return fail(AHIP_EINVAL, "bad precision");
    if (net->initialized) return fail(AHIP_ESTATE, "network already initialized");
    net->dt = dt;
    net->precision = precision;
    int rc = dev_alloc(net->t_dev, sizeof(int64_t));
    if (rc) return rc;
    net->initialized = true;
    return 0;
}

int ahip_set_seed(ahip_net* net, uint64_t seed) {
    if (!net) return fail(AHIP_EINVAL, "null network");
    net->seed = seed;
    invalidate(net);
    return 0;
}

int ahip_set_graph_batch(ahip_net* net, int32_t steps) {
    if (!net || steps < 0) return fail(AHIP_EINVAL, "bad graph batch");
    net->graph_batch = steps;
    invalidate(net);
    return 0;
}

int ahip_set_graph_enabled(ahip_net* net, int32_t enabled) {
    if (!net) return fail(AHIP_EINVAL, "null network");
    net->graph_enabled = enabled != 0;
    invalidate(net);
    return 0;
}

int ahip_graph_prepare(ahip_net* net) {
    if (!net) return fail(AHIP_EINVAL, "null network");
    if (!net->graph_enabled) return fail(AHIP_ESTATE, "graph capture is disabled");
    HIPCHK(hipSetDevice(net->device));
    int rc = prepare(net);
    if (rc) return rc;
    std::vector<int32_t> zb;
    assign_accumulate(net, zb);
    const int32_t G = auto_batch(net);
    if (net->dirty || !net->graph_exec || net->graph_steps != G) {
        if ((rc = build_graph(net, G, zb))) return rc;
    }
    if ((rc = check_graph_tables(net))) return rc;
    HIPCHK(hipGraphUpload(net->graph_exec, net->stream));
    HIPCHK(hipStreamSynchronize(net->stream));
    return 0;
}

int ahip_get_time(ahip_net* net, int64_t* t) {
    if (!net || !t) return fail(AHIP_EINVAL, "null argument");
    *t = net->t;
    return 0;
}

int ahip_set_time(ahip_net* net, int64_t t) {
    if (!net || !net->initialized) return fail(AHIP_ESTATE, "network not initialized");
    HIPCHK(hipStreamSynchronize(net->stream));
    HIPCHK(hipMemcpy(net->t_dev.p, &t, sizeof(int64_t), hipMemcpyHostToDevice));
    net->t = t;
    return 0;
}

int ahip_run(ahip_net* net, int64_t nsteps) {
    if (!net) return fail(AHIP_EINVAL, "null network");
    if (nsteps < 0) return fail(AHIP_EINVAL, "negative step count");
    HIPCHK(hipSetDevice(net->device));
    int rc = prepare(net);
    if (rc) return rc;
    std::vector<int32_t> zb;
    assign_accumulate(net, zb);
    int64_t left = nsteps;
    const int32_t G = auto_batch(net);
    if (net->graph_enabled && left >= G) {
        if (net->dirty || !net->graph_exec || net->graph_steps != G) {
            rc = build_graph(net, G, zb);
            if (rc) return rc;
        }
        if ((rc = check_graph_tables(net))) return rc;
        while (left >= G) {
            HIPCHK(hipGraphLaunch(net->graph_exec, net->stream));
            left -= G;
        }
    }
    while (left > 0) {
        const int32_t n = (int32_t)std::min<int64_t>(left, G);
        rc = enqueue_batch(net, n, zb);
        if (rc) return rc;
        left -= n;
    }
    HIPCHK(hipStreamSynchronize(net->stream));
    net->t += nsteps;
    return 0;
}

int ahip_step(ahip_net* net) { return ahip_run(net, 1); }

// stop_condition() of one population after a step
static int stop_reached(ahip_net* net, Pop& p, bool* out) {
    int32_t c = 0;
    if (p.stop_mode >= 3) {
        HIPCHK(hipMemcpy(&c, p.spike_count.p, sizeof(int32_t), hipMemcpyDeviceToHost));
        *out = p.stop_mode == 3 ? c > 0 : c == p.size;
        return 0;
    }
    auto fn = net->precision == AHIP_F64 ? p.stop_f64 : p.stop_f32;
    if (!fn) return fail(AHIP_EMODEL, "population %d has no compiled stop condition", p.id);
    if (!p.stop_count.p) {
        int rc = dev_alloc(p.stop_count, sizeof(int32_t));
        if (rc) return rc;
    }
    HIPCHK(hipMemsetAsync(p.stop_count.p, 0, sizeof(int32_t), net->stream));
    ahip_pop_args a = make_args(net, p, 0);
    int rc = fn(&a, p.stop_count.as<int32_t>(), (void*)net->stream);
    if (rc) return fail(AHIP_EHIP, "stop condition of population %d failed: %s", p.id,
                        hipGetErrorString((hipError_t)rc));
    HIPCHK(hipMemcpyAsync(&c, p.stop_count.p, sizeof(int32_t), hipMemcpyDeviceToHost, net->stream));
    HIPCHK(hipStreamSynchronize(net->stream));
    *out = p.stop_mode == 1 ? c > 0 : c == p.size;
    return 0;
}

int ahip_pop_set_stop_condition(ahip_net* net, int32_t pop, int32_t mode) {
    Pop* p = get_pop(net, pop);
    if (!p || mode < 0 || mode > 4) return fail(AHIP_EINVAL, "bad population or stop mode");
    if ((mode == 1 || mode == 2) && !(net->precision == AHIP_F64 ? p->stop_f64 : p->stop_f32))
        return fail(AHIP_EMODEL, "the model library of population %d has no stop condition", pop);
    if (mode >= 3 && !p->model->spiking) return fail(AHIP_EINVAL, "\"spiked\" needs a spiking population");
    p->stop_mode = mode;
    return 0;
}

int ahip_run_until(ahip_net* net, int64_t max_steps, const int32_t* pops, int32_t npops,
                   int32_t and_or, int64_t* done) {
    if (!net || !done || (npops > 0 && !pops)) return fail(AHIP_EINVAL, "null argument");
    for (int32_t k = 0; k < npops; k++)
        if (!get_pop(net, pops[k])) return fail(AHIP_EINVAL, "bad population %d", pops[k]);
    // without any stop condition in the network: run(steps) (OpenMPBaseTemplate.py:485-489)
    bool any_cond = false;
    for (auto& pp : net->pops) any_cond = any_cond || pp->stop_mode != 0;
    if (!any_cond) {
        int rc = ahip_run(net, max_steps);
        if (rc) return rc;
        *done = max_steps;
        return 0;
    }
    // step by step; stop starts as and_or and folds the listed populations'
    // conditions with && (and_or = 1) or || (OpenMPBaseTemplate.py:490-517)
    int64_t n = 0;
    while (n < max_steps) {
        int rc = ahip_run(net, 1);
        if (rc) return rc;
        n++;
        bool stop = and_or != 0;
        for (int32_t k = 0; k < npops; k++) {
            Pop& p = *net->pops[pops[k]];
            if (p.stop_mode == 0) continue;
            bool c = false;
            if ((rc = stop_reached(net, p, &c))) return rc;
            stop = and_or != 0 ? (stop && c) : (stop || c);
        }
        if (stop && npops > 0) break;
    }
    *done = n;
    return 0;
}

// ---- populations -----------------------------------------------------------

int ahip_pop_create(ahip_net* net, int32_t size, const char* model_lib, int32_t* pop_id) {
    if (!net || !model_lib || !pop_id) return fail(AHIP_EINVAL, "null argument");
    if (!net->initialized) return fail(AHIP_ESTATE, "initialize the network first");
    if (size <= 0) return fail(AHIP_EINVAL, "population size must be positive");
    void* dl = dlopen(model_lib, RTLD_NOW | RTLD_LOCAL);
    if (!dl) return fail(AHIP_EMODEL, "cannot load model library %s: %s", model_lib, dlerror());
    auto describe = (ahip_model_describe_fn)dlsym(dl, "ahip_model_describe");
    if (!describe) return fail(AHIP_EMODEL, "%s does not export ahip_model_describe", model_lib);
    const ahip_model_desc* d = describe();
    if (!d || d->abi != AHIP_MODEL_ABI)
        return fail(AHIP_EMODEL, "%s: model ABI %d, runtime expects %d", model_lib, d ? d->abi : -1,
                    AHIP_MODEL_ABI);
    if (d->n_slots > AHIP_MAX_SLOTS || d->n_rand > AHIP_MAX_RAND)
        return fail(AHIP_EMODEL, "%s: too many attributes", model_lib);
    auto p = std::make_unique<Pop>();
    p->id = (int32_t)net->pops.size();
    p->size = size;
    p->lib_path = model_lib;
    p->dl = dl;
    p->model = d;
    // optional: the population's stop condition compiled into the library
    p->stop_f64 = (int (*)(const ahip_pop_args*, int32_t*, void*))dlsym(dl, "ahip_model_stop_f64");
    p->stop_f32 = (int (*)(const ahip_pop_args*, int32_t*, void*))dlsym(dl, "ahip_model_stop_f32");
    const size_t es = net->elem();
    p->slots.resize(d->n_slots);
    p->slot_host_global.assign(d->n_slots, 0.0);
    for (int32_t k = 0; k < d->n_slots; k++) {
        const int64_t n = d->slot_local[k] ? size : 1;
        int rc = dev_alloc(p->slots[k], n * es);
        if (rc) return rc;
        LCHK(launch_fill(p->slots[k].p, n, (int32_t)es, d->slot_init[k], net->stream));
        if (!d->slot_local[k]) p->slot_host_global[k] = d->slot_init[k];
    }
    if (d->spiking) {
        int rc;
        if ((rc = dev_alloc(p->refractory, size * sizeof(int32_t)))) return rc;
        if ((rc = dev_alloc(p->refractory_remaining, size * sizeof(int32_t)))) return rc;
        if ((rc = dev_alloc(p->last_spike, size * sizeof(int64_t)))) return rc;
        if ((rc = dev_alloc(p->spike_mask, (size_t)p->nwords() * sizeof(uint64_t)))) return rc;
        if ((rc = dev_alloc(p->block_counts, (size_t)p->nblocks() * sizeof(int32_t)))) return rc;
        if ((rc = dev_alloc(p->spiked, size * sizeof(int32_t)))) return rc;
        if ((rc = dev_alloc(p->spike_count, sizeof(int32_t)))) return rc;
        LCHK(launch_fill_i64(p->last_spike.as<int64_t>(), size, -10000, net->stream));
    }
    HIPCHK(hipStreamSynchronize(net->stream));
    *pop_id = p->id;
    net->pops.push_back(std::move(p));
    invalidate(net);
    return 0;
}

int ahip_pop_size(ahip_net* net, int32_t pop, int32_t* size) {
    Pop* p = get_pop(net, pop);
    if (!p || !size) return fail(AHIP_EINVAL, "bad population");
    *size = p->size;
    return 0;
}

int ahip_pop_set(ahip_net* net, int32_t pop, const char* attr, const double* values, int64_t n) {
    Pop* p = get_pop(net, pop);
    if (!p || !attr || !values) return fail(AHIP_EINVAL, "bad population or attribute");
    int k = p->slot_index(attr);
    if (k < 0) return fail(AHIP_EINVAL, "population %d has no attribute '%s'", pop, attr);
    const int64_t want = p->model->slot_local[k] ? p->size : 1;
    if (n != want) return fail(AHIP_EINVAL, "attribute '%s' expects %lld values, got %lld", attr,
                               (long long)want, (long long)n);
    HIPCHK(hipStreamSynchronize(net->stream));
    if (net->precision == AHIP_F64) {
        HIPCHK(hipMemcpy(p->slots[k].p, values, n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(n);
        for (int64_t i = 0; i < n; i++) f[i] = (float)values[i];
        HIPCHK(hipMemcpy(p->slots[k].p, f.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
    if (!p->model->slot_local[k]) p->slot_host_global[k] = values[0];
    return 0;
}

int ahip_pop_get(ahip_net* net, int32_t pop, const char* attr, double* values, int64_t n) {
    Pop* p = get_pop(net, pop);
    if (!p || !attr || !values) return fail(AHIP_EINVAL, "bad population or attribute");
    int k = p->slot_index(attr);
    if (k < 0) return fail(AHIP_EINVAL, "population %d has no attribute '%s'", pop, attr);
    const int64_t want = p->model->slot_local[k] ? p->size : 1;
    if (n != want) return fail(AHIP_EINVAL, "attribute '%s' holds %lld values, got room for %lld",
                               attr, (long long)want, (long long)n);
    HIPCHK(hipStreamSynchronize(net->stream));
    if (net->precision == AHIP_F64) {
        HIPCHK(hipMemcpy(values, p->slots[k].p, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        to_host_t<float>(p->slots[k].p, values, n);
    }
    return 0;
}

static DevBuf* int_attr(Pop* p, const char* attr) {
    if (!p->model->spiking) return nullptr;
    if (!strcmp(attr, "refractory")) return &p->refractory;
    if (!strcmp(attr, "refractory_remaining")) return &p->refractory_remaining;
    return nullptr;
}

int ahip_pop_set_int(ahip_net* net, int32_t pop, const char* attr, const int32_t* values, int64_t n) {
    Pop* p = get_pop(net, pop);
    if (!p || !attr || !values) return fail(AHIP_EINVAL, "bad population or attribute");
    DevBuf* b = int_attr(p, attr);
    if (!b) return fail(AHIP_EINVAL, "population %d has no integer attribute '%s'", pop, attr);
    if (n != p->size) return fail(AHIP_EINVAL, "'%s' expects %d values", attr, p->size);
    HIPCHK(hipStreamSynchronize(net->stream));
    HIPCHK(hipMemcpy(b->p, values, n * sizeof(int32_t), hipMemcpyHostToDevice));
    return 0;
}

int ahip_pop_get_int(ahip_net* net, int32_t pop, const char* attr, int32_t* values, int64_t n) {
    Pop* p = get_pop(net, pop);
    if (!p || !attr || !values) return fail(AHIP_EINVAL, "bad population or attribute");
    DevBuf* b = int_attr(p, attr);
    if (!b) return fail(AHIP_EINVAL, "population %d has no integer attribute '%s'", pop, attr);
    if (n != p->size) return fail(AHIP_EINVAL, "'%s' holds %d values", attr, p->size);
    HIPCHK(hipStreamSynchronize(net->stream));
    HIPCHK(hipMemcpy(values, b->p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return 0;
}

int ahip_pop_get_last_spike(ahip_net* net, int32_t pop, int64_t* values, int64_t n) {
    Pop* p = get_pop(net, pop);
    if (!p || !values || !p->model->spiking) return fail(AHIP_EINVAL, "bad spiking population");
    if (n != p->size) return fail(AHIP_EINVAL, "last_spike holds %d values", p->size);
    HIPCHK(hipStreamSynchronize(net->stream));
    HIPCHK(hipMemcpy(values, p->last_spike.p, n * sizeof(int64_t), hipMemcpyDeviceToHost));
    return 0;
}

int ahip_pop_spikes(ahip_net* net, int32_t pop, int32_t* out, int32_t cap, int32_t* n) {
    Pop* p = get_pop(net, pop);
    if (!p || !n || !p->model->spiking) return fail(AHIP_EINVAL, "bad spiking population");
    HIPCHK(hipStreamSynchronize(net->stream));
    int32_t c = 0;
    HIPCHK(hipMemcpy(&c, p->spike_count.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    *n = c;
    if (c > 0 && out) {
        if (c > cap) return fail(AHIP_EINVAL, "spike buffer too small (%d > %d)", c, cap);
        HIPCHK(hipMemcpy(out, p->spiked.p, c * sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    return 0;
}

int ahip_pop_reset_spiking(ahip_net* net, int32_t pop) {
    Pop* p = get_pop(net, pop);
    if (!p || !p->model->spiking) return fail(AHIP_EINVAL, "bad spiking population");
    LCHK(launch_fill_i32(p->refractory_remaining.as<int32_t>(), p->size, 0, net->stream));
    LCHK(launch_fill_i64(p->last_spike.as<int64_t>(), p->size, -10000, net->stream));
    LCHK(launch_fill_i32(p->spike_count.as<int32_t>(), 1, 0, net->stream));
    if (p->fr_window > 0) {
        // the spike history of the mean firing rate is emptied (reset_FR)
        LCHK(launch_fill_i32(p->fr_bits.as<int32_t>(), (int64_t)((p->fr_window + 31) / 32) * p->size, 0,
                             net->stream));
        LCHK(launch_fill_i32(p->fr_count.as<int32_t>(), p->size, 0, net->stream));
    }
    HIPCHK(hipStreamSynchronize(net->stream));
    return 0;
}

int ahip_pop_set_active(ahip_net* net, int32_t pop, int32_t active) {
    Pop* p = get_pop(net, pop);
    if (!p) return fail(AHIP_EINVAL, "bad population");
    HIPCHK(hipStreamSynchronize(net->stream));
    p->active = active != 0;
    invalidate(net);
    return 0;
}

int ahip_pop_size_in_bytes(ahip_net* net, int32_t pop, int64_t* bytes) {
    Pop* p = get_pop(net, pop);
    if (!p || !bytes) return fail(AHIP_EINVAL, "bad population");
    size_t b = 0;
    for (const DevBuf& d : p->slots) b += d.bytes;
    for (const DevBuf* d : std::initializer_list<const DevBuf*>{&p->refractory, &p->refractory_remaining, &p->last_spike, &p->spike_mask,
                            &p->block_counts, &p->spiked, &p->spike_count, &p->hist, &p->hist_count,
                            &p->ring_tmp, &p->fr_bits, &p->fr_count})
        b += d->bytes;
    for (const DevBuf& d : p->rand_inject) b += d.bytes;
    *bytes = (int64_t)b;
    return 0;
}

int ahip_proj_size_in_bytes(ahip_net* net, int32_t proj, int64_t* bytes) {
    Proj* p = get_proj(net, proj);
    if (!p || !bytes) return fail(AHIP_EINVAL, "bad projection");
    const RateLayout& L = p->rate;
    size_t b = 0;
    for (const DevBuf* d : std::initializer_list<const DevBuf*>{&L.ptr, &L.idx, &L.val, &L.rl, &L.perm, &L.aux, &L.tiles, &p->col_ptr, &p->row_idx,
                            &p->w_col, &p->delay_dev, &p->ring, &p->xd, &p->syn_rec, &p->last_event,
                            &p->fwd2col, &p->row_ptr_dev, &p->col_dev})
        b += d->bytes;
    for (const DevBuf& d : p->syn_attr) b += d.bytes;
    *bytes = (int64_t)b;
    return 0;
}

int ahip_pop_compute_firing_rate(ahip_net* net, int32_t pop, double window_ms) {
    Pop* p = get_pop(net, pop);
    if (!p || !p->model->spiking) return fail(AHIP_EINVAL, "compute_firing_rate applies to spiking populations");
    if (p->slot_index("r") < 0) return fail(AHIP_EMODEL, "the neuron model has no variable r");
    if (!(window_ms > 0.0)) return 0;   // as the reference: a non-positive window leaves the state alone
    const int64_t W = (int64_t)(window_ms / net->dt);
    if (W < 1 || W > (1 << 20)) return fail(AHIP_EINVAL, "the window must span 1 .. 2^20 steps");
    HIPCHK(hipSetDevice(net->device));
    HIPCHK(hipStreamSynchronize(net->stream));
    if (W != p->fr_window) {
        int rc;
        const int64_t words = (W + 31) / 32;
        if ((rc = dev_alloc(p->fr_bits, (size_t)(words * p->size) * sizeof(uint32_t)))) return rc;
        if ((rc = dev_alloc(p->fr_count, (size_t)p->size * sizeof(int32_t)))) return rc;
        p->fr_window = (int32_t)W;
    }
    p->fr_rate = 1000.0 / window_ms;
    invalidate(net);
    return 0;
}

int ahip_pop_inject_random(ahip_net* net, int32_t pop, int32_t k, const double* buf,
                           int64_t nsteps) {
    Pop* p = get_pop(net, pop);
    if (!p || k < 0 || k >= p->model->n_rand) return fail(AHIP_EINVAL, "bad population or random variable");
    HIPCHK(hipStreamSynchronize(net->stream));
    invalidate(net);
    if (!buf) {
        p->rand_inject[k].release();
        p->rand_inject_steps[k] = 0;
        return 0;
    }
    if (nsteps <= 0) return fail(AHIP_EINVAL, "nsteps must be positive");
    const int64_t n = nsteps * p->size;
    int rc = dev_alloc(p->rand_inject[k], n * net->elem());
    if (rc) return rc;
    if (net->precision == AHIP_F64) {
        HIPCHK(hipMemcpy(p->rand_inject[k].p, buf, n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(n);
        for (int64_t i = 0; i < n; i++) f[i] = (float)buf[i];
        HIPCHK(hipMemcpy(p->rand_inject[k].p, f.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
    p->rand_inject_steps[k] = nsteps;
    p->rand_t0 = net->t;
    return 0;
}

// ---- projections -----------------------------------------------------------

int ahip_proj_create(ahip_net* net, int32_t pre, int32_t post, const char* target,
                     int32_t spiking, int32_t* proj_id) {
    if (!net || !target || !proj_id) return fail(AHIP_EINVAL, "null argument");
    Pop* a = get_pop(net, pre);
    Pop* b = get_pop(net, post);
    if (!a || !b) return fail(AHIP_EINVAL, "bad pre/post population");
    auto p = std::make_unique<Proj>();
    p->id = (int32_t)net->projs.size();
    p->pre = pre;
    p->post = post;
    p->target = target;
    p->spiking = spiking != 0;
    p->nnz = -1;
    if (p->spiking) {
        if (!a->model->spiking) return fail(AHIP_EINVAL, "spiking projection needs a spiking pre-synaptic population");
        p->target_slot = b->slot_index(std::string("g_") + target);
        if (p->target_slot < 0 || !b->model->slot_local[p->target_slot])
            return fail(AHIP_EINVAL, "post-synaptic population has no conductance g_%s", target);
    } else {
        p->target_slot = b->slot_index(std::string("_sum_") + target);
        if (p->target_slot < 0) return fail(AHIP_EINVAL, "post-synaptic population has no target sum(%s)", target);
        p->pre_slot = a->slot_index("r");
        if (p->pre_slot < 0 || !a->model->slot_local[p->pre_slot])
            return fail(AHIP_EINVAL, "rate projection needs a pre-synaptic variable r");
    }
    *proj_id = p->id;
    net->projs.push_back(std::move(p));
    invalidate(net);
    return 0;
}

static int build_backward(ahip_net* net, Proj& p) {
    // CSRC backward view (include/CSRCMatrix.hpp:164-210): per pre neuron the
    // post ranks ascending; weights are stored in column order.
    std::vector<int64_t> col_ptr(p.n_cols + 1, 0);
    for (int64_t j = 0; j < p.nnz; j++) col_ptr[p.h_col[j] + 1]++;
    for (int32_t c = 0; c < p.n_cols; c++) col_ptr[c + 1] += col_ptr[c];
    int64_t max_len = 0;
    for (int32_t c = 0; c < p.n_cols; c++) max_len = std::max(max_len, col_ptr[c + 1] - col_ptr[c]);
    p.scatter_chunks = (int32_t)std::min<int64_t>(16, std::max<int64_t>(1, (max_len + 255) / 256));
    std::vector<int64_t> cur(col_ptr.begin(), col_ptr.end() - 1);
    std::vector<int32_t> row_idx(p.nnz);
    p.h_inv_idx.assign(p.nnz, 0);
    for (int32_t r = 0; r < p.n_rows; r++)
        for (int64_t j = p.h_row_ptr[r]; j < p.h_row_ptr[r + 1]; j++) {
            int64_t q = cur[p.h_col[j]]++;
            row_idx[q] = r;
            p.h_inv_idx[q] = j;
        }
    int rc;
    if ((rc = dev_alloc(p.col_ptr, col_ptr.size() * sizeof(int64_t)))) return rc;
    HIPCHK(hipMemcpy(p.col_ptr.p, col_ptr.data(), col_ptr.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    if ((rc = dev_alloc(p.row_idx, row_idx.size() * sizeof(int32_t)))) return rc;
    if (p.nnz) HIPCHK(hipMemcpy(p.row_idx.p, row_idx.data(), row_idx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    p.w_col.release();
    if (!p.single_w) {
        if ((rc = dev_alloc(p.w_col, p.nnz * net->elem()))) return rc;
        if (net->precision == AHIP_F64) {
            std::vector<double> w(p.nnz);
            for (int64_t q = 0; q < p.nnz; q++) w[q] = p.h_w[p.h_inv_idx[q]];
            if (p.nnz) HIPCHK(hipMemcpy(p.w_col.p, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {
            std::vector<float> w(p.nnz);
            for (int64_t q = 0; q < p.nnz; q++) w[q] = (float)p.h_w[p.h_inv_idx[q]];
            if (p.nnz) HIPCHK(hipMemcpy(p.w_col.p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    return 0;
}

int ahip_proj_init_csr(ahip_net* net, int32_t proj, int32_t n_rows, int32_t n_cols,
                       const int64_t* row_ptr, const int32_t* col_idx, const double* w,
                       double w_const, const ahip_format_opts* opts) {
    Proj* p = get_proj(net, proj);
    if (!p || !row_ptr || (!col_idx && row_ptr[n_rows] > 0)) return fail(AHIP_EINVAL, "bad projection or arrays");
    Pop& pre = *net->pops[p->pre];
    Pop& post = *net->pops[p->post];
    if (n_rows != post.size) return fail(AHIP_EINVAL, "n_rows (%d) must equal the post size (%d)", n_rows, post.size);
    if (n_cols != pre.size) return fail(AHIP_EINVAL, "n_cols (%d) must equal the pre size (%d)", n_cols, pre.size);
    if (row_ptr[0] != 0) return fail(AHIP_EINVAL, "row_ptr[0] must be 0");
    for (int32_t r = 0; r < n_rows; r++)
        if (row_ptr[r + 1] < row_ptr[r]) return fail(AHIP_EINVAL, "row_ptr is not monotonic at row %d", r);
    const int64_t nnz = row_ptr[n_rows];
    for (int64_t j = 0; j < nnz; j++)
        if (col_idx[j] < 0 || col_idx[j] >= n_cols)
            return fail(AHIP_EINVAL, "column index %d out of range at %lld", col_idx[j], (long long)j);
    HIPCHK(hipStreamSynchronize(net->stream));
    p->n_rows = n_rows;
    p->n_cols = n_cols;
    p->nnz = nnz;
    p->h_row_ptr.assign(row_ptr, row_ptr + n_rows + 1);
    p->h_col.assign(col_idx, col_idx + nnz);
    p->single_w = (w == nullptr);
    p->w_const = w_const;
    // new connectivity: no delays until ahip_proj_set_delays
    p->delay_uniform = 1;
    p->delay_per_syn = false;
    p->max_delay = 1;
    p->h_delay.clear();
    p->delay_dev.release();
    p->ring.release();
    p->xd.release();
    if (w)
        p->h_w.assign(w, w + nnz);
    else
        p->h_w.clear();
    if (p->syn && !w) {
        // a plastic projection has one weight per synapse
        p->single_w = false;
        p->h_w.assign(nnz, w_const);
    }
    invalidate(net);
    if (p->spiking) {
        int rc = build_backward(net, *p);
        if (rc || !p->syn) return rc;
        return init_syn_state(net, *p);
    }
    ahip_format_opts o{AHIP_FMT_CSR, 64, 0, 1, 0};
    if (opts) o = *opts;
    if (p->syn && o.format != AHIP_FMT_CSR && o.format != AHIP_FMT_CSR_VEC) {
        // a plastic rate-coded projection updates its forward-order weights in place
        o = ahip_format_opts{AHIP_FMT_CSR, 64, 0, 1, 0};
    }
    std::string err;
    int rc = build_layout(*p, net->precision, o, p->rate, err);
    if (rc) return fail(rc, "%s", err.c_str());
    return p->syn ? init_syn_state(net, *p) : 0;
}

int ahip_proj_generate(ahip_net* net, int32_t proj, const ahip_connect_spec* c, const ahip_format_opts* opts) {
    Proj* p = get_proj(net, proj);
    if (!p || !c) return fail(AHIP_EINVAL, "bad projection or spec");
    Pop& pre = *net->pops[p->pre];
    Pop& post = *net->pops[p->post];
    if (c->pre_begin < 0 || c->pre_end > pre.size || c->pre_begin > c->pre_end || c->post_begin < 0 ||
        c->post_end > post.size || c->post_begin > c->post_end)
        return fail(AHIP_EINVAL, "rank ranges outside the populations");
    if (c->w_kind < 0 || c->w_kind > 2) return fail(AHIP_EINVAL, "bad weight kind");
    const int32_t n_pre = c->pre_end - c->pre_begin;
    const int32_t no_self = (p->pre == p->post && !c->allow_self) ? 1 : 0;
    const int32_t n_rows = post.size;
    std::vector<int64_t> row_ptr(n_rows + 1, 0);
    HIPCHK(hipSetDevice(net->device));
    HIPCHK(hipStreamSynchronize(net->stream));
    DevBuf d_rp, d_counts, d_col, d_w;
    int rc;
    if (c->kind == AHIP_CONN_FIXED_NUMBER_PRE) {
        if (c->number < 0 || c->number > 2016) return fail(AHIP_EINVAL, "number must be 0..2016");
        if (c->number > n_pre - no_self)
            return fail(AHIP_EINVAL, "number (%d) exceeds the %d pre-synaptic candidates", c->number, n_pre - no_self);
        for (int32_t r = 0; r < n_rows; r++)
            row_ptr[r + 1] = row_ptr[r] + ((r >= c->post_begin && r < c->post_end) ? c->number : 0);
    } else if (c->kind == AHIP_CONN_FIXED_PROBABILITY) {
        if (!(c->probability >= 0.0 && c->probability <= 1.0)) return fail(AHIP_EINVAL, "probability outside [0, 1]");
        if ((rc = dev_alloc(d_counts, (size_t)n_rows * sizeof(int64_t)))) return rc;
        LCHK(launch_gen_fixed_probability(c->post_begin, c->post_end, c->pre_begin, n_pre, c->probability, no_self,
                                          c->seed, 0, d_counts.as<int64_t>(), nullptr, nullptr, nullptr, 0, 0.0,
                                          0.0, net->stream));
        std::vector<int64_t> counts(n_rows);
        HIPCHK(hipStreamSynchronize(net->stream));
        HIPCHK(hipMemcpy(counts.data(), d_counts.p, counts.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
        for (int32_t r = 0; r < n_rows; r++) row_ptr[r + 1] = row_ptr[r] + counts[r];
    } else {
        return fail(AHIP_EINVAL, "unknown connector kind");
    }
    const int64_t nnz = row_ptr[n_rows];
    if ((rc = dev_alloc(d_rp, row_ptr.size() * sizeof(int64_t)))) return rc;
    HIPCHK(hipMemcpy(d_rp.p, row_ptr.data(), row_ptr.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    if ((rc = dev_alloc(d_col, (size_t)std::max<int64_t>(nnz, 1) * sizeof(int32_t)))) return rc;
    if (c->w_kind != 0 && (rc = dev_alloc(d_w, (size_t)std::max<int64_t>(nnz, 1) * sizeof(double)))) return rc;
    if (c->kind == AHIP_CONN_FIXED_NUMBER_PRE)
        LCHK(launch_gen_fixed_number_pre(c->post_begin, c->post_end, c->pre_begin, n_pre, c->number, no_self, c->seed,
                                         d_rp.as<int64_t>(), d_col.as<int32_t>(), d_w.as<double>(), c->w_kind, c->w_a,
                                         c->w_b, net->stream));
    else
        LCHK(launch_gen_fixed_probability(c->post_begin, c->post_end, c->pre_begin, n_pre, c->probability, no_self,
                                          c->seed, 1, nullptr, d_rp.as<int64_t>(), d_col.as<int32_t>(),
                                          d_w.as<double>(), c->w_kind, c->w_a, c->w_b, net->stream));
    HIPCHK(hipStreamSynchronize(net->stream));
    std::vector<int32_t> col(nnz);
    std::vector<double> w(c->w_kind != 0 ? nnz : 0);
    if (nnz) HIPCHK(hipMemcpy(col.data(), d_col.p, nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (nnz && c->w_kind != 0) HIPCHK(hipMemcpy(w.data(), d_w.p, nnz * sizeof(double), hipMemcpyDeviceToHost));
    d_col.release();
    d_w.release();
    return ahip_proj_init_csr(net, proj, n_rows, pre.size, row_ptr.data(), col.data(),
                              c->w_kind != 0 ? w.data() : nullptr, c->w_a, opts);
}

int ahip_proj_get_csr(ahip_net* net, int32_t proj, int64_t* row_ptr, int32_t* col, int64_t nnz) {
    Proj* p = get_proj(net, proj);
    if (!p || p->nnz < 0) return fail(AHIP_EINVAL, "bad projection");
    if (row_ptr) std::memcpy(row_ptr, p->h_row_ptr.data(), p->h_row_ptr.size() * sizeof(int64_t));
    if (col) {
        if (nnz < p->nnz) return fail(AHIP_EINVAL, "buffer too small (%lld synapses)", (long long)p->nnz);
        std::memcpy(col, p->h_col.data(), p->nnz * sizeof(int32_t));
    }
    return 0;
}

int ahip_proj_set_format(ahip_net* net, int32_t proj, const ahip_format_opts* opts) {
    Proj* p = get_proj(net, proj);
    if (!p || !opts || p->nnz < 0) return fail(AHIP_EINVAL, "bad projection");
    if (p->spiking) return fail(AHIP_EINVAL, "spiking projections use the CSRC backward view");
    if (p->syn && opts->format != AHIP_FMT_CSR && opts->format != AHIP_FMT_CSR_VEC)
        return fail(AHIP_EINVAL, "a plastic rate-coded projection keeps a csr format");
    if (p->delay_per_syn && opts->format != AHIP_FMT_CSR)
        return fail(AHIP_EINVAL, "per-synapse delays of a rate-coded projection need the csr format");
    HIPCHK(hipStreamSynchronize(net->stream));
    invalidate(net);
    if (p->syn) {
        // the learned weights are on the device: refresh the host copy first
        int rc = syn_field_read(net, *p, -1, p->h_w.data());
        if (rc) return rc;
    }
    std::string err;
    RateLayout L;
    int rc = build_layout(*p, net->precision, *opts, L, err);
    if (rc) return fail(rc, "%s", err.c_str());
    p->rate = std::move(L);
    return 0;
}

int ahip_proj_get_format(ahip_net* net, int32_t proj, ahip_format_opts* opts) {
    Proj* p = get_proj(net, proj);
    if (!p || !opts) return fail(AHIP_EINVAL, "bad projection");
    opts->format = p->rate.format;
    opts->lanes_per_row = p->rate.lanes;
    opts->chunk = p->rate.chunk;
    opts->sigma = p->rate.sigma;
    opts->panel_cols = p->rate.panel_cols;
    return 0;
}

int ahip_proj_set_delays(ahip_net* net, int32_t proj, const int32_t* steps, int64_t n) {
    Proj* p = get_proj(net, proj);
    if (!p || !steps || p->nnz < 0) return fail(AHIP_EINVAL, "bad projection or delays");
    if (n != 1 && n != p->nnz)
        return fail(AHIP_EINVAL, "give one delay or one per synapse (%lld), not %lld", (long long)p->nnz, (long long)n);
    int32_t lo = INT32_MAX, hi = 0;
    for (int64_t j = 0; j < n; j++) {
        lo = std::min(lo, steps[j]);
        hi = std::max(hi, steps[j]);
    }
    if (n > 0 && (lo < 1 || hi > 32767)) return fail(AHIP_EINVAL, "delays must be 1..32767 steps");
    if (p->syn && !(n == 1 || lo == hi)) return fail(AHIP_EINVAL, "a plastic projection takes one uniform delay");
    if (p->syn && (!p->spiking || !p->dep_pre.empty()) && n > 0 && hi > 1)
        return fail(AHIP_EINVAL, "delays are not supported for plastic projections reading pre-synaptic "
                                 "attributes or rate-coded plastic projections");
    HIPCHK(hipStreamSynchronize(net->stream));
    invalidate(net);
    p->delay_dev.release();
    p->h_delay.clear();
    if (n == 1 || lo == hi) {
        // one value for every synapse (core/Projection.py:653-657)
        p->delay_per_syn = false;
        p->delay_uniform = n > 0 ? lo : 1;
        p->max_delay = std::max(p->max_delay, p->delay_uniform);
        p->ring.release();
        return 0;
    }
    if (!p->spiking && p->rate.format != AHIP_FMT_CSR)
        return fail(AHIP_EINVAL, "per-synapse delays of a rate-coded projection need the csr format");
    p->delay_per_syn = true;
    p->delay_uniform = -1;
    p->max_delay = std::max(p->max_delay, hi);
    p->h_delay.assign(steps, steps + n);
    std::vector<int16_t> d16(p->nnz);
    for (int64_t q = 0; q < p->nnz; q++)
        d16[q] = (int16_t)(p->spiking ? steps[p->h_inv_idx[q]] : steps[q]);
    int rc = dev_alloc(p->delay_dev, d16.size() * sizeof(int16_t));
    if (rc) return rc;
    HIPCHK(hipMemcpy(p->delay_dev.p, d16.data(), d16.size() * sizeof(int16_t), hipMemcpyHostToDevice));
    return 0;
}

int ahip_proj_set_max_delay(ahip_net* net, int32_t proj, int32_t max_delay) {
    Proj* p = get_proj(net, proj);
    if (!p || max_delay < 1 || max_delay > 32767) return fail(AHIP_EINVAL, "bad projection or max_delay");
    if (max_delay > p->max_delay) {
        HIPCHK(hipStreamSynchronize(net->stream));
        p->max_delay = max_delay;
        invalidate(net);
    }
    return 0;
}

int ahip_proj_get_delays(ahip_net* net, int32_t proj, int32_t* uniform, int32_t* max_delay, int32_t* steps,
                         int64_t n) {
    Proj* p = get_proj(net, proj);
    if (!p || !uniform || !max_delay) return fail(AHIP_EINVAL, "bad projection or arguments");
    *uniform = p->delay_per_syn ? -1 : p->delay_uniform;
    *max_delay = p->max_delay;
    if (steps) {
        if (n < p->nnz) return fail(AHIP_EINVAL, "buffer too small (%lld synapses)", (long long)p->nnz);
        for (int64_t j = 0; j < p->nnz; j++) steps[j] = p->delay_per_syn ? p->h_delay[j] : p->delay_uniform;
    }
    return 0;
}

int ahip_pop_reset_delays(ahip_net* net, int32_t pop) {
    Pop* p = get_pop(net, pop);
    if (!p) return fail(AHIP_EINVAL, "bad population");
    HIPCHK(hipSetDevice(net->device));
    int rc = ensure_delay_buffers(net);
    if (rc) return rc;
    if (p->depth > 1) {
        if (p->model->spiking) {
            LCHK(launch_fill_i32(p->hist_count.as<int32_t>(), p->depth, 0, net->stream));
        } else {
            const int32_t rs = p->slot_index("r");
            if (rs >= 0) LCHK(launch_fill_hist(net, p->slots[rs].p, p->hist.p, p->size, p->depth, net->stream));
        }
    }
    // pending events of per-synapse spiking delays from this population
    for (auto& pp : net->projs) {
        Proj& q = *pp;
        if (q.pre == pop && q.spiking && q.delay_per_syn && q.ring.p)
            LCHK(launch_fill_i32(q.ring.as<int32_t>(), (int64_t)(q.ring.bytes / sizeof(int32_t)), 0, net->stream));
    }
    HIPCHK(hipStreamSynchronize(net->stream));
    return 0;
}

int ahip_proj_set_launch(ahip_net* net, int32_t proj, int32_t lanes_per_row) {
    Proj* p = get_proj(net, proj);
    if (!p || p->spiking || p->rate.format != AHIP_FMT_CSR) return fail(AHIP_EINVAL, "launch config applies to CSR rate projections");
    if (lanes_per_row < 1 || lanes_per_row > 64 || (lanes_per_row & (lanes_per_row - 1)))
        return fail(AHIP_EINVAL, "lanes_per_row must be a power of two <= 64");
    p->rate.lanes = lanes_per_row;
    invalidate(net);
    return 0;
}

int ahip_proj_nb_synapses(ahip_net* net, int32_t proj, int64_t* nnz) {
    Proj* p = get_proj(net, proj);
    if (!p || !nnz) return fail(AHIP_EINVAL, "bad projection");
    *nnz = p->nnz < 0 ? 0 : p->nnz;
    return 0;
}

// Per-synapse field of a plastic projection (field -1 = w, k = attribute k,
// -2 = _last_event) between the device and a host array in forward order.
// Spiking: fields of the column-order records, copied in chunks of records;
// rate-coded: forward-order arrays.
static int syn_field_io(ahip_net* net, Proj& p, int field, double* dbl, int64_t* i64, bool write) {
    const size_t es = net->elem();
    HIPCHK(hipStreamSynchronize(net->stream));
    if (!p.spiking) {
        void* dev = field == -1 ? p.rate.val.p : field == -2 ? p.last_event.p : p.syn_attr[field].p;
        const size_t vb = field == -2 ? 8 : es;
        std::vector<char> buf((size_t)p.nnz * vb);
        if (write) {
            for (int64_t q = 0; q < p.nnz; q++) {
                if (field == -2) std::memcpy(&buf[q * 8], &i64[q], 8);
                else if (es == 8) std::memcpy(&buf[q * 8], &dbl[q], 8);
                else { const float f = (float)dbl[q]; std::memcpy(&buf[q * 4], &f, 4); }
            }
            if (p.nnz) HIPCHK(hipMemcpy(dev, buf.data(), buf.size(), hipMemcpyHostToDevice));
        } else {
            if (p.nnz) HIPCHK(hipMemcpy(buf.data(), dev, buf.size(), hipMemcpyDeviceToHost));
            for (int64_t q = 0; q < p.nnz; q++) {
                if (field == -2) std::memcpy(&i64[q], &buf[q * 8], 8);
                else if (es == 8) std::memcpy(&dbl[q], &buf[q * 8], 8);
                else { float f; std::memcpy(&f, &buf[q * 4], 4); dbl[q] = f; }
            }
        }
        return 0;
    }
    const int32_t off = field == -1 ? p.rec_off_w : field == -2 ? p.rec_off_last : p.rec_off[field];
    const int64_t chunk = 1 << 20;
    std::vector<char> buf((size_t)std::min<int64_t>(chunk, std::max<int64_t>(p.nnz, 1)) * p.rec_bytes);
    for (int64_t q0 = 0; q0 < p.nnz; q0 += chunk) {
        const int64_t m = std::min(chunk, p.nnz - q0);
        char* dev = (char*)p.syn_rec.p + (size_t)q0 * p.rec_bytes;
        HIPCHK(hipMemcpy(buf.data(), dev, (size_t)m * p.rec_bytes, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < m; i++) {
            char* r = buf.data() + (size_t)i * p.rec_bytes + off;
            const int64_t j = p.h_inv_idx[q0 + i];
            if (field == -2) {
                if (write) std::memcpy(r, &i64[j], 8); else std::memcpy(&i64[j], r, 8);
            } else if (es == 8) {
                if (write) std::memcpy(r, &dbl[j], 8); else std::memcpy(&dbl[j], r, 8);
            } else {
                float f;
                if (write) { f = (float)dbl[j]; std::memcpy(r, &f, 4); }
                else { std::memcpy(&f, r, 4); dbl[j] = f; }
            }
        }
        if (write) HIPCHK(hipMemcpy(dev, buf.data(), (size_t)m * p.rec_bytes, hipMemcpyHostToDevice));
    }
    return 0;
}

static int syn_field_read(ahip_net* net, Proj& p, int field, double* out) {
    return syn_field_io(net, p, field, out, nullptr, false);
}

// Per-post-neuron / per-projection attribute (n values) between device and host.
static int syn_plain_io(ahip_net* net, void* dev, double* v, int64_t n, bool write) {
    HIPCHK(hipStreamSynchronize(net->stream));
    if (net->precision == AHIP_F64) {
        if (n) HIPCHK(hipMemcpy(write ? dev : (void*)v, write ? (void*)v : dev, n * 8,
                                write ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost));
    } else {
        std::vector<float> f(n);
        if (write) {
            for (int64_t q = 0; q < n; q++) f[q] = (float)v[q];
            if (n) HIPCHK(hipMemcpy(dev, f.data(), n * 4, hipMemcpyHostToDevice));
        } else {
            if (n) HIPCHK(hipMemcpy(f.data(), dev, n * 4, hipMemcpyDeviceToHost));
            for (int64_t q = 0; q < n; q++) v[q] = f[q];
        }
    }
    return 0;
}

int ahip_proj_get_w(ahip_net* net, int32_t proj, double* w, int64_t nnz) {
    Proj* p = get_proj(net, proj);
    if (!p || !w || nnz != p->nnz) return fail(AHIP_EINVAL, "bad projection or size");
    if (p->syn) return syn_field_read(net, *p, -1, w);
    for (int64_t j = 0; j < nnz; j++) w[j] = p->single_w ? p->w_const : p->h_w[j];
    return 0;
}

int ahip_proj_set_w(ahip_net* net, int32_t proj, const double* w, int64_t nnz) {
    Proj* p = get_proj(net, proj);
    if (!p || !w || nnz != p->nnz) return fail(AHIP_EINVAL, "bad projection or size");
    HIPCHK(hipStreamSynchronize(net->stream));
    p->single_w = false;
    p->h_w.assign(w, w + nnz);
    invalidate(net);
    if (p->syn) return syn_field_io(net, *p, -1, const_cast<double*>(w), nullptr, true);
    if (p->spiking) return build_backward(net, *p);
    ahip_format_opts o{p->rate.format, p->rate.lanes, p->rate.chunk, p->rate.sigma, p->rate.panel_cols};
    std::string err;
    RateLayout L;
    int rc = build_layout(*p, net->precision, o, L, err);
    if (rc) return fail(rc, "%s", err.c_str());
    p->rate = std::move(L);
    return 0;
}

int ahip_proj_set_synapse(ahip_net* net, int32_t proj, const char* synapse_lib) {
    Proj* p = get_proj(net, proj);
    if (!p || !synapse_lib) return fail(AHIP_EINVAL, "bad projection or library");
    if (p->syn) return fail(AHIP_ESTATE, "projection %d already has a synapse model", p->id);
    void* dl = dlopen(synapse_lib, RTLD_NOW | RTLD_LOCAL);
    if (!dl) return fail(AHIP_EMODEL, "cannot load synapse library %s: %s", synapse_lib, dlerror());
    auto describe = (const ahip_synapse_desc* (*)(void))dlsym(dl, "ahip_synapse_describe");
    if (!describe) return fail(AHIP_EMODEL, "%s does not export ahip_synapse_describe", synapse_lib);
    const ahip_synapse_desc* d = describe();
    if (!d || d->abi != AHIP_SYNAPSE_ABI)
        return fail(AHIP_EMODEL, "%s: synapse ABI %d, runtime expects %d", synapse_lib, d ? d->abi : -1,
                    AHIP_SYNAPSE_ABI);
    if (d->n_attr < 0 || d->n_attr > AHIP_MAX_SYN_ATTRS || d->n_pre_deps < 0 || d->n_pre_deps > AHIP_MAX_SYN_DEPS ||
        d->n_post_deps < 0 || d->n_post_deps > AHIP_MAX_SYN_DEPS)
        return fail(AHIP_EMODEL, "%s: too many attributes", synapse_lib);
    if (!d->pre_f64 || !d->pre_f32 || !d->post_f64 || !d->post_f32 || !d->update_f64 || !d->update_f32)
        return fail(AHIP_EMODEL, "%s: missing launchers", synapse_lib);
    if ((d->spiking != 0) != p->spiking)
        return fail(AHIP_EINVAL, "%s synapse model on a %s projection", d->spiking ? "a spiking" : "a rate-coded",
                    p->spiking ? "spiking" : "rate-coded");
    Pop& pre = *net->pops[p->pre];
    Pop& post = *net->pops[p->post];
    if (p->spiking && !post.model->spiking)
        return fail(AHIP_EINVAL, "spiking synapse models need a spiking post-synaptic population");
    std::vector<int32_t> dep_pre, dep_post;
    for (int32_t k = 0; k < d->n_pre_deps; k++) {
        const int s = pre.slot_index(d->pre_deps[k]);
        if (s < 0) return fail(AHIP_EINVAL, "the pre-synaptic neuron has no attribute %s", d->pre_deps[k]);
        dep_pre.push_back(s);
    }
    for (int32_t k = 0; k < d->n_post_deps; k++) {
        const int s = post.slot_index(d->post_deps[k]);
        if (s < 0) return fail(AHIP_EINVAL, "the post-synaptic neuron has no attribute %s", d->post_deps[k]);
        dep_post.push_back(s);
    }
    if (p->delay_per_syn || (p->delay_uniform > 1 && (!p->spiking || !dep_pre.empty())))
        return fail(AHIP_EINVAL, "this plastic projection cannot have delays");
    HIPCHK(hipSetDevice(net->device));
    HIPCHK(hipStreamSynchronize(net->stream));
    p->syn_dl = dl;  // stays mapped, as the neuron model libraries
    p->syn_path = synapse_lib;
    p->syn = d;
    p->dep_pre = dep_pre;
    p->dep_post = dep_post;
    invalidate(net);
    if (p->nnz < 0) return 0;
    if (p->single_w) {
        p->single_w = false;
        p->h_w.assign(p->nnz, p->w_const);
    }
    int rc;
    if (p->spiking) {
        if ((rc = build_backward(net, *p))) return rc;
    } else {
        const bool csr = p->rate.format == AHIP_FMT_CSR || p->rate.format == AHIP_FMT_CSR_VEC;
        ahip_format_opts o{csr ? p->rate.format : AHIP_FMT_CSR, csr ? p->rate.lanes : 64, 0, 1, 0};
        std::string err;
        RateLayout L;
        if ((rc = build_layout(*p, net->precision, o, L, err))) return fail(rc, "%s", err.c_str());
        p->rate = std::move(L);
    }
    return init_syn_state(net, *p);
}

static int syn_attr(Proj* p, const char* attr) {
    if (!p || !p->syn || !attr) return -1;
    for (int32_t k = 0; k < p->syn->n_attr; k++)
        if (std::strcmp(p->syn->attr_names[k], attr) == 0) return k;
    return -1;
}

static int64_t syn_attr_size(const Proj& p, int k) {
    const int32_t loc = p.syn->attr_locality[k];
    return loc == AHIP_SYN_LOCAL ? p.nnz : loc == AHIP_SYN_SEMIGLOBAL ? p.n_rows : 1;
}

int ahip_proj_syn_get(ahip_net* net, int32_t proj, const char* attr, double* values, int64_t n) {
    Proj* p = get_proj(net, proj);
    const int k = syn_attr(p, attr);
    if (k < 0 || !values || p->nnz < 0) return fail(AHIP_EINVAL, "no synaptic attribute %s", attr ? attr : "");
    const int64_t size = syn_attr_size(*p, k);
    if (n != size) return fail(AHIP_EINVAL, "%s holds %lld values", attr, (long long)size);
    if (p->syn->attr_locality[k] == AHIP_SYN_LOCAL) return syn_field_read(net, *p, k, values);
    return syn_plain_io(net, p->syn_attr[k].p, values, n, false);
}

int ahip_proj_syn_set(ahip_net* net, int32_t proj, const char* attr, const double* values, int64_t n) {
    Proj* p = get_proj(net, proj);
    const int k = syn_attr(p, attr);
    if (k < 0 || !values || p->nnz < 0) return fail(AHIP_EINVAL, "no synaptic attribute %s", attr ? attr : "");
    const int64_t size = syn_attr_size(*p, k);
    const bool local = p->syn->attr_locality[k] == AHIP_SYN_LOCAL;
    if (n == 1 && size != 1) {
        if (local) {
            std::vector<double> v(size, values[0]);
            return syn_field_io(net, *p, k, v.data(), nullptr, true);
        }
        HIPCHK(hipStreamSynchronize(net->stream));
        LCHK(launch_fill(p->syn_attr[k].p, size, (int32_t)net->elem(), values[0], net->stream));
        HIPCHK(hipStreamSynchronize(net->stream));
        return 0;
    }
    if (n != size) return fail(AHIP_EINVAL, "%s holds %lld values", attr, (long long)size);
    if (local) return syn_field_io(net, *p, k, const_cast<double*>(values), nullptr, true);
    return syn_plain_io(net, p->syn_attr[k].p, const_cast<double*>(values), n, true);
}

// set_<constant>(value) of the reference module (core/Constant.py:98-120):
// every population slot and synaptic attribute that mirrors the constant
int ahip_net_set_constant(ahip_net* net, const char* name, double value, int32_t* n_set) {
    if (!net || !name) return fail(AHIP_EINVAL, "bad network or constant name");
    int32_t count = 0;
    for (auto& pp : net->pops) {
        Pop& p = *pp;
        const int k = p.slot_index(name);
        if (k < 0 || p.model->slot_kind[k] != 4) continue;   // absent, or an attribute hiding the constant
        int rc = ahip_pop_set(net, p.id, name, &value, 1);
        if (rc) return rc;
        count++;
    }
    for (auto& pp : net->projs) {
        Proj& p = *pp;
        if (!p.syn || p.nnz < 0) continue;
        const int k = syn_attr(&p, name);
        if (k < 0 || p.syn->attr_locality[k] != AHIP_SYN_CONSTANT) continue;
        int rc = ahip_proj_syn_set(net, p.id, name, &value, 1);
        if (rc) return rc;
        count++;
    }
    if (n_set) *n_set = count;
    return 0;
}

int ahip_proj_get_last_event(ahip_net* net, int32_t proj, int64_t* values, int64_t n) {
    Proj* p = get_proj(net, proj);
    if (!p || !p->syn || !values || n != p->nnz) return fail(AHIP_EINVAL, "bad plastic projection or size");
    return syn_field_io(net, *p, -2, nullptr, values, false);
}

int ahip_proj_set_transmission(ahip_net* net, int32_t proj, int32_t transmission) {
    Proj* p = get_proj(net, proj);
    if (!p) return fail(AHIP_EINVAL, "bad projection");
    HIPCHK(hipStreamSynchronize(net->stream));
    p->transmission = transmission != 0;
    invalidate(net);
    return 0;
}

int ahip_proj_set_learning(ahip_net* net, int32_t proj, int32_t update, int32_t plasticity, int32_t period,
                           int32_t offset) {
    Proj* p = get_proj(net, proj);
    if (!p || !p->syn) return fail(AHIP_EINVAL, "bad plastic projection");
    if (period < 1 || offset < 0 || offset >= period) return fail(AHIP_EINVAL, "need period >= 1 and 0 <= offset < period");
    HIPCHK(hipStreamSynchronize(net->stream));
    p->syn_update = update != 0;
    p->syn_plasticity = plasticity != 0;
    p->syn_period = period;
    p->syn_offset = offset;
    invalidate(net);  // the flags are launch arguments of the captured step
    return 0;
}

int ahip_proj_psp_bytes(ahip_net* net, int32_t proj, double* bytes) {
    Proj* p = get_proj(net, proj);
    if (!p || !bytes || p->nnz < 0) return fail(AHIP_EINVAL, "bad projection");
    *bytes = psp_bytes(*p, net->precision);
    return 0;
}

int ahip_proj_time_psp(ahip_net* net, int32_t proj, int32_t iters, double* avg_ms) {
    Proj* p = get_proj(net, proj);
    if (!p || !avg_ms || iters <= 0 || p->nnz < 0) return fail(AHIP_EINVAL, "bad projection or iteration count");
    if (p->syn && p->spiking) return fail(AHIP_EINVAL, "a plastic spiking projection: use ahip_proj_time_syn");
    HIPCHK(hipSetDevice(net->device));
    int rc = prepare(net);
    if (rc) return rc;
    std::vector<int32_t> zb;
    assign_accumulate(net, zb);
    Pop& pre = *net->pops[p->pre];
    Pop& post = *net->pops[p->post];
    void* target = post.slots[p->target_slot].p;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // one untimed launch, then `iters` timed ones on the network stream
    auto one = [&]() -> int {
        if (p->spiking) return launch_spike_psp(net, *p, pre, target, net->stream);
        return launch_rate_psp(net, *p, pre.slots[p->pre_slot].p, target, net->stream);
    };
    LCHK(one());
    HIPCHK(hipEventRecord(e0, net->stream));
    for (int32_t i = 0; i < iters; i++) LCHK(one());
    HIPCHK(hipEventRecord(e1, net->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *avg_ms = (double)ms / iters;
    return 0;
}

int ahip_time_kernel(ahip_net* net, int32_t what, int32_t index, int32_t iters, double* avg_ms) {
    if (!net || !avg_ms || iters <= 0) return fail(AHIP_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(net->device));
    int rc = prepare(net);
    if (rc) return rc;
    switch (what) {
        case AHIP_PROBE_PRE_EVENTS:
        case AHIP_PROBE_SYN_UPDATE:
        case AHIP_PROBE_POST_EVENTS: {
            Proj* p = get_proj(net, index);
            if (!p || !p->syn || p->nnz <= 0) return fail(AHIP_EINVAL, "bad plastic projection");
            if ((what != AHIP_PROBE_SYN_UPDATE && !p->spiking) ||
                (what == AHIP_PROBE_SYN_UPDATE && (!p->syn->has_update || !p->syn_update)) ||
                (what == AHIP_PROBE_POST_EVENTS && !p->syn->has_post))
                return fail(AHIP_EINVAL, "the projection launches no such kernel");
            break;
        }
        case AHIP_PROBE_UPDATE:
        case AHIP_PROBE_COMPACT: {
            Pop* p = get_pop(net, index);
            if (!p || (what == AHIP_PROBE_COMPACT && !p->model->spiking)) return fail(AHIP_EINVAL, "bad population");
            break;
        }
        case AHIP_PROBE_SCATTER:
            if (index < 0 || index >= (int32_t)net->scatter_groups.size()) return fail(AHIP_EINVAL, "bad scatter group");
            break;
        case AHIP_PROBE_DELIVER:
            if (index != 0 || net->deliver_waves == 0)
                return fail(AHIP_EINVAL, "no separate ordered delivery launch (index 0): the updates deliver");
            break;
        case AHIP_PROBE_PERSIST: {
            ahip_persist_args pa;
            size_t lds = 0;
            int32_t pid = -1;
            net->probe_kind = AHIP_PROBE_PERSIST;
            const bool yes = index == 0 && persist_plan(net, pa, lds, pid);
            net->probe_kind = -1;
            if (!yes) return fail(AHIP_EINVAL, "the network does not run persistent steps");
            break;
        }
        default:
            return fail(AHIP_EINVAL, "unknown kernel kind %d", what);
    }
    std::vector<int32_t> zb;
    assign_accumulate(net, zb);
    HIPCHK(hipEventCreate(&net->probe_ev[0]));
    HIPCHK(hipEventCreate(&net->probe_ev[1]));
    net->probe_proj = index;
    net->probe_kind = what;
    // `iters` ordinary steps (persistent steps: launches of a graph batch),
    // launched eagerly, with the events around the probed kernel: its time
    // in the context of a real step
    const int32_t per_iter = what == AHIP_PROBE_PERSIST ? auto_batch(net) : 1;
    double total = 0.0;
    for (int32_t i = 0; i < iters && rc == 0; i++) {
        rc = enqueue_batch(net, per_iter, zb);
        if (rc) break;
        hipError_t e = hipStreamSynchronize(net->stream);
        if (e != hipSuccess) { rc = fail(AHIP_EHIP, "step failed: %s", hipGetErrorString(e)); break; }
        net->t += per_iter;
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, net->probe_ev[0], net->probe_ev[1]);
        if (e != hipSuccess) { rc = fail(AHIP_EHIP, "probe events: %s", hipGetErrorString(e)); break; }
        total += ms;
    }
    net->probe_proj = -1;
    net->probe_kind = -1;
    hipEventDestroy(net->probe_ev[0]);
    hipEventDestroy(net->probe_ev[1]);
    net->probe_ev[0] = net->probe_ev[1] = nullptr;
    if (rc) return rc;
    *avg_ms = total / iters;
    return 0;
}

int ahip_net_set_persistent(ahip_net* net, int32_t enabled) {
    if (!net) return fail(AHIP_EINVAL, "null network");
    if (net->persist_enabled != (enabled != 0)) {
        HIPCHK(hipStreamSynchronize(net->stream));
        net->persist_enabled = enabled != 0;
        invalidate(net);
    }
    return 0;
}

int ahip_net_uses_persistent(ahip_net* net, int32_t* yes) {
    if (!net || !yes) return fail(AHIP_EINVAL, "null argument");
    if (!net->initialized) return fail(AHIP_ESTATE, "network not initialized");
    int rc = prepare(net);
    if (rc) return rc;
    ahip_persist_args pa;
    size_t lds = 0;
    int32_t pid = -1;
    *yes = persist_plan(net, pa, lds, pid) ? 1 : 0;
    return 0;
}

int ahip_net_set_spike_delivery(ahip_net* net, int32_t mode) {
    if (!net || (mode != AHIP_DELIVERY_ORDERED && mode != AHIP_DELIVERY_ATOMIC))
        return fail(AHIP_EINVAL, "bad network or delivery mode");
    if (net->delivery != mode) {
        HIPCHK(hipStreamSynchronize(net->stream));
        net->delivery = mode;
        invalidate(net);
    }
    return 0;
}

int ahip_ordered_targets(ahip_net* net, int32_t* n) {
    if (!net || !n) return fail(AHIP_EINVAL, "bad arguments");
    int rc = prepare(net);
    if (rc) return rc;
    *n = (int32_t)net->ordered.size();
    return 0;
}

int ahip_scatter_groups(ahip_net* net, int32_t* n_groups) {
    if (!net || !n_groups) return fail(AHIP_EINVAL, "bad arguments");
    int rc = prepare(net);
    if (rc) return rc;
    *n_groups = (int32_t)net->scatter_groups.size();
    return 0;
}

// ---- monitors ---------------------------------------------------------------

int ahip_monitor_create(ahip_net* net, int32_t pop, const char* attr, int64_t capacity,
                        int32_t* mon_id) {
    Pop* p = get_pop(net, pop);
    if (!p || !attr || !mon_id || capacity <= 0) return fail(AHIP_EINVAL, "bad monitor arguments");
    auto m = std::make_unique<Monitor>();
    m->id = (int32_t)net->mons.size();
    m->pop = pop;
    m->attr = attr;
    m->capacity = capacity;
    m->t_start = net->t;
    int rc;
    if (!strcmp(attr, "spike")) {
        if (!p->model->spiking) return fail(AHIP_EINVAL, "spike recording needs a spiking population");
        m->spikes = true;
        // worst case every neuron every step, bounded to 256 Mi indices
        int64_t cap = std::min<int64_t>((int64_t)p->size * capacity, (int64_t)1 << 28);
        m->idx_capacity = cap;
        if ((rc = dev_alloc(m->data, cap * sizeof(int32_t)))) return rc;
        if ((rc = dev_alloc(m->counts, capacity * sizeof(int32_t)))) return rc;
        if ((rc = dev_alloc(m->fill, 2 * sizeof(int64_t)))) return rc;
    } else {
        m->slot = p->slot_index(attr);
        if (m->slot < 0 || !p->model->slot_local[m->slot])
            return fail(AHIP_EINVAL, "population %d has no local attribute '%s'", pop, attr);
        const int kind = p->model->slot_kind[m->slot];
        m->after_psp = (kind == 2) || (p->model->spiking && !strncmp(attr, "g_", 2));
        if ((rc = dev_alloc(m->data, (size_t)capacity * p->size * net->elem()))) return rc;
    }
    *mon_id = m->id;
    net->mons.push_back(std::move(m));
    invalidate(net);
    return 0;
}

// first step >= t at which (step - phase) mod period == 0
static int64_t first_record_step(int64_t t, int64_t phase, int32_t period) {
    const int64_t r = ((t - phase) % period + period) % period;
    return r == 0 ? t : t + (period - r);
}

// rows a value monitor holds after the steps before `now`
static int64_t recorded_rows(const Monitor& m, int64_t now) {
    int64_t n = m.seg_base;
    if (!m.paused && now > m.t_start) n += (now - 1 - m.t_start) / m.period + 1;
    return std::max<int64_t>(0, std::min<int64_t>(n, m.capacity));
}

int ahip_monitor_steps(ahip_net* net, int32_t mon, int64_t* steps) {
    Monitor* m = get_mon(net, mon);
    if (!m || !steps) return fail(AHIP_EINVAL, "bad monitor");
    if (m->spikes) {
        // t_start is the effective start: rows recorded before the current
        // segment are counted in it
        int64_t s = m->paused ? m->seg_base : net->t - m->t_start;
        *steps = std::max<int64_t>(0, std::min<int64_t>(s, m->capacity));
    } else {
        *steps = recorded_rows(*m, net->t);
    }
    return 0;
}

int ahip_monitor_set_period(ahip_net* net, int32_t mon, int32_t period, int64_t phase) {
    Monitor* m = get_mon(net, mon);
    if (!m || period < 1) return fail(AHIP_EINVAL, "bad monitor or period");
    if (m->spikes && period != 1) return fail(AHIP_EINVAL, "spike monitors record every step");
    HIPCHK(hipStreamSynchronize(net->stream));
    if (!m->spikes) {
        m->seg_base = recorded_rows(*m, net->t);
        m->period = period;
        m->phase = phase;
        m->t_start = first_record_step(net->t, phase, period);
    }
    invalidate(net);
    return 0;
}

int ahip_monitor_pause(ahip_net* net, int32_t mon) {
    Monitor* m = get_mon(net, mon);
    if (!m) return fail(AHIP_EINVAL, "bad monitor");
    if (m->paused) return 0;
    HIPCHK(hipStreamSynchronize(net->stream));
    if (m->spikes) {
        int64_t rows = 0;
        ahip_monitor_steps(net, mon, &rows);
        m->seg_base = rows;
    } else {
        m->seg_base = recorded_rows(*m, net->t);
    }
    m->paused = true;
    invalidate(net);
    return 0;
}

int ahip_monitor_resume(ahip_net* net, int32_t mon) {
    Monitor* m = get_mon(net, mon);
    if (!m) return fail(AHIP_EINVAL, "bad monitor");
    if (!m->paused) return 0;
    HIPCHK(hipStreamSynchronize(net->stream));
    m->paused = false;
    // spikes: the row of step t is t - t_start, so the rows go on from
    // seg_base (the fill counts of row seg_base were left by the last
    // recorded step); values: the next step on the period grid
    m->t_start = m->spikes ? net->t - m->seg_base : first_record_step(net->t, m->phase, m->period);
    invalidate(net);
    return 0;
}

int ahip_monitor_get(ahip_net* net, int32_t mon, double* out, int64_t n) {
    Monitor* m = get_mon(net, mon);
    if (!m || !out || m->spikes) return fail(AHIP_EINVAL, "bad value monitor");
    int64_t steps;
    ahip_monitor_steps(net, mon, &steps);
    const int64_t size = net->pops[m->pop]->size;
    if (n != steps * size) return fail(AHIP_EINVAL, "monitor holds %lld values", (long long)(steps * size));
    HIPCHK(hipStreamSynchronize(net->stream));
    if (n == 0) return 0;
    if (net->precision == AHIP_F64) {
        HIPCHK(hipMemcpy(out, m->data.p, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        to_host_t<float>(m->data.p, out, n);
    }
    return 0;
}

int ahip_monitor_get_spikes(ahip_net* net, int32_t mon, int32_t* counts, int64_t steps,
                            int32_t* idx, int64_t idx_cap, int64_t* total) {
    Monitor* m = get_mon(net, mon);
    if (!m || !m->spikes || !total) return fail(AHIP_EINVAL, "bad spike monitor");
    int64_t have;
    ahip_monitor_steps(net, mon, &have);
    if (steps != have) return fail(AHIP_EINVAL, "monitor holds %lld steps", (long long)have);
    HIPCHK(hipStreamSynchronize(net->stream));
    int64_t fills[2] = {0, 0};
    HIPCHK(hipMemcpy(fills, m->fill.p, 2 * sizeof(int64_t), hipMemcpyDeviceToHost));
    const int64_t fill = have > 0 ? fills[have & 1] : 0;   // indices before recorded step `have`
    *total = fill;
    if (fill > m->idx_capacity) return fail(AHIP_ESTATE, "spike monitor overflowed its buffer");
    if (counts && steps > 0) HIPCHK(hipMemcpy(counts, m->counts.p, steps * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (idx) {
        if (fill > idx_cap) return fail(AHIP_EINVAL, "index buffer too small (%lld)", (long long)fill);
        if (fill > 0) HIPCHK(hipMemcpy(idx, m->data.p, fill * sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    return 0;
}

int ahip_monitor_clear(ahip_net* net, int32_t mon) {
    Monitor* m = get_mon(net, mon);
    if (!m) return fail(AHIP_EINVAL, "bad monitor");
    HIPCHK(hipStreamSynchronize(net->stream));
    if (m->spikes) {
        m->t_start = net->t;
        m->seg_base = 0;
        LCHK(launch_fill_i64(m->fill.as<int64_t>(), 2, 0, net->stream));
        HIPCHK(hipStreamSynchronize(net->stream));
    } else {
        m->seg_base = 0;
        m->t_start = first_record_step(net->t, m->phase, m->period);
    }
    invalidate(net);
    return 0;
}

}  // extern "C"
