/* annarchy_hip_synapse.h -- plug-in ABI of generated synapse models.
 *
 * A synapse model (the plasticity code of Projection/OpenMPGenerator.py:
 * 1057-1230, :1482-1815) is compiled into a shared library that exports
 * ahip_synapse_describe().  The runtime owns the per-synapse state and
 * launches three kinds of kernels:
 *   pre    -- spiking: compute_psp of step t, for the spikes of step t-d of
 *             the pre population (d = the uniform delay, 1 without delay):
 *             event-driven decay to t-1, _last_event = t, g_target += <expr>
 *             (w by default), the pre_spike statements
 *             (Projection/OpenMP/CSR.py:870-900);
 *   update -- update_synapse after the neuron updates of step t: the
 *             per-projection equations, then per post neuron its
 *             per-post-neuron equations and the per-synapse equations of
 *             its row (Projection/OpenMP/CSR.py:833-865), every
 *             update_period steps (with dt scaled by the period);
 *   post   -- spiking: after the update, for the spikes of step t of the
 *             post population: event-driven decay to t, _last_event = t,
 *             the post_spike statements (Projection/OpenMP/CSR.py:903-918).
 * Every event or update of one synapse is handled by one lane, so the state
 * updates are race-free; the g_target additions are device atomics.
 */
#ifndef ANNARCHY_HIP_SYNAPSE_H
#define ANNARCHY_HIP_SYNAPSE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AHIP_SYNAPSE_ABI 2
#define AHIP_MAX_SYN_ATTRS 16
#define AHIP_MAX_SYN_DEPS 8

/* attribute locality */
#define AHIP_SYN_LOCAL 0       /* one value per synapse (state order) */
#define AHIP_SYN_SEMIGLOBAL 1  /* one value per post-synaptic neuron (": postsynaptic") */
#define AHIP_SYN_GLOBAL 2      /* one value per projection (": projection") */
#define AHIP_SYN_CONSTANT 3    /* a network constant read by the model: one value per
                                  projection, written by ahip_net_set_constant */

typedef struct ahip_syn_args {
    /* pre events: spike list of step t-d of the pre population.  pre_depth
     * 1: the list itself; else the ring of pre_depth lists of pre_stride
     * entries, the list of step u in slot u mod pre_depth. */
    const int32_t* pre_spiked;
    const int32_t* pre_count;
    int64_t pre_stride;
    int32_t pre_depth;
    int32_t pre_delay;
    int32_t n_cols;
    const int64_t* col_ptr;   /* CSRC backward view (spiking) */
    const int32_t* row_idx;
    /* post events: spike list of step t of the post population */
    const int32_t* post_spiked;
    const int32_t* post_count;
    int32_t n_rows;
    const int64_t* row_ptr;   /* forward view: row pointer and pre rank of */
    const int32_t* col_idx;   /*   every forward synapse */
    const int64_t* fwd2col;   /* state position of forward synapse j; NULL: j itself */
    /* step of the last spike of every pre / post neuron (t_pre, t_post) */
    const int64_t* pre_last_spike;
    const int64_t* post_last_spike;
    void* g;                  /* g_<target> of the post population (spiking) */
    /* Per-synapse values: value of synapse q at ptr[q * state_stride]
     * (state_stride8 for _last_event, in int64 units).  A spiking projection
     * keeps one record per synapse (w, the per-synapse attributes,
     * _last_event), so these pointers point into the same records; a
     * rate-coded one keeps separate arrays (strides 1). */
    void* w;                  /* weights (network precision) */
    int64_t* last_event;      /* -10000 initially */
    int32_t state_stride;     /* in values of the network precision */
    int32_t state_stride8;    /* in int64 */
    void* attr[AHIP_MAX_SYN_ATTRS];   /* attribute k: per synapse as above, n_rows or 1 values */
    /* neuron attributes read as pre.<name> / post.<name>: element i of
     * neuron n is ptr[n * stride] (stride 0 for a population-wide value) */
    const void* pre_attr[AHIP_MAX_SYN_DEPS];
    int32_t pre_attr_stride[AHIP_MAX_SYN_DEPS];
    const void* post_attr[AHIP_MAX_SYN_DEPS];
    int32_t post_attr_stride[AHIP_MAX_SYN_DEPS];
    const int64_t* t_base;    /* simulation step at the start of the batch */
    int32_t step;
    double dt;
    int32_t pre_chunks;       /* waves per pre spike (64 synapses each, strided) */
    int32_t post_chunks;      /* waves per post spike */
    int32_t plasticity;       /* 0: w is left unchanged (_plasticity) */
    int32_t update_period;    /* update every period steps ... */
    int32_t update_offset;    /* ... when (t - offset) mod period == 0 */
} ahip_syn_args;

typedef int (*ahip_syn_launch_fn)(const ahip_syn_args* a, void* stream);

typedef struct ahip_synapse_desc {
    int32_t abi;
    const char* name;
    int32_t spiking;                  /* 1: pre_spike / post_spike model of a spiking projection */
    int32_t n_attr;                   /* attributes besides w */
    const char* const* attr_names;
    const int32_t* attr_locality;     /* AHIP_SYN_LOCAL / _SEMIGLOBAL / _GLOBAL */
    const double* attr_init;
    int32_t n_pre_deps;               /* pre.<name> attributes read */
    const char* const* pre_deps;
    int32_t n_post_deps;              /* post.<name> attributes read */
    const char* const* post_deps;
    int32_t has_pre, has_post, has_update;
    ahip_syn_launch_fn pre_f64, pre_f32, post_f64, post_f32, update_f64, update_f32;
    const char* source_hash;
} ahip_synapse_desc;

const ahip_synapse_desc* ahip_synapse_describe(void);

#ifdef __cplusplus
}
#endif

#endif /* ANNARCHY_HIP_SYNAPSE_H */
