/* annarchy_hip_synapse.h -- plug-in ABI of generated spiking-synapse models.
 *
 * A synapse model with pre_spike / post_spike statements and event-driven
 * variables (the plasticity code of Projection/OpenMPGenerator.py:
 * 1057-1230 and :1482-1815) is compiled into a shared library that exports
 * ahip_synapse_describe().  The runtime owns the per-synapse state (column
 * order of the CSRC view) and launches the two event kernels:
 *   pre  -- compute_psp of step t, for the spikes of step t-d of the pre
 *           population (d = the uniform delay, 1 without delay): event-driven
 *           decay to t-1, _last_event = t, g_target += w, the pre_spike
 *           statements (Projection/OpenMP/CSR.py:870-900);
 *   post -- after the neuron updates of step t, for the spikes of step t of
 *           the post population: event-driven decay to t, _last_event = t,
 *           the post_spike statements (Projection/OpenMP/CSR.py:903-918).
 * Every event of one synapse is handled by one lane, so the state updates
 * are race-free; the g_target additions are device atomics.
 */
#ifndef ANNARCHY_HIP_SYNAPSE_H
#define ANNARCHY_HIP_SYNAPSE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AHIP_SYNAPSE_ABI 1
#define AHIP_MAX_SYN_ATTRS 16

typedef struct ahip_syn_args {
    /* pre events: spike list of step t-d of the pre population (d = the
     * projection's uniform delay in steps, 1 without delay).  pre_depth 1:
     * the list itself; else the ring of pre_depth lists of pre_stride
     * entries, the list of step u in slot u mod pre_depth. */
    const int32_t* pre_spiked;
    const int32_t* pre_count;
    int64_t pre_stride;
    int32_t pre_depth;
    int32_t pre_delay;
    int32_t n_cols;
    const int64_t* col_ptr;   /* CSRC backward view */
    const int32_t* row_idx;
    /* post events: spike list of step t of the post population */
    const int32_t* post_spiked;
    const int32_t* post_count;
    int32_t n_rows;
    const int64_t* row_ptr;   /* forward view: row pointer, pre rank and column position */
    const int32_t* col_idx;   /*   of every forward synapse */
    const int64_t* fwd2col;
    /* step of the last spike of every pre / post neuron (t_pre, t_post) */
    const int64_t* pre_last_spike;
    const int64_t* post_last_spike;
    void* g;                  /* g_<target> of the post population (network precision) */
    void* w;                  /* weights, column order */
    int64_t* last_event;      /* column order, -10000 initially */
    void* local[AHIP_MAX_SYN_ATTRS];      /* per-synapse attribute k (column order), else NULL */
    double global[AHIP_MAX_SYN_ATTRS];    /* per-projection attribute k */
    const int64_t* t_base;    /* simulation step at the start of the batch */
    int32_t step;
    double dt;
} ahip_syn_args;

typedef int (*ahip_syn_launch_fn)(const ahip_syn_args* a, void* stream);

typedef struct ahip_synapse_desc {
    int32_t abi;
    const char* name;
    int32_t n_attr;                   /* attributes besides w */
    const char* const* attr_names;
    const int32_t* attr_local;        /* 1: per synapse, 0: per projection */
    const double* attr_init;
    int32_t has_pre, has_post, pre_g; /* pre_g: pre events add w to g_target */
    ahip_syn_launch_fn pre_f64, pre_f32, post_f64, post_f32;
    const char* source_hash;
} ahip_synapse_desc;

const ahip_synapse_desc* ahip_synapse_describe(void);

#ifdef __cplusplus
}
#endif

#endif /* ANNARCHY_HIP_SYNAPSE_H */
