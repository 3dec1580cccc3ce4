/* annarchy_hip_model.h -- plug-in ABI between the simulation runtime
 * (libannarchy_hip.so) and a generated neuron-model library.
 *
 * A model library is produced from a neuron description: the equation
 * snippets that the sympy front end emits are spliced into a HIP kernel
 * template (one thread per neuron, fused ODE + bounds + threshold + reset +
 * refractory + spike flags).  This mirrors the reference's per-population
 * generated code: PopStruct::update()/spike_gather()
 * (generator/Population/OpenMPGenerator.py:765-1237,
 *  generator/Population/CUDATemplates.py:543-720), which the reference emits
 * as text per network; here it is one shared object per model that the
 * runtime loads with dlopen() and launches from inside its captured step.
 *
 * Plain pointers and sizes only; the stream is passed as an opaque pointer.
 */
#ifndef ANNARCHY_HIP_MODEL_H
#define ANNARCHY_HIP_MODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AHIP_MODEL_ABI 10
#define AHIP_MAX_SLOTS 48
#define AHIP_MAX_RAND 8
#define AHIP_UPDATE_BLOCK 256 /* threads per block of every update kernel */

/* Ordered spike delivery folded into the update of the post population.
 * head[i] = (base << AHIP_HEAD_SHIFT) | count: neuron i's events of this
 * step are bucket records base .. base + count (head[i + 1] >> SHIFT bounds
 * the range).  A record is a key (the projection's key offset + the
 * synapse's column-order index, so keys sort projection by projection and in
 * ascending pre-synaptic rank: the reference's summation order) and the
 * event's weight: uint32 key, uint32 padding, double weight in FP64 (16
 * bytes); uint32 key, float weight in FP32 (8 bytes).  The weights are added
 * to g[i] in ascending key before the equations read g; the count is
 * cleared. */
typedef struct ahip_deliver_desc {
    void* g;
    uint64_t* head;                     /* size + 1 words */
    const void* bucket;                 /* records */
    const void* reserved;
} ahip_deliver_desc;
#define AHIP_HEAD_SHIFT 24
#define AHIP_HEAD_COUNT ((1ull << AHIP_HEAD_SHIFT) - 1ull)
#define AHIP_MAX_DELIVER 4 /* ordered conductances one update delivers */
#define AHIP_MAX_PSUM 2    /* psp sums one update completes from panel partials */

/* Arguments of one population update launch.  Device pointers unless noted. */
typedef struct ahip_pop_args {
    int32_t size;                       /* number of neurons */
    int32_t step;                       /* step index inside the launched batch */
    const int64_t* t_base;              /* simulation step at the start of the batch */
    double dt;                          /* host value */
    void* slot[AHIP_MAX_SLOTS];         /* local: size elements, global: 1 element */
    const void* rand_inject[AHIP_MAX_RAND]; /* injected draws [steps][size] or NULL */
    int64_t rand_inject_steps;          /* rows available in the injected buffers */
    int64_t rand_t0;                    /* simulation step of row 0 of the injected buffers */
    uint64_t seed;                      /* counter-based generator key */
    uint32_t stream_id;                 /* population id, second key word */
    int32_t* refractory;                /* steps, per neuron */
    int32_t* refractory_remaining;
    int64_t* last_spike;
    uint64_t* spike_mask;               /* ceil(size/64) words, bit i = neuron spiked */
    int32_t* block_counts;              /* ceil(size/256) spike counts per update block */
    const ahip_deliver_desc* deliver;   /* ordered conductances delivered by this update, or NULL */
    int32_t n_deliver;
    /* psp sums the update completes: slot psum_slot[k] (a _sum_<target>) is
     * the sum over the panels q < psum_panels[k] of psum_part[k][q *
     * psum_stride[k] + i], added in ascending q (compensated in FP32), plus
     * the slot's own value when psum_acc[k]; the update stores it to the
     * slot before its statements read it */
    int32_t n_psum;
    int32_t psum_slot[AHIP_MAX_PSUM];
    int32_t psum_panels[AHIP_MAX_PSUM];
    int32_t psum_acc[AHIP_MAX_PSUM];
    int32_t psum_stride[AHIP_MAX_PSUM];
    const void* psum_part[AHIP_MAX_PSUM];
} ahip_pop_args;

typedef int (*ahip_launch_fn)(const ahip_pop_args* args, void* stream);

/* ---- persistent steps of a small network --------------------------------
 * A network of one spiking population whose projections all start and end
 * in it (constant weight each, no delays, no synapse model) can run a whole
 * batch of steps in ONE launch of one workgroup: per step the scatter of the
 * previous step's spikes into the conductances, the update of every neuron
 * and the ascending spike list with its recording.  The state the
 * statements touch (the model's persist_slots) and the refractory counters
 * live in LDS for the whole launch.  The steps are the ordinary ones (the
 * same statements, the same order), so spike trains and state equal those of
 * the per-kernel step. */
#define AHIP_PERSIST_MAX_PROJ 8
#define AHIP_PERSIST_THREADS 1024
#define AHIP_PERSIST_ITEMS 512      /* (projection, spike) items per scatter pass */
typedef struct ahip_persist_proj {
    const int64_t* col_ptr;             /* CSRC backward view (pre rank -> synapse range) */
    const int32_t* row_idx;             /* post rank of every synapse, column order */
    double w;                           /* the projection's one weight */
    int32_t n_cols;                     /* pre ranks covered by col_ptr */
    int32_t target;                     /* index of the conductance in the model's persist_slots */
    int32_t round;                      /* scatter round: same-target projections in creation order */
    int32_t lds_cols;                   /* offset of its column ranges in the LDS table (-1: col_ptr) */
} ahip_persist_proj;
typedef struct ahip_persist_args {
    int32_t steps;                      /* steps of this launch */
    int32_t n_proj, n_rounds;
    int32_t n_lds_cols;                 /* entries of the LDS column-range table (0: none) */
    ahip_persist_proj proj[AHIP_PERSIST_MAX_PROJ];
    int32_t* spiked;                    /* the population's spike list and count (kept up to date) */
    int32_t* spike_count;
    /* the population's spike monitor (rec_counts == NULL: none), as filled
     * by the compaction of the per-kernel step */
    int32_t* rec_counts;
    int32_t* rec_idx;
    int64_t* rec_fill;
    int64_t rec_idx_capacity;
    int64_t rec_t_start;
    int64_t rec_capacity;
} ahip_persist_args;
typedef int (*ahip_persist_fn)(const ahip_pop_args* args, const ahip_persist_args* pa, size_t lds_bytes,
                               void* stream);
/* Host-side evaluation of the hoisted step sizes (exponential Euler
 * "pre_loop" values, parser/Equation.py:559-604) from the global slots; writes
 * one value per pre-loop slot.  Host libm is used on purpose so the values
 * equal those of a CPU build. */
typedef void (*ahip_preloop_fn)(const double* globals_host, double dt, double* out);

typedef struct ahip_model_desc {
    int32_t abi;                        /* AHIP_MODEL_ABI */
    const char* name;
    int32_t spiking;                    /* 1: spiking (threshold/reset), 0: rate-coded */
    int32_t n_slots;
    const char* const* slot_names;      /* attributes, _sum_<target>, pre-loop step sizes */
    const int32_t* slot_local;          /* 1: one value per neuron, 0: population-wide */
    const double* slot_init;            /* initial value of each slot */
    const int32_t* slot_kind;           /* 0 parameter, 1 variable, 2 psp sum, 3 pre-loop,
                                           4 network constant (population-wide, written by
                                           ahip_net_set_constant) */
    int32_t n_rand;                     /* random variables drawn per step */
    const int32_t* rand_dist;           /* 0 Uniform(a,b), 1 Normal(mu,sigma) */
    const double* rand_args;            /* two per random variable */
    int32_t has_refractory;
    int32_t has_global_update;          /* a one-thread kernel runs before the local one */
    ahip_launch_fn launch_f64;
    ahip_launch_fn launch_f32;
    ahip_preloop_fn preloop;            /* NULL when there is no pre-loop value */
    const char* source_hash;
    ahip_persist_fn persist_f64;        /* NULL for rate-coded models or with a global update */
    ahip_persist_fn persist_f32;
    int32_t persist_n_slots;            /* local slots the persistent steps hold in LDS, in LDS order: */
    const int32_t* persist_slots;       /* those the statements touch and the conductances */
} ahip_model_desc;

/* Every model library exports this symbol. */
typedef const ahip_model_desc* (*ahip_model_describe_fn)(void);

#ifdef __cplusplus
}
#endif

#endif
