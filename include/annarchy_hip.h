/* annarchy_hip.h -- C-ABI of the MI355X simulation backend (libannarchy_hip.so).
 *
 * This is the drop-in boundary for ANNarchy's simulation step.  In the
 * reference, Python drives a generated nanobind module ANNarchyWrapper<id>
 * (generator/NanoBind/BaseTemplate.py:6-45) loaded by
 * generator/Compiler.py:772-793 and instantiated by Compiler.py:796-1000; the
 * entry points below replace the module functions and the per-object
 * wrappers of that module one for one (each comment names the reference
 * interface).  Conventions:
 *   - every function returns 0 on success or a negative AHIP_E* code; the
 *     message of the last failure is available from ahip_last_error()
 *     (the reference returns bool and prints to std::cerr, Messages.py:94-112);
 *   - host buffers are owned by the caller and only read/written during the
 *     call; device memory is owned by the network;
 *   - a network is not thread-safe; different networks may be used from
 *     different threads.
 */
#ifndef ANNARCHY_HIP_H
#define ANNARCHY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AHIP_OK 0
#define AHIP_EINVAL -1   /* bad argument (index, size, name) */
#define AHIP_EHIP -2     /* HIP runtime error */
#define AHIP_ENOMEM -3   /* device allocation failed */
#define AHIP_ESTATE -4   /* call not valid in the current state */
#define AHIP_EMODEL -5   /* model library could not be loaded */

/* Storage formats of a rate-coded projection (core/Projection.py:691-789
 * chooses among csr/ell/ellr/sell/bsr/dense in the reference). */
#define AHIP_FMT_CSR 0    /* row-wise, L lanes per row (generator/Projection/CUDA/CSR.py:299-334) */
#define AHIP_FMT_ELL 1    /* column-major ELLPACK (CUDA/ELL.py:240-262) */
#define AHIP_FMT_ELLR 2   /* ELLPACK-R, explicit row lengths (CUDA/ELLR.py:239-255) */
#define AHIP_FMT_SELL 3   /* SELL-C-sigma with row sorting inside sigma windows (CUDA/SELL.py:230-246 has no sigma) */
#define AHIP_FMT_BSR 4    /* block sparse rows, square tiles (CUDA/BSR.py:229-252) */
#define AHIP_FMT_DENSE 5  /* row-major dense (CUDA/Dense.py:287-316) */
#define AHIP_FMT_CSR_PANEL 6 /* row-block x column-panel tiles, pre-synaptic panel staged in LDS */
#define AHIP_FMT_CSR_VEC 7   /* CSR read with 16-byte loads, 4 synapses per lane per chunk */
#define AHIP_FMT_SELL_PANEL 8 /* CSR_PANEL tiles stored as sorted 64-row slices, uint16 panel columns */

/* Precision of a network (core/CTypes.py:30-74, ConfigManagement.py:25). */
#define AHIP_F64 0
#define AHIP_F32 1        /* FP32 state and weights, Kahan-compensated psp sums */

typedef struct ahip_net ahip_net;

typedef struct ahip_format_opts {
    int32_t format;           /* AHIP_FMT_* */
    int32_t lanes_per_row;    /* CSR: 1..64 (power of two); SELL: threads per row; CSR_PANEL: 4..32;
                                 SELL_PANEL: 4-synapse steps in flight per lane (1, 2, 4, 8) */
    int32_t chunk;            /* SELL: rows per slice (C); BSR: tile size; CSR_PANEL / SELL_PANEL: rows per
                                 tile (SELL_PANEL 0 = the tallest of 2048 ... 256 rows that still
                                 makes at least 256 tiles) */
    int32_t sigma;            /* SELL: sorting window in rows (1 = no sorting); SELL_PANEL: waves per
                                 workgroup (4 or 8; 0 or 1 = 4) */
    int32_t panel_cols;       /* CSR_PANEL / SELL_PANEL: at most this many columns per panel (0 =
                                 automatic: 20479 for SELL_PANEL); the panels are then made equally wide.  A panel
                                 is staged in LDS: at most 64 KiB (CSR_PANEL) or 160 KiB
                                 (SELL_PANEL, one workgroup per CU at that size) */
} ahip_format_opts;

/* ---- errors / device ------------------------------------------------------ */
const char* ahip_last_error(void);
/* set_device(int) of the CUDA wrapper (NanoBind/Generator.py:41-46). */
int ahip_set_device(int32_t device);
int ahip_device_count(int32_t* count);
/* Library version string and the gfx target the kernels were built for. */
const char* ahip_build_info(void);
/* Measurement helper with no counterpart in the reference: times `iters`
 * launches of a streaming copy of `bytes` bytes (16-byte non-temporal loads
 * and stores) with HIP events; *avg_ms = mean launch time.  Used to calibrate
 * HBM byte counters and to state the achievable bandwidth. */
int ahip_probe_copy(int64_t bytes, int32_t iters, double* avg_ms);

/* ---- network (module functions, NanoBind/BaseTemplate.py:17-28) ---------- */
/* pyx_create() */
int ahip_net_create(ahip_net** out);
int ahip_net_destroy(ahip_net* net);
/* pyx_initialize(dt) + dtype: FP64 or FP32 for the whole network. */
int ahip_net_initialize(ahip_net* net, double dt, int32_t precision);
/* set_seed(seed, num_sources, use_seed_seq): key of the on-device generator. */
int ahip_set_seed(ahip_net* net, uint64_t seed);
/* set_<constant>(value) of the generated module (core/Constant.py:98-120,
 * Network.py:482-512): writes `value` into every population slot of kind 4
 * and every synaptic attribute of locality AHIP_SYN_CONSTANT named `name`
 * (the models that read the constant and do not hide it with an attribute
 * of their own); *n_set (may be NULL) = the number of objects written. */
int ahip_net_set_constant(ahip_net* net, const char* name, double value, int32_t* n_set);
/* run(nbSteps): whole steps are captured in a HIP graph and replayed. */
int ahip_run(ahip_net* net, int64_t nsteps);
/* step() */
int ahip_step(ahip_net* net);
/* run_until(steps, populations, or_and) (Simulate.py:173-266,
 * OpenMPBaseTemplate.py:485-517): without a stop condition anywhere in the
 * network, run(steps); otherwise step by step until the stop conditions of
 * the listed populations, folded with && (and_or = 1) or || (and_or = 0),
 * hold after a step.  *done = steps simulated. */
int ahip_run_until(ahip_net* net, int64_t max_steps, const int32_t* pops, int32_t npops,
                   int32_t and_or, int64_t* done);
/* stop_condition of a population (Population/OpenMPGenerator.py:488-555):
 * mode 0 none, 1 / 2 the condition compiled into its model library
 * (ahip_model_stop_f64/_f32) holds for any / all neurons, 3 / 4
 * "spiked : any" / "spiked : all". */
int ahip_pop_set_stop_condition(ahip_net* net, int32_t pop, int32_t mode);
/* get_time() / set_time(t) */
int ahip_get_time(ahip_net* net, int64_t* t);
int ahip_set_time(ahip_net* net, int64_t t);
/* Number of steps captured per graph launch (0 = automatic). */
int ahip_set_graph_batch(ahip_net* net, int32_t steps);
/* Disable (0) or enable (1) graph capture; eager launches otherwise. */
int ahip_set_graph_enabled(ahip_net* net, int32_t enabled);
/* Captures, instantiates and uploads the graph of one batch of steps now
 * (run() would do so at its first full batch), so that a timed run only
 * replays it.  Simulates nothing.  No counterpart in the reference. */
int ahip_graph_prepare(ahip_net* net);

/* ---- populations (pop<id>_wrapper, NanoBind/Population.py:6-72) --------- */
/* pop<id>_wrapper(size, max_delay): model_lib is the path of a generated
 * neuron-model library (annarchy_hip_model.h). */
int ahip_pop_create(ahip_net* net, int32_t size, const char* model_lib, int32_t* pop_id);
/* Attribute property get/set (vector<double> copies in the reference,
 * NanoBind/Population.py:27-54).  n = size for local attributes, 1 for
 * population-wide ones.  Values are converted to the network precision. */
int ahip_pop_set(ahip_net* net, int32_t pop, const char* attr, const double* values, int64_t n);
int ahip_pop_get(ahip_net* net, int32_t pop, const char* attr, double* values, int64_t n);
/* refractory (rw int vector, in steps) and refractory_remaining */
int ahip_pop_set_int(ahip_net* net, int32_t pop, const char* attr, const int32_t* values, int64_t n);
int ahip_pop_get_int(ahip_net* net, int32_t pop, const char* attr, int32_t* values, int64_t n);
/* last_spike (int64 per neuron, -10000 initially, OpenMPTemplates.py:283-287) */
int ahip_pop_get_last_spike(ahip_net* net, int32_t pop, int64_t* values, int64_t n);
/* Spikes emitted in the last step, ascending (the `spiked` vector). */
int ahip_pop_spikes(ahip_net* net, int32_t pop, int32_t* out, int32_t cap, int32_t* n);
/* reset(): refractory_remaining = 0, last_spike = -10000, spiked cleared. */
int ahip_pop_reset_spiking(ahip_net* net, int32_t pop);
/* activate(bool) (OpenMPTemplates.py:55, core/Population.py:355-372
 * enable/disable): an inactive population is neither updated nor gathered
 * (its last spike list stays as it is), and the projections into it do not
 * transmit. */
int ahip_pop_set_active(ahip_net* net, int32_t pop, int32_t active);
/* size_in_bytes() (NanoBind/Population.py, OpenMPTemplates.py): here the
 * device memory the population holds (attributes, spike structures, delay
 * ring, firing-rate history, injected draws). */
int ahip_pop_size_in_bytes(ahip_net* net, int32_t pop, int64_t* bytes);
/* size_in_bytes() of a projection (CSRCMatrix.hpp size_in_bytes): the device
 * memory of its psp layout, backward view, delays and synaptic state. */
int ahip_proj_size_in_bytes(ahip_net* net, int32_t proj, int64_t* bytes);
/* compute_firing_rate(window) (core/Population.py:700-723,
 * Population/OpenMPGenerator.py:562-618): after every step r = 1000/window x
 * the neuron's spikes in the last window/dt steps (spiking populations; r is
 * 0 until enabled; window <= 0 leaves the state unchanged). */
int ahip_pop_compute_firing_rate(ahip_net* net, int32_t pop, double window_ms);
/* reset() of the delay queues (Population/OpenMPTemplates.py:193-197, the
 * "reset" code of attribute_delayed; Projection reset_ring_buffer,
 * core/Network.py:652-655): every slot of a rate-coded population's queue
 * takes the current rates, a spiking population's queue of spike lists is
 * emptied, and so are the pending events of per-synapse delays from it.  A
 * queue is zero (no spikes) until then, as after the reference's init. */
int ahip_pop_reset_delays(ahip_net* net, int32_t pop);
int ahip_pop_size(ahip_net* net, int32_t pop, int32_t* size);
/* Injected random streams: buf holds nsteps*size draws of random variable k
 * (rand_k), row 0 used at the current time step; replaces the on-device
 * generator so a run can consume the same stream as a CPU run.  buf == NULL
 * returns to the on-device generator. */
int ahip_pop_inject_random(ahip_net* net, int32_t pop, int32_t k, const double* buf,
                           int64_t nsteps);

/* ---- projections (proj<id>_wrapper, NanoBind/Projection.py:6-95) -------- */
/* proj<id>_wrapper(): pre/post population ids and the target name (exc,
 * inh, ...); spiking = 1 makes an event-driven projection into g_<target>. */
int ahip_proj_create(ahip_net* net, int32_t pre, int32_t post, const char* target,
                     int32_t spiking, int32_t* proj_id);
/* init_from_lil(post_ranks, pre_ranks, w, delays, requires_sorting)
 * (ProjectionGenerator.py:743-797) as a bulk CSR: rows are post ranks
 * (n_rows = post size, empty rows allowed), columns ascending inside a row.
 * w: nnz weights in row order, or NULL for one constant weight w_const
 * (core/Projection.py:823-839). */
int ahip_proj_init_csr(ahip_net* net, int32_t proj, int32_t n_rows, int32_t n_cols,
                       const int64_t* row_ptr, const int32_t* col_idx, const double* w,
                       double w_const, const ahip_format_opts* opts);
/* Connectivity drawn on the device (no counterpart call in the reference,
 * whose connectors draw on the host, cython_ext/Connector.pyx:306-380; same
 * distributions, a counter-based random stream instead of numpy's).
 * fixed_number_pre: `number` distinct pre ranks per post neuron (at most
 * 2016); fixed_probability: each pre rank with probability p.  Pre ranks
 * [pre_begin, pre_end) of the pre population connect to post ranks
 * [post_begin, post_end); when pre and post are the same population and
 * allow_self is 0, rank i does not connect to itself.  Weights: w_kind 0 =
 * the constant w_a, 1 = Uniform(w_a, w_b), 2 = Normal(w_a, w_b).  The result
 * replaces the projection's connectivity as ahip_proj_init_csr would. */
#define AHIP_CONN_FIXED_NUMBER_PRE 0
#define AHIP_CONN_FIXED_PROBABILITY 1
typedef struct ahip_connect_spec {
    int32_t kind;
    int32_t number;
    double probability;
    int32_t pre_begin, pre_end;
    int32_t post_begin, post_end;
    int32_t allow_self;
    int32_t w_kind;
    double w_a, w_b;
    uint64_t seed;
} ahip_connect_spec;
int ahip_proj_generate(ahip_net* net, int32_t proj, const ahip_connect_spec* spec,
                       const ahip_format_opts* opts);
/* The forward CSR view of a projection: row_ptr (n_rows + 1) and col (nnz);
 * either may be NULL. */
int ahip_proj_get_csr(ahip_net* net, int32_t proj, int64_t* row_ptr, int32_t* col, int64_t nnz);
/* Re-layout an initialised rate projection in another format. */
int ahip_proj_set_format(ahip_net* net, int32_t proj, const ahip_format_opts* opts);
int ahip_proj_get_format(ahip_net* net, int32_t proj, ahip_format_opts* opts);
/* nb_synapses(), nb_dendrites() */
int ahip_proj_nb_synapses(ahip_net* net, int32_t proj, int64_t* nnz);
/* get/set_local_attribute_all("w") in forward (row) order (a plastic
 * projection reads and writes its device weights). */
int ahip_proj_get_w(ahip_net* net, int32_t proj, double* w, int64_t nnz);
int ahip_proj_set_w(ahip_net* net, int32_t proj, const double* w, int64_t nnz);
/* Plasticity: attach a generated synapse model (a shared library exporting
 * ahip_synapse_describe(), include/annarchy_hip_synapse.h) to a projection,
 * as the reference's Projection(synapse=...) with the generated compute_psp,
 * update_synapse and post_event code (Projection/OpenMPGenerator.py:
 * 1057-1230, 1482-1815).  The projection then keeps one weight per synapse
 * (a rate-coded one in a csr format, updated in place) and the model's
 * attributes at their initial values, with _last_event = -10000;
 * re-initialising the connectivity re-initialises them.  Spiking: pre events
 * replace the plain spike scatter.  After every population update of the
 * step: update_synapse, then (spiking) the post events.  The model's
 * pre.<x> / post.<x> must be attributes of the populations.  Delays: a
 * spiking projection takes one uniform delay when it reads no pre.<x>. */
int ahip_proj_set_synapse(ahip_net* net, int32_t proj, const char* synapse_lib);
/* get/set_local_attribute_all / get/set_semiglobal_attribute_all /
 * get/set_global_attribute of a synaptic attribute: n = nnz values in
 * forward order per synapse, n_rows per post neuron, 1 per projection (set
 * also takes n = 1, one value for all) (Projection/OpenMP/CSR.py accessor
 * templates). */
int ahip_proj_syn_get(ahip_net* net, int32_t proj, const char* attr, double* values, int64_t n);
int ahip_proj_syn_set(ahip_net* net, int32_t proj, const char* attr, const double* values, int64_t n);
/* _last_event of every synapse (forward order), in steps. */
int ahip_proj_get_last_event(ahip_net* net, int32_t proj, int64_t* values, int64_t n);
/* _transmission (core/Projection.py:1066-1100): 0 skips the projection's
 * compute_psp, its pre/post events and update_synapse; its target sum still
 * reads 0 at the step, as the reference clears sums every step. */
int ahip_proj_set_transmission(ahip_net* net, int32_t proj, int32_t transmission);
/* enable_learning / disable_learning (core/Projection.py:1440-1510): update
 * = _update (0 skips update_synapse), plasticity = _plasticity (0 leaves w
 * unchanged in events and updates), update_synapse every `period` steps when
 * (t - offset) mod period == 0, with dt scaled by the period. */
int ahip_proj_set_learning(ahip_net* net, int32_t proj, int32_t update, int32_t plasticity, int32_t period,
                           int32_t offset);
/* Measurement helper with no counterpart in the reference: runs `iters`
 * ordinary simulation steps (they advance the simulation, eagerly launched)
 * with HIP events on the network stream around one kernel inside each step;
 * *avg_ms = mean time of that kernel within real steps.  what / index:
 * plastic projection `index`: its pre events, update_synapse or post
 * events; population `index`: its update or the spike-compaction launch
 * (one launch compacts every spiking population updated in the step); fused
 * spike-scatter launch `index` (0 .. ahip_scatter_groups - 1). */
#define AHIP_PROBE_PRE_EVENTS 0
#define AHIP_PROBE_SYN_UPDATE 1
#define AHIP_PROBE_POST_EVENTS 2
#define AHIP_PROBE_UPDATE 3
#define AHIP_PROBE_COMPACT 4
#define AHIP_PROBE_SCATTER 5
int ahip_time_kernel(ahip_net* net, int32_t what, int32_t index, int32_t iters, double* avg_ms);
#define AHIP_PROBE_DELIVER 6   /* the ordered delivery launch of every conductance (index 0) */
#define AHIP_PROBE_PERSIST 7   /* the persistent steps (index 0): each of the `iters` runs one launch of
                                  a graph batch of steps; *avg_ms = mean launch duration */
/* Number of fused spike-scatter launches per step (spiking projections
 * without per-synapse delays or synapse models, grouped by target). */
int ahip_scatter_groups(ahip_net* net, int32_t* n_groups);
/* Spike delivery of projections with per-synapse weights.  ORDERED (the
 * default) adds every post neuron's events projection by projection and in
 * ascending pre-synaptic rank, the order of the reference's CSRC loop
 * (generator/Projection/OpenMP/CSR.py:870-885), so conductances are
 * bit-identical to the reference's; the scatter appends each event to the
 * post neuron's bucket (and the neuron, at its first event of the step, to
 * the conductance's touched list), and one more kernel adds the buckets of
 * the touched neurons of every such conductance.
 * ATOMIC adds the weights with device atomics in arbitrary order (one
 * rounding of difference per event; runs can differ from each other).
 * Projections with a single weight are order-free either way.  No
 * counterpart in the reference, which has only the ordered form. */
#define AHIP_DELIVERY_ORDERED 0
#define AHIP_DELIVERY_ATOMIC 1
int ahip_net_set_spike_delivery(ahip_net* net, int32_t mode);
/* Persistent steps of a small network (on by default): a network of one
 * active spiking population whose projections all start and end in it, each
 * with one weight, no delay and no synapse model, recorded at most by one
 * spike monitor, with its conductances and spike list within one
 * workgroup's LDS, runs a whole batch of steps in one launch of one
 * workgroup (the same statements in the same order as the per-kernel step;
 * the reference's single-thread loop, SingleThread/CSR.py:867-888, is the
 * shape it restates).  No counterpart in the reference.  enabled = 0 keeps
 * the per-kernel step. */
int ahip_net_set_persistent(ahip_net* net, int32_t enabled);
/* *yes = 1 when the next run would use the persistent steps. */
int ahip_net_uses_persistent(ahip_net* net, int32_t* yes);
/* Number of conductances delivered in order (all in one launch, probed as
 * AHIP_PROBE_DELIVER index 0). */
int ahip_ordered_targets(ahip_net* net, int32_t* n);
/* Bytes the psp kernel must move per step (algorithmic, no padding). */
int ahip_proj_psp_bytes(ahip_net* net, int32_t proj, double* bytes);
/* Times `iters` back-to-back launches of the projection's psp kernel with HIP
 * events recorded on the network stream; *avg_ms = mean launch duration. */
int ahip_proj_time_psp(ahip_net* net, int32_t proj, int32_t iters, double* avg_ms);
/* update_launch_config(nb, tpb) equivalent (Projection.py:2125-2150). */
int ahip_proj_set_launch(ahip_net* net, int32_t proj, int32_t lanes_per_row);
/* Synaptic delays in steps: set_delay(int) / set_delay(vector<vector<int>>)
 * of the uniform and per-synapse delay templates (Projection/OpenMP/CSR.py:
 * 67-101).  n = 1: one delay for every synapse; n = nnz: one per synapse in
 * the forward order of ahip_proj_init_csr.  1 <= delay <= 32767; a delay of 1
 * is the plain transmission (values of the previous step).  A rate-coded
 * projection with differing delays needs the csr format.  Re-initialising
 * the connectivity clears the delays. */
int ahip_proj_set_delays(ahip_net* net, int32_t proj, const int32_t* steps, int64_t n);
/* set_max_delay / update_max_delay (Projection/OpenMP/CSR.py:85-101): the
 * queues read by this projection hold at least max_delay steps (delays drawn
 * from a distribution size them by its maximum, core/Projection.py:661-679).
 * max_delay only grows; growing keeps the queued values. */
int ahip_proj_set_max_delay(ahip_net* net, int32_t proj, int32_t max_delay);
/* *uniform = the common delay or -1 if the delays differ, *max_delay; steps
 * (optional, n >= nnz) receives the delay of every synapse (forward order). */
int ahip_proj_get_delays(ahip_net* net, int32_t proj, int32_t* uniform, int32_t* max_delay, int32_t* steps,
                         int64_t n);

/* ---- monitors (device-side recording) ------------------------------------ */
/* Record attr of pop each step into a device buffer of `capacity` steps
 * (targets g_* and _sum_* are recorded after compute_psp, other variables
 * after the update, OpenMPBaseTemplate.py:235-305).  attr "spike" records the
 * spike list.  Recording starts at the current time step. */
int ahip_monitor_create(ahip_net* net, int32_t pop, const char* attr, int64_t capacity,
                        int32_t* mon_id);
/* Steps recorded so far (<= capacity). */
int ahip_monitor_steps(ahip_net* net, int32_t mon, int64_t* steps);
/* Values: steps*size doubles, step-major. */
int ahip_monitor_get(ahip_net* net, int32_t mon, double* out, int64_t n);
/* Spikes: counts[steps] and the ascending indices of every step, concatenated;
 * *total receives the number of indices (capacity check on idx_cap). */
int ahip_monitor_get_spikes(ahip_net* net, int32_t mon, int32_t* counts, int64_t steps,
                            int32_t* idx, int64_t idx_cap, int64_t* total);
/* Clear the recorded data and restart recording at the current step. */
int ahip_monitor_clear(ahip_net* net, int32_t mon);
/* period / period_offset of a value monitor (generator/Monitor/
 * OpenMPTemplates.py:77, core/Monitor.py:131-166, :323-332): it records at
 * the steps t with (t - phase) mod period == 0 (phase = offset +
 * period_offset in steps), one row per recorded step.  Spike monitors
 * record every step (period 1 only). */
int ahip_monitor_set_period(ahip_net* net, int32_t mon, int32_t period, int64_t phase);
/* pause() / resume() of a monitor (core/Monitor.py:436-500): no row is
 * written while paused; rows go on from where they stopped (for spikes one
 * row per step of the recorded segments). */
int ahip_monitor_pause(ahip_net* net, int32_t mon);
int ahip_monitor_resume(ahip_net* net, int32_t mon);

/* ---- batched instances (parallel_run, core/Network.py:751-926) ------------
 * Many rate-network instances of uneven size advanced together: instance i
 * has N_i neurons  tau dr/dt = -r + sum(exc) + gain_i*sum(inp), r >= 0, a
 * recurrent sparse projection (block-diagonal CSR over all instances, FP32
 * with compensated sums) and an input population of n_in neurons
 * r = 1 + 0.5*sin(omega*t*dt + phase) feeding the first min(N_i, n_in)
 * neurons through one dense n_in x n_in weight matrix shared by every
 * instance (a GEMM on MFMA, split-BF16 operands, FP32 accumulation).  The
 * reference runs such sweeps as one process per instance. */
typedef struct ahip_batch ahip_batch;
const char* ahip_batch_last_error(void);
int ahip_batch_create(int32_t n_inst, const int32_t* sizes, int32_t n_in, double dt, double tau,
                      ahip_batch** out);
int ahip_batch_destroy(ahip_batch* b);
/* row_ptr[total+1], col (global row indices of the same instance), w (nnz) */
int ahip_batch_set_recurrent(ahip_batch* b, const int64_t* row_ptr, const int32_t* col,
                             const double* w, int32_t lanes);
/* w_in: n_in x n_in row-major (row = post neuron); phase: n_in x n_inst */
int ahip_batch_set_input(ahip_batch* b, const double* w_in, const double* phase, double omega);
int ahip_batch_set_gain(ahip_batch* b, const double* gain);
int ahip_batch_set_rates(ahip_batch* b, const double* r);
int ahip_batch_get_rates(ahip_batch* b, double* r);
/* Y = W_in . X of the last step, n_in x n_inst */
int ahip_batch_get_input_sums(ahip_batch* b, double* y);
int ahip_batch_run(ahip_batch* b, int64_t steps, int32_t graph_batch);
/* which: 0 input rates of one chunk, 1 GEMM of one chunk, 2 SpMV + update
 * of one step; mean launch time (ms).  The input rates depend on t alone, so
 * the input psp of a chunk of steps (one graph batch) is one GEMM
 * Y^T = X^T W^T whose rows are (step, instance). */
int ahip_batch_time(ahip_batch* b, int32_t which, int32_t iters, double* avg_ms);
/* Steps per input chunk (the GEMM of one launch covers that many steps). */
int ahip_batch_chunk_steps(ahip_batch* b, int32_t* steps);
/* 2 (default when every instance has fewer than 16384 neurons): chunks of up
 * to 512 rows of one instance in sorted 64-row slices, one lane per row,
 * uint16 instance-local columns, instance rates staged in LDS; 1: the
 * instance's rates staged in LDS, L lanes per row over the CSR; 0: rates
 * gathered from global memory. */
int ahip_batch_set_mode(ahip_batch* b, int32_t mode);
int ahip_batch_get_mode(ahip_batch* b, int32_t* mode);
int ahip_batch_info(ahip_batch* b, int64_t* nnz, int32_t* total, double* sparse_bytes,
                    double* dense_flops);

/* ---- learned storage-format selector --------------------------------------
 * The reference picks a format with a decision tree on density and mean row
 * length (core/Projection.py:691-789).  The selector here is an MLP
 * 16-64-64-12 (ReLU, softmax) trained on the device from our own autotuner
 * timings.  X: n x 16 features, T: n x 12 target distributions, params: 6028
 * floats [W1 b1 W2 b2 W3 b3] (updated in place by training).
 * epochs > 0: full-batch Adam; loss_curve (optional) receives the loss of
 * every epoch.  epochs == 0: one forward/backward pass, grad_out receives
 * dLoss/dparams and loss_curve[0] the loss. */
int ahip_mlp_train(const float* X, const float* T, int32_t n, float* params, int32_t epochs,
                   float lr, float weight_decay, float* loss_curve, float* grad_out);
/* Class (argmax) and logits (optional, n x 12) of every sample. */
int ahip_mlp_predict(const float* X, int32_t n, const float* params, int32_t* classes,
                     float* logits);
/* ahip_mlp_train without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  X (n x 16), T
 * (n x 12), params (6028, updated in place), loss_curve (optional, one
 * float per epoch) and grad_out (optional, 6028) are DEVICE pointers owned
 * by the caller; the kernels are enqueued on `stream` (a hipStream_t, NULL:
 * the null stream) and the call returns without waiting.  The scratch
 * (gradient, Adam moments, per-sample activations) is one static device
 * array of the library, about 2.3 MB per GPU, held until the process ends:
 * n <= 2048, larger n is rejected with AHIP_EINVAL.  Every call on one GPU
 * uses that GPU's array, so calls that overlap on one GPU race. */
int ahip_mlp_train_noalloc(const float* X, const float* T, int32_t n, float* params, int32_t epochs,
                           float lr, float weight_decay, float* loss_curve, float* grad_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif
