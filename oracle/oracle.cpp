// oracle.cpp -- CPU FP64 restatement of ANNarchy's generated OpenMP simulation step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker against which the HIP
// backend is compared; it is never linked into, loaded by or called from the
// product path.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg may use it.
//
// Every function names the reference file:line it follows (paths relative to
// the ANNarchy source tree).  Build flags: -O3 -march=native -fopenmp
// -ffp-contract=off (no FMA contraction, so FP64 neuron updates are
// reproducible bit for bit by a device build that also disables contraction).
//
// Pinning of each function (what checks it besides retyping it):
//   orc_csr_from_lil      tests_cpp golden CSR layouts (6x8, 8x6, empty rows)
//   orc_csrc_inverse      brute-force transpose on random matrices + 6x8 golden
//   orc_csrc_inverse_fast equal to orc_csrc_inverse on random matrices
//   orc_ell_from_lil      tests_cpp golden ELL/ELLR row- and column-major
//   orc_sell_from_lil     tests_cpp golden SELL (S=3) layout
//   orc_bsr_from_lil      tests_cpp golden BSR (tile 2) layout
//   orc_dense_from_lil    tests_cpp golden dense row-major layout
//   orc_spmv_csr          closed form W.r (numpy dense product) as in
//                         tests_python/RateSynapse/test_RateDelays.py:99-153
//   orc_spmv_csr_f32      the FP32 summation error bound n 2^-24 sum|w r|
//                         against the exact product of the FP32 inputs
//   orc_rate_run          closed-form leaky integration at constant input and
//                         the explicit-Euler recurrence evaluated by numpy
//   orc_delayed_psp_run,  tests_python/RateSynapse/test_RateDelays.py uniform
//   orc_rate_run_delayed  (10 / 5 ms) and per-synapse ([[3],[5],[2]]) known
//                         answers; delay 1 equals the undelayed run; a delayed
//                         queue is the shifted input sequence
//   orc_spiking_run       tests_python/Neuron/test_SpikingNeuron.py and
//                         test_Monitor.py:294-333 trajectories (counter model),
//                         test_SpikingTransmission.py:14-78 (5 spikes x w=1)
//                         and :86-253 (uniform and per-synapse delays),
//                         COBA: free-membrane exponential relaxation invariant
//   orc_sweep_instance_run the geometric approach of each input row to
//                         g (W_in u)_i / (1 - beta) at constant input with
//                         self-weights beta; rows beyond the input stay 0
//   orc_mlp_*             finite differences on the CPU (numpy)
//   orc_rate_plastic_run  closed forms: geometric Hebbian growth of one
//                         synapse at a constant rate, Oja's fixed point
//                         1/sqrt(alpha), the exponential relaxation of
//                         IBCM's threshold at a fixed post rate, the update
//                         period (fewer, larger steps)
//   orc_plastic_run       hand-derived closed forms of single-synapse spike
//                         pairings (trace decays between events, the order
//                         g += w before the weight change, clipping, t_pre /
//                         t_post, STP's depression and facilitation), the
//                         independence of synapses (an N x M run equals its
//                         1 x 1 runs), a uniform delay

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>
#include <omp.h>

extern "C" {

// ---------------------------------------------------------------------------
// Containers
// ---------------------------------------------------------------------------

// CSR forward view from a LIL (post_ranks ascending, one list of pre ranks per
// dendrite).  Rows are indexed by post rank: empty rows get an empty range.
// Follows include/CSRMatrix.hpp:142-193.  Returns nnz.
int64_t orc_csr_from_lil(int32_t num_rows, int32_t n_lil, const int32_t* post_ranks,
                         const int64_t* lil_ptr, const int32_t* lil_cols,
                         int64_t* row_begin, int32_t* col_idx) {
    int64_t nnz = 0;
    int32_t k = 0;
    for (int32_t r = 0; r < num_rows; r++) {
        row_begin[r] = nnz;
        if (k == n_lil) continue;
        if (r == post_ranks[k]) {
            for (int64_t j = lil_ptr[k]; j < lil_ptr[k + 1]; j++) col_idx[nnz++] = lil_cols[j];
            k++;
        }
    }
    row_begin[num_rows] = nnz;
    return (k == n_lil) ? nnz : -1;
}

// CSRC backward (column) view: per pre neuron, the post ranks in ascending
// order and the forward index of each synapse.
// Follows include/CSRCMatrix.hpp:164-210.
// The same view by a counting sort, O(nnz): rows are visited in ascending
// order, so each column receives its rows ascending, as the map-based
// restatement below.  Used to build views of large networks; checked equal to
// orc_csrc_inverse on random matrices.
void orc_csrc_inverse_fast(int32_t num_rows, int32_t num_cols, const int64_t* row_begin,
                           const int32_t* col_idx, int64_t* col_ptr, int32_t* row_idx,
                           int64_t* inv_idx) {
    const int64_t nnz = row_begin[num_rows];
    std::vector<int64_t> cur(num_cols + 1, 0);
    for (int64_t s = 0; s < nnz; s++) cur[col_idx[s] + 1]++;
    for (int32_t c = 0; c < num_cols; c++) cur[c + 1] += cur[c];
    for (int32_t c = 0; c <= num_cols; c++) col_ptr[c] = cur[c];
    for (int32_t i = 0; i < num_rows; i++)
        for (int64_t s = row_begin[i]; s < row_begin[i + 1]; s++) {
            const int64_t q = cur[col_idx[s]]++;
            row_idx[q] = i;
            inv_idx[q] = s;
        }
}

void orc_csrc_inverse(int32_t num_rows, int32_t num_cols, const int64_t* row_begin,
                      const int32_t* col_idx, int64_t* col_ptr, int32_t* row_idx,
                      int64_t* inv_idx) {
    std::map<int32_t, std::vector<int32_t>> inv_post;
    std::map<int32_t, std::vector<int64_t>> inv;
    for (int32_t i = 0; i < num_rows; i++) {
        for (int64_t s = row_begin[i]; s < row_begin[i + 1]; s++) {
            inv_post[col_idx[s]].push_back(i);
            inv[col_idx[s]].push_back(s);
        }
    }
    int64_t off = 0;
    for (int32_t c = 0; c < num_cols; c++) {
        col_ptr[c] = off;
        auto it = inv_post.find(c);
        if (it != inv_post.end()) {
            const auto& rows = it->second;
            const auto& idx = inv[c];
            for (size_t q = 0; q < rows.size(); q++) {
                row_idx[off + q] = rows[q];
                inv_idx[off + q] = idx[q];
            }
            off += (int64_t)rows.size();
        }
    }
    col_ptr[num_cols] = off;
}

// ELLPACK over the LIL rows (empty rows are not represented), padding column
// index = INT_MAX, padding value 0.  row_major selects [row][k] vs [k][row].
// Follows include/ELLMatrix.hpp:290-349 (ELLR adds the row lengths,
// include/ELLRMatrix.hpp).  Call with col_out == NULL to get maxnzr only.
int32_t orc_ell_from_lil(int32_t n_lil, const int64_t* lil_ptr, const int32_t* lil_cols,
                         const double* lil_vals, int32_t row_major, int32_t* col_out,
                         double* val_out, int32_t* rl_out) {
    int32_t maxnzr = 0;
    for (int32_t k = 0; k < n_lil; k++)
        maxnzr = std::max<int32_t>(maxnzr, (int32_t)(lil_ptr[k + 1] - lil_ptr[k]));
    if (!col_out) return maxnzr;
    int64_t total = (int64_t)maxnzr * n_lil;
    for (int64_t e = 0; e < total; e++) { col_out[e] = INT_MAX; val_out[e] = 0.0; }
    for (int32_t r = 0; r < n_lil; r++) {
        int32_t c = 0;
        if (rl_out) rl_out[r] = (int32_t)(lil_ptr[r + 1] - lil_ptr[r]);
        for (int64_t j = lil_ptr[r]; j < lil_ptr[r + 1]; j++, c++) {
            int64_t pos = row_major ? (int64_t)r * maxnzr + c : (int64_t)c * n_lil + r;
            col_out[pos] = lil_cols[j];
            val_out[pos] = lil_vals ? lil_vals[j] : 1.0;
        }
    }
    return maxnzr;
}

// Sliced ELLPACK without sorting: slices of block_size consecutive rows of
// the full matrix (rows absent from the LIL have length 0), each padded to its
// longest row; padding column 0, value 0, mask 0.
// Follows include/SELLMatrix.hpp:285-396.  Call with col_out == NULL to get
// the number of stored entries; row_ptr receives num_blocks+1 offsets.
int64_t orc_sell_from_lil(int32_t num_rows, int32_t n_lil, const int32_t* post_ranks,
                          const int64_t* lil_ptr, const int32_t* lil_cols,
                          const double* lil_vals, int32_t block_size, int32_t row_major,
                          int64_t* row_ptr, int32_t* col_out, double* val_out,
                          int8_t* mask_out) {
    std::vector<int32_t> rl(num_rows, 0), lil_of(num_rows, -1);
    for (int32_t k = 0; k < n_lil; k++) {
        rl[post_ranks[k]] = (int32_t)(lil_ptr[k + 1] - lil_ptr[k]);
        lil_of[post_ranks[k]] = k;
    }
    int32_t nb = (num_rows + block_size - 1) / block_size;
    int64_t total = 0;
    std::vector<int32_t> bl(nb, 0);
    for (int32_t b = 0; b < nb; b++) {
        for (int32_t j = 0; j < block_size; j++) {
            int32_t r = b * block_size + j;
            if (r >= num_rows) break;
            bl[b] = std::max(bl[b], rl[r]);
        }
        total += (int64_t)bl[b] * block_size;
    }
    if (!col_out) return total;
    int64_t off = 0;
    for (int32_t b = 0; b < nb; b++) {
        row_ptr[b] = off;
        int64_t n = (int64_t)bl[b] * block_size;
        for (int64_t e = 0; e < n; e++) { col_out[off + e] = 0; val_out[off + e] = 0.0; mask_out[off + e] = 0; }
        for (int32_t j = 0; j < block_size; j++) {
            int32_t r = b * block_size + j;
            if (r >= num_rows) break;
            int32_t k = lil_of[r];
            if (k < 0) continue;
            int32_t c = 0;
            for (int64_t q = lil_ptr[k]; q < lil_ptr[k + 1]; q++, c++) {
                int64_t pos = off + (row_major ? (int64_t)j * bl[b] + c : (int64_t)c * block_size + j);
                col_out[pos] = lil_cols[q];
                val_out[pos] = lil_vals ? lil_vals[q] : 1.0;
                mask_out[pos] = 1;
            }
        }
        off += n;
    }
    row_ptr[nb] = off;
    return total;
}

// Block sparse row: tiles of tile x tile, a tile is stored (dense, row-major
// inside the tile) when it holds at least one synapse; tiles in a block row
// are ordered by block column.  Follows include/BSRMatrix.hpp:149-280.
// Call with bci_out == NULL to get the number of tiles.
int64_t orc_bsr_from_lil(int32_t num_rows, int32_t num_cols, int32_t tile, int32_t n_lil,
                         const int32_t* post_ranks, const int64_t* lil_ptr,
                         const int32_t* lil_cols, const double* lil_vals, int32_t* brp_out,
                         int32_t* bci_out, double* val_out) {
    int32_t nbr = (num_rows + tile - 1) / tile;
    int32_t nbc = (num_cols + tile - 1) / tile;
    std::vector<std::map<int32_t, std::vector<std::pair<int32_t, double>>>> rows(nbr);
    for (int32_t k = 0; k < n_lil; k++) {
        int32_t r = post_ranks[k];
        for (int64_t q = lil_ptr[k]; q < lil_ptr[k + 1]; q++) {
            int32_t c = lil_cols[q];
            rows[r / tile][c / tile].push_back({(r % tile) * tile + (c % tile), lil_vals ? lil_vals[q] : 1.0});
        }
    }
    (void)nbc;
    int64_t ntiles = 0;
    for (int32_t b = 0; b < nbr; b++) ntiles += (int64_t)rows[b].size();
    if (!bci_out) return ntiles;
    int64_t t = 0;
    for (int32_t b = 0; b < nbr; b++) {
        brp_out[b] = (int32_t)t;
        for (auto& kv : rows[b]) {
            bci_out[t] = kv.first;
            double* tilev = val_out + t * tile * tile;
            for (int32_t e = 0; e < tile * tile; e++) tilev[e] = 0.0;
            for (auto& pv : kv.second) tilev[pv.first] = pv.second;
            t++;
        }
    }
    brp_out[nbr] = (int32_t)t;
    return ntiles;
}

// Dense matrix (row-major when row_major != 0).  include/DenseMatrix.hpp:44.
void orc_dense_from_lil(int32_t num_rows, int32_t num_cols, int32_t n_lil,
                        const int32_t* post_ranks, const int64_t* lil_ptr,
                        const int32_t* lil_cols, const double* lil_vals, int32_t row_major,
                        double* out) {
    std::memset(out, 0, sizeof(double) * (size_t)num_rows * num_cols);
    for (int32_t k = 0; k < n_lil; k++) {
        int32_t r = post_ranks[k];
        for (int64_t q = lil_ptr[k]; q < lil_ptr[k + 1]; q++) {
            int32_t c = lil_cols[q];
            size_t pos = row_major ? (size_t)r * num_cols + c : (size_t)c * num_rows + r;
            out[pos] = lil_vals ? lil_vals[q] : 1.0;
        }
    }
}

// ---------------------------------------------------------------------------
// Rate-coded psp: generated ProjStruct::compute_psp for CSR, sum operation.
// Follows generator/Projection/OpenMP/CSR.py:178-196 (psp text
// "w[j]*pre.r[col_idx[j]]", generator/Projection/OpenMPGenerator.py:751-754):
// for each dendrite i (post rank rk), sum from 0.0 over j ascending, then
// target[rk] += sum.  Rows are split statically over threads.
// ---------------------------------------------------------------------------
void orc_spmv_csr(int32_t n_post, const int32_t* post_ranks, const int64_t* row_begin,
                  const int32_t* col_idx, const double* w, const double* r, double* target,
                  int32_t nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int32_t i = 0; i < n_post; i++) {
        int32_t rk = post_ranks ? post_ranks[i] : i;
        double sum = 0.0;
        for (int64_t j = row_begin[rk]; j < row_begin[rk + 1]; j++) sum += w[j] * r[col_idx[j]];
        target[rk] += sum;
    }
}

// FP32 single-precision variant of the same loop (dtype=float32 in the
// reference, core/CTypes.py:30-74); used to state the error of the plain FP32
// form next to the compensated one.
void orc_spmv_csr_f32(int32_t n_post, const int64_t* row_begin, const int32_t* col_idx,
                      const float* w, const float* r, float* target, int32_t nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int32_t i = 0; i < n_post; i++) {
        float sum = 0.0f;
        for (int64_t j = row_begin[i]; j < row_begin[i + 1]; j++) sum += w[j] * r[col_idx[j]];
        target[i] += sum;
    }
}

// ---------------------------------------------------------------------------
// Rate network step loop (one population, one recurrent projection "exc"):
//
//   tau * dr/dt = -r + sum(exc) + Normal(0, sigma)   : min = 0.0
//
// Step order of generator/Template/OpenMPBaseTemplate.py:217-325:
// reset _sum_exc (Population/OpenMPTemplates.py:269-276), compute_psp,
// draw the random variable (Population/OpenMPGenerator.py:675-760), update
// (Population/OpenMPGenerator.py:765-868), t++.  The update statement is the
// explicit-Euler text the sympy front end emits for this equation
// (parser/Equation.py:276-307):
//     double _r = (_sum_exc[i] - r[i] + rand_0[i])/tau;
//     r[i] += dt*_r ;
//     if(r[i] < 0.0) r[i] = 0.0;          (generator/Utils.py:48-66)
// The per-step normal draws are injected (noise[step*N + i]) so that a device
// run can consume the same stream; gain multiplies sum(exc) when non-NULL
// (batched parameter sweep).  trace receives r after each step when non-NULL.
// ---------------------------------------------------------------------------
void orc_rate_run(int32_t N, const int64_t* row_begin, const int32_t* col_idx,
                  const double* w, double* r, double* sum_exc, double tau, double dt,
                  const double* noise, int32_t nsteps, double* trace, int32_t nthreads) {
    for (int32_t s = 0; s < nsteps; s++) {
        for (int32_t i = 0; i < N; i++) sum_exc[i] = 0.0;
        orc_spmv_csr(N, nullptr, row_begin, col_idx, w, r, sum_exc, nthreads);
        const double* rand_0 = noise ? noise + (int64_t)s * N : nullptr;
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int32_t i = 0; i < N; i++) {
            double rd = rand_0 ? rand_0[i] : 0.0;
            double _r = (sum_exc[i] - r[i] + rd) / tau;
            r[i] += dt * _r;
            if (r[i] < 0.0) r[i] = 0.0;
        }
        if (trace) std::memcpy(trace + (int64_t)s * N, r, sizeof(double) * N);
    }
}

// ---------------------------------------------------------------------------
// Rate-coded delays (Population/OpenMPTemplates.py:181-220 and
// Projection/OpenMP/CSR.py:67-101, :962-963).  The pre population keeps a
// queue of its last max_delay rate vectors, hist[k] = slot k of
// `_delayed_r` (k = 0 is the newest); compute_psp reads hist[d-1][col] with
// d the uniform delay or the synapse's delay, and after the population
// updates the new rates are pushed to the front (push_front + pop_back).
// The queue is zero at initialisation and holds the current rates after a
// reset; the caller passes its content in hist (max_delay * n_pre, updated).
//
// orc_delayed_psp_run: psp of one projection for a given sequence of pre
// rates (pre_seq[s] = rates at the end of step s), out[s] = psp of step s.
// orc_rate_run_delayed: orc_rate_run's recurrent network with a delayed
// projection.  delay == NULL means the uniform delay d_uniform (>= 1).
// ---------------------------------------------------------------------------
static void delayed_psp(int32_t n_post, const int64_t* row_begin, const int32_t* col_idx,
                        const double* w, const int32_t* delay, int32_t d_uniform, int32_t n_pre,
                        const double* hist, double* out) {
    for (int32_t i = 0; i < n_post; i++) {
        double sum = 0.0;
        for (int64_t j = row_begin[i]; j < row_begin[i + 1]; j++) {
            const int32_t d = delay ? delay[j] : d_uniform;
            sum += w[j] * hist[(int64_t)(d - 1) * n_pre + col_idx[j]];
        }
        out[i] = sum;
    }
}

static void push_front(double* hist, int32_t max_delay, int32_t n, const double* r) {
    for (int32_t k = max_delay - 1; k > 0; k--)
        std::memcpy(hist + (int64_t)k * n, hist + (int64_t)(k - 1) * n, sizeof(double) * n);
    std::memcpy(hist, r, sizeof(double) * n);
}

void orc_delayed_psp_run(int32_t n_post, const int64_t* row_begin, const int32_t* col_idx,
                         const double* w, const int32_t* delay, int32_t d_uniform, int32_t max_delay,
                         int32_t n_pre, const double* pre_seq, double* hist, int32_t nsteps,
                         double* out) {
    for (int32_t s = 0; s < nsteps; s++) {
        delayed_psp(n_post, row_begin, col_idx, w, delay, d_uniform, n_pre, hist, out + (int64_t)s * n_post);
        push_front(hist, max_delay, n_pre, pre_seq + (int64_t)s * n_pre);
    }
}

void orc_rate_run_delayed(int32_t N, const int64_t* row_begin, const int32_t* col_idx,
                          const double* w, const int32_t* delay, int32_t d_uniform,
                          int32_t max_delay, double* hist, double* r, double* sum_exc, double tau,
                          double dt, const double* noise, int32_t nsteps, double* trace) {
    for (int32_t s = 0; s < nsteps; s++) {
        delayed_psp(N, row_begin, col_idx, w, delay, d_uniform, N, hist, sum_exc);
        const double* rand_0 = noise ? noise + (int64_t)s * N : nullptr;
        for (int32_t i = 0; i < N; i++) {
            double rd = rand_0 ? rand_0[i] : 0.0;
            double _r = (sum_exc[i] - r[i] + rd) / tau;
            r[i] += dt * _r;
            if (r[i] < 0.0) r[i] = 0.0;
        }
        push_front(hist, max_delay, N, r);
        if (trace) std::memcpy(trace + (int64_t)s * N, r, sizeof(double) * N);
    }
}

// ---------------------------------------------------------------------------
// Spiking network step loop.
//
// One spiking population P ("main") with model selected by `model`, plus an
// optional Poisson input population Q.  Projections deliver spikes of P or Q
// into g_exc / g_inh of P through the CSRC backward view:
//   for _pre in pre.spiked (ascending): for syn in col_ptr[_pre]..[_pre+1]:
//       g_target[row_idx[syn]] += w
// (generator/Projection/OpenMP/CSR.py:870-885, SingleThread/CSR.py:867-888;
// w is the column-ordered weight w[inv_idx[syn]] or one constant weight,
// core/Projection.py:823-839).
//
// Step order (generator/Template/OpenMPBaseTemplate.py:217-325):
//   compute_psp (spikes of step t-1) -> record targets -> draw RNG ->
//   update -> spike_gather -> record -> t++
// Update (Population/OpenMPGenerator.py:870-997, generator/Utils.py:139-265):
//   in_ref[i] = refractory_remaining[i] > 0 ? 0 : 1;
//   all ODE increments first, then the switches; non-g_ switches are
//   multiplied by in_ref[i] (Utils.py:69-77); refractory_remaining[i] -=
//   (1 - in_ref[i]).
// Gather (Population/OpenMPGenerator.py:999-1237): skip if in_ref[i]==0; if
// the condition holds: reset statements, push i (ascending), last_spike[i]=t,
// refractory_remaining[i] = refractory[i].
//
// Models:
//   0 COBA LIF (sympy explicit-Euler text):
//     _v = (El + Erev_exc*g_exc[i] + Erev_inh*g_inh[i] + I - g_exc[i]*v[i]
//           - g_inh[i]*v[i] - v[i])/tau;
//     _g_exc = -g_exc[i]/tau_exc;   _g_inh = -g_inh[i]/tau_inh;
//     v[i] += dt*_v* in_ref[i];  g_exc[i] += dt*_g_exc;  g_inh[i] += dt*_g_inh;
//     spike: v[i] > Vt     reset: v[i] = Vr;
//   1 counter neuron (tests_python/Neuron/test_SpikingNeuron.py):
//     if (in_ref[i]) { v[i] = v[i] + 1.0; }   spike: v[i] == Vt   reset: v[i] = Vr;
// Poisson input (inputs/PoissonPopulation.py:171-185):
//     p[i] = 1000.0*rand_0[i]/dt;   spike: p[i] < rates
// ---------------------------------------------------------------------------
struct orc_proj {
    int32_t pre_is_input;    // 0: pre = P, 1: pre = Poisson input Q
    int32_t target;          // 0: g_exc, 1: g_inh
    int32_t num_cols;        // pre population size
    const int64_t* col_ptr;  // num_cols + 1
    const int32_t* row_idx;  // post rank per synapse (column order)
    const double* w_col;     // weight per synapse in column order, or NULL
    double w_const;          // used when w_col == NULL
    int32_t delay;           // uniform delay in steps (<= 1: none)
    const int32_t* delay_col;  // per-synapse delays in steps (column order), or NULL
};

struct orc_spiking_cfg {
    int32_t model;           // 0 COBA, 1 counter
    int32_t N;
    double dt;
    int64_t t0;
    // COBA globals
    double El, Vr, Erev_exc, Erev_inh, Vt, tau, tau_exc, tau_inh, I;
    int32_t has_refractory;
    // Poisson input
    int32_t M;               // 0: no input population
    double rates;
    const double* uniforms;  // nsteps*M injected U(0,1) draws
    int32_t n_proj;
    const orc_proj* proj;
};

// Outputs: spike_counts[nsteps] (spikes of P per step) and spike_idx (ascending
// per step, concatenated; capacity spike_cap).  trace_v/trace_gexc (nsteps*N,
// optional): v after the step, g_exc after compute_psp (record_targets).
// Returns the total number of recorded spikes, or -1 if spike_cap overflowed.
int64_t orc_spiking_run(const orc_spiking_cfg* c, double* v, double* g_exc, double* g_inh,
                        const int32_t* refractory, int32_t* refractory_remaining,
                        int64_t* last_spike, int32_t nsteps, int32_t* spike_counts,
                        int32_t* spike_idx, int64_t spike_cap, double* trace_v,
                        double* trace_gexc, int32_t* input_spike_counts) {
    const int32_t N = c->N;
    std::vector<int32_t> spiked, in_spiked;
    std::vector<int16_t> in_ref(N, 1);
    // Delays.  Uniform: the pre population keeps its last D spike lists,
    // hist[0] = the list of the previous step, and the projection reads
    // hist[d-1] (Population/OpenMPGenerator.py:433-450, `_delayed_spike`;
    // Projection/OpenMPGenerator.py:1238-1249).  Per synapse: a ring of
    // per-post event lists indexed by idx_delay, filled with (idx_delay + d-1)
    // % max_delay and drained in post order, then idx_delay advances
    // (Projection/OpenMP/LIL.py:946-1017).
    int32_t Dp = 1, Dq = 1;
    std::vector<int32_t> maxd(c->n_proj, 1), idx_delay(c->n_proj, 0);
    for (int32_t q = 0; q < c->n_proj; q++) {
        const orc_proj& pr = c->proj[q];
        int32_t d = std::max(1, pr.delay);
        if (pr.delay_col)
            for (int64_t syn = 0; syn < pr.col_ptr[pr.num_cols]; syn++) d = std::max(d, pr.delay_col[syn]);
        maxd[q] = d;
        if (pr.pre_is_input) Dq = std::max(Dq, d); else Dp = std::max(Dp, d);
    }
    std::vector<std::vector<int32_t>> hist_p(Dp), hist_q(Dq);
    std::vector<std::vector<std::vector<std::vector<int64_t>>>> ring(c->n_proj);
    for (int32_t q = 0; q < c->n_proj; q++)
        if (c->proj[q].delay_col) ring[q].assign(maxd[q], std::vector<std::vector<int64_t>>(N));
    std::vector<double> p(c->M > 0 ? c->M : 1, 0.0);
    int64_t t = c->t0;
    int64_t rec = 0;
    bool overflow = false;
    const double dt = c->dt;
    for (int32_t s = 0; s < nsteps; s++) {
        // compute_psp: spikes emitted at the previous step (or d-1 steps before)
        for (int32_t q = 0; q < c->n_proj; q++) {
            const orc_proj& pr = c->proj[q];
            const std::vector<std::vector<int32_t>>& hist = pr.pre_is_input ? hist_q : hist_p;
            double* g = pr.target == 0 ? g_exc : g_inh;
            if (!pr.delay_col) {
                const std::vector<int32_t>& pre = hist[std::max(1, pr.delay) - 1];
                for (size_t k = 0; k < pre.size(); k++) {
                    int32_t _pre = pre[k];
                    if (_pre >= pr.num_cols) continue;
                    for (int64_t syn = pr.col_ptr[_pre]; syn < pr.col_ptr[_pre + 1]; syn++)
                        g[pr.row_idx[syn]] += pr.w_col ? pr.w_col[syn] : pr.w_const;
                }
            } else {
                const std::vector<int32_t>& pre = hist[0];
                for (size_t k = 0; k < pre.size(); k++) {
                    int32_t _pre = pre[k];
                    if (_pre >= pr.num_cols) continue;
                    for (int64_t syn = pr.col_ptr[_pre]; syn < pr.col_ptr[_pre + 1]; syn++) {
                        const int32_t slot = (idx_delay[q] + pr.delay_col[syn] - 1) % maxd[q];
                        ring[q][slot][pr.row_idx[syn]].push_back(syn);
                    }
                }
                std::vector<std::vector<int64_t>>& now = ring[q][idx_delay[q]];
                for (int32_t i = 0; i < N; i++) {
                    for (int64_t syn : now[i]) g[i] += pr.w_col ? pr.w_col[syn] : pr.w_const;
                    now[i].clear();
                }
                idx_delay[q] = (idx_delay[q] + 1) % maxd[q];
            }
        }
        if (trace_gexc) std::memcpy(trace_gexc + (int64_t)s * N, g_exc, sizeof(double) * N);
        // Poisson input: draw + update + gather (ascending)
        in_spiked.clear();
        if (c->M > 0) {
            const double* rand_0 = c->uniforms + (int64_t)s * c->M;
            for (int32_t i = 0; i < c->M; i++) p[i] = 1000.0 * rand_0[i] / dt;
            for (int32_t i = 0; i < c->M; i++)
                if (p[i] < c->rates) in_spiked.push_back(i);
            if (input_spike_counts) input_spike_counts[s] = (int32_t)in_spiked.size();
        }
        // update of P
        if (c->has_refractory)
            for (int32_t i = 0; i < N; i++) in_ref[i] = (refractory_remaining[i] > 0) ? 0 : 1;
        if (c->model == 0) {
            const double El = c->El, Erev_exc = c->Erev_exc, Erev_inh = c->Erev_inh, I = c->I;
            const double tau = c->tau, tau_exc = c->tau_exc, tau_inh = c->tau_inh;
            for (int32_t i = 0; i < N; i++) {
                double _v = (El + Erev_exc * g_exc[i] + Erev_inh * g_inh[i] + I - g_exc[i] * v[i] -
                             g_inh[i] * v[i] - v[i]) / tau;
                double _g_exc = -g_exc[i] / tau_exc;
                double _g_inh = -g_inh[i] / tau_inh;
                v[i] += dt * _v * in_ref[i];
                g_exc[i] += dt * _g_exc;
                g_inh[i] += dt * _g_inh;
                if (c->has_refractory) refractory_remaining[i] -= (1 - in_ref[i]);
            }
        } else {
            for (int32_t i = 0; i < N; i++) {
                if (in_ref[i]) { v[i] = v[i] + 1.0; }
                if (c->has_refractory) refractory_remaining[i] -= (1 - in_ref[i]);
            }
        }
        // spike gather of P
        spiked.clear();
        for (int32_t i = 0; i < N; i++) {
            if (c->has_refractory && in_ref[i] == 0) continue;
            bool cond = (c->model == 0) ? (v[i] > c->Vt) : (v[i] == c->Vt);
            if (cond) {
                v[i] = c->Vr;
                spiked.push_back(i);
                last_spike[i] = t;
                if (c->has_refractory) refractory_remaining[i] = refractory[i];
            }
        }
        spike_counts[s] = (int32_t)spiked.size();
        for (size_t k = 0; k < spiked.size(); k++) {
            if (rec < spike_cap) spike_idx[rec] = spiked[k]; else overflow = true;
            rec++;
        }
        if (trace_v) std::memcpy(trace_v + (int64_t)s * N, v, sizeof(double) * N);
        // delay queues: push the lists of this step to the front
        std::rotate(hist_p.rbegin(), hist_p.rbegin() + 1, hist_p.rend());
        hist_p[0] = spiked;
        std::rotate(hist_q.rbegin(), hist_q.rbegin() + 1, hist_q.rend());
        hist_q[0] = in_spiked;
        t++;
    }
    return overflow ? -1 : rec;
}

// ---------------------------------------------------------------------------
// Plastic spiking projection driven by given spike trains of its pre and post
// populations (steps t0 .. t0+steps-1; no spike before t0).
//   rule 0  STDP, models/Synapses.py:232-310; prm = tau_plus, tau_minus,
//           A_plus, A_minus, w_min, w_max; x, y event-driven traces.
//   rule 1  the t_pre / t_post rule of tests_python/SpikeSynapse/
//           test_SpikingSynapse.py (test_TimeDependentUpdate); prm = tau_pre,
//           tau_post, cApre, cApost, wmax; no event-driven variable.
//   rule 2  STP, models/Synapses.py:131-230; prm = tau_rec, tau_facil, U;
//           x (in x) and u (in y) event-driven, g_target += w*u*x.
// Step t:
//   compute_psp (Projection/OpenMP/CSR.py:870-885, OpenMPGenerator.py:
//     1182-1230): for every pre spike of step t-1 in ascending rank, every
//     synapse of its column in ascending post rank: event-driven decay to
//     t-1, _last_event = t, g[post] += w, the pre_spike statements;
//   neuron updates: last_spike of the neurons spiking at t becomes t
//     (t_pre = last_spike[pre]*dt, t_post = last_spike[post]*dt,
//     parser/ParserTemplate.py:18-19; t = t*dt);
//   post_event (Projection/OpenMP/CSR.py:903-918, OpenMPGenerator.py:
//     1575-1600): for every post spike of step t, its row: event-driven decay
//     to t, _last_event = t, the post_spike statements.
// State arrays (w, x, y, last_event) in forward (row) order, updated in
// place; pre_last / post_last in place.  g_out[s*n_rows + i]: conductance
// added to post neuron i by the pre events of step t0+s.  With a uniform
// delay d the pre events of step t are the spikes of step t-d.
// ---------------------------------------------------------------------------
static inline double orc_clip(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

void orc_plastic_run(int32_t rule, int32_t n_rows, int32_t n_cols, const int64_t* row_ptr, const int32_t* col,
                     double* w, double* x, double* y, int64_t* last_event, const double* prm, double dt,
                     int64_t t0, int32_t steps, const int64_t* pre_ptr, const int32_t* pre_idx,
                     const int64_t* post_ptr, const int32_t* post_idx, int64_t* pre_last, int64_t* post_last,
                     int32_t delay, double* g_out) {
    // CSRC view: per pre neuron, its synapses in ascending post rank
    // (include/CSRCMatrix.hpp:164-210)
    const int64_t nnz = row_ptr[n_rows];
    std::vector<int64_t> col_ptr(n_cols + 1, 0), cur;
    for (int64_t j = 0; j < nnz; j++) col_ptr[col[j] + 1]++;
    for (int32_t c = 0; c < n_cols; c++) col_ptr[c + 1] += col_ptr[c];
    cur.assign(col_ptr.begin(), col_ptr.end() - 1);
    std::vector<int64_t> syn_of(nnz);
    std::vector<int32_t> post_of(nnz);
    for (int32_t r = 0; r < n_rows; r++)
        for (int64_t j = row_ptr[r]; j < row_ptr[r + 1]; j++) {
            const int64_t q = cur[col[j]]++;
            syn_of[q] = j;
            post_of[q] = r;
        }
    for (int32_t s = 0; s < steps; s++) {
        const int64_t t = t0 + s;
        double* g = g_out + (int64_t)s * n_rows;
        for (int32_t i = 0; i < n_rows; i++) g[i] = 0.0;
        // compute_psp: spikes of step t-d (uniform delay d >= 1; the
        // event-driven decay still runs to t-1, OpenMPGenerator.py:1207-1211)
        if (s >= delay) {
            for (int64_t k = pre_ptr[s - delay]; k < pre_ptr[s - delay + 1]; k++) {
                const int32_t pre = pre_idx[k];
                for (int64_t q = col_ptr[pre]; q < col_ptr[pre + 1]; q++) {
                    const int64_t j = syn_of[q];
                    const int32_t post = post_of[q];
                    if (rule == 0) {
                        x[j] *= std::exp(dt * (double)(last_event[j] - (t - 1)) / (prm[0]));
                        y[j] *= std::exp(dt * (double)(last_event[j] - (t - 1)) / (prm[1]));
                        last_event[j] = t;
                        g[post] += w[j];
                        x[j] += prm[2] * prm[5];
                        w[j] = orc_clip(w[j] + y[j], prm[4], prm[5]);
                    } else if (rule == 2) {
                        // STP: x -> 1 with tau_rec, u -> U with tau_facil between events
                        x[j] = 1.0 + (x[j] - 1.0) * std::exp(dt * (double)(last_event[j] - (t - 1)) / (prm[0]));
                        y[j] = prm[2] + (y[j] - prm[2]) * std::exp(dt * (double)(last_event[j] - (t - 1)) / (prm[1]));
                        last_event[j] = t;
                        g[post] += w[j] * y[j] * x[j];
                        x[j] *= 1.0 - y[j];
                        y[j] += prm[2] * (1.0 - y[j]);
                    } else {
                        g[post] += w[j];
                        const double t_post = (double)post_last[post] * dt;
                        w[j] = orc_clip(w[j] - prm[3] * std::exp((t_post - (double)t * dt) / prm[1]), 0.0, prm[4]);
                    }
                }
            }
        }
        // neuron updates: the spikes of step t
        for (int64_t k = pre_ptr[s]; k < pre_ptr[s + 1]; k++) pre_last[pre_idx[k]] = t;
        for (int64_t k = post_ptr[s]; k < post_ptr[s + 1]; k++) post_last[post_idx[k]] = t;
        // post_event: spikes of step t
        for (int64_t k = post_ptr[s]; k < post_ptr[s + 1]; k++) {
            const int32_t i = post_idx[k];
            for (int64_t j = row_ptr[i]; j < row_ptr[i + 1]; j++) {
                if (rule == 2) continue;
                if (rule == 0) {
                    x[j] *= std::exp(dt * (double)(last_event[j] - t) / (prm[0]));
                    y[j] *= std::exp(dt * (double)(last_event[j] - t) / (prm[1]));
                    last_event[j] = t;
                    y[j] -= prm[3] * prm[5];
                    w[j] = orc_clip(w[j] + x[j], prm[4], prm[5]);
                } else {
                    const double t_pre = (double)pre_last[col[j]] * dt;
                    w[j] = orc_clip(w[j] + prm[2] * std::exp((t_pre - (double)t * dt) / prm[0]), 0.0, prm[4]);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Rate-coded plastic projection pre -> post with the post neuron r = sum(exc)
// and given pre-synaptic rates (r_pre0 before step 0, r_pre_seq[t] after the
// update of step t).  Step t (OpenMPBaseTemplate.py:217-325):
//   compute_psp (Projection/OpenMP/CSR.py:178-196): sum[i] = sum over the
//     row, j ascending, of w[j]*r_pre[col[j]] with the rates of step t-1;
//   neuron updates: r_post[i] = sum[i]; r_pre = r_pre_seq[t];
//   update_synapse (Projection/OpenMP/CSR.py:833-865), when t % period == 0,
//     with _dt = dt*period (OpenMPGenerator.py:1722-1725): per row the
//     per-post-neuron equation, then per synapse the weight equation
//     (explicit Euler, Equation.py:276-307, min = 0):
//   rule 0 Hebb (models/Synapses.py:33-61)  dw/dt = eta*pre.r*post.r
//   rule 1 Oja  (:63-94)   dw/dt = eta*(pre.r*post.r - alpha*post.r^2*w)
//   rule 2 IBCM (:96-129)  tau*dtheta/dt + theta = post.r^2 (exponential
//          Euler, Equation.py:559-604), dw/dt = eta*post.r*(post.r-theta)*pre.r
// prm = eta, alpha (Oja) / tau (IBCM).  sum_out[t*n_rows + i] = sum of step t.
// ---------------------------------------------------------------------------
void orc_rate_plastic_run(int32_t rule, int32_t n_rows, int32_t n_cols, const int64_t* row_ptr,
                          const int32_t* col, double* w, double* theta, const double* prm, double dt,
                          int32_t period, int32_t steps, const double* r_pre0, const double* r_pre_seq,
                          double* r_post, double* sum_out) {
    std::vector<double> r_pre(r_pre0, r_pre0 + n_cols);
    const double eta = prm[0];
    for (int32_t t = 0; t < steps; t++) {
        double* sum = sum_out + (int64_t)t * n_rows;
        for (int32_t i = 0; i < n_rows; i++) {
            double s = 0.0;
            for (int64_t j = row_ptr[i]; j < row_ptr[i + 1]; j++) s += w[j] * r_pre[col[j]];
            sum[i] = s;
        }
        for (int32_t i = 0; i < n_rows; i++) r_post[i] = sum[i];
        for (int32_t c = 0; c < n_cols; c++) r_pre[c] = r_pre_seq[(int64_t)t * n_cols + c];
        if (t % period != 0) continue;
        const double _dt = dt * period;
        for (int32_t i = 0; i < n_rows; i++) {
            const double post_r = r_post[i];
            if (rule == 2) {
                const double stepsize = 1.0 - std::exp(-_dt / prm[1]);
                theta[i] += stepsize * (std::pow(post_r, 2) - theta[i]);
            }
            for (int64_t j = row_ptr[i]; j < row_ptr[i + 1]; j++) {
                const double pre_r = r_pre[col[j]];
                double dw;
                if (rule == 0)
                    dw = eta * pre_r * post_r;
                else if (rule == 1)
                    dw = eta * (pre_r * post_r - prm[1] * std::pow(post_r, 2) * w[j]);
                else
                    dw = eta * post_r * (post_r - theta[i]) * pre_r;
                w[j] += _dt * dw;
                if (w[j] < 0.0) w[j] = 0.0;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// One instance of the batched parameter sweep, FP64, in the reference's step
// order (OpenMPBaseTemplate.py:217-325): psp of both projections from the
// rates of step t-1, then the updates of the input and exc populations.
//   input (n_in):  r_in[k] = 1.0 + 0.5*sin(omega*(t*dt) + phase[k])
//   exc (N):       _r = (_sum_exc[i] + gain*_sum_inp[i] - r[i])/tau;
//                  r[i] += dt*_r;  if (r[i] < 0.0) r[i] = 0.0;
//   exc -> exc:    CSR (row_begin, col_idx, w), sum from 0, j ascending
//   inp -> exc:    dense rows 0 .. min(N, n_in)-1 of the shared n_in x n_in W_in
// r_in starts at the value of step 0's update (the input population has no
// state); trace (optional) receives r after every step.
// ---------------------------------------------------------------------------
void orc_sweep_instance_run(int32_t N, const int64_t* row_begin, const int32_t* col_idx,
                            const double* w, int32_t n_in, const double* w_in,
                            const double* phase, double omega, double gain, double tau, double dt,
                            double* r, int32_t nsteps, int64_t t0, double* trace) {
    std::vector<double> r_in(n_in, 0.0), s_exc(N), s_inp(N);
    const int32_t rows_in = std::min(N, n_in);
    int64_t t = t0;
    for (int32_t s = 0; s < nsteps; s++) {
        for (int32_t i = 0; i < N; i++) {
            double sum = 0.0;
            for (int64_t j = row_begin[i]; j < row_begin[i + 1]; j++) sum += w[j] * r[col_idx[j]];
            s_exc[i] = sum;
            double si = 0.0;
            if (i < rows_in)
                for (int32_t k = 0; k < n_in; k++) si += w_in[(int64_t)i * n_in + k] * r_in[k];
            s_inp[i] = si;
        }
        for (int32_t k = 0; k < n_in; k++) r_in[k] = 1.0 + 0.5 * std::sin(omega * ((double)t * dt) + phase[k]);
        for (int32_t i = 0; i < N; i++) {
            double _r = (s_exc[i] + gain * s_inp[i] - r[i]) / tau;
            r[i] += dt * _r;
            if (r[i] < 0.0) r[i] = 0.0;
        }
        if (trace) std::memcpy(trace + (int64_t)s * N, r, sizeof(double) * N);
        t++;
    }
}

// ---------------------------------------------------------------------------
// MLP 16-64-64-12 (ReLU, softmax cross-entropy) used by the learned format
// selector: forward and backward on the CPU in FP64, the gradient oracle for
// the device training kernels.  Layout: W_l row-major [out][in], b_l [out].
// ---------------------------------------------------------------------------
static void dense_fwd(int n, int in, int out, const double* W, const double* b, const double* x,
                      double* y, int relu) {
    for (int s = 0; s < n; s++)
        for (int o = 0; o < out; o++) {
            double a = b[o];
            for (int k = 0; k < in; k++) a += W[o * in + k] * x[s * in + k];
            y[s * out + o] = (relu && a < 0.0) ? 0.0 : a;
        }
}

// Logits (n x d_out) of the MLP in FP64.
void orc_mlp_forward(int32_t n, const double* X, const double* params, double* logits, int32_t d_in,
                     int32_t h1, int32_t h2, int32_t d_out) {
    const double* W1 = params;
    const double* b1 = W1 + h1 * d_in;
    const double* W2 = b1 + h1;
    const double* b2 = W2 + h2 * h1;
    const double* W3 = b2 + h2;
    const double* b3 = W3 + d_out * h2;
    std::vector<double> a1(n * h1), a2(n * h2);
    dense_fwd(n, d_in, h1, W1, b1, X, a1.data(), 1);
    dense_fwd(n, h1, h2, W2, b2, a1.data(), a2.data(), 1);
    dense_fwd(n, h2, d_out, W3, b3, a2.data(), logits, 0);
}

// Returns the mean cross-entropy against target distributions T (n x d_out,
// one-hot for hard labels); grads (same layout as params) accumulate the mean
// gradient when non-NULL.  params = [W1 b1 W2 b2 W3 b3].
double orc_mlp_loss_grad(int32_t n, const double* X, const double* T,
                         const double* params, double* grads, int32_t d_in, int32_t h1,
                         int32_t h2, int32_t d_out) {
    const double* W1 = params;
    const double* b1 = W1 + h1 * d_in;
    const double* W2 = b1 + h1;
    const double* b2 = W2 + h2 * h1;
    const double* W3 = b2 + h2;
    const double* b3 = W3 + d_out * h2;
    std::vector<double> a1(n * h1), a2(n * h2), z(n * d_out);
    dense_fwd(n, d_in, h1, W1, b1, X, a1.data(), 1);
    dense_fwd(n, h1, h2, W2, b2, a1.data(), a2.data(), 1);
    dense_fwd(n, h2, d_out, W3, b3, a2.data(), z.data(), 0);
    double loss = 0.0;
    std::vector<double> dz(n * d_out);
    for (int s = 0; s < n; s++) {
        double m = z[s * d_out];
        for (int o = 1; o < d_out; o++) m = std::max(m, z[s * d_out + o]);
        double se = 0.0;
        for (int o = 0; o < d_out; o++) se += std::exp(z[s * d_out + o] - m);
        double lse = m + std::log(se);
        double tsum = 0.0;
        for (int o = 0; o < d_out; o++) {
            loss += T[s * d_out + o] * (lse - z[s * d_out + o]);
            tsum += T[s * d_out + o];
        }
        for (int o = 0; o < d_out; o++)
            dz[s * d_out + o] = (tsum * std::exp(z[s * d_out + o] - lse) - T[s * d_out + o]) / n;
    }
    loss /= n;
    if (!grads) return loss;
    double* gW1 = grads;
    double* gb1 = gW1 + h1 * d_in;
    double* gW2 = gb1 + h1;
    double* gb2 = gW2 + h2 * h1;
    double* gW3 = gb2 + h2;
    double* gb3 = gW3 + d_out * h2;
    std::vector<double> da2(n * h2, 0.0), da1(n * h1, 0.0);
    for (int s = 0; s < n; s++)
        for (int o = 0; o < d_out; o++) {
            double g = dz[s * d_out + o];
            gb3[o] += g;
            for (int k = 0; k < h2; k++) {
                gW3[o * h2 + k] += g * a2[s * h2 + k];
                da2[s * h2 + k] += g * W3[o * h2 + k];
            }
        }
    for (int s = 0; s < n; s++)
        for (int o = 0; o < h2; o++) {
            double g = a2[s * h2 + o] > 0.0 ? da2[s * h2 + o] : 0.0;
            gb2[o] += g;
            for (int k = 0; k < h1; k++) {
                gW2[o * h1 + k] += g * a1[s * h1 + k];
                da1[s * h1 + k] += g * W2[o * h1 + k];
            }
        }
    for (int s = 0; s < n; s++)
        for (int o = 0; o < h1; o++) {
            double g = a1[s * h1 + o] > 0.0 ? da1[s * h1 + o] : 0.0;
            gb1[o] += g;
            for (int k = 0; k < d_in; k++) gW1[o * d_in + k] += g * X[s * d_in + k];
        }
    return loss;
}

}  // extern "C"
