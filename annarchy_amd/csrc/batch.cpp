// batch.cpp -- C-ABI of the batched multi-instance rate sweep (see k_batched.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"


using namespace ahip;

struct ahip_batch {
    BatchDev B;
    hipStream_t stream = nullptr;
    int32_t parity = 0;            // which rate buffer holds the current rates
    int64_t t = 0;
    int64_t nnz = 0;
    bool ready_sparse = false, ready_input = false;
    hipGraphExec_t graph = nullptr;
    int32_t graph_steps = 0;
    int32_t graph_parity = 0;
    int32_t last_chunk = 0;        // steps of the last chunk enqueued (its Y^T is in B.y)
    std::vector<int32_t> sizes;
    std::vector<int64_t> offsets;
};

namespace {
thread_local std::string g_berr;
int bfail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_berr = buf;
    return code;
}
#define BCHK(expr)                                                                                \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess) return bfail(AHIP_EHIP, "%s: %s", #expr, hipGetErrorString(_e));   \
    } while (0)
#define BLCHK(expr)                                                                               \
    do {                                                                                          \
        int _e = (expr);                                                                          \
        if (_e) return bfail(AHIP_EHIP, "launch %s: %s", #expr, hipGetErrorString((hipError_t)_e)); \
    } while (0)

int balloc(DevBuf& b, size_t bytes, const void* src = nullptr) {
    b.release();
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) { b.p = nullptr; return bfail(AHIP_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e)); }
    b.bytes = bytes;
    if (src) BCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    else BCHK(hipMemset(b.p, 0, bytes));
    // the batch stream is non-blocking: finish the clear before its kernels run
    BCHK(hipDeviceSynchronize());
    return 0;
}

uint16_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    uint32_t r = ((u >> 16) & 1u) + 0x7FFFu;
    return (uint16_t)((u + r) >> 16);
}
float bf16_f(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

void drop_graph(ahip_batch* b) {
    if (b->graph) hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_steps = 0;
}

// X and Y buffers for chunks of up to `steps` steps (reallocated only to grow).
int ensure_chunk(ahip_batch* b, int32_t steps) {
    BatchDev& B = b->B;
    if (steps <= B.cap_steps) return 0;
    BCHK(hipStreamSynchronize(b->stream));
    drop_graph(b);
    const size_t frag = (size_t)B.k_steps * 64 * 8 * sizeof(uint16_t);
    const size_t c_tiles = (size_t)steps * B.n_tiles;
    int rc;
    if ((rc = balloc(B.x_hi, frag * c_tiles))) return rc;
    if ((rc = balloc(B.x_lo, frag * c_tiles))) return rc;
    if ((rc = balloc(B.y, c_tiles * 32 * (size_t)B.m_tiles * 32 * sizeof(float)))) return rc;
    B.cap_steps = steps;
    return 0;
}

int enqueue_steps(ahip_batch* b, int32_t n, int32_t parity0) {
    // reference step order: both psp first, from the rates of step t-1 (the
    // input rates of step t-1 are a function of t alone, so the input psp of
    // a whole chunk of steps is computed up front), then the update of the
    // exc population, fused into the SpMV
    const int32_t cap = b->B.cap_steps;
    for (int32_t j0 = 0; j0 < n; j0 += cap) {
        const int32_t m = std::min(cap, n - j0);
        BLCHK(batch_launch_input(b->B, j0, m, b->stream));
        BLCHK(batch_launch_gemm(b->B, m, b->stream));
        for (int32_t j = 0; j < m; j++) BLCHK(batch_launch_spmv(b->B, (parity0 + j0 + j) & 1, j, b->stream));
        b->last_chunk = m;
    }
    BLCHK(batch_launch_advance(b->B, n, b->stream));
    return 0;
}
}  // namespace

extern "C" {

const char* ahip_batch_last_error(void) { return g_berr.c_str(); }

int ahip_batch_create(int32_t n_inst, const int32_t* sizes, int32_t n_in, double dt, double tau,
                      ahip_batch** out) {
    if (n_inst <= 0 || !sizes || n_in <= 0 || !out || !(dt > 0) || !(tau > 0))
        return bfail(AHIP_EINVAL, "bad batch arguments");
    auto* b = new ahip_batch();
    b->B.n_inst = n_inst;
    b->B.n_in = n_in;
    b->B.dt = dt;
    b->B.tau = tau;
    b->sizes.assign(sizes, sizes + n_inst);
    b->offsets.assign(n_inst + 1, 0);
    for (int32_t i = 0; i < n_inst; i++) {
        if (sizes[i] <= 0) { delete b; return bfail(AHIP_EINVAL, "instance sizes must be positive"); }
        b->offsets[i + 1] = b->offsets[i] + sizes[i];
    }
    if (b->offsets[n_inst] > INT32_MAX) { delete b; return bfail(AHIP_EINVAL, "too many neurons"); }
    b->B.total = (int32_t)b->offsets[n_inst];
    b->B.m_tiles = (n_in + 31) / 32;
    b->B.n_tiles = (n_inst + 31) / 32;
    b->B.k_steps = (n_in + 15) / 16;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete b; return bfail(AHIP_EHIP, "stream: %s", hipGetErrorString(e)); }
    int rc;
    BatchDev& B = b->B;
    std::vector<int32_t> inst(B.total), local(B.total);
    for (int32_t i = 0; i < n_inst; i++)
        for (int32_t n = 0; n < sizes[i]; n++) {
            inst[b->offsets[i] + n] = i;
            local[b->offsets[i] + n] = n;
        }
    if ((rc = balloc(B.inst_of_row, inst.size() * 4, inst.data()))) return rc;
    if ((rc = balloc(B.local_of_row, local.size() * 4, local.data()))) return rc;
    if ((rc = balloc(B.r[0], (size_t)B.total * 4))) return rc;
    if ((rc = balloc(B.r[1], (size_t)B.total * 4))) return rc;
    if ((rc = balloc(B.sum_exc, (size_t)B.total * 4))) return rc;
    std::vector<float> g(n_inst, 1.0f);
    if ((rc = balloc(B.gain, g.size() * 4, g.data()))) return rc;
    const size_t frag = (size_t)B.k_steps * 64 * 8 * sizeof(uint16_t);
    if ((rc = balloc(B.w_hi, frag * B.m_tiles))) return rc;
    if ((rc = balloc(B.w_lo, frag * B.m_tiles))) return rc;
    if ((rc = balloc(B.phase, (size_t)n_in * n_inst * 8))) return rc;
    if ((rc = balloc(B.t_dev, 8))) return rc;
    if ((rc = ensure_chunk(b, 1))) return rc;
    *out = b;
    return 0;
}

int ahip_batch_destroy(ahip_batch* b) {
    if (!b) return 0;
    hipStreamSynchronize(b->stream);
    drop_graph(b);
    hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int ahip_batch_set_recurrent(ahip_batch* b, const int64_t* row_ptr, const int32_t* col,
                             const double* w, int32_t lanes) {
    if (!b || !row_ptr || !col || !w) return bfail(AHIP_EINVAL, "null argument");
    if (lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
        return bfail(AHIP_EINVAL, "lanes must be 4..64");
    BatchDev& B = b->B;
    const int32_t R = B.total;
    if (row_ptr[0] != 0) return bfail(AHIP_EINVAL, "row_ptr[0] != 0");
    for (int32_t r = 0; r < R; r++)
        if (row_ptr[r + 1] < row_ptr[r]) return bfail(AHIP_EINVAL, "row_ptr not monotonic");
    const int64_t nnz = row_ptr[R];
    // block-diagonal: every column must stay inside the row's instance
    std::vector<int32_t> inst(R);
    for (int32_t i = 0; i < B.n_inst; i++)
        for (int64_t r = b->offsets[i]; r < b->offsets[i + 1]; r++) inst[r] = i;
    for (int32_t r = 0; r < R; r++) {
        const int64_t lo = b->offsets[inst[r]], hi = b->offsets[inst[r] + 1];
        for (int64_t j = row_ptr[r]; j < row_ptr[r + 1]; j++)
            if (col[j] < lo || col[j] >= hi) return bfail(AHIP_EINVAL, "column %d of row %d leaves its instance", col[j], r);
    }
    std::vector<float> wf(nnz);
    for (int64_t j = 0; j < nnz; j++) wf[j] = (float)w[j];
    BCHK(hipStreamSynchronize(b->stream));
    drop_graph(b);
    int rc;
    if ((rc = balloc(B.row_ptr, (size_t)(R + 1) * 8, row_ptr))) return rc;
    if ((rc = balloc(B.col, (size_t)nnz * 4, col))) return rc;
    if ((rc = balloc(B.w, (size_t)nnz * 4, wf.data()))) return rc;
    B.lanes = lanes;
    b->nnz = nnz;
    // work chunks for the LDS kernel: rows of one instance, ~32k synapses each
    std::vector<int32_t> ch;
    int32_t max_n = 0;
    for (int32_t i = 0; i < B.n_inst; i++) {
        max_n = std::max<int32_t>(max_n, b->sizes[i]);
        int64_t r = b->offsets[i];
        while (r < b->offsets[i + 1]) {
            int64_t e = r;
            while (e < b->offsets[i + 1] && (e == r || row_ptr[e + 1] - row_ptr[r] <= 32768)) e++;
            ch.push_back(i);
            ch.push_back((int32_t)r);
            ch.push_back((int32_t)e);
            r = e;
        }
    }
    B.max_n = max_n;
    B.n_chunks = (int32_t)(ch.size() / 3);
    if ((rc = balloc(B.chunks, ch.size() * 4, ch.data()))) return rc;
    if ((rc = balloc(B.inst_base, b->offsets.size() * 8, b->offsets.data()))) return rc;
    B.mode = (size_t)max_n * sizeof(float) <= 64 * 1024 ? 1 : 0;
    // mode 2: chunks of <= 512 rows of one instance, rows sorted by length
    // (descending, stable) into 64-row slices, width rounded up to 4
    if (max_n < 65535 && (size_t)(max_n + 1) * sizeof(float) <= 64 * 1024) {
        std::vector<int32_t> sch, sch_sl(1, 0), sl_row;
        std::vector<int64_t> sl_off(1, 0);
        std::vector<int32_t> order;
        for (int32_t i = 0; i < B.n_inst; i++) {
            for (int64_t r0 = b->offsets[i]; r0 < b->offsets[i + 1]; r0 += 512) {
                const int64_t r1 = std::min<int64_t>(r0 + 512, b->offsets[i + 1]);
                order.resize(r1 - r0);
                for (int64_t r = r0; r < r1; r++) order[r - r0] = (int32_t)r;
                std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t c) {
                    return row_ptr[a + 1] - row_ptr[a] > row_ptr[c + 1] - row_ptr[c];
                });
                for (size_t q = 0; q < order.size(); q += 64) {
                    const int64_t len0 = row_ptr[order[q] + 1] - row_ptr[order[q]];
                    for (size_t l = 0; l < 64; l++) sl_row.push_back(q + l < order.size() ? order[q + l] : -1);
                    sl_off.push_back(sl_off.back() + (len0 + 3) / 4 * 4 * 64);
                }
                sch.push_back(i);
                sch.push_back((int32_t)r0);
                sch.push_back((int32_t)r1);
                sch_sl.push_back((int32_t)(sl_off.size() - 1));
            }
        }
        const int64_t stored = sl_off.back();
        std::vector<uint16_t> scol(stored);
        std::vector<float> sw(stored, 0.0f);
        const size_t nsl = sl_off.size() - 1;
        for (size_t sl = 0; sl < nsl; sl++) {
            // padding: column = the instance size (the zero slot), weight 0
            const int32_t r_any = sl_row[sl * 64];
            const int32_t ii = inst[r_any];
            const uint16_t pad = (uint16_t)b->sizes[ii];
            for (int64_t e = sl_off[sl]; e < sl_off[sl + 1]; e++) scol[e] = pad;
            for (int32_t l = 0; l < 64; l++) {
                const int32_t row = sl_row[sl * 64 + l];
                if (row < 0) continue;
                const int64_t lo = b->offsets[ii];
                for (int64_t j = row_ptr[row], k = 0; j < row_ptr[row + 1]; j++, k++) {
                    const int64_t pos = sl_off[sl] + (k / 4) * 256 + 4 * l + k % 4;
                    scol[pos] = (uint16_t)(col[j] - lo);
                    sw[pos] = wf[j];
                }
            }
        }
        B.n_schunks = (int32_t)(sch.size() / 3);
        B.sl_stored = stored;
        if ((rc = balloc(B.schunks, sch.size() * 4, sch.data()))) return rc;
        if ((rc = balloc(B.schunk_slices, sch_sl.size() * 4, sch_sl.data()))) return rc;
        if ((rc = balloc(B.sl_off, sl_off.size() * 8, sl_off.data()))) return rc;
        if ((rc = balloc(B.sl_row, sl_row.size() * 4, sl_row.data()))) return rc;
        if ((rc = balloc(B.sl_col, (size_t)std::max<int64_t>(stored, 1) * 2, stored ? scol.data() : nullptr))) return rc;
        if ((rc = balloc(B.sl_w, (size_t)std::max<int64_t>(stored, 1) * 4, stored ? sw.data() : nullptr))) return rc;
        B.mode = 2;
    }
    b->ready_sparse = true;
    return 0;
}

int ahip_batch_set_mode(ahip_batch* b, int32_t mode) {
    if (!b || mode < 0 || mode > 2)
        return bfail(AHIP_EINVAL, "mode must be 0 (global gathers), 1 (LDS) or 2 (sliced chunks)");
    if (mode == 1 && (size_t)b->B.max_n * sizeof(float) > 64 * 1024)
        return bfail(AHIP_EINVAL, "an instance is too large for the LDS kernel");
    if (mode == 2 && b->B.n_schunks == 0)
        return bfail(AHIP_EINVAL, "no sliced layout (an instance is larger than 16383 neurons)");
    BCHK(hipStreamSynchronize(b->stream));
    drop_graph(b);
    b->B.mode = mode;
    return 0;
}

int ahip_batch_get_mode(ahip_batch* b, int32_t* mode) {
    if (!b || !mode) return bfail(AHIP_EINVAL, "null argument");
    *mode = b->B.mode;
    return 0;
}

int ahip_batch_set_input(ahip_batch* b, const double* w_in, const double* phase, double omega) {
    if (!b || !w_in || !phase) return bfail(AHIP_EINVAL, "null argument");
    BatchDev& B = b->B;
    const int32_t n = B.n_in;
    // B operand W^T: element (k, m) of tile (mt, ks) at lane m%32 + 32*((k%16)/8), item k%8
    std::vector<uint16_t> hi((size_t)B.m_tiles * B.k_steps * 64 * 8, 0), lo(hi.size(), 0);
    for (int32_t mt = 0; mt < B.m_tiles; mt++)
        for (int32_t ks = 0; ks < B.k_steps; ks++)
            for (int lane = 0; lane < 64; lane++)
                for (int j = 0; j < 8; j++) {
                    const int32_t m = mt * 32 + lane % 32;
                    const int32_t k = ks * 16 + 8 * (lane / 32) + j;
                    float v = (m < n && k < n) ? (float)w_in[(size_t)m * n + k] : 0.0f;
                    const size_t pos = (((size_t)mt * B.k_steps + ks) * 64 + lane) * 8 + j;
                    hi[pos] = bf16_rne(v);
                    lo[pos] = bf16_rne(v - bf16_f(hi[pos]));
                }
    BCHK(hipStreamSynchronize(b->stream));
    BCHK(hipMemcpy(B.w_hi.p, hi.data(), hi.size() * 2, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(B.w_lo.p, lo.data(), lo.size() * 2, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(B.phase.p, phase, (size_t)n * B.n_inst * 8, hipMemcpyHostToDevice));
    B.omega = omega;
    drop_graph(b);
    b->ready_input = true;
    return 0;
}

int ahip_batch_set_gain(ahip_batch* b, const double* gain) {
    if (!b || !gain) return bfail(AHIP_EINVAL, "null argument");
    std::vector<float> g(b->B.n_inst);
    for (int32_t i = 0; i < b->B.n_inst; i++) g[i] = (float)gain[i];
    BCHK(hipStreamSynchronize(b->stream));
    BCHK(hipMemcpy(b->B.gain.p, g.data(), g.size() * 4, hipMemcpyHostToDevice));
    return 0;
}

int ahip_batch_set_rates(ahip_batch* b, const double* r) {
    if (!b || !r) return bfail(AHIP_EINVAL, "null argument");
    std::vector<float> f(b->B.total);
    for (int32_t i = 0; i < b->B.total; i++) f[i] = (float)r[i];
    BCHK(hipStreamSynchronize(b->stream));
    BCHK(hipMemcpy(b->B.r[b->parity].p, f.data(), f.size() * 4, hipMemcpyHostToDevice));
    return 0;
}

int ahip_batch_get_rates(ahip_batch* b, double* r) {
    if (!b || !r) return bfail(AHIP_EINVAL, "null argument");
    std::vector<float> f(b->B.total);
    BCHK(hipStreamSynchronize(b->stream));
    BCHK(hipMemcpy(f.data(), b->B.r[b->parity].p, f.size() * 4, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < b->B.total; i++) r[i] = f[i];
    return 0;
}

int ahip_batch_get_input_sums(ahip_batch* b, double* y) {
    if (!b || !y) return bfail(AHIP_EINVAL, "null argument");
    const BatchDev& B = b->B;
    if (b->last_chunk < 1) return bfail(AHIP_ESTATE, "no step has run");
    // the last step of the last chunk: rows [instance] x columns [input neuron]
    const int32_t ld = B.m_tiles * 32;
    const size_t block = (size_t)B.n_tiles * 32 * ld;
    std::vector<float> f(block);
    BCHK(hipStreamSynchronize(b->stream));
    BCHK(hipMemcpy(f.data(), B.y.as<float>() + (size_t)(b->last_chunk - 1) * block, block * 4,
                   hipMemcpyDeviceToHost));
    for (int32_t m = 0; m < B.n_in; m++)
        for (int32_t i = 0; i < B.n_inst; i++) y[(size_t)m * B.n_inst + i] = f[(size_t)i * ld + m];
    return 0;
}

int ahip_batch_run(ahip_batch* b, int64_t steps, int32_t graph_batch) {
    if (!b || steps < 0) return bfail(AHIP_EINVAL, "bad arguments");
    if (!b->ready_sparse || !b->ready_input) return bfail(AHIP_ESTATE, "set the recurrent and input projections first");
    int32_t G = graph_batch > 0 ? graph_batch : 50;
    if (G % 2) G++;  // even batches keep the ping-pong parity of the captured graph fixed
    // one input chunk per graph batch: X and Y for G steps
    int rc0 = ensure_chunk(b, G);
    if (rc0) return rc0;
    int64_t left = steps;
    if (left >= G) {
        if (!b->graph || b->graph_steps != G || b->graph_parity != b->parity) {
            drop_graph(b);
            BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_steps(b, G, b->parity);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) { if (g) hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return bfail(AHIP_EHIP, "capture: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) { b->graph = nullptr; return bfail(AHIP_EHIP, "instantiate: %s", hipGetErrorString(e)); }
            b->graph_steps = G;
            b->graph_parity = b->parity;
        }
        while (left >= G) {
            BCHK(hipGraphLaunch(b->graph, b->stream));
            left -= G;
        }
    }
    if (left > 0) {
        int rc = enqueue_steps(b, (int32_t)left, b->parity);
        if (rc) return rc;
        b->parity ^= (int32_t)(left & 1);
        drop_graph(b);
    }
    BCHK(hipStreamSynchronize(b->stream));
    b->t += steps;
    return 0;
}

// which: 0 = input update, 1 = GEMM, 2 = SpMV + update.  Mean launch time.
int ahip_batch_time(ahip_batch* b, int32_t which, int32_t iters, double* avg_ms) {
    if (!b || !avg_ms || iters <= 0 || which < 0 || which > 2) return bfail(AHIP_EINVAL, "bad arguments");
    if (!b->ready_sparse || !b->ready_input) return bfail(AHIP_ESTATE, "not ready");
    hipEvent_t e0, e1;
    BCHK(hipEventCreate(&e0));
    BCHK(hipEventCreate(&e1));
    // input and GEMM: one launch over a full chunk (cap_steps steps; every run
    // recomputes its chunks, so the buffers they overwrite carry no state)
    const int32_t n = b->B.cap_steps;
    auto one = [&]() -> int {
        if (which == 0) return batch_launch_input(b->B, 0, n, b->stream);
        if (which == 1) return batch_launch_gemm(b->B, n, b->stream);
        // reads the current rates, writes only the other buffer and the psp
        // record, both rewritten by the next step
        return batch_launch_spmv(b->B, b->parity, 0, b->stream);
    };
    BLCHK(one());
    BCHK(hipEventRecord(e0, b->stream));
    for (int32_t i = 0; i < iters; i++) BLCHK(one());
    BCHK(hipEventRecord(e1, b->stream));
    BCHK(hipEventSynchronize(e1));
    float ms = 0;
    BCHK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *avg_ms = (double)ms / iters;
    return 0;
}

int ahip_batch_chunk_steps(ahip_batch* b, int32_t* steps) {
    if (!b || !steps) return bfail(AHIP_EINVAL, "null argument");
    *steps = b->B.cap_steps;
    return 0;
}

int ahip_batch_info(ahip_batch* b, int64_t* nnz, int32_t* total, double* sparse_bytes,
                    double* dense_flops) {
    if (!b) return bfail(AHIP_EINVAL, "null batch");
    if (nnz) *nnz = b->nnz;
    if (total) *total = b->B.total;
    // algorithmic: nnz*(4+4) + rows*(4 + 4 + 4) (row_ptr as int32, r read, r written)
    if (sparse_bytes) *sparse_bytes = (double)b->nnz * 8.0 + (double)b->B.total * 12.0;
    if (dense_flops) *dense_flops = 2.0 * b->B.n_in * (double)b->B.n_in * b->B.n_inst;
    return 0;
}

}  // extern "C"
