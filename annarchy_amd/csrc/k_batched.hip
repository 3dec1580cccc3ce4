// k_batched.hip -- many rate-network instances of uneven size stepped together.
//
// Replaces the reference's multi-instance path (core/Network.py:751-926,
// parallel_run: one process and one compiled network per instance) for the
// parameter-sweep family
//     input population (n_in neurons per instance):  r = 1 + 0.5*sin(omega*t*dt + phase)
//     exc population (N_i neurons):  tau dr/dt = -r + sum(exc) + gain_i*sum(inp), r >= 0
//     exc -> exc  sparse (block-diagonal CSR over all instances, FP32 + Kahan)
//     inp -> exc  dense n_in x n_in weights shared by every instance, applied to
//                 the first min(N_i, n_in) neurons of instance i
// The input rates depend on t only, so the input projection of a whole
// chunk of steps (a captured graph batch) is one GEMM:
//   1. input: X of every step of the chunk, split in BF16 hi + lo;
//   2. Y^T = X^T W^T on MFMA (v_mfma_f32_32x32x16_bf16), FP32 accumulation;
//      W and X are each split in two BF16 terms and the three significant
//      products (hi.hi + hi.lo + lo.hi) are accumulated, which gives ~16
//      mantissa bits per operand instead of 8;
//   3. per step: block-diagonal SpMV (FP32, compensated) fused with the exc
//      update (rates ping-pong between two buffers, so every row reads
//      step-t rates), reading the step's row block of Y^T.
#include <hip/hip_bfloat16.h>

#include "ahip_device.hpp"
#include "internal.hpp"

namespace ahip {

typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));



namespace {

__device__ inline short bf16_bits(float f) {
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    uint32_t r = ((u >> 16) & 1u) + 0x7FFFu;
    return (short)((u + r) >> 16);
}
__device__ inline float bf16_to_f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

// Input rates of `n` consecutive steps as split-BF16 A fragments of X^T.
// Step j of a chunk that starts at step `step0` of the current batch reads
// the input population's rates of the step before it,
//   X_j[k][i] = 1 + 0.5*sin(omega*((t - 1)*dt) + phase[k][i]),  t = t_base + step0 + j,
// and zero for t = 0 (the input rates start at 0 before the first update).
// Row c of X^T is (step j, instance i) = (c / NP, c % NP); fragment element
// (c, k) of tile (ct, ks) sits at lane c%32 + 32*((k%16)/8), item k%8 (the
// 32x32x16 BF16 operand layout).  One thread writes one lane's 8 items.
__global__ __launch_bounds__(256) void batch_input(int32_t n_in, int32_t n_inst, int32_t np, int32_t c_tiles,
                                                   int32_t k_steps, const double* __restrict__ phase,
                                                   double omega, double dt, const int64_t* __restrict__ t_base,
                                                   int32_t step0, v8s* __restrict__ xh, v8s* __restrict__ xl) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (int64_t)c_tiles * k_steps * 64) return;
    const int32_t lane = (int32_t)(g & 63);
    const int64_t tile = g >> 6;
    const int32_t ks = (int32_t)(tile % k_steps), ct = (int32_t)(tile / k_steps);
    const int32_t c = ct * 32 + (lane & 31);
    const int32_t j = c / np, i = c % np;
    const int32_t k0 = ks * 16 + 8 * (lane >> 5);
    const int64_t tprev = *t_base + step0 + j - 1;
    v8s h, l;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int32_t k = k0 + e;
        float v = 0.0f;
        if (k < n_in && i < n_inst && tprev >= 0)
            v = (float)(1.0 + 0.5 * sin(omega * ((double)tprev * dt) + phase[(int64_t)k * n_inst + i]));
        h[e] = bf16_bits(v);
        l[e] = bf16_bits(v - bf16_to_f(h[e]));
    }
    xh[g] = h;
    xl[g] = l;
}

// C = X^T W^T for a chunk of steps (C[c][m] = sum_k X[k][c] W[m][k]), on
// v_mfma_f32_32x32x16_bf16 with FP32 accumulation.  Each operand is split in
// two BF16 terms and the three significant products lo.hi + hi.lo + hi.hi
// are accumulated (about 16 mantissa bits per operand).  A wave owns
// GM x GN tiles of 32 x 32 (rows: steps x instances, columns: input
// neurons); the fragments of the next k step are loaded while the current
// one is multiplied.  Tiles past the edge load a valid tile and store nothing.
template <int GM, int GN>
__global__ __launch_bounds__(256) void batch_gemm(int32_t c_tiles, int32_t m_tiles, int32_t k_steps,
                                                  const v8bf* __restrict__ xh, const v8bf* __restrict__ xl,
                                                  const v8bf* __restrict__ wh, const v8bf* __restrict__ wl,
                                                  float* __restrict__ y, int32_t ldy) {
    const int lane = threadIdx.x & 63;
    const int32_t mblocks = (m_tiles + GN - 1) / GN;
    const int32_t wave = blockIdx.x * 4 + threadIdx.x / 64;
    const int32_t cb = wave / mblocks, mb = wave % mblocks;
    if (cb * GM >= c_tiles) return;
    const v8bf* ah[GM];
    const v8bf* al[GM];
    const v8bf* bh[GN];
    const v8bf* bl[GN];
#pragma unroll
    for (int a = 0; a < GM; a++) {
        const int32_t ct = min(cb * GM + a, c_tiles - 1);
        ah[a] = xh + (int64_t)ct * k_steps * 64 + lane;
        al[a] = xl + (int64_t)ct * k_steps * 64 + lane;
    }
#pragma unroll
    for (int b = 0; b < GN; b++) {
        const int32_t mt = min(mb * GN + b, m_tiles - 1);
        bh[b] = wh + (int64_t)mt * k_steps * 64 + lane;
        bl[b] = wl + (int64_t)mt * k_steps * 64 + lane;
    }
    v16f acc[GM][GN];
#pragma unroll
    for (int a = 0; a < GM; a++)
#pragma unroll
        for (int b = 0; b < GN; b++) acc[a][b] = v16f{0};
    v8bf fah[GM], fal[GM], fbh[GN], fbl[GN];
#pragma unroll
    for (int a = 0; a < GM; a++) { fah[a] = ah[a][0]; fal[a] = al[a][0]; }
#pragma unroll
    for (int b = 0; b < GN; b++) { fbh[b] = bh[b][0]; fbl[b] = bl[b][0]; }
    for (int32_t ks = 0; ks < k_steps; ks++) {
        const int64_t nx = (int64_t)min(ks + 1, k_steps - 1) * 64;
        v8bf nah[GM], nal[GM], nbh[GN], nbl[GN];
#pragma unroll
        for (int a = 0; a < GM; a++) { nah[a] = ah[a][nx]; nal[a] = al[a][nx]; }
#pragma unroll
        for (int b = 0; b < GN; b++) { nbh[b] = bh[b][nx]; nbl[b] = bl[b][nx]; }
#pragma unroll
        for (int a = 0; a < GM; a++)
#pragma unroll
            for (int b = 0; b < GN; b++) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fal[a], fbh[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fah[a], fbl[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fah[a], fbh[b], acc[a][b], 0, 0, 0);
            }
#pragma unroll
        for (int a = 0; a < GM; a++) { fah[a] = nah[a]; fal[a] = nal[a]; }
#pragma unroll
        for (int b = 0; b < GN; b++) { fbh[b] = nbh[b]; fbl[b] = nbl[b]; }
    }
    // D[i][j] of a 32 x 32 tile: lane j + 32*((i/4)%2), item (i%4) + 4*(i/8)
#pragma unroll
    for (int a = 0; a < GM; a++) {
        const int32_t ct = cb * GM + a;
        if (ct >= c_tiles) continue;
#pragma unroll
        for (int b = 0; b < GN; b++) {
            const int32_t mt = mb * GN + b;
            if (mt >= m_tiles) continue;
            const int32_t col = mt * 32 + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int32_t row = ct * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
                y[(int64_t)row * ldy + col] = acc[a][b][r];
            }
        }
    }
}

// Block-diagonal SpMV (L lanes per row, FP32 compensated) fused with the
// update of the exc neurons:  _r = (sum_exc + gain*sum_inp - r)/tau;
// r += dt*_r; r = max(r, 0).
template <int L>
__global__ __launch_bounds__(256) void batch_spmv_update(
    int32_t total, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const float* __restrict__ w, const int32_t* __restrict__ inst_of_row,
    const int32_t* __restrict__ local_of_row, const float* __restrict__ gain,
    const float* __restrict__ y, int32_t ld_y, int32_t n_in, const float* __restrict__ r_in,
    float* __restrict__ r_out, float* __restrict__ sum_out, float dt, float tau) {
    const int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) / L;
    const int lane = threadIdx.x & (L - 1);
    if (row >= total) return;
    const int64_t b = row_ptr[row], e = row_ptr[row + 1];
    Kahan k0{0.f, 0.f}, k1{0.f, 0.f};
    int64_t j = b + lane;
    for (; j + L < e; j += 2 * L) {
        const int32_t c0 = __builtin_nontemporal_load(col + j);
        const int32_t c1 = __builtin_nontemporal_load(col + j + L);
        const float w0 = __builtin_nontemporal_load(w + j);
        const float w1 = __builtin_nontemporal_load(w + j + L);
        k0.add(w0 * r_in[c0]);
        k1.add(w1 * r_in[c1]);
    }
    if (j < e) k0.add(__builtin_nontemporal_load(w + j) * r_in[__builtin_nontemporal_load(col + j)]);
    k0.add(k1.value());
    const float s = group_sum_kahan<L>(k0.s, k0.c);
    if (lane == 0) {
        const int32_t inst = inst_of_row[row];
        const int32_t n = local_of_row[row];
        const float s_in = n < n_in ? y[(int64_t)inst * ld_y + n] : 0.0f;
        const float r = r_in[row];
        const float _r = (s + gain[inst] * s_in - r) / tau;
        float rn = r + dt * _r;
        if (rn < 0.0f) rn = 0.0f;
        r_out[row] = rn;
        sum_out[row] = s;
    }
}

// Same step, one workgroup per chunk of rows of one instance: the instance's
// rates (N_i <= 16384 floats) are staged in LDS once, the chunk's synapses
// are read with 16-byte loads (4 per lane per chunk, head/tail scalar) and
// every gather is served from LDS.  chunks[3*c] = (instance, first row, end row).
typedef int v4i_b __attribute__((ext_vector_type(4)));
typedef float v4f_b __attribute__((ext_vector_type(4)));

template <int L>
__global__ __launch_bounds__(256) void batch_spmv_lds(
    const int32_t* __restrict__ chunks, const int64_t* __restrict__ inst_base,
    const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col, const float* __restrict__ w,
    const float* __restrict__ gain, const float* __restrict__ y, int32_t ld_y, int32_t n_in,
    const float* __restrict__ r_in, float* __restrict__ r_out, float* __restrict__ sum_out, float dt,
    float tau) {
    extern __shared__ float xs[];
    const int32_t inst = chunks[3 * blockIdx.x];
    const int32_t r0 = chunks[3 * blockIdx.x + 1], r1 = chunks[3 * blockIdx.x + 2];
    const int64_t base = inst_base[inst];
    const int32_t n = (int32_t)(inst_base[inst + 1] - base);
    for (int32_t k = threadIdx.x; k < n; k += 256) xs[k] = r_in[base + k];
    __syncthreads();
    const int lane = threadIdx.x & (L - 1);
    const int grp = threadIdx.x / L;
    constexpr int GROUPS = 256 / L;
    for (int32_t row = r0 + grp; row < r1; row += GROUPS) {
        const int64_t beg = row_ptr[row], end = row_ptr[row + 1];
        Kahan k0{0.f, 0.f}, k1{0.f, 0.f};
        const int64_t body = min(end, (beg + 3) & ~(int64_t)3);
        if (beg + lane < body) k0.add(w[beg + lane] * xs[col[beg + lane] - base]);
        int64_t j = body + 4 * lane;
        for (; j + 4 <= end; j += 4 * L) {
            const v4i_b c = __builtin_nontemporal_load(reinterpret_cast<const v4i_b*>(col + j));
            const v4f_b v = __builtin_nontemporal_load(reinterpret_cast<const v4f_b*>(w + j));
            k0.add(v.x * xs[c.x - base]);
            k1.add(v.y * xs[c.y - base]);
            k0.add(v.z * xs[c.z - base]);
            k1.add(v.w * xs[c.w - base]);
        }
        for (; j < end; j++) k0.add(w[j] * xs[col[j] - base]);
        k0.add(k1.value());
        const float s = group_sum_kahan<L>(k0.s, k0.c);
        if (lane == 0) {
            const int32_t nl = (int32_t)(row - base);
            const float s_in = nl < n_in ? y[(int64_t)inst * ld_y + nl] : 0.0f;
            const float r = xs[nl];
            const float _r = (s + gain[inst] * s_in - r) / tau;
            float rn = r + dt * _r;
            if (rn < 0.0f) rn = 0.0f;
            r_out[row] = rn;
            sum_out[row] = s;
        }
    }
}

// Sliced chunks (mode 2): a chunk holds up to 512 rows of one instance,
// sorted by length and packed into slices of 64 rows, one lane per row.  Step
// m of lane l is the 4 synapses at off + 256*m + 4*l: columns as uint16
// relative to the instance (padding points at the zero slot xs[n]), weights
// as one 16-byte load.  With U steps in flight and no cross-lane reduction,
// the lane that owns a row finishes its Kahan sum and applies the update.
typedef unsigned short v4u16_b __attribute__((ext_vector_type(4)));

template <int U>
__global__ __launch_bounds__(512) void batch_sell(
    const int32_t* __restrict__ chunks, const int32_t* __restrict__ chunk_slices,
    const int64_t* __restrict__ inst_base, const int64_t* __restrict__ off, const int32_t* __restrict__ slice_row,
    const uint16_t* __restrict__ col, const float* __restrict__ w, const float* __restrict__ gain,
    const float* __restrict__ y, int32_t ld_y, int32_t n_in, const float* __restrict__ r_in,
    float* __restrict__ r_out, float* __restrict__ sum_out, float dt, float tau) {
    extern __shared__ float xs[];
    const int32_t inst = chunks[3 * blockIdx.x];
    const int64_t base = inst_base[inst];
    const int32_t n = (int32_t)(inst_base[inst + 1] - base);
    for (int32_t k = threadIdx.x; k < n; k += 512) xs[k] = r_in[base + k];
    if (threadIdx.x == 0) xs[n] = 0.0f;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int32_t s_end = chunk_slices[blockIdx.x + 1];
    for (int32_t sl = chunk_slices[blockIdx.x] + wave; sl < s_end; sl += 8) {
        const int64_t b = off[sl];
        const int64_t nsteps = (off[sl + 1] - b) >> 8;
        const uint16_t* cp = col + b + 4 * lane;
        const float* wp = w + b + 4 * lane;
        Kahan k0{0.f, 0.f}, k1{0.f, 0.f};
        int64_t m = 0;
        for (; m + U <= nsteps; m += U) {
            v4u16_b c[U];
            v4f_b v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                c[u] = __builtin_nontemporal_load(reinterpret_cast<const v4u16_b*>(cp + (m + u) * 256));
                v[u] = __builtin_nontemporal_load(reinterpret_cast<const v4f_b*>(wp + (m + u) * 256));
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                k0.add(v[u].x * xs[c[u].x]);
                k1.add(v[u].y * xs[c[u].y]);
                k0.add(v[u].z * xs[c[u].z]);
                k1.add(v[u].w * xs[c[u].w]);
            }
        }
        for (; m < nsteps; m++) {
            const v4u16_b c = __builtin_nontemporal_load(reinterpret_cast<const v4u16_b*>(cp + m * 256));
            const v4f_b v = __builtin_nontemporal_load(reinterpret_cast<const v4f_b*>(wp + m * 256));
            k0.add(v.x * xs[c.x]);
            k1.add(v.y * xs[c.y]);
            k0.add(v.z * xs[c.z]);
            k1.add(v.w * xs[c.w]);
        }
        k0.add(k1.value());
        const int32_t row = slice_row[(int64_t)sl * kWave + lane];
        if (row >= 0) {
            const float sum = k0.value();
            const int32_t nl = (int32_t)(row - base);
            const float s_in = nl < n_in ? y[(int64_t)inst * ld_y + nl] : 0.0f;
            const float r = xs[nl];
            const float _r = (sum + gain[inst] * s_in - r) / tau;
            float rn = r + dt * _r;
            if (rn < 0.0f) rn = 0.0f;
            r_out[row] = rn;
            sum_out[row] = sum;
        }
    }
}

__global__ void batch_advance(int64_t* t, int64_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *t += n;
}

}  // namespace

int batch_launch_input(const BatchDev& B, int32_t step0, int32_t n, hipStream_t s) {
    if (n < 1 || n > B.cap_steps) return (int)hipErrorInvalidValue;
    const int32_t c_tiles = n * B.n_tiles;
    const int64_t nb = ((int64_t)c_tiles * B.k_steps * 64 + 255) / 256;
    hipLaunchKernelGGL(batch_input, dim3(nb), dim3(256), 0, s, B.n_in, B.n_inst, B.n_tiles * 32, c_tiles,
                       B.k_steps, B.phase.as<double>(), B.omega, B.dt, B.t_dev.as<int64_t>(), step0,
                       B.x_hi.as<v8s>(), B.x_lo.as<v8s>());
    return (int)hipGetLastError();
}

constexpr int kGemmGM = 2, kGemmGN = 4;

int batch_launch_gemm(const BatchDev& B, int32_t n, hipStream_t s) {
    if (n < 1 || n > B.cap_steps) return (int)hipErrorInvalidValue;
    const int32_t c_tiles = n * B.n_tiles;
    const int64_t waves = (int64_t)((c_tiles + kGemmGM - 1) / kGemmGM) * ((B.m_tiles + kGemmGN - 1) / kGemmGN);
    const int64_t nb = (waves + 3) / 4;
    hipLaunchKernelGGL((batch_gemm<kGemmGM, kGemmGN>), dim3(nb), dim3(256), 0, s, c_tiles, B.m_tiles, B.k_steps,
                       B.x_hi.as<v8bf>(), B.x_lo.as<v8bf>(), B.w_hi.as<v8bf>(), B.w_lo.as<v8bf>(),
                       B.y.as<float>(), B.m_tiles * 32);
    return (int)hipGetLastError();
}

int batch_launch_spmv(const BatchDev& B, int32_t parity, int32_t j, hipStream_t s) {
    if (j < 0 || j >= B.cap_steps) return (int)hipErrorInvalidValue;
    // row block of step j in Y^T: [instance][input neuron], leading dimension m_tiles*32
    const int32_t ld_y = B.m_tiles * 32;
    const float* yj = B.y.as<float>() + (int64_t)j * B.n_tiles * 32 * ld_y;
    const int64_t threads = (int64_t)B.total * B.lanes;
    const int64_t nb = (threads + 255) / 256;
    const float* rin = B.r[parity].as<float>();
    float* rout = B.r[parity ^ 1].as<float>();
    if (B.mode == 2) {
        const size_t lds = (size_t)(B.max_n + 1) * sizeof(float);
        hipLaunchKernelGGL(batch_sell<2>, dim3(B.n_schunks), dim3(512), lds, s, B.schunks.as<int32_t>(),
                           B.schunk_slices.as<int32_t>(), B.inst_base.as<int64_t>(), B.sl_off.as<int64_t>(),
                           B.sl_row.as<int32_t>(), B.sl_col.as<uint16_t>(), B.sl_w.as<float>(), B.gain.as<float>(),
                           yj, ld_y, B.n_in, rin, rout, B.sum_exc.as<float>(), (float)B.dt,
                           (float)B.tau);
        return (int)hipGetLastError();
    }
    if (B.mode == 1) {
        const size_t lds = (size_t)B.max_n * sizeof(float);
#define AHIP_BLDS_CASE(N)                                                                         \
    case N:                                                                                       \
        hipLaunchKernelGGL(batch_spmv_lds<N>, dim3(B.n_chunks), dim3(256), lds, s,                \
                           B.chunks.as<int32_t>(), B.inst_base.as<int64_t>(), B.row_ptr.as<int64_t>(), \
                           B.col.as<int32_t>(), B.w.as<float>(), B.gain.as<float>(), yj, \
                           ld_y, B.n_in, rin, rout, B.sum_exc.as<float>(), (float)B.dt,     \
                           (float)B.tau);                                                          \
        break;
        switch (B.lanes) {
            AHIP_BLDS_CASE(4)
            AHIP_BLDS_CASE(8)
            AHIP_BLDS_CASE(16)
            AHIP_BLDS_CASE(32)
            AHIP_BLDS_CASE(64)
            default:
                return (int)hipErrorInvalidValue;
        }
#undef AHIP_BLDS_CASE
        return (int)hipGetLastError();
    }
#define AHIP_BATCH_CASE(N)                                                                         \
    case N:                                                                                        \
        hipLaunchKernelGGL(batch_spmv_update<N>, dim3(nb), dim3(256), 0, s, B.total,               \
                           B.row_ptr.as<int64_t>(), B.col.as<int32_t>(), B.w.as<float>(),          \
                           B.inst_of_row.as<int32_t>(), B.local_of_row.as<int32_t>(),             \
                           B.gain.as<float>(), yj, ld_y, B.n_in, rin, rout, \
                           B.sum_exc.as<float>(), (float)B.dt, (float)B.tau);                      \
        break;
    switch (B.lanes) {
        AHIP_BATCH_CASE(4)
        AHIP_BATCH_CASE(8)
        AHIP_BATCH_CASE(16)
        AHIP_BATCH_CASE(32)
        AHIP_BATCH_CASE(64)
        default:
            return (int)hipErrorInvalidValue;
    }
#undef AHIP_BATCH_CASE
    return (int)hipGetLastError();
}

int batch_launch_advance(const BatchDev& B, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(batch_advance, dim3(1), dim3(64), 0, s, B.t_dev.as<int64_t>(), n);
    return (int)hipGetLastError();
}

}  // namespace ahip
