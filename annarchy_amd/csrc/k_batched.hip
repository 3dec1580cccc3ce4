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
// One step = three kernels captured in a graph:
//   1. input update: X[k][i] for all instances, split in BF16 hi + lo
//   2. Y = W_in . X on MFMA (v_mfma_f32_32x32x8_bf16), FP32 accumulation;
//      W and X are each split in two BF16 terms and the three significant
//      products (hi.hi + hi.lo + lo.hi) are accumulated, which gives ~16
//      mantissa bits per operand instead of 8
//   3. block-diagonal SpMV (FP32, compensated) fused with the exc update
//      (rates ping-pong between two buffers, so every row reads step-t rates)
#include <hip/hip_bfloat16.h>

#include "ahip_device.hpp"
#include "internal.hpp"

namespace ahip {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));



namespace {

__device__ inline short bf16_bits(float f) {
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    uint32_t r = ((u >> 16) & 1u) + 0x7FFFu;
    return (short)((u + r) >> 16);
}
__device__ inline float bf16_to_f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

// X[k][i] = 1 + 0.5*sin(omega*(t*dt) + phase[k][i]), stored as B fragments:
// element (k, n) of tile (ntile, kstep) sits at lane = n%32 + 32*((k%8)/4),
// register (k%4).
__global__ __launch_bounds__(256) void batch_input(int32_t n_in, int32_t n_inst, int32_t n_tiles,
                                                   int32_t k_steps, const double* __restrict__ phase,
                                                   double omega, double dt, const int64_t* __restrict__ t_base,
                                                   int32_t step, short* __restrict__ xh,
                                                   short* __restrict__ xl) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t total = (int64_t)k_steps * 8 * n_tiles * 32;
    if (idx >= total) return;
    const int32_t n = (int32_t)(idx % (n_tiles * 32));
    const int32_t k = (int32_t)(idx / (n_tiles * 32));
    float v = 0.0f;
    if (k < n_in && n < n_inst) {
        const double t = (double)(*t_base + step);
        v = (float)(1.0 + 0.5 * sin(omega * (t * dt) + phase[(int64_t)k * n_inst + n]));
    }
    const short h = bf16_bits(v);
    const short l = bf16_bits(v - bf16_to_f(h));
    const int32_t nt = n / 32, ks = k / 8, kk = k % 8;
    const int32_t lane = (n % 32) + 32 * (kk / 4);
    const int64_t pos = (((int64_t)nt * k_steps + ks) * 64 + lane) * 4 + (kk % 4);
    xh[pos] = h;
    xl[pos] = l;
}

// One wave per 32 x 32 tile of Y; 4 waves per block take 4 column tiles.
__global__ __launch_bounds__(256) void batch_gemm(int32_t m_tiles, int32_t n_tiles, int32_t k_steps,
                                                  const v4s* __restrict__ wh, const v4s* __restrict__ wl,
                                                  const v4s* __restrict__ xh, const v4s* __restrict__ xl,
                                                  float* __restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int32_t wave = blockIdx.x * 4 + threadIdx.x / 64;
    if (wave >= m_tiles * n_tiles) return;
    const int32_t mt = wave / n_tiles, nt = wave % n_tiles;
    v16f acc = {0};
    const v4s* a_h = wh + (int64_t)mt * k_steps * 64 + lane;
    const v4s* a_l = wl + (int64_t)mt * k_steps * 64 + lane;
    const v4s* b_h = xh + (int64_t)nt * k_steps * 64 + lane;
    const v4s* b_l = xl + (int64_t)nt * k_steps * 64 + lane;
    for (int32_t ks = 0; ks < k_steps; ks++) {
        const v4s ah = a_h[ks * 64], al = a_l[ks * 64];
        const v4s bh = b_h[ks * 64], bl = b_l[ks * 64];
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(al, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bh, acc, 0, 0, 0);
    }
    const int32_t ld = n_tiles * 32;
    const int32_t col = nt * 32 + (lane % 32);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int32_t row = mt * 32 + 8 * (i / 4) + 4 * (lane / 32) + (i % 4);
        y[(int64_t)row * ld + col] = acc[i];
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
        const float s_in = n < n_in ? y[(int64_t)n * ld_y + inst] : 0.0f;
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
            const float s_in = nl < n_in ? y[(int64_t)nl * ld_y + inst] : 0.0f;
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
            const float s_in = nl < n_in ? y[(int64_t)nl * ld_y + inst] : 0.0f;
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

int batch_launch_input(const BatchDev& B, int32_t step, hipStream_t s) {
    const int64_t total = (int64_t)B.k_steps * 8 * B.n_tiles * 32;
    const int64_t nb = (total + 255) / 256;
    hipLaunchKernelGGL(batch_input, dim3(nb), dim3(256), 0, s, B.n_in, B.n_inst, B.n_tiles, B.k_steps,
                       B.phase.as<double>(), B.omega, B.dt, B.t_dev.as<int64_t>(), step,
                       B.x_hi.as<short>(), B.x_lo.as<short>());
    return (int)hipGetLastError();
}

int batch_launch_gemm(const BatchDev& B, hipStream_t s) {
    const int64_t waves = (int64_t)B.m_tiles * B.n_tiles;
    const int64_t nb = (waves + 3) / 4;
    hipLaunchKernelGGL(batch_gemm, dim3(nb), dim3(256), 0, s, B.m_tiles, B.n_tiles, B.k_steps,
                       B.w_hi.as<v4s>(), B.w_lo.as<v4s>(), B.x_hi.as<v4s>(), B.x_lo.as<v4s>(),
                       B.y.as<float>());
    return (int)hipGetLastError();
}

int batch_launch_spmv(const BatchDev& B, int32_t parity, hipStream_t s) {
    const int64_t threads = (int64_t)B.total * B.lanes;
    const int64_t nb = (threads + 255) / 256;
    const float* rin = B.r[parity].as<float>();
    float* rout = B.r[parity ^ 1].as<float>();
    if (B.mode == 2) {
        const size_t lds = (size_t)(B.max_n + 1) * sizeof(float);
        hipLaunchKernelGGL(batch_sell<2>, dim3(B.n_schunks), dim3(512), lds, s, B.schunks.as<int32_t>(),
                           B.schunk_slices.as<int32_t>(), B.inst_base.as<int64_t>(), B.sl_off.as<int64_t>(),
                           B.sl_row.as<int32_t>(), B.sl_col.as<uint16_t>(), B.sl_w.as<float>(), B.gain.as<float>(),
                           B.y.as<float>(), B.n_tiles * 32, B.n_in, rin, rout, B.sum_exc.as<float>(), (float)B.dt,
                           (float)B.tau);
        return (int)hipGetLastError();
    }
    if (B.mode == 1) {
        const size_t lds = (size_t)B.max_n * sizeof(float);
#define AHIP_BLDS_CASE(N)                                                                         \
    case N:                                                                                       \
        hipLaunchKernelGGL(batch_spmv_lds<N>, dim3(B.n_chunks), dim3(256), lds, s,                \
                           B.chunks.as<int32_t>(), B.inst_base.as<int64_t>(), B.row_ptr.as<int64_t>(), \
                           B.col.as<int32_t>(), B.w.as<float>(), B.gain.as<float>(), B.y.as<float>(), \
                           B.n_tiles * 32, B.n_in, rin, rout, B.sum_exc.as<float>(), (float)B.dt,     \
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
                           B.gain.as<float>(), B.y.as<float>(), B.n_tiles * 32, B.n_in, rin, rout, \
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
