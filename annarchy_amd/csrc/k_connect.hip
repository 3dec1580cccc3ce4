// k_connect.hip -- connectivity generated on the device.
//
// The reference draws connectivity on the host, row by row, from numpy
// (cython_ext/Connector.pyx:306-380): fixed_number_pre takes `number`
// distinct pre-synaptic ranks per post-synaptic neuron (rng.choice without
// replacement, redrawn while it contains the post rank when self-connections
// are not allowed), fixed_probability keeps each pre rank with probability p.
// Here the same distributions are drawn from a counter-based stream
// (Philox4x32-10), so a row depends only on (seed, row) and not on the
// launch shape; the random stream differs from numpy's.
//
// Row r's candidate k is word k%4 of philox(counter = (r, k/4, 0, tag)) mapped
// to [0, n_pre) by a 32x32 multiply-high.  fixed_number_pre keeps the first
// `number` distinct candidates that are not the post rank (when excluded),
// in draw order, then sorts them.  fixed_probability walks the row with
// geometric gaps floor(log(u) / log(1-p)), u from word 0,1 of philox(r, k,
// 1, tag) as a 53-bit uniform; columns come out ascending.  Weights: a
// constant, or Uniform(a, b) / Normal(a, b) of synapse k of row r drawn as
// the neuron-model draws (ahip::draw) with stream tag 0x57.
#include "ahip_device.hpp"
#include "internal.hpp"

namespace ahip {
namespace {

constexpr uint32_t kTagCand = 0x434f4e4eu;  // stream tag of the candidates
constexpr uint32_t kTagW = 0x57u;

__device__ inline uint32_t cand_word(uint64_t seed, int32_t row, int64_t k) {
    u32x4 c{(uint32_t)row, (uint32_t)(k >> 2), 0u, kTagCand};
    u32x4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const int w = (int)(k & 3);
    return w == 0 ? r.x : w == 1 ? r.y : w == 2 ? r.z : r.w;
}

__device__ inline double gap_uniform(uint64_t seed, int32_t row, int64_t k) {
    u32x4 c{(uint32_t)row, (uint32_t)k, 1u, kTagCand};
    u32x4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return 1.0 - u01_53(r.x, r.y);  // (0, 1]
}

__device__ inline void write_weight(double* w, int64_t at, int32_t w_kind, double a, double b, uint64_t seed,
                                    int32_t row, int32_t k) {
    if (w_kind == 0) return;
    w[at] = draw(seed, kTagW, row, k, 0, w_kind - 1, a, b);
}

// One wave (workgroup of 64) per post row.  LDS: hash keys / first-draw
// indices of TSZ entries; the selected values are compacted into `idx` and
// sorted there with a bitonic network.
template <int TSZ>
__global__ __launch_bounds__(kWave) void gen_fixed_number_pre(
    int32_t post_begin, int32_t post_end, int32_t pre_begin, int32_t n_pre, int32_t number, int32_t no_self,
    uint64_t seed, const int64_t* __restrict__ row_ptr, int32_t* __restrict__ col, double* __restrict__ w,
    int32_t w_kind, double wa, double wb) {
    __shared__ int32_t keys[TSZ];
    __shared__ int32_t idx[TSZ];
    const int lane = threadIdx.x;
    for (int32_t r = post_begin + blockIdx.x; r < post_end; r += gridDim.x) {
        for (int i = lane; i < TSZ; i += kWave) {
            keys[i] = -1;
            idx[i] = INT32_MAX;
        }
        __syncthreads();
        int32_t count = 0;
        int64_t last = -1;
        for (int64_t d0 = 0; count < number; d0 += kWave) {
            const int64_t d = d0 + lane;
            const uint32_t word = cand_word(seed, r, d);
            const int32_t c = pre_begin + (int32_t)(((uint64_t)word * (uint64_t)n_pre) >> 32);
            const bool valid = !(no_self && c == r);
            int32_t h = -1;
            if (valid) {
                h = (int32_t)((uint32_t)c * 2654435761u) & (TSZ - 1);
                for (;;) {
                    const int32_t old = atomicCAS(&keys[h], -1, c);
                    if (old == -1 || old == c) break;
                    h = (h + 1) & (TSZ - 1);
                }
                atomicMin(&idx[h], (int32_t)d);
            }
            __syncthreads();
            const bool first = valid && idx[h] == (int32_t)d;
            const uint64_t mask = __ballot(first);
            const int before = __popcll(mask & ((1ull << lane) - 1ull));
            const bool accept = first && count + before < number;
            const uint64_t acc = __ballot(accept);
            if (acc) last = d0 + 63 - __clzll((long long)acc);
            count += __popcll(acc);
            __syncthreads();
        }
        // compact the selected values (first draw <= last) into idx[0..number)
        int32_t pos = 0;
        for (int base = 0; base < TSZ; base += kWave) {
            const int32_t key = keys[base + lane];
            const int32_t di = idx[base + lane];
            const bool keep = key >= 0 && di <= last;
            const uint64_t m = __ballot(keep);
            const int before = __popcll(m & ((1ull << lane) - 1ull));
            __syncthreads();
            if (keep) idx[pos + before] = key;
            pos += __popcll(m);
            __syncthreads();
        }
        int32_t n2 = 1;
        while (n2 < number) n2 <<= 1;
        for (int i = number + lane; i < n2; i += kWave) idx[i] = INT32_MAX;
        __syncthreads();
        // bitonic sort of idx[0..n2)
        for (int32_t size = 2; size <= n2; size <<= 1) {
            for (int32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (int32_t i = lane; i < n2; i += kWave) {
                    const int32_t j = i ^ stride;
                    if (j > i) {
                        const bool up = (i & size) == 0;
                        const int32_t a = idx[i], b = idx[j];
                        if ((a > b) == up) {
                            idx[i] = b;
                            idx[j] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
        const int64_t base_out = row_ptr[r];
        for (int32_t k = lane; k < number; k += kWave) {
            col[base_out + k] = idx[k];
            write_weight(w, base_out + k, w_kind, wa, wb, seed, r, k);
        }
        __syncthreads();
    }
}

// fixed_probability: one thread per post row; pass 0 counts, pass 1 writes.
__global__ __launch_bounds__(256) void gen_fixed_probability(
    int32_t post_begin, int32_t post_end, int32_t pre_begin, int32_t n_pre, double p, int32_t no_self,
    uint64_t seed, int32_t pass, int64_t* __restrict__ counts, const int64_t* __restrict__ row_ptr,
    int32_t* __restrict__ col, double* __restrict__ w, int32_t w_kind, double wa, double wb) {
    const int32_t r = post_begin + (int32_t)(blockIdx.x * 256 + threadIdx.x);
    if (r >= post_end) return;
    int64_t n = 0;
    int64_t at = pass ? row_ptr[r] : 0;
    if (p >= 1.0) {
        for (int32_t j = 0; j < n_pre; j++) {
            const int32_t c = pre_begin + j;
            if (no_self && c == r) continue;
            if (pass) {
                col[at] = c;
                write_weight(w, at, w_kind, wa, wb, seed, r, (int32_t)n);
                at++;
            }
            n++;
        }
    } else if (p > 0.0) {
        const double lq = log1p(-p);
        int64_t j = -1;
        for (int64_t k = 0;; k++) {
            const double u = gap_uniform(seed, r, k);
            const double g = floor(log(u) / lq);
            if (g >= (double)n_pre) break;
            j += (int64_t)g + 1;
            if (j >= n_pre) break;
            const int32_t c = pre_begin + (int32_t)j;
            if (no_self && c == r) continue;
            if (pass) {
                col[at] = c;
                write_weight(w, at, w_kind, wa, wb, seed, r, (int32_t)n);
                at++;
            }
            n++;
        }
    }
    if (!pass) counts[r] = n;
}

}  // namespace

int launch_gen_fixed_number_pre(int32_t post_begin, int32_t post_end, int32_t pre_begin, int32_t n_pre,
                                int32_t number, int32_t no_self, uint64_t seed, const int64_t* row_ptr,
                                int32_t* col, double* w, int32_t w_kind, double wa, double wb, hipStream_t s) {
    const int32_t rows = post_end - post_begin;
    if (rows <= 0 || number <= 0) return 0;
    const unsigned nb = (unsigned)std::min<int32_t>(rows, 16384);
    // the hash holds at most number + 63 distinct values: keep it at most
    // half full
    int tsz = 64;
    while (tsz < 2 * number + kWave) tsz <<= 1;
#define AHIP_GEN_CASE(S)                                                                                  \
    case S:                                                                                               \
        hipLaunchKernelGGL((gen_fixed_number_pre<S>), dim3(nb), dim3(kWave), 0, s, post_begin, post_end, \
                           pre_begin, n_pre, number, no_self, seed, row_ptr, col, w, w_kind, wa, wb);    \
        break;
    switch (tsz) {
        AHIP_GEN_CASE(256)
        AHIP_GEN_CASE(512)
        AHIP_GEN_CASE(1024)
        AHIP_GEN_CASE(2048)
        AHIP_GEN_CASE(4096)
        case 64:
        case 128:
            hipLaunchKernelGGL((gen_fixed_number_pre<256>), dim3(nb), dim3(kWave), 0, s, post_begin, post_end,
                               pre_begin, n_pre, number, no_self, seed, row_ptr, col, w, w_kind, wa, wb);
            break;
        default:
            return (int)hipErrorInvalidValue;
    }
#undef AHIP_GEN_CASE
    return (int)hipGetLastError();
}

int launch_gen_fixed_probability(int32_t post_begin, int32_t post_end, int32_t pre_begin, int32_t n_pre,
                                 double p, int32_t no_self, uint64_t seed, int32_t pass, int64_t* counts,
                                 const int64_t* row_ptr, int32_t* col, double* w, int32_t w_kind, double wa,
                                 double wb, hipStream_t s) {
    const int32_t rows = post_end - post_begin;
    if (rows <= 0) return 0;
    const unsigned nb = (unsigned)((rows + 255) / 256);
    hipLaunchKernelGGL(gen_fixed_probability, dim3(nb), dim3(256), 0, s, post_begin, post_end, pre_begin, n_pre,
                       p, no_self, seed, pass, counts, row_ptr, col, w, w_kind, wa, wb);
    return (int)hipGetLastError();
}

}  // namespace ahip
