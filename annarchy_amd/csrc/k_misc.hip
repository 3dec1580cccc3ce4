// k_misc.hip -- small kernels of the step: fills, time advance, recording.
// Fills are kernels (not memset nodes) so that they are safe inside a captured
// graph that is replayed many times.
#include "ahip_device.hpp"
#include "internal.hpp"

namespace ahip {
namespace {

template <typename T>
__global__ void fill_kernel(T* __restrict__ dst, int64_t n, T v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

__global__ void advance_time(int64_t* t, int64_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *t += n;
}

template <typename T>
__global__ void record_values(const T* __restrict__ src, int32_t size, T* __restrict__ dst,
                              const int64_t* __restrict__ t_base, int32_t step, int64_t t_start,
                              int32_t period, int64_t base, int64_t capacity) {
    // every period steps from t_start on, into rows base, base + 1, ...
    const int64_t d = *t_base + step - t_start;
    if (d < 0 || d % period != 0) return;
    const int64_t slot = base + d / period;
    if (slot >= capacity) return;
    T* out = dst + slot * (int64_t)size;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < size; i += gridDim.x * blockDim.x)
        out[i] = src[i];
}

template <typename S, typename D>
__global__ void convert_kernel(const S* __restrict__ src, D* __restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (D)src[i];
}

inline int grid_for(int64_t n) {
    int64_t nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    if (nb < 1) nb = 1;
    return (int)nb;
}

}  // namespace

int launch_fill(void* dst, int64_t n, int32_t elem, double value, hipStream_t s) {
    if (n <= 0) return 0;
    if (elem == 8)
        hipLaunchKernelGGL(fill_kernel<double>, dim3(grid_for(n)), dim3(256), 0, s, (double*)dst, n,
                           value);
    else
        hipLaunchKernelGGL(fill_kernel<float>, dim3(grid_for(n)), dim3(256), 0, s, (float*)dst, n,
                           (float)value);
    return (int)hipGetLastError();
}

int launch_fill_i32(int32_t* dst, int64_t n, int32_t value, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(fill_kernel<int32_t>, dim3(grid_for(n)), dim3(256), 0, s, dst, n, value);
    return (int)hipGetLastError();
}

int launch_fill_i64(int64_t* dst, int64_t n, int64_t value, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(fill_kernel<int64_t>, dim3(grid_for(n)), dim3(256), 0, s, dst, n, value);
    return (int)hipGetLastError();
}

int launch_advance_time(int64_t* t, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(advance_time, dim3(1), dim3(64), 0, s, t, n);
    return (int)hipGetLastError();
}

int launch_record_values(const Monitor& m, const void* src, int32_t size, int32_t elem,
                         const int64_t* t_base, int32_t step, hipStream_t s) {
    const int nb = grid_for(size);
    if (elem == 8)
        hipLaunchKernelGGL(record_values<double>, dim3(nb), dim3(256), 0, s, (const double*)src,
                           size, m.data.as<double>(), t_base, step, m.t_start, m.period, m.seg_base,
                           m.capacity);
    else
        hipLaunchKernelGGL(record_values<float>, dim3(nb), dim3(256), 0, s, (const float*)src,
                           size, m.data.as<float>(), t_base, step, m.t_start, m.period, m.seg_base,
                           m.capacity);
    return (int)hipGetLastError();
}

int launch_convert(const void* src, int32_t src_f64, void* dst, int32_t dst_f64, int64_t n,
                   hipStream_t s) {
    if (n <= 0) return 0;
    const int nb = grid_for(n);
    if (src_f64 && dst_f64)
        hipLaunchKernelGGL((convert_kernel<double, double>), dim3(nb), dim3(256), 0, s,
                           (const double*)src, (double*)dst, n);
    else if (src_f64)
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(nb), dim3(256), 0, s,
                           (const double*)src, (float*)dst, n);
    else if (dst_f64)
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(nb), dim3(256), 0, s,
                           (const float*)src, (double*)dst, n);
    else
        hipLaunchKernelGGL((convert_kernel<float, float>), dim3(nb), dim3(256), 0, s,
                           (const float*)src, (float*)dst, n);
    return (int)hipGetLastError();
}

}  // namespace ahip

namespace ahip {
namespace {
// Streaming copy used to calibrate HBM counters and the achievable bandwidth:
// 16-byte non-temporal loads and stores, grid-stride.
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void copy16(const v4u* __restrict__ src, v4u* __restrict__ dst,
                                              int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        v4u v = __builtin_nontemporal_load(src + i);
        __builtin_nontemporal_store(v, dst + i);
    }
}
}  // namespace

int launch_copy16(const void* src, void* dst, int64_t n16, hipStream_t s) {
    hipLaunchKernelGGL(copy16, dim3(256 * 8 * 4), dim3(256), 0, s, (const v4u*)src, (v4u*)dst, n16);
    return (int)hipGetLastError();
}
}  // namespace ahip
