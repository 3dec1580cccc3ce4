// k_rate_psp.hip -- rate-coded compute_psp: y = W . r over the device layouts.
//
// Semantics: generator/Projection/OpenMP/CSR.py:178-196 -- for each dendrite
// (post rank), sum from 0 of w[j]*pre.r[col[j]], then target[post] += sum
// after the per-step reset of the target (Population/OpenMPTemplates.py:
// 269-276).  A projection that is the only (first) one into its target
// overwrites the target instead of reset + add, which is the same value.
// FP64: plain sums (row order differs from the sequential loop: results agree
// to rounding).  FP32: Neumaier-compensated lane sums and a compensated lane
// reduction.
#include "ahip_device.hpp"
#include "internal.hpp"

namespace ahip {
namespace {

constexpr int kBlock = 256;

template <typename T>
struct Acc;
template <>
struct Acc<double> {
    double s = 0.0;
    __device__ inline void add(double x) { s += x; }
    template <int W>
    __device__ inline double reduce() { return group_sum<W>(s); }
};
template <>
struct Acc<float> {
    Kahan k{0.0f, 0.0f};
    __device__ inline void add(float x) { k.add(x); }
    template <int W>
    __device__ inline float reduce() { return group_sum_kahan<W>(k.s, k.c); }
};

template <typename T>
__device__ inline T ldnt(const T* p) {
    return __builtin_nontemporal_load(p);
}

template <typename T>
__device__ inline void store_target(T* y, int64_t r, T v, bool acc) {
    if (acc)
        y[r] += v;
    else
        y[r] = v;
}

// CSR, L lanes per row, rows indexed by post rank (row_ptr over all post
// neurons).  Four independent loads in flight per lane.
template <typename T, int L>
__global__ __launch_bounds__(kBlock) void csr_rows(int32_t n_rows, const int64_t* __restrict__ row_ptr,
                                                   const int32_t* __restrict__ col,
                                                   const T* __restrict__ w,
                                                   const T* __restrict__ x, T* __restrict__ y,
                                                   int32_t acc) {
    const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t row = gid / L;
    const int lane = threadIdx.x & (L - 1);
    if (row >= n_rows) return;  // whole groups leave together (L divides kBlock)
    const int64_t beg = row_ptr[row], end = row_ptr[row + 1];
    Acc<T> a0, a1;
    int64_t j = beg + lane;
    for (; j + 3 * L < end; j += 4 * L) {
        const int32_t c0 = ldnt(col + j), c1 = ldnt(col + j + L);
        const int32_t c2 = ldnt(col + j + 2 * L), c3 = ldnt(col + j + 3 * L);
        const T w0 = ldnt(w + j), w1 = ldnt(w + j + L);
        const T w2 = ldnt(w + j + 2 * L), w3 = ldnt(w + j + 3 * L);
        a0.add(w0 * x[c0]);
        a1.add(w1 * x[c1]);
        a0.add(w2 * x[c2]);
        a1.add(w3 * x[c3]);
    }
    for (; j < end; j += L) a0.add(ldnt(w + j) * x[ldnt(col + j)]);
    a0.add(a1.template reduce<1>());  // fold the second accumulator in
    const T s = a0.template reduce<L>();
    if (lane == 0) store_target(y, row, s, acc != 0);
}

// CSR with 16-byte loads: every lane takes 4 consecutive synapses (one int4
// of column indices, 32 or 16 bytes of weights) and keeps two chunks in
// flight, so a wave has 2 x 64 x (16 + 4*sizeof(T)) bytes outstanding.  The
// row head up to the next multiple-of-4 index is done with scalar loads,
// the straddling tail chunk by the one lane that owns it.
typedef int v4i __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Quad;
template <>
struct Quad<double> {
    double v[4];
    __device__ inline void load(const double* p) {
        const v2d a = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(p));
        const v2d b = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(p) + 1);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    }
};
template <>
struct Quad<float> {
    float v[4];
    __device__ inline void load(const float* p) {
        const v4f a = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p));
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    }
};

template <typename T, int L>
__global__ __launch_bounds__(kBlock) void csr_vec(int32_t n_rows, const int64_t* __restrict__ row_ptr,
                                                  const int32_t* __restrict__ col,
                                                  const T* __restrict__ w, const T* __restrict__ x,
                                                  T* __restrict__ y, int32_t acc) {
    const int64_t row = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / L;
    const int lane = threadIdx.x & (L - 1);
    if (row >= n_rows) return;
    const int64_t beg = row_ptr[row], end = row_ptr[row + 1];
    Acc<T> a0, a1;
    const int64_t body = min(end, (beg + 3) & ~(int64_t)3);
    if (beg + lane < body) a0.add(w[beg + lane] * x[col[beg + lane]]);
    int64_t j = body + 4 * lane;
    for (; j + 4 * L + 4 <= end; j += 8 * L) {
        const v4i c0 = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(col + j));
        const v4i c1 = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(col + j + 4 * L));
        Quad<T> w0, w1;
        w0.load(w + j);
        w1.load(w + j + 4 * L);
        a0.add(w0.v[0] * x[c0.x]);
        a1.add(w0.v[1] * x[c0.y]);
        a0.add(w0.v[2] * x[c0.z]);
        a1.add(w0.v[3] * x[c0.w]);
        a0.add(w1.v[0] * x[c1.x]);
        a1.add(w1.v[1] * x[c1.y]);
        a0.add(w1.v[2] * x[c1.z]);
        a1.add(w1.v[3] * x[c1.w]);
    }
    for (; j + 4 <= end; j += 4 * L) {
        const v4i c0 = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(col + j));
        Quad<T> w0;
        w0.load(w + j);
        a0.add(w0.v[0] * x[c0.x]);
        a1.add(w0.v[1] * x[c0.y]);
        a0.add(w0.v[2] * x[c0.z]);
        a1.add(w0.v[3] * x[c0.w]);
    }
    for (; j < end; j++) a0.add(w[j] * x[col[j]]);
    a0.add(a1.template reduce<1>());
    const T s = a0.template reduce<L>();
    if (lane == 0) store_target(y, row, s, acc != 0);
}

// ELLPACK, column-major over the stored rows (rows of the LIL); padding column
// = INT32_MAX ends a row (CUDA/ELL.py:240-262).  ELLR reads the row length.
template <typename T, bool R>
__global__ __launch_bounds__(kBlock) void ell_rows(int32_t n_stored, int32_t width,
                                                   const int32_t* __restrict__ post,
                                                   const int32_t* __restrict__ rl,
                                                   const int32_t* __restrict__ col,
                                                   const T* __restrict__ w,
                                                   const T* __restrict__ x, T* __restrict__ y,
                                                   int32_t acc) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n_stored) return;
    Acc<T> a;
    const int32_t len = R ? rl[r] : width;
    for (int32_t k = 0; k < len; k++) {
        const int64_t pos = (int64_t)k * n_stored + r;
        const int32_t c = ldnt(col + pos);
        if (!R && c == INT32_MAX) break;
        a.add(ldnt(w + pos) * x[c]);
    }
    T s = a.template reduce<1>();
    store_target(y, post[r], s, acc != 0);
}

// SELL-C-sigma with TPR threads per row: slice of C = 64/TPR rows, entry
// (row r, k) at off + (k/TPR)*64 + r*TPR + k%TPR; slice width padded to a
// multiple of TPR with column 0 / weight 0.  One wave per slice.
template <typename T, int TPR>
__global__ __launch_bounds__(kBlock) void sell_slices(int32_t n_slices, int32_t n_rows,
                                                      const int64_t* __restrict__ off,
                                                      const int32_t* __restrict__ perm,
                                                      const int32_t* __restrict__ col,
                                                      const T* __restrict__ w,
                                                      const T* __restrict__ x, T* __restrict__ y,
                                                      int32_t acc) {
    constexpr int C = kWave / TPR;
    const int64_t slice = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (slice >= n_slices) return;
    const int64_t b = off[slice], e = off[slice + 1];
    Acc<T> a0, a1;
    int64_t p = b + lane;
    for (; p + kWave < e; p += 2 * kWave) {
        const int32_t c0 = ldnt(col + p), c1 = ldnt(col + p + kWave);
        const T w0 = ldnt(w + p), w1 = ldnt(w + p + kWave);
        a0.add(w0 * x[c0]);
        a1.add(w1 * x[c1]);
    }
    if (p < e) a0.add(ldnt(w + p) * x[ldnt(col + p)]);
    a0.add(a1.template reduce<1>());
    const T s = a0.template reduce<TPR>();
    const int r = lane / TPR;
    const int64_t srow = slice * C + r;
    if ((lane & (TPR - 1)) == 0 && srow < n_rows) {
        const int32_t pr = perm[srow];
        if (pr >= 0) store_target(y, pr, s, acc != 0);
    }
}

// BSR with BS x BS tiles (row-major inside a tile); one wave per block row,
// 64/(BS*BS) tiles in flight per wave.
template <typename T, int BS>
__global__ __launch_bounds__(kBlock) void bsr_rows(int32_t n_brows, int32_t n_rows, int32_t n_cols,
                                                   const int32_t* __restrict__ brp,
                                                   const int32_t* __restrict__ bci,
                                                   const T* __restrict__ val,
                                                   const T* __restrict__ x, T* __restrict__ y,
                                                   int32_t acc) {
    constexpr int TT = BS * BS;
    constexpr int SLOTS = kWave / TT;
    const int64_t br = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (br >= n_brows) return;
    const int slot = lane / TT, e = lane % TT, tr = e / BS, tc = e % BS;
    Acc<T> a;
    for (int32_t t = brp[br] + slot; t < brp[br + 1]; t += SLOTS) {
        const int32_t c = bci[t] * BS + tc;
        const T v = ldnt(val + (int64_t)t * TT + e);
        if (c < n_cols) a.add(v * x[c]);
    }
    // sum over the BS columns of a tile row (adjacent lanes), then over slots
    T s;
    if constexpr (sizeof(T) == 8) {
        s = group_sum<BS>(a.s);
#pragma unroll
        for (int o = TT; o < kWave; o <<= 1) s += __shfl_xor(s, o, kWave);
    } else {
        float ss = a.k.s, cc = a.k.c;
        float v = group_sum_kahan<BS>(ss, cc);
#pragma unroll
        for (int o = TT; o < kWave; o <<= 1) v += __shfl_xor(v, o, kWave);
        s = v;
    }
    const int64_t row = br * BS + tr;
    if (slot == 0 && tc == 0 && row < n_rows) store_target(y, row, s, acc != 0);
}

// Dense row-major W [n_rows x n_cols]: one wave per row.
template <typename T>
__global__ __launch_bounds__(kBlock) void dense_rows(int32_t n_rows, int32_t n_cols,
                                                     const T* __restrict__ W,
                                                     const T* __restrict__ x, T* __restrict__ y,
                                                     int32_t acc) {
    const int64_t row = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (row >= n_rows) return;
    const T* wr = W + row * (int64_t)n_cols;
    Acc<T> a0, a1;
    int32_t k = lane;
    for (; k + kWave < n_cols; k += 2 * kWave) {
        a0.add(ldnt(wr + k) * x[k]);
        a1.add(ldnt(wr + k + kWave) * x[k + kWave]);
    }
    if (k < n_cols) a0.add(ldnt(wr + k) * x[k]);
    a0.add(a1.template reduce<1>());
    const T s = a0.template reduce<kWave>();
    if (lane == 0) store_target(y, row, s, acc != 0);
}

// Row-block x column-panel tiles.  The gather of the pre-synaptic rates is
// the limit of the row kernels above (every w*x[c] is a random 8-byte read
// through the vector L1, one cache line per lane), so here a workgroup owns
// one tile: it stages the panel's slice of x (P values) in LDS with coalesced
// loads, then streams the tile's synapses (contiguous, rows ascending) with L
// lanes per row and serves every x[c] from LDS.  Tile t = q*n_rb + rb writes
// the partial sums of its rows into part[q][row]; tile_reduce adds the panels
// in ascending order, so the result does not depend on scheduling.
template <typename T, int L>
__global__ __launch_bounds__(kBlock) void tile_psp(int32_t n_rows, int32_t n_cols, int32_t panel_cols,
                                                   int32_t rb_rows, int32_t n_rb,
                                                   const int64_t* __restrict__ base,
                                                   const int32_t* __restrict__ roff,
                                                   const int32_t* __restrict__ col,
                                                   const T* __restrict__ w,
                                                   const T* __restrict__ x, T* __restrict__ part) {
    extern __shared__ unsigned char lds_raw[];
    T* xs = reinterpret_cast<T*>(lds_raw);
    const int32_t t = blockIdx.x;
    const int32_t q = t / n_rb, rb = t % n_rb;
    const int32_t c0 = q * panel_cols;
    const int32_t nc = min(panel_cols, n_cols - c0);
    for (int32_t k = threadIdx.x; k < nc; k += kBlock) xs[k] = x[c0 + k];
    __syncthreads();
    const int lane = threadIdx.x & (L - 1);
    const int grp = threadIdx.x / L;
    constexpr int GROUPS = kBlock / L;
    const int32_t row0 = rb * rb_rows;
    const int32_t nr = min(rb_rows, n_rows - row0);
    const int64_t b0 = base[t];
    const int32_t* ro = roff + (int64_t)t * (rb_rows + 1);
    T* pq = part + (int64_t)q * n_rows + row0;
    for (int32_t lr = grp; lr < nr; lr += GROUPS) {
        const int64_t beg = b0 + ro[lr], end = b0 + ro[lr + 1];
        Acc<T> a0, a1;
        if (L >= 4) {
            // 16-byte loads: scalar head up to a multiple-of-4 index, then 4
            // synapses per lane per chunk, scalar tail in the straddling lane
            const int64_t body = min(end, (beg + 3) & ~(int64_t)3);
            if (beg + lane < body) a0.add(ldnt(w + beg + lane) * xs[ldnt(col + beg + lane) - c0]);
            int64_t j = body + 4 * lane;
            for (; j + 4 <= end; j += 4 * L) {
                const v4i c = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(col + j));
                Quad<T> q4;
                q4.load(w + j);
                a0.add(q4.v[0] * xs[c.x - c0]);
                a1.add(q4.v[1] * xs[c.y - c0]);
                a0.add(q4.v[2] * xs[c.z - c0]);
                a1.add(q4.v[3] * xs[c.w - c0]);
            }
            for (; j < end; j++) a0.add(ldnt(w + j) * xs[ldnt(col + j) - c0]);
        } else {
            int64_t j = beg + lane;
            for (; j + L < end; j += 2 * L) {
                const int32_t cA = ldnt(col + j), cB = ldnt(col + j + L);
                const T wA = ldnt(w + j), wB = ldnt(w + j + L);
                a0.add(wA * xs[cA - c0]);
                a1.add(wB * xs[cB - c0]);
            }
            if (j < end) a0.add(ldnt(w + j) * xs[ldnt(col + j) - c0]);
        }
        a0.add(a1.template reduce<1>());
        const T s = a0.template reduce<L>();
        if (lane == 0) pq[lr] = s;
    }
}

// Sliced tiles: the tiles of tile_psp with their rows sorted by length and
// packed into slices of 64 lanes (uint16 panel-relative columns, 4 synapses
// per lane per step).  A wave streams one slice with U steps (U x 64 x
// (8 + 4*sizeof(T)) bytes) in flight and no per-row set-up; a row of a
// slice owns g consecutive lanes (g = 1 unless the row is long) and its
// first lane adds the g lane sums in lane order.  Padding reads the zero slot
// xs[P].  Tile t = q*n_rb + rb writes part[q][row]; tile_reduce adds the
// panels in ascending order.  (Adding them in the last tile of a row block
// to finish needs a device-scope release per tile, which on this chip writes
// back the L2 of the tile's XCD: three times slower in total.)
typedef unsigned short v4u16 __attribute__((ext_vector_type(4)));

template <typename T, int U, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void sell_panel_psp(int32_t n_rows, int32_t n_cols, int32_t panel_cols,
                                                         int32_t rb_rows, int32_t n_rb,
                                                         const int64_t* __restrict__ off,
                                                         const int32_t* __restrict__ rowof,
                                                         const int32_t* __restrict__ tile_slices,
                                                         const int32_t* __restrict__ rb_row0,
                                                         const uint16_t* __restrict__ col,
                                                         const T* __restrict__ w,
                                                         const T* __restrict__ x, T* __restrict__ part) {
    extern __shared__ unsigned char lds_raw[];
    T* xs = reinterpret_cast<T*>(lds_raw);
    const int32_t t = blockIdx.x;
    const int32_t q = t / n_rb, rb = t % n_rb;
    const int32_t c0 = q * panel_cols;
    const int32_t nc = min(panel_cols, n_cols - c0);
    // stage the panel of x: each thread loads 8 values before it stores
    // them, so a panel costs a few load latencies instead of one per value
    constexpr int32_t kB = 8, kT = WAVES * kWave;
    for (int32_t k0 = 0; k0 < nc; k0 += kB * kT) {
        T v[kB];
#pragma unroll
        for (int i = 0; i < kB; i++) {
            const int32_t k = k0 + i * kT + (int32_t)threadIdx.x;
            v[i] = k < nc ? x[c0 + k] : T(0);
        }
#pragma unroll
        for (int i = 0; i < kB; i++) {
            const int32_t k = k0 + i * kT + (int32_t)threadIdx.x;
            if (k < nc) xs[k] = v[i];
        }
    }
    if (threadIdx.x == 0) xs[panel_cols] = T(0);
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    (void)rb_rows;
    T* pq = part + (int64_t)q * n_rows + rb_row0[rb];
    const int32_t s0 = tile_slices[t], s_end = tile_slices[t + 1];
    // slices are sorted by length (descending); in every round of 2*WAVES
    // slices wave w takes the w-th and the (2*WAVES-1-w)-th, so each wave's
    // pair of a long and a short slice is about as long as any other's
    for (int32_t k = 0;; k++) {
        const int32_t r = k >> 1;
        const int32_t gs = s0 + r * 2 * WAVES + ((k & 1) ? 2 * WAVES - 1 - wave : wave);
        if (s0 + r * 2 * WAVES >= s_end) break;
        if (gs >= s_end) continue;
        const int64_t b = off[gs];
        const int64_t nsteps = (off[gs + 1] - b) >> 8;
        const uint16_t* cp = col + b + 4 * lane;
        const T* wp = w + b + 4 * lane;
        Acc<T> a0, a1;
        int64_t m = 0;
        for (; m + U <= nsteps; m += U) {
            v4u16 c[U];
            Quad<T> v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                c[u] = __builtin_nontemporal_load(reinterpret_cast<const v4u16*>(cp + (m + u) * 256));
                v[u].load(wp + (m + u) * 256);
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                a0.add(v[u].v[0] * xs[c[u].x]);
                a1.add(v[u].v[1] * xs[c[u].y]);
                a0.add(v[u].v[2] * xs[c[u].z]);
                a1.add(v[u].v[3] * xs[c[u].w]);
            }
        }
        for (; m < nsteps; m++) {
            const v4u16 c = __builtin_nontemporal_load(reinterpret_cast<const v4u16*>(cp + m * 256));
            Quad<T> v;
            v.load(wp + m * 256);
            a0.add(v.v[0] * xs[c.x]);
            a1.add(v.v[1] * xs[c.y]);
            a0.add(v.v[2] * xs[c.z]);
            a1.add(v.v[3] * xs[c.w]);
        }
        a0.add(a1.template reduce<1>());
        T v = a0.template reduce<1>();
        const int32_t ro = rowof[(int64_t)gs * kWave + lane];
        const int g = ro >= 0 ? (ro >> 16) : 0;
        if (__any(g > 1)) {
            Acc<T> r;
            r.add(v);
            for (int i = 1; i < kWave; i++) {
                const T o = __shfl(v, (lane + i) & (kWave - 1), kWave);
                if (i < g) r.add(o);
                if (!__any(i + 1 < g)) break;
            }
            v = r.template reduce<1>();
        }
        if (ro >= 0) pq[ro & 0xFFFF] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void tile_reduce(int32_t n_rows, int32_t n_panels,
                                                      const T* __restrict__ part, T* __restrict__ y,
                                                      int32_t acc) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n_rows) return;
    store_target(y, r, panel_sum(part, n_panels, n_rows, r), acc != 0);
}

template <typename T>
int launch_rate_t(const ahip_net* net, const Proj& p, const T* x, T* y, hipStream_t s, bool reduce) {
    const RateLayout& L = p.rate;
    const int32_t acc = p.accumulate ? 1 : 0;
    (void)net;
    switch (L.format) {
        case AHIP_FMT_CSR: {
            const int64_t threads = (int64_t)p.n_rows * L.lanes;
            const int64_t nb = (threads + kBlock - 1) / kBlock;
            const int64_t* rp = L.ptr.as<int64_t>();
            const int32_t* ci = L.idx.as<int32_t>();
            const T* w = L.val.as<T>();
            if (nb == 0) return 0;
            switch (L.lanes) {
#define AHIP_CSR_CASE(N)                                                                     \
    case N:                                                                                  \
        hipLaunchKernelGGL((csr_rows<T, N>), dim3(nb), dim3(kBlock), 0, s, p.n_rows, rp, ci, \
                           w, x, y, acc);                                                    \
        break;
                AHIP_CSR_CASE(1)
                AHIP_CSR_CASE(2)
                AHIP_CSR_CASE(4)
                AHIP_CSR_CASE(8)
                AHIP_CSR_CASE(16)
                AHIP_CSR_CASE(32)
                AHIP_CSR_CASE(64)
#undef AHIP_CSR_CASE
                default:
                    return (int)hipErrorInvalidValue;
            }
            return (int)hipGetLastError();
        }
        case AHIP_FMT_CSR_VEC: {
            const int64_t threads = (int64_t)p.n_rows * L.lanes;
            const int64_t nb = (threads + kBlock - 1) / kBlock;
            if (nb == 0) return 0;
            const int64_t* rp = L.ptr.as<int64_t>();
            const int32_t* ci = L.idx.as<int32_t>();
            const T* w = L.val.as<T>();
            switch (L.lanes) {
#define AHIP_VEC_CASE(N)                                                                    \
    case N:                                                                                 \
        hipLaunchKernelGGL((csr_vec<T, N>), dim3(nb), dim3(kBlock), 0, s, p.n_rows, rp, ci, \
                           w, x, y, acc);                                                   \
        break;
                AHIP_VEC_CASE(4)
                AHIP_VEC_CASE(8)
                AHIP_VEC_CASE(16)
                AHIP_VEC_CASE(32)
                AHIP_VEC_CASE(64)
#undef AHIP_VEC_CASE
                default:
                    return (int)hipErrorInvalidValue;
            }
            return (int)hipGetLastError();
        }
        case AHIP_FMT_ELL:
        case AHIP_FMT_ELLR: {
            const int32_t n_stored = L.n_units;
            if (n_stored == 0) return 0;
            const int64_t nb = (n_stored + kBlock - 1) / kBlock;
            if (L.format == AHIP_FMT_ELL)
                hipLaunchKernelGGL((ell_rows<T, false>), dim3(nb), dim3(kBlock), 0, s, n_stored,
                                   L.width, L.perm.as<int32_t>(), L.rl.as<int32_t>(),
                                   L.idx.as<int32_t>(), L.val.as<T>(), x, y, acc);
            else
                hipLaunchKernelGGL((ell_rows<T, true>), dim3(nb), dim3(kBlock), 0, s, n_stored,
                                   L.width, L.perm.as<int32_t>(), L.rl.as<int32_t>(),
                                   L.idx.as<int32_t>(), L.val.as<T>(), x, y, acc);
            return (int)hipGetLastError();
        }
        case AHIP_FMT_SELL: {
            const int64_t nb = ((int64_t)L.n_units * kWave + kBlock - 1) / kBlock;
            if (nb == 0) return 0;
            const int64_t* off = L.ptr.as<int64_t>();
            const int32_t* perm = L.perm.as<int32_t>();
            const int32_t* ci = L.idx.as<int32_t>();
            const T* w = L.val.as<T>();
            const int32_t n_srows = L.n_units * (kWave / L.lanes);
            switch (L.lanes) {
#define AHIP_SELL_CASE(N)                                                                    \
    case N:                                                                                  \
        hipLaunchKernelGGL((sell_slices<T, N>), dim3(nb), dim3(kBlock), 0, s, L.n_units,     \
                           n_srows, off, perm, ci, w, x, y, acc);                            \
        break;
                AHIP_SELL_CASE(1)
                AHIP_SELL_CASE(2)
                AHIP_SELL_CASE(4)
                AHIP_SELL_CASE(8)
                AHIP_SELL_CASE(16)
                AHIP_SELL_CASE(32)
                AHIP_SELL_CASE(64)
#undef AHIP_SELL_CASE
                default:
                    return (int)hipErrorInvalidValue;
            }
            return (int)hipGetLastError();
        }
        case AHIP_FMT_BSR: {
            const int64_t nb = ((int64_t)L.n_units * kWave + kBlock - 1) / kBlock;
            if (nb == 0) return 0;
            if (L.chunk == 2)
                hipLaunchKernelGGL((bsr_rows<T, 2>), dim3(nb), dim3(kBlock), 0, s, L.n_units,
                                   p.n_rows, p.n_cols, L.ptr.as<int32_t>(), L.idx.as<int32_t>(),
                                   L.val.as<T>(), x, y, acc);
            else if (L.chunk == 4)
                hipLaunchKernelGGL((bsr_rows<T, 4>), dim3(nb), dim3(kBlock), 0, s, L.n_units,
                                   p.n_rows, p.n_cols, L.ptr.as<int32_t>(), L.idx.as<int32_t>(),
                                   L.val.as<T>(), x, y, acc);
            else if (L.chunk == 8)
                hipLaunchKernelGGL((bsr_rows<T, 8>), dim3(nb), dim3(kBlock), 0, s, L.n_units,
                                   p.n_rows, p.n_cols, L.ptr.as<int32_t>(), L.idx.as<int32_t>(),
                                   L.val.as<T>(), x, y, acc);
            else
                return (int)hipErrorInvalidValue;
            return (int)hipGetLastError();
        }
        case AHIP_FMT_DENSE: {
            const int64_t nb = ((int64_t)p.n_rows * kWave + kBlock - 1) / kBlock;
            if (nb == 0) return 0;
            hipLaunchKernelGGL((dense_rows<T>), dim3(nb), dim3(kBlock), 0, s, p.n_rows, p.n_cols,
                               L.val.as<T>(), x, y, acc);
            return (int)hipGetLastError();
        }
        case AHIP_FMT_CSR_PANEL: {
            const int64_t ntiles = (int64_t)L.n_units * L.width;
            if (ntiles == 0 || p.n_rows == 0) return 0;
            const size_t lds = (size_t)L.panel_cols * sizeof(T);
            T* part = L.perm.as<T>();
            switch (L.lanes) {
#define AHIP_TILE_CASE(N)                                                                      \
    case N:                                                                                    \
        hipLaunchKernelGGL((tile_psp<T, N>), dim3(ntiles), dim3(kBlock), lds, s, p.n_rows,      \
                           p.n_cols, L.panel_cols, L.chunk, L.width, L.ptr.as<int64_t>(),       \
                           L.aux.as<int32_t>(), L.idx.as<int32_t>(), L.val.as<T>(), x, part);   \
        break;
                AHIP_TILE_CASE(4)
                AHIP_TILE_CASE(8)
                AHIP_TILE_CASE(16)
                AHIP_TILE_CASE(32)
#undef AHIP_TILE_CASE
                default:
                    return (int)hipErrorInvalidValue;
            }
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return (int)e;
            // reduce = false: the post population's update adds the panels
            if (!reduce) return 0;
            hipLaunchKernelGGL((tile_reduce<T>), dim3((p.n_rows + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                               p.n_rows, L.n_units, part, y, acc);
            return (int)hipGetLastError();
        }
        case AHIP_FMT_SELL_PANEL: {
            const int64_t ntiles = (int64_t)L.n_units * L.width;
            if (ntiles == 0 || p.n_rows == 0) return 0;
            const size_t lds = (size_t)(L.panel_cols + 1) * sizeof(T);
            T* part = L.perm.as<T>();
            switch (L.lanes * 16 + L.sigma) {
#define AHIP_SP_CASE(N, W)                                                                          \
    case N * 16 + W:                                                                                \
        hipLaunchKernelGGL((sell_panel_psp<T, N, W>), dim3(ntiles), dim3(W * kWave), lds, s,        \
                           p.n_rows, p.n_cols, L.panel_cols, L.chunk, L.width,                      \
                           L.ptr.as<int64_t>(), L.aux.as<int32_t>(), L.tiles.as<int32_t>(),         \
                           L.rl.as<int32_t>(), L.idx.as<uint16_t>(), L.val.as<T>(), x, part);       \
        break;
                AHIP_SP_CASE(1, 4)
                AHIP_SP_CASE(2, 4)
                AHIP_SP_CASE(4, 4)
                AHIP_SP_CASE(8, 4)
                AHIP_SP_CASE(2, 8)
                AHIP_SP_CASE(4, 8)
                AHIP_SP_CASE(8, 8)
#undef AHIP_SP_CASE
                default:
                    return (int)hipErrorInvalidValue;
            }
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return (int)e;
            // reduce = false: the post population's update adds the panels
            if (!reduce) return 0;
            hipLaunchKernelGGL((tile_reduce<T>), dim3((p.n_rows + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                               p.n_rows, L.n_units, part, y, acc);
            return (int)hipGetLastError();
        }
        default:
            return (int)hipErrorInvalidValue;
    }
}

}  // namespace

int launch_rate_psp(const ahip_net* net, const Proj& p, const void* pre_r, void* target,
                    hipStream_t s, bool reduce) {
    if (net->precision == AHIP_F64)
        return launch_rate_t<double>(net, p, (const double*)pre_r, (double*)target, s, reduce);
    return launch_rate_t<float>(net, p, (const float*)pre_r, (float*)target, s, reduce);
}

}  // namespace ahip
