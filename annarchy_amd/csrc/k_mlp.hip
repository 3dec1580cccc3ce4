// k_mlp.hip -- the learned storage-format selector: an MLP
// d_in -> h1 -> h2 -> d_out (ReLU, softmax cross-entropy against target
// distributions), FP32, trained and evaluated on the device.
//
// The reference chooses formats with a decision tree on density and mean row
// length (core/Projection.py:691-789).  Here the choice is learned from the
// timings of our own exhaustive autotuner.  One training epoch (full batch):
//   1. forward + deltas, one thread per sample: activations a1, a2 and the
//      back-propagated deltas g1, g2, dz are written to global memory;
//   2. parameter gradients, one thread per parameter: dW[o][k] =
//      sum_s delta[s][o] * act[s][k], summed over samples in index order (no
//      atomics: the training run is deterministic);
//   3. Adam update of the parameters.
// Layout of the parameter vector: [W1 (h1 x d_in), b1, W2 (h2 x h1), b2,
// W3 (d_out x h2), b3], row-major.
#include <mutex>

#include "ahip_device.hpp"
#include "internal.hpp"

namespace ahip {

namespace {

constexpr int DIN = 16, H1 = 64, H2 = 64, DOUT = 12;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H1 * DIN, OFF_W2 = OFF_B1 + H1, OFF_B2 = OFF_W2 + H2 * H1,
              OFF_W3 = OFF_B2 + H2, OFF_B3 = OFF_W3 + DOUT * H2, NPAR = OFF_B3 + DOUT;
static_assert(NPAR == 6028, "16-64-64-12 has 6028 parameters");

__device__ inline void forward(const float* __restrict__ P, const float* x, float* a1, float* a2, float* z) {
    for (int o = 0; o < H1; o++) {
        float a = P[OFF_B1 + o];
        for (int k = 0; k < DIN; k++) a += P[OFF_W1 + o * DIN + k] * x[k];
        a1[o] = a > 0.f ? a : 0.f;
    }
    for (int o = 0; o < H2; o++) {
        float a = P[OFF_B2 + o];
        for (int k = 0; k < H1; k++) a += P[OFF_W2 + o * H1 + k] * a1[k];
        a2[o] = a > 0.f ? a : 0.f;
    }
    for (int o = 0; o < DOUT; o++) {
        float a = P[OFF_B3 + o];
        for (int k = 0; k < H2; k++) a += P[OFF_W3 + o * H2 + k] * a2[k];
        z[o] = a;
    }
}

// 1. forward + deltas per sample; loss_s written per sample.
__global__ __launch_bounds__(64) void mlp_deltas(int32_t n, const float* __restrict__ X,
                                                 const float* __restrict__ T, const float* __restrict__ P,
                                                 float* __restrict__ A1, float* __restrict__ A2,
                                                 float* __restrict__ G1, float* __restrict__ G2,
                                                 float* __restrict__ DZ, float* __restrict__ loss) {
    const int32_t s = blockIdx.x * 64 + threadIdx.x;
    if (s >= n) return;
    float x[DIN], a1[H1], a2[H2], z[DOUT];
    for (int k = 0; k < DIN; k++) x[k] = X[(int64_t)s * DIN + k];
    forward(P, x, a1, a2, z);
    float m = z[0];
    for (int o = 1; o < DOUT; o++) m = fmaxf(m, z[o]);
    float se = 0.f;
    for (int o = 0; o < DOUT; o++) se += expf(z[o] - m);
    const float lse = m + logf(se);
    float l = 0.f, tsum = 0.f;
    float dz[DOUT];
    for (int o = 0; o < DOUT; o++) {
        const float t = T[(int64_t)s * DOUT + o];
        l += t * (lse - z[o]);
        tsum += t;
    }
    for (int o = 0; o < DOUT; o++) dz[o] = (tsum * expf(z[o] - lse) - T[(int64_t)s * DOUT + o]) / (float)n;
    loss[s] = l / (float)n;
    float g2[H2];
    for (int k = 0; k < H2; k++) {
        float d = 0.f;
        for (int o = 0; o < DOUT; o++) d += dz[o] * P[OFF_W3 + o * H2 + k];
        g2[k] = a2[k] > 0.f ? d : 0.f;
    }
    for (int k = 0; k < H1; k++) {
        float d = 0.f;
        for (int o = 0; o < H2; o++) d += g2[o] * P[OFF_W2 + o * H1 + k];
        G1[(int64_t)s * H1 + k] = a1[k] > 0.f ? d : 0.f;
        A1[(int64_t)s * H1 + k] = a1[k];
    }
    for (int k = 0; k < H2; k++) {
        G2[(int64_t)s * H2 + k] = g2[k];
        A2[(int64_t)s * H2 + k] = a2[k];
    }
    for (int o = 0; o < DOUT; o++) DZ[(int64_t)s * DOUT + o] = dz[o];
}

// 2. parameter gradients: one thread per parameter, samples in order.
__global__ __launch_bounds__(256) void mlp_grads(int32_t n, const float* __restrict__ X,
                                                 const float* __restrict__ A1, const float* __restrict__ A2,
                                                 const float* __restrict__ G1, const float* __restrict__ G2,
                                                 const float* __restrict__ DZ, float* __restrict__ grad) {
    const int32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NPAR) return;
    const float *delta, *act;
    int dd, da, o, k;
    if (p < OFF_B1) { o = p / DIN; k = p % DIN; delta = G1; dd = H1; act = X; da = DIN; }
    else if (p < OFF_W2) { o = p - OFF_B1; k = -1; delta = G1; dd = H1; act = nullptr; da = 0; }
    else if (p < OFF_B2) { o = (p - OFF_W2) / H1; k = (p - OFF_W2) % H1; delta = G2; dd = H2; act = A1; da = H1; }
    else if (p < OFF_W3) { o = p - OFF_B2; k = -1; delta = G2; dd = H2; act = nullptr; da = 0; }
    else if (p < OFF_B3) { o = (p - OFF_W3) / H2; k = (p - OFF_W3) % H2; delta = DZ; dd = DOUT; act = A2; da = H2; }
    else { o = p - OFF_B3; k = -1; delta = DZ; dd = DOUT; act = nullptr; da = 0; }
    float g = 0.f;
    if (k < 0) {
        for (int32_t s = 0; s < n; s++) g += delta[(int64_t)s * dd + o];
    } else {
        for (int32_t s = 0; s < n; s++) g += delta[(int64_t)s * dd + o] * act[(int64_t)s * da + k];
    }
    grad[p] = g;
}

// 3. Adam (Kingma & Ba 2015), bias-corrected.
__global__ __launch_bounds__(256) void mlp_adam(float* __restrict__ P, const float* __restrict__ g,
                                                float* __restrict__ m, float* __restrict__ v, float lr,
                                                float b1, float b2, float eps, float c1, float c2,
                                                float wd) {
    const int32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NPAR) return;
    const float gp = g[p] + wd * P[p];
    const float mp = b1 * m[p] + (1.f - b1) * gp;
    const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    P[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

__global__ __launch_bounds__(64) void mlp_predict(int32_t n, const float* __restrict__ X,
                                                  const float* __restrict__ P, int32_t* __restrict__ cls,
                                                  float* __restrict__ logits) {
    const int32_t s = blockIdx.x * 64 + threadIdx.x;
    if (s >= n) return;
    float x[DIN], a1[H1], a2[H2], z[DOUT];
    for (int k = 0; k < DIN; k++) x[k] = X[(int64_t)s * DIN + k];
    forward(P, x, a1, a2, z);
    int best = 0;
    for (int o = 1; o < DOUT; o++)
        if (z[o] > z[best]) best = o;
    cls[s] = best;
    if (logits)
        for (int o = 0; o < DOUT; o++) logits[(int64_t)s * DOUT + o] = z[o];
}

// Sum of the per-sample losses of one epoch into loss_curve[e]: one
// workgroup, FP64 partial sums in a fixed order (deterministic).
__global__ __launch_bounds__(256) void mlp_loss_sum(int32_t n, const float* __restrict__ L,
                                                    float* __restrict__ out) {
    __shared__ double part[256];
    double a = 0.0;
    for (int32_t s = threadIdx.x; s < n; s += 256) a += (double)L[s];
    part[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (float)part[0];
}

__global__ __launch_bounds__(256) void mlp_zero(float* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

__global__ __launch_bounds__(256) void mlp_copy(float* __restrict__ dst, const float* __restrict__ src, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ---- training without allocations -----------------------------------------
// Scratch of ahip_mlp_train_noalloc: the gradient and the Adam moments
// (3 x NPAR floats) and the per-sample activations, deltas and losses of up
// to kNoallocSamples samples.  Each GPU holds its own copy of the array for
// the life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that device.
constexpr int32_t kNoallocSamples = 2048;
constexpr int64_t kPerSample = 2 * H1 + 2 * H2 + DOUT + 1;
constexpr int64_t kScratchFloats = 3 * (int64_t)NPAR + kNoallocSamples * kPerSample;
static __device__ __align__(256) float scratch[kScratchFloats];

}  // namespace

struct MlpWork {
    DevBuf X, T, P, G, M, V, A1, A2, G1, G2, DZ, L;
};

static int mlp_alloc(DevBuf& b, size_t bytes) {
    b.release();
    if (bytes == 0) bytes = 8;
    if (hipMalloc(&b.p, bytes) != hipSuccess) { b.p = nullptr; return AHIP_ENOMEM; }
    b.bytes = bytes;
    return hipMemset(b.p, 0, bytes) == hipSuccess ? 0 : AHIP_EHIP;
}

int mlp_run(int32_t n, const float* X, const float* T, float* params, int32_t epochs, float lr,
            float wd, float* loss_curve, float* grad_out) {
    MlpWork w;
    int rc;
    const size_t fs = sizeof(float);
    if ((rc = mlp_alloc(w.X, (size_t)n * DIN * fs))) return rc;
    if ((rc = mlp_alloc(w.T, (size_t)n * DOUT * fs))) return rc;
    if ((rc = mlp_alloc(w.P, NPAR * fs))) return rc;
    if ((rc = mlp_alloc(w.G, NPAR * fs))) return rc;
    if ((rc = mlp_alloc(w.M, NPAR * fs))) return rc;
    if ((rc = mlp_alloc(w.V, NPAR * fs))) return rc;
    if ((rc = mlp_alloc(w.A1, (size_t)n * H1 * fs))) return rc;
    if ((rc = mlp_alloc(w.A2, (size_t)n * H2 * fs))) return rc;
    if ((rc = mlp_alloc(w.G1, (size_t)n * H1 * fs))) return rc;
    if ((rc = mlp_alloc(w.G2, (size_t)n * H2 * fs))) return rc;
    if ((rc = mlp_alloc(w.DZ, (size_t)n * DOUT * fs))) return rc;
    if ((rc = mlp_alloc(w.L, (size_t)n * fs))) return rc;
    if (hipMemcpy(w.X.p, X, (size_t)n * DIN * fs, hipMemcpyHostToDevice) != hipSuccess) return AHIP_EHIP;
    if (hipMemcpy(w.T.p, T, (size_t)n * DOUT * fs, hipMemcpyHostToDevice) != hipSuccess) return AHIP_EHIP;
    if (hipMemcpy(w.P.p, params, NPAR * fs, hipMemcpyHostToDevice) != hipSuccess) return AHIP_EHIP;
    const int nbs = (n + 63) / 64, nbp = (NPAR + 255) / 256;
    std::vector<float> lh(n);
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.f, c2 = 1.f;
    const int32_t iters = epochs > 0 ? epochs : 1;
    for (int32_t e = 0; e < iters; e++) {
        hipLaunchKernelGGL(mlp_deltas, dim3(nbs), dim3(64), 0, 0, n, w.X.as<float>(), w.T.as<float>(),
                           w.P.as<float>(), w.A1.as<float>(), w.A2.as<float>(), w.G1.as<float>(),
                           w.G2.as<float>(), w.DZ.as<float>(), w.L.as<float>());
        hipLaunchKernelGGL(mlp_grads, dim3(nbp), dim3(256), 0, 0, n, w.X.as<float>(), w.A1.as<float>(),
                           w.A2.as<float>(), w.G1.as<float>(), w.G2.as<float>(), w.DZ.as<float>(),
                           w.G.as<float>());
        if (loss_curve) {
            if (hipMemcpy(lh.data(), w.L.p, (size_t)n * fs, hipMemcpyDeviceToHost) != hipSuccess) return AHIP_EHIP;
            double l = 0;
            for (int32_t s = 0; s < n; s++) l += lh[s];
            loss_curve[e] = (float)l;
        }
        if (epochs <= 0) break;  // gradient only
        c1 *= b1;
        c2 *= b2;
        hipLaunchKernelGGL(mlp_adam, dim3(nbp), dim3(256), 0, 0, w.P.as<float>(), w.G.as<float>(),
                           w.M.as<float>(), w.V.as<float>(), lr, b1, b2, 1e-8f, 1.f - c1, 1.f - c2, wd);
        if (hipGetLastError() != hipSuccess) return AHIP_EHIP;
    }
    if (hipDeviceSynchronize() != hipSuccess) return AHIP_EHIP;
    if (grad_out && hipMemcpy(grad_out, w.G.p, NPAR * fs, hipMemcpyDeviceToHost) != hipSuccess) return AHIP_EHIP;
    if (epochs > 0 && hipMemcpy(params, w.P.p, NPAR * fs, hipMemcpyDeviceToHost) != hipSuccess) return AHIP_EHIP;
    return 0;
}

int mlp_predict_run(int32_t n, const float* X, const float* params, int32_t* cls, float* logits) {
    DevBuf dX, dP, dC, dL;
    int rc;
    if ((rc = mlp_alloc(dX, (size_t)n * DIN * 4))) return rc;
    if ((rc = mlp_alloc(dP, NPAR * 4))) return rc;
    if ((rc = mlp_alloc(dC, (size_t)n * 4))) return rc;
    if ((rc = mlp_alloc(dL, (size_t)n * DOUT * 4))) return rc;
    if (hipMemcpy(dX.p, X, (size_t)n * DIN * 4, hipMemcpyHostToDevice) != hipSuccess) return AHIP_EHIP;
    if (hipMemcpy(dP.p, params, NPAR * 4, hipMemcpyHostToDevice) != hipSuccess) return AHIP_EHIP;
    hipLaunchKernelGGL(mlp_predict, dim3((n + 63) / 64), dim3(64), 0, 0, n, dX.as<float>(), dP.as<float>(),
                       dC.as<int32_t>(), dL.as<float>());
    if (hipDeviceSynchronize() != hipSuccess) return AHIP_EHIP;
    if (hipMemcpy(cls, dC.p, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) return AHIP_EHIP;
    if (logits && hipMemcpy(logits, dL.p, (size_t)n * DOUT * 4, hipMemcpyDeviceToHost) != hipSuccess) return AHIP_EHIP;
    return 0;
}

// The device address of `scratch` on the current device, looked up once per
// device (hipGetSymbolAddress with that device current).
static float* noalloc_scratch(int* rc) {
    static std::mutex mu;
    static float* ptr[64] = {nullptr};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        *rc = AHIP_EHIP;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            *rc = AHIP_EHIP;
            return nullptr;
        }
        ptr[dev] = static_cast<float*>(p);
    }
    *rc = 0;
    return ptr[dev];
}

int mlp_run_noalloc(int32_t n, const float* X, const float* T, float* P, int32_t epochs, float lr, float wd,
                    float* loss_curve, float* grad_out, hipStream_t st) {
    int rc = 0;
    float* base = noalloc_scratch(&rc);
    if (!base) return rc;
    float* G = base;
    float* M = G + NPAR;
    float* V = M + NPAR;
    float* A1 = V + NPAR;
    float* A2 = A1 + (int64_t)n * H1;
    float* G1 = A2 + (int64_t)n * H2;
    float* G2 = G1 + (int64_t)n * H1;
    float* DZ = G2 + (int64_t)n * H2;
    float* L = DZ + (int64_t)n * DOUT;
    // the previous call's contents stay in the array: clear what starts at 0
    // (the moments; every other region is written before it is read)
    hipLaunchKernelGGL(mlp_zero, dim3((2 * NPAR + 255) / 256), dim3(256), 0, st, M, (int64_t)2 * NPAR);
    const int nbs = (n + 63) / 64, nbp = (NPAR + 255) / 256;
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.f, c2 = 1.f;
    const int32_t iters = epochs > 0 ? epochs : 1;
    for (int32_t e = 0; e < iters; e++) {
        hipLaunchKernelGGL(mlp_deltas, dim3(nbs), dim3(64), 0, st, n, X, T, P, A1, A2, G1, G2, DZ, L);
        hipLaunchKernelGGL(mlp_grads, dim3(nbp), dim3(256), 0, st, n, X, A1, A2, G1, G2, DZ, G);
        if (loss_curve) hipLaunchKernelGGL(mlp_loss_sum, dim3(1), dim3(256), 0, st, n, L, loss_curve + e);
        if (epochs <= 0) break;  // gradient only
        c1 *= b1;
        c2 *= b2;
        hipLaunchKernelGGL(mlp_adam, dim3(nbp), dim3(256), 0, st, P, G, M, V, lr, b1, b2, 1e-8f, 1.f - c1,
                           1.f - c2, wd);
    }
    if (grad_out) hipLaunchKernelGGL(mlp_copy, dim3(nbp), dim3(256), 0, st, grad_out, G, (int64_t)NPAR);
    return hipGetLastError() == hipSuccess ? 0 : AHIP_EHIP;
}

}  // namespace ahip

extern "C" {

// Training as ahip_mlp_train with caller-owned device buffers and no device
// allocation: the scratch is the library's static array (up to 2048
// samples); kernels only, enqueued on `stream`.  Calls that overlap on one
// GPU share that array and race.
int ahip_mlp_train_noalloc(const float* X, const float* T, int32_t n, float* params, int32_t epochs, float lr,
                           float weight_decay, float* loss_curve, float* grad_out, void* stream) {
    if (!X || !T || !params || n <= 0 || epochs < 0) return AHIP_EINVAL;
    if (n > ahip::kNoallocSamples)
        return ahip::report_error(AHIP_EINVAL, "ahip_mlp_train_noalloc: more than 2048 samples (the size of "
                                               "its static scratch); use ahip_mlp_train");
    return ahip::mlp_run_noalloc(n, X, T, params, epochs, lr, weight_decay, loss_curve, grad_out,
                                 (hipStream_t)stream);
}


// Full-batch Adam training (epochs > 0) or one gradient evaluation
// (epochs == 0: grad_out receives dLoss/dparams, loss_curve[0] the loss).
int ahip_mlp_train(const float* X, const float* T, int32_t n, float* params, int32_t epochs, float lr,
                   float weight_decay, float* loss_curve, float* grad_out) {
    if (!X || !T || !params || n <= 0 || epochs < 0) return AHIP_EINVAL;
    return ahip::mlp_run(n, X, T, params, epochs, lr, weight_decay, loss_curve, grad_out);
}

int ahip_mlp_predict(const float* X, int32_t n, const float* params, int32_t* classes, float* logits) {
    if (!X || !params || !classes || n <= 0) return AHIP_EINVAL;
    return ahip::mlp_predict_run(n, X, params, classes, logits);
}

}  // extern "C"
