// Learned initial depth: training-set generation, training (forward,
// backward, Adam) and inference kernels for the 5 -> 32 -> 32 -> 1 MLP of
// trmc_mlp.h, plus their C-ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "trmc.h"
#include "trmc_device.h"
#include "trmc_mlp.h"
#include "trmc_solver.h"
#include "trmc_util.h"

namespace trmc {

// ------------------------------------------------------------ dataset
__device__ __forceinline__ unsigned long long dmix(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double dunif(unsigned long long key, unsigned long long i)
{
    return ((double)(dmix(key ^ dmix(i)) >> 11) + 0.5) * 0x1.0p-53;
}

__device__ __forceinline__ float f32r(double x) { return (float)x; }

// Training pair i: Q log-uniform in [1e-3, 1e4] m3/s and channel parameters
// from the workload distributions (float-representable, as the generator
// writes them); target y = log(h_normal / h_wide) with h_normal the FP64
// Manning normal depth and h_wide the FP32 wide-channel scale used at
// inference.
__global__ void k_make_dataset(long long n, unsigned long long seed, float *X, float *Y, double *raw)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long key = dmix(seed * 0x2545F4914F6CDD1Dull + 0x51ull);
    const unsigned long long base = (unsigned long long)i * 16ull;
    const double Q = f32r(pow(10.0, -3.0 + 7.0 * dunif(key, base + 0)));
    const double dx = f32r(500.0 + 4500.0 * dunif(key, base + 1));
    const double nch = f32r(0.03 + 0.05 * dunif(key, base + 2));
    const double bw = f32r(1.0 + 49.0 * dunif(key, base + 3));
    const double cs = f32r(1.0 + 2.0 * dunif(key, base + 4));
    const double tw = f32r(bw + 2.0 + 18.0 * dunif(key, base + 5));
    const double twcc = f32r(tw + 5.0 + 95.0 * dunif(key, base + 6));
    const double ncc = f32r(nch * (1.2 + 0.8 * dunif(key, base + 7)));
    const double u1 = dunif(key, base + 8), u2 = dunif(key, base + 9);
    const double gauss = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    const double s0 = f32r(fmin(1e-2, fmax(1e-4, exp(log(1e-3) + gauss))));
    Seg c;
    c.dx = dx; c.bw = bw; c.twcc = twcc; c.n_ch = nch; c.ncc = ncc; c.s0 = s0;
    c.z = 1.0 / cs;
    c.bfd = (bw > tw) ? bw / 0.00001 : (bw == tw) ? bw / (2.0 * c.z) : (tw - bw) / (2.0 * c.z);
    c.ks = sqrt(s0) / nch;
    c.sq = sqrt(1.0 + c.z * c.z);
    c.sqs0 = sqrt(s0);
    const double h = normal_depth(Q, c);
    const float hw = mlp_wide_depth((float)Q, (float)nch, (float)bw, (float)s0);
    float f[kMlpIn];
    mlp_features((float)Q, (float)nch, (float)bw, (float)cs, (float)s0, f);
    for (int k = 0; k < kMlpIn; ++k) X[(size_t)i * kMlpIn + k] = f[k];
    Y[i] = (float)log(h / (double)hw);
    if (raw) {
        double *r = raw + (size_t)i * 9;
        r[0] = Q; r[1] = dx; r[2] = bw; r[3] = tw; r[4] = twcc; r[5] = nch; r[6] = ncc; r[7] = cs; r[8] = s0;
    }
}

// ------------------------------------------------------------ training
// One block = one wave of 64 lanes, each lane one sample per chunk.
// Activations and deltas are staged in LDS, then every lane accumulates
// 21 of the 1281 gradient components over the chunk's 64 samples (no atomics).
// Per-block partial gradients go to partial[block][1281]; k_mlp_reduce_adam
// sums them in block order, so training is deterministic.
constexpr int kCompPerLane = (kMlpParams + 63) / 64;   // 21

__global__ void __launch_bounds__(64) k_mlp_grad(const float *W, const float *X, const float *Y, long long i0,
                                                 int chunks, float scale, float *partial, float *loss_part,
                                                 float *pred)
{
    __shared__ float sW[kMlpParams];
    __shared__ float sx[kMlpIn][64], sa1[kMlpHid][64], sa2[kMlpHid][64], sd1[kMlpHid][64], sd2[kMlpHid][64];
    __shared__ float sdy[64];
    const int l = threadIdx.x;
    for (int c = l; c < kMlpParams; c += 64) sW[c] = W[c];
    __syncthreads();
    float g[kCompPerLane];
#pragma unroll
    for (int q = 0; q < kCompPerLane; ++q) g[q] = 0.0f;
    float loss = 0.0f;
    for (int ch = 0; ch < chunks; ++ch) {
        const long long i = i0 + ((long long)blockIdx.x * chunks + ch) * 64 + l;
        float x[kMlpIn], a1[kMlpHid], a2[kMlpHid];
#pragma unroll
        for (int f = 0; f < kMlpIn; ++f) x[f] = X[(size_t)i * kMlpIn + f];
        const float y = mlp_forward_scalar(sW, x, a1, a2);
        if (pred) pred[i] = y;
        const float e = y - Y[i];
        loss += 0.5f * e * e;
        const float dy = e * scale;
        float d2[kMlpHid];
#pragma unroll
        for (int o = 0; o < kMlpHid; ++o) d2[o] = dy * sW[kOffW3 + o] * (1.0f - a2[o] * a2[o]);
#pragma unroll
        for (int in = 0; in < kMlpHid; ++in) {
            float s = 0.0f;
#pragma unroll
            for (int o = 0; o < kMlpHid; ++o) s += d2[o] * sW[kOffW2 + o * kMlpHid + in];
            sd1[in][l] = s * (1.0f - a1[in] * a1[in]);
        }
#pragma unroll
        for (int h = 0; h < kMlpHid; ++h) { sa1[h][l] = a1[h]; sa2[h][l] = a2[h]; sd2[h][l] = d2[h]; }
#pragma unroll
        for (int f = 0; f < kMlpIn; ++f) sx[f][l] = x[f];
        sdy[l] = dy;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kCompPerLane; ++q) {
            const int c = l + 64 * q;
            float acc = 0.0f;
            if (c < kOffB1) {                       // W1[h][f]
                const int h = c / kMlpIn, f = c % kMlpIn;
                for (int s = 0; s < 64; ++s) acc += sd1[h][s] * sx[f][s];
            } else if (c < kOffW2) {                // b1
                const int h = c - kOffB1;
                for (int s = 0; s < 64; ++s) acc += sd1[h][s];
            } else if (c < kOffB2) {                // W2[o][i]
                const int o = (c - kOffW2) / kMlpHid, in = (c - kOffW2) % kMlpHid;
                for (int s = 0; s < 64; ++s) acc += sd2[o][s] * sa1[in][s];
            } else if (c < kOffW3) {                // b2
                const int o = c - kOffB2;
                for (int s = 0; s < 64; ++s) acc += sd2[o][s];
            } else if (c < kOffB3) {                // W3
                const int h = c - kOffW3;
                for (int s = 0; s < 64; ++s) acc += sdy[s] * sa2[h][s];
            } else if (c == kOffB3) {
                for (int s = 0; s < 64; ++s) acc += sdy[s];
            }
            g[q] += acc;
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < kCompPerLane; ++q) {
        const int c = l + 64 * q;
        if (c < kMlpParams) partial[(size_t)blockIdx.x * kMlpParams + c] = g[q];
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (l == 0) loss_part[blockIdx.x] = loss;
}

__global__ void k_mlp_reduce_adam(float *W, float *m, float *v, const float *partial, const float *loss_part,
                                  int nblocks, float lr, float b1, float b2, float eps, int step, float *grad_out,
                                  float *loss_out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0 && loss_out) {
        float L = 0.0f;
        for (int b = 0; b < nblocks; ++b) L += loss_part[b];
        *loss_out += L;
    }
    if (c >= kMlpParams) return;
    float g = 0.0f;
    for (int b = 0; b < nblocks; ++b) g += partial[(size_t)b * kMlpParams + c];
    if (grad_out) grad_out[c] = g;
    if (lr > 0.0f) {
        const float mt = b1 * m[c] + (1.0f - b1) * g;
        const float vt = b2 * v[c] + (1.0f - b2) * g * g;
        m[c] = mt;
        v[c] = vt;
        const float mh = mt / (1.0f - powf(b1, (float)step));
        const float vh = vt / (1.0f - powf(b2, (float)step));
        W[c] -= lr * mh / (sqrtf(vh) + eps);
    }
}

// ------------------------------------------------------------ inference
__global__ void __launch_bounds__(64) k_mlp_predict_mfma(const float *W, const float *X, long long n, float *out)
{
    __shared__ float sW[kMlpParams];
    for (int c = threadIdx.x; c < kMlpParams; c += 64) sW[c] = W[c];
    __syncthreads();
    const long long i = (long long)blockIdx.x * 64 + threadIdx.x;
    float x[kMlpIn];
    for (int f = 0; f < kMlpIn; ++f) x[f] = (i < n) ? X[(size_t)i * kMlpIn + f] : 0.0f;
    const float y = mlp_forward_wave(sW, x);      // all 64 lanes participate
    if (i < n) out[i] = y;
}

// The routing kernel's packed form: lane i of the grid holds the sample pair
// (X[2i], X[2i+1]) and wants outputs when i % 3 != 1; results go to
// out[2i], out[2i+1] for wanting lanes, NaN for the others.
__global__ void __launch_bounds__(64) k_mlp_predict_pairs(const float *W, const float *X, long long n, float *out)
{
    __shared__ float sW[kMlpParams];
    for (int c = threadIdx.x; c < kMlpParams; c += 64) sW[c] = W[c];
    __syncthreads();
    const long long i = (long long)blockIdx.x * 64 + threadIdx.x;
    const bool want = (2 * i + 1 < n) && (i % 3 != 1);
    float xn[kMlpIn], xo[kMlpIn];
    for (int f = 0; f < kMlpIn; ++f) {
        xn[f] = want ? X[(size_t)(2 * i) * kMlpIn + f] : 0.0f;
        xo[f] = want ? X[(size_t)(2 * i + 1) * kMlpIn + f] : 0.0f;
    }
    float yn = NAN, yo = NAN;
    mlp_wave_pairs(sW, want, xn, xo, yn, yo);
    if (2 * i + 1 < n) {
        out[2 * i] = want ? yn : NAN;
        out[2 * i + 1] = want ? yo : NAN;
    }
}

__global__ void __launch_bounds__(64) k_mlp_predict_scalar(const float *W, const float *X, long long n, float *out)
{
    const long long i = (long long)blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    float x[kMlpIn], a1[kMlpHid], a2[kMlpHid];
    for (int f = 0; f < kMlpIn; ++f) x[f] = X[(size_t)i * kMlpIn + f];
    out[i] = mlp_forward_scalar(W, x, a1, a2);
}

// ------------------------------------------------------------ host side
namespace {
struct Dataset {
    float *X = nullptr, *Y = nullptr;
    long long n = 0;
} g_ds;

int hip_fail(hipError_t e, const char *what)
{
    return fail(TRMC_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define MLPCHK(expr)                                  \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

int ensure_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(TRMC_ERR_NO_DEVICE, "no HIP device visible");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(TRMC_ERR_NO_DEVICE, "device 0 is not gfx950");
    return TRMC_OK;
}

// grad pass over samples [i0, i0 + nb) of X/Y (device); nb multiple of 64
int grad_pass(const float *dW, const float *dX, const float *dY, long long i0, long long nb, float *d_partial,
              float *d_loss_part, float *d_pred, int *nblocks_out)
{
    const long long chunks_total = nb / 64;
    int nblocks = (int)std::min<long long>(256, chunks_total);
    while (chunks_total % nblocks) --nblocks;
    const int chunks = (int)(chunks_total / nblocks);
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(64), 0, nullptr, dW, dX, dY, i0, chunks,
                       1.0f / (float)nb, d_partial, d_loss_part, d_pred);
    *nblocks_out = nblocks;
    return hipGetLastError() == hipSuccess ? TRMC_OK : fail(TRMC_ERR_HIP, "k_mlp_grad launch failed");
}
}  // namespace

}  // namespace trmc

using namespace trmc;

extern "C" {

int trmc_mlp_make_dataset(int n, uint64_t seed, float *features, float *target, double *raw)
{
    if (n <= 0) return fail(TRMC_ERR_ARG, "n must be positive");
    int rc = ensure_device();
    if (rc) return rc;
    if (g_ds.X) { (void)hipFree(g_ds.X); g_ds.X = nullptr; }
    if (g_ds.Y) { (void)hipFree(g_ds.Y); g_ds.Y = nullptr; }
    g_ds.n = 0;
    MLPCHK(hipMalloc((void **)&g_ds.X, (size_t)n * kMlpIn * sizeof(float)));
    MLPCHK(hipMalloc((void **)&g_ds.Y, (size_t)n * sizeof(float)));
    double *d_raw = nullptr;
    if (raw) MLPCHK(hipMalloc((void **)&d_raw, (size_t)n * 9 * sizeof(double)));
    hipLaunchKernelGGL(k_make_dataset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, (long long)n,
                       (unsigned long long)seed, g_ds.X, g_ds.Y, d_raw);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess && features) e = hipMemcpy(features, g_ds.X, (size_t)n * kMlpIn * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && target) e = hipMemcpy(target, g_ds.Y, (size_t)n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && raw) e = hipMemcpy(raw, d_raw, (size_t)n * 9 * 8, hipMemcpyDeviceToHost);
    if (d_raw) (void)hipFree(d_raw);
    if (e != hipSuccess) return hip_fail(e, "dataset generation");
    g_ds.n = n;
    return TRMC_OK;
}

void trmc_mlp_init_weights(uint64_t seed, float *weights)
{
    if (!weights) return;
    // Glorot-uniform weights from a counter-based stream, zero biases
    auto u = [&](uint64_t i) {
        uint64_t z = seed * 0x9E3779B97F4A7C15ull + i * 0xD1B54A32D192ED03ull + 0x1234567ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
    };
    std::fill(weights, weights + kMlpParams, 0.0f);
    const double a1 = std::sqrt(6.0 / (kMlpIn + kMlpHid)), a2 = std::sqrt(6.0 / (2 * kMlpHid)),
                 a3 = std::sqrt(6.0 / (kMlpHid + 1));
    for (int c = 0; c < kOffB1; ++c) weights[kOffW1 + c] = (float)(a1 * u(c));
    for (int c = 0; c < kMlpHid * kMlpHid; ++c) weights[kOffW2 + c] = (float)(a2 * u(1000 + c));
    for (int c = 0; c < kMlpHid; ++c) weights[kOffW3 + c] = (float)(a3 * u(5000 + c));
}

int trmc_mlp_forward_backward(const float *weights, const float *features, const float *target, int n,
                              float *pred, float *grad)
{
    if (!weights || !features || !target || n <= 0 || n % 64 != 0)
        return fail(TRMC_ERR_ARG, "need weights, features, target and n a positive multiple of 64");
    int rc = ensure_device();
    if (rc) return rc;
    float *dW = nullptr, *dX = nullptr, *dY = nullptr, *dP = nullptr, *dpart = nullptr, *dlp = nullptr,
          *dg = nullptr;
    auto cleanup = [&]() {
        for (float *p : {dW, dX, dY, dP, dpart, dlp, dg})
            if (p) (void)hipFree(p);
    };
    hipError_t e = hipMalloc((void **)&dW, kMlpParams * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dX, (size_t)n * kMlpIn * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dY, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dP, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dpart, (size_t)256 * kMlpParams * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dlp, 256 * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dg, kMlpParams * 4);
    if (e == hipSuccess) e = hipMemcpy(dW, weights, kMlpParams * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dX, features, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dY, target, (size_t)n * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) { cleanup(); return hip_fail(e, "forward_backward setup"); }
    int nblocks = 0;
    rc = grad_pass(dW, dX, dY, 0, n, dpart, dlp, dP, &nblocks);
    if (!rc) {
        hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, nullptr, dW,
                           (float *)nullptr, (float *)nullptr, dpart, dlp, nblocks, 0.0f, 0.9f, 0.999f, 1e-8f, 1,
                           dg, (float *)nullptr);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess && pred) e = hipMemcpy(pred, dP, (size_t)n * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess && grad) e = hipMemcpy(grad, dg, kMlpParams * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "forward_backward");
    }
    cleanup();
    return rc;
}

int trmc_mlp_train(float *weights, int epochs, int batch, float lr, uint64_t seed, float *loss_curve)
{
    (void)seed;
    if (!weights || epochs <= 0 || batch <= 0 || batch % 64 != 0)
        return fail(TRMC_ERR_ARG, "need weights, epochs > 0 and batch a positive multiple of 64");
    if (!g_ds.X || g_ds.n < batch) return fail(TRMC_ERR_STATE, "no training set (trmc_mlp_make_dataset)");
    int rc = ensure_device();
    if (rc) return rc;
    float *dW = nullptr, *dm = nullptr, *dv = nullptr, *dpart = nullptr, *dlp = nullptr, *dloss = nullptr;
    auto cleanup = [&]() {
        for (float *p : {dW, dm, dv, dpart, dlp, dloss})
            if (p) (void)hipFree(p);
    };
    hipError_t e = hipMalloc((void **)&dW, kMlpParams * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dm, kMlpParams * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dv, kMlpParams * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dpart, (size_t)256 * kMlpParams * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dlp, 256 * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dloss, 4);
    if (e == hipSuccess) e = hipMemcpy(dW, weights, kMlpParams * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(dm, 0, kMlpParams * 4);
    if (e == hipSuccess) e = hipMemset(dv, 0, kMlpParams * 4);
    if (e != hipSuccess) { cleanup(); return hip_fail(e, "train setup"); }
    const long long per_epoch = g_ds.n / batch;
    int step = 0;
    for (int ep = 0; ep < epochs && !rc; ++ep) {
        // step decay: full rate, then x0.3 after 60 % and again after 85 % of the epochs
        const float frac = (float)ep / (float)epochs;
        const float lr_ep = lr * (frac < 0.6f ? 1.0f : frac < 0.85f ? 0.3f : 0.09f);
        e = hipMemset(dloss, 0, 4);
        for (long long b = 0; b < per_epoch && e == hipSuccess && !rc; ++b) {
            int nblocks = 0;
            rc = grad_pass(dW, g_ds.X, g_ds.Y, b * batch, batch, dpart, dlp, nullptr, &nblocks);
            if (rc) break;
            ++step;
            hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, nullptr, dW, dm,
                               dv, dpart, dlp, nblocks, lr_ep, 0.9f, 0.999f, 1e-8f, step, (float *)nullptr, dloss);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) { rc = hip_fail(e, "training"); break; }
        if (loss_curve) {
            float L = 0.0f;
            e = hipMemcpy(&L, dloss, 4, hipMemcpyDeviceToHost);
            loss_curve[ep] = L / (float)(per_epoch * batch);
        }
    }
    if (!rc) {
        e = hipMemcpy(weights, dW, kMlpParams * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "train readback");
    }
    cleanup();
    return rc;
}

int trmc_mlp_predict(const float *weights, const float *features, int n, float *out, int use_mfma)
{
    if (!weights || !features || !out || n <= 0) return fail(TRMC_ERR_ARG, "bad argument");
    if (use_mfma == 2 && n % 2 != 0) return fail(TRMC_ERR_ARG, "the packed form takes sample pairs (n even)");
    int rc = ensure_device();
    if (rc) return rc;
    float *dW = nullptr, *dX = nullptr, *dO = nullptr;
    hipError_t e = hipMalloc((void **)&dW, kMlpParams * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dX, (size_t)n * kMlpIn * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dO, (size_t)n * 4);
    if (e == hipSuccess) e = hipMemcpy(dW, weights, kMlpParams * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dX, features, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const dim3 grid((unsigned)((n + 63) / 64));
        if (use_mfma == 2)
            hipLaunchKernelGGL(k_mlp_predict_pairs, dim3((unsigned)((n / 2 + 63) / 64)), dim3(64), 0, nullptr,
                               dW, dX, (long long)n, dO);
        else if (use_mfma)
            hipLaunchKernelGGL(k_mlp_predict_mfma, grid, dim3(64), 0, nullptr, dW, dX, (long long)n, dO);
        else
            hipLaunchKernelGGL(k_mlp_predict_scalar, grid, dim3(64), 0, nullptr, dW, dX, (long long)n, dO);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, dO, (size_t)n * 4, hipMemcpyDeviceToHost);
    for (float *p : {dW, dX, dO})
        if (p) (void)hipFree(p);
    return e == hipSuccess ? TRMC_OK : hip_fail(e, "predict");
}

}  // extern "C"
