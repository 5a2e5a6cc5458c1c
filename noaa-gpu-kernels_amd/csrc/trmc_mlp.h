// 5 -> 32 -> 32 -> 1 tanh MLP proposing the initial depth of the converged
// secant, FP32.
//
// Inputs (normalised in mlp_features): log10 Q, Manning n, bottom width,
// side slope cs, log10 S0.  Output y: the log ratio of the Manning normal
// depth to the wide-channel estimate h_wide = (Q n / (bw sqrt S0))^(3/5), so
// the guess is h_wide * exp(y).
//
// Weight layout (TRMC_MLP_NPARAM = 1281 floats):
//   W1[32][5] at 0, b1[32] at 160, W2[32][32] at 192 (W2[o][i]), b2[32] at 1216,
//   W3[32] at 1248, b3 at 1280.
#pragma once

#include <hip/hip_runtime.h>

#include "trmc_solver.h"

namespace trmc {

constexpr int kMlpIn = 5, kMlpHid = 32;
constexpr int kOffW1 = 0, kOffB1 = 160, kOffW2 = 192, kOffB2 = 1216, kOffW3 = 1248, kOffB3 = 1280;
constexpr int kMlpParams = 1281;

__host__ __device__ inline void mlp_features(float Q, float n, float bw, float cs, float s0, float *f)
{
    f[0] = (log10f(Q) - 0.5f) * 0.5f;
    f[1] = (n - 0.055f) * 40.0f;
    f[2] = (bw - 25.5f) * (1.0f / 24.5f);
    f[3] = cs - 2.0f;
    f[4] = log10f(s0) + 3.0f;
}

// wide-channel Manning depth, the scale the network corrects
__host__ __device__ inline float mlp_wide_depth(float Q, float n, float bw, float s0)
{
    return powf(Q * n / (bw * sqrtf(s0)), 0.6f);
}

__device__ __forceinline__ float fast_tanh(float x)
{
    // 1 - 2 / (1 + e^{2x}) with the hardware exp2 and reciprocal (about one
    // ulp each): e^{2x} = 2^{2x / ln 2}.  Saturates to +-1 (exp2 -> inf or 0).
    const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
    return fmaf(-2.0f, __builtin_amdgcn_rcpf(1.0f + e), 1.0f);
}

// Scalar forward for one sample (training kernels); a1, a2 receive the
// hidden activations.
__device__ __forceinline__ float mlp_forward_scalar(const float *W, const float *x, float *a1, float *a2)
{
#pragma unroll
    for (int h = 0; h < kMlpHid; ++h) {
        float z = W[kOffB1 + h];
#pragma unroll
        for (int f = 0; f < kMlpIn; ++f) z += W[kOffW1 + h * kMlpIn + f] * x[f];
        a1[h] = fast_tanh(z);
    }
#pragma unroll
    for (int o = 0; o < kMlpHid; ++o) {
        float z = W[kOffB2 + o];
#pragma unroll
        for (int i = 0; i < kMlpHid; ++i) z += W[kOffW2 + o * kMlpHid + i] * a1[i];
        a2[o] = fast_tanh(z);
    }
    float y = W[kOffB3];
#pragma unroll
    for (int h = 0; h < kMlpHid; ++h) y += W[kOffW3 + h] * a2[h];
    return y;
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Wave-wide forward with v_mfma_f32_16x16x4f32.  Every lane passes its
// sample's 5 features and gets its own output; must be called by all 64
// lanes of the wave together.  W may point to LDS or global memory.
//
// Operand layouts of that instruction for lane l, r = 0..3:
//   A (16x4): A[l % 16][l / 16];   B (4x16): B[l / 16][l % 16];
//   D (16x16): D[4 (l / 16) + r][l % 16].
// The network is evaluated transposed, hidden units as rows and the wave's
// 64 samples as 4 column tiles nj of 16; one tile at a time:
//   layer 1: H1^T (32x16) = W1 (32x8, features padded to 8) X^T (8x16):
//            A = W1[16 mi + l%16][4 kk + l/16], B = feature 4 kk + l/16 of
//            sample 16 nj + l%16 (shuffled from its lane)
//   layer 2: H2^T = W2 H1^T, with k-steps kappa = (mi, r): lane l's D
//            register r of row tile mi is exactly the B operand
//            H1^T[16 mi + 4 (l/16) + r][l%16], so A = W2[16 mo + l%16][16 mi + 4 (l/16) + r]
//   layer 3: y = W3 . H2^T + b3, partial dot per lane, summed over the four
//            lane groups that hold the same sample column.
// Only the first `tiles` column tiles (wave-uniform, 1..4) are evaluated:
// lanes 16 tiles .. 63 get an unspecified output.
__device__ __forceinline__ float mlp_forward_wave(const float *W, const float *x, int tiles = 4)
{
    const int l = threadIdx.x & 63, li = l & 15, lg = l >> 4;
    float mine = 0.0f;
#pragma unroll
    for (int nj = 0; nj < 4; ++nj) {
        if (nj >= tiles) break;
        const int src = 16 * nj + li;
        float xs[kMlpIn];
#pragma unroll
        for (int f = 0; f < kMlpIn; ++f) xs[f] = __shfl(x[f], src, 64);
        const float b0 = (lg == 0) ? xs[0] : (lg == 1) ? xs[1] : (lg == 2) ? xs[2] : xs[3];
        const float b1v = (lg == 0) ? xs[4] : 0.0f;
        floatx4 h1[2];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) {
            const int row = 16 * mi + 4 * lg;
            floatx4 acc = {W[kOffB1 + row], W[kOffB1 + row + 1], W[kOffB1 + row + 2], W[kOffB1 + row + 3]};
            const float a0 = W[kOffW1 + (16 * mi + li) * kMlpIn + lg];
            const float a1 = (lg == 0) ? W[kOffW1 + (16 * mi + li) * kMlpIn + 4] : 0.0f;
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1v, acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[r] = fast_tanh(acc[r]);
            h1[mi] = acc;
        }
        float p = 0.0f;
#pragma unroll
        for (int mo = 0; mo < 2; ++mo) {
            const int row = 16 * mo + 4 * lg;
            floatx4 acc = {W[kOffB2 + row], W[kOffB2 + row + 1], W[kOffB2 + row + 2], W[kOffB2 + row + 3]};
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(
                        W[kOffW2 + (16 * mo + li) * kMlpHid + 16 * mi + 4 * lg + r], h1[mi][r], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) p = fmaf(W[kOffW3 + row + r], fast_tanh(acc[r]), p);
        }
        p += __shfl_xor(p, 16, 64);
        p += __shfl_xor(p, 32, 64);
        if (lg == nj) mine = p;
    }
    return mine + W[kOffB3];
}

// Network inputs for flow Q at a segment (normalised features of
// mlp_features; cs = 1 / z).
__device__ __forceinline__ void mlp_segment_inputs(double Q, const Seg &c, float *x)
{
    mlp_features((float)fmax(Q, 1e-6), (float)c.n_ch, (float)c.bw, (float)(1.0 / c.z), (float)c.s0, x);
}

// The LEARNED guess: from the anchor depth d, whose Manning flow is q_old,
// along the network's rating curve h(Q) = h_wide(Q) e^{y(Q)} to the routed
// flow q_new:  h = d (q_new / q_old)^(3/5) e^{y_new - y_old}.  The ratio
// cancels the per-segment bias of h(Q) (the floodplain geometry is not among
// the network's inputs).
__device__ __forceinline__ double mlp_guess_from(double d, double q_new, double q_old, float y_new, float y_old)
{
    const float r = (float)(fmax(q_new, 1e-6) / fmax(q_old, 1e-6));
    return d * (double)__expf(0.6f * __logf(r) + (y_new - y_old));
}

// Bit index of the n-th (0-based) set bit of m, for n < popcount(m):
// binary search on popcounts of the low half of the remaining window.
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int n)
{
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const int cnt = __popcll((m >> pos) & ((1ull << w) - 1ull));
        if (n >= cnt) {
            n -= cnt;
            pos += w;
        }
    }
    return pos;
}

// Network outputs for the lanes with `want`, two samples per lane (xn, xo):
// lanes 2j and 2j + 1 of a forward evaluate the new and the old sample of the
// j-th wanting lane, so one wave forward serves 32 wanting lanes.  Must be called by all
// 64 lanes together; returns the number of forwards run.
__device__ __forceinline__ int mlp_wave_pairs(const float *W, bool want, const float *xn, const float *xo,
                                              float &yn, float &yo)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long mask = __ballot(want);
    const int nwant = __popcll(mask);
    const int rank = __popcll(mask & ((1ull << lane) - 1ull));
    int calls = 0;
    for (int base = 0; base < nwant; base += 32) {    // wave-uniform
        const int j = base + (lane >> 1);
        const int src = (j < nwant) ? nth_set_bit(mask, j) : lane;
        float x[kMlpIn];
#pragma unroll
        for (int f = 0; f < kMlpIn; ++f) {
            const float vn = __shfl(xn[f], src, 64), vo = __shfl(xo[f], src, 64);
            x[f] = (lane & 1) ? vo : vn;
        }
        // samples sit in lanes 0 .. 2 cnt - 1: only their tiles are run
        const int cnt = (nwant - base < 32) ? nwant - base : 32;
        const float y = mlp_forward_wave(W, x, (2 * cnt + 15) >> 4);
        const int r = rank - base;
        const bool mine = want && r >= 0 && r < 32;
        const int at = mine ? 2 * r : 0;
        const float y0 = __shfl(y, at, 64), y1 = __shfl(y, at + 1, 64);
        if (mine) {
            yn = y0;
            yo = y1;
        }
        ++calls;
    }
    return calls;
}

// One lane's guess with the scalar forward (schedules that solve one
// segment per lane at a time); W in global memory.
__device__ __forceinline__ double mlp_guess_scalar(const float *W, double d, double q_new, double q_old,
                                                   const Seg &c)
{
    float xn[kMlpIn], xo[kMlpIn], a1[kMlpHid], a2[kMlpHid];
    mlp_segment_inputs(q_new, c, xn);
    mlp_segment_inputs(q_old, c, xo);
    const float yn = mlp_forward_scalar(W, xn, a1, a2);
    const float yo = mlp_forward_scalar(W, xo, a1, a2);
    return mlp_guess_from(d, q_new, q_old, yn, yo);
}

}  // namespace trmc
