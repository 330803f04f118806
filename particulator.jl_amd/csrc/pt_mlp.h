// pt_mlp.h -- learned inverse-CDF samplers for the collision step on gfx950.
//
// A sampler is an MLP  (log10 T, one-hot target Z in {7, 8, 18}, u1, u2, phase)
//                        -> 3 hidden layers x 64, GELU -> 2 outputs
// (phase: position of T in its Seltzer-Berger energy-grid cell, 0 for Moller)
// trained so that for uniform u1, u2 its outputs are quantiles of the exact sampler's
// distribution at (T, Z):
//   bremsstrahlung (SeltzerBerger + modified Tsai angle): y0 = log(k/T) (quantile u1),
//                                                         y1 = log(1 - cos theta) (u2);
//   Moller: y0 = log(eps/eps0) / log(0.5/eps0), eps0 = tcut/T (quantile u1); y1 unused.
// Training runs in FP32 on the device (forward, backward and Adam kernels below);
// inference runs with BF16 weights and activations on the matrix cores
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation), one wave evaluating 64 samples.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pt {

constexpr int MLP_IN = 16;   // inputs padded to one MFMA K step (7 used)
constexpr int MLP_H = 64;
constexpr int MLP_OUT = 2;
constexpr int MLP_OUTP = 32;  // output rows padded to one 32-wide MFMA tile

// FP32 parameter vector used by training: W[o][i] row-major, then the bias of each layer
constexpr int P_W1 = 0;
constexpr int P_B1 = P_W1 + MLP_H * MLP_IN;
constexpr int P_W2 = P_B1 + MLP_H;
constexpr int P_B2 = P_W2 + MLP_H * MLP_H;
constexpr int P_W3 = P_B2 + MLP_H;
constexpr int P_B3 = P_W3 + MLP_H * MLP_H;
constexpr int P_W4 = P_B3 + MLP_H;
constexpr int P_B4 = P_W4 + MLP_OUT * MLP_H;
constexpr int MLP_NPARAM = P_B4 + MLP_OUT;

// BF16 inference image: W1[64][16], W2[64][64], W3[64][64], W4[32][64] (rows >= 2 zero);
// then FP32 biases b1[64], b2[64], b3[64], b4[32]
constexpr int I_W1 = 0;
constexpr int I_W2 = I_W1 + MLP_H * MLP_IN;
constexpr int I_W3 = I_W2 + MLP_H * MLP_H;
constexpr int I_W4 = I_W3 + MLP_H * MLP_H;
constexpr int I_NW = I_W4 + MLP_OUTP * MLP_H;  // bf16 elements
constexpr int I_NB = 3 * MLP_H + MLP_OUTP;     // floats
constexpr int MLP_IMAGE_BYTES = I_NW * 2 + I_NB * 4;
// per-wave LDS scratch of the inference: activations [64][64] bf16 + outputs [64][2] f32
constexpr int MLP_WAVE_BYTES = 64 * MLP_H * 2 + 64 * MLP_OUT * 4;

enum { LEARN_SB = 0, LEARN_MOLLER = 1, NLEARN = 2 };

struct MlpDev {
    const uint16_t* w;  // BF16 image (global memory)
    const float* b;
    int on;
    int pad;
    double emin, emax;  // energies (J) the step uses the network for; the rejection sampler elsewhere
};

// input features of a sample
__host__ __device__ inline void mlp_features(double T_eV, int Z, double u1, double u2, double phase,
                                             float x[MLP_IN]) {
    for (int q = 0; q < MLP_IN; q++) x[q] = 0.0f;
    x[0] = (float)((log10(T_eV) - 6.0) * 0.5);
    x[1] = Z == 7 ? 1.0f : 0.0f;
    x[2] = Z == 8 ? 1.0f : 0.0f;
    x[3] = Z == 18 ? 1.0f : 0.0f;
    x[4] = (float)(2.0 * u1 - 1.0);
    x[5] = (float)(2.0 * u2 - 1.0);
    x[6] = (float)(2.0 * phase - 1.0);
}

// The network predicts residuals around a closed-form baseline of its outputs, a function
// of the (FP32) features, so that the BF16 error of the network applies to a small residual:
// bremsstrahlung  log(k/T)         ~ (1 - u1) log(100 eV / T)  (dsigma/dk ~ 1/k),
//                 log(1 - cos th)  ~ -2 log(2 (1 + T / mc^2))   (the Tsai scale umax);
// Moller: no baseline.  output = network + baseline.
constexpr float LN10 = 2.302585093f;
constexpr float MC2_EV = 510998.95f;
__host__ __device__ inline void mlp_baseline(int which, const float x[MLP_IN], float& b0, float& b1) {
    if (which != LEARN_SB) {
        b0 = 0.0f;
        b1 = 0.0f;
        return;
    }
    const float lgT = LN10 * (2.0f * x[0] + 6.0f);  // ln(T / eV)
    const float u1 = 0.5f * (x[4] + 1.0f);
    b0 = (1.0f - u1) * (4.605170186f - lgT);          // ln(100) - ln T
    b1 = -2.0f * logf(2.0f * (1.0f + expf(lgT) / MC2_EV));
}

// GELU, tanh form: 0.5 z (1 + tanh(c (z + 0.044715 z^3))), c = sqrt(2/pi)
constexpr float GELU_C = 0.7978845608028654f;
constexpr float GELU_A = 0.044715f;

__device__ __forceinline__ float gelu_tanh(float a, bool fast) {
    // tanh(a) = 1 - 2 / (exp(2a) + 1): saturates correctly at both ends
    float e = fast ? __expf(2.0f * a) : expf(2.0f * a);
    return 1.0f - 2.0f / (e + 1.0f);
}

__device__ __forceinline__ float gelu(float z, bool fast) {
    float a = GELU_C * (z + GELU_A * z * z * z);
    return 0.5f * z * (1.0f + gelu_tanh(a, fast));
}

__device__ __forceinline__ float gelu_grad(float z) {
    float a = GELU_C * (z + GELU_A * z * z * z);
    float th = gelu_tanh(a, false);
    return 0.5f * (1.0f + th) + 0.5f * z * (1.0f - th * th) * GELU_C * (1.0f + 3.0f * GELU_A * z * z);
}

// float -> bf16, round to nearest even (finite inputs)
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ bf16x8_t lds_frag(const uint16_t* p) {
    uint4 v = *(const uint4*)p;  // 8 consecutive bf16, 16-byte aligned
    return __builtin_bit_cast(bf16x8_t, v);
}

// v_mfma_f32_32x32x16_bf16 operands: lane l supplies row (l % 32) of A and column (l % 32)
// of B, with k = 8 (l / 32) + 0..7 of the 16-wide K step; result element r of lane l is
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
__device__ __forceinline__ int mfma_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// one dense layer on the wave's 64 rows: D = act[64][K] * W[N][K]^T, N = 64 (two column
// tiles) or 32 (one); K = 16 or 64.  acc[mi][nj] receives the 32x32 tiles.
template <int K, int NT>
__device__ __forceinline__ void mlp_layer(const uint16_t* act, const uint16_t* W, int lane, f32x16_t acc[2][2]) {
    const int r32 = lane & 31, kh = 8 * (lane >> 5);
    for (int mi = 0; mi < 2; mi++)
        for (int nj = 0; nj < NT; nj++) acc[mi][nj] = f32x16_t{};
#pragma unroll
    for (int kk = 0; kk < K / 16; kk++) {
        bf16x8_t a0 = lds_frag(act + (size_t)(r32)*MLP_H + 16 * kk + kh);
        bf16x8_t a1 = lds_frag(act + (size_t)(32 + r32) * MLP_H + 16 * kk + kh);
#pragma unroll
        for (int nj = 0; nj < NT; nj++) {
            bf16x8_t b = lds_frag(W + (size_t)(32 * nj + r32) * K + 16 * kk + kh);
            acc[0][nj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, acc[0][nj], 0, 0, 0);
            acc[1][nj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, acc[1][nj], 0, 0, 0);
        }
    }
}

// LDS written by some lanes of the wave and read by others: the wave's LDS accesses run
// in order, so only the compiler must keep them in program order
__device__ __forceinline__ void wave_lds_order() { __asm__ volatile("" ::: "memory"); }

// Evaluate the MLP for the 64 samples of a wave (every lane calls it; idle lanes pass
// zeros).  wl / bl: the BF16 weights and FP32 biases of the image (global memory or
// LDS); act / outb: this wave's LDS scratch.
__device__ __forceinline__ void mlp_wave(const uint16_t* wl, const float* bl, uint16_t* act, float* outb,
                                         const float x[MLP_IN], float& y0, float& y1) {
    const int lane = threadIdx.x & 63;
    // input row of this lane's sample (columns 0..15 of the activation buffer)
    wave_lds_order();
    for (int q = 0; q < MLP_IN; q++) act[lane * MLP_H + q] = f2bf(x[q]);
    wave_lds_order();
    f32x16_t acc[2][2];
    mlp_layer<MLP_IN, 2>(act, wl + I_W1, lane, acc);
    const float* bias = bl;
    for (int layer = 0; layer < 3; layer++) {
        // bias + GELU, written back as the next layer's input
        wave_lds_order();
        for (int mi = 0; mi < 2; mi++)
            for (int nj = 0; nj < 2; nj++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    int row = 32 * mi + mfma_row(r, lane), col = 32 * nj + (lane & 31);
                    act[row * MLP_H + col] = f2bf(gelu(acc[mi][nj][r] + bias[col], true));
                }
        bias += MLP_H;
        wave_lds_order();
        if (layer == 0) mlp_layer<MLP_H, 2>(act, wl + I_W2, lane, acc);
        else if (layer == 1) mlp_layer<MLP_H, 2>(act, wl + I_W3, lane, acc);
    }
    mlp_layer<MLP_H, 1>(act, wl + I_W4, lane, acc);
    wave_lds_order();
    if ((lane & 31) < MLP_OUT) {
        int col = lane & 31;
        for (int mi = 0; mi < 2; mi++)
#pragma unroll
            for (int r = 0; r < 16; r++) outb[(32 * mi + mfma_row(r, lane)) * MLP_OUT + col] = acc[mi][0][r] + bias[col];
    }
    wave_lds_order();
    y0 = outb[lane * MLP_OUT];
    y1 = outb[lane * MLP_OUT + 1];
    wave_lds_order();
}

// stage a BF16 image into LDS (block-wide; caller synchronises)
__device__ __forceinline__ void mlp_stage(const MlpDev& m, uint16_t* wl, float* bl) {
    for (int q = threadIdx.x; q < I_NW / 8; q += blockDim.x) ((uint4*)wl)[q] = ((const uint4*)m.w)[q];
    for (int q = threadIdx.x; q < I_NB; q += blockDim.x) bl[q] = m.b[q];
}

}  // namespace pt
