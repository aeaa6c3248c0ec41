// mlp.h -- the learned eddy-viscosity closure: a per-node MLP
// 7 -> 64 -> 64 -> 64 -> 1 (GELU hidden layers, softplus output), BF16
// weights and activations, FP32 accumulation on the matrix cores.
//
// Every GEMM is a wavefront-level product over 64 rows (64 nodes or 64
// training samples) built from v_mfma_f32_16x16x16_bf16: operand lanes
// l = 16 q + r hold row/column r and the four k = 4q..4q+3 of a 16x16x16
// step, and the 4 accumulator lanes hold rows 4q..4q+3 of column r.  The
// operands are read through callables, so the same routine serves the
// forward pass (H W), the data gradient (dZ W^T) and the weight gradient
// (H^T dZ) straight from LDS in whatever layout they sit.
//
// No reference counterpart: the reference's closure is constant-Cs
// Smagorinsky (MRTtest/DynInlet3.cu:628-735); its Pi_neq (:722-729) and
// StrainRate (:367-393) are where this model's inputs come from.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "storage.h"

namespace lbm {

constexpr int MLP_IN = 7;    // inputs
constexpr int MLP_INP = 16;  // inputs padded to one MFMA k-step
constexpr int MLP_H = 64;    // hidden width
// flat FP32 parameter vector: W1[7][64] b1[64] W2[64][64] b2 W3[64][64] b3 w4[64] b4
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + MLP_IN * MLP_H;
constexpr int OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;
constexpr int OFF_W3 = OFF_B2 + MLP_H;
constexpr int OFF_B3 = OFF_W3 + MLP_H * MLP_H;
constexpr int OFF_W4 = OFF_B3 + MLP_H;
constexpr int OFF_B4 = OFF_W4 + MLP_H;
constexpr int MLP_NPARAM = OFF_B4 + 1; // 8897

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf(float v) { return (short)float_to_bf16_bits(v); }
__device__ __forceinline__ float fb(short h) { return bf16_bits_to_float((uint16_t)h); }

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x)
{
    return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}
__device__ __forceinline__ float softplus(float z) { return fmaxf(z, 0.f) + log1pf(expf(-fabsf(z))); }
__device__ __forceinline__ float sigmoid(float z) { return 1.f / (1.f + expf(-z)); }

// C[M x N] = A[M x K] B[K x N] for one wavefront; a(m,k), b(k,n) return BF16
// bits, out(m, n, v) receives each FP32 result.  M, N, K multiples of 16.
// Must run with all 64 lanes active.
template <int M, int N, int K, class FA, class FB, class FO>
__device__ __forceinline__ void wave_gemm(FA&& a, FB&& b, FO&& out)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, q = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < M / 16; ++mt)
#pragma unroll
        for (int nt = 0; nt < N / 16; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < K / 16; ++kt) {
                bf16x4 av, bv;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    av[j] = a(mt * 16 + r, kt * 16 + 4 * q + j);
                    bv[j] = b(kt * 16 + 4 * q + j, nt * 16 + r);
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) out(mt * 16 + 4 * q + j, nt * 16 + r, acc[j]);
        }
}

// Model weights staged in LDS: BF16 matrices (in x out, row-major), FP32
// biases and output layer.
struct MlpLds {
    short w1[MLP_INP * MLP_H];
    short w2[MLP_H * MLP_H];
    short w3[MLP_H * MLP_H];
    float b1[MLP_H], b2[MLP_H], b3[MLP_H], w4[MLP_H];
    float b4;
};

// Feature normalisation and validity range that travel with the weights.
struct MlpNorm {
    float mu[MLP_IN], inv_sd[MLP_IN];
    float lo[MLP_IN], hi[MLP_IN]; // training range of the raw inputs
    float nu_scale;               // output = nu_scale * softplus(.)
};

// cooperative load of the BF16 weight copies (wbf: W1 padded [16][64], W2, W3)
// and the FP32 vectors from the flat parameter vector
__device__ __forceinline__ void mlp_stage(MlpLds& s, const short* wbf, const float* p)
{
    for (int i = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
         i < MLP_INP * MLP_H + 2 * MLP_H * MLP_H; i += blockDim.x * blockDim.y * blockDim.z) {
        if (i < MLP_INP * MLP_H) s.w1[i] = wbf[i];
        else if (i < MLP_INP * MLP_H + MLP_H * MLP_H) s.w2[i - MLP_INP * MLP_H] = wbf[i];
        else s.w3[i - MLP_INP * MLP_H - MLP_H * MLP_H] = wbf[i];
    }
    for (int i = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z); i < MLP_H;
         i += blockDim.x * blockDim.y * blockDim.z) {
        s.b1[i] = p[OFF_B1 + i];
        s.b2[i] = p[OFF_B2 + i];
        s.b3[i] = p[OFF_B3 + i];
        s.w4[i] = p[OFF_W4 + i];
    }
    if (threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) s.b4 = p[OFF_B4];
}

// First layer (inference): buf[64][64] = BF16(GELU(x W1 + b1)), x BF16 [64][16].
__device__ __forceinline__ void mlp_layer1(const MlpLds& s, const short* x, short* buf)
{
    wave_gemm<64, 64, 16>([&](int m, int k) { return x[m * MLP_INP + k]; },
                          [&](int k, int n) { return s.w1[k * MLP_H + n]; },
                          [&](int m, int n, float v) { buf[m * MLP_H + n] = bf(gelu(fb(bf(v + s.b1[n])))); });
}

// Hidden layer in place (inference): buf holds activations h = BF16(GELU(z))
// and becomes BF16(GELU(BF16(buf W + b))) -- the same roundings as the
// training forward pass, which keeps z in BF16.  The 64x64 results are
// held in registers until every lane of the wavefront has read its operands.
// Contains block barriers: every wavefront of the workgroup must call it.
__device__ __forceinline__ void mlp_layer_inplace(short* buf, const short* w, const float* b)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, q = lane >> 4;
    float res[16][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                bf16x4 av, bv;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    av[j] = buf[(mt * 16 + r) * MLP_H + kt * 16 + 4 * q + j];
                    bv[j] = w[(kt * 16 + 4 * q + j) * MLP_H + nt * 16 + r];
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) res[mt * 4 + nt][j] = acc[j];
        }
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                buf[(mt * 16 + 4 * q + j) * MLP_H + nt * 16 + r] =
                    bf(gelu(fb(bf(res[mt * 4 + nt][j] + b[nt * 16 + r]))));
    __syncthreads();
}

// Output layer (64 -> 1) on the vector ALUs: this lane's row of the
// activations buf.
__device__ __forceinline__ float mlp_head(const MlpLds& s, const short* buf)
{
    const int lane = threadIdx.x & 63;
    float zo = s.b4;
#pragma unroll 8
    for (int i = 0; i < MLP_H; ++i) zo += fb(buf[lane * MLP_H + i]) * s.w4[i];
    return zo;
}

// Inference for the 64 rows of one wavefront (x: BF16 [64][16], buf: BF16
// [64][64] scratch in LDS); returns this lane's output pre-activation.
__device__ __forceinline__ float mlp_infer(const MlpLds& s, const short* x, short* buf)
{
    mlp_layer1(s, x, buf);
    __syncthreads();
    mlp_layer_inplace(buf, s.w2, s.b2);
    mlp_layer_inplace(buf, s.w3, s.b3);
    return mlp_head(s, buf);
}

} // namespace lbm
