// mlp.h -- the learned eddy-viscosity closure: a per-node MLP
// 7 -> 64 -> 64 -> 64 -> 1 (GELU hidden layers, softplus output), BF16
// weights and activations, FP32 accumulation on the matrix cores.
//
// Training: every GEMM is a wavefront-level product over 64 samples built
// from v_mfma_f32_16x16x16_bf16 (wave_gemm): operand lanes l = 16 q + r hold
// row/column r and the four k = 4q..4q+3 of a 16x16x16 step, and the
// accumulator lanes hold rows 4q..4q+3 of column r.  The operands are read
// through callables, so the same routine serves the forward pass (H W), the
// data gradient (dZ W^T) and the weight gradient (H^T dZ) straight from LDS.
//
// Inference (mlp_chain_infer): the transposed products Z^T = W^T H^T, 16
// nodes per MFMA column block.  An accumulator lane then holds hidden units
// 4q..4q+3 of node r -- exactly the B-operand slots of the next layer -- so
// activations stay in registers through all three layers; the weights are
// pre-arranged (MlpChain) so each A operand is one 8- or 16-byte LDS read.
// Layers 2 and 3 use gfx950's v_mfma_f32_16x16x32_bf16, whose 32-long k
// slots 8q..8q+7 take hidden units {4q..4q+3} of one 16-unit block and
// {4q..4q+3} of the next (the pairing the accumulator layout produces).
//
// GELU is the tanh form 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))) in
// training and inference alike.
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

// GELU (tanh form) through the logistic function: 0.5 (1 + tanh u) =
// 1 / (1 + exp(-2u)) with u = sqrt(2/pi) (x + 0.044715 x^3), so
// gelu(x) = x s and gelu'(x) = s + 2 sqrt(2/pi) x s (1 - s) (1 + 3 0.044715 x^2).
// s costs one v_exp_f32 and one v_rcp_f32 (the exponent -2u log2(e) is one
// multiply-add); the limits s -> 0 (exp2 -> inf) and s -> 1 are exact.
__device__ __forceinline__ float gelu_s(float x)
{
    const float K1 = -2.3022081981443250f;  // -2 sqrt(2/pi) log2(e)
    const float K2 = -0.1029432395800235f;  // K1 * 0.044715
    const float v = x * fmaf(x * x, K2, K1);
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(v));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_s(x); }
__device__ __forceinline__ float gelu_grad(float x)
{
    const float s = gelu_s(x);
    return s + 1.5957691216057308f * x * s * (1.f - s) * fmaf(0.134145f, x * x, 1.f);
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

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// Weights arranged for the register chain (built by k_mlp_pack_chain):
// A-operand lane l = 16 q + r of output block mt holds
//   w1[mt][l][j]     = W1[4q + j][16 mt + r]              (inputs padded to 16)
//   w2[mt][kt][l][j] = W2[h(kt, q, j)][16 mt + r], h = 32 kt + 4q + j (j < 4),
//                                                      32 kt + 16 + 4q + j - 4 (j >= 4)
// and b*[mt][q] = b[16 mt + 4q .. +3], w4[mt][q] likewise.
struct MlpChain {
    bf16x4 w1[4][64];
    bf16x8 w2[4][2][64];
    bf16x8 w3[4][2][64];
    f32x4 b1[4][4], b2[4][4], b3[4][4], w4[4][4];
    float b4;
    float pad[3];
};

__device__ __forceinline__ int chain_h(int kt, int q, int j) { return 32 * kt + (j < 4 ? 4 * q + j : 12 + 4 * q + j); }

__global__ void k_mlp_pack_chain(const float* p, MlpChain* c)
{
    const int l = threadIdx.x; // 64 lanes
    const int q = l >> 4, r = l & 15;
    for (int mt = 0; mt < 4; ++mt) {
        const int col = 16 * mt + r;
        for (int j = 0; j < 4; ++j) {
            const int in = 4 * q + j;
            c->w1[mt][l][j] = in < MLP_IN ? bf(p[OFF_W1 + in * MLP_H + col]) : (short)0;
        }
        for (int kt = 0; kt < 2; ++kt)
            for (int j = 0; j < 8; ++j) {
                const int h = chain_h(kt, q, j);
                c->w2[mt][kt][l][j] = (__bf16)p[OFF_W2 + h * MLP_H + col];
                c->w3[mt][kt][l][j] = (__bf16)p[OFF_W3 + h * MLP_H + col];
            }
    }
    if (l < 16) {
        const int mt = l >> 2, qq = l & 3;
        for (int j = 0; j < 4; ++j) {
            const int h = 16 * mt + 4 * qq + j;
            c->b1[mt][qq][j] = p[OFF_B1 + h];
            c->b2[mt][qq][j] = p[OFF_B2 + h];
            c->b3[mt][qq][j] = p[OFF_B3 + h];
            c->w4[mt][qq][j] = p[OFF_W4 + h];
        }
    }
    if (l == 0) c->b4 = p[OFF_B4];
}

// cooperative copy of the chain weights into LDS
__device__ __forceinline__ void chain_stage(MlpChain& s, const MlpChain* g)
{
    const int n = (int)(sizeof(MlpChain) / 16);
    const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nt = blockDim.x * blockDim.y * blockDim.z;
    const f32x4* src = reinterpret_cast<const f32x4*>(g);
    f32x4* dst = reinterpret_cast<f32x4*>(&s);
    for (int i = t; i < n; i += nt) dst[i] = src[i];
}

// z = BF16(acc), h = BF16(GELU(z)) for the four units this lane holds; the
// accumulator started from the layer's bias, so acc = b + W h.  z is
// rounded two units at a time (one packed conversion, then each half
// widened back by a shift or a mask).
__device__ __forceinline__ void chain_act(const f32x4& acc, __bf16 h[4])
{
#pragma unroll
    for (int j = 0; j < 4; j += 2) {
        const f32x2 a = {acc[j], acc[j + 1]};
        const uint32_t zb = __builtin_bit_cast(uint32_t, __builtin_convertvector(a, bf16x2));
        h[j] = (__bf16)gelu(__builtin_bit_cast(float, zb << 16));
        h[j + 1] = (__bf16)gelu(__builtin_bit_cast(float, zb & 0xffff0000u));
    }
}

// The standardised inputs of one node as BF16 pairs: word k holds input 2k
// in its low half and input 2k+1 in its high half (input 7 is padding, 0).
struct MlpIn {
    uint32_t d[4];
};
__device__ __forceinline__ MlpIn mlp_pack_inputs(const float x[MLP_IN])
{
    MlpIn in;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t lo = (uint16_t)bf(x[2 * k]);
        const uint32_t hi = 2 * k + 1 < MLP_IN ? (uint16_t)bf(x[2 * k + 1]) : 0u;
        in.d[k] = lo | (hi << 16);
    }
    return in;
}

// The network for the 64 nodes of one wavefront; lane l passes the inputs
// of its own node.  Returns the output pre-activation of this lane's node.
// All 64 lanes must take part.
__device__ __forceinline__ float mlp_chain_infer(const MlpChain& c, const MlpIn& in)
{
    const int lane = threadIdx.x & 63;
    const int q = lane >> 4, r = lane & 15;
    float mine = 0.f;
#pragma unroll 1
    for (int nt = 0; nt < 4; ++nt) {
        // keeps the weight reads inside the loop (re-read from LDS per node
        // block) instead of 128 VGPRs of hoisted weights
        asm volatile("" ::: "memory");
        // first layer's B operand: lane (q, r) takes inputs 4q..4q+3 of node
        // 16 nt + r, gathered across lanes (inputs 8..15 are the zero padding)
        const int src = (16 * nt + r) * 4;
        const uint32_t a0 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)in.d[0]);
        const uint32_t a1 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)in.d[1]);
        const uint32_t a2 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)in.d[2]);
        const uint32_t a3 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)in.d[3]);
        const u32x2 xw = q == 0 ? u32x2{a0, a1} : (q == 1 ? u32x2{a2, a3} : u32x2{0u, 0u});
        const bf16x4 xin = __builtin_bit_cast(bf16x4, xw);
        f32x4 acc[4];
        bf16x8 hb[2];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(c.w1[mt][lane], xin, c.b1[mt][q], 0, 0, 0);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            __bf16 h[4];
            chain_act(acc[mt], h);
#pragma unroll
            for (int j = 0; j < 4; ++j) hb[mt >> 1][(mt & 1) * 4 + j] = h[j];
        }
#pragma unroll
        for (int layer = 0; layer < 2; ++layer) {
            const bf16x8(*w)[2][64] = layer == 0 ? c.w2 : c.w3;
            const f32x4(*b)[4] = layer == 0 ? c.b2 : c.b3;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                f32x4 a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[mt][0][lane], hb[0], b[mt][q], 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[mt][1][lane], hb[1], a, 0, 0, 0);
                acc[mt] = a;
            }
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                __bf16 h[4];
                chain_act(acc[mt], h);
#pragma unroll
                for (int j = 0; j < 4; ++j) hb[mt >> 1][(mt & 1) * 4 + j] = h[j];
            }
        }
        float part = 0.f;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const f32x4 w4 = c.w4[mt][q];
#pragma unroll
            for (int j = 0; j < 4; ++j) part += (float)hb[mt >> 1][(mt & 1) * 4 + j] * w4[j];
        }
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        if (nt == q) mine = part + c.b4;
    }
    return mine;
}

} // namespace lbm
