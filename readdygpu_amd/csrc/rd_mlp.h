// rd_mlp.h -- BF16 MLP 3 -> 64 -> 64 -> 64 -> 3 (SiLU) on gfx950 MFMA.
//
// One wave evaluates 64 samples with v_mfma_f32_16x16x32_bf16 (FP32
// accumulation).  Lane l = 16 g + j holds, for every sample tile tj (samples
// 16 tj + j) and feature tile ti, the four features 16 ti + 4 g + r, r = 0..3:
// exactly the accumulator layout of the 16x16 MFMA (column j, rows 4g + r).
// The same registers serve as the B operand of the next layer when the
// weights' K order is permuted to match: logical k = 8 g + t of K-block kb is
// feature F(kb, g, t) = (2 kb + t/4) 16 + 4 g + t%4.  Activations therefore
// never leave registers in the forward pass.  Layer 1 (K = 3) runs on the
// VALU in FP32 (positions need more than BF16's 8 bits); layer 4 runs on MFMA
// with its 3 output rows padded to 16.
//
// Weight vector (FP32 master, 8771 values):
//   W1[64][3] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3]
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RD_MLP_NPARAM 8771
#define RD_MLP_W1 0
#define RD_MLP_B1 192
#define RD_MLP_W2 256
#define RD_MLP_B2 4352
#define RD_MLP_W3 4416
#define RD_MLP_B3 8512
#define RD_MLP_W4 8576
#define RD_MLP_B4 8768
#define RD_MLP_ROW 8776          // partial-gradient row: 8771 grads + loss, padded

typedef __bf16 rd_bf16x8 __attribute__((ext_vector_type(8)));
typedef float rd_f32x4 __attribute__((ext_vector_type(4)));

// BF16 weights in MFMA A-operand order (16 B per lane per fragment)
struct RdMlpPack {
    rd_bf16x8 w2[4][2][64];      // A = W2, K permuted           [out tile][kb][lane]
    rd_bf16x8 w3[4][2][64];
    rd_bf16x8 w4[2][64];         // A = W4, rows 3..15 zero
    rd_bf16x8 w2t[4][2][64];     // A = W2^T, K permuted (backward)
    rd_bf16x8 w3t[4][2][64];
    rd_bf16x8 w4t[4][64];        // A = W4^T, K = 3 real of 32
    float w1[64 * 3];            // BF16-rounded W1 as FP32
    float b1[64], b2[64], b3[64], b4[4];
    // inference copies: layer 1 and the biases times -log2(e), W4 times
    // -ln 2 (the forward pass then works on -log2(e) z throughout)
    float w1s[64 * 3];
    float b1s[64], b2s[64], b3s[64];
    rd_bf16x8 w4s[2][64];
};

#define RD_MLP_NLOG2E -1.44269504f   // -log2(e)
#define RD_MLP_NLN2 -0.693147181f    // -ln(2) = 1 / -log2(e)

// position normalisation and output scale of a trained surrogate
struct RdMlpNorm {
    float lo[3];
    float sc[3];                 // 2 / (hi - lo)
    float gscale;                // output * gscale = dU/dx
    float pad;
};

// K permutation: feature held by (K-block kb, lane group g, element t)
__host__ __device__ inline int rd_mlp_feat(int kb, int g, int t) { return (2 * kb + t / 4) * 16 + 4 * g + (t & 3); }

// SiLU with the hardware exp2 and reciprocal (v_exp_f32, v_rcp_f32: one
// transcendental each, ~1 ulp), not the IEEE division sequence
__device__ __forceinline__ float rd_sigmoid(float z)
{
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.44269504f));
}
__device__ __forceinline__ float rd_silu(float z) { return z * rd_sigmoid(z); }
__device__ __forceinline__ float rd_dsilu(float z)
{
    const float s = rd_sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

// B operand for sample tile tj, K-block kb, from the lane's activations
__device__ __forceinline__ rd_bf16x8 rd_mlp_bop(const float (&a)[4][4][4], int tj, int kb)
{
    rd_bf16x8 v;
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = (__bf16)a[tj][2 * kb + t / 4][t & 3];
    return v;
}

#define RD_MFMA(A, B, C) __builtin_amdgcn_mfma_f32_16x16x32_bf16((A), (B), (C), 0, 0, 0)

// z = W h + b for a 64x64 layer; A(to, kb) returns the lane's weight fragment
template <class AF>
__device__ __forceinline__ void rd_mlp_layer64(AF&& A, const float* __restrict__ b, const float (&h)[4][4][4],
                                               float (&z)[4][4][4], int lane)
{
    rd_bf16x8 bop[4][2];
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) {
        bop[tj][0] = rd_mlp_bop(h, tj, 0);
        bop[tj][1] = rd_mlp_bop(h, tj, 1);
    }
    const int g = lane >> 4;
#pragma unroll
    for (int to = 0; to < 4; ++to) {
        const rd_bf16x8 a0 = A(to, 0), a1 = A(to, 1);
        rd_f32x4 bias;
#pragma unroll
        for (int r = 0; r < 4; ++r) bias[r] = b[to * 16 + 4 * g + r];
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) {
            rd_f32x4 acc = bias;
            acc = RD_MFMA(a0, bop[tj][0], acc);
            acc = RD_MFMA(a1, bop[tj][1], acc);
#pragma unroll
            for (int r = 0; r < 4; ++r) z[tj][to][r] = acc[r];
        }
    }
}

// layer 1 on the VALU: z1 = W1 xn + b1 for the lane's 4 samples
__device__ __forceinline__ void rd_mlp_layer1(const float* __restrict__ w1, const float* __restrict__ b1,
                                              const float (&xn)[4][3], float (&z)[4][4][4], int lane)
{
    const int g = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int f = ti * 16 + 4 * g + r;
            const float wa = w1[3 * f], wb = w1[3 * f + 1], wc = w1[3 * f + 2], bb = b1[f];
#pragma unroll
            for (int tj = 0; tj < 4; ++tj) {
                float v = wa * xn[tj][0];
                v = v + wb * xn[tj][1];
                v = v + wc * xn[tj][2];
                z[tj][ti][r] = v + bb;
            }
        }
}

// layer 4: y[tj][0..2] valid in lanes 0..15 (sample 16 tj + lane); A(kb)
// returns the lane's fragment of W4 (rows 3..15 zero)
template <class AF>
__device__ __forceinline__ void rd_mlp_layer4(AF&& A, const float* __restrict__ b4, const float (&h)[4][4][4],
                                              float (&y)[4][3], int lane)
{
    const rd_bf16x8 a0 = A(0), a1 = A(1);
    const int g = lane >> 4;
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) {
        rd_f32x4 acc;
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = (g == 0 && r < 3) ? b4[r] : 0.0f;
        acc = RD_MFMA(a0, rd_mlp_bop(h, tj, 0), acc);
        acc = RD_MFMA(a1, rd_mlp_bop(h, tj, 1), acc);
        y[tj][0] = acc[0]; y[tj][1] = acc[1]; y[tj][2] = acc[2];
    }
}

__device__ __forceinline__ void rd_mlp_silu_all(const float (&z)[4][4][4], float (&h)[4][4][4])
{
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) h[a][b][r] = rd_silu(z[a][b][r]);
}

// Move the outputs so that lane l holds sample l (y[tj] of lane l%16 for tj = l/16).
__device__ __forceinline__ float3 rd_mlp_gather_out(const float (&y)[4][3], int lane)
{
    const int src = lane & 15, tj = lane >> 4;
    float o[3] = { 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float v = __shfl(y[t][c], src, 64);
            if (t == tj) o[c] = v;
        }
    return make_float3(o[0], o[1], o[2]);
}

// sum over the 16 lanes of a lane group (lanes with equal lane>>4)
__device__ __forceinline__ float rd_sum16(float v)
{
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
}
