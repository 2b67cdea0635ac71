// pm_mlp.h -- the learned step proposal: a 24 -> 64 -> 64 -> 1 MLP evaluated by
// one wave for 64 samples at a time with BF16 MFMA (v_mfma_f32_16x16x32_bf16)
// and FP32 accumulation.  No counterpart in the reference; the closest
// analogue is the DensityGrid empty-space skip (Dataset.cpp:94-164,
// RayMarcher.cpp:284-304).
//
// Parameter vector (FP32 master copy, PM_MLP_NPARAM floats, this order):
//   W1[64][32]  (inputs 24..31 are padding and stay zero)
//   b1[64]  W2[64][64]  b2[64]  w3[64]  b3
// Activation: act(z) = z * sigmoid(1.702 z) (the sigmoid form of GELU).
// Output y estimates log2(1 + d/h), d = distance along the ray from the sample
// to the first iso crossing (or to the AABB exit on a miss).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pm_device.h"

#define PM_MLP_IN 24
#define PM_MLP_K1 32
#define PM_MLP_H 64
#define PM_MLP_OFF_W1 0
#define PM_MLP_OFF_B1 (PM_MLP_OFF_W1 + PM_MLP_H * PM_MLP_K1)
#define PM_MLP_OFF_W2 (PM_MLP_OFF_B1 + PM_MLP_H)
#define PM_MLP_OFF_B2 (PM_MLP_OFF_W2 + PM_MLP_H * PM_MLP_H)
#define PM_MLP_OFF_W3 (PM_MLP_OFF_B2 + PM_MLP_H)
#define PM_MLP_OFF_B3 (PM_MLP_OFF_W3 + PM_MLP_H)
#define PM_MLP_NPARAM (PM_MLP_OFF_B3 + 1)

typedef __bf16 pm_bf16x8 __attribute__((ext_vector_type(8)));
typedef float pm_f32x4 __attribute__((ext_vector_type(4)));

namespace pm {

// Network as the kernels read it: BF16 copies of W1, W2 and W2^T, FP32 rest.
struct MlpDev {
    const uint16_t *W1;    // [64][32] bf16 bits
    const uint16_t *W2;    // [64][64]
    const uint16_t *W2T;   // [64][64], W2T[i][o] = W2[o][i]
    const float *p;        // FP32 master vector (biases and w3 are read from it)
};

// float -> bf16, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ float mlp_act(float z) { return z / (1.0f + __expf(-1.702f * z)); }
__device__ __forceinline__ float mlp_act_grad(float z)
{
    float s = 1.0f / (1.0f + __expf(-1.702f * z));
    return s + 1.702f * z * s * (1.0f - s);
}

__device__ __forceinline__ pm_bf16x8 load_bf16x8(const uint16_t *p)
{
    uint4 v = *(const uint4 *)p;   // 16-byte aligned by construction
    return __builtin_bit_cast(pm_bf16x8, v);
}

// Wave-level ordering of LDS traffic between lanes (no block barrier needed:
// every wave owns its LDS region).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// C[64 x 16*NT] = A[64 x 32*KT] * B[16*NT x 32*KT]^T, bf16 in, fp32 out.
// A rows (stride lda) and B rows (stride ldb) hold K contiguously.
// MFMA 16x16x32 operand layout: lane l reads row (l % 16) of its A tile and of
// its B tile, K slots 8*(l/16) .. +8; the result element acc[mi][ni][r] is
// C[16 mi + 4 (l/16) + r][16 ni + l % 16].  A and B use the same K slots, so
// the sum does not depend on the order the hardware assigns to them.
template <int NT, int KT>
__device__ __forceinline__ void wave_gemm_abt(const uint16_t *A, int lda, const uint16_t *B, int ldb,
                                              pm_f32x4 (&acc)[4][NT])
{
    const int l = threadIdx.x & 63, row = l & 15, kb = l >> 4;
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int ni = 0; ni < NT; ni++) acc[mi][ni] = pm_f32x4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
    for (int kk = 0; kk < KT; kk++) {
        pm_bf16x8 a[4], b[NT];
#pragma unroll
        for (int mi = 0; mi < 4; mi++) a[mi] = load_bf16x8(A + (16 * mi + row) * lda + 32 * kk + 8 * kb);
#pragma unroll
        for (int ni = 0; ni < NT; ni++) b[ni] = load_bf16x8(B + (16 * ni + row) * ldb + 32 * kk + 8 * kb);
#pragma unroll
        for (int mi = 0; mi < 4; mi++)
#pragma unroll
            for (int ni = 0; ni < NT; ni++)
                acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mi], b[ni], acc[mi][ni], 0, 0, 0);
    }
}

// Forward pass of the wave's 64 samples, streaming one 16 x 16 output tile at a
// time so that only a few accumulators are live (the caller keeps its own ray
// state in registers).  X: LDS [64][32] bf16 sample rows written by the
// caller; H: LDS [64][64] bf16 scratch; Y: LDS [64] floats.  Returns y of
// sample `lane`.  Must be called by all 64 lanes together.
__device__ __forceinline__ float mlp_wave_forward(const MlpDev &net, const uint16_t *X, uint16_t *H, float *Y)
{
    const int l = threadIdx.x & 63, col = l & 15, kb = l >> 4;
    // layer 1: H = act(X W1^T + b1)
#pragma unroll 1
    for (int ni = 0; ni < 4; ni++) {
        const pm_bf16x8 b = load_bf16x8(net.W1 + (16 * ni + col) * PM_MLP_K1 + 8 * kb);
        const float bias = net.p[PM_MLP_OFF_B1 + 16 * ni + col];
#pragma unroll 1
        for (int mi = 0; mi < 4; mi++) {
            pm_bf16x8 a = load_bf16x8(X + (16 * mi + col) * PM_MLP_K1 + 8 * kb);
            pm_f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, pm_f32x4{ 0.f, 0.f, 0.f, 0.f }, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) H[(16 * mi + 4 * kb + r) * PM_MLP_H + 16 * ni + col] = f2bf(mlp_act(acc[r] + bias));
        }
    }
    wave_lds_sync();
    // layer 2 and the output unit: y = w3 . act(H W2^T + b2) + b3
    float part[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int r = 0; r < 4; r++) part[mi][r] = 0.f;
#pragma unroll 1
    for (int ni = 0; ni < 4; ni++) {
        const pm_bf16x8 b0 = load_bf16x8(net.W2 + (16 * ni + col) * PM_MLP_H + 8 * kb);
        const pm_bf16x8 b1 = load_bf16x8(net.W2 + (16 * ni + col) * PM_MLP_H + 32 + 8 * kb);
        const float bias = net.p[PM_MLP_OFF_B2 + 16 * ni + col], w3 = net.p[PM_MLP_OFF_W3 + 16 * ni + col];
#pragma unroll
        for (int mi = 0; mi < 4; mi++) {
            pm_bf16x8 a0 = load_bf16x8(H + (16 * mi + col) * PM_MLP_H + 8 * kb);
            pm_bf16x8 a1 = load_bf16x8(H + (16 * mi + col) * PM_MLP_H + 32 + 8 * kb);
            pm_f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, pm_f32x4{ 0.f, 0.f, 0.f, 0.f }, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) part[mi][r] += w3 * mlp_act(acc[r] + bias);
        }
    }
    // sum over the 16 lanes of a group (the 64 units are spread over col and ni)
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            float v = part[mi][r];
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            v += __shfl_xor(v, 4, 64);
            v += __shfl_xor(v, 8, 64);
            part[mi][r] = v;
        }
    if (col == 0) {
        const float b3 = net.p[PM_MLP_OFF_B3];
#pragma unroll
        for (int mi = 0; mi < 4; mi++)
#pragma unroll
            for (int r = 0; r < 4; r++) Y[16 * mi + 4 * kb + r] = part[mi][r] + b3;
    }
    wave_lds_sync();
    return Y[l];
}

// The same forward pass for LDS-tight callers: the 64 samples go through the
// network in two halves of 32, so the hidden layer needs a [32][64] buffer.
// X: LDS [64][32] bf16, H: LDS [32][64] bf16, Y: LDS [64] floats.
__device__ __forceinline__ float mlp_wave_forward_halves(const MlpDev &net, const uint16_t *X, uint16_t *H, float *Y)
{
    const int l = threadIdx.x & 63, col = l & 15, kb = l >> 4;
#pragma unroll 1
    for (int half = 0; half < 2; half++) {
#pragma unroll 1
        for (int ni = 0; ni < 4; ni++) {
            const pm_bf16x8 b = load_bf16x8(net.W1 + (16 * ni + col) * PM_MLP_K1 + 8 * kb);
            const float bias = net.p[PM_MLP_OFF_B1 + 16 * ni + col];
#pragma unroll
            for (int m2 = 0; m2 < 2; m2++) {
                const int mi = 2 * half + m2;
                pm_bf16x8 a = load_bf16x8(X + (16 * mi + col) * PM_MLP_K1 + 8 * kb);
                pm_f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, pm_f32x4{ 0.f, 0.f, 0.f, 0.f }, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; r++)
                    H[(16 * m2 + 4 * kb + r) * PM_MLP_H + 16 * ni + col] = f2bf(mlp_act(acc[r] + bias));
            }
        }
        wave_lds_sync();
        float part[2][4] = { { 0.f, 0.f, 0.f, 0.f }, { 0.f, 0.f, 0.f, 0.f } };
#pragma unroll 1
        for (int ni = 0; ni < 4; ni++) {
            const pm_bf16x8 b0 = load_bf16x8(net.W2 + (16 * ni + col) * PM_MLP_H + 8 * kb);
            const pm_bf16x8 b1 = load_bf16x8(net.W2 + (16 * ni + col) * PM_MLP_H + 32 + 8 * kb);
            const float bias = net.p[PM_MLP_OFF_B2 + 16 * ni + col], w3 = net.p[PM_MLP_OFF_W3 + 16 * ni + col];
#pragma unroll
            for (int m2 = 0; m2 < 2; m2++) {
                pm_bf16x8 a0 = load_bf16x8(H + (16 * m2 + col) * PM_MLP_H + 8 * kb);
                pm_bf16x8 a1 = load_bf16x8(H + (16 * m2 + col) * PM_MLP_H + 32 + 8 * kb);
                pm_f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, pm_f32x4{ 0.f, 0.f, 0.f, 0.f }, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; r++) part[m2][r] += w3 * mlp_act(acc[r] + bias);
            }
        }
#pragma unroll
        for (int m2 = 0; m2 < 2; m2++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                float v = part[m2][r];
                v += __shfl_xor(v, 1, 64);
                v += __shfl_xor(v, 2, 64);
                v += __shfl_xor(v, 4, 64);
                v += __shfl_xor(v, 8, 64);
                part[m2][r] = v;
            }
        if (col == 0) {
            const float b3 = net.p[PM_MLP_OFF_B3];
#pragma unroll
            for (int m2 = 0; m2 < 2; m2++)
#pragma unroll
                for (int r = 0; r < 4; r++) Y[16 * (2 * half + m2) + 4 * kb + r] = part[m2][r] + b3;
        }
        wave_lds_sync();
    }
    return Y[l];
}

}  // namespace pm
