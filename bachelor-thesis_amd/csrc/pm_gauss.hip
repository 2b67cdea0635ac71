// pm_gauss.hip -- screen-space depth smoothing: GaussRenderPass
// (GaussRenderPass.cpp:8-48 kernel table, 347-358 uniforms; gauss.frag).
//
// out(x, y) = sum_{i=-N..N} sum_{j=-N..N} K(|i|, |j|) depth(x + i, y + j),
// i (the horizontal offset) outer, j inner, as gauss.frag loops; the depth
// texture is addressed clamp-to-edge (BilateralBuffer.cpp:246-248) at texel
// centres (Spread = 1, GaussRenderPass.h:66), so every tap is a texel value.
// One 32x8 output tile per 256-thread workgroup: the (32 + 2N) x (8 + 2N)
// input window (edge-clamped) and the (N + 1)^2 weights are staged in LDS,
// then each thread sums its (2N + 1)^2 taps from LDS (a wave's 64 lanes read
// 32 consecutive floats of two rows: no bank conflicts).  HBM traffic is one
// read of the image (plus the halo) and one write.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pmarch.h"

namespace pm {

static constexpr int kTW = 32, kTH = 8;

__global__ __launch_bounds__(256) void k_depth_smooth(const float *__restrict__ in, float *__restrict__ out,
                                                      uint32_t W, uint32_t H, int N, const float *__restrict__ K)
{
    extern __shared__ float tile[];               // (kTW + 2N) x (kTH + 2N), then (N + 1)^2 weights
    const int tw = kTW + 2 * N, th = kTH + 2 * N;
    float *Ks = tile + tw * th;
    const int x0 = (int)blockIdx.x * kTW, y0 = (int)blockIdx.y * kTH;
    for (int k = (int)threadIdx.x; k < (N + 1) * (N + 1); k += 256) Ks[k] = K[k];
    for (int k = (int)threadIdx.x; k < tw * th; k += 256) {
        const int ty = k / tw, tx = k - ty * tw;
        int gx = x0 + tx - N, gy = y0 + ty - N;
        gx = gx < 0 ? 0 : (gx > (int)W - 1 ? (int)W - 1 : gx);
        gy = gy < 0 ? 0 : (gy > (int)H - 1 ? (int)H - 1 : gy);
        tile[k] = in[(size_t)gy * W + (size_t)gx];
    }
    __syncthreads();
    const int lx = (int)threadIdx.x % kTW, ly = (int)threadIdx.x / kTW;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= (int)W || y >= (int)H) return;
    float s = 0.0f;
    for (int i = -N; i <= N; i++) {
        const float *col = tile + (ly + N) * tw + (lx + N + i);
        const float *krow = Ks + (i < 0 ? -i : i) * (N + 1);
        for (int j = -N; j <= N; j++) s += krow[j < 0 ? -j : j] * col[j * tw];
    }
    out[(size_t)y * W + (size_t)x] = s;
}

hipError_t launch_depth_smooth(const float *in, float *out, uint32_t W, uint32_t H, int N, const float *K,
                               hipStream_t st)
{
    const dim3 grid((W + kTW - 1) / kTW, (H + kTH - 1) / kTH);
    const size_t lds = sizeof(float) * ((size_t)(kTW + 2 * N) * (kTH + 2 * N) + (size_t)(N + 1) * (N + 1));
    hipLaunchKernelGGL(k_depth_smooth, grid, dim3(256), lds, st, in, out, W, H, N, K);
    return hipGetLastError();
}

}  // namespace pm
