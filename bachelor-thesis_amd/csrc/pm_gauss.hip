// pm_gauss.hip -- screen-space depth smoothing: GaussRenderPass
// (GaussRenderPass.cpp:8-48 kernel table, 347-358 uniforms; gauss.frag).
//
// out(x, y) = sum_{i=-N..N} sum_{j=-N..N} K(|i|, |j|) depth(x + i, y + j),
// i (the horizontal offset) outer, j inner, as gauss.frag loops, each term
// accumulated with one fused multiply-add (gauss.frag's `sum += k * t` carries
// no `precise`, so the shader compiler may fuse it; fused here and in the
// oracle, so both round the same way); the depth texture is addressed
// clamp-to-edge (BilateralBuffer.cpp:246-248) at texel centres (Spread = 1,
// GaussRenderPass.h:66), so every tap is a texel value.
//
// N = 8 (the reference's default): one 32x32 output tile per 256-thread
// workgroup, contiguous runs of tiles per XCD; the edge-clamped 48x48 input
// window is staged in LDS and each
// thread computes 4 vertically adjacent outputs, so a column's 20 values are
// read from LDS once per horizontal offset and used by all 4 (85 LDS reads per
// output instead of 289); the weights are uniform and come through scalar
// loads.  Other N: one 32x8 tile per workgroup, each thread one output, the
// window and the (N + 1)^2 weights in LDS.  HBM traffic is one read of the
// image (plus the halo) and one write.
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
        for (int j = -N; j <= N; j++) s = __builtin_fmaf(krow[j < 0 ? -j : j], col[j * tw], s);
    }
    out[(size_t)y * W + (size_t)x] = s;
}

// Compile-time N: 32 x (8 R) outputs per workgroup, R per thread down a column.
// The horizontal offsets are a loop, not unrolled: unrolled, the compiler
// hoists every column's LDS reads ahead of the arithmetic (~350 VGPRs, one
// wave per SIMD).
template <int N, int R>
__global__ __launch_bounds__(256) void k_depth_smooth_n(const float *__restrict__ in, float *__restrict__ out,
                                                        uint32_t W, uint32_t H, const float *__restrict__ K)
{
    constexpr int TH = 8 * R, tw = kTW + 2 * N, th = TH + 2 * N;
    __shared__ float tile[tw * th];
    // Workgroups are dispatched to the 8 XCDs round-robin (blockIdx % 8 names
    // the XCD): each XCD takes a contiguous run of row-major tiles, so the
    // halos that neighbouring tiles share are read once into that XCD's L2.
    const uint32_t tiles_x = (W + kTW - 1) / kTW, T = tiles_x * ((H + TH - 1) / TH), per = (T + 7) / 8;
    const uint32_t t = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (t >= T) return;                            // whole workgroup, before the barrier
    const int x0 = (int)(t % tiles_x) * kTW, y0 = (int)(t / tiles_x) * TH;
    for (int k = (int)threadIdx.x; k < tw * th; k += 256) {
        const int ty = k / tw, tx = k - ty * tw;
        int gx = x0 + tx - N, gy = y0 + ty - N;
        gx = gx < 0 ? 0 : (gx > (int)W - 1 ? (int)W - 1 : gx);
        gy = gy < 0 ? 0 : (gy > (int)H - 1 ? (int)H - 1 : gy);
        tile[k] = in[(size_t)gy * W + (size_t)gx];
    }
    __syncthreads();
    const int lx = (int)threadIdx.x % kTW, ly = (int)threadIdx.x / kTW;
    float s[R];
#pragma unroll
    for (int m = 0; m < R; m++) s[m] = 0.0f;
#pragma unroll 1
    for (int i = -N; i <= N; i++) {
        // rows ly R .. ly R + R - 1 + 2N of window column lx + N + i
        float v[R + 2 * N];
        const float *col = tile + (ly * R) * tw + (lx + N + i);
#pragma unroll
        for (int k = 0; k < R + 2 * N; k++) v[k] = col[k * tw];
        const float *kr = K + (i < 0 ? -i : i) * (N + 1);     // uniform: scalar loads
        float kw[N + 1];
#pragma unroll
        for (int q = 0; q <= N; q++) kw[q] = kr[q];
#pragma unroll
        for (int m = 0; m < R; m++)
#pragma unroll
            for (int j = -N; j <= N; j++) s[m] = __builtin_fmaf(kw[j < 0 ? -j : j], v[m + j + N], s[m]);
    }
    const int x = x0 + lx;
    if (x >= (int)W) return;
#pragma unroll
    for (int m = 0; m < R; m++) {
        const int y = y0 + ly * R + m;
        if (y < (int)H) out[(size_t)y * W + (size_t)x] = s[m];
    }
}

hipError_t launch_depth_smooth(const float *in, float *out, uint32_t W, uint32_t H, int N, const float *K,
                               hipStream_t st)
{
    if (N == 8) {
        constexpr int R = 4;
        const uint32_t T = ((W + kTW - 1) / kTW) * ((H + 8 * R - 1) / (8 * R));
        hipLaunchKernelGGL((k_depth_smooth_n<8, R>), dim3(8 * ((T + 7) / 8)), dim3(256), 0, st, in, out, W, H, K);
        return hipGetLastError();
    }
    const dim3 grid((W + kTW - 1) / kTW, (H + kTH - 1) / kTH);
    const size_t lds = sizeof(float) * ((size_t)(kTW + 2 * N) * (kTH + 2 * N) + (size_t)(N + 1) * (N + 1));
    hipLaunchKernelGGL(k_depth_smooth, grid, dim3(256), lds, st, in, out, W, H, N, K);
    return hipGetLastError();
}

}  // namespace pm
