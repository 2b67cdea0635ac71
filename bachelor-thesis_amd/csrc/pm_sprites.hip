// pm_sprites.hip -- the depth-sprite pass that gives the reference march its
// ray starts: DepthRenderPass (DepthRenderPass.cpp:243-250 uniforms View,
// Projection, Radius), the sprite quads of AdvancedRenderer::CollectRenderData
// (AdvancedRenderer.cpp:455-474: p + R (+-x +-y) with x, y the first two rows
// of View, CameraController3D.cpp:80; UV (-1, -1) at p + R (-x + y)), depth.vert
// and depth.frag, depth test LESS against a clear of 1.0.
//
// Fragment at the centre of pixel (px, py), NDC ((px + 0.5) 2 / W - 1, ...):
// the pixel's ray from the camera meets the quad's plane at P; for a planar quad
// the perspective-correct interpolation of the vertices' view positions and
// UVs is View P and ((P - p).x / R, -(P - p).y / R).  depth.frag: discard when
// l2 = |UV|^2 > 1, else pView = View P - R (0, 0, cos(pi/2 sqrt(l2))) and the
// depth is (Projection pView).z / .w.  Fragments whose plane depth lies
// outside [0, 1] are clipped, written depths are clamped to [0, 1], and a
// particle with a quad corner behind the camera (w <= 0) is not drawn.  The
// minimum over all fragments of a pixel is kept (atomicMin on the bits of
// non-negative floats: the order of the particles does not matter), with an
// early depth test: at 1080p the dam break's 1M sprites put ~600 fragments on
// each drawn pixel, and only the nearer ones reach the atomic.
// cos(pi/2 s) is the Taylor polynomial through s^12 (error < 7e-9 on [0, 1]).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pm_internal.h"

namespace pm {

// glm mat4 * vec4 (x, y, z, 1), column-major M: (m0 x + m1 y) + (m2 z + m3).
__device__ __forceinline__ void sp_xform(const float *M, float x, float y, float z, float o[4])
{
#pragma unroll
    for (int r = 0; r < 4; r++) o[r] = (M[r] * x + M[4 + r] * y) + (M[8 + r] * z + M[12 + r] * 1.0f);
}

__device__ __forceinline__ float sp_cos_half_pi(float s)
{
    const float t = 1.57079632679f * s, z = t * t;
    float c = 1.0f / 479001600.0f;
    c = c * z + -1.0f / 3628800.0f;
    c = c * z + 1.0f / 40320.0f;
    c = c * z + -1.0f / 720.0f;
    c = c * z + 1.0f / 24.0f;
    c = c * z + -0.5f;
    return c * z + 1.0f;
}

// A sprite's screen footprint and the plane of its quad.
struct SpriteBox {
    int x0, x1, y0, y1;          // pixel range, clamped to the image; empty when x0 > x1 or y0 > y1
    float ax, ay, az, bx, by, bz;  // rows 0 and 1 of View: the quad's x and y
    float pn;                    // (p - eye) . (x cross y)
    float nx, ny, nz;            // x cross y
};

// False when the particle draws nothing (a quad corner behind the eye, or a
// footprint outside the image).
__device__ __forceinline__ bool sp_box(const SpriteLaunch &S, const float4 q, SpriteBox &B)
{
    const float R = S.radius;
    const float *V = S.view, *P = S.proj;
    B.ax = V[0]; B.ay = V[4]; B.az = V[8];      // row 0 of View: the quad's x
    B.bx = V[1]; B.by = V[5]; B.bz = V[9];      // row 1: its y
    float lo_x = 3.0e38f, hi_x = -3.0e38f, lo_y = 3.0e38f, hi_y = -3.0e38f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float sx = (k & 1) ? 1.0f : -1.0f, sy = (k & 2) ? 1.0f : -1.0f;
        const float cx = q.x + R * (sx * B.ax + sy * B.bx), cy = q.y + R * (sx * B.ay + sy * B.by),
                    cz = q.z + R * (sx * B.az + sy * B.bz);
        float v[4], c[4];
        sp_xform(V, cx, cy, cz, v);
        sp_xform(P, v[0], v[1], v[2], c);
        if (!(c[3] > 0.0f)) return false;
        const float fx = (c[0] / c[3] + 1.0f) * 0.5f * (float)S.W, fy = (c[1] / c[3] + 1.0f) * 0.5f * (float)S.H;
        lo_x = fminf(lo_x, fx); hi_x = fmaxf(hi_x, fx);
        lo_y = fminf(lo_y, fy); hi_y = fmaxf(hi_y, fy);
    }
    // pixel centres px + 0.5 inside [lo, hi], clamped to the image
    const float x0f = fmaxf(ceilf(lo_x - 0.5f), 0.0f), x1f = fminf(floorf(hi_x - 0.5f), (float)S.W - 1.0f);
    const float y0f = fmaxf(ceilf(lo_y - 0.5f), 0.0f), y1f = fminf(floorf(hi_y - 0.5f), (float)S.H - 1.0f);
    if (!(x0f <= x1f && y0f <= y1f)) return false;
    B.x0 = (int)x0f; B.x1 = (int)x1f; B.y0 = (int)y0f; B.y1 = (int)y1f;
    // plane of the quad: normal x cross y
    B.nx = B.ay * B.bz - B.az * B.by; B.ny = B.az * B.bx - B.ax * B.bz; B.nz = B.ax * B.by - B.ay * B.bx;
    B.pn = ((q.x - S.cam[0]) * B.nx + (q.y - S.cam[1]) * B.ny) + (q.z - S.cam[2]) * B.nz;
    return true;
}

// The fragment of sprite q at pixel (px, py), if any, into the depth image.
__device__ __forceinline__ void sp_fragment(const SpriteLaunch &S, const float4 q, const SpriteBox &B, int px, int py,
                                            uint32_t *depth)
{
    const float R = S.radius, invR = 1.0f / R;
    const float *V = S.view, *P = S.proj;
    const float o0 = S.cam[0], o1 = S.cam[1], o2 = S.cam[2];
    const float ndx = ((float)px + 0.5f) * (2.0f / (float)S.W) - 1.0f;
    const float ndy = ((float)py + 0.5f) * (2.0f / (float)S.H) - 1.0f;
    float w[4];
    sp_xform(S.inv_pv, ndx, ndy, 1.0f, w);
    // the ray direction scaled by w (no division): t d is unchanged; the
    // far point of a perspective projection has w = 1 / far > 0
    if (!(w[3] > 0.0f)) return;
    const float d0 = w[0] - o0 * w[3], d1 = w[1] - o1 * w[3], d2 = w[2] - o2 * w[3];
    const float den = (d0 * B.nx + d1 * B.ny) + d2 * B.nz;
    if (den == 0.0f) return;
    const float t = B.pn / den;
    if (!(t > 0.0f)) return;
    const float X = o0 + t * d0, Y = o1 + t * d1, Z = o2 + t * d2;
    const float rx = X - q.x, ry = Y - q.y, rz = Z - q.z;
    const float u = ((rx * B.ax + ry * B.ay) + rz * B.az) * invR;
    const float v = -(((rx * B.bx + ry * B.by) + rz * B.bz) * invR);
    const float l2 = u * u + v * v;
    if (l2 > 1.0f) return;                                     // depth.frag: discard
    float vp[4], cp[4], cs[4];
    sp_xform(V, X, Y, Z, vp);
    sp_xform(P, vp[0], vp[1], vp[2], cp);
    const float zp = cp[2] / cp[3];
    if (!(zp >= 0.0f && zp <= 1.0f)) return;                   // clipped
    const float off = sp_cos_half_pi(sqrtf(l2));
    sp_xform(P, vp[0], vp[1], vp[2] - R * off, cs);
    float z = cs[2] / cs[3];
    if (z != z) return;
    z = z <= 0.0f ? 0.0f : (z > 1.0f ? 1.0f : z);           // -0 -> +0: the bits order as the values
    // early depth test: only a nearer fragment pays for the atomic (a stale
    // read costs an atomic that changes nothing, never a wrong result)
    uint32_t *dst = &depth[(size_t)py * S.W + (size_t)px];
    const uint32_t zb = __float_as_uint(z);
    if (zb < __atomic_load_n(dst, __ATOMIC_RELAXED)) atomicMin(dst, zb);
}

// Footprints larger than this many pixels (a particle close to the eye: one
// just past the near plane covers the whole image) are not drawn by their own
// thread but listed for k_depth_sprites_big, where a wave shares the pixels.
constexpr uint32_t kSpriteBigPixels = 1024;

// One thread per particle: its footprint's pixel centres, or the big list.
// big[0] counts the listed sprites, big[1 .. cap] hold their indices; a sprite
// past the list's capacity is drawn here.  The depth is the minimum over all
// fragments (atomicMin), so who draws a sprite does not change the image.
__global__ __launch_bounds__(256) void k_depth_sprites(SpriteLaunch S, const float4 *__restrict__ pts, uint32_t n,
                                                       uint32_t *depth, uint32_t *big, uint32_t cap)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float4 q = pts[i];
    SpriteBox B;
    if (!sp_box(S, q, B)) return;
    const uint32_t area = (uint32_t)(B.x1 - B.x0 + 1) * (uint32_t)(B.y1 - B.y0 + 1);
    if (area > kSpriteBigPixels && big) {
        const uint32_t slot = atomicAdd(big, 1u);
        if (slot < cap) {
            big[1 + slot] = i;
            return;
        }
    }
    for (int py = B.y0; py <= B.y1; py++)
        for (int px = B.x0; px <= B.x1; px++) sp_fragment(S, q, B, px, py, depth);
}

// The listed sprites, one per wave at a time: the lanes take the footprint's
// pixels in turn.
__global__ __launch_bounds__(256) void k_depth_sprites_big(SpriteLaunch S, const float4 *__restrict__ pts,
                                                           uint32_t *depth, const uint32_t *big, uint32_t cap)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nlist = big[0] < cap ? big[0] : cap;
    const uint32_t waves = gridDim.x * 4u;
    for (uint32_t j = blockIdx.x * 4u + (threadIdx.x >> 6); j < nlist; j += waves) {
        const float4 q = pts[big[1 + j]];
        SpriteBox B;
        if (!sp_box(S, q, B)) continue;
        const uint32_t bw = (uint32_t)(B.x1 - B.x0 + 1);
        const uint32_t area = bw * (uint32_t)(B.y1 - B.y0 + 1);
        for (uint32_t k = lane; k < area; k += 64u)
            sp_fragment(S, q, B, B.x0 + (int)(k % bw), B.y0 + (int)(k / bw), depth);
    }
}

__global__ void k_zero_big(uint32_t *big)
{
    if (threadIdx.x == 0) big[0] = 0u;
}

__global__ void k_fill_one(uint32_t *p, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = 0x3f800000u;                                 // 1.0f: the depth clear value
}

hipError_t launch_depth_sprites(const SpriteLaunch &S, const float4 *pts, uint32_t n, uint32_t *depth, uint32_t *big,
                                uint32_t cap, hipStream_t st)
{
    const uint64_t npx = (uint64_t)S.W * S.H;
    hipLaunchKernelGGL(k_fill_one, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, st, depth, npx);
    if (!n) return hipGetLastError();
    if (big) hipLaunchKernelGGL(k_zero_big, dim3(1), dim3(64), 0, st, big);
    hipLaunchKernelGGL(k_depth_sprites, dim3((n + 255) / 256), dim3(256), 0, st, S, pts, n, depth, big, cap);
    if (big) hipLaunchKernelGGL(k_depth_sprites_big, dim3(1024), dim3(256), 0, st, S, pts, depth, (const uint32_t *)big, cap);
    return hipGetLastError();
}

}  // namespace pm
