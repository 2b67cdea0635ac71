// pm_device.h -- device-side data layout and the per-sample SPH density.
//
// Data layout in HBM (one frame):
//   sorted[n]        float4 (x, y, z, bits of the input index), cell-major with x
//                    fastest, and by ascending input index inside a cell
//   cell_start[C+1]  u32 exclusive scan of per-cell counts, C = dims.x*dims.y*dims.z
// A density sample reads the 27 cells around its own cell as 9 x-rows of three
// consecutive cells: row (dy, dz) is the contiguous range
//   [cell_start[lin(cx-1, cy+dy, cz+dz)], cell_start[lin(cx+1, cy+dy, cz+dz) + 1]).
//
// All float arithmetic is written in the order of the reference kernel
// (src/app/Kernel.cpp:16-52) and compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PM_PI_F 3.14159265358979323846f
#define PM_PI_D 3.14159265358979323846

struct GridDev {
    const float4 *sorted;
    const uint32_t *cell_start;
    int32_t ox, oy, oz;       // key of dense cell (0,0,0)
    int32_t nx, ny, nz;       // dense dims
    float h, inv_h, h2, sig;  // support, 1/h, h*h, 8/(pi h^3)
    double h_d, h2_d, hinv_d, sig_d;
    float bmin[3], bmax[3];   // padded AABB
    const uint8_t *empty27;   // per cell: 1 if the cell and its 26 neighbours are empty
    const uint64_t *bound;    // per cell: bit l set when sub-box l (4x4x4) stays below iso
    // reference DensityGrid over the padded AABB (Dataset.cpp:94-164)
    const uint8_t *dg_flag;   // flag per node for the selected skip rule
    int32_t dg_nx, dg_ny, dg_nz;
};

// CompactNSearch cell key: int(x / h), minus one for negative x.
__device__ __forceinline__ int32_t pm_key(float x, float inv_h)
{
    float v = inv_h * x;
    return x >= 0.0f ? (int32_t)v : (int32_t)v - 1;
}

// Kernel.cpp:16-32 on a squared distance known to be < h^2.
__device__ __forceinline__ float pm_W_in(float r2, float inv_h, float sig)
{
    float q = sqrtf(r2) * inv_h;
    if (q >= 0.5f) {
        float qq = 1.0f - q;
        return sig * (2.0f * qq * qq * qq);
    }
    return sig * (6.0f * (q * q * q - q * q) + 1.0f);
}

__device__ __forceinline__ double pm_W_in_d(double r2, double hinv, double sig)
{
    double q = sqrt(r2) * hinv;
    if (q >= 0.5) {
        double qq = 1.0 - q;
        return sig * (2.0 * qq * qq * qq);
    }
    return sig * (6.0 * (q * q * q - q * q) + 1.0);
}

// Neumaier-compensated accumulation (exact same branches as the oracle).
__device__ __forceinline__ void pm_neu_add(float &s, float &c, float w)
{
    float t = s + w;
    bool big = fabsf(s) >= fabsf(w);
    float a = big ? s : w;
    float b = big ? w : s;
    c += (a - t) + b;
    s = t;
}

// Dense coordinates of the stencil centre; false if the stencil would leave
// the grid (then no particle lies within h of the point).
__device__ __forceinline__ bool pm_stencil(const GridDev &g, float x, float y, float z,
                                           int32_t &cx, int32_t &cy, int32_t &cz)
{
    cx = pm_key(x, g.inv_h) - g.ox;
    cy = pm_key(y, g.inv_h) - g.oy;
    cz = pm_key(z, g.inv_h) - g.oz;
    return cx >= 1 && cx <= g.nx - 2 && cy >= 1 && cy <= g.ny - 2 && cz >= 1 && cz <= g.nz - 2;
}

// rho(q) = sum_j W(x_j - q), FP32, canonical order; adds the number of
// stencil candidates to cand.
__device__ __forceinline__ float pm_density32(const GridDev &g, float qx, float qy, float qz,
                                              bool compensated, uint32_t &cand)
{
    int32_t cx, cy, cz;
    float s = 0.0f, c = 0.0f;
    if (!pm_stencil(g, qx, qy, qz, cx, cy, cz)) return 0.0f;
    for (int dz = -1; dz <= 1; dz++) {
        for (int dy = -1; dy <= 1; dy++) {
            uint32_t base = (uint32_t)(cx - 1) + (uint32_t)g.nx * ((uint32_t)(cy + dy) + (uint32_t)g.ny * (uint32_t)(cz + dz));
            uint32_t b = g.cell_start[base], e = g.cell_start[base + 3];
            cand += e - b;
            for (uint32_t i = b; i < e; i++) {
                float4 p = g.sorted[i];
                float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
                float r2 = (rx * rx + ry * ry) + rz * rz;
                if (r2 < g.h2) {
                    float w = pm_W_in(r2, g.inv_h, g.sig);
                    if (compensated) pm_neu_add(s, c, w);
                    else s += w;
                }
            }
        }
    }
    return compensated ? s + c : s;
}

// Same sum as pm_density32, with the in-support candidates compacted through a
// per-lane LDS buffer before W is evaluated.  A wave walks its rows together;
// the distance test is cheap and the W evaluation (sqrt, branch, compensated
// add) then runs over each lane's own short list, so lanes whose candidates are
// mostly out of support no longer hold the wave in the expensive branch.  The
// buffer is drained in order, so the summation order -- and the result -- is
// exactly that of pm_density32.
//   buf: this lane's slot 0; slot j is buf[j * 64] (lane-interleaved: no bank
//   conflicts whatever slot each lane is at).
template <int CAP>
__device__ __forceinline__ void pm_drain(const GridDev &g, const float *buf, int n, float &s, float &c)
{
    for (int j = 0; j < n; j++) {
        float w = pm_W_in(buf[j * 64], g.inv_h, g.sig);
        pm_neu_add(s, c, w);
    }
}

// Distance test of candidate p against q; appends |p - q|^2 to the lane's
// buffer when it is inside the support.
__device__ __forceinline__ void pm_test_append(const GridDev &g, float4 p, float qx, float qy, float qz, float *buf,
                                               int &n)
{
    float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
    float r2 = (rx * rx + ry * ry) + rz * rz;
    if (r2 < g.h2) { buf[n * 64] = r2; n++; }
}

// UNROLL candidates are loaded before any is tested, so each lane keeps UNROLL
// independent gathers in flight instead of one dependent load per iteration.
template <int CAP, int UNROLL>
__device__ __forceinline__ float pm_density32_lds(const GridDev &g, float qx, float qy, float qz, float *buf,
                                                  uint32_t &cand)
{
    static_assert(UNROLL >= 1 && UNROLL <= CAP, "unroll");
    int32_t cx, cy, cz;
    float s = 0.0f, c = 0.0f;
    bool in = pm_stencil(g, qx, qy, qz, cx, cy, cz);
    int n = 0;
    if (in) {
        for (int dz = -1; dz <= 1; dz++) {
            for (int dy = -1; dy <= 1; dy++) {
                uint32_t base = (uint32_t)(cx - 1) + (uint32_t)g.nx * ((uint32_t)(cy + dy) + (uint32_t)g.ny * (uint32_t)(cz + dz));
                uint32_t b = g.cell_start[base], e = g.cell_start[base + 3];
                cand += e - b;
                if (__ballot(n + (int)(e - b) > CAP) != 0) {   // some lane might overflow:
                    pm_drain<CAP>(g, buf, n, s, c);           // all drain together
                    n = 0;
                }
                if ((int)(e - b) <= CAP) {
                    uint32_t i = b;
                    for (; i + UNROLL <= e; i += UNROLL) {
                        float4 p[UNROLL];
#pragma unroll
                        for (int u = 0; u < UNROLL; u++) p[u] = g.sorted[i + u];
#pragma unroll
                        for (int u = 0; u < UNROLL; u++) pm_test_append(g, p[u], qx, qy, qz, buf, n);
                    }
                    for (; i < e; i++) pm_test_append(g, g.sorted[i], qx, qy, qz, buf, n);
                } else {                                // a row longer than the buffer
                    for (uint32_t i = b; i < e; i++) {
                        pm_test_append(g, g.sorted[i], qx, qy, qz, buf, n);
                        if (n == CAP) { pm_drain<CAP>(g, buf, n, s, c); n = 0; }
                    }
                }
            }
        }
    }
    pm_drain<CAP>(g, buf, n, s, c);
    return s + c;
}

// FP64 density; the stencil is chosen from the FP32-rounded point.
__device__ __forceinline__ double pm_density64(const GridDev &g, double qx, double qy, double qz)
{
    int32_t cx, cy, cz;
    double acc = 0.0;
    if (!pm_stencil(g, (float)qx, (float)qy, (float)qz, cx, cy, cz)) return 0.0;
    for (int dz = -1; dz <= 1; dz++) {
        for (int dy = -1; dy <= 1; dy++) {
            uint32_t base = (uint32_t)(cx - 1) + (uint32_t)g.nx * ((uint32_t)(cy + dy) + (uint32_t)g.ny * (uint32_t)(cz + dz));
            uint32_t b = g.cell_start[base], e = g.cell_start[base + 3];
            for (uint32_t i = b; i < e; i++) {
                float4 p = g.sorted[i];
                double rx = (double)p.x - qx, ry = (double)p.y - qy, rz = (double)p.z - qz;
                double r2 = (rx * rx + ry * ry) + rz * rz;
                if (r2 < g.h2_d) acc += pm_W_in_d(r2, g.hinv_d, g.sig_d);
            }
        }
    }
    return acc;
}
