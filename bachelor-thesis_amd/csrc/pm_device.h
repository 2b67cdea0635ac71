// pm_device.h -- device-side data layout and the per-sample SPH density.
//
// Data layout in HBM (one frame), two grids over the same particles:
//   search grid (cells of h, CompactNSearch's keys): sorted[n] float4 (x, y, z,
//     bits of the input index), cell-major with x fastest and by ascending
//     input index inside a cell; cell_start[C+1] u32 exclusive scan of the
//     per-cell counts.  Neighbour counts, skip structures and features use it.
//   march grid (cells of h/2, the same key rule with inv_hf = 2 inv_h, four
//     empty cells of margin): fsorted[n], fcell_start[Cf+1], same layout.
// Every density sum walks the 5x5x5 fine cells around the sample's fine cell
// as x-rows (dz outer, then dy), each row the contiguous range of its fine
// cells; rows and row ends farther than h from the sample are dropped
// (pm_fine_row), which leaves about 40 % of the 27-h-cell candidates.
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
    // march grid
    const float4 *fsorted;
    const uint32_t *fcell_start;
    int32_t fox, foy, foz;    // fine key of fine cell (0,0,0)
    int32_t fnx, fny, fnz;
    float inv_hf;             // 2 / h
};

// ---- march-grid stencil ----
// In fine-cell units u = inv_hf x the sample sits at offset f in [0, 1] inside
// its fine cell (f = 1 only for the negative-integer key).  The distance from
// the sample to the cell at offset d along an axis is d - f (d > 0),
// f - (d + 1) (d < 0), 0 (d = 0); a row or row end is kept while its squared
// distance is below 4 (1 + 2^-10), h being 2 fine cells -- the margin only
// keeps cells, so every particle within h stays a candidate.
#define PM_FINE_LIM 4.00390625f

// Per-sample walk state: the centre cell's linear index, the strides and the
// squared distances to the x-offsets -2, -1, +1, +2 (the same for every row).
struct FineWalk {
    uint32_t base, sy, sz;
    float fy, fz;
    float xm2, xm1, xp1, xp2;
};

__device__ __forceinline__ float pm_fine_off(int d, float f)
{
    return d > 0 ? (float)d - f : (d < 0 ? f - (float)(d + 1) : 0.0f);
}

// Fine stencil centre; false if the 5x5x5 stencil would leave the march grid
// (then no particle lies within h of the point).
__device__ __forceinline__ bool pm_fine_begin(const GridDev &g, float x, float y, float z, FineWalk &w)
{
    const float ux = g.inv_hf * x, uy = g.inv_hf * y, uz = g.inv_hf * z;
    const int32_t kx = x >= 0.0f ? (int32_t)ux : (int32_t)ux - 1;
    const int32_t ky = y >= 0.0f ? (int32_t)uy : (int32_t)uy - 1;
    const int32_t kz = z >= 0.0f ? (int32_t)uz : (int32_t)uz - 1;
    const float fx = ux - (float)kx;
    w.fy = uy - (float)ky;
    w.fz = uz - (float)kz;
    const int32_t cx = kx - g.fox, cy = ky - g.foy, cz = kz - g.foz;
    w.sy = (uint32_t)g.fnx;
    w.sz = (uint32_t)g.fnx * (uint32_t)g.fny;
    w.base = (uint32_t)cx + w.sy * (uint32_t)cy + w.sz * (uint32_t)cz;
    float t;
    t = pm_fine_off(-2, fx); w.xm2 = t * t;
    t = pm_fine_off(-1, fx); w.xm1 = t * t;
    t = pm_fine_off(1, fx); w.xp1 = t * t;
    t = pm_fine_off(2, fx); w.xp2 = t * t;
    return cx >= 2 && cx <= g.fnx - 3 && cy >= 2 && cy <= g.fny - 3 && cz >= 2 && cz <= g.fnz - 3;
}

__device__ __forceinline__ float pm_fine_tz2(const FineWalk &w, int dz)
{
    const float t = pm_fine_off(dz, w.fz);
    return t * t;
}

// Stored range [b, e) of row (dy, dz) (tz2 = its squared z-offset); an empty
// range when the row is dropped.
__device__ __forceinline__ void pm_fine_row(const GridDev &g, const FineWalk &w, int dy, int dz, float tz2,
                                            uint32_t &b, uint32_t &e)
{
    const float ty = pm_fine_off(dy, w.fy);
    const float r2 = ty * ty + tz2;
    b = e = 0;
    if (!(r2 < PM_FINE_LIM)) return;
    const int lo = (w.xm2 + r2 < PM_FINE_LIM) ? -2 : ((w.xm1 + r2 < PM_FINE_LIM) ? -1 : 0);
    const int hi = (w.xp2 + r2 < PM_FINE_LIM) ? 2 : ((w.xp1 + r2 < PM_FINE_LIM) ? 1 : 0);
    const uint32_t row = w.base + (uint32_t)dy * w.sy + (uint32_t)dz * w.sz;
    b = g.fcell_start[row + (uint32_t)lo];
    e = g.fcell_start[row + (uint32_t)hi + 1];
}

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
    FineWalk fw;
    float s = 0.0f, c = 0.0f;
    if (!pm_fine_begin(g, qx, qy, qz, fw)) return 0.0f;
    _Pragma("unroll 1") for (int dz = -2; dz <= 2; dz++) {
        const float tz2 = pm_fine_tz2(fw, dz);
        _Pragma("unroll 1") for (int dy = -2; dy <= 2; dy++) {
            uint32_t b, e;
            pm_fine_row(g, fw, dy, dz, tz2, b, e);
            cand += e - b;
            for (uint32_t i = b; i < e; i++) {
                float4 p = g.fsorted[i];
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
// buffer when it is inside the support.  The store is unconditional (slot n
// is free and in bounds: the caller keeps n + remaining candidates <= CAP) and
// only the count depends on the test, so the loop has no branch.
__device__ __forceinline__ void pm_test_append(const GridDev &g, float4 p, float qx, float qy, float qz, float *buf,
                                               int &n)
{
    float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
    float r2 = (rx * rx + ry * ry) + rz * rz;
    buf[n * 64] = r2;
    n += r2 < g.h2 ? 1 : 0;
}

// UNROLL candidates are loaded before any is tested, so each lane keeps UNROLL
// independent gathers in flight instead of one dependent load per iteration.
template <int CAP, int UNROLL>
__device__ __forceinline__ float pm_density32_lds(const GridDev &g, float qx, float qy, float qz, float *buf,
                                                  uint32_t &cand)
{
    static_assert(UNROLL >= 1 && UNROLL <= CAP, "unroll");
    FineWalk fw;
    float s = 0.0f, c = 0.0f;
    bool in = pm_fine_begin(g, qx, qy, qz, fw);
    int n = 0;
    if (in) {
        _Pragma("unroll 1") for (int dz = -2; dz <= 2; dz++) {
            const float tz2 = pm_fine_tz2(fw, dz);
            _Pragma("unroll 1") for (int dy = -2; dy <= 2; dy++) {
                uint32_t b, e;
                pm_fine_row(g, fw, dy, dz, tz2, b, e);
                cand += e - b;
                if (__ballot(e > b) == 0) continue;         // no lane has candidates in this row
                if (__ballot(n + (int)(e - b) > CAP) != 0) {   // some lane might overflow:
                    pm_drain<CAP>(g, buf, n, s, c);           // all drain together
                    n = 0;
                }
                if ((int)(e - b) <= CAP) {
                    const float4 *p = g.fsorted + b;
                    const float4 *const pe = g.fsorted + e;
                    for (; p + UNROLL <= pe; p += UNROLL) {
                        float4 v[UNROLL];
#pragma unroll
                        for (int u = 0; u < UNROLL; u++) v[u] = p[u];
#pragma unroll
                        for (int u = 0; u < UNROLL; u++) pm_test_append(g, v[u], qx, qy, qz, buf, n);
                    }
                    for (; p < pe; p++) pm_test_append(g, *p, qx, qy, qz, buf, n);
                } else {                                // a row longer than the buffer
                    for (uint32_t i = b; i < e; i++) {
                        pm_test_append(g, g.fsorted[i], qx, qy, qz, buf, n);
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
    FineWalk fw;
    double acc = 0.0;
    if (!pm_fine_begin(g, (float)qx, (float)qy, (float)qz, fw)) return 0.0;
    _Pragma("unroll 1") for (int dz = -2; dz <= 2; dz++) {
        const float tz2 = pm_fine_tz2(fw, dz);
        _Pragma("unroll 1") for (int dy = -2; dy <= 2; dy++) {
            uint32_t b, e;
            pm_fine_row(g, fw, dy, dz, tz2, b, e);
            for (uint32_t i = b; i < e; i++) {
                float4 p = g.fsorted[i];
                double rx = (double)p.x - qx, ry = (double)p.y - qy, rz = (double)p.z - qz;
                double r2 = (rx * rx + ry * ry) + rz * rz;
                if (r2 < g.h2_d) acc += pm_W_in_d(r2, g.hinv_d, g.sig_d);
            }
        }
    }
    return acc;
}
