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
    const uint64_t *bound;    // per cell 8 words: the 512 eighth-box bits (set: the box stays below iso)
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

// Row r of the canonical order (dz = r / 5 - 2 outer, dy = r % 5 - 2 inner),
// without a branch: both bounds are loaded for every row (a dropped row loads
// its first cell's bounds, in range whenever pm_fine_begin accepted the
// centre) and a dropped row then gets the empty range.  r is wave-uniform.
__device__ __forceinline__ void pm_fine_row_at(const GridDev &g, const FineWalk &w, int r, uint32_t &b, uint32_t &e)
{
    const int dz = r / 5 - 2, dy = r - 5 * (r / 5) - 2;
    const float tz = pm_fine_off(dz, w.fz), ty = pm_fine_off(dy, w.fy);
    const float r2 = ty * ty + tz * tz;
    const bool kept = r2 < PM_FINE_LIM;
    const int lo = !kept ? 0 : ((w.xm2 + r2 < PM_FINE_LIM) ? -2 : ((w.xm1 + r2 < PM_FINE_LIM) ? -1 : 0));
    const int hi = !kept ? -1 : ((w.xp2 + r2 < PM_FINE_LIM) ? 2 : ((w.xp1 + r2 < PM_FINE_LIM) ? 1 : 0));
    const uint32_t row = w.base + (uint32_t)(dy * (int)w.sy + dz * (int)w.sz);
    const uint32_t lb = g.fcell_start[row + (uint32_t)lo];
    const uint32_t le = g.fcell_start[row + (uint32_t)(hi + 1)];
    b = kept ? lb : 0u;
    e = kept ? le : 0u;
}

// CompactNSearch cell key: int(x / h), minus one for negative x.
__device__ __forceinline__ int32_t pm_key(float x, float inv_h)
{
    float v = inv_h * x;
    return x >= 0.0f ? (int32_t)v : (int32_t)v - 1;
}

// sqrtf(x) for the kernel weight, x >= 0: the hardware square root (within one
// ulp) corrected to the correctly rounded result by the signs of the residuals
// x - s' s of its neighbours s' = s -+ 1 ulp -- the expansion the compiler emits
// for sqrtf at x >= 2^-96, without its scaling of smaller inputs and its
// special-value fix-up.  Below 2^-96 the hardware result is kept: there
// q = sqrt(x) / h < 2^-14 for any h > 2^-34, and W(q) = sigma (6 (q^3 - q^2) + 1)
// rounds to sigma exactly whatever the last bit of the root.
__device__ __forceinline__ float pm_sqrt_w(float x)
{
    float s = __builtin_amdgcn_sqrtf(x);
    if (x >= 0x1p-96f) {
        const float sm = __uint_as_float(__float_as_uint(s) - 1u);
        const float sp = __uint_as_float(__float_as_uint(s) + 1u);
        const float rm = __builtin_fmaf(-sm, s, x);
        const float rp = __builtin_fmaf(-sp, s, x);
        s = rm <= 0.0f ? sm : s;
        s = rp > 0.0f ? sp : s;
    }
    return s;
}

// Kernel.cpp:16-32 on a squared distance known to be < h^2.
__device__ __forceinline__ float pm_W_in(float r2, float inv_h, float sig)
{
    float q = pm_sqrt_w(r2) * inv_h;
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
// stencil candidates to cand and of particles within h to nsup.
__device__ __forceinline__ float pm_density32(const GridDev &g, float qx, float qy, float qz,
                                              bool compensated, uint32_t &cand, uint32_t &nsup)
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
            // groups of 4 loads, the last group's unused slots re-reading its first candidate
            const int m = (int)(e - b);
            const float4 *const p0 = g.fsorted + b;
            for (int i = 0; i < m; i += 4) {
                float4 v[4];
#pragma unroll
                for (int u = 0; u < 4; u++) v[u] = p0[i + u < m ? i + u : i];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    float rx = v[u].x - qx, ry = v[u].y - qy, rz = v[u].z - qz;
                    float r2 = (rx * rx + ry * ry) + rz * rz;
                    if (i + u < m && r2 < g.h2) {
                        float w = pm_W_in(r2, g.inv_h, g.sig);
                        if (compensated) pm_neu_add(s, c, w);
                        else s += w;
                        nsup++;
                    }
                }
            }
        }
    }
    return compensated ? s + c : s;
}

__device__ __forceinline__ float pm_density32(const GridDev &g, float qx, float qy, float qz,
                                              bool compensated, uint32_t &cand)
{
    uint32_t nsup = 0;
    return pm_density32(g, qx, qy, qz, compensated, cand, nsup);
}

// Candidates of q's own trimmed fine stencil (what pm_density32 visits), with no
// particle read: the per-sample count of instrumented launches.
__device__ __forceinline__ uint32_t pm_stencil_cand(const GridDev &g, float qx, float qy, float qz)
{
    FineWalk fw;
    if (!pm_fine_begin(g, qx, qy, qz, fw)) return 0u;
    uint32_t cand = 0;
    _Pragma("unroll 1") for (int dz = -2; dz <= 2; dz++) {
        const float tz2 = pm_fine_tz2(fw, dz);
        _Pragma("unroll 1") for (int dy = -2; dy <= 2; dy++) {
            uint32_t b, e;
            pm_fine_row(g, fw, dy, dz, tz2, b, e);
            cand += e - b;
        }
    }
    return cand;
}

// Particles stored in the 27 search cells (cells of h) around q's cell, the
// candidates of the reference's neighbour query (CompactNSearch's 27-cell
// scan, Dataset.cpp:272-280): 0 when the 3x3x3 block leaves the grid.
__device__ __forceinline__ uint32_t pm_cand27(const GridDev &g, float qx, float qy, float qz)
{
    int32_t cx, cy, cz;
    if (!pm_stencil(g, qx, qy, qz, cx, cy, cz)) return 0u;
    uint32_t cand = 0;
    _Pragma("unroll 1") for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++) {
            const uint32_t base = (uint32_t)(cx - 1) + (uint32_t)g.nx * ((uint32_t)(cy + dy) + (uint32_t)g.ny * (uint32_t)(cz + dz));
            cand += g.cell_start[base + 3] - g.cell_start[base];
        }
    return cand;
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
                                                  uint32_t &cand, uint32_t &nsup)
{
    static_assert(UNROLL >= 1 && UNROLL <= CAP, "unroll");
    FineWalk fw;
    float s = 0.0f, c = 0.0f;
    bool in = pm_fine_begin(g, qx, qy, qz, fw);
    int n = 0;
    uint32_t drained = 0;
    if (in) {
        _Pragma("unroll 1") for (int dz = -2; dz <= 2; dz++) {
            const float tz2 = pm_fine_tz2(fw, dz);
            _Pragma("unroll 1") for (int dy = -2; dy <= 2; dy++) {
                uint32_t b, e;
                pm_fine_row(g, fw, dy, dz, tz2, b, e);
                cand += e - b;
                if (__ballot(e > b) == 0) continue;         // no lane has candidates in this row
                if (__ballot(n + (int)(e - b) >= CAP) != 0) {  // some lane might fill the buffer:
                    pm_drain<CAP>(g, buf, n, s, c);           // all drain together
                    drained += (uint32_t)n;
                    n = 0;
                }
                if ((int)(e - b) < CAP) {
                    // groups of UNROLL loads, the last group's unused slots re-reading
                    // the group's first candidate; slot n stays free for their stores
                    const int m = (int)(e - b);
                    const float4 *const p0 = g.fsorted + b;
                    for (int i = 0; i < m; i += UNROLL) {
                        float4 v[UNROLL];
#pragma unroll
                        for (int u = 0; u < UNROLL; u++) v[u] = p0[i + u < m ? i + u : i];
#pragma unroll
                        for (int u = 0; u < UNROLL; u++) {
                            const float rx = v[u].x - qx, ry = v[u].y - qy, rz = v[u].z - qz;
                            const float r2 = (rx * rx + ry * ry) + rz * rz;
                            buf[n * 64] = r2;
                            n += (i + u < m && r2 < g.h2) ? 1 : 0;
                        }
                    }
                } else {
                    // a row longer than the buffer (dense clusters hold hundreds of
                    // particles per row): in chunks that fit, loaded UNROLL at a time
                    const float4 *p = g.fsorted + b;
                    const float4 *const pe = g.fsorted + e;
                    while (p < pe) {
                        const float4 *const ce = p + (CAP - n) < pe ? p + (CAP - n) : pe;
                        for (; p + UNROLL <= ce; p += UNROLL) {
                            float4 v[UNROLL];
#pragma unroll
                            for (int u = 0; u < UNROLL; u++) v[u] = p[u];
#pragma unroll
                            for (int u = 0; u < UNROLL; u++) pm_test_append(g, v[u], qx, qy, qz, buf, n);
                        }
                        for (; p < ce; p++) pm_test_append(g, *p, qx, qy, qz, buf, n);
                        if (p < pe) { pm_drain<CAP>(g, buf, n, s, c); drained += (uint32_t)n; n = 0; }
                    }
                }
            }
        }
    }
    pm_drain<CAP>(g, buf, n, s, c);
    nsup += drained + (uint32_t)n;
    return s + c;
}

template <int CAP, int UNROLL>
__device__ __forceinline__ float pm_density32_lds(const GridDev &g, float qx, float qy, float qz, float *buf,
                                                  uint32_t &cand)
{
    uint32_t nsup = 0;
    return pm_density32_lds<CAP, UNROLL>(g, qx, qy, qz, buf, cand, nsup);
}

// ---- fused central differences ----
// The six samples q +- e along x, y, z (in this order: +x, -x, +y, -y, +z,
// -z) lie in the box [q - e, q + e].  One walk over the march-grid cells near
// that box -- per axis from the fine cell of the box's low end to that of its
// high end, +-2 cells; rows and row ends kept by their distance to the box with
// the limit of pm_fine_row -- visits in the canonical order every cell that any
// of the six single-sample stencils keeps (the box is nearer to each cell than
// any of the six samples).  A candidate within (h + e)(1 + 2^-10) of q is kept
// and tested against each sample with that sample's own r^2 expression, so
// every sample's sum is its single-sample sum, bit for bit.
__device__ __forceinline__ float pm_box_off(float c, float ulo, float uhi, float kl, float kh)
{
    return c > kh ? c - uhi : (c < kl ? ulo - (c + 1.0f) : 0.0f);
}

struct BoxWalk {
    float ulo[3], uhi[3];     // box in fine-cell units
    float kl[3], kh[3];       // fine keys of the box ends
    uint32_t base, sy, sz;    // dense index of cell kl - 2 (per axis); strides
    int nx, ny, nz;           // cells per axis (kh - kl + 5)
    float xm2, xm1, xp1, xp2; // squared x-distances of the four trimmable end cells
};

// false when the cells leave the march grid (the caller then takes the six
// samples one at a time).
__device__ __forceinline__ bool pm_box_begin(const GridDev &g, const float lo[3], const float hi[3], BoxWalk &w)
{
    const int32_t fo[3] = { g.fox, g.foy, g.foz }, fn[3] = { g.fnx, g.fny, g.fnz };
    int32_t c0[3];
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        w.ulo[a] = g.inv_hf * lo[a];
        w.uhi[a] = g.inv_hf * hi[a];
        const int32_t kl = lo[a] >= 0.0f ? (int32_t)w.ulo[a] : (int32_t)w.ulo[a] - 1;
        const int32_t kh = hi[a] >= 0.0f ? (int32_t)w.uhi[a] : (int32_t)w.uhi[a] - 1;
        w.kl[a] = (float)kl;
        w.kh[a] = (float)kh;
        c0[a] = kl - 2 - fo[a];
        ok = ok && c0[a] >= 0 && kh + 2 - fo[a] <= fn[a] - 1;
        if (a == 0) w.nx = kh - kl + 5;
        if (a == 1) w.ny = kh - kl + 5;
        if (a == 2) w.nz = kh - kl + 5;
    }
    w.sy = (uint32_t)g.fnx;
    w.sz = (uint32_t)g.fnx * (uint32_t)g.fny;
    w.base = (uint32_t)c0[0] + w.sy * (uint32_t)c0[1] + w.sz * (uint32_t)c0[2];
    float t;
    t = pm_box_off(w.kl[0] - 2.0f, w.ulo[0], w.uhi[0], w.kl[0], w.kh[0]); w.xm2 = t * t;
    t = pm_box_off(w.kl[0] - 1.0f, w.ulo[0], w.uhi[0], w.kl[0], w.kh[0]); w.xm1 = t * t;
    t = pm_box_off(w.kh[0] + 1.0f, w.ulo[0], w.uhi[0], w.kl[0], w.kh[0]); w.xp1 = t * t;
    t = pm_box_off(w.kh[0] + 2.0f, w.ulo[0], w.uhi[0], w.kl[0], w.kh[0]); w.xp2 = t * t;
    return ok;
}

// Range [b, e) of row (jy, jz) (cell offsets from kl - 2), empty when the row
// is dropped or jz is past the box's rows; both bounds are always loaded.
__device__ __forceinline__ void pm_box_row(const GridDev &g, const BoxWalk &w, int jy, int jz, uint32_t &b,
                                           uint32_t &e)
{
    const bool valid = jz < w.nz;
    const float ty = pm_box_off(w.kl[1] - 2.0f + (float)jy, w.ulo[1], w.uhi[1], w.kl[1], w.kh[1]);
    const float tz = pm_box_off(w.kl[2] - 2.0f + (float)jz, w.ulo[2], w.uhi[2], w.kl[2], w.kh[2]);
    const float r2 = ty * ty + tz * tz;
    const bool kept = valid && r2 < PM_FINE_LIM;
    const int lo = !kept ? 0 : ((w.xm2 + r2 < PM_FINE_LIM) ? 0 : ((w.xm1 + r2 < PM_FINE_LIM) ? 1 : 2));
    const int hi = !kept ? -1 : ((w.xp2 + r2 < PM_FINE_LIM) ? w.nx - 1 : ((w.xp1 + r2 < PM_FINE_LIM) ? w.nx - 2 : w.nx - 3));
    const uint32_t row = valid ? w.base + (uint32_t)jy * w.sy + (uint32_t)jz * w.sz : w.base;
    const uint32_t lb = g.fcell_start[row + (uint32_t)lo];
    const uint32_t le = g.fcell_start[row + (uint32_t)(hi + 1)];
    b = kept ? lb : 0u;
    e = kept ? le : 0u;
}

struct Cdiff6 {
    float v[6];      // densities at +x, -x, +y, -y, +z, -z
    uint32_t cand;   // candidates of the walk; 0xffffffff: no fused walk
};

struct Acc6 {
    float s[6], c[6];
};

// Adds candidate p to the six sums (each sample's own r^2 expression, as in
// pm_density32: (rx rx + ry ry) + rz rz).
__device__ __forceinline__ void pm_acc6(const GridDev &g, float px, float py, float pz, const float x[3],
                                        const float xp[3], const float xm[3], bool comp, Acc6 &a)
{
    const float rx0 = px - x[0], ry0 = py - x[1], rz0 = pz - x[2];
    const float sx0 = rx0 * rx0, sy0 = ry0 * ry0, sz0 = rz0 * rz0;
    const float rxp = px - xp[0], rxm = px - xm[0];
    const float ryp = py - xp[1], rym = py - xm[1];
    const float rzp = pz - xp[2], rzm = pz - xm[2];
    float r2[6];
    r2[0] = (rxp * rxp + sy0) + sz0;
    r2[1] = (rxm * rxm + sy0) + sz0;
    r2[2] = (sx0 + ryp * ryp) + sz0;
    r2[3] = (sx0 + rym * rym) + sz0;
    const float sxy = sx0 + sy0;
    r2[4] = sxy + rzp * rzp;
    r2[5] = sxy + rzm * rzm;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        if (r2[k] < g.h2) {
            const float w = pm_W_in(r2[k], g.inv_h, g.sig);
            if (comp) pm_neu_add(a.s[k], a.c[k], w);
            else a.s[k] += w;
        }
    }
}

// The six central-difference densities of q = (x[0], x[1], x[2]) in one walk.
// A kept candidate is added to the six sums at once (about nine in ten pass the
// union test).
__device__ __forceinline__ Cdiff6 pm_cdiff6(const GridDev &g, float qx, float qy, float qz, float eps, bool comp)
{
    Cdiff6 out;
    const float x[3] = { qx, qy, qz };
    const float xp[3] = { qx + eps, qy + eps, qz + eps }, xm[3] = { qx - eps, qy - eps, qz - eps };
    BoxWalk w;
    if (!pm_box_begin(g, xm, xp, w)) {
        out.cand = 0xffffffffu;
#pragma unroll
        for (int k = 0; k < 6; k++) out.v[k] = 0.0f;
        return out;
    }
    const float he = (g.h + eps) * 1.0009765625f;
    const float R2 = he * he;
    Acc6 a;
#pragma unroll
    for (int k = 0; k < 6; k++) { a.s[k] = 0.0f; a.c[k] = 0.0f; }
    uint32_t cand = 0;
    int jy = 0, jz = 0;
    uint32_t nb, ne;
    pm_box_row(g, w, 0, 0, nb, ne);
    const int nr = w.ny * w.nz;
    _Pragma("unroll 1") for (int r = 0; __ballot(r < nr) != 0; r++) {
        const uint32_t b = nb, e = ne;
        cand += e - b;
        const bool any = __ballot(e > b) != 0;
        if (++jy == w.ny) { jy = 0; jz++; }
        pm_box_row(g, w, jy, jz, nb, ne);                 // next row in flight
        if (!any) continue;
        const int m = (int)(e - b);
        const float4 *const p0 = g.fsorted + b;
        for (int i = 0; i < m; i += 4) {
            float4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) v[u] = p0[i + u < m ? i + u : i];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const float rx = v[u].x - qx, ry = v[u].y - qy, rz = v[u].z - qz;
                if (i + u < m && (rx * rx + ry * ry) + rz * rz < R2) pm_acc6(g, v[u].x, v[u].y, v[u].z, x, xp, xm, comp, a);
            }
        }
    }
    out.cand = cand;
#pragma unroll
    for (int k = 0; k < 6; k++) out.v[k] = comp ? a.s[k] + a.c[k] : a.s[k];
    return out;
}

// The same six densities with the kept candidates compacted first through the
// lane's LDS slots (buf: lane-interleaved, CAP floats; entries of three floats)
// and added from there: better where every lane of the wave finishes a hit at
// once (the finish kernel), worse inside the march kernel's own finishing.
template <int CAP>
__device__ __forceinline__ Cdiff6 pm_cdiff6_lds(const GridDev &g, float qx, float qy, float qz, float eps, float *buf,
                                             bool comp)
{
    constexpr int CAPE = CAP / 3;
    Cdiff6 out;
    const float x[3] = { qx, qy, qz };
    const float xp[3] = { qx + eps, qy + eps, qz + eps }, xm[3] = { qx - eps, qy - eps, qz - eps };
    BoxWalk w;
    if (!pm_box_begin(g, xm, xp, w)) {
        out.cand = 0xffffffffu;
#pragma unroll
        for (int k = 0; k < 6; k++) out.v[k] = 0.0f;
        return out;
    }
    const float he = (g.h + eps) * 1.0009765625f;
    const float R2 = he * he;
    Acc6 a;
#pragma unroll
    for (int k = 0; k < 6; k++) { a.s[k] = 0.0f; a.c[k] = 0.0f; }
    uint32_t cand = 0;
    int n = 0;
    int jy = 0, jz = 0;
    uint32_t nb, ne;
    pm_box_row(g, w, 0, 0, nb, ne);
    const int nr = w.ny * w.nz;
    _Pragma("unroll 1") for (int r = 0; __ballot(r < nr) != 0; r++) {
        const uint32_t b = nb, e = ne;
        cand += e - b;
        const bool any = __ballot(e > b) != 0;
        if (++jy == w.ny) { jy = 0; jz++; }
        pm_box_row(g, w, jy, jz, nb, ne);                 // next row in flight
        if (!any) continue;
        const int m = (int)(e - b);
        const float4 *const p0 = g.fsorted + b;
        if (__ballot(n + m >= CAPE) != 0) {              // drain all lanes together
            for (int j = 0; j < n; j++)
                pm_acc6(g, buf[(3 * j) * 64], buf[(3 * j + 1) * 64], buf[(3 * j + 2) * 64], x, xp, xm, comp, a);
            n = 0;
        }
        if (m < CAPE) {
            for (int i = 0; i < m; i += 4) {
                float4 v[4];
#pragma unroll
                for (int u = 0; u < 4; u++) v[u] = p0[i + u < m ? i + u : i];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const float rx = v[u].x - qx, ry = v[u].y - qy, rz = v[u].z - qz;
                    buf[(3 * n) * 64] = v[u].x;
                    buf[(3 * n + 1) * 64] = v[u].y;
                    buf[(3 * n + 2) * 64] = v[u].z;
                    n += (i + u < m && (rx * rx + ry * ry) + rz * rz < R2) ? 1 : 0;
                }
            }
        } else {                                          // a row longer than the buffer
            for (int i = 0; i < m; i += 4) {
                float4 v[4];
#pragma unroll
                for (int u = 0; u < 4; u++) v[u] = p0[i + u < m ? i + u : i];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const float rx = v[u].x - qx, ry = v[u].y - qy, rz = v[u].z - qz;
                    if (i + u < m && (rx * rx + ry * ry) + rz * rz < R2) pm_acc6(g, v[u].x, v[u].y, v[u].z, x, xp, xm, comp, a);
                }
            }
        }
    }
    for (int j = 0; j < n; j++)
        pm_acc6(g, buf[(3 * j) * 64], buf[(3 * j + 1) * 64], buf[(3 * j + 2) * 64], x, xp, xm, comp, a);
    out.cand = cand;
#pragma unroll
    for (int k = 0; k < 6; k++) out.v[k] = comp ? a.s[k] + a.c[k] : a.s[k];
    return out;
}

// FP64 density; the stencil is chosen from the FP32-rounded point.
__device__ __forceinline__ double pm_density64(const GridDev &g, double qx, double qy, double qz)
{
    FineWalk fw;
    double acc = 0.0;
    if (!pm_fine_begin(g, (float)qx, (float)qy, (float)qz, fw)) return 0.0;
    // candidates are loaded in groups of 4 (the last group re-reads one of
    // the row's candidates in its unused slots) with the next row's bounds in
    // flight: one round trip per row while it has <= 4 candidates
    uint32_t nb, ne;
    pm_fine_row_at(g, fw, 0, nb, ne);
    _Pragma("unroll 1") for (int r = 0; r < 25; r++) {
        const uint32_t b = nb, e = ne;
        if (r < 24) pm_fine_row_at(g, fw, r + 1, nb, ne);
        const float4 *const p0 = g.fsorted + b;
        const int m = (int)(e - b);
        for (int i = 0; i < m; i += 4) {
            float4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) v[u] = p0[i + u < m ? i + u : i];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                double rx = (double)v[u].x - qx, ry = (double)v[u].y - qy, rz = (double)v[u].z - qz;
                double r2 = (rx * rx + ry * ry) + rz * rz;
                if (i + u < m && r2 < g.h2_d) acc += pm_W_in_d(r2, g.hinv_d, g.sig_d);
            }
        }
    }
    return acc;
}
