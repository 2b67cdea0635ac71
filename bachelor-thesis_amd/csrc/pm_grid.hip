// pm_grid.hip -- GPU uniform grid: the replacement of CompactNSearch's
// add_point_set / z_sort / sort_field / update_point_sets (Dataset.cpp:49-76)
// and of Frame::ComputeAABB (Dataset.cpp:78-92).
//
// Pipeline (one stream, no host round trip except the extent read-back):
//   bounds reduce -> cell id + histogram (atomics) -> exclusive scan ->
//   scatter into cell-major float4 -> per-cell sort by input index.
// The per-cell sort makes the stored order deterministic (cell, input index),
// which is the canonical neighbour order the density sums follow.
// All passes are HBM-bound streaming passes over n particles or C cells.
#include "pm_internal.h"

#include <float.h>

namespace pm {

static constexpr int kThreads = 256;

uint32_t bounds_partials(uint64_t n)
{
    uint64_t b = (n + kThreads - 1) / kThreads;
    if (b > 1024) b = 1024;
    if (b < 1) b = 1;
    return (uint32_t)b;
}

__global__ __launch_bounds__(kThreads) void k_bounds_partial(const float *__restrict__ xyz, uint64_t n,
                                                             float *__restrict__ partial)
{
    float mn[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, mx[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        for (int a = 0; a < 3; a++) {
            float v = xyz[3 * i + a];
            mn[a] = fminf(mn[a], v);
            mx[a] = fmaxf(mx[a], v);
        }
    }
    __shared__ float red[6][kThreads];
    for (int a = 0; a < 3; a++) { red[a][threadIdx.x] = mn[a]; red[3 + a][threadIdx.x] = mx[a]; }
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int a = 0; a < 3; a++) {
                red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(kThreads) void k_bounds_final(const float *__restrict__ partial, uint32_t np,
                                                           float *__restrict__ out)
{
    __shared__ float red[6][kThreads];
    float v[6] = { FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX };
    for (uint32_t i = threadIdx.x; i < np; i += kThreads)
        for (int a = 0; a < 3; a++) {
            v[a] = fminf(v[a], partial[6 * i + a]);
            v[3 + a] = fmaxf(v[3 + a], partial[6 * i + 3 + a]);
        }
    for (int a = 0; a < 6; a++) red[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 3; a++) {
                red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = red[threadIdx.x][0];
}

hipError_t launch_bounds_reduce(const float *xyz, uint64_t n, float *partial, uint32_t npartial, float *out,
                                hipStream_t st)
{
    hipLaunchKernelGGL(k_bounds_partial, dim3(npartial), dim3(kThreads), 0, st, xyz, n, partial);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kThreads), 0, st, partial, npartial, out);
    return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint64_t n, uint32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ __launch_bounds__(kThreads) void k_cell_count(GridBuildArgs a)
{
    uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= a.n) return;
    int32_t cx = pm_key(a.xyz[3 * i + 0], a.inv_h) - a.ox;
    int32_t cy = pm_key(a.xyz[3 * i + 1], a.inv_h) - a.oy;
    int32_t cz = pm_key(a.xyz[3 * i + 2], a.inv_h) - a.oz;
    uint32_t c = (uint32_t)cx + (uint32_t)a.nx * ((uint32_t)cy + (uint32_t)a.ny * (uint32_t)cz);
    a.cell_of[i] = c;
    atomicAdd(&a.counts[c], 1u);
}

// ---- exclusive scan of u32 (len up to 2^31), 1024 elements per block ----
static constexpr int kScanPer = 4;
static constexpr int kScanBlock = kThreads * kScanPer;

size_t scan_tmp_size(uint32_t len)
{
    return (size_t)((len + kScanBlock - 1) / kScanBlock) + 1;
}

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *sh, uint32_t &total)
{
    // Hillis-Steele over 256 thread totals in LDS.
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        uint32_t add = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    total = sh[kThreads - 1];
    uint32_t incl = sh[threadIdx.x];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kThreads) void k_scan_blocks(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                          uint32_t len, uint32_t *__restrict__ block_sums)
{
    __shared__ uint32_t sh[kThreads];
    uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
    uint32_t v[kScanPer], sum = 0;
    for (int j = 0; j < kScanPer; j++) {
        v[j] = base + j < len ? in[base + j] : 0u;
        sum += v[j];
    }
    uint32_t total;
    uint32_t ex = block_excl_scan(sum, sh, total);
    for (int j = 0; j < kScanPer; j++) {
        if (base + j < len) out[base + j] = ex;
        ex += v[j];
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kThreads) void k_scan_sums(uint32_t *__restrict__ sums, uint32_t nb)
{
    __shared__ uint32_t sh[kThreads];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += kThreads) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t total;
        uint32_t ex = block_excl_scan(v, sh, total);
        if (i < nb) sums[i] = ex + carry;
        carry += total;
    }
}

__global__ __launch_bounds__(kThreads) void k_scan_add(uint32_t *__restrict__ out, uint32_t len,
                                                       const uint32_t *__restrict__ sums)
{
    uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
    uint32_t add = sums[blockIdx.x];
    for (int j = 0; j < kScanPer; j++)
        if (base + j < len) out[base + j] += add;
}

__global__ __launch_bounds__(kThreads) void k_scatter(GridBuildArgs a)
{
    uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= a.n) return;
    uint32_t c = a.cell_of[i];
    uint32_t slot = a.cell_start[c] + atomicAdd(&a.cursor[c], 1u);
    a.scattered[slot] = make_float4(a.xyz[3 * i + 0], a.xyz[3 * i + 1], a.xyz[3 * i + 2], __uint_as_float((uint32_t)i));
}

// Order inside each cell by input index: the particle at scattered slot s goes
// to cell_start[c] + (number of particles of its cell with a smaller index).
// One thread per particle, read-only over its own cell: no dependent swaps and
// no atomics, so dense cells cost O(k) per particle in parallel rather than an
// O(k^2) serial insertion sort.
__global__ __launch_bounds__(kThreads) void k_cell_rank(GridBuildArgs a)
{
    uint64_t s = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (s >= a.n) return;
    const float4 v = a.scattered[s];
    const uint32_t key = __float_as_uint(v.w);
    const uint32_t c = a.cell_of[key];
    const uint32_t b = a.cell_start[c], e = a.cell_start[c + 1];
    uint32_t rank = 0;
    for (uint32_t j = b; j < e; j++) rank += __float_as_uint(a.scattered[j].w) < key ? 1u : 0u;
    a.sorted[b + rank] = v;
}

hipError_t launch_scan_u32(const uint32_t *in, uint32_t *out, uint32_t len, uint32_t *tmp, hipStream_t st)
{
    uint32_t nb = (len + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kThreads), 0, st, in, out, len, tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kThreads), 0, st, tmp, nb);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kThreads), 0, st, out, len, tmp);
    return hipGetLastError();
}

hipError_t launch_grid_build(const GridBuildArgs &a, hipStream_t st)
{
    uint32_t len = a.ncells + 1;
    uint32_t fill_blocks = (uint32_t)((len + kThreads - 1) / kThreads);
    if (fill_blocks > 65535) fill_blocks = 65535;
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks), dim3(kThreads), 0, st, a.counts, (uint64_t)len, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks), dim3(kThreads), 0, st, a.cursor, (uint64_t)a.ncells, 0u);
    uint32_t pblocks = (uint32_t)((a.n + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_cell_count, dim3(pblocks), dim3(kThreads), 0, st, a);
    uint32_t nb = (len + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kThreads), 0, st, a.counts, a.cell_start, len, a.scan_tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kThreads), 0, st, a.scan_tmp, nb);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kThreads), 0, st, a.cell_start, len, a.scan_tmp);
    hipLaunchKernelGGL(k_scatter, dim3(pblocks), dim3(kThreads), 0, st, a);
    hipLaunchKernelGGL(k_cell_rank, dim3(pblocks), dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

// empty27[c] = 1 when cells c-1..c+1 on every axis hold no particle (border
// cells count missing neighbours as empty).
__global__ __launch_bounds__(kThreads) void k_empty27(const uint32_t *__restrict__ cs, int32_t nx, int32_t ny, int32_t nz,
                                                      uint8_t *__restrict__ out)
{
    uint32_t c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= (uint32_t)nx * ny * nz) return;
    int32_t x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
    int32_t x0 = x > 0 ? x - 1 : 0, x1 = x < nx - 1 ? x + 1 : nx - 1;
    uint32_t cnt = 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++) {
            int32_t yy = y + dy, zz = z + dz;
            if (yy < 0 || yy >= ny || zz < 0 || zz >= nz) continue;
            uint32_t row = (uint32_t)nx * ((uint32_t)yy + (uint32_t)ny * (uint32_t)zz);
            cnt += cs[row + x1 + 1] - cs[row + x0];
        }
    out[c] = cnt == 0;
}

hipError_t launch_empty27(const uint32_t *cs, int32_t nx, int32_t ny, int32_t nz, uint8_t *out, hipStream_t st)
{
    uint32_t n = (uint32_t)nx * ny * nz;
    hipLaunchKernelGGL(k_empty27, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, st, cs, nx, ny, nz, out);
    return hipGetLastError();
}

// Density bound per sub-box.  Every cell of the dense grid is split into
// 8x8x8 sub-boxes (eighth boxes); sub-box (fx, fy, fz) holds the points whose
// u = inv_h * x lies in [key + f/8, key + (f+1)/8) on every axis, and quarter
// box s = f / 2 (4x4x4 per cell) the eight sub-boxes with the same s.  For a
// box B widened by delta (which covers the rounding of u and of the march's
// sample positions), U = sum over the cell's 27-cell stencil of W(dist(x_j, B))
// bounds the density at every point of B, since W falls with r and the
// stencil holds every particle within h.  A sub-box's bit is set when U of its
// quarter box, or else its own U, is below the limit iso (1 - 2^-9); the margin
// covers the rounding of both sums.  Two rules spare sub-box sums that would
// find no bit: a cell whose centre row (dz = dy = 0) alone takes every quarter
// box's U past the limit lies inside the fluid and gets no bit, and neither
// does a quarter box whose U reaches 4 x the limit (a clear bit only means the
// sample is evaluated) -- except for the zero-density bound (iso = PM_ZERO_ISO),
// where a quarter box's U only says that some particle lies within h of it,
// not near which of its sub-boxes.  A cell's 512 bits are 64 bytes: byte
// l = sx + 4 sy + 16 sz belongs to quarter box s and its bit
// (fx & 1) + 2 (fy & 1) + 4 (fz & 1) to the sub-box inside it, so a byte of
// 0xff marks a quarter box the march can cross in one jump.  Each U sums the
// centre row (dz = dy = 0) first, then the other eight in (dz, dy) order,
// particles in storage order.  Border cells (no stencil, zero density) and
// cells with an empty 27-neighbourhood get every bit.  One wave per cell: lane
// l takes quarter box l, all lanes read the same particle; the quarter boxes'
// sums come first (one W per particle), the eight sub-box sums only for the
// quarter boxes between the limit and 4 x the limit, and each walk stops once
// every box it sums has reached the value that decides it.
// Squared distance along one axis from x to [lo, hi] (0 inside).
__device__ __forceinline__ float bound_e2(float lo, float hi, float x)
{
    const float e = fmaxf(fmaxf(lo - x, x - hi), 0.0f);
    return e * e;
}

// W at the distance from particle p to the box [l, h] (0 outside the support).
__device__ __forceinline__ float bound_term(const GridDev &g, float4 p, float lx, float hx, float ly, float hy, float lz,
                                            float hz)
{
    const float d2 = (bound_e2(lx, hx, p.x) + bound_e2(ly, hy, p.y)) + bound_e2(lz, hz, p.z);
    return d2 < g.h2 ? pm_W_in(d2, g.inv_h, g.sig) : 0.0f;
}

// The eight sub-box bounds of one quarter box, one particle at a time.
struct BoundEighths {
    float lo[3][2], hi[3][2];   // per axis, the two sub-boxes' widened ends
    float U[8];

    __device__ void init(const GridDev &g, const int32_t k[3], const int32_t s[3], float delta)
    {
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                lo[a][b] = ((float)k[a] + 0.125f * (float)(2 * s[a] + b)) * g.h - delta;
                hi[a][b] = ((float)k[a] + 0.125f * (float)(2 * s[a] + b + 1)) * g.h + delta;
            }
#pragma unroll
        for (int f = 0; f < 8; f++) U[f] = 0.0f;
    }
    __device__ void add(const GridDev &g, float4 p)
    {
        const float ex0 = bound_e2(lo[0][0], hi[0][0], p.x), ex1 = bound_e2(lo[0][1], hi[0][1], p.x);
        const float ey0 = bound_e2(lo[1][0], hi[1][0], p.y), ey1 = bound_e2(lo[1][1], hi[1][1], p.y);
        const float ez0 = bound_e2(lo[2][0], hi[2][0], p.z), ez1 = bound_e2(lo[2][1], hi[2][1], p.z);
        // no sub-box within h: all eight terms are 0
        if (!((fminf(ex0, ex1) + fminf(ey0, ey1)) + fminf(ez0, ez1) < g.h2)) return;
#pragma unroll
        for (int f = 0; f < 8; f++) {
            const float d2 = ((f & 1 ? ex1 : ex0) + (f & 2 ? ey1 : ey0)) + (f & 4 ? ez1 : ez0);
            U[f] += d2 < g.h2 ? pm_W_in(d2, g.inv_h, g.sig) : 0.0f;
        }
    }
    __device__ bool below(float lim) const
    {
        bool b = false;
#pragma unroll
        for (int f = 0; f < 8; f++) b = b || U[f] < lim;
        return b;
    }
};

// Row r of the bound walk's order: the centre row, then (dz, dy) in order.
__device__ __forceinline__ uint32_t bound_row(const GridDev &g, int32_t x, int32_t y, int32_t z, int r)
{
    const int dz = r == 0 ? 0 : (r - 1 + (r > 4)) / 3 - 1, dy = r == 0 ? 0 : (r - 1 + (r > 4)) % 3 - 1;
    return (uint32_t)x - 1u + (uint32_t)g.nx * ((uint32_t)(y + dy) + (uint32_t)g.ny * (uint32_t)(z + dz));
}

// The byte of quarter box `lane` of cell c (one wave; each lane stores its own).
// quarter_only: every byte 0xff or 0 (the quarter boxes' bounds alone).
__device__ __forceinline__ uint32_t bound_cell(const GridDev &g, float iso, bool quarter_only, uint32_t c, uint32_t lane)
{
    const uint32_t nx = (uint32_t)g.nx, ny = (uint32_t)g.ny;
    const int32_t x = (int32_t)(c % nx), y = (int32_t)((c / nx) % ny), z = (int32_t)(c / (nx * ny));
    if (x < 1 || y < 1 || z < 1 || x > g.nx - 2 || y > g.ny - 2 || z > g.nz - 2 || g.empty27[c]) return 0xffu;
    const int32_t k[3] = { x + g.ox, y + g.oy, z + g.oz };
    int32_t km = abs(k[0]) > abs(k[1]) ? abs(k[0]) : abs(k[1]);
    km = (km > abs(k[2]) ? km : abs(k[2])) + 1;
    const float delta = g.h * (0x1p-10f + (float)km * 0x1p-18f);
    const int32_t sq[3] = { (int32_t)(lane & 3u), (int32_t)((lane >> 2) & 3u), (int32_t)(lane >> 4) };
    const float lim = iso * (1.0f - 0x1p-9f), lim4 = iso == PM_ZERO_ISO ? INFINITY : 4.0f * lim;
    // the quarter box's bound; inside the fluid the centre row alone takes
    // every quarter box past the limit
    const float lx = ((float)k[0] + 0.25f * (float)sq[0]) * g.h - delta, hx = ((float)k[0] + 0.25f * (float)(sq[0] + 1)) * g.h + delta;
    const float ly = ((float)k[1] + 0.25f * (float)sq[1]) * g.h - delta, hy = ((float)k[1] + 0.25f * (float)(sq[1] + 1)) * g.h + delta;
    const float lz = ((float)k[2] + 0.25f * (float)sq[2]) * g.h - delta, hz = ((float)k[2] + 0.25f * (float)(sq[2] + 1)) * g.h + delta;
    float Uq = 0.0f;
    for (int r = 0; r < 9; r++) {
        const uint32_t row = bound_row(g, x, y, z, r);
        const uint32_t b = g.cell_start[row], e = g.cell_start[row + 3];
        // all lanes read the same particles (broadcast); 8 loads in flight
        uint32_t j = b;
        for (; j + 8 <= e; j += 8) {
            float4 p[8];
#pragma unroll
            for (int u = 0; u < 8; u++) p[u] = g.sorted[j + u];
#pragma unroll
            for (int u = 0; u < 8; u++) Uq += bound_term(g, p[u], lx, hx, ly, hy, lz, hz);
        }
        for (; j < e; j++) Uq += bound_term(g, g.sorted[j], lx, hx, ly, hy, lz, hz);
        if (r == 0 && __ballot(Uq < lim) == 0) return 0u;       // inside the fluid
        if (__ballot(Uq < (quarter_only ? lim : lim4)) == 0) break;
    }
    const bool whole = Uq < lim;
    if (quarter_only) return whole ? 0xffu : 0u;
    const bool eighths = !whole && Uq < lim4;
    if (__ballot(eighths) == 0) return whole ? 0xffu : 0u;
    // the eight sub-boxes of the quarter boxes left at or above the limit
    BoundEighths B;
    B.init(g, k, sq, delta);
    for (int r = 0; r < 9; r++) {
        if (eighths) {
            const uint32_t row = bound_row(g, x, y, z, r);
            const uint32_t b = g.cell_start[row], e = g.cell_start[row + 3];
            uint32_t j = b;
            for (; j + 4 <= e; j += 4) {
                float4 p[4];
#pragma unroll
                for (int u = 0; u < 4; u++) p[u] = g.sorted[j + u];
#pragma unroll
                for (int u = 0; u < 4; u++) B.add(g, p[u]);
            }
            for (; j < e; j++) B.add(g, g.sorted[j]);
        }
        if (__ballot(eighths && B.below(lim)) == 0) break;
    }
    if (whole) return 0xffu;
    if (!eighths) return 0u;
    uint32_t m = 0;
#pragma unroll
    for (int f = 0; f < 8; f++) m |= (B.U[f] < lim ? 1u : 0u) << f;
    return m;
}

// One wave per cell; bits[64 c + l] is the byte of quarter box l.
__global__ __launch_bounds__(256) void k_bound_grid(GridDev g, float iso, int quarter_only, uint8_t *__restrict__ out)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= (uint32_t)g.nx * (uint32_t)g.ny * (uint32_t)g.nz) return;   // whole wave
    out[64 * (uint64_t)c + lane] = (uint8_t)bound_cell(g, iso, quarter_only != 0, c, lane);
}

hipError_t launch_bound_grid(const GridDev &g, float iso, bool quarter_only, uint64_t *out, hipStream_t st)
{
    uint32_t n = (uint32_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_bound_grid, dim3((n + 3) / 4), dim3(256), 0, st, g, iso, quarter_only ? 1 : 0, (uint8_t *)out);
    return hipGetLastError();
}

// Reference DensityGrid node flags (Dataset.cpp:94-164).  Node (x,y,z) covers
// [bmin + (x,y,z) h, +h); its centre is bmin + ((x,y,z) + 0.5) h.  Two flags:
//   box  : some particle with |x_j - centre|_inf < h and count * W(0) > 1
//          (one reading of find_neighbors_box, whose source is absent);
//   cons : some particle with |x_j - centre|_inf < 1.5 h (1.001 margin), i.e.
//          a cleared node has zero density everywhere inside it.
__global__ __launch_bounds__(kThreads) void k_densitygrid(GridDev g, int32_t dnx, int32_t dny, int32_t dnz,
                                                          uint8_t *__restrict__ flag_box, uint8_t *__restrict__ flag_cons,
                                                          uint32_t *__restrict__ count_box)
{
    uint32_t id = blockIdx.x * kThreads + threadIdx.x;
    if (id >= (uint32_t)dnx * dny * dnz) return;
    int32_t k[3] = { (int32_t)(id % dnx), (int32_t)((id / dnx) % dny), (int32_t)(id / (dnx * dny)) };
    float c[3];
    const float cw = 1.0f * g.h;
    for (int a = 0; a < 3; a++) c[a] = g.bmin[a] + ((float)k[a] + 0.5f) * cw;
    const int32_t org[3] = { g.ox, g.oy, g.oz }, dim[3] = { g.nx, g.ny, g.nz };
    uint32_t res[2] = { 0u, 0u };
    const float R[2] = { g.h, 1.5f * g.h * 1.001f };
    for (int w = 0; w < 2; w++) {
        int32_t lo[3], hi[3];
        bool empty = false;
        for (int a = 0; a < 3; a++) {
            lo[a] = pm_key(c[a] - R[w], g.inv_h) - org[a];
            hi[a] = pm_key(c[a] + R[w], g.inv_h) - org[a];
            if (lo[a] < 0) lo[a] = 0;
            if (hi[a] > dim[a] - 1) hi[a] = dim[a] - 1;
            if (lo[a] > hi[a]) empty = true;
        }
        if (empty) continue;
        uint32_t cnt = 0;
        for (int32_t z = lo[2]; z <= hi[2]; z++)
            for (int32_t y = lo[1]; y <= hi[1]; y++) {
                uint32_t row = (uint32_t)g.nx * ((uint32_t)y + (uint32_t)g.ny * (uint32_t)z);
                uint32_t b = g.cell_start[row + lo[0]], e = g.cell_start[row + hi[0] + 1];
                for (uint32_t i = b; i < e; i++) {
                    float4 p = g.sorted[i];
                    if (fabsf(p.x - c[0]) < R[w] && fabsf(p.y - c[1]) < R[w] && fabsf(p.z - c[2]) < R[w]) cnt++;
                }
            }
        res[w] = cnt;
    }
    count_box[id] = res[0];
    flag_box[id] = (float)res[0] * g.sig > 1.0f;
    flag_cons[id] = res[1] > 0;
}

hipError_t launch_densitygrid(const GridDev &g, int32_t dnx, int32_t dny, int32_t dnz, uint8_t *flag_box, uint8_t *flag_cons,
                              uint32_t *count_box, hipStream_t st)
{
    uint32_t n = (uint32_t)dnx * dny * dnz;
    hipLaunchKernelGGL(k_densitygrid, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, st, g, dnx, dny, dnz, flag_box,
                       flag_cons, count_box);
    return hipGetLastError();
}

// Streaming copy (read n float4, write n float4): the measured-bandwidth
// reference point for the roofline and for scaling HBM byte counters.
__global__ __launch_bounds__(kThreads) void k_stream_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                          uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads)
        dst[i] = src[i];
}

hipError_t launch_stream_copy(const void *src, void *dst, uint64_t bytes, hipStream_t st)
{
    uint64_t n = bytes / 16;
    uint64_t blocks = (n + kThreads - 1) / kThreads;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_stream_copy, dim3((uint32_t)blocks), dim3(kThreads), 0, st, (const float4 *)src, (float4 *)dst, n);
    return hipGetLastError();
}

}  // namespace pm
